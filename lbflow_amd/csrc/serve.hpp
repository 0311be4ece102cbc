// serve.hpp — batched step of many resident cases (serving tier).
//
// Every active case owns a slot in a device table: its step arguments (the
// same KArgs its single-case launches use), the steps it still has to run and
// the step index whose launch starts from natural slots. A tile is one
// workgroup's worth of nodes of one case: (slot << 24) | block. Per step and
// case type (lattice x storage x operator), two launches walk that type's tile
// list, one per parity class of its cases, and a third runs the last step of
// the cases that end at that step; each uses a fixed, tunable number of
// workgroups of a tunable tile size (TILE nodes, one per thread), so the
// launch shape does not depend on which cases are resident and a captured
// graph of K steps stays valid while cases come and go: admission and
// retirement only rewrite the tables.
//
// Each tile runs exactly the per-node code of the single-case AA kernel
// (aa_node) for the case's parity, with its fresh flag read from its slot, so
// a case's populations are bitwise the same whether it runs alone, in a mix,
// or through lbf_main_steps. The last step of a case also stores moments, as
// the single-case path does at the end of a batch.
#pragma once

#include "dispatch.hpp"
#include "fast.hpp"

namespace lbf {

#ifndef LBF_SERVE_WAVES_16_EVEN
#define LBF_SERVE_WAVES_16_EVEN 7
#endif
#ifndef LBF_SERVE_WAVES_16_ODD
#define LBF_SERVE_WAVES_16_ODD 5
#endif

constexpr int kServeSlotBits = 8;
constexpr int kServeMaxSlots = 1 << kServeSlotBits;
constexpr uint32_t kServeBlockMask = (1u << (32 - kServeSlotBits)) - 1;
// tile sizes (nodes per tile = threads per workgroup) the autotuner picks from
constexpr int kNumTileSizes = 3;
constexpr int kTileSizes[kNumTileSizes] = {128, 256, 512};
constexpr int kDefaultTileIdx = 1;  // 256, the single-case kernels' workgroup

struct ServeSlot {
  KArgs a;                        // single-case step arguments (flags: store moments only)
  long long remaining;            // steps left; 0 = slot idle
  unsigned long long fresh_step;  // this step index reads natural slots (~0: none)
  unsigned long long step;        // the case's step count (equal to a.ctrl->step)
  uint32_t type_bit;              // 1 << case type (serve_advance's last-step flags)
  uint32_t pad;
};

// XCD-aware walk order: workgroups are dealt round-robin to the 8 XCDs. When
// the grid is a multiple of 8 G, permute the ids within each group of 8 G so
// that every XCD walks runs of G consecutive tiles (neighbouring rows share its
// L2); the walk still visits each tile exactly once.
__device__ __forceinline__ uint32_t serve_first_tile() {
  constexpr uint32_t G = 4;
  uint32_t first = blockIdx.x;
  if (gridDim.x % (8 * G) == 0) {
    const uint32_t w = blockIdx.x >> 3;
    first = ((w / G) * 8u + (blockIdx.x & 7u)) * G + w % G;
  }
  return first;
}

// One step of the cases of one parity class of one type. The type's tile list
// holds the class-0 tiles (cases whose step count was even when the tables
// were written) in [0, counts[0]) and the class-1 tiles in [counts[0],
// counts[0] + counts[1]); the host launches each class with the parity its
// cases have at that step of the graph, so the step body is compiled for one
// parity and without the moment stores, like the single-case streaming
// kernels (and at their occupancy). A case's last step (remaining == 1) is
// left to serve_last.
template <class L, class St, int OP, int TILE, bool ODD>
__global__ void __launch_bounds__(TILE)
    __attribute__((amdgpu_waves_per_eu(sizeof(typename St::S) == 2
                                           ? (ODD ? LBF_SERVE_WAVES_16_ODD : LBF_SERVE_WAVES_16_EVEN)
                                           : (ODD ? LBF_WAVES_32_ODD : LBF_WAVES_32_EVEN))))
    serve_step(const ServeSlot* __restrict__ slots, const uint32_t* __restrict__ tiles,
               const uint32_t* __restrict__ counts, int cls) {
  const uint32_t n0 = counts[0];
  const uint32_t begin = cls ? n0 : 0u, end = cls ? n0 + counts[1] : n0;
  double acc[5] = {0, 0, 0, 0, 0};
  for (uint32_t t = begin + serve_first_tile(); t < end; t += gridDim.x) {
    const uint32_t tile = tiles[t];
    const ServeSlot& s = slots[tile >> (32 - kServeSlotBits)];
    if (s.remaining <= 1) continue;  // finished within this replay, or its last step
    const unsigned long long step = s.step;
    const uint32_t n = (tile & kServeBlockMask) * TILE + threadIdx.x;
    aa_node<L, St, OP, ODD, false>(s.a, n, step, step == s.fresh_step, acc);
  }
}

// The last step of every case of one type that ends at this step: the same
// node code with the moment stores, as the single-case path runs the last
// step of a batch. serve_advance sets the type's bit in *last_flags when some
// case of that type has one step left; otherwise every workgroup returns at
// once.
template <class L, class St, int OP, int TILE>
__global__ void __launch_bounds__(TILE)
    serve_last(const ServeSlot* __restrict__ slots, const uint32_t* __restrict__ tiles,
               const uint32_t* __restrict__ counts, const uint32_t* __restrict__ last_flags,
               uint32_t type_bit) {
  if (!(*last_flags & type_bit)) return;
  const uint32_t end = counts[0] + counts[1];
  double acc[5] = {0, 0, 0, 0, 0};
  for (uint32_t t = serve_first_tile(); t < end; t += gridDim.x) {
    const uint32_t tile = tiles[t];
    const ServeSlot& s = slots[tile >> (32 - kServeSlotBits)];
    if (s.remaining != 1) continue;
    const unsigned long long step = s.step;
    const bool fresh = step == s.fresh_step;
    const uint32_t n = (tile & kServeBlockMask) * TILE + threadIdx.x;
    if (step & 1)
      aa_node<L, St, OP, true, true>(s.a, n, step, fresh, acc);
    else
      aa_node<L, St, OP, false, true>(s.a, n, step, fresh, acc);
  }
}

template <class L, class St, int OP, int TILE>
hipError_t serve_launch(const ServeSlot* slots, const uint32_t* tiles, const uint32_t* counts,
                        int cls, int odd, int grid, hipStream_t s) {
  if (odd)
    hipLaunchKernelGGL((serve_step<L, St, OP, TILE, true>), dim3(grid), dim3(TILE), 0, s, slots,
                       tiles, counts, cls);
  else
    hipLaunchKernelGGL((serve_step<L, St, OP, TILE, false>), dim3(grid), dim3(TILE), 0, s, slots,
                       tiles, counts, cls);
  return hipGetLastError();
}

template <class L, class St, int OP, int TILE>
hipError_t serve_launch_last(const ServeSlot* slots, const uint32_t* tiles,
                             const uint32_t* counts, const uint32_t* last_flags,
                             uint32_t type_bit, int grid, hipStream_t s) {
  hipLaunchKernelGGL((serve_last<L, St, OP, TILE>), dim3(grid), dim3(TILE), 0, s, slots, tiles,
                     counts, last_flags, type_bit);
  return hipGetLastError();
}

}  // namespace lbf

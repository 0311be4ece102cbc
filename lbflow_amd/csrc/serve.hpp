// serve.hpp — batched step of many resident cases (serving tier).
//
// Every active case owns a slot in a device table: its step arguments (the
// same KArgs its single-case launches use), the steps it still has to run and
// the step index whose launch starts from natural slots. A tile is one
// workgroup's worth of nodes of one case: (slot << 24) | block. One launch per
// case type (lattice x storage x operator) walks that type's tile list with a
// fixed, tunable number of workgroups of a tunable tile size (TILE nodes, one
// per thread), so the launch shape does not depend on which cases are
// resident and a captured graph of K steps stays valid while cases come and
// go: admission and retirement only rewrite the tables.
//
// Each tile runs exactly the per-node code of the single-case AA kernel
// (aa_node), with the case's own parity and fresh flag read from its slot, so
// a case's populations are bitwise the same whether it runs alone, in a mix,
// or through lbf_main_steps. The last step of a case also stores moments, as
// the single-case path does at the end of a batch.
#pragma once

#include "dispatch.hpp"
#include "fast.hpp"

namespace lbf {

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
};

template <class L, class St, int OP, int TILE>
__global__ void __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(5)))
    serve_step(const ServeSlot* __restrict__ slots, const uint32_t* __restrict__ tiles,
               const uint32_t* __restrict__ ntiles) {
  const uint32_t nt = *ntiles;
  double acc[5] = {0, 0, 0, 0, 0};
  // Workgroups are dealt round-robin to the 8 XCDs. When the grid is a
  // multiple of 8 G, permute the ids within each group of 8 G so that every
  // XCD walks runs of G consecutive tiles (neighbouring rows share its L2);
  // the walk still visits each tile exactly once.
  constexpr uint32_t G = 4;
  uint32_t first = blockIdx.x;
  if (gridDim.x % (8 * G) == 0) {
    const uint32_t w = blockIdx.x >> 3;
    first = ((w / G) * 8u + (blockIdx.x & 7u)) * G + w % G;
  }
  for (uint32_t t = first; t < nt; t += gridDim.x) {
    const uint32_t tile = tiles[t];
    const ServeSlot& s = slots[tile >> (32 - kServeSlotBits)];
    const long long rem = s.remaining;
    if (rem <= 0) continue;  // finished within this graph replay
    const unsigned long long step = s.a.ctrl->step;
    const bool fresh = step == s.fresh_step;
    const uint32_t n = (tile & kServeBlockMask) * TILE + threadIdx.x;
    if (rem == 1) {
      if (step & 1)
        aa_node<L, St, OP, true, true>(s.a, n, step, fresh, acc);
      else
        aa_node<L, St, OP, false, true>(s.a, n, step, fresh, acc);
    } else {
      if (step & 1)
        aa_node<L, St, OP, true, false>(s.a, n, step, fresh, acc);
      else
        aa_node<L, St, OP, false, false>(s.a, n, step, fresh, acc);
    }
  }
}

template <class L, class St, int OP, int TILE>
hipError_t serve_launch(const ServeSlot* slots, const uint32_t* tiles, const uint32_t* ntiles,
                        int grid, hipStream_t s) {
  hipLaunchKernelGGL((serve_step<L, St, OP, TILE>), dim3(grid), dim3(TILE), 0, s, slots, tiles,
                     ntiles);
  return hipGetLastError();
}

}  // namespace lbf

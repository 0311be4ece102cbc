// dispatch.hpp — per-lattice launcher tables (one translation unit per lattice
// so the kernels of the four velocity sets compile in parallel).
#pragma once

#include "engine.hpp"

namespace lbf {

using LaunchFn = hipError_t (*)(const KArgs&, int grid, hipStream_t);

struct ServeSlot;  // serve.hpp
// batched step of one parity class of a tile list (serving tier): slots,
// tiles, counts (class 0, class 1), class, odd, grid
using ServeLaunchFn = hipError_t (*)(const ServeSlot*, const uint32_t*, const uint32_t*, int, int,
                                     int, hipStream_t);
// last steps over the same list: slots, tiles, counts, last-step flags, type bit, grid
using ServeLastFn = hipError_t (*)(const ServeSlot*, const uint32_t*, const uint32_t*,
                                   const uint32_t*, uint32_t, int, hipStream_t);

struct Launchers {
  LaunchFn mom_step = nullptr;  // fused pull-stream / faces / collide
  LaunchFn mom_init = nullptr;  // f = f_eq
  LaunchFn sca_step = nullptr;
  LaunchFn sca_init = nullptr;
  LaunchFn aa_step = nullptr;   // in-place AA-pattern step (momentum only)
  LaunchFn aa_init = nullptr;
  LaunchFn aa_readback = nullptr;
  // two steps per pass over the two-array scheme (3-D lattices, FP64 / FP32),
  // tiles of tb_tx x tb_ty columns in tile shape KArgs::tb_shape (1: rows of
  // 64 nodes where they fit)
  LaunchFn tb_step = nullptr;
  int tb_tx[2] = {0, 0}, tb_ty[2] = {0, 0};
  // [BGK, TRT][tile size 128, 256, 512] (D2Q9 / D3Q19, AA storages)
  ServeLaunchFn serve[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  ServeLastFn serve_last[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
};

// precision: LBF_F64 .. LBF_BF16_SHIFTED; returns false when not built
bool launchers_d2q9(int precision, Launchers* out);
bool launchers_d3q15(int precision, Launchers* out);
bool launchers_d3q19(int precision, Launchers* out);
bool launchers_d3q27(int precision, Launchers* out);

// lattice-independent kernels (misc.hip)
hipError_t launch_ctrl_set(Ctrl* ctrl, int what, long long value, hipStream_t s);
// one-workgroup launch that ends a step (residual reduction, stop rule,
// step counter); nblocks = workgroups of the step's field launches
hipError_t launch_advance(const KArgs& a, long long nblocks, hipStream_t s);
hipError_t launch_reduce(const void* values, int is_double, const uint32_t* cls, int need_link,
                         long long N, double* partials, int grid, double* out, int count_nonfinite,
                         hipStream_t s);

struct ClassArgs {
  int nx, ny, nz, q;
  long long N, ld;
  int c[27][3];
};
hipError_t launch_build_classes(const ClassArgs& a, const uint8_t* type, uint32_t* cls,
                                hipStream_t s);
hipError_t launch_fill(void* p, int is_double, long long n_valid, long long n_total, double value,
                       hipStream_t s);
hipError_t launch_qads_init(const uint32_t* cls, double* q, long long ld, double q0,
                            hipStream_t s);

hipError_t launch_nonfinite(const void* rho, const void* u, int is_double, int d, long long ld,
                            const uint32_t* cls, long long N, double* partials, int grid,
                            double* out, hipStream_t s);

enum CtrlOp : int {
  kCtrlFresh = 0,     // value: field mask to mark fresh
  kCtrlResetStop = 1,
  kCtrlMaxIter = 2,   // value: max_iter
  kCtrlAddSteps = 3,  // value: steps run since the counter last advanced (unless stopped)
};

}  // namespace lbf

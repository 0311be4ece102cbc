// lattice_launch.hpp — instantiates the step kernels of one velocity set.
#pragma once

#include "dispatch.hpp"
#include "fast.hpp"
#include "kernels.hpp"
#include "serve.hpp"
#include "tb.hpp"

namespace lbf {

template <class K>
static hipError_t launch(K kernel, const KArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <class L, class St>
static hipError_t mom_step_plain(const KArgs& a, int grid, hipStream_t s) {
  if (a.op == 0) return launch(momentum_step_plain<L, St, 0>, a, grid, s);
  if (a.op == 1) return launch(momentum_step_plain<L, St, 1>, a, grid, s);
  if constexpr (L::has_mrt) return launch(momentum_step_plain<L, St, 2>, a, grid, s);
  return hipErrorInvalidValue;
}
template <class L, class St>
static hipError_t mom_init_t(const KArgs& a, int grid, hipStream_t s) {
  return launch(momentum_init<L, St>, a, grid, s);
}
template <class L, class T>
static hipError_t sca_step_t(const KArgs& a, int grid, hipStream_t s) {
  if (a.op == 0) return launch(scalar_step<L, T, 0>, a, grid, s);
  if (a.op == 1) return launch(scalar_step<L, T, 1>, a, grid, s);
  if constexpr (L::has_mrt) return launch(scalar_step<L, T, 2>, a, grid, s);
  return hipErrorInvalidValue;
}
template <class L, class T>
static hipError_t sca_init_t(const KArgs& a, int grid, hipStream_t s) {
  return launch(scalar_init<L, T>, a, grid, s);
}
template <class L, class St, bool ODD>
static hipError_t aa_step_p(const KArgs& a, int grid, hipStream_t s) {
  const bool mom = (a.flags & (kFlagStoreMoments | kFlagResiduals)) != 0;
  if (mom) {
    if (a.op == 0) return launch(momentum_step_aa<L, St, 0, ODD, true>, a, grid, s);
    if (a.op == 1) return launch(momentum_step_aa<L, St, 1, ODD, true>, a, grid, s);
    if constexpr (L::has_mrt) return launch(momentum_step_aa<L, St, 2, ODD, true>, a, grid, s);
  } else {
    if (a.op == 0) return launch(momentum_step_aa<L, St, 0, ODD, false>, a, grid, s);
    if (a.op == 1) return launch(momentum_step_aa<L, St, 1, ODD, false>, a, grid, s);
    if constexpr (L::has_mrt) return launch(momentum_step_aa<L, St, 2, ODD, false>, a, grid, s);
  }
  return hipErrorInvalidValue;
}
template <class L, class St>
static hipError_t aa_step_t(const KArgs& a, int grid, hipStream_t s) {
  return a.aa_odd ? aa_step_p<L, St, true>(a, grid, s) : aa_step_p<L, St, false>(a, grid, s);
}
template <class L, class St>
static hipError_t aa_readback_t(const KArgs& a, int grid, hipStream_t s) {
  return a.aa_odd ? launch(momentum_readback_aa<L, St, true>, a, grid, s)
                  : launch(momentum_readback_aa<L, St, false>, a, grid, s);
}
template <class L, class St>
static hipError_t aa_init_t(const KArgs& a, int grid, hipStream_t s) {
  return launch(momentum_init_aa<L, St>, a, grid, s);
}

template <class L, class St, int SHAPE>
static hipError_t tb_step_s(const KArgs& a, int grid, hipStream_t s) {
  constexpr int nt = TbShape<L, St, SHAPE>::NT;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(nt), 0, s, a);
    return hipGetLastError();
  };
  if (a.flags & kFlagStoreMoments) {
    if (a.op == 0) return go(momentum_step_tb2<L, St, 0, true, SHAPE>);
    if (a.op == 1) return go(momentum_step_tb2<L, St, 1, true, SHAPE>);
    if constexpr (L::has_mrt) return go(momentum_step_tb2<L, St, 2, true, SHAPE>);
  } else {
    if (a.op == 0) return go(momentum_step_tb2<L, St, 0, false, SHAPE>);
    if (a.op == 1) return go(momentum_step_tb2<L, St, 1, false, SHAPE>);
    if constexpr (L::has_mrt) return go(momentum_step_tb2<L, St, 2, false, SHAPE>);
  }
  return hipErrorInvalidValue;
}
template <class L, class St>
static hipError_t tb_step_t(const KArgs& a, int grid, hipStream_t s) {
  return a.tb_shape ? tb_step_s<L, St, 1>(a, grid, s) : tb_step_s<L, St, 0>(a, grid, s);
}
template <class L, class St>
static void tb_launchers(Launchers* out) {
  if constexpr (L::D == 3) {
    out->tb_step = tb_step_t<L, St>;
    out->tb_tx[0] = TbShape<L, St, 0>::TX, out->tb_ty[0] = TbShape<L, St, 0>::TY;
    out->tb_tx[1] = TbShape<L, St, 1>::TX, out->tb_ty[1] = TbShape<L, St, 1>::TY;
  }
}

// serving-tier batched steps exist for the lattices of the serving mix
template <class L, class St>
static void serve_launchers(Launchers* out) {
  if constexpr (L::id == 1 || L::id == 3) {
    static_assert(kNumTileSizes == 3, "one launcher per tile size");
    out->serve[0][0] = serve_launch<L, St, 0, kTileSizes[0]>;
    out->serve[0][1] = serve_launch<L, St, 0, kTileSizes[1]>;
    out->serve[0][2] = serve_launch<L, St, 0, kTileSizes[2]>;
    out->serve[1][0] = serve_launch<L, St, 1, kTileSizes[0]>;
    out->serve[1][1] = serve_launch<L, St, 1, kTileSizes[1]>;
    out->serve[1][2] = serve_launch<L, St, 1, kTileSizes[2]>;
    out->serve_last[0][0] = serve_launch_last<L, St, 0, kTileSizes[0]>;
    out->serve_last[0][1] = serve_launch_last<L, St, 0, kTileSizes[1]>;
    out->serve_last[0][2] = serve_launch_last<L, St, 0, kTileSizes[2]>;
    out->serve_last[1][0] = serve_launch_last<L, St, 1, kTileSizes[0]>;
    out->serve_last[1][1] = serve_launch_last<L, St, 1, kTileSizes[1]>;
    out->serve_last[1][2] = serve_launch_last<L, St, 1, kTileSizes[2]>;
  }
}

template <class L>
static bool make_launchers(int precision, Launchers* out) {
  switch (precision) {
    case 0:  // LBF_F64: parity path, two arrays or in place (same bits)
      out->mom_step = mom_step_plain<L, StF64>;
      out->mom_init = mom_init_t<L, StF64>;
      out->sca_step = sca_step_t<L, double>;
      out->sca_init = sca_init_t<L, double>;
      out->aa_step = aa_step_t<L, StF64>;
      out->aa_init = aa_init_t<L, StF64>;
      out->aa_readback = aa_readback_t<L, StF64>;
      tb_launchers<L, StF64>(out);
      return true;
    case 1:  // LBF_F32
      out->mom_step = mom_step_plain<L, StF32>;
      out->mom_init = mom_init_t<L, StF32>;
      out->sca_step = sca_step_t<L, float>;
      out->sca_init = sca_init_t<L, float>;
      out->aa_step = aa_step_t<L, StF32>;
      out->aa_init = aa_init_t<L, StF32>;
      out->aa_readback = aa_readback_t<L, StF32>;
      tb_launchers<L, StF32>(out);
      serve_launchers<L, StF32>(out);
      return true;
    case 2:  // LBF_F32_SHIFTED
      out->aa_step = aa_step_t<L, StF32S>;
      out->aa_init = aa_init_t<L, StF32S>;
      out->aa_readback = aa_readback_t<L, StF32S>;
      // two arrays with shifted storage: the two-step scheme and its single steps
      out->mom_step = mom_step_plain<L, StF32S>;
      out->mom_init = mom_init_t<L, StF32S>;
      tb_launchers<L, StF32S>(out);
      serve_launchers<L, StF32S>(out);
      return true;
    case 3:  // LBF_F16_SHIFTED
      out->aa_step = aa_step_t<L, StF16S>;
      out->aa_init = aa_init_t<L, StF16S>;
      out->aa_readback = aa_readback_t<L, StF16S>;
      // two arrays with shifted storage: the two-step scheme and its single steps
      out->mom_step = mom_step_plain<L, StF16S>;
      out->mom_init = mom_init_t<L, StF16S>;
      tb_launchers<L, StF16S>(out);
      serve_launchers<L, StF16S>(out);
      return true;
    case 4:  // LBF_BF16_SHIFTED
      out->aa_step = aa_step_t<L, StBF16S>;
      out->aa_init = aa_init_t<L, StBF16S>;
      out->aa_readback = aa_readback_t<L, StBF16S>;
      // two arrays with shifted storage: the two-step scheme and its single steps
      out->mom_step = mom_step_plain<L, StBF16S>;
      out->mom_init = mom_init_t<L, StBF16S>;
      tb_launchers<L, StBF16S>(out);
      serve_launchers<L, StBF16S>(out);
      return true;
    default:
      return false;
  }
}

}  // namespace lbf

#define LBF_LATTICE_TU(Lat, name)                              \
  namespace lbf {                                              \
  bool launchers_##name(int precision, Launchers* out) {       \
    return make_launchers<Lat>(precision, out);                \
  }                                                            \
  }

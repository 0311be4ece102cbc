// misc.hip — lattice-independent kernels: control-block updates, the
// advance-only step (frozen momentum without scalars) and FP64 reductions.
#include <cmath>

#include "dispatch.hpp"
#include "kernels.hpp"

namespace lbf {

__global__ void ctrl_set_kernel(Ctrl* ctrl, int what, long long value) {
  if (threadIdx.x != 0) return;
  switch (what) {
    case kCtrlFresh:
      for (int f = 0; f < 5; f++)
        if (value >> f & 1) ctrl->fresh[f] = 1;
      break;
    case kCtrlResetStop:
      ctrl->stopped = 0;
      ctrl->done = 0;
      break;
    case kCtrlMaxIter:
      ctrl->max_iter = (unsigned long long)value;
      break;
    case kCtrlAddSteps:
      if (!ctrl->stopped) {
        ctrl->step += (unsigned long long)value;
        for (int f = 0; f < 5; f++) ctrl->fresh[f] = 0;
      }
      break;
  }
}

hipError_t launch_ctrl_set(Ctrl* ctrl, int what, long long value, hipStream_t s) {
  hipLaunchKernelGGL(ctrl_set_kernel, dim3(1), dim3(64), 0, s, ctrl, what, value);
  return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) advance_kernel(KArgs a, long long nblocks) {
  if (a.ctrl->stopped) return;
  const unsigned long long step = a.ctrl->step;
  advance_core(a, step, nblocks);
}

// First level of a step's residual reduction: workgroup g sums, column by
// column and in a fixed order, the partial rows of the step's workgroups
// [g * kResidualFold, (g + 1) * kResidualFold) into row nblocks + g of the
// same array. The advancing workgroup then sums those few rows, instead of
// reading every step workgroup's row alone (a latency-bound walk over
// 65,536 rows at 256^3).
__global__ void __launch_bounds__(kBlock) fold_partials(double* p, long long nblocks,
                                                        uint32_t mask) {
  // every column in one pass over the rows (each thread sums its rows in
  // order, then the workgroup sums in block_sum's order: the same bits as
  // one block_sum per column)
  __shared__ double lds[kMaxCols][kBlock / 64];
  const long long b0 = (long long)blockIdx.x * kResidualFold;
  double s[kMaxCols];
  for (int k = 0; k < kMaxCols; k++) s[k] = 0.0;
  for (int i = threadIdx.x; i < kResidualFold; i += kBlock) {
    const long long b = b0 + i;
    if (b < nblocks)
      for (int k = 0; k < kMaxCols; k++)
        if (mask >> k & 1u) s[k] += p[b * kMaxCols + k];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = 0; k < kMaxCols; k++) {
    const double v = wave_sum(s[k]);
    if (lane == 0) lds[k][wid] = v;
  }
  __syncthreads();
  const int k = threadIdx.x;
  if (k < kMaxCols && (mask >> k & 1u)) {
    double r = lds[k][0];
    for (int w = 1; w < kBlock / 64; w++) r += lds[k][w];
    p[(nblocks + blockIdx.x) * kMaxCols + k] = r;
  }
}

hipError_t launch_advance(const KArgs& a, long long nblocks, hipStream_t s) {
  if ((a.flags & kFlagResiduals) && nblocks > kResidualFold) {
    const long long folds = (nblocks + kResidualFold - 1) / kResidualFold;
    hipLaunchKernelGGL(fold_partials, dim3((unsigned)folds), dim3(kBlock), 0, s, a.partials,
                       nblocks, a.res_mask);
    KArgs b = a;
    b.partials = a.partials + nblocks * kMaxCols;
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(kBlock), 0, s, b, folds);
  } else {
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(kBlock), 0, s, a, nblocks);
  }
  return hipGetLastError();
}

// Sum of values[n] over fluid nodes (need_link: only nodes with a solid
// neighbour). count_nonfinite: counts non-finite values instead. Two passes
// with fixed association order, so the result does not depend on timing.
template <class T>
__global__ void __launch_bounds__(kBlock) reduce_pass1(const T* values, const uint32_t* cls,
                                                       int need_link, long long N, int count_nf,
                                                       double* partials) {
  __shared__ double lds[kBlock / 64];
  double acc = 0.0;
  for (long long n = (long long)blockIdx.x * kBlock + threadIdx.x; n < N;
       n += (long long)gridDim.x * kBlock) {
    const uint32_t c = cls[n];
    if (c & kSolidBit) continue;
    if (need_link && !(c & kLinkMask)) continue;
    const double v = (double)values[n];
    acc += count_nf ? (std::isfinite(v) ? 0.0 : 1.0) : v;
  }
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) reduce_pass2(const double* partials, int n, double* out) {
  __shared__ double lds[kBlock / 64];
  double acc = 0.0;
  for (int b = threadIdx.x; b < n; b += kBlock) acc += partials[b];
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) *out = s;
}

// number of fluid nodes with a non-finite density or velocity component
template <class T>
__global__ void __launch_bounds__(kBlock) nonfinite_pass1(const T* rho, const T* u, int d,
                                                          long long ld, const uint32_t* cls,
                                                          long long N, double* partials) {
  __shared__ double lds[kBlock / 64];
  double acc = 0.0;
  for (long long n = (long long)blockIdx.x * kBlock + threadIdx.x; n < N;
       n += (long long)gridDim.x * kBlock) {
    if (cls[n] & kSolidBit) continue;
    bool bad = !std::isfinite((double)rho[n]);
    for (int x = 0; x < d; x++) bad = bad || !std::isfinite((double)u[x * ld + n]);
    acc += bad ? 1.0 : 0.0;
  }
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

hipError_t launch_nonfinite(const void* rho, const void* u, int is_double, int d, long long ld,
                            const uint32_t* cls, long long N, double* partials, int grid,
                            double* out, hipStream_t s) {
  if (is_double)
    hipLaunchKernelGGL(nonfinite_pass1<double>, dim3(grid), dim3(kBlock), 0, s,
                       (const double*)rho, (const double*)u, d, ld, cls, N, partials);
  else
    hipLaunchKernelGGL(nonfinite_pass1<float>, dim3(grid), dim3(kBlock), 0, s, (const float*)rho,
                       (const float*)u, d, ld, cls, N, partials);
  hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(kBlock), 0, s, partials, grid, out);
  return hipGetLastError();
}

hipError_t launch_reduce(const void* values, int is_double, const uint32_t* cls, int need_link,
                         long long N, double* partials, int grid, double* out, int count_nonfinite,
                         hipStream_t s) {
  if (is_double)
    hipLaunchKernelGGL(reduce_pass1<double>, dim3(grid), dim3(kBlock), 0, s,
                       (const double*)values, cls, need_link, N, count_nonfinite, partials);
  else
    hipLaunchKernelGGL(reduce_pass1<float>, dim3(grid), dim3(kBlock), 0, s, (const float*)values,
                       cls, need_link, N, count_nonfinite, partials);
  hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(kBlock), 0, s, partials, grid, out);
  return hipGetLastError();
}

// Plain device copy (16-byte vectors), used to calibrate memory counters:
// a copy of B bytes moves exactly B bytes in and B bytes out.
// One vector per thread over an exact grid: the fastest plain copy found on
// MI355X (about 6.1 TB/s read + write), i.e. the attainable streaming rate.
__global__ void __launch_bounds__(kBlock) copy_kernel(const uint4* __restrict__ src,
                                                      uint4* __restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ---- case set-up kernels ---------------------------------------------------

// Node classes (src/momentum/lattice.rs:127-191): for a fluid node, bit k is
// set when its neighbour in direction k, wrapped periodically on every axis
// as the reference wraps it (lattice.rs:139), is solid; solid nodes and the
// padding up to ld get kSolidBit. type == nullptr: every node is fluid.
__global__ void __launch_bounds__(kBlock) build_classes_kernel(ClassArgs a, const uint8_t* type,
                                                               uint32_t* cls) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ld) return;
  if (i >= a.N || (type && type[i])) {
    cls[i] = kSolidBit;
    return;
  }
  uint32_t m = 0;
  if (type) {
    const long long nxy = (long long)a.nx * a.ny;
    const int z = (int)(i / nxy);
    const int y = (int)((i - (long long)z * nxy) / a.nx);
    const int x = (int)(i - (long long)z * nxy - (long long)y * a.nx);
    for (int k = 0; k < a.q; k++) {
      int xx = x + a.c[k][0], yy = y + a.c[k][1], zz = z + a.c[k][2];
      xx = xx < 0 ? xx + a.nx : (xx >= a.nx ? xx - a.nx : xx);
      yy = yy < 0 ? yy + a.ny : (yy >= a.ny ? yy - a.ny : yy);
      zz = zz < 0 ? zz + a.nz : (zz >= a.nz ? zz - a.nz : zz);
      if (type[xx + a.nx * ((long long)yy + (long long)a.ny * zz)]) m |= 1u << k;
    }
  }
  cls[i] = m;
}

hipError_t launch_build_classes(const ClassArgs& a, const uint8_t* type, uint32_t* cls,
                                hipStream_t s) {
  const long long blocks = (a.ld + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(build_classes_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, a, type,
                     cls);
  return hipGetLastError();
}

// p[i] = value for i < n_valid, 0 for n_valid <= i < n_total
template <class T>
__global__ void __launch_bounds__(kBlock) fill_kernel(T* p, long long n_valid, long long n_total,
                                                      double value) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n_total) p[i] = i < n_valid ? (T)value : (T)0;
}

hipError_t launch_fill(void* p, int is_double, long long n_valid, long long n_total, double value,
                       hipStream_t s) {
  const long long blocks = (n_total + kBlock - 1) / kBlock;
  if (blocks == 0) return hipSuccess;
  if (is_double)
    hipLaunchKernelGGL(fill_kernel<double>, dim3((unsigned)blocks), dim3(kBlock), 0, s,
                       (double*)p, n_valid, n_total, value);
  else
    hipLaunchKernelGGL(fill_kernel<float>, dim3((unsigned)blocks), dim3(kBlock), 0, s, (float*)p,
                       n_valid, n_total, value);
  return hipGetLastError();
}

// q_ads initial value on momentum bounce-back nodes (fluid nodes with a solid
// neighbour), 0 elsewhere (src/passive_scalar/node.rs:77-85)
__global__ void __launch_bounds__(kBlock) qads_init_kernel(const uint32_t* cls, double* q,
                                                           long long ld, double q0) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= ld) return;
  const uint32_t c = cls[i];
  q[i] = (!(c & kSolidBit) && (c & kLinkMask)) ? q0 : 0.0;
}

hipError_t launch_qads_init(const uint32_t* cls, double* q, long long ld, double q0,
                            hipStream_t s) {
  const long long blocks = (ld + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(qads_init_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, cls, q, ld, q0);
  return hipGetLastError();
}

// The in-place kernel's access pattern without its arithmetic or bounce-back:
// every lane reads its element of each of Q planes and writes it to the
// paired plane (plane j -> plane Q-j, plane 0 onto itself), one element per
// lane, like the even step, with the same buffer addressing. Its rate bounds
// what the in-place kernel can reach.
template <class S, int Q>
__global__ void __launch_bounds__(kBlock) sweep_kernel(void* base, uint32_t plane_bytes, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)0xffffffff, 0x00020000);
  const uint32_t own = i * (uint32_t)sizeof(S);
  S v[Q];
  sfor<Q>([&](auto J) {
    constexpr int j = decltype(J)::value;
    if constexpr (sizeof(S) == 8)
      v[j] = __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b64(r, own, j * plane_bytes, 0));
    else if constexpr (sizeof(S) == 4)
      v[j] = __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b32(r, own, j * plane_bytes, 0));
    else
      v[j] = __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b16(r, own, j * plane_bytes, 0));
  });
  sfor<Q>([&](auto J) {
    constexpr int j = decltype(J)::value;
    constexpr int jb = j == 0 ? 0 : Q - j;
    if constexpr (sizeof(S) == 8) {
      typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v[j]), r, own, jb * plane_bytes, 0);
    } else if constexpr (sizeof(S) == 4) {
      __builtin_amdgcn_raw_buffer_store_b32(v[j], r, own, jb * plane_bytes, 0);
    } else {
      __builtin_amdgcn_raw_buffer_store_b16(v[j], r, own, jb * plane_bytes, 0);
    }
  });
}

template <class S>
hipError_t launch_sweep(void* base, int q, uint32_t plane_bytes, uint32_t n, hipStream_t s) {
  const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
  switch (q) {
    case 9: hipLaunchKernelGGL((sweep_kernel<S, 9>), g, b, 0, s, base, plane_bytes, n); break;
    case 15: hipLaunchKernelGGL((sweep_kernel<S, 15>), g, b, 0, s, base, plane_bytes, n); break;
    case 19: hipLaunchKernelGGL((sweep_kernel<S, 19>), g, b, 0, s, base, plane_bytes, n); break;
    case 27: hipLaunchKernelGGL((sweep_kernel<S, 27>), g, b, 0, s, base, plane_bytes, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace lbf

// Benchmark hook (not part of the reference-equivalent ABI): in-place sweep
// of q (9, 15, 19 or 27) planes of `ld` elements of es (2, 4 or 8) bytes, the
// first n elements of each, on `stream` (the array must stay below 4 GiB).
extern "C" int lbfx_sweep(void* base, int q, long long ld, int es, long long n, void* stream) {
  if (!base || n < 0 || n > ld || (es != 2 && es != 4 && es != 8)) return -1;
  if ((unsigned long long)q * ld * es >= (1ull << 32)) return -1;
  if (n == 0) return 0;
  const uint32_t pb = (uint32_t)(ld * es);
  hipError_t e;
  if (es == 2)
    e = lbf::launch_sweep<uint16_t>(base, q, pb, (uint32_t)n, (hipStream_t)stream);
  else if (es == 4)
    e = lbf::launch_sweep<uint32_t>(base, q, pb, (uint32_t)n, (hipStream_t)stream);
  else
    e = lbf::launch_sweep<uint64_t>(base, q, pb, (uint32_t)n, (hipStream_t)stream);
  return e == hipSuccess ? 0 : (e == hipErrorInvalidValue ? -1 : -3);
}

// Benchmark hook (not part of the reference-equivalent ABI): copies `bytes`
// (a multiple of 16) from src to dst on `stream`.
extern "C" int lbfx_copy(void* dst, const void* src, size_t bytes, void* stream) {
  const size_t n = bytes / 16;
  const size_t blocks = (n + lbf::kBlock - 1) / lbf::kBlock;
  if (blocks == 0) return 0;
  if (blocks > 0x7fffffffu) return -2;
  hipLaunchKernelGGL(lbf::copy_kernel, dim3((unsigned)blocks), dim3(lbf::kBlock), 0,
                     (hipStream_t)stream, (const uint4*)src, (uint4*)dst, n);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

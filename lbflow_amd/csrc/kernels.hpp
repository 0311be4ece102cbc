// kernels.hpp — fused lattice step kernels (two-array pull scheme).
//
// One launch = one time step of one field for every fluid node:
//   gather   f_j(x) = f*_j(x - c_j) with periodic wrap (src/lib.rs:141-152,
//            neighbour table src/momentum/lattice.rs:127-154); for a link
//            whose neighbour is solid, f_j(x) = f*_jbar(x) (inner bounce-back,
//            src/lib.rs:154-165);
//   faces    the face conditions of the faces the node lies on, in
//            FACES_2D/FACES_3D order (src/momentum/bc.rs:29-149);
//   moments  rho, u (+F/2rho) (src/momentum/node.rs:418-457);
//   collide  BGK(+Guo)/TRT/MRT (src/kernel.rs, src/momentum/node.rs:311-333);
//   store    f* into the other buffer.
// This is the reference's step t-1 second half (stream, bounce-back, faces)
// followed by step t first half (moments, equilibrium, collision): the state
// kept between launches is f*, and `fresh` marks a buffer that instead holds
// post-stream f (after initialize_nodes / set_pops), whose first launch only
// collides. Populations live in structure-of-arrays planes [q][ld].
//
// Each step ends with a one-workgroup launch (advance_core) that reduces the
// per-workgroup residual partials in a fixed order, applies the stop rule
// (src/momentum/lattice.rs:407-418) and advances the device step counter, so
// the host never synchronises inside a run and a captured sequence of
// launches can be replayed.
#pragma once

#include <type_traits>

#include "engine.hpp"
#include "node_math.hpp"

namespace lbf {

// ------------------------------------------------------------- storage traits
struct StF64 {
  using S = double;
  using T = double;
  static constexpr bool shifted = false;
};
struct StF32 {
  using S = float;
  using T = float;
  static constexpr bool shifted = false;
};
struct StF32S {
  using S = float;
  using T = float;
  static constexpr bool shifted = true;
};
struct StF16S {
  using S = _Float16;
  using T = float;
  static constexpr bool shifted = true;
};
struct StBF16S {
  using S = __bf16;
  using T = float;
  static constexpr bool shifted = true;
};

template <class St>
__device__ __forceinline__ typename St::T ldp(const typename St::S* p) {
  return static_cast<typename St::T>(*p);
}
template <class St>
__device__ __forceinline__ void stp(typename St::S* p, typename St::T v) {
  *p = static_cast<typename St::S>(v);
}

// ------------------------------------------------------------- geometry
struct Geo {
  int x, y, z;
  long long offx[3], offy[3], offz[3];  // index offsets for shifts -1, 0, +1 (wrapped)
  __device__ __forceinline__ void init(long long n, int nx, int ny, int nz) {
    const long long nxy = (long long)nx * ny;
    z = (int)(n / nxy);
    const long long r = n - (long long)z * nxy;
    y = (int)(r / nx);
    x = (int)(r - (long long)y * nx);
    offx[0] = (x == 0) ? (nx - 1) : -1;
    offx[1] = 0;
    offx[2] = (x == nx - 1) ? -(long long)(nx - 1) : 1;
    offy[0] = (y == 0) ? (long long)(ny - 1) * nx : -(long long)nx;
    offy[1] = 0;
    offy[2] = (y == ny - 1) ? -(long long)(ny - 1) * nx : (long long)nx;
    offz[0] = (z == 0) ? (long long)(nz - 1) * nxy : -nxy;
    offz[1] = 0;
    offz[2] = (z == nz - 1) ? -(long long)(nz - 1) * nxy : nxy;
  }
  // offset of node x + (sx, sy, sz) with periodic wrap
  __device__ __forceinline__ long long off(int sx, int sy, int sz) const {
    return offx[sx + 1] + offy[sy + 1] + offz[sz + 1];
  }
  __device__ __forceinline__ uint32_t faces(int nx, int ny, int nz, int D) const {
    uint32_t m = 0;
    m |= (x == 0) ? 1u : 0u;
    m |= (x == nx - 1) ? 2u : 0u;
    m |= (y == 0) ? 4u : 0u;
    m |= (y == ny - 1) ? 8u : 0u;
    if (D == 3) {
      m |= (z == 0) ? 16u : 0u;
      m |= (z == nz - 1) ? 32u : 0u;
    }
    return m;
  }
};

// ------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Residual partials of one workgroup: columns 0 (rho), 1..3 (u) and 8
// (non-finite count) of its row, the five sums taken together (one wave
// reduction each, one barrier) in block_sum's order, so the bits are the
// same as five block_sum calls.
__device__ __forceinline__ void store_residual_partials(const double* acc, double* row) {
  __shared__ double lds[5][kBlock / 64];
  double v[5];
  for (int k = 0; k < 5; k++) v[k] = wave_sum(acc[k]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 5; k++) lds[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x < 5) {
    const int k = threadIdx.x;
    double r = lds[k][0];
    for (int w = 1; w < kBlock / 64; w++) r += lds[k][w];
    row[k < 4 ? k : 8] = r;
  }
}

// deterministic workgroup sum; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = lds[0];
    for (int w = 1; w < kBlock / 64; w++) r += lds[w];
  }
  return r;
}

// Ends a step; run by the single workgroup of the advance launch that
// follows the step's field launches (stream order makes their writes
// visible). Reduces the residual partials of the step's `nblocks`
// workgroups (columns in a.res_mask, summed in workgroup order), applies
// the stop rule (src/momentum/lattice.rs:407-418) and advances the device
// step counter. Keeping this out of the streaming kernels avoids a
// device-scope release fence per workgroup, which would write back L2
// lines that other waves have only partly filled.
__device__ __forceinline__ void advance_core(const KArgs& a, unsigned long long step,
                                             long long nblocks) {
  __shared__ double lds[kBlock / 64];
  Ctrl* ctrl = a.ctrl;
  double res[kMaxCols];
  for (int k = 0; k < kMaxCols; k++) res[k] = 0.0;
  const bool resid = (a.flags & kFlagResiduals) != 0;
  if (resid) {
    for (int k = 0; k < kMaxCols; k++) {
      if (!(a.res_mask >> k & 1u)) continue;
      double s = 0.0;
      for (long long b = threadIdx.x; b < nblocks; b += kBlock) s += a.partials[b * kMaxCols + k];
      res[k] = block_sum(s, lds);
    }
  }
  if (threadIdx.x == 0) {
    bool conv = true;
    for (int k = 0; k < 8; k++) {
      res[k] = sqrt(res[k]);
      if (a.res_mask >> k & 1u) conv = conv && (res[k] <= 1e-7);
    }
    if (resid) {
      for (int k = 0; k < kMaxCols; k++) {
        ctrl->res[k] = res[k];
        ctrl->hist[step % kHist][k] = res[k];
      }
    }
    bool stop = false;
    if (a.flags & kFlagStopCheck) stop = (step > 10ull && conv) || step > ctrl->max_iter;
    ctrl->step = step + 1;
    for (int f = 0; f < 5; f++) ctrl->fresh[f] = 0;
    if (stop) ctrl->stopped = 1;
  }
}

// ------------------------------------------------------------- momentum
template <class L, class St>
__device__ __forceinline__ void apply_momentum_faces(const KArgs& a, uint32_t on,
                                                     typename St::T* f,
                                                     const typename St::S* __restrict__ src,
                                                     long long n, const Geo& g) {
  using T = typename St::T;
  using k = K<L>;
  using Dv = Derived<L>;
  constexpr int D = L::D;
  sfor<2 * D>([&](auto FI) {
    constexpr int F = decltype(FI)::value;
    if (!(on >> F & 1u)) return;
    const FaceK& bc = a.face[F];
    if (bc.kind == 1 /*NoSlip*/ || bc.kind == 2 /*BounceBack*/ || bc.kind == 3 /*AntiBB*/) {
      T vel[3] = {T(0), T(0), T(0)};
      T udu = T(0);
      if (bc.kind == 3) {
        // 1.5 u(x) - 0.5 u(x + n_in), src/momentum/bc.rs:93-101
        constexpr int nd = Dv::normal(F);
        const long long nb = n + g.off(k::cx(nd), k::cy(nd), k::cz(nd));
        const int pv = (int)((a.ctrl->step + a.step_offset + 1) & 1);  // previous step's moments
        const T* up = (const T*)a.u[pv];
        sfor<D>([&](auto A) {
          constexpr int ax = decltype(A)::value;
          vel[ax] = T(1.5) * up[ax * a.ld + n] - T(0.5) * up[ax * a.ld + nb];
        });
        udu = udotu<L>(vel);
      } else if (bc.kind == 2) {
        sfor<D>([&](auto A) {
          constexpr int ax = decltype(A)::value;
          vel[ax] = T(bc.vel[ax]);
        });
      }
      sfor<L::NQF>([&](auto KI) {
        constexpr int i = Dv::qface(F, decltype(KI)::value);
        constexpr int ib = k::qbar(i);
        const T fsi = ldp<St>(src + i * a.ld + n);
        if (bc.kind == 1) {
          f[ib] = fsi;  // bc.rs:119-129
        } else if (bc.kind == 2) {
          const T udc = udotc<L, i>(vel);  // bc.rs:62-84
          // shifted storage (f - w) * scale: the correction scales, w_ib = w_i
          const T dv = T(2.0) * T(k::w(i)) * T(bc.density) * T(3.0) * udc;
          f[ib] = fsi - (St::shifted ? T(a.scale) * dv : dv);
        } else {
          const T udc = udotc<L, i>(vel);  // bc.rs:86-117
          f[ib] = -fsi + T(2.0) * T(k::w(i)) * T(bc.density) *
                             (T(1.0) + T(0.5) * udc * udc * T(9.0) - T(0.5) * udu * T(3.0));
        }
      });
    } else if (bc.kind == 4 /*ZouHe*/) {
      // combinations the reference panics on are rejected by lbf_create
      if constexpr (L::has_zou_he) {
        (void)zou_he_d2q9<T>(F, f, bc.known, T(bc.density), T(bc.vel[0]), T(bc.vel[1]));
      } else {
        const T vel[3] = {T(bc.vel[0]), T(bc.vel[1]), T(bc.vel[2])};
        zou_he_general<L, T, F>(f, bc.known, T(bc.density), vel);
      }
    }
  });
}

// Node block of this workgroup. Workgroups are dealt round-robin to the 8
// XCDs (each with its own L2), so XCD k runs workgroups k, k+8, ...; with
// xcd_group G > 0 its t-th workgroup takes block ((t / G) * 8 + k) * G + t % G:
// each XCD sweeps runs of G consecutive blocks (neighbouring rows share its
// L2) while the 8 XCDs together still sweep one contiguous window. A
// permutation of [0, gridDim) when gridDim % (8 G) == 0 (the host checks).
__device__ __forceinline__ uint32_t block_index(const KArgs& a) {
  uint32_t blk = blockIdx.x;
  if (a.xcd_group) {
    const uint32_t G = (uint32_t)a.xcd_group, t = blk >> 3;
    blk = ((t / G) * 8u + (blk & 7u)) * G + t % G;
  }
  return blk;
}

// Stream-only read-back of one node's post-stream populations as f64: SoA
// out[j][ld] over every node, or AoS out[(n - out_n0) * Q + j] (solid nodes
// written as zeros) when a.out_aos is set; then the launch only visits
// [out_n0, out_n1) (readback_range).
template <int Q, class T>
__device__ __forceinline__ void readback_store(const KArgs& a, long long n, const T (&f)[Q]) {
  double* __restrict__ outd = (double*)a.out;
  sfor<Q>([&](auto J) {
    constexpr int j = decltype(J)::value;
    if (a.out_aos)
      outd[(n - a.out_n0) * Q + j] = (double)f[j];
    else
      outd[j * a.ld + n] = (double)f[j];
  });
}
// a solid node of an AoS read-back range reads as zeros
template <int Q>
__device__ __forceinline__ void readback_solid(const KArgs& a, long long n) {
  if (!a.out_aos) return;
  double* __restrict__ outd = (double*)a.out;
  sfor<Q>([&](auto J) { outd[(n - a.out_n0) * Q + decltype(J)::value] = 0.0; });
}

// Node range a launch visits: every node, or [out_n0, out_n1) for an AoS
// read-back.
__device__ __forceinline__ void readback_range(const KArgs& a, long long* n0, long long* n1) {
  const bool r = a.mode == kModeStreamOnly && a.out_aos;
  *n0 = r ? a.out_n0 : 0;
  *n1 = r ? a.out_n1 : a.N;
}

// pair-form moments and collision of the FP32 kernels (fast.hpp)
template <class L, class St, int OP, class GetForce, class Moments>
__device__ __forceinline__ void pair_collide(const KArgs& a, const typename St::T* f,
                                             GetForce&& get_force, Moments&& moments,
                                             typename St::T* fs);

// Plain-form momentum step (storage holds f), two-array scheme: FP64 and
// FP32 cases the in-place scheme does not take (Shan-Chen, anti-bounce-back,
// Zou-He edges in FP64), or on request.
template <class L, class St, int OP>
__global__ void __launch_bounds__(kBlock) momentum_step_plain(KArgs a) {
  using T = typename St::T;
  using S = typename St::S;
  using k = K<L>;
  constexpr int Q = L::Q, D = L::D;
  Ctrl* ctrl = a.ctrl;
  if (a.mode != kModeStreamOnly && ctrl->stopped) return;
  // a launch of a batch whose counter has not advanced yet counts on from it;
  // only the batch's first step can read post-stream f
  const unsigned long long step = ctrl->step + a.step_offset;
  const int par = (int)(step & 1);
  const bool fresh = a.step_offset == 0 && ctrl->fresh[0] != 0;
  T* phi = (T*)a.phi;
  const S* __restrict__ src = (const S*)a.pop[par];
  S* __restrict__ dst = (S*)a.pop[par ^ 1];
  T* rho_c = (T*)a.rho[par];
  T* u_c = (T*)a.u[par];
  const T* rho_p = (const T*)a.rho[par ^ 1];
  const T* u_p = (const T*)a.u[par ^ 1];
  const bool store_m = (a.flags & (kFlagStoreMoments | kFlagResiduals)) != 0;
  const bool resid = (a.flags & kFlagResiduals) != 0;
  const T omega = T(a.omega), omega_p = T(a.omega_p), wp = T(a.wp), wm = T(a.wm),
          coeff_b = T(a.coeff_b);
  double acc[5] = {0, 0, 0, 0, 0};
  const long long ld = a.ld;
  long long nbeg, nend;
  readback_range(a, &nbeg, &nend);
  for (long long n = nbeg + (long long)block_index(a) * kBlock + threadIdx.x; n < nend;
       n += (long long)gridDim.x * kBlock) {
    const uint32_t cls = a.cls[n];
    if (cls & kSolidBit) {
      if (a.mode == kModeStreamOnly) readback_solid<Q>(a, n);
      continue;
    }
    T f[Q];
    if (fresh) {
      sfor<Q>([&](auto J) {
        constexpr int j = decltype(J)::value;
        f[j] = ldp<St>(src + j * ld + n);
      });
    } else {
      Geo g;
      g.init(n, a.nx, a.ny, a.nz);
      sfor<Q>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr int jb = k::qbar(j);
        if (cls >> jb & 1u)
          f[j] = ldp<St>(src + jb * ld + n);
        else
          f[j] = ldp<St>(src + j * ld + n + g.off(-k::cx(j), -k::cy(j), -k::cz(j)));
      });
      const uint32_t on = g.faces(a.nx, a.ny, a.nz, D) & a.face_active;
      if (on) apply_momentum_faces<L, St>(a, on, f, src, n, g);
    }
    if (a.mode == kModeStreamOnly) {
      if constexpr (St::shifted) {
        double fd[Q];
        sfor<Q>([&](auto J) {
          constexpr int j = decltype(J)::value;
          fd[j] = (double)f[j] * a.inv_scale + k::w(j);
        });
        readback_store<Q>(a, n, fd);
      } else {
        readback_store<Q>(a, n, f);
      }
      continue;
    }
    // FP32 storage collides in pair form (pair_collide, shared with the
    // in-place and two-step kernels, so the three schemes give the same bits);
    // FP64 in the reference's order
    constexpr bool pair = !std::is_same<St, StF64>::value;
    T rho, u[3];
    if (!pair || a.mode == kModePhi) moments<L>(f, rho, u);
    if (a.mode == kModePhi) {
      // compute_phi (src/momentum/multiphase.rs:48-54): psi = 1 - exp(-rho)
      phi[n] = T(1) * (T(1) - exp(-rho / T(1)));
      continue;
    }
    T F[3] = {T(0), T(0), T(0)};
    const bool has_force = a.force_kind || a.multiphase;
    if (has_force) {
      // body force (+ Shan-Chen force, node.rs:388-404; multiphase.rs:56-93)
      T B[3] = {T(0), T(0), T(0)};
      if (a.force_kind)
        sfor<D>([&](auto A) {
          constexpr int ax = decltype(A)::value;
          B[ax] = a.force_kind == 1 ? T(a.F[ax]) : T(a.src_field[ax * ld + n]);
        });
      if (a.multiphase) {
        Geo g;
        g.init(n, a.nx, a.ny, a.nz);
        T ph[Q];
        sfor<Q>([&](auto I) {
          constexpr int i = decltype(I)::value;
          ph[i] = (cls >> i & 1u) ? T(a.wall_phi)
                                  : phi[n + g.off(k::cx(i), k::cy(i), k::cz(i))];
        });
        const T node_phi = phi[n];
        sfor<D>([&](auto A) {
          constexpr int ax = decltype(A)::value;
          T sum = T(0);
          sfor<Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            sum = sum + T(k::w(i)) * T(k::ca(i, ax)) * ph[i];
          });
          const T sc = -T(a.sc_g) * node_phi * sum;
          F[ax] = a.force_kind ? B[ax] + sc : sc;
        });
      } else {
        sfor<D>([&](auto A) { F[decltype(A)::value] = B[decltype(A)::value]; });
      }
      if constexpr (!pair)
        sfor<D>([&](auto A) {
          constexpr int ax = decltype(A)::value;
          u[ax] += T(0.5) * F[ax] / rho;  // node.rs:450-455
        });
    }
    T fs[Q];
    if constexpr (pair)
      pair_collide<L, St, OP>(
          a, f,
          [&](T* Fv) {
            sfor<D>([&](auto A) { Fv[decltype(A)::value] = F[decltype(A)::value]; });
            return has_force;
          },
          [&](T r, const T* uu) {
            rho = r;
            sfor<D>([&](auto A) { u[decltype(A)::value] = uu[decltype(A)::value]; });
          },
          fs);
    if (store_m) {
      rho_c[n] = rho;
      sfor<D>([&](auto A) {
        constexpr int ax = decltype(A)::value;
        u_c[ax * ld + n] = u[ax];
      });
      if (resid) {
        const double dr = (double)rho - (double)rho_p[n];
        acc[0] += dr * dr;
        bool bad = !isfinite((double)rho);
        sfor<D>([&](auto A) {
          constexpr int ax = decltype(A)::value;
          const double du = (double)u[ax] - (double)u_p[ax * ld + n];
          acc[1 + ax] += du * du;
          bad = bad || !isfinite((double)u[ax]);
        });
        acc[4] += bad ? 1.0 : 0.0;
      }
    }
    if constexpr (!pair) {
      T feq[Q];
      equilibrium<L>(rho, u, feq);
      if constexpr (OP == 0) {
        bgk<L>(f, feq, omega, omega_p, fs);
        if (has_force) add_guo<L>(u, F, coeff_b, fs);
      } else if constexpr (OP == 1) {
        trt<L>(f, feq, wp, wm, fs);
      } else {
        T s[Q];
        sfor<Q>([&](auto KI) {
          constexpr int kk = decltype(KI)::value;
          s[kk] = T(a.s[kk]);
        });
        mrt<L>(rho, u, f, feq, s, fs);
      }
    }
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      stp<St>(dst + j * ld + n, fs[j]);
    });
  }
  if (a.mode != kModePull) return;
  // residual partials: columns 0 rho, 1..3 u, 8 non-finite count
  if (a.flags & kFlagResiduals) store_residual_partials(acc, a.partials + (long long)blockIdx.x * kMaxCols);
}

// The parity kernel's arithmetic after streaming, for the in-place kernel
// with FP64 storage (no Shan-Chen force there): the same calls in the same
// order, compiled without contraction, so both schemes give the same bits.
// Moments, then the body-force shift of u (src/momentum/node.rs:450-455).
template <class L>
__device__ __forceinline__ void ref_moments_force(const KArgs& a, uint32_t n, const double* f,
                                                  double& rho, double* u, double* F) {
  constexpr int D = L::D;
  moments<L>(f, rho, u);
  F[0] = F[1] = F[2] = 0.0;
  if (a.force_kind) {
    sfor<D>([&](auto A) {
      constexpr int ax = decltype(A)::value;
      F[ax] = a.force_kind == 1 ? a.F[ax] : a.src_field[(size_t)ax * (size_t)a.ld + n];
    });
    sfor<D>([&](auto A) {
      constexpr int ax = decltype(A)::value;
      u[ax] += 0.5 * F[ax] / rho;
    });
  }
}

// Equilibrium and collision (src/kernel.rs:3-120, src/momentum/node.rs:311-333).
template <class L, int OP>
__device__ __forceinline__ void ref_collide(const KArgs& a, const double* f, double rho,
                                            const double* u, const double* F, double* fs) {
  constexpr int Q = L::Q;
  double feq[Q];
  equilibrium<L>(rho, u, feq);
  if constexpr (OP == 0) {
    bgk<L>(f, feq, a.omega, a.omega_p, fs);
    if (a.force_kind) add_guo<L>(u, F, a.coeff_b, fs);
  } else if constexpr (OP == 1) {
    trt<L>(f, feq, a.wp, a.wm, fs);
  } else {
    double s[Q];
    sfor<Q>([&](auto KI) {
      constexpr int kk = decltype(KI)::value;
      s[kk] = a.s[kk];
    });
    mrt<L>(rho, u, f, feq, s, fs);
  }
}

// ------------------------------------------------------------- scalar
// Gathers the post-stream, post-inner-condition scalar populations of node m
// (src/passive_scalar/lattice.rs:274-297, node.rs:281-294).
template <class L, class T>
__device__ __forceinline__ void scalar_gather(const KArgs& a, const T* __restrict__ src, long long m,
                                              uint32_t cls, const Geo& g, const T* cbuf, T* gp) {
  using k = K<L>;
  const long long ld = a.ld;
  sfor<L::Q>([&](auto J) {
    constexpr int j = decltype(J)::value;
    constexpr int jb = k::qbar(j);
    if (cls >> jb & 1u) {
      if (a.inner_bc == 1) {
        gp[j] = src[jb * ld + m];
      } else {
        // -g*_i + 2 w_i c_wall, with c_wall the solid neighbour's (constant) value
        const long long sn = m + g.off(k::cx(jb), k::cy(jb), k::cz(jb));
        gp[j] = -src[jb * ld + m] + T(2.0) * T(k::w(jb)) * cbuf[sn];
      }
    } else {
      gp[j] = src[j * ld + m + g.off(-k::cx(j), -k::cy(j), -k::cz(j))];
    }
  });
}

template <class L, class T>
__device__ __forceinline__ bool scalar_source(const KArgs& a, long long n, bool bb, T cval,
                                              double& q, T& S) {
  // get_source_value + Langmuir adsorption (node.rs:258-266, adsorption.rs:50-92)
  const bool has_user = a.force_kind != 0;
  T user = T(0), ads = T(0);
  if (has_user) user = a.force_kind == 1 ? T(a.src_uniform) : T(a.src_field[n]);
  if (a.has_ads) {
    if (bb) {
      const T qq = T(q);
      const T r = T(a.k_a) * cval * (T(a.q_max) - qq) - T(a.k_d) * qq;
      ads = -r;
      T qn = qq + r;
      if (qn < T(0.0))
        qn = T(0.0);
      else if (qn > T(a.q_max))
        qn = T(a.q_max);
      q = (double)qn;
    } else {
      ads = T(0.0);
    }
  }
  if (has_user && a.has_ads)
    S = user + ads;
  else if (has_user)
    S = user;
  else if (a.has_ads)
    S = ads;
  else
    return false;
  return true;
}

template <class L, class T, int OP>
__global__ void __launch_bounds__(kBlock) scalar_step(KArgs a) {
  using k = K<L>;
  using Dv = Derived<L>;
  constexpr int Q = L::Q, D = L::D;
  Ctrl* ctrl = a.ctrl;
  if (a.mode == kModePull && ctrl->stopped) return;
  const unsigned long long step = ctrl->step + a.step_offset;
  const int par = (int)(step & 1);
  const bool fresh = a.step_offset == 0 && ctrl->fresh[a.field] != 0;
  const T* __restrict__ src = (const T*)a.pop[par];
  T* __restrict__ dst = (T*)a.pop[par ^ 1];
  T* c_c = (T*)a.c[par];
  const T* c_p = (const T*)a.c[par ^ 1];
  const T* u_c = (const T*)a.u[par];
  const T* u_p = (const T*)a.u[par ^ 1];
  const long long ld = a.ld;
  double acc = 0.0;
  const T omega = T(a.omega), omega_p = T(a.omega_p), wp = T(a.wp), wm = T(a.wm),
          coeff_b = T(a.coeff_b);
  long long nbeg, nend;
  readback_range(a, &nbeg, &nend);
  for (long long n = nbeg + (long long)block_index(a) * kBlock + threadIdx.x; n < nend;
       n += (long long)gridDim.x * kBlock) {
    const uint32_t cls = a.cls[n];
    if (cls & kSolidBit) {
      if (a.mode == kModeStreamOnly) readback_solid<Q>(a, n);
      continue;
    }
    T gv[Q];
    Geo g;
    g.init(n, a.nx, a.ny, a.nz);
    if (fresh) {
      sfor<Q>([&](auto J) {
        constexpr int j = decltype(J)::value;
        gv[j] = src[j * ld + n];
      });
    } else {
      scalar_gather<L, T>(a, src, n, cls, g, c_p, gv);
      const uint32_t on = g.faces(a.nx, a.ny, a.nz, D) & a.face_active;
      if (on) {
        // scalar face conditions, src/passive_scalar/bc.rs:38-219
        sfor<2 * D>([&](auto FI) {
          constexpr int F = decltype(FI)::value;
          if (!(on >> F & 1u)) return;
          const SFaceK& bc = a.sface[F];
          constexpr int nd = Dv::normal(F);
          if (bc.kind == 1 || bc.kind == 2) {  // anti-bounce-back / AntiBBNoFlux
            T vel[3] = {T(0), T(0), T(0)};
            const FaceK& mb = a.face[F];
            bool known = mb.kind == 1 || mb.kind == 2;
            if (mb.kind == 4) {
              const unsigned need = 2u | 4u | (D == 3 ? 8u : 0u);
              known = (mb.known & need) == need;
            }
            if (known) {
              if (mb.kind != 1)
                sfor<D>([&](auto A) {
                  constexpr int ax = decltype(A)::value;
                  vel[ax] = T(mb.vel[ax]);
                });
            } else {
              const long long nb = n + g.off(k::cx(nd), k::cy(nd), k::cz(nd));
              sfor<D>([&](auto A) {
                constexpr int ax = decltype(A)::value;
                vel[ax] = T(1.5) * u_p[ax * ld + n] - T(0.5) * u_p[ax * ld + nb];
              });
            }
            const T value = bc.kind == 1 ? T(bc.value) : c_p[n];
            const T udu = udotu<L>(vel);
            sfor<L::NQF>([&](auto KI) {
              constexpr int i = Dv::qface(F, decltype(KI)::value);
              constexpr int ib = k::qbar(i);
              const T udc = udotc<L, i>(vel);
              T t = T(1.0) + udc * T(3.0);
              t = t + T(0.5) * udc * udc * T(9.0);
              t = t - T(0.5) * udu * T(3.0);
              const T geq = T(k::w(i)) * value * t;
              gv[ib] = -src[i * ld + n] + T(2.0) * geq;
            });
          } else if (bc.kind == 3) {  // BBNoFlux
            sfor<L::NQF>([&](auto KI) {
              constexpr int i = Dv::qface(F, decltype(KI)::value);
              gv[k::qbar(i)] = src[i * ld + n];
            });
          } else if (bc.kind == 4 || bc.kind == 5) {  // zeroth / second order no-flux
            const long long nb = n + g.off(k::cx(nd), k::cy(nd), k::cz(nd));
            T gn[Q];
            const uint32_t cn = a.cls[nb];
            if (cn & kSolidBit) {
              sfor<Q>([&](auto J) { gn[decltype(J)::value] = T(0); });
            } else {
              Geo g2;
              g2.init(nb, a.nx, a.ny, a.nz);
              scalar_gather<L, T>(a, src, nb, cn, g2, c_p, gn);
            }
            if (bc.kind == 5) {
              Geo g2;
              g2.init(nb, a.nx, a.ny, a.nz);
              const long long nn = nb + g2.off(k::cx(nd), k::cy(nd), k::cz(nd));
              T gnn[Q];
              const uint32_t cnn = a.cls[nn];
              if (cnn & kSolidBit) {
                sfor<Q>([&](auto J) { gnn[decltype(J)::value] = T(0); });
              } else {
                Geo g3;
                g3.init(nn, a.nx, a.ny, a.nz);
                scalar_gather<L, T>(a, src, nn, cnn, g3, c_p, gnn);
              }
              sfor<Q>([&](auto J) {
                constexpr int j = decltype(J)::value;
                gn[j] = T(2.0) * gn[j] - gnn[j];
              });
            }
            if (bc.unknown_only) {
              sfor<L::NQF>([&](auto KI) {
                constexpr int i = Dv::qface(F, decltype(KI)::value);
                gv[k::qbar(i)] = gn[k::qbar(i)];
              });
            } else {
              sfor<Q>([&](auto J) {
                constexpr int j = decltype(J)::value;
                gv[j] = gn[j];
              });
            }
          }
        });
      }
    }
    if (a.mode == kModeStreamOnly) {
      readback_store<Q>(a, n, gv);
      continue;
    }
    const bool bb = (cls & kLinkMask) != 0;
    double q = a.has_ads ? a.q_ads[n] : 0.0;
    // compute_scalar_value, node.rs:272-279
    T cval = gv[0];
    sfor<Q - 1>([&](auto J) { cval = cval + gv[decltype(J)::value + 1]; });
    T S;
    if (scalar_source<L, T>(a, n, bb, c_p[n], q, S)) cval += T(0.5) * S;
    c_c[n] = cval;
    if (a.flags & kFlagResiduals) {
      const double dc = (double)cval - (double)c_p[n];
      acc += dc * dc;
    }
    T u[3];
    sfor<D>([&](auto A) {
      constexpr int ax = decltype(A)::value;
      u[ax] = u_c[ax * ld + n];
    });
    T geq[Q], gs[Q];
    equilibrium<L>(cval, u, geq);
    if constexpr (OP == 0) {  // node.rs:195-213
      bgk<L>(gv, geq, omega, omega_p, gs);
      if (scalar_source<L, T>(a, n, bb, cval, q, S)) {
        sfor<Q>([&](auto J) {
          constexpr int j = decltype(J)::value;
          gs[j] += coeff_b * T(k::w(j)) * S;
        });
      }
    } else if constexpr (OP == 1) {
      trt<L>(gv, geq, wp, wm, gs);
    } else {
      T s[Q];
      sfor<Q>([&](auto KI) {
        constexpr int kk = decltype(KI)::value;
        s[kk] = T(a.s[kk]);
      });
      mrt<L>(cval, u, gv, geq, s, gs);
    }
    if (a.has_ads) a.q_ads[n] = q;
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      dst[j * ld + n] = gs[j];
    });
  }
  if (a.mode == kModeStreamOnly) return;
  if (a.flags & kFlagResiduals) {
    __shared__ double lds[kBlock / 64];
    const double s = block_sum(acc, lds);
    if (threadIdx.x == 0) a.partials[(long long)blockIdx.x * kMaxCols + 4 + (a.field - 1)] = s;
  }
}

// ------------------------------------------------------------- init
// f = f_eq(rho, u) on fluid nodes (src/momentum/lattice.rs:322-328); 0 on solids.
template <class L, class St>
__global__ void __launch_bounds__(kBlock) momentum_init(KArgs a) {
  using T = typename St::T;
  using S = typename St::S;
  constexpr int Q = L::Q, D = L::D;
  const int cur = (int)((a.ctrl->step + 1) & 1);
  const int par = (int)(a.ctrl->step & 1);
  S* dst = (S*)a.pop[par];
  const T* rho = (const T*)a.rho[cur];
  const T* uu = (const T*)a.u[cur];
  for (long long n = (long long)blockIdx.x * kBlock + threadIdx.x; n < a.N;
       n += (long long)gridDim.x * kBlock) {
    T feq[Q];
    if (a.cls[n] & kSolidBit) {
      sfor<Q>([&](auto J) { feq[decltype(J)::value] = T(0); });
    } else {
      T u[3];
      sfor<D>([&](auto A) {
        constexpr int ax = decltype(A)::value;
        u[ax] = uu[ax * a.ld + n];
      });
      if constexpr (St::shifted) {
        // stored units scale * (f_eq - w)
        equilibrium_shifted<L>(rho[n] - T(1), rho[n], u, feq);
        sfor<Q>([&](auto J) { feq[decltype(J)::value] *= T(a.scale); });
      } else {
        equilibrium<L>(rho[n], u, feq);
      }
    }
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      stp<St>(dst + j * a.ld + n, feq[j]);
    });
  }
}

// g = g_eq(c, u) on fluid nodes (src/passive_scalar/lattice.rs:241-246)
template <class L, class T>
__global__ void __launch_bounds__(kBlock) scalar_init(KArgs a) {
  constexpr int Q = L::Q, D = L::D;
  const int cur = (int)((a.ctrl->step + 1) & 1);
  const int par = (int)(a.ctrl->step & 1);
  T* dst = (T*)a.pop[par];
  const T* cc = (const T*)a.c[cur];
  const T* uu = (const T*)a.u[cur];
  for (long long n = (long long)blockIdx.x * kBlock + threadIdx.x; n < a.N;
       n += (long long)gridDim.x * kBlock) {
    T geq[Q];
    if (a.cls[n] & kSolidBit) {
      sfor<Q>([&](auto J) { geq[decltype(J)::value] = T(0); });
    } else {
      T u[3];
      sfor<D>([&](auto A) {
        constexpr int ax = decltype(A)::value;
        u[ax] = uu[ax * a.ld + n];
      });
      equilibrium<L>(cc[n], u, geq);
    }
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      dst[j * a.ld + n] = geq[j];
    });
  }
}

}  // namespace lbf

// tb.hpp — two time steps per pass over the two-array scheme, the first
// step's results held in LDS.
//
// A workgroup owns a tile of TX x TY node columns and a segment [z0, z1) of
// the z axis and marches along z. For each plane p = z0-1 .. z1 it runs the
// first step (pull from the source array, bounce-back, faces, collide) on the
// tile plus a one-node ring around it and keeps the post-collision values in
// an LDS ring of three planes. It then runs the second step on plane p-1 of
// the tile itself, pulling every population from LDS, and stores f* to the
// destination array. Per two steps a node's Q values are read about
// (1 + ring share) times and written once, where two single steps read and
// write them twice each; neighbouring tiles share their ring reads through
// L2 (tiles of one z segment run on one XCD, see tb_tile).
//
// Arithmetic: the two-array kernel's (FP64: node_math.hpp calls in the
// reference's order without contraction; FP32: pair_collide), with the same
// pull / bounce-back / face order, so one pass gives the bits of two
// momentum_step_plain launches in either storage.
// Scope: 3-D lattices, FP64 / FP32 storage, faces periodic, no-slip or moving
// bounce-back (opposite faces both periodic or both not), no scalars, no
// Shan-Chen force, no residuals (those steps run the single-step kernel).
#pragma once

#include "fast.hpp"
#include "kernels.hpp"

namespace lbf {

// Tile shape by bytes per node (Q * storage size): the LDS ring of three
// planes of (TX + 2) x (TY + 2) nodes, plus their class words, fits one CU's
// 160 KiB. SHAPE 1 takes rows twice as wide (64 nodes) where that fits.
template <class L, class St, int SHAPE>
struct TbShape {
  static constexpr int BPN = L::Q * (int)sizeof(typename St::S);
  static constexpr int TX = BPN > 152 ? 16 : (SHAPE == 1 && BPN <= 80 ? 64 : 32);
  static constexpr int TY = BPN <= 80 ? (SHAPE == 1 ? 8 : 16) : 8;
  static constexpr int HX = TX + 2, HY = TY + 2, HN = HX * HY;
  static constexpr int NT = (HN + 63) / 64 * 64;  // one thread per ring-plane node
  static constexpr int CORE = TX * TY;
  // waves per SIMD of the one workgroup on a CU (registers are budgeted for them)
  static constexpr int WAVES_PER_SIMD = (NT / 64 + 3) / 4;
  static constexpr long LDS = 3L * HN * (BPN + 4);
  static_assert(NT <= 1024 && CORE <= NT, "tile shape");
  static_assert(LDS <= 160L * 1024, "LDS ring exceeds a CU");
};

// face conditions of kinds no-slip (1) and moving bounce-back (2) in face
// order, as apply_momentum_faces; own(i) is the node's own f*_i. The gather
// already put f_ibar = f*_i on every link that crosses a non-periodic face.
// own is called with std::integral_constant<int, i>.
template <class L, class T, class Own>
__device__ __forceinline__ void tb_faces(const KArgs& a, uint32_t on, T* f, Own&& own) {
  using k = K<L>;
  using Dv = Derived<L>;
  sfor<2 * L::D>([&](auto FI) {
    constexpr int F = decltype(FI)::value;
    if (!(on >> F & 1u)) return;
    const FaceK& bc = a.face[F];
    T vel[3] = {T(0), T(0), T(0)};
    if (bc.kind == 2)
      sfor<L::D>([&](auto A) { vel[decltype(A)::value] = T(bc.vel[decltype(A)::value]); });
    sfor<L::NQF>([&](auto KI) {
      constexpr int i = Dv::qface(F, decltype(KI)::value);
      constexpr int ib = k::qbar(i);
      const T fsi = own(std::integral_constant<int, i>{});
      if (bc.kind == 1) {
        f[ib] = fsi;  // bc.rs:119-129
      } else if (bc.kind == 2) {
        const T udc = udotc<L, i>(vel);  // bc.rs:62-84
        f[ib] = fsi - T(2.0) * T(k::w(i)) * T(bc.density) * T(3.0) * udc;
      }
    });
  });
}

// moments, force, collision of one node: the two-array kernel's sequence
// (kernels.hpp momentum_step_plain) for a node without Shan-Chen force;
// FP32 in pair form (pair_collide), FP64 in the reference's order
template <class L, class St, int OP>
__device__ __forceinline__ void tb_collide(const KArgs& a, uint32_t n, const typename St::T* f,
                                           typename St::T* fs, typename St::T& rho,
                                           typename St::T* u) {
  using T = typename St::T;
  constexpr int Q = L::Q, D = L::D;
  T F[3] = {T(0), T(0), T(0)};
  const bool has_force = a.force_kind != 0;
  if (has_force)
    sfor<D>([&](auto A) {
      constexpr int ax = decltype(A)::value;
      F[ax] = a.force_kind == 1 ? T(a.F[ax]) : T(a.src_field[ax * a.ld + n]);
    });
  if constexpr (!std::is_same<St, StF64>::value) {
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
    return;
  }
  moments<L>(f, rho, u);
  if (has_force)
    sfor<D>([&](auto A) {
      constexpr int ax = decltype(A)::value;
      u[ax] += T(0.5) * F[ax] / rho;  // node.rs:450-455
    });
  T feq[Q];
  equilibrium<L>(rho, u, feq);
  if constexpr (OP == 0) {
    bgk<L>(f, feq, T(a.omega), T(a.omega_p), fs);
    if (has_force) add_guo<L>(u, F, T(a.coeff_b), fs);
  } else if constexpr (OP == 1) {
    trt<L>(f, feq, T(a.wp), T(a.wm), fs);
  } else {
    T s[Q];
    sfor<Q>([&](auto KI) { s[decltype(KI)::value] = T(a.s[decltype(KI)::value]); });
    mrt<L>(rho, u, f, feq, s, fs);
  }
}

// Tile and z segment of this workgroup. Workgroups are dealt round-robin to
// the 8 XCDs; when the grid is a multiple of 8, XCD k takes the k-th eighth
// of the (segment, tile) order, so the tiles running together on one XCD are
// neighbours in one segment and share their ring reads in its L2.
__device__ __forceinline__ uint32_t tb_tile(uint32_t b, uint32_t nb) {
  return (nb & 7u) ? b : (b & 7u) * (nb >> 3) + (b >> 3);
}

// One pass = steps t and t+1 (t = the device counter + step_offset): reads
// pop[t & 1] (f* of step t-1, or post-stream f when fresh), writes f* of
// step t+1 into pop[(t & 1) ^ 1]; the host swaps the two pointers after
// enqueueing the pass, so the state is again in pop[(t + 2) & 1]. MOM stores
// the second step's moments into rho / u[(t + 1) & 1].
template <class L, class St, int OP, bool MOM, int SHAPE>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, TbShape<L, St, SHAPE>::NT),
                               amdgpu_waves_per_eu(TbShape<L, St, SHAPE>::WAVES_PER_SIMD)))
    momentum_step_tb2(KArgs a) {
  using T = typename St::T;
  using S = typename St::S;
  using k = K<L>;
  using Sh = TbShape<L, St, SHAPE>;
  constexpr int Q = L::Q, TX = Sh::TX, TY = Sh::TY, HX = Sh::HX, HN = Sh::HN;
  static_assert(L::D == 3, "two-step passes march along z");
  __shared__ S ring[3][Q][HN];
  __shared__ uint32_t cring[3][HN];

  const Ctrl* ctrl = a.ctrl;
  if (ctrl->stopped) return;
  const unsigned long long step = ctrl->step + a.step_offset;
  const int par = (int)(step & 1);
  const bool fresh = a.step_offset == 0 && ctrl->fresh[0] != 0;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(a.pop[par], (short)0, (int)0xffffffff, 0x00020000);
  const __amdgpu_buffer_rsrc_t rd =
      __builtin_amdgcn_make_buffer_rsrc(a.pop[par ^ 1], (short)0, (int)0xffffffff, 0x00020000);
  const uint32_t pb = a.plane_bytes;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t nxy = (uint32_t)nx * (uint32_t)ny;
  const uint32_t act = a.face_active;
  const bool per_x = !(act & 3u), per_y = !(act & 12u), per_z = !(act & 48u);

  const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
  const uint32_t tiles = (uint32_t)(ntx * nty);
  const uint32_t g = tb_tile(blockIdx.x, gridDim.x);
  const int seg = (int)(g / tiles), tile = (int)(g % tiles);
  const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * TY;
  const int wx = min(TX, nx - x0), wy = min(TY, ny - y0);
  const int z0 = seg * a.tb_zs, z1 = min(z0 + a.tb_zs, nz);
  if (z0 >= nz) return;  // whole workgroup: no barrier is skipped by a part of it

  // ---- this thread's node of the first step's ring plane
  const int t = threadIdx.x;
  const int lx = t % HX, ly = t / HX;
  bool h_ok = t < HN && lx <= wx + 1 && ly <= wy + 1;
  int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
  if (gx < 0 || gx >= nx) {
    h_ok = h_ok && per_x;
    gx = gx < 0 ? gx + nx : gx - nx;
  }
  if (gy < 0 || gy >= ny) {
    h_ok = h_ok && per_y;
    gy = gy < 0 ? gy + ny : gy - ny;
  }
  // wrapped neighbour offsets for shifts -1, 0, +1 (x, y; z per plane)
  int32_t ox[3], oy[3];
  ox[0] = gx == 0 ? nx - 1 : -1, ox[1] = 0, ox[2] = gx == nx - 1 ? -(nx - 1) : 1;
  oy[0] = gy == 0 ? (ny - 1) * nx : -nx, oy[1] = 0, oy[2] = gy == ny - 1 ? -(ny - 1) * nx : nx;
  const uint32_t on_xy = (gx == 0 ? 1u : 0u) | (gx == nx - 1 ? 2u : 0u) | (gy == 0 ? 4u : 0u) |
                         (gy == ny - 1 ? 8u : 0u);
  const uint32_t nxy_node = (uint32_t)gx + (uint32_t)nx * (uint32_t)gy;

  // plane p of the first step for this thread: node index and validity
  auto plane_node = [&](int p, uint32_t& n, int& gz) -> bool {
    bool ok = h_ok;
    gz = p;
    if (p < 0 || p >= nz) {
      ok = ok && per_z;
      gz = p < 0 ? p + nz : p - nz;
    }
    n = nxy_node + (uint32_t)gz * nxy;
    return ok;
  };
  auto class_of = [&](int p) -> uint32_t {
    uint32_t n;
    int gz;
    return plane_node(p, n, gz) ? a.cls[n] : kSolidBit;
  };

  // loaded first-step populations of one plane (in flight across the
  // second step of the previous plane)
  S v[Q];
  uint32_t v_cls = kSolidBit, v_n = 0, v_on = 0;
  auto issue = [&](int p, uint32_t cls) {
    uint32_t n;
    int gz;
    const bool ok = plane_node(p, n, gz);
    v_cls = ok ? cls : kSolidBit;
    v_n = n;
    const uint32_t on = (on_xy | (gz == 0 ? 16u : 0u) | (gz == nz - 1 ? 32u : 0u)) & act;
    v_on = on;
    if (v_cls & kSolidBit) return;
    const uint32_t own = n * (uint32_t)sizeof(S);
    if (fresh) {
      sfor<Q>([&](auto J) {
        constexpr int j = decltype(J)::value;
        v[j] = bload<S>(rs, own, j * pb);
      });
      return;
    }
    const int32_t oz[3] = {gz == 0 ? (int32_t)((nz - 1) * nxy) : -(int32_t)nxy, 0,
                           gz == nz - 1 ? -(int32_t)((nz - 1) * nxy) : (int32_t)nxy};
    const uint32_t sel = (v_cls & kLinkMask) | face_mask<L>(on);
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      constexpr int jb = k::qbar(j);
      // f_j(x) = f*_j(x - c_j), or f*_jb(x) across a solid link or a face
      const int32_t d = ox[1 - k::cx(j)] + oy[1 - k::cy(j)] + oz[1 - k::cz(j)];
      const uint32_t nb = (uint32_t)((int32_t)n + d) * (uint32_t)sizeof(S) + j * pb;
      const uint32_t bb = own + jb * pb;
      v[j] = bload<S>(rs, opaque((sel >> jb & 1u) ? bb : nb), 0u);
    });
  };
  // first step of a taken plane (post-stream f, class word, node) into ring
  // slot `slot`
  auto first = [&](int slot, const T* f, uint32_t cls, uint32_t n) {
    if (t < HN) cring[slot][t] = cls;
    if (cls & kSolidBit) return;
    T fs[Q], rho, u[3];
    tb_collide<L, St, OP>(a, n, f, fs, rho, u);
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      ring[slot][j][t] = static_cast<S>(fs[j]);
    });
  };
  // second step on plane p2 of the core; s1 = its slot, sm / sp = the slots
  // of planes p2 - 1 / p2 + 1
  T* rho_c = (T*)a.rho[par ^ 1];
  T* u_c = (T*)a.u[par ^ 1];
  auto second = [&](int p2, int sm, int s1, int sp) {
    if (t >= TX * TY) return;
    const int cx = t % TX, cy = t / TX;
    if (cx >= wx || cy >= wy) return;
    const int h = (cy + 1) * HX + cx + 1;
    const uint32_t cls = cring[s1][h];
    if (cls & kSolidBit) return;
    const int x = x0 + cx, y = y0 + cy;
    const uint32_t on = ((x == 0 ? 1u : 0u) | (x == nx - 1 ? 2u : 0u) | (y == 0 ? 4u : 0u) |
                         (y == ny - 1 ? 8u : 0u) | (p2 == 0 ? 16u : 0u) |
                         (p2 == nz - 1 ? 32u : 0u)) & act;
    const uint32_t sel = (cls & kLinkMask) | face_mask<L>(on);
    T f[Q], fs[Q], rho, u[3];
    const S* flat = &ring[0][0][0];
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      constexpr int jb = k::qbar(j);
      constexpr int cz = k::cz(j);
      const int sl = cz > 0 ? sm : (cz < 0 ? sp : s1);
      // one LDS read per direction: the address selects pull or bounce-back
      const int pulled = (sl * Q + j) * HN + h - k::cx(j) - k::cy(j) * HX;
      const int back = (s1 * Q + jb) * HN + h;
      f[j] = T(flat[(sel >> jb & 1u) ? back : pulled]);
    });
    if (on)
      tb_faces<L, T>(a, on, f, [&](auto I) { return T(ring[s1][decltype(I)::value][h]); });
    const uint32_t n = (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)p2);
    tb_collide<L, St, OP>(a, n, f, fs, rho, u);
    if constexpr (MOM) {
      rho_c[n] = rho;
      sfor<3>([&](auto A) {
        constexpr int ax = decltype(A)::value;
        u_c[ax * a.ld + n] = u[ax];
      });
    }
    const uint32_t own = n * (uint32_t)sizeof(S);
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      bstore<S>(static_cast<S>(fs[j]), rd, own, j * pb);
    });
  };

  // ---- march: iteration i runs the first step on plane z0 - 1 + i and the
  // second step on plane z0 - 2 + i (from i = 2). The next plane's loads go
  // out after the first barrier (kEarly: before the first step's arithmetic,
  // which measured 5 % slower in FP32 and spills in FP64).
  constexpr bool kEarly = false;
  const int np = (z1 - z0) + 2;
  uint32_t c_next = class_of(z0);
  issue(z0 - 1, class_of(z0 - 1));
  for (int i = 0; i < np; i++) {
    const int slot = i % 3;
    // take this plane's loaded values (post-stream f after the faces)
    auto take = [&](T* f) {
      if (v_cls & kSolidBit) return;
      sfor<Q>([&](auto J) { f[decltype(J)::value] = T(v[decltype(J)::value]); });
      // the gather loaded f*_i into v[ibar] on every face-crossing link i
      if (!fresh && v_on)
        tb_faces<L, T>(a, v_on, f, [&](auto I) { return T(v[k::qbar(decltype(I)::value)]); });
    };
    auto next = [&]() {
      if (i + 1 < np) {
        const uint32_t c = c_next;
        if (i + 2 < np) c_next = class_of(z0 + i + 1);
        issue(z0 + i, c);
      }
    };
    if constexpr (kEarly) {
      // the next plane's loads are in flight across this plane's first step,
      // the barriers and the second step
      T f[Q];
      const uint32_t cls = v_cls, n1 = v_n;
      take(f);
      next();
      first(slot, f, cls, n1);
    } else {
      [&]() {
        T f[Q];
        take(f);
        first(slot, f, v_cls, v_n);
      }();
    }
    __syncthreads();  // ring slot `slot` complete
    if constexpr (!kEarly) next();
    if (i >= 2) second(z0 - 2 + i, (i + 1) % 3, (i + 2) % 3, slot);
    __syncthreads();  // second step's reads done before the next first step overwrites
  }
}

}  // namespace lbf

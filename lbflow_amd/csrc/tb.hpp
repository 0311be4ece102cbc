// tb.hpp — two time steps per pass over the two-array scheme, the first
// step's results held in LDS.
//
// A workgroup owns a tile of TX x TY node columns and a segment [z0, z1) of
// the z axis and marches along z. For each plane p = z0-1 .. z1 it runs the
// first step (pull from the source array, bounce-back, faces, collide) on the
// tile plus a one-node ring around it and keeps the post-collision values in
// an LDS ring of planes. It then runs the second step on plane p-1 of the
// tile itself, pulling every population from LDS, and stores f* to the
// destination array. Per two steps a node's Q values are read about
// (1 + ring share) times and written once, where two single steps read and
// write them twice each; neighbouring tiles share their ring reads through
// L2 (tiles of one z segment run on one XCD, see tb_tile).
//
// Arithmetic: the two-array kernel's (FP64: node_math.hpp calls in the
// reference's order without contraction; FP32 and shifted storage:
// pair_collide), with the same pull / bounce-back / face order, so one pass
// gives the bits of two momentum_step_plain launches in every storage.
// Scope: 3-D lattices, faces periodic, no-slip or moving bounce-back
// (opposite faces both periodic or both not), no scalars, no Shan-Chen
// force, no residuals (those steps run the single-step kernel).
#pragma once

#include "fast.hpp"
#include "kernels.hpp"

namespace lbf {

// The LDS ring holds the first step's post-collision values of the tile and
// its one-node ring, HN = (TX + 2) x (TY + 2) nodes per plane. The second
// step on plane p reads the directions with c_z = +1 from plane p - 1, c_z = 0
// and every bounce-back value from plane p, and c_z = -1 from plane p + 1. So
// the c_z = +1 directions live in a ring of four planes and the others in a
// ring of three: the first step of the next plane then writes nothing the
// second step of this one still reads, and one barrier per plane suffices.
template <class L>
constexpr int tb_up_dirs() {
  int n = 0;
  for (int i = 0; i < L::Q; i++) n += L::c[i][2] > 0 ? 1 : 0;
  return n;
}
// position of direction j in its ring (c_z = +1 ring or the other)
template <class L>
constexpr int tb_ring_index(int j) {
  int n = 0;
  for (int i = 0; i < j; i++) n += (L::c[i][2] > 0) == (L::c[j][2] > 0) ? 1 : 0;
  return n;
}
template <class L, class St>
constexpr long tb_lds_bytes(int tx, int ty) {
  return (long)(4 * tb_up_dirs<L>() + 3 * (L::Q - tb_up_dirs<L>())) * (tx + 2) * (ty + 2) *
         (long)sizeof(typename St::S);
}
constexpr long kTbLds = 160L * 1024;
// Tile shape: the first of these (TX, TY) whose ring fits one CU's 160 KiB.
// Wide rows first (the loads of a wavefront then touch fewer cache lines;
// 256^3 D3Q19 FP32: 64 x 8 417 us per step, 32 x 16 432 us), then square
// tiles of 128-byte rows for 8-byte storage.
constexpr int kTbShapes[2][8][2] = {
    {{32, 16}, {32, 8}, {16, 16}, {16, 8}, {32, 4}, {16, 4}, {16, 2}, {8, 2}},
    {{64, 8}, {32, 8}, {16, 16}, {16, 8}, {32, 4}, {16, 4}, {16, 2}, {8, 2}}};
template <class L, class St, int SHAPE>
constexpr int tb_shape_index() {
  for (int k = 0; k < 8; k++)
    if (tb_lds_bytes<L, St>(kTbShapes[SHAPE][k][0], kTbShapes[SHAPE][k][1]) <= kTbLds) return k;
  return 7;
}
template <class L, class St, int SHAPE>
struct TbShape {
  static constexpr int TX = kTbShapes[SHAPE][tb_shape_index<L, St, SHAPE>()][0];
  static constexpr int TY = kTbShapes[SHAPE][tb_shape_index<L, St, SHAPE>()][1];
  static constexpr int HX = TX + 2, HY = TY + 2, HN = HX * HY;
  static constexpr int NT = (HN + 63) / 64 * 64;  // one thread per ring-plane node
  static constexpr int CORE = TX * TY;
  static constexpr int NUP = tb_up_dirs<L>();    // directions with c_z = +1
  static constexpr int NREST = L::Q - NUP;
  // waves per SIMD of the one workgroup on a CU (registers are budgeted for them)
  static constexpr int WAVES_PER_SIMD = (NT / 64 + 3) / 4;
  static constexpr long LDS = tb_lds_bytes<L, St>(TX, TY);
  static_assert(NT <= 1024 && CORE <= NT, "tile shape");
  static_assert(LDS <= kTbLds, "LDS ring exceeds a CU");
};

// face conditions of kinds no-slip (1) and moving bounce-back (2) in face
// order, as apply_momentum_faces; own(i) is the node's own f*_i. The gather
// already put f_ibar = f*_i on every link that crosses a non-periodic face.
// own is called with std::integral_constant<int, i>.
template <class L, class T, bool SHIFTED, class Own>
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
        // shifted storage (f - w) * scale: the correction scales, w_ib = w_i
        const T dv = T(2.0) * T(k::w(i)) * T(bc.density) * T(3.0) * udc;
        f[ib] = fsi - (SHIFTED ? T(a.scale) * dv : dv);
      }
    });
  });
}

// moments, force, collision of one node: the two-array kernel's sequence
// (kernels.hpp momentum_step_plain) for a node without Shan-Chen force;
// FP32 and shifted storage in pair form (pair_collide), FP64 in the
// reference's order
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
  constexpr int Q = L::Q, TX = Sh::TX, TY = Sh::TY, HX = Sh::HX, HY = Sh::HY, HN = Sh::HN;
  static_assert(L::D == 3, "two-step passes march along z");
  constexpr int NUP = Sh::NUP, NREST = Sh::NREST;
  // ring of the c_z = +1 directions (4 planes), then the others (3 planes)
  __shared__ S ring[4 * NUP * HN + 3 * NREST * HN];
  // element index of direction J (std::integral_constant) of ring node h in
  // the slot of march iteration `it` (the iteration that wrote that plane)
  auto at = [&](auto J, int it, int h) -> int {
    constexpr int j = decltype(J)::value;
    constexpr int idx = tb_ring_index<L>(j);
    if constexpr (L::c[j][2] > 0)
      return ((it & 3) * NUP + idx) * HN + h;
    else
      return 4 * NUP * HN + ((it % 3) * NREST + idx) * HN + h;
  };

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

  // ---- this thread's node of the first step's ring plane: threads
  // [0, TX * HY) take the tile's own columns of the ring rows, one row of TX
  // nodes after another (a wavefront's loads then cover whole rows), the
  // next 2 HY threads the ring's two side columns
  const int t = threadIdx.x;
  int lx, ly;
  if (t < TX * HY) {
    ly = t / TX, lx = 1 + t % TX;
  } else {
    const int e = t - TX * HY;
    ly = e >> 1, lx = (e & 1) ? TX + 1 : 0;
  }
  const int hh = ly * HX + lx;  // ring index of that node
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
  // byte offsets (within a plane) of the node's 9 in-plane neighbours
  // x + (sx, sy), sx, sy in {-1, 0, 1}, wrapped: a load's address is then one
  // of these plus a per-plane, per-direction scalar
  uint32_t nb_xy[9];
  sfor<9>([&](auto K9) {
    constexpr int kk = decltype(K9)::value;
    nb_xy[kk] = (uint32_t)((int32_t)nxy_node + ox[kk % 3] + oy[kk / 3]) * (uint32_t)sizeof(S);
  });

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
    // byte offset of the plane gz + sz (sz in {-1, 0, 1}, wrapped): uniform
    const uint32_t pz = (uint32_t)gz * nxy * (uint32_t)sizeof(S);
    const uint32_t zb[3] = {
        (uint32_t)(gz == 0 ? nz - 1 : gz - 1) * nxy * (uint32_t)sizeof(S), pz,
        (uint32_t)(gz == nz - 1 ? 0 : gz + 1) * nxy * (uint32_t)sizeof(S)};
    const uint32_t sel = (v_cls & kLinkMask) | face_mask<L>(on);
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      constexpr int jb = k::qbar(j);
      // f_j(x) = f*_j(x - c_j), or f*_jb(x) across a solid link or a face
      constexpr int kk = (1 - k::cx(j)) + 3 * (1 - k::cy(j));
      const uint32_t nb = nb_xy[kk] + (zb[1 - k::cz(j)] + j * pb);
      const uint32_t bb = own + jb * pb;
      v[j] = bload<S>(rs, opaque((sel >> jb & 1u) ? bb : nb), 0u);
    });
  };
  // first step of a taken plane (post-stream f, class word, node) into ring
  // slot `slot`
  auto first = [&](int it, const T* f, uint32_t cls, uint32_t n) {
    if (cls & kSolidBit) return;
    T fs[Q], rho, u[3];
    tb_collide<L, St, OP>(a, n, f, fs, rho, u);
    sfor<Q>([&](auto J) { ring[at(J, it, hh)] = static_cast<S>(fs[decltype(J)::value]); });
  };
  // this thread's node of the second step (core column) and its class word
  const int ccx = t % TX, ccy = t / TX;
  const bool core_ok = t < TX * TY && ccx < wx && ccy < wy;
  const uint32_t core_xy = (uint32_t)(x0 + ccx) + (uint32_t)nx * (uint32_t)(y0 + ccy);
  auto core_class = [&](int p) -> uint32_t {
    return core_ok ? a.cls[core_xy + (uint32_t)p * nxy] : kSolidBit;
  };
  // second step in march iteration `it` on plane p2 of the core, whose
  // first-step values the iteration before wrote (it - 2: the plane below,
  // it: the plane above)
  T* rho_c = (T*)a.rho[par ^ 1];
  T* u_c = (T*)a.u[par ^ 1];
  auto second = [&](int p2, int it, uint32_t cls) {
    if (cls & kSolidBit) return;
    const int cx = ccx, cy = ccy;
    const int h = (cy + 1) * HX + cx + 1;
    const int x = x0 + cx, y = y0 + cy;
    const uint32_t on = ((x == 0 ? 1u : 0u) | (x == nx - 1 ? 2u : 0u) | (y == 0 ? 4u : 0u) |
                         (y == ny - 1 ? 8u : 0u) | (p2 == 0 ? 16u : 0u) |
                         (p2 == nz - 1 ? 32u : 0u)) & act;
    const uint32_t sel = (cls & kLinkMask) | face_mask<L>(on);
    T f[Q], fs[Q], rho, u[3];
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      constexpr int jb = k::qbar(j);
      constexpr int cz = k::cz(j);
      const int src = cz > 0 ? it - 2 : (cz < 0 ? it : it - 1);
      const int ip = at(J, src, h - k::cx(j) - k::cy(j) * HX);
      const int ib = at(std::integral_constant<int, jb>{}, it - 1, h);
      if constexpr (sizeof(S) <= 4) {
        // both candidates read at immediate offsets from a per-slot base,
        // then the value is selected (cheaper than selecting the address:
        // 256^3 FP32 413 -> 401 us per step)
        const S pulled = ring[ip], back = ring[ib];
        f[j] = T((sel >> jb & 1u) ? back : pulled);
      } else {
        // 8-byte values: one read at the selected address (reading both
        // measured 1136 against 1067 us per FP64 step)
        f[j] = T(ring[(sel >> jb & 1u) ? ib : ip]);
      }
    });
    if (on && !a.faces_noslip_only)
      tb_faces<L, T, St::shifted>(a, on, f, [&](auto I) { return T(ring[at(I, it - 1, h)]); });
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
  // out after the barrier (issuing them before the first step's arithmetic
  // measured 5 % slower in FP32 and spills in FP64).
  const int np = (z1 - z0) + 2;
#ifdef LBF_TB_POISON
  // check builds: the ring starts as NaN, so a read of a slot that no first
  // step of this launch wrote shows in the results
  for (int e = t; e < (int)(sizeof(ring) / sizeof(S)); e += Sh::NT) ring[e] = S(__builtin_nanf(""));
  __syncthreads();
#endif
  uint32_t c_next = class_of(z0);
  uint32_t c2_next = core_class(z0);  // the first second step's plane
  issue(z0 - 1, class_of(z0 - 1));
  for (int i = 0; i < np; i++) {
    // take this plane's loaded values (post-stream f after the faces)
    auto take = [&](T* f) {
      if (v_cls & kSolidBit) return;
      sfor<Q>([&](auto J) { f[decltype(J)::value] = T(v[decltype(J)::value]); });
      // the gather loaded f*_i into v[ibar] on every face-crossing link i,
      // which is all a no-slip face asks
      if (!fresh && v_on && !a.faces_noslip_only)
        tb_faces<L, T, St::shifted>(a, v_on, f, [&](auto I) { return T(v[k::qbar(decltype(I)::value)]); });
    };
    [&]() {
      T f[Q];
      take(f);
      first(i, f, v_cls, v_n);
    }();
    // the only barrier of the iteration: this plane's ring slot is complete,
    // and every wave has left the previous iteration's second step, whose
    // ring slots the next first step may overwrite
    __syncthreads();
    uint32_t c2 = kSolidBit;
    if (i >= 2) {
      c2 = c2_next;
      if (i + 1 < np) c2_next = core_class(z0 - 1 + i);
    }
    // the next plane's loads, in flight across the second step
    if (i + 1 < np) {
      const uint32_t c = c_next;
      if (i + 2 < np) c_next = class_of(z0 + i + 1);
      issue(z0 + i, c);
    }
    if (i >= 2) second(z0 - 2 + i, i, c2);
  }
}

}  // namespace lbf

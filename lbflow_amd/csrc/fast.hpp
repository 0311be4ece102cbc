// fast.hpp — in-place AA-pattern step kernels (FP64 in the parity kernel's
// arithmetic, FP32 and shifted storage in pair form).
//
// One population array F[q][ld]. Even steps read the node's own slots
// F[j][x] (= post-stream f_j), collide, and write f*_j to the swapped own slot
// F[jbar][x]. Odd steps read f_j(x) from F[jbar][x - c_j] (the neighbour's
// swapped slot), collide, and write f*_j to F[j][x + c_j]. Every slot is read
// and written by the same thread within a step, so the update is in place:
// half the memory of a two-array scheme at the same 2*Q*s bytes per update.
//
// Bounce-back: the population arriving through a link whose neighbour is
// solid, or that crosses a face with a non-periodic condition, is
// f_ibar = f*_i (the node's own post-collision value sent back).
//  - Links across non-periodic faces: the odd step writes f*_i to the node's
//    own slot F[ibar][x] instead of F[i][x + c_i] (which is another node's).
//  - Links to a solid node y = x + c_i: the odd step writes f*_i to
//    F[i][y] as for a fluid neighbour (y never runs, so only x writes that
//    slot), and the following even step reads f_ibar(x) from there. Every
//    store of both steps then goes to a cache line that one wavefront writes
//    whole. Stores redirected into lines that another wavefront also writes
//    cost up to 1.6x the write traffic on a 30 %-solid lattice (measured);
//    the extra loads of this layout cost less.
//  - The even step writes only own slots, so the odd step reads a solid
//    link's value from the node's own slot F[i][x].
// Face conditions are then applied in face order from those values
// (src/momentum/bc.rs:29-129). This requires opposite faces to be both
// periodic or both not, which the host checks before choosing this scheme.
// At an edge where a Zou-He face meets another non-periodic face the closure
// therefore sees the other face's bounce-back values, where the reference,
// applying faces one after another, can still hold wrapped pull values; the
// host's automatic choice keeps such FP64 cases on the two-array scheme.
//
// Addressing: one buffer resource over the whole population array, a 32-bit
// per-thread byte offset (shared by all planes of a neighbour) and a
// per-plane scalar offset, so each load/store needs one VGPR of address. The
// host uses this scheme only while the array is below 4 GiB.
//
// Storage: plain f (f64 / f32) or shifted delta_i = (f_i - w_i) * scale in
// f32, fp16 or bf16 with all arithmetic in f32 on delta (node_math.hpp).
#pragma once

#include <type_traits>

#include "engine.hpp"
#include "kernels.hpp"
#include "node_math.hpp"

// Waves per SIMD the streaming variant of the in-place kernel is compiled
// for (D3Q19 and smaller lattices), by storage width and step parity.
#ifndef LBF_WAVES_16_EVEN
#define LBF_WAVES_16_EVEN 8
#endif
#ifndef LBF_WAVES_16_ODD
#define LBF_WAVES_16_ODD 6
#endif
#ifndef LBF_WAVES_32_EVEN
#define LBF_WAVES_32_EVEN 8
#endif
#ifndef LBF_WAVES_32_ODD
#define LBF_WAVES_32_ODD 6
#endif
// waves per SIMD the moment-storing variant (last step of a batch, residual
// steps) is compiled for at least
#ifndef LBF_WAVES_MOM
#define LBF_WAVES_MOM 2
#endif

namespace lbf {

// equilibrium moments minus M w, in delta form, for shifted MRT:
// meq_k(rho,u) - (M w)_k = drho g_k(0) + rho h_k(u) + (g_k(0) - (M w)_k)
// with meq_k(rho, u) = rho g_k(u) the reference's closed forms and
// h_k(u) = g_k(u) - g_k(0) written out so no constant cancels in f32.
// Shifted storage units (scale * delta): G = scale * drho, B = scale * rho,
// and deq holds scale * delta_eq.
template <class L>
__device__ __forceinline__ void meq_shifted(float G, float B, float sc, const float* u,
                                            const float* deq, float* m) {
  using Dv = Derived<L>;
  auto mw = [](int k) {
    double s = 0;
    for (int i = 0; i < L::Q; i++) s += Dv::M(k, i) * Dv::w(i);
    return s;
  };
  if constexpr (L::id == 1) {
    const float ux = u[0], uy = u[1], u2 = ux * ux + uy * uy;
    const float h[9] = {0.f, -3.f * u2, 9.f * ux * ux * uy * uy - 3.f * u2, ux,
                        3.f * ux * ux * ux - ux, uy, 3.f * uy * uy * uy - uy, u2, ux * uy};
    constexpr double g0[9] = {1, 1, 1, 0, 0, 0, 0, 0, 0};
    sfor<9>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr float c = float(g0[k] - mw(k));
      m[k] = G * float(g0[k]) + B * h[k] + sc * c;
    });
  } else if constexpr (L::id == 3) {
    const float ux = u[0], uy = u[1], uz = u[2], u2 = ux * ux + uy * uy + uz * uz;
    const float h9 = 2.f * ux * ux - (uy * uy + uz + uz);
    const float h11 = uy * uy - uz * uz;
    const float h[19] = {0.f, 19.f * u2, -5.5f * u2, ux, (-2.f / 3.f) * ux, uy, (-2.f / 3.f) * uy,
                         uz, (-2.f / 3.f) * uz, h9, -0.5f * h9, h11, -0.5f * h11, ux * uy,
                         uy * uz, ux * uz, 0.f, 0.f, 0.f};
    constexpr double g0[19] = {1, -11, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    sfor<19>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr float c = float(g0[k] - mw(k));
      m[k] = G * float(g0[k]) + B * h[k] + sc * c;
    });
  } else if constexpr (L::id == 2) {
    const float ux = u[0], uy = u[1], uz = u[2], u2 = ux * ux + uy * uy + uz * uz;
    const float h[15] = {0.f, u2, -5.f * u2, ux, (-7.f / 3.f) * ux, uy, (-7.f / 3.f) * uy, uz,
                         (-7.f / 3.f) * uz, 2.f * ux * ux - (uy * uy + uz * uz), uy * uy - uz * uz,
                         ux * uy, uy * uz, ux * uz, 0.f};
    constexpr double g0[15] = {1, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    sfor<15>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr float c = float(g0[k] - mw(k));
      m[k] = G * float(g0[k]) + B * h[k] + sc * c;
    });
  } else {
    // no closed form: m_eq = M f_eq, so m_eq - M w = M delta_eq
    sfor<L::Q>([&](auto K) {
      constexpr int k = decltype(K)::value;
      m[k] = mrt_dot<L, k>(deq);
    });
  }
}

// A device-memory pointer as a global-address-space one: pointers read from
// memory (the serving tier's slot table) are otherwise generic and loaded
// through flat instructions; kernel-argument pointers are global already.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* global_ptr(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* global_ptr(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// ------------------------------------------------------------ buffer I/O
template <class S>
__device__ __forceinline__ S bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(S) == 8)
    return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else if constexpr (sizeof(S) == 4)
    return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else
    return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
}
template <class S>
__device__ __forceinline__ void bstore(S v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(S) == 8) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
  }
  else if constexpr (sizeof(S) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, voff, soff, 0);
}

// A per-lane byte offset the compiler cannot see through: without it,
// load(c ? a : b) is turned into branches around two loads with a wait
// inside each, which serialises the gather on mixed wavefronts.
__device__ __forceinline__ uint32_t opaque(uint32_t x) {
  asm volatile("" : "+v"(x));
  return x;
}

// directions crossing face F (c . n_out = +1)
template <class L>
constexpr uint32_t face_dirs(int F) {
  uint32_t m = 0;
  for (int i = 0; i < L::Q; i++)
    if (Derived<L>::in_face(F, i)) m |= 1u << i;
  return m;
}

// i leads the opposite pair (i, ibar): the rest direction and the second
// member of each pair are not leads
template <class L>
constexpr bool pair_lead(int i) {
  return i != 0 && i < K<L>::qbar(i);
}

// u . F over the lattice's dimensions
template <class L, class T>
__device__ __forceinline__ T udotu_F(const T* u, const T* F) {
  T s = u[0] * F[0] + u[1] * F[1];
  if constexpr (L::D == 3) s = s + u[2] * F[2];
  return s;
}

// Parity of row k of the MRT matrix under c -> -c: +1 if M(k, ibar) =
// M(k, i) for all i, -1 if M(k, ibar) = -M(k, i), 0 otherwise. Rows of
// fixed parity act on pair sums (even) or pair differences (odd) only.
template <class L>
constexpr int mrt_row_parity(int kr) {
  bool even = true, odd = true;
  for (int i = 0; i < L::Q; i++) {
    const double a = Derived<L>::M(kr, i), b = Derived<L>::M(kr, K<L>::qbar(i));
    if (a != b) even = false;
    if (a != -b) odd = false;
  }
  return even ? 1 : (odd ? -1 : 0);
}
template <class L>
constexpr bool mrt_rows_have_parity() {
  for (int kr = 0; kr < L::Q; kr++)
    if (mrt_row_parity<L>(kr) == 0) return false;
  return true;
}

// (M f)_k from the rest value f0 and the pair sums sp / differences dm of
// the pair leads (half the products of a dense row)
template <class L, int Krow, class T>
__device__ __forceinline__ T mrt_dot_pairs(T f0, const T* sp, const T* dm) {
  constexpr int par = mrt_row_parity<L>(Krow);
  static_assert(par != 0, "MRT row without parity");
  T s = T(0);
  bool first = true;
  auto add = [&](T t) {
    if (first) {
      s = t;
      first = false;
    } else {
      s = s + t;
    }
  };
  if constexpr (par > 0 && Derived<L>::M(Krow, 0) != 0.0) add(T(Derived<L>::M(Krow, 0)) * f0);
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (pair_lead<L>(i)) {
      constexpr double m = Derived<L>::M(Krow, i);
      if constexpr (m != 0.0) add(T(m) * (par > 0 ? sp[i] : dm[i]));
    }
  });
  return s;
}

// f = M^-1 ms over pairs: f_i, f_ibar = S_i +- A_i with S_i from the even
// rows and A_i from the odd rows (M^-1(ibar, k) = parity_k M^-1(i, k))
template <class L, class T>
__device__ __forceinline__ void mrt_inv_pairs(const T* ms, T* fs) {
  using Dv = Derived<L>;
  auto part = [&](auto I, auto W) -> T {
    constexpr int i = decltype(I)::value;
    T s = T(0);
    bool first = true;
    sfor<L::Q>([&](auto Kc) {
      constexpr int kk = decltype(Kc)::value;
      constexpr double m = Dv::Minv(i, kk);
      if constexpr (m != 0.0) {
        if constexpr (mrt_row_parity<L>(kk) == decltype(W)::value) {
          const T t = T(m) * ms[kk];
          if (first) {
            s = t;
            first = false;
          } else {
            s = s + t;
          }
        }
      }
    });
    return s;
  };
  using even = std::integral_constant<int, 1>;
  using odd = std::integral_constant<int, -1>;
  fs[0] = part(std::integral_constant<int, 0>{}, even{});
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (pair_lead<L>(i)) {
      const T S = part(I, even{}), A = part(I, odd{});
      fs[i] = S + A;
      fs[K<L>::qbar(i)] = S - A;
    }
  });
}

// D3Q27's own MRT basis (lattice.hpp) has tensor-product rows
// p_o1(cx) p_o2(cy) p_o3(cz) with p_0 = 1, p_1(a) = a, p_2(a) = 3a^2 - 2, so
// M x is three passes of the 3-point transform along x, y and z over the
// directions arranged as a 3x3x3 cube, and M^-1 likewise with p_o(a) / |p_o|^2
// (|p_o|^2 = 3, 2, 6): about 250 operations instead of two dense 27x27
// products. Returns dfs = M^-1 (s . (M neq)).
template <class L>
constexpr int cube_dir(int x, int y, int z) {
  for (int i = 0; i < L::Q; i++)
    if (L::c[i][0] == x - 1 && L::c[i][1] == y - 1 && L::c[i][2] == z - 1) return i;
  return -1;
}
template <class L>
constexpr int tensor_row(int ox, int oy, int oz) {
  for (int k = 0; k < L::Q; k++)
    if (L::ord[k][0] == ox && L::ord[k][1] == oy && L::ord[k][2] == oz) return k;
  return -1;
}
template <class L, class T>
__device__ __forceinline__ void mrt_tensor_relax(const T* neq, const T* s, T* dfs) {
  static_assert(L::Q == 27 && L::p(0, -1) == 1 && L::p(1, -1) == -1 && L::p(2, -1) == 1 &&
                    L::p(2, 0) == -2 && L::p(1, 1) == 1 && L::p(2, 1) == 1,
                "tensor-product MRT basis p0 = 1, p1 = a, p2 = 3a^2 - 2");
  T g[27];  // g[x + 3 (y + 3 z)], x, y, z = c + 1
  sfor<27>([&](auto I) {
    constexpr int q = decltype(I)::value;
    constexpr int i = cube_dir<L>(q % 3, (q / 3) % 3, q / 9);
    static_assert(i >= 0, "D3Q27 direction");
    g[q] = neq[i];
  });
  // forward along axis A (stride 1, 3, 9): (a0, a1, a2) -> (a0 + a1 + a2,
  // a2 - a0, a0 + a2 - 2 a1), order index in place of the coordinate
  auto fwd = [&](auto SI) {
    constexpr int st = decltype(SI)::value;
    sfor<9>([&](auto LI) {
      constexpr int l = decltype(LI)::value;
      constexpr int b = st == 1 ? 3 * l : (st == 3 ? (l % 3) + 9 * (l / 3) : l);
      const T a0 = g[b], a1 = g[b + st], a2 = g[b + 2 * st];
      const T e = a0 + a2;
      g[b] = e + a1;
      g[b + st] = a2 - a0;
      g[b + 2 * st] = e - T(2) * a1;
    });
  };
  auto inv = [&](auto SI) {
    constexpr int st = decltype(SI)::value;
    sfor<9>([&](auto LI) {
      constexpr int l = decltype(LI)::value;
      constexpr int b = st == 1 ? 3 * l : (st == 3 ? (l % 3) + 9 * (l / 3) : l);
      const T m0 = g[b] * T(1.0 / 3.0), m1 = g[b + st] * T(0.5), m2 = g[b + 2 * st] * T(1.0 / 6.0);
      const T e = m0 + m2;
      g[b] = e - m1;
      g[b + st] = m0 - T(2) * m2;
      g[b + 2 * st] = e + m1;
    });
  };
  fwd(std::integral_constant<int, 1>{});
  fwd(std::integral_constant<int, 3>{});
  fwd(std::integral_constant<int, 9>{});
  sfor<27>([&](auto I) {
    constexpr int q = decltype(I)::value;
    constexpr int kk = tensor_row<L>(q % 3, (q / 3) % 3, q / 9);
    static_assert(kk >= 0, "tensor row");
    g[q] = s[kk] * g[q];
  });
  inv(std::integral_constant<int, 9>{});
  inv(std::integral_constant<int, 3>{});
  inv(std::integral_constant<int, 1>{});
  sfor<27>([&](auto I) {
    constexpr int q = decltype(I)::value;
    dfs[cube_dir<L>(q % 3, (q / 3) % 3, q / 9)] = g[q];
  });
}

// 1 / x: the hardware reciprocal (1 ulp) in f32, a division in f64
template <class T>
__device__ __forceinline__ T recip(T x) {
  if constexpr (sizeof(T) == 4)
    return __builtin_amdgcn_rcpf(x);
  else
    return T(1) / x;
}

struct NodePos {
  uint32_t on;   // faces (with a non-periodic condition) the node lies on
  int32_t ox[3], oy[3], oz[3];  // wrapped neighbour offsets for shifts -1, 0, +1
};

template <class L>
__device__ __forceinline__ void node_pos(const KArgs& a, uint32_t n, bool need_offsets,
                                         NodePos& p) {
  const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny, nz = (uint32_t)a.nz;
  const uint32_t nxy = nx * ny;
  const uint32_t z = n / nxy;
  const uint32_t r = n - z * nxy;
  const uint32_t y = r / nx;
  const uint32_t x = r - y * nx;
  uint32_t on = 0;
  on |= (x == 0) ? 1u : 0u;
  on |= (x == nx - 1) ? 2u : 0u;
  on |= (y == 0) ? 4u : 0u;
  on |= (y == ny - 1) ? 8u : 0u;
  if constexpr (L::D == 3) {
    on |= (z == 0) ? 16u : 0u;
    on |= (z == nz - 1) ? 32u : 0u;
  }
  p.on = on & a.face_active;
  if (need_offsets) {
    p.ox[0] = x == 0 ? (int32_t)(nx - 1) : -1;
    p.ox[1] = 0;
    p.ox[2] = x == nx - 1 ? -(int32_t)(nx - 1) : 1;
    p.oy[0] = y == 0 ? (int32_t)((ny - 1) * nx) : -(int32_t)nx;
    p.oy[1] = 0;
    p.oy[2] = y == ny - 1 ? -(int32_t)((ny - 1) * nx) : (int32_t)nx;
    p.oz[0] = z == 0 ? (int32_t)((nz - 1) * nxy) : -(int32_t)nxy;
    p.oz[1] = 0;
    p.oz[2] = z == nz - 1 ? -(int32_t)((nz - 1) * nxy) : (int32_t)nxy;
  }
}

// directions crossing the non-periodic faces the node lies on
template <class L>
__device__ __forceinline__ uint32_t face_mask(uint32_t on) {
  uint32_t m = 0;
  if (on) {
    sfor<2 * L::D>([&](auto FI) {
      constexpr int Fc = decltype(FI)::value;
      if (on >> Fc & 1u) m |= face_dirs<L>(Fc);
    });
  }
  return m;
}

// Post-stream populations of node n in the compute type (shifted storage:
// scale * (f - w), the units the step runs in). links: bit i = neighbour
// x + c_i (periodic wrap) is solid; fm: face_mask.
template <class L, class St, bool ODD>
__device__ __forceinline__ void aa_gather(__amdgpu_buffer_rsrc_t rs, uint32_t plane_bytes,
                                          uint32_t n, bool fresh, uint32_t links, uint32_t fm,
                                          const NodePos& p, typename St::T* f) {
  using S = typename St::S;
  using T = typename St::T;
  using k = K<L>;
  const uint32_t own = n * (uint32_t)sizeof(S);
  // odd: links whose value is in the node's own slot; even: solid links
  // whose value the previous odd step left in the solid neighbour's slot
  const uint32_t sel = ODD ? (links | fm) : (links & ~fm);
  sfor<L::Q>([&](auto J) {
    constexpr int j = decltype(J)::value;
    constexpr int jb = k::qbar(j);
    S v;
    if (fresh) {
      v = bload<S>(rs, own, j * plane_bytes);
    } else if constexpr (!ODD) {
      // f_j(x) = f*_jb(x), stored at F[jb][x + c_jb] when that node is solid
      const int32_t d = p.ox[1 + k::cx(jb)] + p.oy[1 + k::cy(jb)] + p.oz[1 + k::cz(jb)];
      const uint32_t nb = (uint32_t)((int32_t)n + d) * (uint32_t)sizeof(S) + jb * plane_bytes;
      const uint32_t mine = own + j * plane_bytes;
      v = bload<S>(rs, opaque((sel >> jb & 1u) ? nb : mine), 0u);
    } else {
      // one load per direction: the redirect only selects the lane's address
      const int32_t d = p.ox[1 - k::cx(j)] + p.oy[1 - k::cy(j)] + p.oz[1 - k::cz(j)];
      const uint32_t nb = (uint32_t)((int32_t)n + d) * (uint32_t)sizeof(S) + jb * plane_bytes;
      const uint32_t mine = own + j * plane_bytes;
      v = bload<S>(rs, opaque((sel >> jb & 1u) ? mine : nb), 0u);
    }
    f[j] = T(v);
  });
}

// face conditions from the redirected values (f_ibar = f*_i on crossing links)
// (shifted storage: f holds scale * (f - w); sc / isc convert)
template <class L, class T, bool SHIFTED>
__device__ __forceinline__ void aa_faces(const KArgs& a, uint32_t on, T sc, T isc, T* f) {
  using k = K<L>;
  using Dv = Derived<L>;
  T fsv[L::Q];
  sfor<L::Q>([&](auto J) { fsv[decltype(J)::value] = f[decltype(J)::value]; });
  sfor<2 * L::D>([&](auto FI) {
    constexpr int Fc = decltype(FI)::value;
    if (!(on >> Fc & 1u)) return;
    const FaceK& bc = a.face[Fc];
    if (bc.kind == 4) {
      // Zou-He from the node's post-stream populations (the unknown ones hold
      // the redirected bounce-back values and are replaced); delta-form
      // storage converts to f around the closure
      if constexpr (SHIFTED)
        sfor<L::Q>([&](auto J) {
          constexpr int j = decltype(J)::value;
          f[j] = f[j] * isc + T(k::w(j));
        });
      if constexpr (L::has_zou_he) {
        (void)zou_he_d2q9<T>(Fc, f, bc.known, T(bc.density), T(bc.vel[0]), T(bc.vel[1]));
      } else {
        const T v3[3] = {T(bc.vel[0]), T(bc.vel[1]), T(bc.vel[2])};
        zou_he_general<L, T, Fc>(f, bc.known, T(bc.density), v3);
      }
      if constexpr (SHIFTED)
        sfor<L::Q>([&](auto J) {
          constexpr int j = decltype(J)::value;
          f[j] = (f[j] - T(k::w(j))) * sc;
        });
      return;
    }
    T vel[3] = {T(0), T(0), T(0)};
    if (bc.kind == 2)
      sfor<L::D>([&](auto A) { vel[decltype(A)::value] = T(bc.vel[decltype(A)::value]); });
    sfor<L::NQF>([&](auto KI) {
      constexpr int i = Dv::qface(Fc, decltype(KI)::value);
      constexpr int ib = k::qbar(i);
      if (bc.kind == 1) {
        f[ib] = fsv[ib];
      } else if (bc.kind == 2) {
        const T udc = udotc<L, i>(vel);
        const T dv = T(2.0) * T(k::w(i)) * T(bc.density) * T(3.0) * udc;
        f[ib] = fsv[ib] - (SHIFTED ? sc * dv : dv);
      }
    });
  });
}

// Moments and collision of one node in pair form (FP32 arithmetic on plain
// or shifted storage): the in-place, two-array and two-step FP32 kernels all
// call this, so they give the same bits. f: post-stream populations in the
// compute units (shifted storage: scale * (f - w)); get_force(Fv) fills the
// node's force and returns whether there is one (called once the velocity
// sums are formed); moments(rho, u) receives rho and u shifted by F / 2 rho;
// f* is returned in fs.
template <class L, class St, int OP, class GetForce, class Moments>
__device__ __forceinline__ void pair_collide(const KArgs& a, const typename St::T* f,
                                             GetForce&& get_force, Moments&& moments,
                                             typename St::T* fs) {
#pragma clang fp contract(fast)
  using T = typename St::T;
  using k = K<L>;
  constexpr int Q = L::Q, D = L::D;
  const T omega = T(a.omega), omega_p = T(a.omega_p), wp = T(a.wp), wm = T(a.wm),
          coeff_b = T(a.coeff_b);
  const T sc = T(a.scale), isc = T(a.inv_scale);
#include "pair_collide.inc"
}

// One node of the AA-pattern fused step (OP 0 BGK, 1 TRT, 2 MRT; ODD selects
// the step parity) with its class word `cls` (kSolidBit past the last node).
// `step` is the case's completed-step count, `fresh` marks an array that holds
// post-stream f in natural slots. MOM: store moments (and accumulate residual
// terms into acc) as the flags ask.
template <class L, class St, int OP, bool ODD, bool MOM>
__device__ __forceinline__ void aa_node_cls(const KArgs& a, uint32_t n, uint32_t cls,
                                            unsigned long long step, bool fresh, double* acc) {
#pragma clang fp contract(fast)
  using T = typename St::T;
  using S = typename St::S;
  using k = K<L>;
  constexpr int Q = L::Q, D = L::D;
  static_assert(k::qbar(0) == 0 && !k::any_c(0), "direction 0 is the rest direction");
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(a.pop[0], (short)0, (int)0xffffffff, 0x00020000);
  const uint32_t pb = a.plane_bytes;
  const int par = (int)(step & 1);
  __attribute__((address_space(1))) T* rho_c = global_ptr((T*)a.rho[par]);
  __attribute__((address_space(1))) T* u_c = global_ptr((T*)a.u[par]);
  const __attribute__((address_space(1))) T* rho_p = global_ptr((const T*)a.rho[par ^ 1]);
  const __attribute__((address_space(1))) T* u_p = global_ptr((const T*)a.u[par ^ 1]);
  const bool store_m = MOM && (a.flags & (kFlagStoreMoments | kFlagResiduals)) != 0;
  const bool resid = MOM && (a.flags & kFlagResiduals) != 0;
  const T omega = T(a.omega), omega_p = T(a.omega_p), wp = T(a.wp), wm = T(a.wm),
          coeff_b = T(a.coeff_b);
  const T sc = T(a.scale), isc = T(a.inv_scale);
  const uint32_t ld = (uint32_t)a.ld;
  if (!(cls & kSolidBit)) {
    NodePos p;
    node_pos<L>(a, n, true, p);
    const uint32_t fm = face_mask<L>(p.on);
    T f[Q];
    aa_gather<L, St, ODD>(rs, pb, n, fresh, cls & kLinkMask, fm, p, f);
    // shifted storage: f holds scale * (f - w) and the step runs in those units
    if (p.on && !fresh && !a.faces_noslip_only) aa_faces<L, T, St::shifted>(a, p.on, sc, isc, f);
    // moments (and residual terms) of this step, as the flags ask
    auto store_moments = [&](T rho, const T* u) {
      if (!store_m) return;
      rho_c[n] = rho;
      sfor<D>([&](auto A) { u_c[(size_t)decltype(A)::value * ld + n] = u[decltype(A)::value]; });
      if (resid) {
        const double dr = (double)rho - (double)rho_p[n];
        acc[0] += dr * dr;
        bool bad = !isfinite((double)rho);
        sfor<D>([&](auto A) {
          constexpr int ax = decltype(A)::value;
          const double du = (double)u[ax] - (double)u_p[(size_t)ax * ld + n];
          acc[1 + ax] += du * du;
          bad = bad || !isfinite((double)u[ax]);
        });
        acc[4] += bad ? 1.0 : 0.0;
      }
    };
    T fs[Q];
    if constexpr (std::is_same<St, StF64>::value) {
      // FP64 storage: the parity kernel's arithmetic in its order, without
      // contraction, so this scheme and the two-array scheme give the same bits
      double rho, u[3], Fv[3];
      ref_moments_force<L>(a, n, f, rho, u, Fv);
      store_moments(rho, u);
      ref_collide<L, OP>(a, f, rho, u, Fv, fs);
    } else {
      // the uniform force and the per-node field are read in separate branches:
      // a select between their addresses would load the uniform value per lane
      auto get_force = [&](T* Fv) {
        if (a.force_kind == 1) {
          sfor<D>([&](auto A) { Fv[decltype(A)::value] = T(a.F[decltype(A)::value]); });
        } else if (a.force_kind) {
          sfor<D>([&](auto A) {
            constexpr int ax = decltype(A)::value;
            Fv[ax] = T(global_ptr(a.src_field)[(size_t)ax * ld + n]);
          });
        }
        return a.force_kind != 0;
      };
      auto& moments = store_moments;
#include "pair_collide.inc"
    }
    // ---- scatter
    const uint32_t own = n * (uint32_t)sizeof(S);
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      constexpr int jb = k::qbar(j);
      const S sv = static_cast<S>(fs[j]);
      if (!ODD) {
        bstore<S>(sv, rs, own, jb * pb);
      } else {
        const int32_t d = p.ox[1 + k::cx(j)] + p.oy[1 + k::cy(j)] + p.oz[1 + k::cz(j)];
        const uint32_t nb = (uint32_t)((int32_t)n + d) * (uint32_t)sizeof(S) + j * pb;
        const uint32_t mine = own + jb * pb;
        bstore<S>(sv, rs, opaque((fm >> j & 1u) ? mine : nb), 0u);
      }
    });
  }
}

// class word of node n of a case (kSolidBit past its last node)
__device__ __forceinline__ uint32_t node_class(const KArgs& a, uint32_t n) {
  return n < (uint32_t)a.N ? global_ptr(a.cls)[n] : kSolidBit;
}

template <class L, class St, int OP, bool ODD, bool MOM>
__device__ __forceinline__ void aa_node(const KArgs& a, uint32_t n, unsigned long long step,
                                        bool fresh, double* acc) {
  aa_node_cls<L, St, OP, ODD, MOM>(a, n, node_class(a, n), step, fresh, acc);
}

// AA-pattern fused step over one case. MOM: this launch may store moments /
// residual partials (the last step of a batch, or steps with residuals); the
// streaming steps in between use the MOM=false instantiation, which is
// compiled for full occupancy.
template <class L, class St, int OP, bool ODD, bool MOM>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(MOM ? LBF_WAVES_MOM
                                       : (sizeof(typename St::T) == 8 && (L::Q > 19 || L::Q == 15)) ? 2
                                       : (sizeof(typename St::T) == 8 && OP == 0) ? 3
                                       : (sizeof(typename St::T) == 8 || L::Q > 19) ? 4
                                       : sizeof(typename St::S) == 2
                                           ? (ODD ? LBF_WAVES_16_ODD : LBF_WAVES_16_EVEN)
                                           : (ODD ? LBF_WAVES_32_ODD : LBF_WAVES_32_EVEN))))
    momentum_step_aa(KArgs a) {
  Ctrl* ctrl = a.ctrl;
  if (ctrl->stopped) return;
  const unsigned long long step = ctrl->step + a.step_offset;
  double acc[5] = {0, 0, 0, 0, 0};
  // one node per thread (the host launches ceil(N / kBlock) workgroups),
  // node blocks in XCD runs (block_index)
  const uint32_t n = block_index(a) * kBlock + threadIdx.x;
  aa_node<L, St, OP, ODD, MOM>(a, n, step, a.fresh != 0, acc);
  if (a.flags & kFlagResiduals) store_residual_partials(acc, a.partials + (long long)blockIdx.x * kMaxCols);
}


// Post-stream populations (f, not delta) into out[q][ld] as f64: the first
// half of the next step, for read-back.
template <class L, class St, bool ODD>
__global__ void __launch_bounds__(kBlock) momentum_readback_aa(KArgs a) {
  using T = typename St::T;
  using k = K<L>;
  constexpr int Q = L::Q;
  const bool fresh = a.fresh != 0;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(a.pop[0], (short)0, (int)0xffffffff, 0x00020000);
  long long nbeg, nend;
  readback_range(a, &nbeg, &nend);
  for (uint32_t n = (uint32_t)nbeg + blockIdx.x * kBlock + threadIdx.x; n < (uint32_t)nend;
       n += gridDim.x * kBlock) {
    const uint32_t cls = a.cls[n];
    if (cls & kSolidBit) {
      readback_solid<Q>(a, n);
      continue;
    }
    NodePos p;
    node_pos<L>(a, n, true, p);
    T f[Q];
    aa_gather<L, St, ODD>(rs, a.plane_bytes, n, fresh, cls & kLinkMask, face_mask<L>(p.on), p, f);
    if (p.on && !fresh && !a.faces_noslip_only)
      aa_faces<L, T, St::shifted>(a, p.on, T(a.scale), T(a.inv_scale), f);
    double fd[Q];
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      fd[j] = St::shifted ? (double)f[j] * a.inv_scale + k::w(j) : (double)f[j];
    });
    readback_store<Q>(a, n, fd);
  }
}

// f = f_eq (or delta = delta_eq scaled) into the natural slots F[j][x]
template <class L, class St>
__global__ void __launch_bounds__(kBlock) momentum_init_aa(KArgs a) {
  using T = typename St::T;
  using S = typename St::S;
  constexpr int Q = L::Q, D = L::D;
  const int cur = (int)((a.ctrl->step + 1) & 1);
  S* F = (S*)a.pop[0];
  const T* rho = (const T*)a.rho[cur];
  const T* uu = (const T*)a.u[cur];
  for (long long n = (long long)blockIdx.x * kBlock + threadIdx.x; n < a.N;
       n += (long long)gridDim.x * kBlock) {
    T feq[Q];
    if (a.cls[n] & kSolidBit) {
      sfor<Q>([&](auto J) { feq[decltype(J)::value] = T(0); });
    } else {
      T u[3];
      sfor<D>([&](auto A) { u[decltype(A)::value] = uu[decltype(A)::value * a.ld + n]; });
      if constexpr (St::shifted)
        equilibrium_shifted<L>(rho[n] - T(1), rho[n], u, feq);
      else
        equilibrium<L>(rho[n], u, feq);
    }
    sfor<Q>([&](auto J) {
      constexpr int j = decltype(J)::value;
      T v = feq[j];
      if constexpr (St::shifted) v = v * T(a.scale);
      F[(size_t)j * a.ld + n] = static_cast<S>(v);
    });
  }
}

}  // namespace lbf

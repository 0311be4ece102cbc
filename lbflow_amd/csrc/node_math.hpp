// node_math.hpp — per-node lattice-Boltzmann arithmetic in registers.
//
// "Plain" form (storage holds f): every expression keeps the association
// order of the reference's per-node functions (src/kernel.rs, cited per
// function) so the f64 path reproduces the reference bit for bit when built
// without contraction. Sums over directions skip the terms whose lattice
// coefficient is zero; in IEEE arithmetic adding those (+-0) terms cannot
// change a sum's value, so the result is the same as the reference's dense
// fold.
//
// "Shifted" form (storage holds delta_i = f_i - w_i, FP32 arithmetic): the
// same operators rewritten in delta so that no large constant cancels in
// single precision: rho = 1 + sum(delta), j = sum(delta c), and
// delta* = delta - omega (delta - delta_eq) with
// delta_eq = w (drho + rho (3 u.c + 4.5 (u.c)^2 - 1.5 u.u)).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "lattice.hpp"

namespace lbf {

template <class F, int... I>
__host__ __device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
// compile-time loop: f(integral_constant<int, i>) for i in [0, N)
template <int N, class F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

template <class L>
struct K {
  using Dv = Derived<L>;
  static constexpr int Q = L::Q, D = L::D, NQF = L::NQF;
  static constexpr int cx(int i) { return L::c[i][0]; }
  static constexpr int cy(int i) { return L::c[i][1]; }
  static constexpr int cz(int i) { return L::c[i][2]; }
  static constexpr int ca(int i, int a) { return L::c[i][a]; }
  static constexpr int qbar(int i) { return Dv::qbar(i); }
  static constexpr double w(int i) { return Dv::w(i); }
  static constexpr bool any_c(int i) { return ca(i, 0) != 0 || ca(i, 1) != 0 || (D == 3 && ca(i, 2) != 0); }
};

// u . c_i summed over nonzero components in axis order (kernel.rs:12-17)
template <class L, int I, class T>
__device__ __forceinline__ T udotc(const T* u) {
  using k = K<L>;
  T s = T(0);
  bool first = true;
  sfor<L::D>([&](auto A) {
    constexpr int a = decltype(A)::value;
    constexpr int c = k::ca(I, a);
    if constexpr (c != 0) {
      T t = (c > 0) ? u[a] : -u[a];
      if (first) {
        s = t;
        first = false;
      } else {
        s = s + t;
      }
    }
  });
  return s;
}

template <class L, class T>
__device__ __forceinline__ T udotu(const T* u) {
  T s = u[0] * u[0];
  sfor<L::D - 1>([&](auto A) {
    constexpr int a = decltype(A)::value + 1;
    s = s + u[a] * u[a];
  });
  return s;
}

// ---------------------------------------------------------------- plain form

// equilibrium, src/kernel.rs:3-26
template <class L, class T>
__device__ __forceinline__ void equilibrium(T value, const T* u, T* feq) {
  using k = K<L>;
  const T udu = udotu<L>(u);
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    const T udc = udotc<L, i>(u);
    T t = T(1.0) + udc * T(3.0);
    t = t + T(0.5) * udc * udc * T(9.0);
    t = t - T(0.5) * udu * T(3.0);
    feq[i] = T(k::w(i)) * value * t;
  });
}

// rho = sum f, u = sum f c / rho (+ F/(2 rho)), src/momentum/node.rs:418-457
template <class L, class T>
__device__ __forceinline__ void moments(const T* f, T& rho, T* u) {
  using k = K<L>;
  T r = f[0];
  sfor<L::Q - 1>([&](auto I) {
    constexpr int i = decltype(I)::value + 1;
    r = r + f[i];
  });
  rho = r;
  sfor<L::D>([&](auto A) {
    constexpr int a = decltype(A)::value;
    T s = T(0);
    bool first = true;
    sfor<L::Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      constexpr int c = k::ca(i, a);
      if constexpr (c != 0) {
        T t = (c > 0) ? f[i] : -f[i];
        if (first) {
          s = t;
          first = false;
        } else {
          s = s + t;
        }
      }
    });
    u[a] = s / r;
  });
}

// BGK, src/kernel.rs:28-42 (omega = 1/tau, omega' = 1 - omega precomputed)
template <class L, class T>
__device__ __forceinline__ void bgk(const T* f, const T* feq, T omega, T omega_p, T* fs) {
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    fs[i] = omega_p * f[i] + omega * feq[i];
  });
}

// TRT, src/kernel.rs:44-65
template <class L, class T>
__device__ __forceinline__ void trt(const T* f, const T* feq, T wp, T wm, T* fs) {
  using k = K<L>;
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr int ib = k::qbar(i);
    const T fp = T(0.5) * (f[i] + f[ib]);
    const T fep = T(0.5) * (feq[i] + feq[ib]);
    const T fm = T(0.5) * (f[i] - f[ib]);
    const T fem = T(0.5) * (feq[i] - feq[ib]);
    fs[i] = f[i] - wp * (fp - fep) - wm * (fm - fem);
  });
}

// sum_i M[k][i] x[i] over nonzero entries in index order
template <class L, int Krow, class T>
__device__ __forceinline__ T mrt_dot(const T* x) {
  using Dv = Derived<L>;
  T s = T(0);
  bool first = true;
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr double m = Dv::M(Krow, i);
    if constexpr (m != 0.0) {
      T t;
      if constexpr (m == 1.0)
        t = x[i];
      else if constexpr (m == -1.0)
        t = -x[i];
      else
        t = T(m) * x[i];
      if (first) {
        s = t;
        first = false;
      } else {
        s = s + t;
      }
    }
  });
  return s;
}

template <class L, int Irow, class T>
__device__ __forceinline__ T mrt_inv_dot(const T* ms) {
  using Dv = Derived<L>;
  T s = T(0);
  bool first = true;
  sfor<L::Q>([&](auto Kc) {
    constexpr int kk = decltype(Kc)::value;
    constexpr double m = Dv::Minv(Irow, kk);
    if constexpr (m != 0.0) {
      T t = T(m) * ms[kk];
      if (first) {
        s = t;
        first = false;
      } else {
        s = s + t;
      }
    }
  });
  return s;
}

// closed-form equilibrium moments: d2q9.rs:247-264, d3q15.rs:365-396,
// d3q19.rs:550-587 (the `velocity[2] + velocity[2]` terms of d3q19.rs:575,577
// are reproduced); D3Q27 has none and uses M f_eq (kernel.rs:116-126).
template <class L, class T>
__device__ __forceinline__ void meq_closed(T rho, const T* u, const T* feq, T* m) {
  if constexpr (L::id == 1) {
    const T ux = u[0], uy = u[1];
    m[0] = rho;
    m[1] = rho - T(3.0) * rho * (ux * ux + uy * uy);
    m[2] = T(9.0) * rho * ux * ux * uy * uy - T(3.0) * rho * (ux * ux + uy * uy) + rho;
    m[3] = rho * ux;
    m[4] = T(3.0) * rho * ux * ux * ux - rho * ux;
    m[5] = rho * uy;
    m[6] = T(3.0) * rho * uy * uy * uy - rho * uy;
    m[7] = rho * (ux * ux + uy * uy);
    m[8] = rho * ux * uy;
  } else if constexpr (L::id == 2) {
    const T ux = u[0], uy = u[1], uz = u[2];
    const T u2 = ux * ux + uy * uy + uz * uz;
    m[0] = rho;
    m[1] = -rho + rho * u2;
    m[2] = rho - T(5.0) * rho * u2;
    m[3] = rho * ux;
    m[4] = T(-7.0 / 3.0) * rho * ux;
    m[5] = rho * uy;
    m[6] = T(-7.0 / 3.0) * rho * uy;
    m[7] = rho * uz;
    m[8] = T(-7.0 / 3.0) * rho * uz;
    m[9] = T(2.0) * rho * ux * ux - rho * (uy * uy + uz * uz);
    m[10] = rho * (uy * uy - uz * uz);
    m[11] = rho * ux * uy;
    m[12] = rho * uy * uz;
    m[13] = rho * ux * uz;
    m[14] = T(0);
  } else if constexpr (L::id == 3) {
    const T ux = u[0], uy = u[1], uz = u[2];
    const T u2 = ux * ux + uy * uy + uz * uz;
    m[0] = rho;
    m[1] = T(-11.0) * rho + T(19.0) * rho * u2;
    m[2] = T(3.0) * rho - T(5.5) * rho * u2;
    m[3] = rho * ux;
    m[4] = T(-2.0 / 3.0) * rho * ux;
    m[5] = rho * uy;
    m[6] = T(-2.0 / 3.0) * rho * uy;
    m[7] = rho * uz;
    m[8] = T(-2.0 / 3.0) * rho * uz;
    m[9] = T(2.0) * rho * ux * ux - rho * (uy * uy + uz + uz);
    m[10] = T(-0.5) * (T(2.0) * rho * ux * ux - rho * (uy * uy + uz + uz));
    m[11] = rho * (uy * uy - uz * uz);
    m[12] = T(-0.5) * rho * (uy * uy - uz * uz);
    m[13] = rho * ux * uy;
    m[14] = rho * uy * uz;
    m[15] = rho * ux * uz;
    m[16] = T(0);
    m[17] = T(0);
    m[18] = T(0);
  } else {
    sfor<L::Q>([&](auto Kr) {
      constexpr int kk = decltype(Kr)::value;
      m[kk] = mrt_dot<L, kk>(feq);
    });
  }
}

// MRT, src/kernel.rs:67-120 (s[k]*DELTA_T == s[k])
template <class L, class T>
__device__ __forceinline__ void mrt(T value, const T* u, const T* f, const T* feq, const T* s,
                                    T* fs) {
  T m[L::Q], meq[L::Q], ms[L::Q];
  sfor<L::Q>([&](auto Kr) {
    constexpr int kk = decltype(Kr)::value;
    m[kk] = mrt_dot<L, kk>(f);
  });
  meq_closed<L>(value, u, feq, meq);
  sfor<L::Q>([&](auto Kr) {
    constexpr int kk = decltype(Kr)::value;
    ms[kk] = m[kk] - s[kk] * (m[kk] - meq[kk]);
  });
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    fs[i] = mrt_inv_dot<L, i>(ms);
  });
}

// Guo source term, src/kernel.rs:122-154 (coeff_b = 1 - 0.5/tau precomputed)
template <class L, class T>
__device__ __forceinline__ void add_guo(const T* u, const T* F, T coeff_b, T* fs) {
  using k = K<L>;
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    const T udc = udotc<L, i>(u);
    T acc = T(0);
    sfor<L::D>([&](auto A) {
      constexpr int a = decltype(A)::value;
      constexpr T c = T(k::ca(i, a));
      const T term = (T(3.0) * c - T(3.0) * u[a] + T(9.0) * udc * c) * F[a];
      if constexpr (a == 0)
        acc = term;
      else
        acc = acc + term;
    });
    fs[i] += coeff_b * T(k::w(i)) * acc;
  });
}

// D2Q9 Zou-He, src/velocity_set/d2q9.rs:165-245. `known` uses LBF_KNOWN_*
// bits; returns false for the combinations on which the reference panics.
template <class T>
__device__ __forceinline__ bool zou_he_d2q9(int face, T* f, unsigned known, T density, T vx,
                                            T vy) {
  const bool hd = known & 1u, hx = known & 2u, hy = known & 4u;
  T ux = vx, uy = vy, rho;
  switch (face) {
    case kWest:
      if (!hd && hx && hy)
        rho = T(1.0) / (T(1.0) - ux) * (f[0] + f[2] + f[4] + T(2.0) * (f[3] + f[6] + f[7]));
      else if (hd && !hx && hy) {
        rho = density;
        ux = T(1.0) - T(1.0) / rho * (f[0] + f[2] + f[4] + T(2.0) * (f[3] + f[6] + f[7]));
      } else
        return false;
      f[1] = f[3] + T(2.0 / 3.0) * rho * ux;
      f[5] = f[7] - T(0.5) * (f[2] - f[4]) + T(0.5) * rho * uy + T(1.0 / 6.0) * rho * ux;
      f[8] = f[6] + T(0.5) * (f[2] - f[4]) - T(0.5) * rho * uy + T(1.0 / 6.0) * rho * ux;
      return true;
    case kEast:
      if (!hd && hx && hy)
        rho = T(1.0) / (T(1.0) + ux) * (f[0] + f[2] + f[4] + T(2.0) * (f[1] + f[5] + f[8]));
      else if (hd && !hx && hy) {
        rho = density;
        ux = T(1.0) / rho * (f[0] + f[2] + f[4] + T(2.0) * (f[1] + f[5] + f[8])) - T(1.0);
      } else
        return false;
      f[3] = f[1] - T(2.0 / 3.0) * rho * ux;
      f[7] = f[5] + T(0.5) * (f[2] - f[4]) - T(0.5) * rho * uy - T(1.0 / 6.0) * rho * ux;
      f[6] = f[8] - T(0.5) * (f[2] - f[4]) + T(0.5) * rho * uy - T(1.0 / 6.0) * rho * ux;
      return true;
    case kSouth:
      if (!hd && hx && hy)
        rho = T(1.0) / (T(1.0) - uy) * (f[0] + f[1] + f[3] + T(2.0) * (f[4] + f[7] + f[8]));
      else if (hd && hx && !hy) {
        rho = density;
        uy = T(1.0) - T(1.0) / rho * (f[0] + f[1] + f[3] + T(2.0) * (f[4] + f[7] + f[8]));
      } else
        return false;
      f[2] = f[4] + T(2.0 / 3.0) * rho * uy;
      f[5] = f[7] - T(0.5) * (f[1] - f[3]) + T(0.5) * rho * ux + T(1.0 / 6.0) * rho * uy;
      f[6] = f[8] + T(0.5) * (f[1] - f[3]) - T(0.5) * rho * ux + T(1.0 / 6.0) * rho * uy;
      return true;
    case kNorth:
      if (!hd && hx && hy)
        rho = T(1.0) / (T(1.0) + uy) * (f[0] + f[1] + f[3] + T(2.0) * (f[2] + f[5] + f[6]));
      else if (hd && hx && !hy) {
        rho = density;
        uy = T(1.0) / rho * (f[0] + f[1] + f[3] + T(2.0) * (f[2] + f[5] + f[6])) - T(1.0);
      } else
        return false;
      f[4] = f[2] - T(2.0 / 3.0) * rho * uy;
      f[7] = f[5] + T(0.5) * (f[1] - f[3]) - T(0.5) * rho * ux - T(1.0 / 6.0) * rho * uy;
      f[8] = f[6] - T(0.5) * (f[1] - f[3]) + T(0.5) * rho * ux - T(1.0 / 6.0) * rho * uy;
      return true;
  }
  return false;
}

// Zou-He closure for any lattice and face (this engine's own; the reference
// has one only for D2Q9, src/velocity_set/d2q9.rs:165-245, and panics for
// the 3-D sets). Face Fc with inward normal sign s along axis a:
//  - S0 = sum of f_i with c_ia = 0, So = sum of f_i with s c_ia < 0;
//  - velocity given: rho = (S0 + 2 So) / (1 - s u_a); density given:
//    s u_a = 1 - (S0 + 2 So) / rho (tangential velocity given);
//  - unknowns (s c_ia > 0): f_i = f_ibar + 6 w_i rho (c_i . u), which sets
//    rho and the normal momentum (sum of w_i over the unknowns is 1/6);
//  - tangential axis t: lambda_t = (rho u_t - sum_i c_it f_i) / sum_U c_it^2
//    and f_i += c_it lambda_t over the unknowns, which sets the tangential
//    momentum and leaves rho and the other components unchanged (the unknown
//    set is symmetric in every tangential axis).
// On D2Q9 this reproduces the reference's closed forms.
template <class L, int Fc>
struct ZouHeFace {
  static constexpr int a = Fc / 2;
  static constexpr int s = (Fc % 2 == 0) ? 1 : -1;
  static constexpr int cn(int i) { return L::c[i][a] * s; }
  static constexpr int den(int t) {
    int n = 0;
    for (int i = 0; i < L::Q; i++)
      if (cn(i) > 0) n += L::c[i][t] * L::c[i][t];
    return n;
  }
};

template <class L, class T, int Fc>
__device__ __forceinline__ void zou_he_general(T* f, unsigned known, T density, const T* vel) {
  using k = K<L>;
  using Z = ZouHeFace<L, Fc>;
  constexpr int a = Z::a, D = L::D;
  T s0 = T(0), so = T(0);
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (Z::cn(i) == 0) s0 = s0 + f[i];
    if constexpr (Z::cn(i) < 0) so = so + f[i];
  });
  T u[3] = {vel[0], vel[1], vel[2]};
  T rho;
  if (known & 1u) {
    rho = density;
    const T un = T(1) - (s0 + T(2) * so) / rho;
    u[a] = Z::s > 0 ? un : -un;
  } else {
    const T un = Z::s > 0 ? u[a] : -u[a];
    rho = (s0 + T(2) * so) / (T(1) - un);
  }
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (Z::cn(i) > 0) f[i] = f[k::qbar(i)] + T(6) * T(k::w(i)) * rho * udotc<L, i>(u);
  });
  T lam[3] = {T(0), T(0), T(0)};
  sfor<D>([&](auto TI) {
    constexpr int t = decltype(TI)::value;
    if constexpr (t != a) {
      T m = T(0);
      sfor<L::Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        if constexpr (L::c[i][t] > 0) m = m + f[i];
        if constexpr (L::c[i][t] < 0) m = m - f[i];
      });
      lam[t] = (rho * u[t] - m) / T(Z::den(t));
    }
  });
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (Z::cn(i) > 0) {
      sfor<D>([&](auto TI) {
        constexpr int t = decltype(TI)::value;
        if constexpr (t != a && L::c[i][t] != 0) f[i] = f[i] + T(L::c[i][t]) * lam[t];
      });
    }
  });
}

// ------------------------------------------------------------- shifted form
// delta-form equilibrium: delta_eq_i = w_i (drho + rho (3 u.c + 4.5 (u.c)^2 - 1.5 u.u))
template <class L>
__device__ __forceinline__ void equilibrium_shifted(float drho, float rho, const float* u,
                                                    float* deq) {
  using k = K<L>;
  const float udu = udotu<L>(u);
  sfor<L::Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    const float udc = udotc<L, i>(u);
    const float t = udc * (3.0f + 4.5f * udc) - 1.5f * udu;
    deq[i] = float(k::w(i)) * (drho + rho * t);
  });
}

}  // namespace lbf

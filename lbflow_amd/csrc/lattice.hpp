// lattice.hpp — compile-time velocity sets for the device kernels.
//
// Directions, weights, opposites and face sets follow the velocity-set tables
// of gilberto-ribeiro/lbflow (src/velocity_set/d2q9.rs:8-42,
// d3q15.rs:7-57, d3q19.rs:7-67, d3q27.rs:7-84); the numbering has to be the
// same because the ABI exchanges populations in that order. Opposites and
// face sets are derived here from the direction vectors rather than tabled:
//   qbar(i)      = the j with c_j = -c_i
//   qfaces(face) = directions with c . n_out = +1, in increasing index order
//   normal(face) = the direction equal to the inward unit normal
// The MRT basis rows are the polynomials of c whose values are the integer
// matrices of the reference (d2q9.rs:44-54, d3q15.rs:59-..., d3q19.rs:69-...),
// and M^-1 = M^T diag(1/|row|^2) because the rows are orthogonal.
#pragma once

#include <cstdint>

namespace lbf {

constexpr int kWest = 0, kEast = 1, kSouth = 2, kNorth = 3, kBottom = 4, kTop = 5;

struct D2Q9 {
  static constexpr int id = 1, D = 2, Q = 9, NQF = 3, NMRT = 9;
  static constexpr int c[Q][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},   {-1, 0, 0}, {0, -1, 0},
                                  {1, 1, 0},  {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
  // weights as exact fractions num/den (4/9, 1/9, 1/36)
  static constexpr int wnum[Q] = {4, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int wden[Q] = {9, 9, 9, 9, 9, 36, 36, 36, 36};
  static constexpr bool has_mrt = true;
  static constexpr bool has_zou_he = true;
  static constexpr double mrt_row(int k, int cx, int cy, int) {
    double x = cx, y = cy, c2 = x * x + y * y;
    switch (k) {
      case 0: return 1;
      case 1: return 3 * c2 - 4;
      case 2: return (9 * c2 * c2 - 21 * c2 + 8) / 2;
      case 3: return x;
      case 4: return (3 * c2 - 5) * x;
      case 5: return y;
      case 6: return (3 * c2 - 5) * y;
      case 7: return x * x - y * y;
      default: return x * y;
    }
  }
};

struct D3Q15 {
  static constexpr int id = 2, D = 3, Q = 15, NQF = 5, NMRT = 15;
  static constexpr int c[Q][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                                  {0, 0, 1},   {0, 0, -1},  {1, 1, 1},  {-1, -1, -1}, {1, 1, -1},
                                  {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};
  static constexpr int wnum[Q] = {2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int wden[Q] = {9, 9, 9, 9, 9, 9, 9, 72, 72, 72, 72, 72, 72, 72, 72};
  static constexpr bool has_mrt = true;
  static constexpr bool has_zou_he = false;
  static constexpr double mrt_row(int k, int cx, int cy, int cz) {
    double x = cx, y = cy, z = cz, c2 = x * x + y * y + z * z;
    switch (k) {
      case 0: return 1;
      case 1: return c2 - 2;
      case 2: return (15 * c2 * c2 - 55 * c2 + 32) / 2;
      case 3: return x;
      case 4: return (5 * c2 - 13) * x / 2;
      case 5: return y;
      case 6: return (5 * c2 - 13) * y / 2;
      case 7: return z;
      case 8: return (5 * c2 - 13) * z / 2;
      case 9: return 3 * x * x - c2;
      case 10: return y * y - z * z;
      case 11: return x * y;
      case 12: return y * z;
      case 13: return x * z;
      default: return x * y * z;
    }
  }
};

struct D3Q19 {
  static constexpr int id = 3, D = 3, Q = 19, NQF = 5, NMRT = 19;
  static constexpr int c[Q][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                                  {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                                  {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
                                  {1, 0, -1}, {-1, 0, 1},  {0, 1, -1}, {0, -1, 1}};
  static constexpr int wnum[Q] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int wden[Q] = {3, 18, 18, 18, 18, 18, 18, 36, 36, 36,
                                  36, 36, 36, 36, 36, 36, 36, 36, 36};
  static constexpr bool has_mrt = true;
  static constexpr bool has_zou_he = false;
  static constexpr double mrt_row(int k, int cx, int cy, int cz) {
    double x = cx, y = cy, z = cz, c2 = x * x + y * y + z * z;
    switch (k) {
      case 0: return 1;
      case 1: return 19 * c2 - 30;
      case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
      case 3: return x;
      case 4: return (5 * c2 - 9) * x;
      case 5: return y;
      case 6: return (5 * c2 - 9) * y;
      case 7: return z;
      case 8: return (5 * c2 - 9) * z;
      case 9: return 3 * x * x - c2;
      case 10: return (3 * c2 - 5) * (3 * x * x - c2);
      case 11: return y * y - z * z;
      case 12: return (3 * c2 - 5) * (y * y - z * z);
      case 13: return x * y;
      case 14: return y * z;
      case 15: return x * z;
      case 16: return (y * y - z * z) * x;
      case 17: return (z * z - x * x) * y;
      default: return (x * x - y * y) * z;
    }
  }
};

// D3Q27 directions/weights (d3q27.rs:7-65). The reference has no MRT basis
// for it (MRT_MATRIX is all ones, the inverse is todo!()), so MRT here uses
// this engine's own orthogonal basis, see d3q27_basis below.
struct D3Q27 {
  static constexpr int id = 4, D = 3, Q = 27, NQF = 9, NMRT = 27;
  static constexpr int c[Q][3] = {
      {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
      {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},  {-1, 0, -1}, {0, 1, 1},   {0, -1, -1}, {1, -1, 0},
      {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1}, {0, 1, -1},  {0, -1, 1},  {1, 1, 1},   {-1, -1, -1},
      {1, 1, -1},  {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};
  static constexpr int wnum[Q] = {8, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1,
                                  1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int wden[Q] = {27, 27, 27, 27, 27, 27, 27, 54, 54, 54, 54, 54, 54, 54,
                                  54, 54, 54, 54, 54, 216, 216, 216, 216, 216, 216, 216, 216};
  static constexpr bool has_mrt = true;
  static constexpr bool has_zou_he = false;
  // Orthogonal basis of tensor-product Hermite-like polynomials in each axis:
  // p0(a)=1, p1(a)=a, p2(a)=3a^2-2 (orthogonal over a in {-1,0,1} with equal
  // weights). Row k = p_{i}(cx) p_{j}(cy) p_{l}(cz) with (i,j,l) enumerated by
  // total order then lexicographically; row 0 is density, rows 1..3 momentum.
  static constexpr double p(int o, double a) { return o == 0 ? 1.0 : (o == 1 ? a : 3 * a * a - 2); }
  static constexpr int ord[27][3] = {
      {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
      {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1},
      {1, 0, 2}, {0, 1, 2}, {1, 1, 1}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1},
      {1, 2, 1}, {1, 1, 2}, {2, 2, 1}, {2, 1, 2}, {1, 2, 2}, {2, 2, 2}};
  static constexpr double mrt_row(int k, int cx, int cy, int cz) {
    return p(ord[k][0], cx) * p(ord[k][1], cy) * p(ord[k][2], cz);
  }
};

// ---------------------------------------------------------------- derived
template <class L>
struct Derived {
  static constexpr int Q = L::Q;
  static constexpr int qbar(int i) {
    for (int j = 0; j < Q; j++)
      if (L::c[j][0] == -L::c[i][0] && L::c[j][1] == -L::c[i][1] && L::c[j][2] == -L::c[i][2])
        return j;
    return -1;
  }
  // outward normal component of a face: West -x, East +x, South -y, ...
  static constexpr int out_axis(int f) { return f / 2; }
  static constexpr int out_sign(int f) { return (f % 2) ? 1 : -1; }
  static constexpr bool in_face(int f, int i) {
    return L::c[i][out_axis(f)] == out_sign(f);
  }
  static constexpr int qface(int f, int k) {
    int n = 0;
    for (int i = 0; i < Q; i++)
      if (in_face(f, i)) {
        if (n == k) return i;
        n++;
      }
    return -1;
  }
  static constexpr int normal(int f) {  // inward unit normal direction
    for (int i = 0; i < Q; i++) {
      bool ok = true;
      for (int a = 0; a < 3; a++) {
        int want = (a == out_axis(f)) ? -out_sign(f) : 0;
        if (L::c[i][a] != want) ok = false;
      }
      if (ok) return i;
    }
    return -1;
  }
  static constexpr double w(int i) { return double(L::wnum[i]) / double(L::wden[i]); }
  static constexpr double M(int k, int i) { return L::mrt_row(k, L::c[i][0], L::c[i][1], L::c[i][2]); }
  static constexpr double norm(int k) {
    double s = 0;
    for (int i = 0; i < Q; i++) s += M(k, i) * M(k, i);
    return s;
  }
  // one correctly rounded division per entry
  static constexpr double Minv(int i, int k) { return M(k, i) / norm(k); }
};

// Table objects usable from device code (constexpr arrays materialised once).
template <class L>
struct Tab {
  static constexpr int Q = L::Q;
  struct Arr {
    int qbar[L::Q];
    int qf[6][L::NQF];
    int normal[6];
    double w[L::Q];
  };
  static constexpr Arr make() {
    Arr a{};
    for (int i = 0; i < Q; i++) {
      a.qbar[i] = Derived<L>::qbar(i);
      a.w[i] = Derived<L>::w(i);
    }
    for (int f = 0; f < 2 * L::D; f++) {
      for (int k = 0; k < L::NQF; k++) a.qf[f][k] = Derived<L>::qface(f, k);
      a.normal[f] = Derived<L>::normal(f);
    }
    return a;
  }
};

}  // namespace lbf

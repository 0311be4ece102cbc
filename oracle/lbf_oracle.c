/*
 * lbf_oracle.c — CPU f64 restatement of lbflow's lattice time step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle: it is linked by the
 * tests, by __graft_entry__.smoke() as the checker, and by bench.py's
 * cpu_baseline leg. The product (lbflow_amd/, include/) never calls it.
 *
 * The reference (gilberto-ribeiro/lbflow, Rust) cannot be compiled here (no
 * rustc/cargo, crates not vendored), so this is a line-by-line restatement of
 * its algorithm in plain C, run with floating-point contraction disabled
 * (-ffp-contract=off, as Rust never fuses multiply-add). Every formula keeps
 * the reference's association order so that results match to the last bit
 * wherever the reference's own order is deterministic. Each function cites
 * the reference file:line it follows.
 *
 * Storage mirrors the reference's per-node vectors: f, f_eq, f_star as
 * [node][q] arrays (src/momentum/node.rs:40-58), g, g_eq, g_star for each
 * scalar (src/passive_scalar/node.rs:40-54). Phases run as separate loops
 * over the fluid nodes with an implicit barrier between them, exactly like
 * the rayon `par_iter` phases of src/momentum/lattice.rs:330-367.
 *
 * Deliberate fixes of reference nondeterminism:
 *  - face conditions run in FACES_2D/FACES_3D order (src/lib.rs:27-29)
 *    instead of HashMap order (src/momentum/bc.rs:30-32);
 *  - residual sums run sequentially in node order (rayon's order varies).
 *
 * Parity status of every function is listed in the test suite header
 * (tests/test_oracle.py); functions without a known-answer test of the
 * reference are pinned by closed forms, invariants and brute force there.
 */
#include "../include/lbflow_hip.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- errors */
static __thread char g_err[512];
static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
const char* lbo_last_error(void) { return g_err; }

/* ------------------------------------------------------- constants.rs:1-21 */
#define CS_2_INV 3.0
#define CS_4_INV 9.0
#define DELTA_T 1.0
#define MIN_ITER 10
#define TOL 1e-7

/* ------------------------------------------------------- velocity sets */
typedef struct vset {
  int id, d, q, nqf;
  int c[27][3];
  double w[27];
  int qbar[27];
  int qfaces[6][9];
  int normal[6]; /* face_normal_directions, src/velocity_set.rs:42-47,66-71,... */
  int has_mrt;
  double M[27][27], Minv[27][27];
} vset;

/* D2Q9: src/velocity_set/d2q9.rs:8-42 */
static const int C_D2Q9[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1},
                                 {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
static const int QBAR_D2Q9[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const int QF_D2Q9[4][3] = {{3, 6, 7}, {1, 5, 8}, {4, 7, 8}, {2, 5, 6}};
/* D3Q15: src/velocity_set/d3q15.rs:7-57 */
static const int C_D3Q15[15][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},
                                   {0, -1, 0}, {0, 0, 1},   {0, 0, -1}, {1, 1, 1},
                                   {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},
                                   {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};
static const int QF_D3Q15[6][5] = {{2, 8, 10, 12, 13}, {1, 7, 9, 11, 14}, {4, 8, 10, 11, 14},
                                   {3, 7, 9, 12, 13},  {6, 8, 9, 12, 14}, {5, 7, 10, 11, 13}};
/* D3Q19: src/velocity_set/d3q19.rs:7-67 */
static const int C_D3Q19[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                                   {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                                   {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
                                   {1, 0, -1}, {-1, 0, 1}, {0, 1, -1}, {0, -1, 1}};
static const int QF_D3Q19[6][5] = {{2, 8, 10, 14, 16}, {1, 7, 9, 13, 15},  {4, 8, 12, 13, 18},
                                   {3, 7, 11, 14, 17}, {6, 10, 12, 15, 17}, {5, 9, 11, 16, 18}};
/* D3Q27: src/velocity_set/d3q27.rs:7-84 */
static const int C_D3Q27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
    {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},  {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0},
    {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1}, {0, 1, -1}, {0, -1, 1},  {1, 1, 1},   {-1, -1, -1},
    {1, 1, -1},  {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
static const int QF_D3Q27[6][9] = {{2, 8, 10, 14, 16, 20, 22, 24, 25},
                                   {1, 7, 9, 13, 15, 19, 21, 23, 26},
                                   {4, 8, 12, 13, 18, 20, 22, 23, 26},
                                   {3, 7, 11, 14, 17, 19, 21, 24, 25},
                                   {6, 10, 12, 15, 17, 20, 21, 24, 26},
                                   {5, 9, 11, 16, 18, 19, 22, 23, 25}};

/* MRT basis rows as polynomials of c. These generate exactly the integer
 * matrices MRT_MATRIX of src/velocity_set/d2q9.rs:44-54,
 * d3q15.rs:59-... and d3q19.rs:69-... (row order kept). */
static double mrt_row(int id, int k, const int* cv) {
  double x = cv[0], y = cv[1], z = (id == LBF_D2Q9) ? 0.0 : cv[2];
  double c2 = x * x + y * y + z * z;
  if (id == LBF_D2Q9) {
    switch (k) {
      case 0: return 1;
      case 1: return 3 * c2 - 4;
      case 2: return (9 * c2 * c2 - 21 * c2 + 8) / 2;
      case 3: return x;
      case 4: return (3 * c2 - 5) * x;
      case 5: return y;
      case 6: return (3 * c2 - 5) * y;
      case 7: return x * x - y * y;
      case 8: return x * y;
    }
  } else if (id == LBF_D3Q15) {
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
      case 14: return x * y * z;
    }
  } else if (id == LBF_D3Q19) {
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
      case 18: return (x * x - y * y) * z;
    }
  } else if (id == LBF_D3Q27) {
    /* The reference has no D3Q27 MRT basis (MRT_MATRIX is an all-ones
     * placeholder, d3q27.rs:86, and the inverse is todo!(),
     * velocity_set.rs:168). This is the engine's own basis: rows are the
     * tensor products p_a(cx) p_b(cy) p_c(cz) of p0 = 1, p1 = a, p2 = 3a^2 - 2
     * (orthogonal over a in {-1, 0, 1}), (a, b, c) in this order. */
    static const int ORD27[27][3] = {
        {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
        {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1},
        {1, 0, 2}, {0, 1, 2}, {1, 1, 1}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1},
        {1, 2, 1}, {1, 1, 2}, {2, 2, 1}, {2, 1, 2}, {1, 2, 2}, {2, 2, 2}};
    const double a[3] = {x, y, z};
    double r = 1.0;
    for (int ax = 0; ax < 3; ax++) {
      const int o = ORD27[k][ax];
      r *= o == 0 ? 1.0 : (o == 1 ? a[ax] : 3.0 * a[ax] * a[ax] - 2.0);
    }
    return r;
  }
  return 0;
}

static int vset_init(int id, vset* v) {
  memset(v, 0, sizeof *v);
  v->id = id;
  switch (id) {
    case LBF_D2Q9:
      v->d = 2, v->q = 9, v->nqf = 3;
      for (int i = 0; i < 9; i++) {
        v->c[i][0] = C_D2Q9[i][0], v->c[i][1] = C_D2Q9[i][1];
        v->qbar[i] = QBAR_D2Q9[i];
        v->w[i] = i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
      }
      for (int f = 0; f < 4; f++)
        for (int j = 0; j < 3; j++) v->qfaces[f][j] = QF_D2Q9[f][j];
      /* src/velocity_set.rs:66-71: West 1, East 3, South 2, North 4 */
      v->normal[0] = 1, v->normal[1] = 3, v->normal[2] = 2, v->normal[3] = 4;
      v->has_mrt = 1;
      break;
    case LBF_D3Q15:
      v->d = 3, v->q = 15, v->nqf = 5;
      for (int i = 0; i < 15; i++) {
        for (int x = 0; x < 3; x++) v->c[i][x] = C_D3Q15[i][x];
        v->w[i] = i == 0 ? 2.0 / 9.0 : (i < 7 ? 1.0 / 9.0 : 1.0 / 72.0);
      }
      for (int f = 0; f < 6; f++)
        for (int j = 0; j < 5; j++) v->qfaces[f][j] = QF_D3Q15[f][j];
      v->has_mrt = 1;
      break;
    case LBF_D3Q19:
      v->d = 3, v->q = 19, v->nqf = 5;
      for (int i = 0; i < 19; i++) {
        for (int x = 0; x < 3; x++) v->c[i][x] = C_D3Q19[i][x];
        v->w[i] = i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
      }
      for (int f = 0; f < 6; f++)
        for (int j = 0; j < 5; j++) v->qfaces[f][j] = QF_D3Q19[f][j];
      v->has_mrt = 1;
      break;
    case LBF_D3Q27:
      v->d = 3, v->q = 27, v->nqf = 9;
      for (int i = 0; i < 27; i++) {
        for (int x = 0; x < 3; x++) v->c[i][x] = C_D3Q27[i][x];
        v->w[i] = i == 0 ? 8.0 / 27.0 : (i < 7 ? 2.0 / 27.0 : (i < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
      }
      for (int f = 0; f < 6; f++)
        for (int j = 0; j < 9; j++) v->qfaces[f][j] = QF_D3Q27[f][j];
      /* MRT_MATRIX is a placeholder and the inverse is todo!() (d3q27.rs:86,
       * velocity_set.rs:168): MRT is not constructible in the reference; it
       * runs here in the engine's own basis (mrt_row). PARITY UNPINNED by the
       * reference: pinned by mathematics in tests/test_oracle.py (M M^-1 = I,
       * uniform rates equal BGK, conservation, one relaxed mode at a time). */
      v->has_mrt = 1;
      break;
    default:
      return -1;
  }
  if (v->d == 3) {
    /* Q_BAR of every 3-D set pairs 2k-1 <-> 2k (d3q15.rs, d3q19.rs, d3q27.rs) */
    v->qbar[0] = 0;
    for (int i = 1; i < v->q; i++) v->qbar[i] = (i % 2) ? i + 1 : i - 1;
    for (int f = 0; f < 6; f++) v->normal[f] = f + 1; /* velocity_set.rs:95-102 */
  }
  if (v->has_mrt) {
    double norm[27];
    for (int k = 0; k < v->q; k++) {
      norm[k] = 0;
      for (int i = 0; i < v->q; i++) {
        v->M[k][i] = mrt_row(id, k, v->c[i]);
        norm[k] += v->M[k][i] * v->M[k][i];
      }
    }
    /* The basis rows are orthogonal, so M^-1 = M^T diag(1/|row|^2). Every
     * entry is one correctly rounded division of exact integers, which is the
     * same double as the reference's literal fractions (e.g. -5.0/399.0). */
    for (int i = 0; i < v->q; i++)
      for (int k = 0; k < v->q; k++) v->Minv[i][k] = v->M[k][i] / norm[k];
  }
  return 0;
}

/* -------------------------------------------- per-node arithmetic (kernel.rs) */

/* src/kernel.rs:3-26 */
static void k_equilibrium(const vset* v, double value, const double* u, double* feq) {
  double udu = u[0] * u[0];
  for (int x = 1; x < v->d; x++) udu = udu + u[x] * u[x];
  for (int i = 0; i < v->q; i++) {
    double udc = u[0] * (double)v->c[i][0];
    for (int x = 1; x < v->d; x++) udc = udc + u[x] * (double)v->c[i][x];
    feq[i] = v->w[i] * value *
             (1.0 + udc * CS_2_INV + 0.5 * udc * udc * CS_4_INV - 0.5 * udu * CS_2_INV);
  }
}

/* src/kernel.rs:28-42 */
static void k_bgk(const vset* v, const double* f, const double* feq, double tau, double* fs) {
  double omega = DELTA_T / tau;
  double omega_prime = 1.0 - omega;
  for (int i = 0; i < v->q; i++) fs[i] = omega_prime * f[i] + omega * feq[i];
}

/* src/kernel.rs:44-65 */
static void k_trt(const vset* v, const double* f, const double* feq, double wp, double wm,
                  double* fs) {
  for (int i = 0; i < v->q; i++) {
    int ib = v->qbar[i];
    double f_plus = 0.5 * (f[i] + f[ib]);
    double feq_plus = 0.5 * (feq[i] + feq[ib]);
    double f_minus = 0.5 * (f[i] - f[ib]);
    double feq_minus = 0.5 * (feq[i] - feq[ib]);
    fs[i] = f[i] - wp * DELTA_T * (f_plus - feq_plus) - wm * DELTA_T * (f_minus - feq_minus);
  }
}

/* Closed-form equilibrium moments: src/velocity_set/d2q9.rs:247-264,
 * d3q15.rs:365-396, d3q19.rs:550-587 (including the `velocity[2] +
 * velocity[2]` of d3q19.rs:575,577, kept as the reference computes it). */
static void k_meq(int id, double rho, const double* u, double* m) {
  if (id == LBF_D2Q9) {
    double ux = u[0], uy = u[1];
    m[0] = rho;
    m[1] = rho - 3.0 * rho * (ux * ux + uy * uy);
    m[2] = 9.0 * rho * ux * ux * uy * uy - 3.0 * rho * (ux * ux + uy * uy) + rho;
    m[3] = rho * ux;
    m[4] = 3.0 * rho * ux * ux * ux - rho * ux;
    m[5] = rho * uy;
    m[6] = 3.0 * rho * uy * uy * uy - rho * uy;
    m[7] = rho * (ux * ux + uy * uy);
    m[8] = rho * ux * uy;
  } else if (id == LBF_D3Q15) {
    double ux = u[0], uy = u[1], uz = u[2];
    double u2 = ux * ux + uy * uy + uz * uz;
    m[0] = rho;
    m[1] = -rho + rho * u2;
    m[2] = rho - 5.0 * rho * u2;
    m[3] = rho * ux;
    m[4] = -7.0 / 3.0 * rho * ux;
    m[5] = rho * uy;
    m[6] = -7.0 / 3.0 * rho * uy;
    m[7] = rho * uz;
    m[8] = -7.0 / 3.0 * rho * uz;
    m[9] = 2.0 * rho * ux * ux - rho * (uy * uy + uz * uz);
    m[10] = rho * (uy * uy - uz * uz);
    m[11] = rho * ux * uy;
    m[12] = rho * uy * uz;
    m[13] = rho * ux * uz;
    m[14] = 0.0;
  } else if (id == LBF_D3Q19) {
    double ux = u[0], uy = u[1], uz = u[2];
    double u2 = ux * ux + uy * uy + uz * uz;
    m[0] = rho;
    m[1] = -11.0 * rho + 19.0 * rho * u2;
    m[2] = 3.0 * rho - 5.5 * rho * u2;
    m[3] = rho * ux;
    m[4] = -2.0 / 3.0 * rho * ux;
    m[5] = rho * uy;
    m[6] = -2.0 / 3.0 * rho * uy;
    m[7] = rho * uz;
    m[8] = -2.0 / 3.0 * rho * uz;
    m[9] = 2.0 * rho * ux * ux - rho * (uy * uy + uz + uz);
    m[10] = -0.5 * (2.0 * rho * ux * ux - rho * (uy * uy + uz + uz));
    m[11] = rho * (uy * uy - uz * uz);
    m[12] = -0.5 * rho * (uy * uy - uz * uz);
    m[13] = rho * ux * uy;
    m[14] = rho * uy * uz;
    m[15] = rho * ux * uz;
    m[16] = 0.0;
    m[17] = 0.0;
    m[18] = 0.0;
  }
}

/* src/kernel.rs:67-120: m_eq from the closed form where the set has one
 * (D2Q9, D3Q15, D3Q19: velocity_set.rs:79-81,110-112,141-143), otherwise
 * m_eq = M f_eq (kernel.rs:92-105, the `None` branch; D3Q27 with the
 * engine's basis). f_eq is recomputed from (value, u), which gives the bits
 * of the node's stored f_eq (the same function of the same inputs). */
static void k_mrt(const vset* v, double value, const double* u, const double* f,
                  const double* s, double* fs) {
  double m[27], meq[27] = {0}, ms[27];
  int q = v->q;
  for (int k = 0; k < q; k++) {
    double acc = v->M[k][0] * f[0];
    for (int i = 1; i < q; i++) acc = acc + v->M[k][i] * f[i];
    m[k] = acc;
  }
  if (v->id == LBF_D3Q27) {
    double feq[27];
    k_equilibrium(v, value, u, feq);
    for (int k = 0; k < q; k++) {
      double acc = v->M[k][0] * feq[0];
      for (int i = 1; i < q; i++) acc = acc + v->M[k][i] * feq[i];
      meq[k] = acc;
    }
  } else {
    k_meq(v->id, value, u, meq);
  }
  for (int k = 0; k < q; k++) ms[k] = m[k] - s[k] * DELTA_T * (m[k] - meq[k]);
  for (int i = 0; i < q; i++) {
    double acc = v->Minv[i][0] * ms[0];
    for (int k = 1; k < q; k++) acc = acc + v->Minv[i][k] * ms[k];
    fs[i] = acc;
  }
}

/* Guo source, src/kernel.rs:122-154 */
static void k_momentum_source(const vset* v, const double* u, const double* F, double tau,
                              double* S) {
  double coeff_b = 1.0 - 0.5 * DELTA_T / tau;
  for (int i = 0; i < v->q; i++) {
    double udc = u[0] * (double)v->c[i][0];
    for (int x = 1; x < v->d; x++) udc = udc + u[x] * (double)v->c[i][x];
    double acc = 0.0;
    for (int x = 0; x < v->d; x++) {
      double cx = (double)v->c[i][x];
      double term = (CS_2_INV * cx - CS_2_INV * u[x] + CS_4_INV * udc * cx) * F[x];
      acc = x == 0 ? term : acc + term;
    }
    S[i] = coeff_b * v->w[i] * acc;
  }
}

/* Zou-He closure for the 3-D lattices: this build's own (the reference has
 * none). The device restates it in node_math.hpp (zou_he_general) with the
 * same operation order: S0 / So sums, rho or the normal velocity, unknowns
 * f_i = f_ibar + 6 w_i rho (c_i . u), then the tangential corrections
 * lambda_t = (rho u_t - sum_i c_it f_i) / sum_U c_it^2 added as c_it lambda_t. */
static int zou_he_3d_ok(int face, unsigned known, int D) {
  const unsigned bits[3] = {LBF_KNOWN_UX, LBF_KNOWN_UY, LBF_KNOWN_UZ};
  int a = face / 2, tang = 1;
  for (int t = 0; t < D; t++)
    if (t != a && !(known & bits[t])) tang = 0;
  if (!(known & LBF_KNOWN_DENSITY)) return tang && (known & bits[a]);
  return tang && !(known & bits[a]);
}

static void k_zou_he_general(const vset* v, int face, double* f, unsigned known, double density,
                             const double* vel) {
  const int a = face / 2, s = (face % 2 == 0) ? 1 : -1, q = v->q, D = v->d;
  double s0 = 0.0, so = 0.0;
  for (int i = 0; i < q; i++) {
    int cn = v->c[i][a] * s;
    if (cn == 0) s0 = s0 + f[i];
    if (cn < 0) so = so + f[i];
  }
  double u[3] = {vel[0], vel[1], vel[2]}, rho;
  if (known & LBF_KNOWN_DENSITY) {
    rho = density;
    double un = 1.0 - (s0 + 2.0 * so) / rho;
    u[a] = s > 0 ? un : -un;
  } else {
    double un = s > 0 ? u[a] : -u[a];
    rho = (s0 + 2.0 * so) / (1.0 - un);
  }
  for (int i = 0; i < q; i++) {
    if (v->c[i][a] * s <= 0) continue;
    double udc = 0.0;
    int first = 1;
    for (int x = 0; x < D; x++) {
      if (v->c[i][x] == 0) continue;
      double t = v->c[i][x] > 0 ? u[x] : -u[x];
      udc = first ? t : udc + t;
      first = 0;
    }
    f[i] = f[v->qbar[i]] + 6.0 * v->w[i] * rho * udc;
  }
  double lam[3] = {0, 0, 0};
  for (int t = 0; t < D; t++) {
    if (t == a) continue;
    double m = 0.0;
    int den = 0;
    for (int i = 0; i < q; i++) {
      if (v->c[i][t] > 0) m = m + f[i];
      if (v->c[i][t] < 0) m = m - f[i];
      if (v->c[i][a] * s > 0) den += v->c[i][t] * v->c[i][t];
    }
    lam[t] = (rho * u[t] - m) / (double)den;
  }
  for (int i = 0; i < q; i++) {
    if (v->c[i][a] * s <= 0) continue;
    for (int t = 0; t < D; t++)
      if (t != a && v->c[i][t] != 0) f[i] = f[i] + (double)v->c[i][t] * lam[t];
  }
}

/* src/kernel.rs:156-169 */
static void k_scalar_source(const vset* v, double source, double tau_g, double* S) {
  double coeff_b = 1.0 - 0.5 / tau_g;
  for (int i = 0; i < v->q; i++) S[i] = coeff_b * v->w[i] * source;
}

/* D2Q9 Zou-He, src/velocity_set/d2q9.rs:165-245. Returns 0 or -1 where the
 * reference panics (ZOUHE_ERROR_MESSAGE). */
static int k_zou_he_d2q9(int face, double* f, unsigned known, double density, const double* vel) {
  int hd = (known & LBF_KNOWN_DENSITY) != 0, hx = (known & LBF_KNOWN_UX) != 0,
      hy = (known & LBF_KNOWN_UY) != 0;
  double ux = vel[0], uy = vel[1], rho;
  switch (face) {
    case LBF_WEST:
      if (!hd && hx && hy) {
        rho = 1.0 / (1.0 - ux) * (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7]));
      } else if (hd && !hx && hy) {
        rho = density;
        ux = 1.0 - 1.0 / rho * (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7]));
      } else
        return -1;
      f[1] = f[3] + (2.0 / 3.0) * rho * ux;
      f[5] = f[7] - 0.5 * (f[2] - f[4]) + 0.5 * rho * uy + (1.0 / 6.0) * rho * ux;
      f[8] = f[6] + 0.5 * (f[2] - f[4]) - 0.5 * rho * uy + (1.0 / 6.0) * rho * ux;
      return 0;
    case LBF_EAST:
      if (!hd && hx && hy) {
        rho = 1.0 / (1.0 + ux) * (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8]));
      } else if (hd && !hx && hy) {
        rho = density;
        ux = 1.0 / rho * (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) - 1.0;
      } else
        return -1;
      f[3] = f[1] - (2.0 / 3.0) * rho * ux;
      f[7] = f[5] + 0.5 * (f[2] - f[4]) - 0.5 * rho * uy - (1.0 / 6.0) * rho * ux;
      f[6] = f[8] - 0.5 * (f[2] - f[4]) + 0.5 * rho * uy - (1.0 / 6.0) * rho * ux;
      return 0;
    case LBF_SOUTH:
      if (!hd && hx && hy) {
        rho = 1.0 / (1.0 - uy) * (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8]));
      } else if (hd && hx && !hy) {
        rho = density;
        uy = 1.0 - 1.0 / rho * (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8]));
      } else
        return -1;
      f[2] = f[4] + (2.0 / 3.0) * rho * uy;
      f[5] = f[7] - 0.5 * (f[1] - f[3]) + 0.5 * rho * ux + (1.0 / 6.0) * rho * uy;
      f[6] = f[8] + 0.5 * (f[1] - f[3]) - 0.5 * rho * ux + (1.0 / 6.0) * rho * uy;
      return 0;
    case LBF_NORTH:
      if (!hd && hx && hy) {
        rho = 1.0 / (1.0 + uy) * (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6]));
      } else if (hd && hx && !hy) {
        rho = density;
        uy = 1.0 / rho * (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6])) - 1.0;
      } else
        return -1;
      f[4] = f[2] - (2.0 / 3.0) * rho * uy;
      f[7] = f[5] + 0.5 * (f[1] - f[3]) - 0.5 * rho * ux - (1.0 / 6.0) * rho * uy;
      f[8] = f[6] - 0.5 * (f[1] - f[3]) + 0.5 * rho * ux - (1.0 / 6.0) * rho * uy;
      return 0;
  }
  return -1;
}

/* Closed-form velocities (explicit_computation = true), d2q9.rs:158-163 and
 * d3q27.rs:88-118; off the step path (lattice.rs:333 passes false) and kept
 * for the reference's known-answer tests. */
static void k_velocity_explicit(int id, double rho, const double* f, double* u) {
  if (id == LBF_D2Q9) {
    u[0] = (1.0 / rho) * (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]);
    u[1] = (1.0 / rho) * (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]);
  } else if (id == LBF_D3Q27) {
    u[0] = (1.0 / rho) * (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[13] - f[14] + f[15] -
                          f[16] + f[19] - f[20] + f[21] - f[22] + f[23] - f[24] - f[25] + f[26]);
    u[1] = (1.0 / rho) * (f[3] - f[4] + f[7] - f[8] + f[11] - f[12] - f[13] + f[14] + f[17] -
                          f[18] + f[19] - f[20] + f[21] - f[22] - f[23] + f[24] + f[25] - f[26]);
    u[2] = (1.0 / rho) * (f[5] - f[6] + f[9] - f[10] + f[11] - f[12] - f[15] + f[16] - f[17] +
                          f[18] + f[19] - f[20] - f[21] + f[22] + f[23] - f[24] + f[25] - f[26]);
  }
}

/* ----------------------------------------------------------- lattice */
typedef struct scalar_lat {
  lbf_scalar_desc desc;
  double *g, *geq, *gs, *c, *c_shallow, *q_ads, *src_field;
  unsigned char* has_q; /* q_ads is Some(...) */
  double residual;
} scalar_lat;

struct lbo_case {
  vset v;
  int d, q;
  long n[3], N;
  lbf_case_desc desc;
  unsigned char* type; /* 0 fluid, 1 solid */
  long* nbr;           /* [N*q] periodic neighbour (lattice.rs:127-154) */
  long* fluid;
  long nfluid;
  long* face[6];
  long nface[6];
  int nfaces;
  long* bb;
  long nbb;
  unsigned* bbmask; /* [N] bit i: neighbour i solid (lattice.rs:177-191) */
  double *f, *feq, *fs, *rho, *u, *rho_sh, *u_sh, *force_field;
  int has_force;    /* a body force (uniform or field) is given */
  int multiphase;   /* Shan-Chen force on (src/momentum/multiphase.rs) */
  double sc_g, wall_phi;
  double* phi;      /* [N] pseudo-potential of the current step */
  scalar_lat sc[LBF_MAX_SCALARS];
  int ns;
  double res_rho, res_u[3];
  unsigned long long steps;
  int stopped;
};
typedef struct lbo_case lbo_case;

/* kernel::pseudo_potential, src/kernel.rs:171-174 (LATTICE_DENSITY = 1) */
static double pseudo_potential(double density) {
  const double ref = 1.0;
  return ref * (1.0 - exp(-density / ref));
}

/* Node::compute_shan_chen_force, src/momentum/multiphase.rs:56-93: every
 * direction's neighbour (periodic wrap on every axis) contributes its psi,
 * a solid neighbour the wall's. The reference computes this force twice per
 * step: in compute_velocity, inside the same parallel loop that computes
 * each node's density and psi (so a neighbour's psi may still be the
 * previous step's, depending on thread timing), and in the BGK collision,
 * after every psi is current. This restatement uses the current psi in both
 * places, which is the collision's value and one of the orders the
 * velocity pass can see. */
static void shan_chen_force(const lbo_case* L, long n, double* F) {
  const int q = L->q, D = L->d;
  const double node_phi = L->phi[n];
  for (int x = 0; x < D; x++) {
    double sum = 0.0;
    for (int i = 0; i < q; i++) {
      long nb = L->nbr[n * q + i];
      double phi_i = L->type[nb] ? L->wall_phi : L->phi[nb];
      sum = sum + L->v.w[i] * (double)L->v.c[i][x] * phi_i;
    }
    F[x] = -L->sc_g * node_phi * sum;
  }
}

/* Node::get_force, src/momentum/node.rs:388-404: body force, Shan-Chen
 * force, or their sum. Returns 0 when the node has no force. */
static int get_force(const lbo_case* L, long n, double* F) {
  if (!L->has_force && !L->multiphase) return 0;
  double B[3] = {0, 0, 0};
  if (L->has_force)
    for (int x = 0; x < L->d; x++)
      B[x] = L->desc.force_kind == LBF_SOURCE_FIELD ? L->force_field[n * L->d + x]
                                                    : L->desc.force[x];
  if (!L->multiphase) {
    for (int x = 0; x < L->d; x++) F[x] = B[x];
    return 1;
  }
  double S[3];
  shan_chen_force(L, n, S);
  for (int x = 0; x < L->d; x++) F[x] = L->has_force ? B[x] + S[x] : S[x];
  return 1;
}

void lbo_destroy(lbo_case* L) {
  if (!L) return;
  free(L->type), free(L->nbr), free(L->fluid), free(L->bb), free(L->bbmask);
  for (int f = 0; f < 6; f++) free(L->face[f]);
  free(L->f), free(L->feq), free(L->fs), free(L->rho), free(L->u), free(L->rho_sh),
      free(L->u_sh), free(L->force_field), free(L->phi);
  for (int s = 0; s < L->ns; s++) {
    scalar_lat* S = &L->sc[s];
    free(S->g), free(S->geq), free(S->gs), free(S->c), free(S->c_shallow), free(S->q_ads),
        free(S->src_field), free(S->has_q);
  }
  free(L);
}

static int check_collision(const vset* v, const lbf_collision* col, const char* what) {
  if (col->op == LBF_BGK) {
    if (!(col->tau > 0)) return fail(LBF_EINVAL, "%s: BGK tau must be > 0", what);
  } else if (col->op == LBF_TRT) {
  } else if (col->op == LBF_MRT) {
    if (!v->has_mrt)
      return fail(LBF_EUNSUPPORTED, "%s: MRT is not constructible for this velocity set", what);
  } else
    return fail(LBF_EINVAL, "%s: unknown collision operator %d", what, col->op);
  return 0;
}

/* Lattice::new, src/momentum/lattice.rs:27-207 (+ scalar lattice
 * src/passive_scalar/lattice.rs:26-181, node q_ads init node.rs:77-85). */
int lbo_create(const lbf_case_desc* d, lbo_case** out) {
  *out = NULL;
  if (!d || d->abi_version != LBF_ABI_VERSION) return fail(LBF_EINVAL, "bad descriptor/ABI version");
  lbo_case* L = (lbo_case*)calloc(1, sizeof *L);
  if (vset_init(d->velocity_set, &L->v)) {
    free(L);
    return fail(LBF_EUNSUPPORTED, "velocity set %d is not constructible", d->velocity_set);
  }
  L->desc = *d;
  L->d = L->v.d, L->q = L->v.q;
  L->n[0] = d->n[0], L->n[1] = d->n[1], L->n[2] = L->d == 3 ? d->n[2] : 1;
  if (L->n[0] < 1 || L->n[1] < 1 || L->n[2] < 1) {
    lbo_destroy(L);
    return fail(LBF_EINVAL, "lattice dimensions must be >= 1");
  }
  L->N = L->n[0] * L->n[1] * L->n[2];
  int rc = check_collision(&L->v, &d->collision, "momentum");
  if (rc) {
    lbo_destroy(L);
    return rc;
  }
  long N = L->N;
  int q = L->q, D = L->d;
  L->nfaces = D == 2 ? 4 : 6;
  for (int f = 0; f < L->nfaces; f++) {
    int k = d->faces[f].kind;
    if (k < LBF_BC_NO_SLIP || k > LBF_BC_PERIODIC) {
      lbo_destroy(L);
      return fail(LBF_EINVAL, "face %d has no boundary condition (Boundary faces not found)", f);
    }
    if (k == LBF_BC_ZOU_HE && L->v.id != LBF_D2Q9) {
      if (!(d->flags & LBF_CASE_ZOU_HE_3D)) {
        lbo_destroy(L);
        return fail(LBF_EUNSUPPORTED, "Zou & He BC computation function not defined.");
      }
      if (!zou_he_3d_ok(f, d->faces[f].known, D)) {
        lbo_destroy(L);
        return fail(LBF_EINVAL, "Zou-He face %d: give u (velocity face) or rho and the "
                                "tangential u (pressure face)", f);
      }
    }
  }
  L->type = (unsigned char*)calloc(N, 1);
  if (d->node_type)
    for (long i = 0; i < N; i++) {
      if (d->node_type[i] > 1) {
        lbo_destroy(L);
        return fail(LBF_EINVAL, "Invalid node type at %ld", i);
      }
      L->type[i] = d->node_type[i];
    }
  /* neighbours: (index + c_i).rem_euclid(n), lattice.rs:127-154 */
  L->nbr = (long*)malloc(sizeof(long) * N * q);
  for (long i = 0; i < N; i++) {
    long x = i % L->n[0], y = (i / L->n[0]) % L->n[1], z = i / (L->n[0] * L->n[1]);
    long idx[3] = {x, y, z};
    for (int k = 0; k < q; k++) {
      long nb[3] = {x, y, z};
      for (int a = 0; a < D; a++) {
        long t = (idx[a] + L->v.c[k][a]) % L->n[a];
        if (t < 0) t += L->n[a];
        nb[a] = t;
      }
      L->nbr[i * q + k] = nb[0] + L->n[0] * (nb[1] + L->n[1] * nb[2]);
    }
  }
  L->fluid = (long*)malloc(sizeof(long) * N);
  for (long i = 0; i < N; i++)
    if (!L->type[i]) L->fluid[L->nfluid++] = i;
  /* face sets: fluid nodes with index[axis] == limit, lattice.rs:163-175 */
  for (int f = 0; f < L->nfaces; f++) {
    int axis = f / 2;
    long limit = (f % 2) ? L->n[axis] - 1 : 0;
    L->face[f] = (long*)malloc(sizeof(long) * (L->nfluid + 1));
    for (long j = 0; j < L->nfluid; j++) {
      long i = L->fluid[j];
      long x = i % L->n[0], y = (i / L->n[0]) % L->n[1], z = i / (L->n[0] * L->n[1]);
      long idx[3] = {x, y, z};
      if (idx[axis] == limit) L->face[f][L->nface[f]++] = i;
    }
  }
  /* bounce-back nodes, lattice.rs:177-191 */
  L->bbmask = (unsigned*)calloc(N, sizeof(unsigned));
  L->bb = (long*)malloc(sizeof(long) * (L->nfluid + 1));
  for (long j = 0; j < L->nfluid; j++) {
    long i = L->fluid[j];
    unsigned m = 0;
    for (int k = 0; k < q; k++)
      if (L->type[L->nbr[i * q + k]]) m |= 1u << k;
    if (m) {
      L->bbmask[i] = m;
      L->bb[L->nbb++] = i;
    }
  }
  L->f = (double*)calloc(N * q, sizeof(double));
  L->feq = (double*)calloc(N * q, sizeof(double));
  L->fs = (double*)calloc(N * q, sizeof(double));
  L->rho = (double*)malloc(sizeof(double) * N);
  L->u = (double*)malloc(sizeof(double) * N * D);
  L->rho_sh = (double*)malloc(sizeof(double) * N);
  L->u_sh = (double*)malloc(sizeof(double) * N * D);
  for (long i = 0; i < N; i++) {
    L->rho[i] = d->initial_density ? d->initial_density[i] : d->initial_density_uniform;
    for (int x = 0; x < D; x++)
      L->u[i * D + x] =
          d->initial_velocity ? d->initial_velocity[i * D + x] : d->initial_velocity_uniform[x];
  }
  memcpy(L->rho_sh, L->rho, sizeof(double) * N);
  memcpy(L->u_sh, L->u, sizeof(double) * N * D);
  L->has_force = d->force_kind != LBF_SOURCE_NONE;
  L->multiphase = d->multiphase != 0;
  if (L->multiphase) {
    L->sc_g = d->g_function;
    L->wall_phi = pseudo_potential(d->wall_density); /* multiphase.rs:10-18 */
    L->phi = (double*)calloc(N, sizeof(double));
    for (long i = 0; i < N; i++) L->phi[i] = pseudo_potential(L->rho[i]);
  }
  if (d->force_kind == LBF_SOURCE_FIELD) {
    if (!d->force_field) {
      lbo_destroy(L);
      return fail(LBF_EINVAL, "force field pointer is NULL");
    }
    L->force_field = (double*)malloc(sizeof(double) * N * D);
    memcpy(L->force_field, d->force_field, sizeof(double) * N * D);
  }
  if (d->n_scalars > LBF_MAX_SCALARS) {
    lbo_destroy(L);
    return fail(LBF_EUNSUPPORTED, "at most %d scalars", LBF_MAX_SCALARS);
  }
  L->ns = (int)d->n_scalars;
  for (int s = 0; s < L->ns; s++) {
    scalar_lat* S = &L->sc[s];
    S->desc = d->scalars[s];
    if (S->desc.velocity_set != L->v.id) {
      lbo_destroy(L);
      return fail(LBF_EUNSUPPORTED, "scalar %d: velocity set must equal the momentum one", s);
    }
    rc = check_collision(&L->v, &S->desc.collision, "scalar");
    if (rc) {
      lbo_destroy(L);
      return rc;
    }
    for (int f = 0; f < L->nfaces; f++) {
      int k = S->desc.faces[f].kind;
      if (k < LBF_SBC_ANTI_BOUNCE_BACK || k > LBF_SBC_PERIODIC) {
        lbo_destroy(L);
        return fail(LBF_EINVAL, "scalar %d face %d has no boundary condition", s, f);
      }
    }
    S->g = (double*)calloc(N * q, sizeof(double));
    S->geq = (double*)calloc(N * q, sizeof(double));
    S->gs = (double*)calloc(N * q, sizeof(double));
    S->c = (double*)malloc(sizeof(double) * N);
    S->c_shallow = (double*)malloc(sizeof(double) * N);
    S->q_ads = (double*)calloc(N, sizeof(double));
    S->has_q = (unsigned char*)calloc(N, 1);
    for (long i = 0; i < N; i++)
      S->c[i] = S->desc.initial_value ? S->desc.initial_value[i] : S->desc.initial_value_uniform;
    memcpy(S->c_shallow, S->c, sizeof(double) * N);
    if (S->desc.source_kind == LBF_SOURCE_FIELD) {
      if (!S->desc.source_field) {
        lbo_destroy(L);
        return fail(LBF_EINVAL, "scalar source field pointer is NULL");
      }
      S->src_field = (double*)malloc(sizeof(double) * N);
      memcpy(S->src_field, S->desc.source_field, sizeof(double) * N);
    }
    if (S->desc.has_adsorption)
      for (long i = 0; i < N; i++) {
        S->has_q[i] = 1;
        /* q_ads = initial on momentum bounce-back nodes, 0 elsewhere (node.rs:77-85) */
        S->q_ads[i] = L->bbmask[i] ? S->desc.initial_q_ads : 0.0;
      }
  }
  *out = L;
  return 0;
}

/* ---------------------------------------------------- momentum phases */

/* Node::compute_density / compute_velocity(false), src/momentum/node.rs:418-457 */
static void node_density(lbo_case* L, long n) {
  int q = L->q;
  const double* f = L->f + n * q;
  double rho = f[0];
  for (int i = 1; i < q; i++) rho = rho + f[i];
  L->rho[n] = rho;
  if (L->multiphase) L->phi[n] = pseudo_potential(rho); /* compute_phi, multiphase.rs:48-54 */
}

static void node_velocity(lbo_case* L, long n) {
  int q = L->q, D = L->d;
  const double* f = L->f + n * q;
  const double rho = L->rho[n];
  for (int x = 0; x < D; x++) {
    double acc = f[0] * (double)L->v.c[0][x];
    for (int i = 1; i < q; i++) acc = acc + f[i] * (double)L->v.c[i][x];
    L->u[n * D + x] = acc / rho;
  }
  double F[3];
  if (get_force(L, n, F))
    for (int x = 0; x < D; x++) L->u[n * D + x] += 0.5 * DELTA_T * F[x] / rho;
}

static void node_moments(lbo_case* L, long n) {
  node_density(L, n);
  node_velocity(L, n);
}

/* BGK (+ Guo source) node.rs:311-333; TRT lib.rs:118-127; MRT lib.rs:129-139 */
static void node_collide(lbo_case* L, long n) {
  int q = L->q;
  const lbf_collision* col = &L->desc.collision;
  double* f = L->f + n * q;
  double* feq = L->feq + n * q;
  double* fs = L->fs + n * q;
  if (col->op == LBF_BGK) {
    k_bgk(&L->v, f, feq, col->tau, fs);
    double F[3], S[27];
    if (get_force(L, n, F)) {
      k_momentum_source(&L->v, L->u + n * L->d, F, col->tau, S);
      for (int i = 0; i < q; i++) fs[i] += S[i];
    }
  } else if (col->op == LBF_TRT) {
    k_trt(&L->v, f, feq, col->omega_plus, col->omega_minus, fs);
  } else {
    k_mrt(&L->v, L->rho[n], L->u + n * L->d, f, col->relaxation, fs);
  }
}

/* NodeLike::compute_streaming, src/lib.rs:141-152 (pull; solid neighbours
 * supply their never-updated zero f_star) */
static void node_stream(long n, int q, const vset* v, const long* nbr, const double* fs,
                        double* f) {
  for (int i = 0; i < q; i++) {
    int ib = v->qbar[i];
    f[n * q + ib] = fs[nbr[n * q + i] * q + ib];
  }
}

/* NodeLike::compute_inner_bounce_back, src/lib.rs:154-165 */
static void node_inner_bb(long n, int q, const vset* v, unsigned mask, const double* fs,
                          double* f) {
  for (int i = 0; i < q; i++)
    if (mask >> i & 1u) f[n * q + v->qbar[i]] = fs[n * q + i];
}

/* momentum face conditions, src/momentum/bc.rs:29-149 */
static int momentum_faces(lbo_case* L) {
  int q = L->q, D = L->d;
  const vset* v = &L->v;
  for (int fc = 0; fc < L->nfaces; fc++) {
    const lbf_face_bc* bc = &L->desc.faces[fc];
    const int* qf = v->qfaces[fc];
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (long j = 0; j < L->nface[fc]; j++) {
      long n = L->face[fc][j];
      double* f = L->f + n * q;
      const double* fs = L->fs + n * q;
      switch (bc->kind) {
        case LBF_BC_NO_SLIP: /* bc.rs:119-129 */
          for (int k = 0; k < v->nqf; k++) f[v->qbar[qf[k]]] = fs[qf[k]];
          break;
        case LBF_BC_BOUNCE_BACK: /* bc.rs:62-84 */
          for (int k = 0; k < v->nqf; k++) {
            int i = qf[k];
            double udc = bc->velocity[0] * (double)v->c[i][0];
            for (int x = 1; x < D; x++) udc = udc + bc->velocity[x] * (double)v->c[i][x];
            f[v->qbar[i]] = fs[i] - 2.0 * v->w[i] * bc->density * CS_2_INV * udc;
          }
          break;
        case LBF_BC_ANTI_BOUNCE_BACK: { /* bc.rs:86-117 */
          long nb = L->nbr[n * q + v->normal[fc]];
          double vel[3] = {0, 0, 0};
          for (int x = 0; x < D; x++) vel[x] = 1.5 * L->u[n * D + x] - 0.5 * L->u[nb * D + x];
          double udu = vel[0] * vel[0];
          for (int x = 1; x < D; x++) udu = udu + vel[x] * vel[x];
          for (int k = 0; k < v->nqf; k++) {
            int i = qf[k];
            double udc = vel[0] * (double)v->c[i][0];
            for (int x = 1; x < D; x++) udc = udc + vel[x] * (double)v->c[i][x];
            f[v->qbar[i]] = -fs[i] + 2.0 * v->w[i] * bc->density *
                                         (1.0 + 0.5 * udc * udc * CS_4_INV - 0.5 * udu * CS_2_INV);
          }
        } break;
        case LBF_BC_ZOU_HE: /* bc.rs:131-149 */
          if (v->id == LBF_D2Q9)
            bad |= k_zou_he_d2q9(fc, f, bc->known, bc->density, bc->velocity) != 0;
          else
            k_zou_he_general(v, fc, f, bc->known, bc->density, bc->velocity);
          break;
        default: /* Periodic: nothing (bc.rs:50) */
          break;
      }
    }
    if (bad) return fail(LBF_EINVAL, "Error: Wrong input of Zou & He (1997) boundary conditions.");
  }
  return 0;
}

/* Lattice::main_steps, src/momentum/lattice.rs:431-441 */
static int momentum_main_steps(lbo_case* L) {
  int q = L->q;
  long nf = L->nfluid;
  /* update_density_and_velocity_step (lattice.rs:330-335); densities and
   * psi of every node first (see shan_chen_force) */
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) node_density(L, L->fluid[j]);
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) node_velocity(L, L->fluid[j]);
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) {
    long n = L->fluid[j];
    k_equilibrium(&L->v, L->rho[n], L->u + n * L->d, L->feq + n * q);
  }
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) node_collide(L, L->fluid[j]);
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) node_stream(L->fluid[j], q, &L->v, L->nbr, L->fs, L->f);
#pragma omp parallel for schedule(static)
  for (long j = 0; j < L->nbb; j++) {
    long n = L->bb[j];
    node_inner_bb(n, q, &L->v, L->bbmask[n], L->fs, L->f);
  }
  return momentum_faces(L);
}

/* ------------------------------------------------------ scalar phases */

/* get_source_value, src/passive_scalar/node.rs:258-266, with the Langmuir
 * adsorption source and its q_ads update, src/passive_scalar/adsorption.rs:50-92 */
static int scalar_source(scalar_lat* S, const lbo_case* L, long n, double* out) {
  int has_user = S->desc.source_kind != LBF_SOURCE_NONE;
  double user = 0.0, ads = 0.0;
  if (has_user)
    user = S->desc.source_kind == LBF_SOURCE_FIELD ? S->src_field[n] : S->desc.source_uniform;
  int has_ads = S->desc.has_adsorption != 0;
  if (has_ads) {
    if (L->bbmask[n]) {
      double c = S->c[n];
      double q = S->q_ads[n];
      double r = S->desc.adsorption_rate * c * (S->desc.q_max - q) - S->desc.desorption_rate * q;
      ads = -r;
      /* update_q_ads, adsorption.rs:73-92 */
      double qn = q + r * DELTA_T;
      if (qn < 0.0)
        qn = 0.0;
      else if (qn > S->desc.q_max)
        qn = S->desc.q_max;
      S->q_ads[n] = qn;
    } else {
      ads = 0.0;
    }
  }
  if (has_user && has_ads)
    *out = user + ads;
  else if (has_user)
    *out = user;
  else if (has_ads)
    *out = ads;
  else
    return 0;
  return 1;
}

/* wall velocity of a scalar face condition, src/passive_scalar/bc.rs:43-59,200-219 */
static void scalar_wall_velocity(const lbo_case* L, int fc, long n, double* vel) {
  const lbf_face_bc* mb = &L->desc.faces[fc];
  int D = L->d;
  int known = 0;
  switch (mb->kind) {
    case LBF_BC_NO_SLIP:
      for (int x = 0; x < D; x++) vel[x] = 0.0;
      known = 1;
      break;
    case LBF_BC_BOUNCE_BACK:
      for (int x = 0; x < D; x++) vel[x] = mb->velocity[x];
      known = 1;
      break;
    case LBF_BC_ZOU_HE: {
      unsigned need = LBF_KNOWN_UX | LBF_KNOWN_UY | (D == 3 ? LBF_KNOWN_UZ : 0u);
      if ((mb->known & need) == need) {
        for (int x = 0; x < D; x++) vel[x] = mb->velocity[x];
        known = 1;
      }
    } break;
    default:
      break;
  }
  if (!known) {
    long nb = L->nbr[n * L->q + L->v.normal[fc]];
    for (int x = 0; x < D; x++) vel[x] = 1.5 * L->u[n * D + x] - 0.5 * L->u[nb * D + x];
  }
}

/* compute_anti_bounce_back_bc, src/passive_scalar/bc.rs:107-135 */
static void scalar_anti_bb(const lbo_case* L, scalar_lat* S, int fc, long n, double value) {
  const vset* v = &L->v;
  int q = L->q, D = L->d;
  double vel[3] = {0, 0, 0};
  scalar_wall_velocity(L, fc, n, vel);
  double udu = vel[0] * vel[0];
  for (int x = 1; x < D; x++) udu = udu + vel[x] * vel[x];
  double* g = S->g + n * q;
  const double* gs = S->gs + n * q;
  for (int k = 0; k < v->nqf; k++) {
    int i = v->qfaces[fc][k];
    double udc = vel[0] * (double)v->c[i][0];
    for (int x = 1; x < D; x++) udc = udc + vel[x] * (double)v->c[i][x];
    double geq =
        v->w[i] * value * (1.0 + udc * CS_2_INV + 0.5 * udc * udc * CS_4_INV - 0.5 * udu * CS_2_INV);
    g[v->qbar[i]] = -gs[i] + 2.0 * geq;
  }
}

static void scalar_faces(lbo_case* L, scalar_lat* S) {
  const vset* v = &L->v;
  int q = L->q;
  for (int fc = 0; fc < L->nfaces; fc++) {
    const lbf_scalar_face_bc* bc = &S->desc.faces[fc];
    const int* qf = v->qfaces[fc];
#pragma omp parallel for schedule(static)
    for (long j = 0; j < L->nface[fc]; j++) {
      long n = L->face[fc][j];
      double* g = S->g + n * q;
      const double* gs = S->gs + n * q;
      switch (bc->kind) {
        case LBF_SBC_ANTI_BOUNCE_BACK:
          scalar_anti_bb(L, S, fc, n, bc->scalar_value);
          break;
        case LBF_SBC_ANTI_BB_NO_FLUX: /* bc.rs:137-139 */
          scalar_anti_bb(L, S, fc, n, S->c[n]);
          break;
        case LBF_SBC_BB_NO_FLUX: /* bc.rs:141-151 */
          for (int k = 0; k < v->nqf; k++) g[v->qbar[qf[k]]] = gs[qf[k]];
          break;
        case LBF_SBC_ZEROTH_ORDER_NO_FLUX: { /* bc.rs:153-173 */
          long nb = L->nbr[n * q + v->normal[fc]];
          const double* ng = S->g + nb * q;
          if (bc->unknown_populations_only)
            for (int k = 0; k < v->nqf; k++) g[v->qbar[qf[k]]] = ng[v->qbar[qf[k]]];
          else
            for (int i = 0; i < q; i++) g[i] = ng[i];
        } break;
        case LBF_SBC_SECOND_ORDER_NO_FLUX: { /* bc.rs:175-198 */
          long nb = L->nbr[n * q + v->normal[fc]];
          long nn = L->nbr[nb * q + v->normal[fc]];
          const double* ng = S->g + nb * q;
          const double* nng = S->g + nn * q;
          if (bc->unknown_populations_only)
            for (int k = 0; k < v->nqf; k++) {
              int i = v->qbar[qf[k]];
              g[i] = 2.0 * ng[i] - nng[i];
            }
          else
            for (int i = 0; i < q; i++) g[i] = 2.0 * ng[i] - nng[i];
        } break;
        default:
          break;
      }
    }
  }
}

/* scalar Lattice::main_steps, src/passive_scalar/lattice.rs:342-350 */
static void scalar_main_steps(lbo_case* L, scalar_lat* S) {
  int q = L->q, D = L->d;
  long nf = L->nfluid;
  const vset* v = &L->v;
  /* update_scalar_value_step, node.rs:272-279 */
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) {
    long n = L->fluid[j];
    const double* g = S->g + n * q;
    double c = g[0];
    for (int i = 1; i < q; i++) c = c + g[i];
    double src;
    if (scalar_source(S, L, n, &src)) c += 0.5 * DELTA_T * src;
    S->c[n] = c;
  }
  /* equilibrium_step with the momentum node's velocity (node.rs:133-135) */
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) {
    long n = L->fluid[j];
    k_equilibrium(v, S->c[n], L->u + n * D, S->geq + n * q);
  }
  /* collision_step, lattice.rs:260-272 */
  const lbf_collision* col = &S->desc.collision;
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) {
    long n = L->fluid[j];
    const double* g = S->g + n * q;
    const double* geq = S->geq + n * q;
    double* gs = S->gs + n * q;
    if (col->op == LBF_BGK) { /* node.rs:195-213 */
      k_bgk(v, g, geq, col->tau, gs);
      double src;
      if (scalar_source(S, L, n, &src)) {
        double st[27];
        k_scalar_source(v, src, col->tau, st);
        for (int i = 0; i < q; i++) gs[i] += st[i];
      }
    } else if (col->op == LBF_TRT) {
      k_trt(v, g, geq, col->omega_plus, col->omega_minus, gs);
    } else {
      k_mrt(v, S->c[n], L->u + n * D, g, col->relaxation, gs);
    }
  }
#pragma omp parallel for schedule(static)
  for (long j = 0; j < nf; j++) node_stream(L->fluid[j], q, v, L->nbr, S->gs, S->g);
  /* inner_boundary_condition_step, lattice.rs:280-297 */
#pragma omp parallel for schedule(static)
  for (long j = 0; j < L->nbb; j++) {
    long n = L->bb[j];
    unsigned m = L->bbmask[n];
    if (S->desc.inner_bc == LBF_INNER_BOUNCE_BACK) {
      node_inner_bb(n, q, v, m, S->gs, S->g);
    } else { /* compute_inner_anti_bounce_back, node.rs:281-294 */
      for (int i = 0; i < q; i++)
        if (m >> i & 1u) {
          long sn = L->nbr[n * q + i];
          S->g[n * q + v->qbar[i]] = -S->gs[n * q + i] + 2.0 * v->w[i] * S->c[sn];
        }
    }
  }
  scalar_faces(L, S);
}

/* ------------------------------------------------- residuals / loop */

/* compute_lattice_residuals, src/momentum/lattice.rs:369-394 and
 * src/passive_scalar/lattice.rs:299-313 (sequential node-order sums) */
static void compute_residuals(lbo_case* L) {
  int D = L->d;
  double sr = 0.0, su[3] = {0, 0, 0};
  for (long j = 0; j < L->nfluid; j++) {
    long n = L->fluid[j];
    double dr = L->rho[n] - L->rho_sh[n];
    sr += dr * dr;
    for (int x = 0; x < D; x++) {
      double du = L->u[n * D + x] - L->u_sh[n * D + x];
      su[x] += du * du;
    }
  }
  L->res_rho = sqrt(sr);
  for (int x = 0; x < D; x++) L->res_u[x] = sqrt(su[x]);
  for (int s = 0; s < L->ns; s++) {
    scalar_lat* S = &L->sc[s];
    double sc = 0.0;
    for (long j = 0; j < L->nfluid; j++) {
      long n = L->fluid[j];
      double dc = S->c[n] - S->c_shallow[n];
      sc += dc * dc;
    }
    S->residual = sqrt(sc);
  }
}

/* update_shallow_nodes, lattice.rs:396-400 */
static void update_shallow(lbo_case* L) {
  int D = L->d;
  for (long j = 0; j < L->nfluid; j++) {
    long n = L->fluid[j];
    L->rho_sh[n] = L->rho[n];
    for (int x = 0; x < D; x++) L->u_sh[n * D + x] = L->u[n * D + x];
    for (int s = 0; s < L->ns; s++) L->sc[s].c_shallow[n] = L->sc[s].c[n];
  }
}

/* stop_condition, lattice.rs:407-418 / passive_scalar/lattice.rs:409-429 */
static int converged(const lbo_case* L) {
  int ok = L->res_rho <= TOL;
  for (int x = 0; x < L->d; x++) ok = ok && L->res_u[x] <= TOL;
  for (int s = 0; s < L->ns; s++) ok = ok && L->sc[s].residual <= TOL;
  return ok;
}

/* initialize_nodes: momentum lattice.rs:322-328, scalars lattice.rs:241-246 */
int lbo_initialize_nodes(lbo_case* L) {
  int q = L->q, D = L->d;
#pragma omp parallel for schedule(static)
  for (long j = 0; j < L->nfluid; j++) {
    long n = L->fluid[j];
    k_equilibrium(&L->v, L->rho[n], L->u + n * D, L->feq + n * q);
    memcpy(L->f + n * q, L->feq + n * q, sizeof(double) * q);
  }
  for (int s = 0; s < L->ns; s++) {
    scalar_lat* S = &L->sc[s];
#pragma omp parallel for schedule(static)
    for (long j = 0; j < L->nfluid; j++) {
      long n = L->fluid[j];
      k_equilibrium(&L->v, S->c[n], L->u + n * D, S->geq + n * q);
      memcpy(S->g + n * q, S->geq + n * q, sizeof(double) * q);
    }
  }
  L->steps = 0;
  L->stopped = 0;
  return 0;
}

/* one iteration of src/passive_scalar.rs:153-176 (momentum.rs:273-292) */
static int one_step(lbo_case* L, unsigned flags, unsigned long long max_iter) {
  int rc = 0;
  if (!(L->desc.flags & LBF_CASE_FREEZE_MOMENTUM)) rc = momentum_main_steps(L);
  if (rc) return rc;
  for (int s = 0; s < L->ns; s++) scalar_main_steps(L, &L->sc[s]);
  if (flags & LBF_STEP_RESIDUALS) compute_residuals(L);
  unsigned long long t = L->steps;
  L->steps++;
  if (flags & LBF_STEP_STOP_ON_CONVERGENCE) {
    int stop = (t > MIN_ITER && converged(L)) || t > max_iter;
    if (stop) {
      L->stopped = 1;
      return 0;
    }
  }
  update_shallow(L);
  return 0;
}

int lbo_main_steps(lbo_case* L, unsigned nsteps, unsigned flags) {
  // every call starts unstopped, as lbf_main_steps does: a stop ends the
  // call's loop, not the case (src/momentum.rs:273-292 only ends the loop)
  L->stopped = 0;
  for (unsigned k = 0; k < nsteps; k++) {
    if (L->stopped) break;
    int rc = one_step(L, flags, ~0ull);
    if (rc) return rc;
  }
  return 0;
}

int lbo_run(lbo_case* L, unsigned long long max_iterations, unsigned long long* steps_done) {
  L->stopped = 0;
  while (!L->stopped) {
    int rc = one_step(L, LBF_STEP_RESIDUALS | LBF_STEP_STOP_ON_CONVERGENCE, max_iterations);
    if (rc) return rc;
  }
  if (steps_done) *steps_done = L->steps;
  return 0;
}

int lbo_residuals(lbo_case* L, double* r_rho, double* r_u, double* r_c) {
  if (r_rho) *r_rho = L->res_rho;
  if (r_u)
    for (int x = 0; x < L->d; x++) r_u[x] = L->res_u[x];
  if (r_c)
    for (int s = 0; s < L->ns; s++) r_c[s] = L->sc[s].residual;
  return 0;
}

int lbo_time_step(lbo_case* L, unsigned long long* t) {
  *t = L->steps;
  return 0;
}

int lbo_reduce(lbo_case* L, int which, double* out) {
  int kind = which & 0xff, s = which >> 8;
  double acc = 0.0;
  switch (kind) {
    case LBF_REDUCE_MEAN_DENSITY: /* src/momentum/post.rs:43-58 */
    case LBF_REDUCE_TOTAL_MASS:
      for (long j = 0; j < L->nfluid; j++) acc += L->rho[L->fluid[j]];
      *out = kind == LBF_REDUCE_MEAN_DENSITY ? acc / (double)L->nfluid : acc;
      return 0;
    case LBF_REDUCE_SCALAR_TOTAL:
      if (s >= L->ns) return fail(LBF_EINVAL, "no scalar %d", s);
      for (long j = 0; j < L->nfluid; j++) acc += L->sc[s].c[L->fluid[j]];
      *out = acc;
      return 0;
    case LBF_REDUCE_ADSORBED:
      if (s >= L->ns) return fail(LBF_EINVAL, "no scalar %d", s);
      for (long j = 0; j < L->nbb; j++) acc += L->sc[s].q_ads[L->bb[j]];
      *out = acc;
      return 0;
    case LBF_REDUCE_NONFINITE:
      for (long j = 0; j < L->nfluid; j++) {
        long n = L->fluid[j];
        int bad = !isfinite(L->rho[n]);
        for (int x = 0; x < L->d; x++) bad |= !isfinite(L->u[n * L->d + x]);
        acc += bad;
      }
      *out = acc;
      return 0;
  }
  return fail(LBF_EINVAL, "unknown reduction %d", which);
}

int lbo_get_fields(lbo_case* L, double* rho, double* u) {
  if (rho) memcpy(rho, L->rho, sizeof(double) * L->N);
  if (u) memcpy(u, L->u, sizeof(double) * L->N * L->d);
  return 0;
}

int lbo_get_scalar(lbo_case* L, unsigned s, double* value, double* q_ads) {
  if ((int)s >= L->ns) return fail(LBF_EINVAL, "no scalar %u", s);
  scalar_lat* S = &L->sc[s];
  if (value) memcpy(value, S->c, sizeof(double) * L->N);
  if (q_ads)
    for (long i = 0; i < L->N; i++) q_ads[i] = S->has_q[i] ? S->q_ads[i] : NAN;
  return 0;
}

static double* field_ptr(lbo_case* L, int field, int which) {
  if (field == 0) return which == 0 ? L->f : which == 1 ? L->fs : L->feq;
  if (field - 1 < L->ns) {
    scalar_lat* S = &L->sc[field - 1];
    return which == 0 ? S->g : which == 1 ? S->gs : S->geq;
  }
  return NULL;
}

int lbo_get_pops(lbo_case* L, int field, double* out) {
  double* p = field_ptr(L, field, 0);
  if (!p) return fail(LBF_EINVAL, "no field %d", field);
  memcpy(out, p, sizeof(double) * L->N * L->q);
  return 0;
}

int lbo_set_pops(lbo_case* L, int field, const double* in) {
  double* p = field_ptr(L, field, 0);
  if (!p) return fail(LBF_EINVAL, "no field %d", field);
  memcpy(p, in, sizeof(double) * L->N * L->q);
  return 0;
}

int lbo_get_link_tables(lbo_case* L, unsigned char* type, unsigned* mask, unsigned char* faces) {
  if (type) memcpy(type, L->type, L->N);
  if (mask) memcpy(mask, L->bbmask, sizeof(unsigned) * L->N);
  if (faces) {
    memset(faces, 0, L->N);
    for (int f = 0; f < L->nfaces; f++)
      for (long j = 0; j < L->nface[f]; j++) faces[L->face[f][j]] |= (unsigned char)(1u << f);
  }
  return 0;
}

void lbo_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int lbo_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------ unit-level test hooks */

int lbo_tables(int id, int* d, int* q, int* c, double* w, int* qbar, int* qfaces, int* nqf,
               int* normal) {
  vset v;
  if (vset_init(id, &v)) return fail(LBF_EUNSUPPORTED, "velocity set %d", id);
  *d = v.d, *q = v.q, *nqf = v.nqf;
  for (int i = 0; i < v.q; i++) {
    for (int x = 0; x < 3; x++) c[i * 3 + x] = x < v.d ? v.c[i][x] : 0;
    w[i] = v.w[i];
    qbar[i] = v.qbar[i];
  }
  for (int f = 0; f < 6; f++) {
    normal[f] = v.normal[f];
    for (int k = 0; k < 9; k++) qfaces[f * 9 + k] = k < v.nqf ? v.qfaces[f][k] : -1;
  }
  return 0;
}

int lbo_mrt_matrices(int id, double* M, double* Minv) {
  vset v;
  if (vset_init(id, &v) || !v.has_mrt) return fail(LBF_EUNSUPPORTED, "no MRT for %d", id);
  for (int k = 0; k < v.q; k++)
    for (int i = 0; i < v.q; i++) M[k * v.q + i] = v.M[k][i], Minv[k * v.q + i] = v.Minv[k][i];
  return 0;
}

int lbo_k_equilibrium(int id, double value, const double* u, double* out) {
  vset v;
  if (vset_init(id, &v)) return -1;
  k_equilibrium(&v, value, u, out);
  return 0;
}
int lbo_k_bgk(int id, const double* f, const double* feq, double tau, double* out) {
  vset v;
  if (vset_init(id, &v)) return -1;
  k_bgk(&v, f, feq, tau, out);
  return 0;
}
int lbo_k_trt(int id, const double* f, const double* feq, double wp, double wm, double* out) {
  vset v;
  if (vset_init(id, &v)) return -1;
  k_trt(&v, f, feq, wp, wm, out);
  return 0;
}
int lbo_k_mrt(int id, double value, const double* u, const double* f, const double* s,
              double* out) {
  vset v;
  if (vset_init(id, &v) || !v.has_mrt) return -1;
  k_mrt(&v, value, u, f, s, out);
  return 0;
}
int lbo_k_meq(int id, double rho, const double* u, double* out) {
  k_meq(id, rho, u, out);
  return 0;
}
int lbo_k_momentum_source(int id, const double* u, const double* F, double tau, double* out) {
  vset v;
  if (vset_init(id, &v)) return -1;
  k_momentum_source(&v, u, F, tau, out);
  return 0;
}
int lbo_k_scalar_source(int id, double source, double tau_g, double* out) {
  vset v;
  if (vset_init(id, &v)) return -1;
  k_scalar_source(&v, source, tau_g, out);
  return 0;
}
int lbo_k_zou_he_general(int vs, int face, double* f, unsigned known, double density,
                         const double* vel) {
  vset v;
  if (vset_init(vs, &v)) return -1;
  k_zou_he_general(&v, face, f, known, density, vel);
  return 0;
}

int lbo_k_zou_he(int face, double* f, unsigned known, double density, const double* vel) {
  return k_zou_he_d2q9(face, f, known, density, vel);
}
int lbo_k_velocity_explicit(int id, double rho, const double* f, double* u) {
  k_velocity_explicit(id, rho, f, u);
  return 0;
}

/* node-level hooks used to restate the reference's node tests */
long lbo_neighbor(lbo_case* L, long node, int i) { return L->nbr[node * L->q + i]; }
int lbo_node_set(lbo_case* L, int field, int which, long node, const double* v) {
  double* p = field_ptr(L, field, which);
  if (!p) return -1;
  memcpy(p + node * L->q, v, sizeof(double) * L->q);
  return 0;
}
int lbo_node_get(lbo_case* L, int field, int which, long node, double* v) {
  double* p = field_ptr(L, field, which);
  if (!p) return -1;
  memcpy(v, p + node * L->q, sizeof(double) * L->q);
  return 0;
}
/* op 0: compute_density+compute_velocity(false); 1: streaming (momentum);
 * 2: no-slip face `arg` on node; 3: compute_equilibrium; 4: collision */
int lbo_node_op(lbo_case* L, int op, long node, int arg) {
  switch (op) {
    case 0:
      node_moments(L, node);
      return 0;
    case 1:
      node_stream(node, L->q, &L->v, L->nbr, L->fs, L->f);
      return 0;
    case 2:
      for (int k = 0; k < L->v.nqf; k++) {
        int i = L->v.qfaces[arg][k];
        L->f[node * L->q + L->v.qbar[i]] = L->fs[node * L->q + i];
      }
      return 0;
    case 3:
      k_equilibrium(&L->v, L->rho[node], L->u + node * L->d, L->feq + node * L->q);
      return 0;
    case 4:
      node_collide(L, node);
      return 0;
  }
  return -1;
}
int lbo_node_moments(lbo_case* L, long node, double* rho, double* u) {
  *rho = L->rho[node];
  for (int x = 0; x < L->d; x++) u[x] = L->u[node * L->d + x];
  return 0;
}
int lbo_node_set_moments(lbo_case* L, long node, double rho, const double* u) {
  L->rho[node] = rho;
  for (int x = 0; x < L->d; x++) L->u[node * L->d + x] = u[x];
  return 0;
}

/* struct sizes/offsets as compiled from include/lbflow_hip.h, so the tests
 * can check the ctypes mirror of the descriptor layout */
size_t lbo_sizeof(int which) {
  switch (which) {
    case 0: return sizeof(lbf_case_desc);
    case 1: return sizeof(lbf_scalar_desc);
    case 2: return sizeof(lbf_face_bc);
    case 3: return sizeof(lbf_collision);
    case 4: return sizeof(lbf_scalar_face_bc);
    case 5: return offsetof(lbf_case_desc, scalars);
    case 6: return offsetof(lbf_case_desc, shift_scale);
    case 7: return offsetof(lbf_scalar_desc, initial_q_ads);
    case 8: return offsetof(lbf_case_desc, force_field);
    case 9: return offsetof(lbf_case_desc, wall_density);
  }
  return 0;
}

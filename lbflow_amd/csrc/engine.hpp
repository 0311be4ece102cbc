// engine.hpp — shared host/device definitions of the step engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbf {

constexpr int kBlock = 256;            // threads per workgroup (4 wave64)
constexpr uint32_t kSolidBit = 1u << 31;  // node class: solid node
constexpr uint32_t kLinkMask = (1u << 27) - 1;  // bits 0..26: neighbour i is solid
constexpr int kMaxCols = 12;  // residual columns: rho, u0..u2, c0..c3, nonfinite, spare
constexpr int kHist = 64;     // residual history ring
// step workgroups whose residual partials one folding workgroup sums
// (misc.hip, launch_advance); the partial array holds
// partial_rows(nblocks) rows of kMaxCols
constexpr long long kResidualFold = 1024;
constexpr long long partial_rows(long long nblocks) {
  return nblocks + (nblocks + kResidualFold - 1) / kResidualFold;
}

enum Mode : int { kModePull = 0, kModeStreamOnly = 1, kModePhi = 2 };
enum Flag : uint32_t {
  kFlagStoreMoments = 1u,
  kFlagResiduals = 2u,
  kFlagStopCheck = 8u,   // evaluate the reference stop rule after the step
};

// Device control block (one per case).
struct Ctrl {
  unsigned long long step;     // completed steps
  unsigned long long max_iter; // stop when step index > max_iter (with kFlagStopCheck)
  int stopped;
  int fresh[5];                // field holds post-stream f (after init / set_pops)
  unsigned done;               // reserved
  unsigned pad;
  double res[kMaxCols];        // residual norms of the last step
  double hist[kHist][kMaxCols];
};

struct FaceK {
  int kind;
  unsigned known;
  double density;
  double vel[3];
};

struct SFaceK {
  int kind;
  int unknown_only;
  double value;
};

// Kernel arguments (passed by value). Pointers to population planes use a
// plane stride `ld` (elements); plane i of a field starts at base + i*ld.
struct KArgs {
  int nx, ny, nz, d;
  long long N, ld;
  int mode;
  uint32_t flags;
  uint32_t res_mask;  // residual columns reduced by the advancing launch
  int field;          // 0 momentum, 1+s scalar
  int nfaces;
  uint32_t face_active;  // faces with a non-periodic condition
  int faces_noslip_only; // every non-periodic face is no-slip (AA: nothing to apply)
  int fresh;             // AA: the array holds post-stream f in natural slots
  int aa_odd;            // AA: parity of the step being launched
  uint32_t step_offset;  // steps enqueued since the device step counter last advanced
  uint32_t plane_bytes;  // AA: bytes per population plane (array < 4 GiB)
  int xcd_group;         // runs of node blocks per XCD (0 = round-robin order)
  int tb_zs;             // two-step passes: z planes per workgroup segment
  int tb_shape;          // two-step passes: tile shape (1: rows of 64 nodes where they fit)
  void* pop[2];          // ping-pong population buffers (two-array scheme)
  void* out;             // stream-only destination: f64 [q][ld] over every node, or
                         // with out_aos f64 [node - out_n0][q] over [out_n0, out_n1)
  int out_aos;
  long long out_n0, out_n1;
  const uint32_t* cls;
  void* rho[2];          // momentum moment buffers (compute type)
  void* u[2];            // [d][ld] each
  void* c[2];            // scalar value buffers of this scalar
  double* q_ads;         // [ld] (scalar with adsorption)
  const double* src_field;  // force [d][ld] or scalar source [ld]
  double* partials;      // [grid][kMaxCols]
  Ctrl* ctrl;
  // collision
  int op;
  double omega, omega_p, wp, wm, coeff_b;
  double s[27];
  // forcing / source
  int force_kind;
  double F[3];
  double src_uniform;
  int has_ads;
  double k_a, k_d, q_max;
  int inner_bc;
  FaceK face[6];   // momentum face table (also used by scalars for the wall velocity)
  SFaceK sface[6];
  double scale, inv_scale;  // shifted storage
  // Shan-Chen force: psi of every fluid node of the current step in phi[ld]
  int multiphase;
  double sc_g, wall_phi;
  void* phi;
};

}  // namespace lbf

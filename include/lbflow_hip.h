/*
 * lbflow_hip.h — C ABI of the MI355X (gfx950) lattice-Boltzmann step engine.
 *
 * The engine replaces the per-node lattice update of gilberto-ribeiro/lbflow
 * (one `Lattice::main_steps` call per time step, src/momentum/lattice.rs:431-441,
 * and the coupled scalar step, src/passive_scalar/lattice.rs:342-350) with
 * device-resident structure-of-arrays populations and hand-written HIP kernels.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative LBF_E* code on error;
 *    lbf_last_error() then holds a thread-local message. Nothing aborts across
 *    the ABI (the reference panics instead, e.g. src/momentum/lattice.rs:49-54).
 *  - Node order is the reference's linear index i = x + nx*(y + ny*z)
 *    (src/momentum/lattice.rs:70-85). Host arrays are owned by the caller and
 *    hold f64 regardless of the storage precision on the device.
 *  - A case handle owns all its device memory and one HIP stream. A handle is
 *    not thread-safe; distinct handles may be driven from distinct threads.
 *  - No callbacks cross the ABI. The reference's user closures
 *    (`force: Fn(&Node)->Vec<Float>`, src/momentum.rs:25; `source_value:
 *    Fn(&Node)->Float`, src/passive_scalar.rs:19) are replaced by a uniform
 *    value or a static per-node field evaluated once on the host.
 */
#ifndef LBFLOW_HIP_H
#define LBFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBF_ABI_VERSION 1u
#define LBF_MAX_Q 27
#define LBF_MAX_SCALARS 4

/* ---- error codes -------------------------------------------------------- */
#define LBF_OK 0
#define LBF_EINVAL (-1)      /* invalid descriptor or argument               */
#define LBF_EUNSUPPORTED (-2) /* valid in the reference, not built here       */
#define LBF_EDEVICE (-3)     /* HIP runtime error                             */
#define LBF_ENOMEM (-4)      /* device or host allocation failed / HBM budget */
#define LBF_ESTATE (-5)      /* call not valid in the handle's current state  */
#define LBF_ENONFINITE (-6)  /* a NaN/Inf was detected in the fields          */
#define LBF_ETIMEOUT (-7)    /* a wait ended before the request completed     */

/* ---- velocity sets: discriminants of `VelocitySet` (src/velocity_set.rs:18-25) */
#define LBF_D2Q5 0 /* not constructible in the reference (todo!, velocity_set.rs:50) */
#define LBF_D2Q9 1
#define LBF_D3Q15 2
#define LBF_D3Q19 3
#define LBF_D3Q27 4

/* ---- collision operators: `CollisionOperator` (src/lib.rs:37-42) -------- */
#define LBF_BGK 0 /* BGK(tau)                                         */
#define LBF_TRT 1 /* TRT(omega_plus, omega_minus)                     */
#define LBF_MRT 2 /* MRT(relaxation vector of length Q)               */

/* ---- faces: `BoundaryFace` (src/lib.rs:17-25); this is also the order in
 *      which face conditions are applied (FACES_2D/FACES_3D, src/lib.rs:27-29) */
#define LBF_WEST 0
#define LBF_EAST 1
#define LBF_SOUTH 2
#define LBF_NORTH 3
#define LBF_BOTTOM 4
#define LBF_TOP 5

/* ---- momentum face conditions: `momentum::bc::BoundaryCondition`
 *      (src/momentum/bc.rs:11-26) ---------------------------------------- */
#define LBF_BC_UNSET 0
#define LBF_BC_NO_SLIP 1
#define LBF_BC_BOUNCE_BACK 2      /* moving wall {density, velocity}          */
#define LBF_BC_ANTI_BOUNCE_BACK 3 /* {density}                                 */
#define LBF_BC_ZOU_HE 4           /* {density?, velocity[d]?}, D2Q9 only        */
#define LBF_BC_PERIODIC 5

/* bits of lbf_face_bc.known for LBF_BC_ZOU_HE (Option<Float> presence) */
#define LBF_KNOWN_DENSITY 1u
#define LBF_KNOWN_UX 2u
#define LBF_KNOWN_UY 4u
#define LBF_KNOWN_UZ 8u

/* ---- scalar face conditions: `passive_scalar::bc::BoundaryCondition`
 *      (src/passive_scalar/bc.rs:9-30) ------------------------------------ */
#define LBF_SBC_UNSET 0
#define LBF_SBC_ANTI_BOUNCE_BACK 1       /* {scalar_value}                    */
#define LBF_SBC_ANTI_BB_NO_FLUX 2
#define LBF_SBC_BB_NO_FLUX 3
#define LBF_SBC_ZEROTH_ORDER_NO_FLUX 4   /* {unknown_populations_only}        */
#define LBF_SBC_SECOND_ORDER_NO_FLUX 5   /* {unknown_populations_only}        */
#define LBF_SBC_PERIODIC 6

/* `InnerBoundaryCondition` (src/passive_scalar/bc.rs:32-35) */
#define LBF_INNER_ANTI_BOUNCE_BACK 0
#define LBF_INNER_BOUNCE_BACK 1

/* ---- population storage on the device ----------------------------------- */
#define LBF_F64 0         /* plain f64, f64 arithmetic: the parity path        */
#define LBF_F32 1         /* plain f32 storage, f32 arithmetic                 */
#define LBF_F32_SHIFTED 2 /* (f_i - w_i)*scale stored in f32, f32 arithmetic   */
#define LBF_F16_SHIFTED 3 /* (f_i - w_i)*scale stored in IEEE fp16             */
#define LBF_BF16_SHIFTED 4 /* (f_i - w_i)*scale stored in bfloat16             */

/* ---- streaming scheme --------------------------------------------------- */
#define LBF_STREAM_AUTO 0      /* AA wherever the case allows it; in FP64 only
                                  where it gives the two-array scheme's bits (not
                                  with a Zou-He face next to another non-periodic
                                  face); else two arrays                         */
#define LBF_STREAM_TWO_ARRAY 1 /* fused pull-stream + collide, ping-pong       */
#define LBF_STREAM_AA 2        /* in-place AA pattern (one population array)   */
#define LBF_STREAM_TWO_STEP 3  /* two arrays, two time steps per pass with the
                                  first step's values held in LDS, where the
                                  steps need no residuals; other steps as
                                  LBF_STREAM_TWO_ARRAY, bit for bit the same
                                  (shifted storage: as LBF_STREAM_AA)
                                  (3-D lattices, any storage, faces no-slip,
                                  bounce-back or periodic, no scalars)        */

/* ---- source kinds (force / scalar source) ------------------------------- */
#define LBF_SOURCE_NONE 0
#define LBF_SOURCE_UNIFORM 1
#define LBF_SOURCE_FIELD 2

/* ---- case flags ---------------------------------------------------------- */
#define LBF_CASE_FREEZE_MOMENTUM 1u /* CLI --freeze-momentum (src/cli.rs:40-42) */
/* Zou-He faces on D3Q15/D3Q19/D3Q27 with this engine's own closure (no
 * counterpart: the reference panics, src/velocity_set.rs:140-160). Velocity
 * faces give every velocity component; pressure faces give the density and
 * the tangential components. */
#define LBF_CASE_ZOU_HE_3D 2u

/* ---- step flags ---------------------------------------------------------- */
#define LBF_STEP_RESIDUALS 1u /* residual norms after every step (lattice.rs:369-394) */
#define LBF_STEP_STOP_ON_CONVERGENCE 2u /* honour the stop rule (lattice.rs:407-418) */

typedef struct lbf_face_bc {
  int32_t kind;     /* LBF_BC_*                                              */
  uint32_t known;   /* LBF_KNOWN_* (Zou-He only)                             */
  double density;   /* wall / Zou-He density                                  */
  double velocity[3];
} lbf_face_bc;

typedef struct lbf_collision {
  int32_t op; /* LBF_BGK / LBF_TRT / LBF_MRT */
  int32_t reserved;
  double tau;                    /* BGK                                       */
  double omega_plus, omega_minus; /* TRT                                      */
  double relaxation[LBF_MAX_Q];  /* MRT, first Q entries used                 */
} lbf_collision;

typedef struct lbf_scalar_face_bc {
  int32_t kind; /* LBF_SBC_* */
  int32_t unknown_populations_only;
  double scalar_value;
} lbf_scalar_face_bc;

typedef struct lbf_scalar_desc {
  lbf_collision collision;
  int32_t velocity_set; /* must equal the momentum velocity set            */
  int32_t inner_bc;     /* LBF_INNER_*                                     */
  lbf_scalar_face_bc faces[6];
  const double* initial_value; /* [N] or NULL for the uniform value        */
  double initial_value_uniform;
  int32_t source_kind;         /* LBF_SOURCE_*                             */
  int32_t has_adsorption;      /* Langmuir model (adsorption.rs:4-10)      */
  double source_uniform;
  const double* source_field;  /* [N]                                      */
  double adsorption_rate, desorption_rate, q_max, initial_q_ads;
} lbf_scalar_desc;

typedef struct lbf_case_desc {
  uint32_t abi_version; /* LBF_ABI_VERSION */
  int32_t velocity_set; /* LBF_D2Q9 ...    */
  uint32_t n[3];        /* n[2] ignored (1) for 2-D lattices */
  uint32_t flags;       /* LBF_CASE_*      */
  lbf_collision collision;
  const uint8_t* node_type; /* [N], 0 fluid / 1 solid (NodeType, lib.rs:31-35); NULL = all fluid */
  lbf_face_bc faces[6];
  int32_t force_kind; /* LBF_SOURCE_* */
  int32_t precision;  /* LBF_F64 ...  */
  double force[3];
  const double* force_field; /* [N*d] */
  const double* initial_density;  /* [N] or NULL */
  double initial_density_uniform;
  const double* initial_velocity; /* [N*d] or NULL */
  double initial_velocity_uniform[3];
  uint32_t n_scalars;
  int32_t streaming;  /* LBF_STREAM_* */
  const lbf_scalar_desc* scalars;
  double shift_scale; /* shifted storage: stored = (f - w)*scale; 0 = automatic */
  /* Shan-Chen pseudo-potential force (`multiphase::Parameters`,
   * src/momentum/multiphase.rs:4-94): F = -G psi(rho) sum_i w_i c_i psi_i with
   * psi(rho) = 1 - exp(-rho) (src/kernel.rs:171-174) and psi(wall_density) for
   * solid neighbours; added to the body force (node.rs:388-404). */
  int32_t multiphase; /* 0 = off */
  int32_t reserved;
  double g_function;   /* G */
  double wall_density; /* rho_w */
} lbf_case_desc;

typedef struct lbf_case lbf_case;

/* ---- case lifecycle ------------------------------------------------------
 * lbf_create: validates the descriptor, builds the node classes (periodic
 * neighbour table, fluid / face / bounce-back sets — src/momentum/lattice.rs:
 * 127-191 and src/passive_scalar/lattice.rs:105-166) and uploads them.
 * Replaces `Lattice::new` (src/momentum/lattice.rs:27-207). */
int lbf_create(const lbf_case_desc* desc, lbf_case** out);
void lbf_destroy(lbf_case* c);

/* f = f_eq(rho0, u0) on fluid nodes, then g = g_eq(c0, u0) for every scalar.
 * Replaces `Lattice::initialize_nodes` (src/momentum/lattice.rs:322-328,
 * src/passive_scalar/lattice.rs:241-246). */
int lbf_initialize_nodes(lbf_case* c);

/* nsteps × (momentum `main_steps` + every scalar's `main_steps`), each step
 * followed by residuals/shallow update when LBF_STEP_RESIDUALS is set.
 * Replaces one iteration of the loop in src/passive_scalar.rs:153-176 /
 * src/momentum.rs:273-292 (without file output). Asynchronous on the case
 * stream; lbf_synchronize waits. */
int lbf_main_steps(lbf_case* c, uint32_t nsteps, uint32_t flags);
int lbf_synchronize(lbf_case* c);

/* The reference time loop with its stop rule: runs until
 * (t > MIN_ITER && all residuals <= 1e-7) || t > max_iterations, i.e. at most
 * max_iterations + 2 steps (src/momentum.rs:273-292, lattice.rs:407-418).
 * Convergence is detected on the device; *steps_done receives the count. */
int lbf_run(lbf_case* c, uint64_t max_iterations, uint64_t* steps_done);

/* Residual norms of the last completed step: |Δrho|_2, |Δu_x|_2.. (d values),
 * |Δc|_2 per scalar (src/momentum/lattice.rs:369-394,
 * src/passive_scalar/lattice.rs:299-313). r_u / r_c may be NULL. */
int lbf_residuals(lbf_case* c, double* r_rho, double* r_u, double* r_c);

/* Number of completed steps (`Lattice::time_step`, lattice.rs:267). */
int lbf_time_step(lbf_case* c, uint64_t* t);

/* FP64 device reductions over fluid nodes. */
#define LBF_REDUCE_MEAN_DENSITY 0 /* compute_mean_density (src/momentum/post.rs:43-58) */
#define LBF_REDUCE_TOTAL_MASS 1   /* sum of rho over fluid nodes                        */
#define LBF_REDUCE_SCALAR_TOTAL 2 /* sum of c over fluid nodes (scalar index `which>>8`) */
#define LBF_REDUCE_ADSORBED 3     /* sum of q_ads over bounce-back nodes                 */
#define LBF_REDUCE_NONFINITE 4    /* number of fluid nodes with a non-finite moment      */
int lbf_reduce(lbf_case* c, int32_t which, double* out);

/* Macroscopic fields in reference node order: rho[N], u[N*d]. Solid nodes keep
 * their initial values as in the reference (they are never updated). */
int lbf_get_fields(lbf_case* c, double* rho, double* u);
/* Scalar value c[N] and adsorbed amount q_ads[N] (NaN where the reference has
 * None) of scalar s; either pointer may be NULL. */
int lbf_get_scalar(lbf_case* c, uint32_t s, double* value, double* q_ads);

/* Post-stream populations (the `f` / `g` vectors of the reference nodes) in
 * reference AoS order out[node*Q + i]. field 0 = momentum, 1+s = scalar s. */
int lbf_get_pops(lbf_case* c, int32_t field, double* out);
/* lbf_get_pops without any device or host allocation. No counterpart in the
 * reference: this is for callers that cannot allocate (code in a tight
 * serving or time loop, harnesses that forbid hipMalloc). The populations
 * pass through a fixed 16 MiB device array of the library (one per GPU, kept
 * until the process ends) in chunks of at most 16 MiB / (8 Q) nodes, so every
 * lattice size is accepted and no request needs more than that array. Calls
 * are serialised inside the process (they share the array); the result is
 * the same as lbf_get_pops. */
int lbf_get_pops_noalloc(lbf_case* c, int32_t field, double* out);
/* Overwrite the post-stream populations (parity tests, restarts). */
int lbf_set_pops(lbf_case* c, int32_t field, const double* in);

/* Node classes as built on the device, for bit-exact checks:
 * node_type[N] (0/1), bb_link_mask[N] (bit i set when the periodic neighbour
 * in direction i is solid, fluid nodes only), face_bits[N] (bit f set when
 * the fluid node lies on face f). Any pointer may be NULL. */
int lbf_get_link_tables(lbf_case* c, uint8_t* node_type, uint32_t* bb_link_mask,
                        uint8_t* face_bits);

/* Bytes of device memory the case holds (for HBM-budget admission). */
int lbf_device_bytes(const lbf_case_desc* desc, uint64_t* bytes);

const char* lbf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* LBFLOW_HIP_H */

/*
 * lbflow_serve.h — serving tier of the MI355X lattice-Boltzmann engine.
 *
 * The reference has no counterpart: lbflow solves one case per process
 * (`lbflow::solve`, src/lib.rs:170). This interface runs many independent
 * cases on one GPU with continuous batching:
 *  - a request (case descriptor + step count) is queued on submit and
 *    admitted while the resident cases' device bytes stay within the HBM
 *    budget and a slot is free; admission happens as earlier cases finish;
 *  - resident cases keep their populations in HBM and advance together in
 *    batched launches (one per case type) captured in HIP graphs of K steps,
 *    whose workgroup count and tile size an autotuner picks by measurement
 *    (choices are kept for the process);
 *  - a finished case's fields (and, on request, populations) are copied to
 *    host memory and its device memory is freed.
 * A case advances exactly as through lbf_main_steps(case, steps, 0), so its
 * result is bitwise the same alone, in any mix, and in the single-case path.
 *
 * Servable cases: D2Q9 or D3Q19; BGK or TRT; storage LBF_F32,
 * LBF_F32_SHIFTED, LBF_F16_SHIFTED or LBF_BF16_SHIFTED (in-place AA scheme);
 * faces periodic, no-slip or bounce-back with opposite faces equally
 * periodic; no scalars. Other descriptors are rejected at submit with
 * LBF_EUNSUPPORTED.
 *
 * Threading: a server owns a worker thread (the GPU loop, on a stream with
 * a hardware queue of its own) and eight helper threads (case creation and
 * initialisation; result download and release). All functions may be called
 * from any thread.
 */
#ifndef LBFLOW_SERVE_H
#define LBFLOW_SERVE_H

#include "lbflow_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbf_server lbf_server;

typedef struct lbf_server_config {
  uint32_t abi_version;         /* LBF_ABI_VERSION                                  */
  uint32_t max_slots;           /* resident cases at once, 1..256 (0 = 256)         */
  uint64_t hbm_budget_bytes;    /* device bytes resident cases may hold (0 = 80 % of
                                   the device's free memory at creation)             */
  uint32_t max_steps_per_round; /* steps per graph replay, power of two <= 256 (0 = 64) */
  int32_t autotune;             /* 1: choose the batched launches' workgroup cap and
                                   tile size by measurement per mix of case types;
                                   finished choices are kept for the process and
                                   reused by later servers                          */
  uint32_t grid;                /* workgroup cap of a batched launch when autotune =
                                   0: a launch takes min(grid, its tile count rounded
                                   up to a power of two) workgroups, each walking
                                   its share of the tiles (0 = no cap: one tile per
                                   workgroup)                                        */
  uint32_t tile;                /* nodes per tile (= threads per workgroup) when
                                   autotune = 0: 128, 256 or 512 (0 = 256)          */
} lbf_server_config;

/* submit flags */
#define LBF_SERVE_KEEP_POPS 1u /* also keep the final post-stream populations */

/* request states */
#define LBF_REQ_QUEUED 0
#define LBF_REQ_PREPARING 1 /* admitted: device memory reserved, case being set up */
#define LBF_REQ_RUNNING 2
#define LBF_REQ_DONE 3
#define LBF_REQ_FAILED 4

typedef struct lbf_request_info {
  int32_t state;          /* LBF_REQ_*                                    */
  int32_t error;          /* error code when FAILED                       */
  uint64_t nodes;         /* lattice nodes                                */
  uint64_t steps;         /* requested steps                              */
  uint64_t device_bytes;  /* HBM the case holds while resident            */
  double t_submit, t_admit, t_start, t_done; /* seconds since server creation */
} lbf_request_info;

typedef struct lbf_server_stats {
  uint64_t submitted, completed, failed, queued, resident;
  uint64_t rounds;            /* graph replays                                   */
  uint64_t graphs_built;      /* graphs captured (one per launch shape of a mix:
                                 types, K, cap, tile, per-class tile buckets)   */
  uint64_t hbm_in_use, hbm_peak, hbm_budget;
  double node_updates;        /* sum of nodes * steps over completed requests     */
  double first_submit, last_done; /* seconds since server creation               */
  double gpu_seconds;         /* time spent in graph replays                      */
  double p50_latency, p99_latency, mean_latency; /* submit -> done, completed requests */
  double cases_per_second;    /* completed / (last_done - first_submit)           */
  double mlups;               /* node_updates / (last_done - first_submit) / 1e6  */
  double prep_seconds;        /* case creation + initialisation, summed over helper threads */
  double retire_seconds;      /* result download + release, summed over helper threads */
  uint32_t tuned_grid;        /* workgroup cap the autotuner settled on last      */
  uint32_t tuned_tile;        /* tile size the autotuner settled on last          */
} lbf_server_stats;

int lbf_server_create(const lbf_server_config* cfg, lbf_server** out);
/* Stops accepting work, waits for nothing: resident and queued requests are
 * dropped and their memory freed. */
void lbf_server_destroy(lbf_server* s);

/* Queues a case for `steps` steps. The descriptor's host arrays are copied,
 * so they may be freed when this returns. */
int lbf_server_submit(lbf_server* s, const lbf_case_desc* desc, uint64_t steps, uint32_t flags,
                      uint64_t* ticket);
int lbf_server_info(lbf_server* s, uint64_t ticket, lbf_request_info* info);
/* Waits until the request is DONE or FAILED (LBF_ETIMEOUT after timeout_s). */
int lbf_server_wait(lbf_server* s, uint64_t ticket, double timeout_s);
/* Copies the result of a DONE request: rho[N], u[N*d] (reference node order)
 * and, with LBF_SERVE_KEEP_POPS, pops[N*Q] (AoS); any pointer may be NULL.
 * release != 0 frees the stored result and forgets the ticket. */
int lbf_server_result(lbf_server* s, uint64_t ticket, double* rho, double* u, double* pops,
                      int32_t release);
/* Waits until every submitted request is DONE or FAILED. */
int lbf_server_drain(lbf_server* s, double timeout_s);
int lbf_server_get_stats(lbf_server* s, lbf_server_stats* out);
/* Seconds since server creation on the server's clock. */
double lbf_server_now(lbf_server* s);

#ifdef __cplusplus
}
#endif

#endif /* LBFLOW_SERVE_H */

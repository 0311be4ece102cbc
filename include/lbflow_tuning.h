/* lbflow_tuning.h — benchmark and tuning hooks of the HIP engine.
 *
 * None of these has a counterpart in the reference (lbflow solves on the CPU
 * and has no FFI); the Rust binding does not need them. bench.py and the
 * tuning probes use them to time kernels on the case's own stream, to measure
 * the device's attainable copy bandwidth and to try launch orders and array
 * layouts. Same conventions as lbflow_hip.h: 0 = OK, negative = error.
 */
#ifndef LBFLOW_TUNING_H
#define LBFLOW_TUNING_H

#include <stddef.h>

#include "lbflow_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Copies `bytes` (a multiple of 16) from src to dst (device pointers) on the
 * HIP stream `stream` (NULL = default stream) with a 16-byte-per-lane kernel:
 * the attainable-bandwidth reference and the calibration kernel for the
 * rocprofv3 FETCH_SIZE / WRITE_SIZE counters. */
int lbfx_copy(void* dst, const void* src, size_t bytes, void* stream);

/* In-place sweep of q (9, 15, 19 or 27) planes of ld elements of es (2, 4
 * or 8) bytes from `base` (device memory, q * ld * es below 4 GiB), the
 * first n elements of each, on `stream`: every lane reads its element of each
 * plane and writes it to the paired plane, with the in-place kernel's buffer
 * addressing. That kernel's access pattern without arithmetic or
 * bounce-back, so its rate bounds that kernel's attainable bandwidth.
 * Returns -1 for an unsupported q, es or size. */
int lbfx_sweep(void* base, int q, long long ld, int es, long long n, void* stream);

/* The HIP stream (hipStream_t) a case enqueues its launches on, so callers can
 * bracket its work with events. */
void* lbfx_case_stream(lbf_case* c);

/* Workgroup order of the in-place kernel: runs of g consecutive node blocks
 * per XCD (0 = plain round-robin). Returns -1 when the case's grid is not a
 * multiple of 8 g. */
int lbfx_set_block_group(lbf_case* c, int g);

/* Pad, in elements (a multiple of 64, at most 2^20), added to the plane
 * stride of cases created after the call (process-wide; 0 by default).
 * Returns -1 for an invalid pad. */
int lbfx_set_plane_pad(long long elements);

/* The engine's MRT matrices of velocity set vs (LBF_D2Q9 .. LBF_D3Q27) as the
 * kernels use them: M[k*Q + i] (moment k of direction i) and
 * Minv[i*Q + k]. D2Q9 / D3Q15 / D3Q19 follow the reference's tables
 * (d2q9.rs:44-156, d3q15.rs:59-363, d3q19.rs:69-548); D3Q27 is this engine's
 * own tensor-product basis (the reference has none, d3q27.rs:86). Host only,
 * no device call. Returns -1 for a set without MRT. */
int lbfx_mrt_matrices(int velocity_set, double* M, double* Minv);

#ifdef __cplusplus
}
#endif

#endif /* LBFLOW_TUNING_H */

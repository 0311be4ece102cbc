// api.hip — host implementation of the C ABI declared in include/lbflow_hip.h.
//
// A case handle owns its device buffers and one HIP stream (default
// priority). Geometry classes are built once on the host, bit-exact with the
// reference's construction (src/momentum/lattice.rs:127-191), and uploaded as
// one 32-bit class word per node. The time loop runs on the device: every
// step is a fixed sequence of launches whose last one advances the device
// step counter, so lbf_main_steps and lbf_run only enqueue work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbflow_hip.h"
#include "dispatch.hpp"
#include "internal.hpp"
#include "engine.hpp"
#include "lattice.hpp"

using namespace lbf;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(LBF_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                 \
  } while (0)

struct VsInfo {
  int d, q;
  bool mrt;
  int c[27][3];
};

bool vs_info(int id, VsInfo* v) {
  static const int c9[9][3] = {{0, 0, 0}, {1, 0, 0},  {0, 1, 0},   {-1, 0, 0}, {0, -1, 0},
                               {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
  static const int c15[15][3] = {{0, 0, 0},   {1, 0, 0},  {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},
                                 {0, 0, 1},   {0, 0, -1}, {1, 1, 1},   {-1, -1, -1}, {1, 1, -1},
                                 {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};
  static const int c27[27][3] = {
      {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
      {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},  {-1, 0, -1}, {0, 1, 1},   {0, -1, -1}, {1, -1, 0},
      {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1}, {0, 1, -1},  {0, -1, 1},  {1, 1, 1},   {-1, -1, -1},
      {1, 1, -1},  {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};
  switch (id) {
    case LBF_D2Q9:
      v->d = 2, v->q = 9, v->mrt = true;
      memcpy(v->c, c9, sizeof c9);
      return true;
    case LBF_D3Q15:
      v->d = 3, v->q = 15, v->mrt = true;
      memcpy(v->c, c15, sizeof c15);
      return true;
    case LBF_D3Q19:  // first 19 directions of the D3Q27 ordering
      v->d = 3, v->q = 19, v->mrt = true;
      memcpy(v->c, c27, sizeof(int) * 3 * 19);
      return true;
    case LBF_D3Q27:
      v->d = 3, v->q = 27, v->mrt = true;  // own basis (no reference MRT)
      memcpy(v->c, c27, sizeof c27);
      return true;
  }
  return false;
}

// The AA scheme redirects links across non-periodic faces to the node's own
// slots, which reproduces no-slip and moving-wall bounce-back exactly, and
// Zou-He replaces those populations from the node's other ones; it needs
// opposite faces to be both periodic or both not, and no scalars.
bool aa_compatible(const lbf_case_desc* d, const VsInfo& v) {
  if (d->n_scalars || d->multiphase) return false;
  for (int f = 0; f < 2 * v.d; f++) {
    const int k = d->faces[f].kind;
    if (k != LBF_BC_NO_SLIP && k != LBF_BC_BOUNCE_BACK && k != LBF_BC_PERIODIC &&
        k != LBF_BC_ZOU_HE)
      return false;
  }
  for (int a = 0; a < v.d; a++)
    if ((d->faces[2 * a].kind == LBF_BC_PERIODIC) != (d->faces[2 * a + 1].kind == LBF_BC_PERIODIC))
      return false;
  return true;
}

// The in-place scheme gives the reference's bits in FP64 unless a Zou-He face
// meets another non-periodic face: at such edges the reference applies the
// faces in order (src/lib.rs:27-29), so the Zou-He closure can still see the
// wrapped pull values of a later face's links, while the in-place gather has
// already put bounce-back values there.
bool aa_reproduces_reference(const lbf_case_desc* d, const VsInfo& v) {
  for (int f = 0; f < 2 * v.d; f++) {
    if (d->faces[f].kind != LBF_BC_ZOU_HE) continue;
    for (int g = 0; g < 2 * v.d; g++)
      if (g / 2 != f / 2 && d->faces[g].kind != LBF_BC_PERIODIC) return false;
  }
  return true;
}

// Two steps per pass (tb.hpp): 3-D lattices in any storage, faces no-slip,
// moving bounce-back or periodic (opposite faces alike), no scalars and no
// Shan-Chen force.
bool tb_compatible(const lbf_case_desc* d, const VsInfo& v) {
  if (v.d != 3 || d->n_scalars || d->multiphase) return false;
  if (d->precision < LBF_F64 || d->precision > LBF_BF16_SHIFTED) return false;
  for (int f = 0; f < 2 * v.d; f++) {
    const int k = d->faces[f].kind;
    if (k != LBF_BC_NO_SLIP && k != LBF_BC_BOUNCE_BACK && k != LBF_BC_PERIODIC) return false;
  }
  for (int a = 0; a < v.d; a++)
    if ((d->faces[2 * a].kind == LBF_BC_PERIODIC) != (d->faces[2 * a + 1].kind == LBF_BC_PERIODIC))
      return false;
  return true;
}

bool get_launchers(int vs, int prec, Launchers* L) {
  switch (vs) {
    case LBF_D2Q9: return launchers_d2q9(prec, L);
    case LBF_D3Q15: return launchers_d3q15(prec, L);
    case LBF_D3Q19: return launchers_d3q19(prec, L);
    case LBF_D3Q27: return launchers_d3q27(prec, L);
  }
  return false;
}

// lattice weight from the squared length of c (velocity-set tables:
// d2q9.rs:20-30, d3q15.rs:25-41, d3q19.rs:29-49, d3q27.rs:37-65)
double lattice_weight(int vs, const int* c) {
  const int c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  switch (vs) {
    case LBF_D2Q9: return c2 == 0 ? 4.0 / 9.0 : (c2 == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
    case LBF_D3Q15: return c2 == 0 ? 2.0 / 9.0 : (c2 == 1 ? 1.0 / 9.0 : 1.0 / 72.0);
    case LBF_D3Q19: return c2 == 0 ? 1.0 / 3.0 : (c2 == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
    case LBF_D3Q27:
      return c2 == 0 ? 8.0 / 27.0 : (c2 == 1 ? 2.0 / 27.0 : (c2 == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
  }
  return 0.0;
}

// float -> bfloat16 bits, round to nearest even
uint16_t to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

void store_shifted(int prec, char* buf, size_t k, double v) {
  if (prec == LBF_F32_SHIFTED) {
    reinterpret_cast<float*>(buf)[k] = (float)v;
  } else if (prec == LBF_F16_SHIFTED) {
    const _Float16 h = (_Float16)(float)v;
    memcpy(buf + 2 * k, &h, 2);
  } else {
    const uint16_t b = to_bf16((float)v);
    memcpy(buf + 2 * k, &b, 2);
  }
}

size_t storage_bytes(int prec) {
  switch (prec) {
    case LBF_F64: return 8;
    case LBF_F32:
    case LBF_F32_SHIFTED: return 4;
    default: return 2;
  }
}

}  // namespace

struct ScalarDev {
  void* pop[2] = {nullptr, nullptr};
  void* c[2] = {nullptr, nullptr};
  double* q_ads = nullptr;
  double* src = nullptr;
  lbf_scalar_desc desc{};
};

// Steps per captured batch, and the largest case that replays batches (a
// 2^21-node step takes about 60 us in FP32, so the few microseconds of launch
// latency per kernel matter only below that).
constexpr uint32_t kGraphSteps = 16;
constexpr long long kGraphMaxNodes = 1ll << 21;

struct lbf_case {
  lbf_case_desc desc{};
  VsInfo v{};
  int d = 0, q = 0, nx = 1, ny = 1, nz = 1;
  long long N = 0, ld = 0;
  int prec = LBF_F64;
  size_t sb = 8, tb = 8;  // storage / compute bytes
  int device = 0;
  int grid = 1;
  hipStream_t stream = nullptr;
  Launchers L{};
  // device memory
  void* pop[2] = {nullptr, nullptr};
  uint32_t* cls = nullptr;
  void* rho[2] = {nullptr, nullptr};
  void* u[2] = {nullptr, nullptr};
  double* force = nullptr;
  void* phi = nullptr;  // Shan-Chen pseudo-potential of the current step
  Ctrl* ctrl = nullptr;
  double* partials = nullptr;
  double* readback = nullptr;
  double* red_out = nullptr;
  ScalarDev sc[LBF_MAX_SCALARS];
  int ns = 0;
  bool aa = false;       // in-place AA scheme (one population array)
  // two-array scheme with two steps per pass (tb.hpp) where the steps allow
  // it: ts_grid workgroups of ts_zs z planes each
  bool two_step = false;
  int ts_grid = 0, ts_zs = 0, ts_shape = 0;
  uint64_t host_step = 0;  // steps enqueued (device count, resynchronised on reads)
  // steps enqueued since the device step counter last advanced
  // (their launches carry it as step_offset; flush_steps advances the counter)
  uint32_t pending_steps = 0;
  bool aa_fresh = false;   // AA: array holds post-stream f (after init / set_pops)
  // AA: the in-place pattern of step t is odd when (t + aa_shift) is odd. A
  // fresh array (post-stream f in natural slots) must start with the even
  // pattern, which reads and writes only the node's own slots; the odd
  // pattern writes the neighbours' slots, which they would still be reading
  // as their own fresh values. initialize_nodes and set_pops set the shift.
  int aa_shift = 0;
  int xcd_group = 0;       // AA: runs of consecutive blocks per XCD (0: round-robin)
  bool may_have_stopped = false;  // steps were enqueued with the stop rule on
  double scale = 1.0;    // shifted storage: stored = (f - w) * scale
  double w[27] = {};
  bool frozen = false;
  bool needs_moments = false;  // some condition reads stored moments every step
  // captured batches of kGraphSteps steps (cases of at most kGraphMaxNodes
  // nodes, where launch latency bounds the step), keyed by step flags
  bool graphs_on = false;
  struct StepGraph {
    uint32_t flags;
    bool every_last;
    hipGraphExec_t exec;
  } graphs[4];
  int ngraphs = 0;
  uint32_t face_active = 0;
  // host copies of the node classes
  std::vector<uint8_t> h_type;  // node types (host copy)
  long long n_fluid = 0;
  KArgs base{};
};

namespace {

// Case buffers come from the device's stream-ordered memory pool when the
// device has one: creating and releasing a case then costs no device-wide
// synchronisation (hipFree waits for the whole device), which matters when a
// server turns cases over while others run. The pool keeps released memory
// for reuse instead of returning it to the driver.
bool use_pool(int device) {
  static std::once_flag once[64];
  static bool ok[64] = {};
  if (device < 0 || device >= 64) return false;
  std::call_once(once[device], [device] {
    int supported = 0;
    if (hipDeviceGetAttribute(&supported, hipDeviceAttributeMemoryPoolsSupported, device) !=
            hipSuccess ||
        !supported)
      return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) != hipSuccess) return;
    uint64_t keep = UINT64_MAX;
    ok[device] = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess;
  });
  return ok[device];
}

hipError_t dev_alloc(lbf_case* c, void** p, size_t bytes) {
  if (c->stream && use_pool(c->device)) return hipMallocAsync(p, bytes, c->stream);
  return hipMalloc(p, bytes);
}

void drop_step_graphs(lbf_case* c) {
  for (int i = 0; i < c->ngraphs; i++) (void)hipGraphExecDestroy(c->graphs[i].exec);
  c->ngraphs = 0;
}

void free_case(lbf_case* c) {
  if (!c) return;
  drop_step_graphs(c);
  const bool pooled = c->stream && use_pool(c->device);
  auto F = [&](void* p) {
    if (!p) return;
    if (pooled)
      (void)hipFreeAsync(p, c->stream);
    else
      (void)hipFree(p);
  };
  F(c->pop[0]), F(c->pop[1]), F(c->cls), F(c->rho[0]), F(c->rho[1]), F(c->u[0]), F(c->u[1]),
      F(c->force), F(c->ctrl), F(c->partials), F(c->readback), F(c->red_out), F(c->phi);
  for (int s = 0; s < c->ns; s++) {
    ScalarDev& S = c->sc[s];
    F(S.pop[0]), F(S.pop[1]), F(S.c[0]), F(S.c[1]), F(S.q_ads), F(S.src);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int check_collision(const VsInfo& v, const lbf_collision& col, const char* what) {
  if (col.op == LBF_BGK) {
    if (!(col.tau > 0)) return fail(LBF_EINVAL, "%s: BGK tau must be > 0", what);
  } else if (col.op == LBF_TRT) {
  } else if (col.op == LBF_MRT) {
    if (!v.mrt) return fail(LBF_EUNSUPPORTED, "%s: MRT not available for this velocity set", what);
  } else {
    return fail(LBF_EINVAL, "%s: unknown collision operator %d", what, col.op);
  }
  return 0;
}

// Zou-He combinations accepted by src/velocity_set/d2q9.rs:173-242
bool zou_he_ok(int face, unsigned known) {
  const bool hd = known & LBF_KNOWN_DENSITY, hx = known & LBF_KNOWN_UX, hy = known & LBF_KNOWN_UY;
  if (face == LBF_WEST || face == LBF_EAST) return (!hd && hx && hy) || (hd && !hx && hy);
  if (face == LBF_SOUTH || face == LBF_NORTH) return (!hd && hx && hy) || (hd && hx && !hy);
  return false;
}

// this engine's own 3-D closure: all velocity components, or the density and
// the tangential components
bool zou_he_3d_ok(int face, unsigned known, int D) {
  const unsigned bits[3] = {LBF_KNOWN_UX, LBF_KNOWN_UY, LBF_KNOWN_UZ};
  const int a = face / 2;
  bool tang = true;
  for (int t = 0; t < D; t++)
    if (t != a && !(known & bits[t])) tang = false;
  if (!(known & LBF_KNOWN_DENSITY)) return tang && (known & bits[a]);
  return tang && !(known & bits[a]);
}

void fill_collision(KArgs& a, const lbf_collision& col, int q) {
  a.op = col.op;
  a.omega = 1.0 / col.tau;        // DELTA_T / tau, src/kernel.rs:35
  a.omega_p = 1.0 - a.omega;      // src/kernel.rs:36
  a.coeff_b = 1.0 - 0.5 / col.tau;  // src/kernel.rs:131 and :163
  a.wp = col.omega_plus;
  a.wm = col.omega_minus;
  for (int k = 0; k < 27; k++) a.s[k] = k < q ? col.relaxation[k] : 0.0;
}

template <class T>
void to_compute(const double* in, size_t n, std::vector<char>& out) {
  out.resize(n * sizeof(T));
  T* o = reinterpret_cast<T*>(out.data());
  for (size_t i = 0; i < n; i++) o[i] = static_cast<T>(in[i]);
}

// Host <-> device copy in the case's stream order, waited for. No copy goes
// through the legacy null stream, so a case never waits for work on other
// streams (a server's graph replays, other cases).
hipError_t copy_sync(lbf_case* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}

int upload_compute(lbf_case* c, void* dst, const double* src, size_t n) {
  std::vector<char> buf;
  if (c->tb == 8)
    to_compute<double>(src, n, buf);
  else
    to_compute<float>(src, n, buf);
  HIPCHK(copy_sync(c, dst, buf.data(), buf.size(), hipMemcpyHostToDevice));
  return 0;
}

int download_compute(lbf_case* c, const void* src, double* out, size_t n) {
  std::vector<char> buf(n * c->tb);
  HIPCHK(copy_sync(c, buf.data(), src, buf.size(), hipMemcpyDeviceToHost));
  if (c->tb == 8)
    memcpy(out, buf.data(), n * 8);
  else
    for (size_t i = 0; i < n; i++) out[i] = reinterpret_cast<const float*>(buf.data())[i];
  return 0;
}

int sync_and_check(lbf_case* c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int read_ctrl(lbf_case* c, Ctrl* h) {
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(copy_sync(c, h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  c->host_step = h->step;
  return 0;
}

int validate(const lbf_case_desc* d, VsInfo* v) {
  if (!d) return fail(LBF_EINVAL, "descriptor is NULL");
  if (d->abi_version != LBF_ABI_VERSION)
    return fail(LBF_EINVAL, "ABI version %u, library expects %u", d->abi_version, LBF_ABI_VERSION);
  if (!vs_info(d->velocity_set, v))
    return fail(LBF_EUNSUPPORTED, "velocity set %d is not constructible", d->velocity_set);
  for (int a = 0; a < v->d; a++)
    if (d->n[a] < 1) return fail(LBF_EINVAL, "lattice dimension %d must be >= 1", a);
  int rc = check_collision(*v, d->collision, "momentum");
  if (rc) return rc;
  const int nf = 2 * v->d;
  for (int f = 0; f < nf; f++) {
    const int k = d->faces[f].kind;
    if (k < LBF_BC_NO_SLIP || k > LBF_BC_PERIODIC)
      return fail(LBF_EINVAL, "face %d has no boundary condition", f);
    if (k == LBF_BC_ZOU_HE) {
      if (d->velocity_set != LBF_D2Q9) {
        if (!(d->flags & LBF_CASE_ZOU_HE_3D))
          return fail(LBF_EUNSUPPORTED, "Zou & He BC computation function not defined.");
        if (!zou_he_3d_ok(f, d->faces[f].known, v->d))
          return fail(LBF_EINVAL, "Zou-He face %d: give u (velocity face) or rho and the "
                                  "tangential u (pressure face)", f);
      } else if (!zou_he_ok(f, d->faces[f].known)) {
        return fail(LBF_EINVAL, "Error: Wrong input of Zou & He (1997) boundary conditions.");
      }
    }
  }
  if (d->precision < LBF_F64 || d->precision > LBF_BF16_SHIFTED)
    return fail(LBF_EINVAL, "unknown precision %d", d->precision);
  if (d->n_scalars > LBF_MAX_SCALARS)
    return fail(LBF_EUNSUPPORTED, "at most %d scalars", LBF_MAX_SCALARS);
  if (d->n_scalars && !d->scalars) return fail(LBF_EINVAL, "scalars pointer is NULL");
  if (d->force_kind == LBF_SOURCE_FIELD && !d->force_field)
    return fail(LBF_EINVAL, "force field pointer is NULL");
  for (uint32_t s = 0; s < d->n_scalars; s++) {
    const lbf_scalar_desc& S = d->scalars[s];
    if (S.velocity_set != d->velocity_set)
      return fail(LBF_EUNSUPPORTED, "scalar %u: velocity set must equal the momentum one", s);
    rc = check_collision(*v, S.collision, "scalar");
    if (rc) return rc;
    for (int f = 0; f < nf; f++) {
      const int k = S.faces[f].kind;
      if (k < LBF_SBC_ANTI_BOUNCE_BACK || k > LBF_SBC_PERIODIC)
        return fail(LBF_EINVAL, "scalar %u face %d has no boundary condition", s, f);
    }
    if (S.source_kind == LBF_SOURCE_FIELD && !S.source_field)
      return fail(LBF_EINVAL, "scalar %u source field pointer is NULL", s);
  }
  return 0;
}

}  // namespace

extern "C" {

const char* lbf_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// Plane stride (elements) of every per-node array: N rounded up to a whole
// wavefront, plus an optional pad set by lbfx_set_plane_pad for cases created
// afterwards (0 by default).
static std::atomic<long long> g_plane_pad{0};
static long long plane_stride(long long n_nodes) {
  return (n_nodes + 63) / 64 * 64 + g_plane_pad.load(std::memory_order_relaxed);
}

extern "C" {

int lbf_device_bytes(const lbf_case_desc* d, uint64_t* bytes) {
  VsInfo v;
  int rc = validate(d, &v);
  if (rc) return rc;
  const long long N = (long long)d->n[0] * d->n[1] * (v.d == 3 ? d->n[2] : 1);
  const long long ld = plane_stride(N);
  const size_t sb = storage_bytes(d->precision), tb = d->precision == LBF_F64 ? 8 : 4;
  const bool aa_ok = aa_compatible(d, v);
  const bool forced =
      (d->precision >= LBF_F32_SHIFTED && d->streaming != LBF_STREAM_TWO_STEP) ||
      d->streaming == LBF_STREAM_AA;
  bool aa = forced || (d->streaming == LBF_STREAM_AUTO && aa_ok &&
                       (d->precision != LBF_F64 || aa_reproduces_reference(d, v)));
  // as lbf_create: the automatic choice falls back to two arrays at 4 GiB
  if (aa && !forced && (unsigned long long)ld * v.q * sb >= (1ull << 32)) aa = false;
  uint64_t b = (aa ? 1ull : 2ull) * v.q * ld * sb + 4ull * ld + 2ull * (1 + v.d) * ld * tb;
  if (d->force_kind == LBF_SOURCE_FIELD) b += 8ull * v.d * ld;
  if (d->multiphase) b += (uint64_t)ld * tb;
  for (uint32_t s = 0; s < d->n_scalars; s++) {
    b += 2ull * v.q * ld * tb + 2ull * ld * tb;
    if (d->scalars[s].has_adsorption) b += 8ull * ld;
    if (d->scalars[s].source_kind == LBF_SOURCE_FIELD) b += 8ull * ld;
  }
  *bytes = b + sizeof(Ctrl) + (uint64_t)partial_rows((N + kBlock - 1) / kBlock) * kMaxCols * 8;
  return 0;
}

int lbf_create(const lbf_case_desc* d, lbf_case** out) {
  if (!out) return fail(LBF_EINVAL, "out is NULL");
  *out = nullptr;
  VsInfo v;
  int rc = validate(d, &v);
  if (rc) return rc;
  lbf_case* c = new lbf_case();
  c->desc = *d;
  c->v = v;
  c->d = v.d, c->q = v.q;
  c->nx = (int)d->n[0], c->ny = (int)d->n[1], c->nz = v.d == 3 ? (int)d->n[2] : 1;
  c->N = (long long)c->nx * c->ny * c->nz;
  c->ld = plane_stride(c->N);
  c->prec = d->precision;
  c->sb = storage_bytes(c->prec);
  c->tb = c->prec == LBF_F64 ? 8 : 4;
  c->frozen = (d->flags & LBF_CASE_FREEZE_MOMENTUM) != 0;
  for (int i = 0; i < c->q; i++) c->w[i] = lattice_weight(d->velocity_set, v.c[i]);
  {
    const bool shifted = c->prec >= LBF_F32_SHIFTED;
    const bool aa_ok = aa_compatible(d, v);
    if (d->streaming == LBF_STREAM_AA && !aa_ok) {
      delete c;
      return fail(LBF_EUNSUPPORTED, "AA streaming needs no scalars and no Shan-Chen force, "
                                    "face conditions in {no-slip, bounce-back, periodic, "
                                    "Zou-He} and matching periodicity of opposite faces");
    }
    if (shifted && !aa_ok) {
      delete c;
      return fail(LBF_EUNSUPPORTED, "shifted storage runs on the AA or two-step kernels only "
                                    "(no scalars, no Shan-Chen force, faces in {no-slip, "
                                    "bounce-back, periodic, Zou-He})");
    }
    if (d->streaming == LBF_STREAM_TWO_STEP) {
      if (!tb_compatible(d, v)) {
        delete c;
        return fail(LBF_EUNSUPPORTED, "two-step passes need a 3-D lattice, no scalars, no "
                                    "Shan-Chen force and faces in {no-slip, bounce-back, "
                                    "periodic} with matching periodicity of opposite faces");
      }
      c->two_step = true;
    }
    if (d->streaming == LBF_STREAM_AA) c->aa = true;
    else if (d->streaming == LBF_STREAM_AUTO)
      c->aa = aa_ok && (c->prec != LBF_F64 || aa_reproduces_reference(d, v));
    if (shifted && !c->two_step) c->aa = true;
    if (c->aa && (unsigned long long)c->ld * c->q * c->sb >= (1ull << 32)) {
      if (shifted || d->streaming == LBF_STREAM_AA) {
        delete c;
        return fail(LBF_EUNSUPPORTED, "AA population array must stay below 4 GiB");
      }
      c->aa = false;
    }
    if (shifted) c->scale = d->shift_scale > 0 ? d->shift_scale
                                               : (c->prec == LBF_F16_SHIFTED ? 4096.0 : 1.0);
  }
  c->ns = (int)d->n_scalars;
  if (!get_launchers(d->velocity_set, c->prec, &c->L)) {
    delete c;
    return fail(LBF_EUNSUPPORTED, "precision %d is not built for velocity set %d", d->precision,
                d->velocity_set);
  }
  const int nf = 2 * c->d;
  for (int f = 0; f < nf; f++)
    if (d->faces[f].kind != LBF_BC_PERIODIC) c->face_active |= 1u << f;
  for (int f = 0; f < nf; f++)
    if (d->faces[f].kind == LBF_BC_ANTI_BOUNCE_BACK) c->needs_moments = true;
  if (c->ns) c->needs_moments = true;

  // ---- node types (the classes themselves are built on the device)
  const long long N = c->N;
  const int q = c->q;
  c->h_type.assign(N, 0);
  if (d->node_type) {
    for (long long i = 0; i < N; i++)
      if (d->node_type[i] > 1) {
        free_case(c);
        return fail(LBF_EINVAL, "Invalid node type at node %lld", i);
      }
    memcpy(c->h_type.data(), d->node_type, N);
  }
  c->n_fluid = 0;
  for (long long i = 0; i < N; i++) c->n_fluid += c->h_type[i] == 0;

  // ---- device memory
  auto cleanup = [&](int code) {
    free_case(c);
    return code;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return cleanup(fail(LBF_EDEVICE, "no HIP device"));
  const long long blocks = (N + kBlock - 1) / kBlock;
  // one node per thread: the streaming kernels have no loop-carried state and
  // every workgroup writes its own residual partials
  c->grid = (int)std::max(1LL, blocks);
  // in-place kernel: runs of 4 node blocks per XCD (neighbouring rows then
  // share an L2); 256^3 D3Q19 FP32 measured 570 -> 535 us per step. The
  // two-array kernels keep the round-robin order: there runs of 1-32 blocks
  // cut HBM reads by up to a fifth yet cost 0-6 % in time (measured).
  c->xcd_group = (c->aa && c->grid % 32 == 0) ? 4 : 0;
  {
    const char* env = getenv("LBF_STEP_GRAPHS");
    c->graphs_on = !c->two_step && N <= kGraphMaxNodes && !(env && env[0] == '0');
  }
  if (c->two_step) {
    if ((unsigned long long)c->ld * q * c->sb >= (1ull << 32) || !c->L.tb_step) {
      free_case(c);
      return fail(LBF_EUNSUPPORTED, "two-step passes need population arrays below 4 GiB");
    }
    // tiles of tb_tx x tb_ty columns (shape 1: rows of 64 nodes where the
    // ring fits; LBF_TB_SHAPE=0 takes 32 x 16 tiles); the z axis is cut into
    // segments so the grid holds about one workgroup per CU (each holds most
    // of a CU's LDS): 256^3 D3Q19 FP32 measured 428 us per step with 256
    // workgroups, 450-460 us with 512 or 1024 (more ring planes re-read and
    // a second round); LBF_TB_WORKGROUPS sets another count
    const char* es = getenv("LBF_TB_SHAPE");
    c->ts_shape = es && es[0] == '0' ? 0 : 1;
    const int tx = c->L.tb_tx[c->ts_shape], ty = c->L.tb_ty[c->ts_shape];
    const long long tiles = (long long)((c->nx + tx - 1) / tx) * ((c->ny + ty - 1) / ty);
    const char* ew = getenv("LBF_TB_WORKGROUPS");
    const char* ez = getenv("LBF_TB_ZS");
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        cus <= 0)
      cus = 256;
    const long long want = ew && atoll(ew) > 0 ? atoll(ew) : cus;
    long long zs = (c->nz * tiles + want - 1) / want;
    if (ez && atoi(ez) > 0) zs = atoi(ez);
    zs = std::max(1LL, std::min<long long>(zs, c->nz));
    c->ts_zs = (int)zs;
    c->ts_grid = (int)(tiles * ((c->nz + zs - 1) / zs));
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(LBF_EDEVICE, "hipStreamCreate failed"));
  // buffers are cleared on the case stream; the stream is synchronised
  // before any synchronous copy writes into them
  auto alloc = [&](void** p, size_t bytes) -> int {
    if (dev_alloc(c, p, bytes) != hipSuccess)
      return fail(LBF_ENOMEM, "device allocation of %zu bytes failed", bytes);
    if (hipMemsetAsync(*p, 0, bytes, c->stream) != hipSuccess)
      return fail(LBF_EDEVICE, "hipMemsetAsync failed");
    return 0;
  };
  // for buffers that synchronous copies fill right away
  auto alloc_sync = [&](void** p, size_t bytes) -> int {
    int r = alloc(p, bytes);
    if (!r && hipStreamSynchronize(c->stream) != hipSuccess)
      r = fail(LBF_EDEVICE, "hipStreamSynchronize failed");
    return r;
  };
  const size_t pop_bytes = (size_t)q * c->ld * c->sb;
  for (int b = 0; b < 2; b++) {
    if ((b == 0 || !c->aa) && (rc = alloc(&c->pop[b], pop_bytes))) return cleanup(rc);
    if ((rc = alloc(&c->rho[b], c->ld * c->tb))) return cleanup(rc);
    if ((rc = alloc(&c->u[b], (size_t)c->d * c->ld * c->tb))) return cleanup(rc);
  }
  if ((rc = alloc((void**)&c->cls, c->ld * 4))) return cleanup(rc);
  {
    // node classes on the device; node_type is staged in rho[1], which the
    // initial moments overwrite afterwards on the same stream
    ClassArgs ca{};
    ca.nx = c->nx, ca.ny = c->ny, ca.nz = c->nz, ca.q = q;
    ca.N = N, ca.ld = c->ld;
    for (int k = 0; k < q; k++)
      for (int a = 0; a < 3; a++) ca.c[k][a] = a < c->d ? v.c[k][a] : 0;
    const uint8_t* type_dev = nullptr;
    if (d->node_type) {
      if (hipMemcpyAsync(c->rho[1], d->node_type, N, hipMemcpyHostToDevice, c->stream) !=
          hipSuccess)
        return cleanup(fail(LBF_EDEVICE, "upload of node types failed"));
      type_dev = (const uint8_t*)c->rho[1];
    }
    if (launch_build_classes(ca, type_dev, c->cls, c->stream) != hipSuccess)
      return cleanup(fail(LBF_EDEVICE, "building node classes failed"));
  }
  if ((rc = alloc((void**)&c->ctrl, sizeof(Ctrl)))) return cleanup(rc);
  if (launch_ctrl_set(c->ctrl, kCtrlMaxIter, -1, c->stream) != hipSuccess)  // max_iter = ~0
    return cleanup(fail(LBF_EDEVICE, "control block initialisation failed"));
  if ((rc = alloc((void**)&c->partials, (size_t)partial_rows(c->grid) * kMaxCols * 8)))
    return cleanup(rc);
  if (d->multiphase && (rc = alloc(&c->phi, c->ld * c->tb))) return cleanup(rc);
  if ((rc = alloc((void**)&c->red_out, 64))) return cleanup(rc);

  // initial moments into both buffers (solid nodes keep them forever)
  if (!d->initial_density && !d->initial_velocity) {
    bool ok = true;
    for (int b = 0; b < 2; b++) {
      ok = ok && launch_fill(c->rho[b], c->tb == 8, N, c->ld, d->initial_density_uniform,
                             c->stream) == hipSuccess;
      for (int a = 0; a < c->d; a++)
        ok = ok && launch_fill((char*)c->u[b] + (size_t)a * c->ld * c->tb, c->tb == 8, N, c->ld,
                               d->initial_velocity_uniform[a], c->stream) == hipSuccess;
    }
    if (!ok) return cleanup(fail(LBF_EDEVICE, "initial moments failed"));
  } else {
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return cleanup(fail(LBF_EDEVICE, "case set-up failed"));
    std::vector<double> rho(c->ld, 0.0), u((size_t)c->d * c->ld, 0.0);
    for (long long i = 0; i < N; i++) {
      rho[i] = d->initial_density ? d->initial_density[i] : d->initial_density_uniform;
      for (int a = 0; a < c->d; a++)
        u[a * c->ld + i] =
            d->initial_velocity ? d->initial_velocity[i * c->d + a] : d->initial_velocity_uniform[a];
    }
    for (int b = 0; b < 2; b++) {
      if ((rc = upload_compute(c, c->rho[b], rho.data(), c->ld))) return cleanup(rc);
      if ((rc = upload_compute(c, c->u[b], u.data(), (size_t)c->d * c->ld))) return cleanup(rc);
    }
  }
  // everything above ran on the case stream; later uploads are synchronous
  if (hipStreamSynchronize(c->stream) != hipSuccess)
    return cleanup(fail(LBF_EDEVICE, "case set-up failed"));
  if (d->force_kind == LBF_SOURCE_FIELD) {
    std::vector<double> F((size_t)c->d * c->ld, 0.0);
    for (long long i = 0; i < N; i++)
      for (int a = 0; a < c->d; a++) F[a * c->ld + i] = d->force_field[i * c->d + a];
    if ((rc = alloc_sync((void**)&c->force, F.size() * 8))) return cleanup(rc);
    if (copy_sync(c, c->force, F.data(), F.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(LBF_EDEVICE, "upload of force field failed"));
  }
  for (int s = 0; s < c->ns; s++) {
    ScalarDev& S = c->sc[s];
    S.desc = d->scalars[s];
    S.desc.initial_value = nullptr;
    S.desc.source_field = nullptr;
    const lbf_scalar_desc& sd = d->scalars[s];
    for (int b = 0; b < 2; b++) {
      if ((rc = alloc_sync(&S.pop[b], (size_t)q * c->ld * c->tb))) return cleanup(rc);
      if ((rc = alloc_sync(&S.c[b], c->ld * c->tb))) return cleanup(rc);
    }
    std::vector<double> cv(c->ld, 0.0);
    for (long long i = 0; i < N; i++)
      cv[i] = sd.initial_value ? sd.initial_value[i] : sd.initial_value_uniform;
    for (int b = 0; b < 2; b++)
      if ((rc = upload_compute(c, S.c[b], cv.data(), c->ld))) return cleanup(rc);
    if (sd.has_adsorption) {
      // q_ads init: initial value on momentum bounce-back nodes, else 0 (node.rs:77-85)
      if ((rc = alloc_sync((void**)&S.q_ads, c->ld * 8))) return cleanup(rc);
      if (launch_qads_init(c->cls, S.q_ads, c->ld, sd.initial_q_ads, c->stream) != hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess)
        return cleanup(fail(LBF_EDEVICE, "q_ads initialisation failed"));
    }
    if (sd.source_kind == LBF_SOURCE_FIELD) {
      std::vector<double> sf(c->ld, 0.0);
      for (long long i = 0; i < N; i++) sf[i] = sd.source_field[i];
      if ((rc = alloc_sync((void**)&S.src, c->ld * 8))) return cleanup(rc);
      if (copy_sync(c, S.src, sf.data(), c->ld * 8, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(LBF_EDEVICE, "upload of scalar source failed"));
    }
  }

  // ---- base kernel arguments
  KArgs& a = c->base;
  a.nx = c->nx, a.ny = c->ny, a.nz = c->nz, a.d = c->d;
  a.N = N, a.ld = c->ld;
  a.nfaces = nf;
  a.face_active = c->face_active;
  a.cls = c->cls;
  a.rho[0] = c->rho[0], a.rho[1] = c->rho[1];
  a.u[0] = c->u[0], a.u[1] = c->u[1];
  a.partials = c->partials;
  a.ctrl = c->ctrl;
  if (d->multiphase) {
    a.multiphase = 1;
    a.sc_g = d->g_function;
    a.wall_phi = 1.0 * (1.0 - std::exp(-d->wall_density / 1.0));  // multiphase.rs:10-18
    a.phi = c->phi;
  }
  for (int f = 0; f < 6; f++) {
    a.face[f].kind = f < nf ? d->faces[f].kind : LBF_BC_PERIODIC;
    a.face[f].known = d->faces[f].known;
    a.face[f].density = d->faces[f].density;
    for (int x = 0; x < 3; x++) a.face[f].vel[x] = d->faces[f].velocity[x];
  }
  c->desc.node_type = nullptr;
  c->desc.force_field = nullptr;
  c->desc.initial_density = nullptr;
  c->desc.initial_velocity = nullptr;
  c->desc.scalars = nullptr;
  *out = c;
  return 0;
}

void lbf_destroy(lbf_case* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  free_case(c);
}

}  // extern "C"

namespace {

// in-place pattern of the next step: 1 = odd
int aa_phase(const lbf_case* c) { return (int)((c->host_step + (uint64_t)c->aa_shift) & 1); }

KArgs momentum_args(lbf_case* c) {
  KArgs a = c->base;
  a.field = 0;
  a.pop[0] = c->pop[0], a.pop[1] = c->pop[1];
  fill_collision(a, c->desc.collision, c->q);
  a.force_kind = c->desc.force_kind;
  for (int x = 0; x < 3; x++) a.F[x] = c->desc.force[x];
  a.src_field = c->force;
  a.scale = c->scale;
  a.plane_bytes = (uint32_t)(c->ld * c->sb);
  a.fresh = c->aa_fresh ? 1 : 0;
  a.aa_odd = aa_phase(c);
  a.xcd_group = c->aa ? c->xcd_group : 0;
  a.tb_zs = c->ts_zs;
  a.tb_shape = c->ts_shape;
  a.inv_scale = 1.0 / c->scale;
  a.faces_noslip_only = 1;
  for (int f = 0; f < 2 * c->d; f++)
    if (c->desc.faces[f].kind != LBF_BC_NO_SLIP && c->desc.faces[f].kind != LBF_BC_PERIODIC)
      a.faces_noslip_only = 0;
  return a;
}

KArgs scalar_args(lbf_case* c, int s) {
  KArgs a = c->base;
  ScalarDev& S = c->sc[s];
  a.xcd_group = 0;
  a.field = 1 + s;
  a.pop[0] = S.pop[0], a.pop[1] = S.pop[1];
  a.c[0] = S.c[0], a.c[1] = S.c[1];
  fill_collision(a, S.desc.collision, c->q);
  a.force_kind = S.desc.source_kind;  // reused as the scalar source kind
  a.src_uniform = S.desc.source_uniform;
  a.src_field = S.src;
  a.has_ads = S.desc.has_adsorption;
  a.k_a = S.desc.adsorption_rate, a.k_d = S.desc.desorption_rate, a.q_max = S.desc.q_max;
  a.q_ads = S.q_ads;
  a.inner_bc = S.desc.inner_bc;
  uint32_t act = 0;
  for (int f = 0; f < 2 * c->d; f++) {
    a.sface[f].kind = S.desc.faces[f].kind;
    a.sface[f].unknown_only = S.desc.faces[f].unknown_populations_only;
    a.sface[f].value = S.desc.faces[f].scalar_value;
    if (S.desc.faces[f].kind != LBF_SBC_PERIODIC) act |= 1u << f;
  }
  a.face_active = act;
  return a;
}

uint32_t residual_mask(lbf_case* c) {
  uint32_t m = 0;
  if (!c->frozen) {
    m |= 1u;
    for (int x = 0; x < c->d; x++) m |= 1u << (1 + x);
    m |= 1u << 8;
  }
  for (int s = 0; s < c->ns; s++) m |= 1u << (4 + s);
  return m;
}

// Advances the device step counter by the steps enqueued without their own
// advancing launch.
int flush_steps(lbf_case* c) {
  if (c->pending_steps) {
    HIPCHK(launch_ctrl_set(c->ctrl, kCtrlAddSteps, c->pending_steps, c->stream));
    c->pending_steps = 0;
  }
  return 0;
}

// Enqueues one time step: momentum launch, then one launch per scalar; the
// last launch advances the step. A step without residuals or the stop rule
// has nothing for that launch to reduce: its launches read the counter plus
// the steps enqueued since it last advanced (step_offset), and flush_steps
// advances the counter once for the batch.
int enqueue_step(lbf_case* c, uint32_t flags, bool last_of_batch) {
  const bool resid = flags & LBF_STEP_RESIDUALS;
  const bool stop = flags & LBF_STEP_STOP_ON_CONVERGENCE;
  const bool defer = !resid && !stop;
  if (!defer) {
    int rc = flush_steps(c);
    if (rc) return rc;
  }
  uint32_t kf = 0;
  if (resid) kf |= kFlagResiduals;
  if (stop) kf |= kFlagStopCheck;
  const uint32_t mask = residual_mask(c);
  if (!c->frozen) {
    KArgs a = momentum_args(c);
    a.mode = kModePull;
    a.flags = kf & ~kFlagStopCheck;
    if (resid || last_of_batch || c->needs_moments) a.flags |= kFlagStoreMoments;
    a.res_mask = mask;
    a.step_offset = defer ? c->pending_steps : 0;
    if (c->aa) {
      HIPCHK(c->L.aa_step(a, c->grid, c->stream));
    } else {
      if (c->desc.multiphase) {
        // psi of every node of this step before any node's force reads it
        KArgs p = a;
        p.mode = kModePhi;
        HIPCHK(c->L.mom_step(p, c->grid, c->stream));
      }
      HIPCHK(c->L.mom_step(a, c->grid, c->stream));
    }
  }
  for (int s = 0; s < c->ns; s++) {
    KArgs a = scalar_args(c, s);
    a.mode = kModePull;
    a.flags = kf & ~kFlagStopCheck;
    a.res_mask = mask;
    a.step_offset = defer ? c->pending_steps : 0;
    HIPCHK(c->L.sca_step(a, c->grid, c->stream));
  }
  if (defer) {
    c->pending_steps++;
  } else {
    // one-workgroup launch: residual reduction, stop rule, step counter
    KArgs a = momentum_args(c);
    a.flags = kf;
    a.res_mask = mask;
    HIPCHK(launch_advance(a, c->grid, c->stream));
  }
  c->host_step++;
  if (!c->frozen) c->aa_fresh = false;
  return 0;
}

// The captured batch of kGraphSteps steps for these step flags (captured on
// first use from an even, non-fresh step: the in-place launches alternate by
// parity). Replays enqueue exactly the launches enqueue_step would.
int step_graph(lbf_case* c, uint32_t flags, bool every_last, hipGraphExec_t* out) {
  for (int i = 0; i < c->ngraphs; i++)
    if (c->graphs[i].flags == flags && c->graphs[i].every_last == every_last) {
      *out = c->graphs[i].exec;
      return 0;
    }
  if (c->ngraphs == (int)(sizeof(c->graphs) / sizeof(c->graphs[0]))) drop_step_graphs(c);
  const uint64_t step0 = c->host_step;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (uint32_t k = 0; k < kGraphSteps && !rc; k++) rc = enqueue_step(c, flags, every_last);
  if (!rc) rc = flush_steps(c);  // the graph advances the counter for its own steps
  c->pending_steps = 0;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->host_step = step0;  // capture enqueued nothing
  if (rc || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : fail(LBF_EDEVICE, "step capture failed: %s", hipGetErrorString(e));
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(LBF_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
  c->graphs[c->ngraphs++] = {flags, every_last, x};
  *out = x;
  return 0;
}

// nsteps steps: captured batches where the case replays them (from an even,
// non-fresh step, keeping the last step of a batch that stores moments out of
// them), single launches otherwise.
// One two-step pass (tb.hpp): steps without residuals or the stop rule, so
// both steps defer their counter advance. The pass writes the other array,
// and swapping the two pointers keeps the state in pop[step & 1].
int enqueue_pass(lbf_case* c, bool last_of_batch) {
  KArgs a = momentum_args(c);
  a.mode = kModePull;
  a.flags = last_of_batch ? kFlagStoreMoments : 0u;
  a.res_mask = residual_mask(c);
  a.step_offset = c->pending_steps;
  HIPCHK(c->L.tb_step(a, c->ts_grid, c->stream));
  std::swap(c->pop[0], c->pop[1]);
  c->pending_steps += 2;
  c->host_step += 2;
  c->aa_fresh = false;
  return 0;
}

int enqueue_steps(lbf_case* c, uint64_t nsteps, uint32_t flags, bool every_last) {
  // passes need both steps' launches deferred and at most the second step's moments
  const bool passes = c->two_step && !c->frozen && !c->needs_moments &&
                      !(flags & (LBF_STEP_RESIDUALS | LBF_STEP_STOP_ON_CONVERGENCE)) &&
                      !every_last;
  for (uint64_t k = 0; k < nsteps;) {
    const uint64_t left = nsteps - k;
    if (passes && left >= 2) {
      int rc = enqueue_pass(c, left == 2);
      if (rc) return rc;
      k += 2;
      continue;
    }
    if (c->graphs_on && !c->aa_fresh && aa_phase(c) == 0 &&
        (every_last ? left >= kGraphSteps : left > kGraphSteps)) {
      hipGraphExec_t g = nullptr;
      int rc = flush_steps(c);  // the graph's launches count from the device counter
      if (!rc) rc = step_graph(c, flags, every_last, &g);
      if (rc) return rc;
      HIPCHK(hipGraphLaunch(g, c->stream));
      c->host_step += kGraphSteps;
      k += kGraphSteps;
    } else {
      int rc = enqueue_step(c, flags, every_last || k + 1 == nsteps);
      if (rc) return rc;
      k++;
    }
  }
  return flush_steps(c);
}

}  // namespace

extern "C" {

int lbf_initialize_nodes(lbf_case* c) {
  if (!c) return fail(LBF_EINVAL, "case is NULL");
  if (c->may_have_stopped) {
    // the host step count may be ahead of a run that stopped early
    Ctrl h;
    int rc = read_ctrl(c, &h);
    if (rc) return rc;
    c->may_have_stopped = false;
  }
  c->aa_shift = (int)(c->host_step & 1);  // the fresh step takes the even pattern
  KArgs a = momentum_args(c);
  if (c->aa)
    HIPCHK(c->L.aa_init(a, c->grid, c->stream));
  else
    HIPCHK(c->L.mom_init(a, c->grid, c->stream));
  c->aa_fresh = true;
  long long fresh = 1;
  for (int s = 0; s < c->ns; s++) {
    KArgs b = scalar_args(c, s);
    HIPCHK(c->L.sca_init(b, c->grid, c->stream));
    fresh |= 1ll << (1 + s);
  }
  HIPCHK(launch_ctrl_set(c->ctrl, kCtrlFresh, fresh, c->stream));
  return 0;
}

int lbf_main_steps(lbf_case* c, uint32_t nsteps, uint32_t flags) {
  if (!c) return fail(LBF_EINVAL, "case is NULL");
  if (c->may_have_stopped) {
    // a previous call could have stopped early: take the step count from the device
    Ctrl h;
    int rc = read_ctrl(c, &h);
    if (rc) return rc;
    c->may_have_stopped = false;
  }
  HIPCHK(launch_ctrl_set(c->ctrl, kCtrlResetStop, 0, c->stream));
  if (flags & LBF_STEP_STOP_ON_CONVERGENCE) {
    HIPCHK(launch_ctrl_set(c->ctrl, kCtrlMaxIter, -1, c->stream));
    c->may_have_stopped = true;
  }
  return enqueue_steps(c, nsteps, flags, false);
}

int lbf_synchronize(lbf_case* c) {
  if (!c) return fail(LBF_EINVAL, "case is NULL");
  return sync_and_check(c);
}

int lbf_run(lbf_case* c, uint64_t max_iterations, uint64_t* steps_done) {
  if (!c) return fail(LBF_EINVAL, "case is NULL");
  HIPCHK(launch_ctrl_set(c->ctrl, kCtrlResetStop, 0, c->stream));
  HIPCHK(launch_ctrl_set(c->ctrl, kCtrlMaxIter, (long long)max_iterations, c->stream));
  const uint32_t flags = LBF_STEP_RESIDUALS | LBF_STEP_STOP_ON_CONVERGENCE;
  Ctrl h;
  int rc = read_ctrl(c, &h);
  if (rc) return rc;
  uint64_t chunk = 64;
  while (true) {
    // the stop rule fires at the latest at step index max_iterations + 1
    const uint64_t limit = max_iterations + 2;
    const uint64_t left = h.step >= limit ? 1 : limit - h.step;
    const uint64_t n = std::min<uint64_t>(chunk, left);
    if ((rc = enqueue_steps(c, n, flags, true))) return rc;
    if ((rc = read_ctrl(c, &h))) return rc;
    if (h.stopped) break;
    chunk = std::min<uint64_t>(chunk * 2, 1024);
  }
  if (steps_done) *steps_done = h.step;
  return 0;
}

int lbf_residuals(lbf_case* c, double* r_rho, double* r_u, double* r_c) {
  if (!c) return fail(LBF_EINVAL, "case is NULL");
  Ctrl h;
  int rc = read_ctrl(c, &h);
  if (rc) return rc;
  if (r_rho) *r_rho = h.res[0];
  if (r_u)
    for (int x = 0; x < c->d; x++) r_u[x] = h.res[1 + x];
  if (r_c)
    for (int s = 0; s < c->ns; s++) r_c[s] = h.res[4 + s];
  return 0;
}

int lbf_time_step(lbf_case* c, uint64_t* t) {
  if (!c || !t) return fail(LBF_EINVAL, "NULL argument");
  Ctrl h;
  int rc = read_ctrl(c, &h);
  if (rc) return rc;
  *t = h.step;
  return 0;
}

int lbf_reduce(lbf_case* c, int32_t which, double* out) {
  if (!c || !out) return fail(LBF_EINVAL, "NULL argument");
  Ctrl h;
  int rc = read_ctrl(c, &h);
  if (rc) return rc;
  const int cur = (int)((h.step + 1) & 1);
  const int kind = which & 0xff, s = which >> 8;
  double r = 0.0;
  switch (kind) {
    case LBF_REDUCE_MEAN_DENSITY:
    case LBF_REDUCE_TOTAL_MASS:
      HIPCHK(launch_reduce(c->rho[cur], c->tb == 8, c->cls, 0, c->N, c->partials, c->grid,
                           c->red_out, 0, c->stream));
      break;
    case LBF_REDUCE_SCALAR_TOTAL:
      if (s >= c->ns) return fail(LBF_EINVAL, "no scalar %d", s);
      HIPCHK(launch_reduce(c->sc[s].c[cur], c->tb == 8, c->cls, 0, c->N, c->partials, c->grid,
                           c->red_out, 0, c->stream));
      break;
    case LBF_REDUCE_ADSORBED:
      if (s >= c->ns) return fail(LBF_EINVAL, "no scalar %d", s);
      if (!c->sc[s].q_ads) {
        *out = 0.0;
        return 0;
      }
      HIPCHK(launch_reduce(c->sc[s].q_ads, 1, c->cls, 1, c->N, c->partials, c->grid, c->red_out,
                           0, c->stream));
      break;
    case LBF_REDUCE_NONFINITE:
      HIPCHK(launch_nonfinite(c->rho[cur], c->u[cur], c->tb == 8, c->d, c->ld, c->cls, c->N,
                              c->partials, c->grid, c->red_out, c->stream));
      break;
    default:
      return fail(LBF_EINVAL, "unknown reduction %d", which);
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(copy_sync(c, &r, c->red_out, 8, hipMemcpyDeviceToHost));
  if (kind == LBF_REDUCE_MEAN_DENSITY) r = r / (double)c->n_fluid;
  *out = r;
  return 0;
}

int lbf_get_fields(lbf_case* c, double* rho, double* u) {
  if (!c) return fail(LBF_EINVAL, "case is NULL");
  Ctrl h;
  int rc = read_ctrl(c, &h);
  if (rc) return rc;
  const int cur = (int)((h.step + 1) & 1);
  if (rho) {
    std::vector<double> r(c->ld);
    if ((rc = download_compute(c, c->rho[cur], r.data(), c->ld))) return rc;
    memcpy(rho, r.data(), c->N * 8);
  }
  if (u) {
    std::vector<double> uu((size_t)c->d * c->ld);
    if ((rc = download_compute(c, c->u[cur], uu.data(), uu.size()))) return rc;
    for (long long i = 0; i < c->N; i++)
      for (int a = 0; a < c->d; a++) u[i * c->d + a] = uu[a * c->ld + i];
  }
  return 0;
}

int lbf_get_scalar(lbf_case* c, uint32_t s, double* value, double* q_ads) {
  if (!c) return fail(LBF_EINVAL, "case is NULL");
  if ((int)s >= c->ns) return fail(LBF_EINVAL, "no scalar %u", s);
  Ctrl h;
  int rc = read_ctrl(c, &h);
  if (rc) return rc;
  const int cur = (int)((h.step + 1) & 1);
  if (value) {
    std::vector<double> v(c->ld);
    if ((rc = download_compute(c, c->sc[s].c[cur], v.data(), c->ld))) return rc;
    memcpy(value, v.data(), c->N * 8);
  }
  if (q_ads) {
    if (c->sc[s].q_ads) {
      HIPCHK(copy_sync(c, q_ads, c->sc[s].q_ads, c->N * 8, hipMemcpyDeviceToHost));
    } else {
      for (long long i = 0; i < c->N; i++) q_ads[i] = NAN;
    }
  }
  return 0;
}

int lbf_get_pops(lbf_case* c, int32_t field, double* out) {
  if (!c || !out) return fail(LBF_EINVAL, "NULL argument");
  if (field < 0 || field > c->ns) return fail(LBF_EINVAL, "no field %d", field);
  const size_t bytes = (size_t)c->q * c->ld * 8;
  {
    Ctrl h;  // synchronises and resynchronises the host step count
    int rc = read_ctrl(c, &h);
    if (rc) return rc;
  }
  if (!c->readback) HIPCHK(dev_alloc(c, (void**)&c->readback, bytes));
  HIPCHK(hipMemsetAsync(c->readback, 0, bytes, c->stream));
  KArgs a = field == 0 ? momentum_args(c) : scalar_args(c, field - 1);
  a.mode = kModeStreamOnly;
  a.flags = 0;
  a.out = c->readback;
  if (field == 0 && c->aa)
    HIPCHK(c->L.aa_readback(a, c->grid, c->stream));
  else if (field == 0)
    HIPCHK(c->L.mom_step(a, c->grid, c->stream));
  else
    HIPCHK(c->L.sca_step(a, c->grid, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<double> soa((size_t)c->q * c->ld);
  HIPCHK(copy_sync(c, soa.data(), c->readback, bytes, hipMemcpyDeviceToHost));
  for (long long i = 0; i < c->N; i++)
    for (int j = 0; j < c->q; j++) out[i * c->q + j] = soa[(size_t)j * c->ld + i];
  return 0;
}

}  // extern "C"

// ---- read-back without allocation (lbf_get_pops_noalloc)
// Device scratch of the allocation-free read-back: one copy per GPU, held
// until the process ends. Host code never passes this name to a kernel or a
// copy (on the host it is only a stand-in for the symbol); it uses the
// per-device address hipGetSymbolAddress returns.
static constexpr size_t kNoAllocScratchBytes = 16u << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / 8];

namespace {
std::mutex g_scratch_mu;          // serialises use of each device's array
double* g_scratch_ptr[64] = {};   // per-device address of `scratch`

// Address of this device's scratch array, looked up on its first use with
// the device current. Returns 0 or an error code.
int scratch_address(int device, double** out) {
  if (device < 0 || device >= 64) return fail(LBF_EDEVICE, "device %d out of range", device);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) HIPCHK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(LBF_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[device] = (double*)p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}
}  // namespace

extern "C" {

int lbf_get_pops_noalloc(lbf_case* c, int32_t field, double* out) {
  if (!c || !out) return fail(LBF_EINVAL, "NULL argument");
  if (field < 0 || field > c->ns) return fail(LBF_EINVAL, "no field %d", field);
  // nodes per chunk: Q f64 values per node must fit the scratch array
  const long long chunk = (long long)(kNoAllocScratchBytes / (8ull * (unsigned)c->q));
  if (chunk < 1)
    return fail(LBF_EINVAL, "%d populations per node exceed the %zu-byte scratch array", c->q,
                kNoAllocScratchBytes);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  double* dev = nullptr;
  int rc = scratch_address(c->device, &dev);
  if (rc) return rc;
  {
    Ctrl h;  // synchronises and resynchronises the host step count
    if ((rc = read_ctrl(c, &h))) return rc;
  }
  for (long long n0 = 0; n0 < c->N; n0 += chunk) {
    const long long n1 = std::min<long long>(c->N, n0 + chunk);
    KArgs a = field == 0 ? momentum_args(c) : scalar_args(c, field - 1);
    a.mode = kModeStreamOnly;
    a.flags = 0;
    a.xcd_group = 0;
    a.out = dev;
    a.out_aos = 1;
    a.out_n0 = n0;
    a.out_n1 = n1;
    // every node of the range writes its Q values (zeros on solid nodes), so
    // nothing has to be cleared first
    const int grid = (int)((n1 - n0 + kBlock - 1) / kBlock);
    if (field == 0 && c->aa)
      HIPCHK(c->L.aa_readback(a, grid, c->stream));
    else if (field == 0)
      HIPCHK(c->L.mom_step(a, grid, c->stream));
    else
      HIPCHK(c->L.sca_step(a, grid, c->stream));
    // the next chunk's launch follows this copy on the same stream
    HIPCHK(hipMemcpyAsync(out + n0 * c->q, dev, (size_t)(n1 - n0) * c->q * 8,
                          hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int lbf_set_pops(lbf_case* c, int32_t field, const double* in) {
  if (!c || !in) return fail(LBF_EINVAL, "NULL argument");
  if (field < 0 || field > c->ns) return fail(LBF_EINVAL, "no field %d", field);
  Ctrl h;
  int rc = read_ctrl(c, &h);
  if (rc) return rc;
  const int par = (int)(h.step & 1);
  std::vector<double> soa((size_t)c->q * c->ld, 0.0);
  for (long long i = 0; i < c->N; i++)
    for (int j = 0; j < c->q; j++) soa[(size_t)j * c->ld + i] = in[i * c->q + j];
  void* dst = field == 0 ? c->pop[c->aa ? 0 : par] : c->sc[field - 1].pop[par];
  const size_t elem = field == 0 ? c->sb : c->tb;
  std::vector<char> buf(soa.size() * elem);
  if (field == 0 && c->prec >= LBF_F32_SHIFTED) {
    for (int j = 0; j < c->q; j++)
      for (long long i = 0; i < c->ld; i++) {
        const size_t k = (size_t)j * c->ld + i;
        const double dv = i < c->N ? (soa[k] - c->w[j]) * c->scale : 0.0;
        store_shifted(c->prec, buf.data(), k, dv);
      }
  } else if (elem == 8) {
    memcpy(buf.data(), soa.data(), buf.size());
  } else {
    float* o = reinterpret_cast<float*>(buf.data());
    for (size_t i = 0; i < soa.size(); i++) o[i] = (float)soa[i];
  }
  HIPCHK(copy_sync(c, dst, buf.data(), buf.size(), hipMemcpyHostToDevice));
  HIPCHK(launch_ctrl_set(c->ctrl, kCtrlFresh, 1ll << field, c->stream));
  if (field == 0) {
    c->aa_fresh = true;
    c->aa_shift = (int)(c->host_step & 1);  // the fresh step takes the even pattern
  }
  return 0;
}

int lbf_get_link_tables(lbf_case* c, uint8_t* node_type, uint32_t* bb_link_mask,
                        uint8_t* face_bits) {
  if (!c) return fail(LBF_EINVAL, "case is NULL");
  if (node_type) memcpy(node_type, c->h_type.data(), c->N);
  if (bb_link_mask || face_bits) {
    std::vector<uint32_t> cls(c->ld);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(copy_sync(c, cls.data(), c->cls, c->ld * 4, hipMemcpyDeviceToHost));
    const long long n3[3] = {c->nx, c->ny, c->nz};
    for (long long i = 0; i < c->N; i++) {
      const bool fluid = !(cls[i] & kSolidBit);
      if (bb_link_mask) bb_link_mask[i] = fluid ? (cls[i] & kLinkMask) : 0u;
      if (face_bits) {
        const long long idx[3] = {i % c->nx, (i / c->nx) % c->ny, i / ((long long)c->nx * c->ny)};
        uint8_t fb = 0;
        for (int f = 0; fluid && f < 2 * c->d; f++) {
          const long long lim = (f % 2) ? n3[f / 2] - 1 : 0;
          if (idx[f / 2] == lim) fb |= (uint8_t)(1u << f);
        }
        face_bits[i] = fb;
      }
    }
  }
  return 0;
}

}  // extern "C"

// Tuning hook (not part of the reference-equivalent ABI): workgroup order of
// the in-place kernel, runs of g consecutive node blocks per XCD (0 = plain
// round-robin). Returns -1 when the grid is not a multiple of 8 g.
extern "C" int lbfx_set_block_group(lbf_case* c, int g) {
  if (!c || g < 0 || (g > 0 && c->grid % (8 * g) != 0)) return -1;
  c->xcd_group = g;
  drop_step_graphs(c);  // captured launches carry the block order
  return 0;
}

namespace {
template <class L>
void mrt_tables(double* M, double* Minv) {
  for (int k = 0; k < L::Q; k++)
    for (int i = 0; i < L::Q; i++) {
      M[k * L::Q + i] = Derived<L>::M(k, i);
      Minv[i * L::Q + k] = Derived<L>::Minv(i, k);
    }
}
}  // namespace

// Inspection hook (not part of the reference-equivalent ABI): the MRT
// matrices the kernels are built with (constexpr tables of lattice.hpp).
extern "C" int lbfx_mrt_matrices(int vs, double* M, double* Minv) {
  if (!M || !Minv) return -1;
  switch (vs) {
    case LBF_D2Q9: mrt_tables<D2Q9>(M, Minv); return 0;
    case LBF_D3Q15: mrt_tables<D3Q15>(M, Minv); return 0;
    case LBF_D3Q19: mrt_tables<D3Q19>(M, Minv); return 0;
    case LBF_D3Q27: mrt_tables<D3Q27>(M, Minv); return 0;
  }
  return -1;
}

// Tuning hook (not part of the reference-equivalent ABI): pad, in elements (a
// multiple of 64, at most 2^20), added to the plane stride of cases created
// after the call. Shifts the 19 (or 27) population planes against each other
// in the HBM address map. Returns -1 for an invalid pad.
extern "C" int lbfx_set_plane_pad(long long elements) {
  if (elements < 0 || elements % 64 != 0 || elements > (1ll << 20)) return -1;
  g_plane_pad.store(elements, std::memory_order_relaxed);
  return 0;
}

// Benchmark hook (not part of the reference-equivalent ABI): the HIP stream
// a case enqueues its launches on, so callers can bracket work with events.
extern "C" void* lbfx_case_stream(lbf_case* c) { return c ? (void*)c->stream : nullptr; }

// ---- internal interface for the serving tier (serve.hip) -----------------
namespace lbf {

int set_error(int code, const char* msg) { return fail(code, "%s", msg); }

int case_serve_info(lbf_case* c, CaseServeInfo* out) {
  if (!c || !out) return fail(LBF_EINVAL, "NULL argument");
  // the server derives each case's in-place pattern from its step count
  if (c->aa_shift)
    return fail(LBF_EUNSUPPORTED, "case restarted at an odd step: its in-place pattern is shifted");
  out->velocity_set = c->desc.velocity_set;
  out->precision = c->prec;
  out->op = c->desc.collision.op;
  out->aa = c->aa;
  out->n_scalars = c->ns;
  out->N = c->N;
  out->grid = c->grid;
  out->stream = c->stream;
  out->fresh = c->aa_fresh;
  out->step = c->host_step;
  out->args = momentum_args(c);
  out->args.flags = kFlagStoreMoments;
  out->args.mode = kModePull;
  return 0;
}

void case_note_external_steps(lbf_case* c, uint64_t steps) {
  if (!c || steps == 0) return;
  c->host_step += steps;
  c->aa_fresh = false;
}

}  // namespace lbf

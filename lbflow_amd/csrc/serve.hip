// serve.hip — serving tier: continuous batching of resident cases on one GPU.
//
// Three kinds of threads share a server:
//  - callers submit requests (descriptor copy + step count) into a FIFO;
//  - the worker thread admits requests while the HBM budget and slots allow
//    (first in first out, with backfill that never eats into what the oldest
//    waiting request needs), activates prepared cases into slots, and runs
//    rounds: one graph replay of K steps of every resident case (with
//    LBF_SERVE_EAGER=1 the same launches enqueued directly), then retires
//    the cases that finished;
//  - eight helper threads create and initialise admitted cases (uploads,
//    node classes, f = f_eq) and download and free finished ones, release
//    first, so the GPU loop never waits on that work.
// Case set-up and release issue synchronous copies, so they never overlap a
// graph capture (hip_mu). A pending capture closes a gate in front of the
// helpers' shared sections, so a steady stream of set-ups and releases cannot
// keep the capture waiting; it waits only for the sections already running.
// Device tables (slots, per-type tile lists) are rewritten only between
// rounds, on the worker's stream, so a captured graph of K steps for a given
// mix of case types stays valid while cases come and go.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lbflow_serve.h"
#include "dispatch.hpp"
#include "internal.hpp"
#include "serve.hpp"

namespace lbf {

// Ends one step of every active slot: the case's step counter and the
// remaining count advance together (one thread per slot). Then *last_flags =
// the type bits of the slots that have one step left, which the next step's
// serve_last launches read.
__global__ void __launch_bounds__(kServeMaxSlots) serve_advance(ServeSlot* slots,
                                                                uint32_t* last_flags) {
  __shared__ uint32_t bits;
  if (threadIdx.x == 0) bits = 0;
  __syncthreads();
  ServeSlot& s = slots[threadIdx.x];
  if (s.remaining > 0) {
    s.a.ctrl->step += 1;
    s.step += 1;
    s.remaining -= 1;
    if (s.remaining == 1) atomicOr(&bits, s.type_bit);
  }
  __syncthreads();
  if (threadIdx.x == 0) *last_flags = bits;
}

namespace {

// case types: lattice (D2Q9, D3Q19) x storage (F32, F32S, F16S, BF16S) x op (BGK, TRT)
constexpr int kTypes = 16;
// device counters: per type the tile counts of parity classes 0 and 1, then
// the last-step flags
constexpr int kCounters = 2 * kTypes + 1;
constexpr int kLastFlags = 2 * kTypes;
constexpr int kHelpers = 8;  // threads that set up and release cases
// caps on the workgroups of one batched launch the autotuner tries, the
// starting shape first; a launch takes min(cap, its tile count rounded up to a
// power of two), so the first cap runs one tile per workgroup
constexpr uint32_t kGridCandidates[] = {1u << 22, 32768, 8192, 2048};
// a measured shape replaces the best one only when it is this much faster:
// the rounds that measure successive shapes also differ in which cases they
// run, so small differences are noise
constexpr double kTuneMargin = 0.97;
// workgroups of a last-step launch at most (it usually returns at once)
constexpr uint32_t kLastGridMax = 2048;
// distinct graphs kept; past this the cache is emptied
constexpr size_t kMaxGraphs = 256;
constexpr int kNumCandidates = sizeof(kGridCandidates) / sizeof(kGridCandidates[0]);

int tile_index_of(uint32_t tile) {
  for (int i = 0; i < kNumTileSizes; i++)
    if ((uint32_t)kTileSizes[i] == tile) return i;
  return -1;
}

int type_index(int vs, int prec, int op) {
  if (vs != LBF_D2Q9 && vs != LBF_D3Q19) return -1;
  if (prec < LBF_F32 || prec > LBF_BF16_SHIFTED) return -1;
  if (op != LBF_BGK && op != LBF_TRT) return -1;
  return ((vs == LBF_D3Q19 ? 4 : 0) + (prec - LBF_F32)) * 2 + op;
}

int err(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return set_error(code, buf);
}

struct Request {
  uint64_t ticket = 0, steps = 0;
  uint32_t flags = 0;
  lbf_case_desc desc{};
  std::vector<uint8_t> node_type;
  std::vector<double> rho0, u0, force;
  uint64_t nodes = 0, bytes = 0;
  int d = 2, q = 9, type = -1;
  int state = LBF_REQ_QUEUED, error = 0;
  std::string errmsg;
  double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
  lbf_case* c = nullptr;
  int slot = -1;
  long long remaining = 0;
  unsigned long long step0 = 0;  // the case's step count when it was activated
  std::vector<double> rho, u, pops;
};

// Launch shape search for one mix of case types: first the workgroup counts at
// the default tile size, then the other tile sizes at the best count; each
// candidate is one round, scored as seconds per node update.
// Launch-shape search for one mix of case types: each challenger shape (the
// other workgroup caps at the best tile size, then the other tile sizes at
// the best cap) runs in the round after one of the incumbent shape, and the
// two are compared only when both rounds ran the same resident cases for the
// same number of steps (a round of other cases measures the cases, not the
// shape).
struct Tune {
  int phase = 0;  // 0: workgroup caps, 1: tile sizes
  int next = 1;   // challenger index in the phase (cap index; tile: 0..1 among the others)
  uint32_t best_grid = kGridCandidates[0];
  int best_tile = kDefaultTileIdx;
  bool have_base = false;        // the previous round measured the incumbent
  double base_score = 0;         // its seconds per algorithmic byte
  uint64_t base_version = 0;     // resident-set version and steps of that round
  int base_k = 0;
  int tries = 0;                 // unpaired attempts of the current challenger
  bool done = false;
};
// unpaired attempts after which a challenger is passed over
constexpr int kTuneTries = 4;

// Finished searches, shared by every server of the process: (device, mix of
// case types) -> (workgroups, tile size index).
std::mutex g_tune_mu;
std::map<uint64_t, std::pair<uint32_t, int>> g_tune_cache;

}  // namespace
}  // namespace lbf

using namespace lbf;

struct lbf_server {
  lbf_server_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::chrono::steady_clock::time_point t0;

  std::mutex mu;
  std::shared_mutex hip_mu;  // graph capture (exclusive) vs case set-up / release
  std::mutex gate_mu;        // capture_wanted: no new shared section may start
  std::condition_variable gate_cv;
  bool capture_wanted = false;
  std::condition_variable cv_work, cv_help, cv_done;
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::deque<Request*> queue, to_prep, ready;
  std::deque<std::pair<Request*, int>> to_finish;
  uint64_t next_ticket = 1;
  uint64_t used = 0, peak = 0;
  int reserved_slots = 0;  // preparing + ready + resident
  bool stop = false;
  std::thread worker;
  std::vector<std::thread> helpers;

  // stats (under mu)
  uint64_t submitted = 0, completed = 0, failed = 0, rounds = 0, graphs_built = 0;
  double node_updates = 0, first_submit = -1, last_done = 0, gpu_seconds = 0;
  double prep_seconds = 0, retire_seconds = 0;
  std::vector<double> latencies;
  uint32_t tuned_grid = 0, tuned_tile = 0;
  int resident = 0;

  // worker-owned
  Request* slot_req[kServeMaxSlots] = {};
  ServeSlot* d_slots = nullptr;
  ServeSlot* h_slots = nullptr;
  uint32_t* d_ntiles = nullptr;
  uint32_t* h_ntiles = nullptr;
  uint32_t* d_tiles[kTypes] = {};
  size_t tiles_cap[kTypes] = {};
  uint32_t* h_tiles = nullptr;
  size_t h_tiles_cap = 0;
  ServeLaunchFn launch[kTypes][kNumTileSizes] = {};
  ServeLastFn launch_last[kTypes][kNumTileSizes] = {};
  int cfg_tile = kDefaultTileIdx;  // tile size index when autotune is off
  int cur_tile = kDefaultTileIdx;  // tile size index of the tile lists on the device
  bool dirty = true;
  // per type and parity class: tile count of the tables on the device and its
  // power-of-two bucket (0: no tile), which sizes that class's launches; the
  // bucket grows past the count at once and shrinks when the count falls to a
  // quarter of it, so graphs stay valid while the mix drifts
  uint32_t count_of[kTypes][2] = {};
  uint32_t bucket[kTypes][2] = {};
  std::map<std::vector<uint32_t>, hipGraphExec_t> graphs;
  // LBF_SERVE_EAGER=1: rounds enqueue their launches directly instead of
  // replaying captured graphs (same kernels, same order; for tracing tools)
  bool eager = false;
  std::map<uint32_t, Tune> tune;
  uint64_t mix_version = 0;  // changes whenever a case joins or leaves the slots
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // ---------------------------------------------------------------- admission
  // FIFO; a later request may start before the oldest waiting one only with
  // memory that request would not need (so the head is never starved).
  void admit_locked() {
    const uint64_t budget = cfg.hbm_budget_bytes;
    uint64_t headroom = budget - used;
    bool head_blocked = false;
    uint64_t head_need = 0;
    for (auto it = queue.begin(); it != queue.end();) {
      Request* r = *it;
      if (reserved_slots >= (int)cfg.max_slots) break;
      const uint64_t avail = head_blocked ? (headroom > head_need ? headroom - head_need : 0)
                                          : headroom;
      if (r->bytes <= avail) {
        used += r->bytes;
        headroom -= r->bytes;
        peak = std::max(peak, used);
        reserved_slots++;
        r->state = LBF_REQ_PREPARING;
        r->t_admit = now();
        to_prep.push_back(r);
        it = queue.erase(it);
        cv_help.notify_one();
      } else {
        if (!head_blocked) {
          head_blocked = true;
          head_need = r->bytes;
        }
        ++it;
      }
    }
  }

  void finish_locked(Request* r, int state, int error) {
    r->state = state;
    r->error = error;
    r->t_done = now();
    used -= r->bytes;
    reserved_slots--;
    if (state == LBF_REQ_DONE) {
      completed++;
      node_updates += (double)r->nodes * (double)r->steps;
      last_done = std::max(last_done, r->t_done);
      latencies.push_back(r->t_done - r->t_submit);
    } else {
      failed++;
    }
    cv_done.notify_all();
    cv_work.notify_one();
  }

  // Shared section of a helper: waits while a capture is pending, then holds
  // hip_mu shared for the caller's scope.
  std::shared_lock<std::shared_mutex> enter_shared() {
    std::unique_lock<std::mutex> g(gate_mu);
    gate_cv.wait(g, [&] { return !capture_wanted; });
    g.unlock();
    return std::shared_lock<std::shared_mutex>(hip_mu);
  }

  // --------------------------------------------------------------- preparation
  void do_prep(Request* r) {
    {
      lbf_case_desc d = r->desc;
      d.node_type = r->node_type.empty() ? nullptr : r->node_type.data();
      d.initial_density = r->rho0.empty() ? nullptr : r->rho0.data();
      d.initial_velocity = r->u0.empty() ? nullptr : r->u0.data();
      d.force_field = r->force.empty() ? nullptr : r->force.data();
      lbf_case* c = nullptr;
      CaseServeInfo info{};
      int rc;
      const double tp = now();
      {
        auto hl = enter_shared();
        rc = lbf_create(&d, &c);
        if (!rc) rc = lbf_initialize_nodes(c);
        if (!rc) rc = lbf_synchronize(c);
        if (!rc) rc = case_serve_info(c, &info);
        if (!rc && (!info.aa || info.n_scalars)) rc = err(LBF_EUNSUPPORTED, "case is not servable");
        if (rc) {
          r->errmsg = lbf_last_error();
          if (c) lbf_destroy(c);
          c = nullptr;
        }
      }
      std::lock_guard<std::mutex> lk(mu);
      prep_seconds += now() - tp;
      if (rc) {
        finish_locked(r, LBF_REQ_FAILED, rc);
      } else {
        r->c = c;
        ready.push_back(r);
        cv_work.notify_one();
      }
    }
  }

  // ------------------------------------------------------------------- worker
  int activate(Request* r) {
    int slot = -1;
    for (int i = 0; i < kServeMaxSlots; i++)
      if (!slot_req[i]) {
        slot = i;
        break;
      }
    if (slot < 0) return err(LBF_ESTATE, "no free slot");
    CaseServeInfo info{};
    int rc = case_serve_info(r->c, &info);
    if (rc) return rc;
    r->slot = slot;
    r->remaining = (long long)r->steps;
    r->step0 = info.step;
    slot_req[slot] = r;
    ServeSlot& s = h_slots[slot];
    s.a = info.args;
    s.remaining = r->remaining;
    s.fresh_step = info.fresh ? info.step : ~0ull;
    s.step = info.step;
    s.type_bit = 1u << r->type;
    dirty = true;
    mix_version++;
    return 0;
  }

  uint32_t type_mask() const {
    uint32_t m = 0;
    for (int i = 0; i < kServeMaxSlots; i++)
      if (slot_req[i]) m |= 1u << slot_req[i]->type;
    return m;
  }

  void drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }

  // parity of the case's step count now (its class in the tile lists)
  static int parity_of(const Request* r) {
    return (int)((r->step0 + r->steps - (unsigned long long)r->remaining) & 1u);
  }

  // tiles of a case at the current tile size (the last one may be partial)
  uint32_t tiles_of(const Request* r) const {
    const uint64_t t = (uint64_t)kTileSizes[cur_tile];
    return (uint32_t)((r->nodes + t - 1) / t);
  }

  int upload() {
    size_t total = 0;
    std::vector<size_t> count(kTypes, 0);
    for (int i = 0; i < kServeMaxSlots; i++)
      if (slot_req[i]) count[slot_req[i]->type] += (size_t)tiles_of(slot_req[i]);
    for (int t = 0; t < kTypes; t++) total += count[t];
    if (total > h_tiles_cap) {
      if (h_tiles) (void)hipHostFree(h_tiles);
      h_tiles_cap = std::max<size_t>(total, 2 * h_tiles_cap);
      if (hipHostMalloc((void**)&h_tiles, h_tiles_cap * 4) != hipSuccess) {
        h_tiles = nullptr;
        h_tiles_cap = 0;
        return err(LBF_ENOMEM, "pinned tile staging (%zu entries)", total);
      }
    }
    // grow the device tile lists first, before this upload queues any copy:
    // no list is freed while the stream still has work (the previous round's
    // replay was waited for)
    for (int t = 0; t < kTypes; t++) {
      if (count[t] > tiles_cap[t]) {
        drop_graphs();  // graphs hold the old tile pointers
        if (d_tiles[t]) (void)hipFree(d_tiles[t]);
        tiles_cap[t] = std::max(count[t], 2 * tiles_cap[t]);
        if (hipMalloc((void**)&d_tiles[t], tiles_cap[t] * 4) != hipSuccess) {
          d_tiles[t] = nullptr;
          tiles_cap[t] = 0;
          return err(LBF_ENOMEM, "tile list (%zu entries)", count[t]);
        }
      }
    }
    // per type: the tiles of the cases at an even step count, then those at an
    // odd one (the graph launches each class with its parity)
    size_t off = 0;
    uint32_t last = 0;
    for (int t = 0; t < kTypes; t++) {
      size_t k = off;
      for (int cls = 0; cls < 2; cls++) {
        const size_t k0 = k;
        for (int i = 0; i < kServeMaxSlots; i++) {
          Request* r = slot_req[i];
          if (!r || r->type != t || parity_of(r) != cls) continue;
          const uint32_t nb = tiles_of(r);
          for (uint32_t b = 0; b < nb; b++)
            h_tiles[k++] = ((uint32_t)i << (32 - kServeSlotBits)) | b;
          if (h_slots[i].remaining == 1) last |= 1u << t;
        }
        h_ntiles[2 * t + cls] = (uint32_t)(k - k0);
        count_of[t][cls] = (uint32_t)(k - k0);
        uint32_t& b = bucket[t][cls];
        const uint32_t c = count_of[t][cls];
        if (!c)
          b = 0;
        else if (c > b || (uint64_t)c * 4 <= b) {
          b = 1;
          while (b < c) b <<= 1;
        }
      }
      if (count[t]) {
        const hipError_t e =
            hipMemcpyAsync(d_tiles[t], h_tiles + off, count[t] * 4, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return err(LBF_EDEVICE, "tile upload failed: %s", hipGetErrorString(e));
      }
      off = k;
    }
    h_ntiles[kLastFlags] = last;
    hipError_t e = hipMemcpyAsync(d_ntiles, h_ntiles, kCounters * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_slots, h_slots, sizeof(ServeSlot) * kServeMaxSlots,
                         hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return err(LBF_EDEVICE, "slot table upload failed: %s", hipGetErrorString(e));
    dirty = false;
    return 0;
  }

  // Graph of K steps of the resident mix: per type the launches of its
  // non-empty parity classes (workgroups min(cap, bucket)), the last-step
  // launch for the types with a case ending within these K steps (`last`),
  // then serve_advance; keyed by everything that shapes it.
  int get_graph(uint32_t mask, uint32_t last, int K, uint32_t cap, int tile, hipGraphExec_t* out) {
    std::vector<uint32_t> key = {mask, last, (uint32_t)K, cap, (uint32_t)tile};
    for (int t = 0; t < kTypes; t++)
      if (mask >> t & 1u) key.insert(key.end(), {bucket[t][0], bucket[t][1]});
    auto it = graphs.find(key);
    if (it != graphs.end()) {
      *out = it->second;
      return 0;
    }
    // case set-up and release on the helper threads issue synchronous
    // copies; keep them out of the capture window. The gate stays closed until
    // this function returns (after hl is released).
    {
      std::lock_guard<std::mutex> g(gate_mu);
      capture_wanted = true;
    }
    struct GateOpen {
      lbf_server* p;
      ~GateOpen() {
        {
          std::lock_guard<std::mutex> g(p->gate_mu);
          p->capture_wanted = false;
        }
        p->gate_cv.notify_all();
      }
    } gate_open{this};
    std::unique_lock<std::shared_mutex> hl(hip_mu);
    hipError_t eb = hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed);
    if (eb != hipSuccess) return err(LBF_EDEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(eb));
    if (graphs.size() >= kMaxGraphs) drop_graphs();
    const bool ok = enqueue_round(mask, last, K, cap, tile);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &g);
    if (!ok || e != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      return err(LBF_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return err(LBF_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    graphs[key] = x;
    {
      std::lock_guard<std::mutex> lk(mu);
      graphs_built++;
    }
    *out = x;
    return 0;
  }

  // The launches of one round on the worker stream (captured by get_graph, or
  // enqueued as they are with LBF_SERVE_EAGER): step k of the round runs, per
  // type, its non-empty parity classes (class c at parity (c + k) & 1), the
  // last-step launch for the types in `last`, then serve_advance.
  bool enqueue_round(uint32_t mask, uint32_t last, int K, uint32_t cap, int tile) {
    bool ok = true;
    for (int k = 0; k < K && ok; k++) {
      for (int t = 0; t < kTypes && ok; t++) {
        if (!(mask >> t & 1u)) continue;
        for (int cls = 0; cls < 2 && ok; cls++)
          if (bucket[t][cls])
            ok = launch[t][tile](d_slots, d_tiles[t], d_ntiles + 2 * t, cls, (cls + k) & 1,
                                 (int)std::min(cap, bucket[t][cls]), stream) == hipSuccess;
        if (ok && (last >> t & 1u)) {
          const uint32_t all = std::max(bucket[t][0], bucket[t][1]);
          ok = launch_last[t][tile](d_slots, d_tiles[t], d_ntiles + 2 * t,
                                    d_ntiles + kLastFlags, 1u << t,
                                    (int)std::min({cap, kLastGridMax, 2 * all}),
                                    stream) == hipSuccess;
        }
      }
      if (ok) {
        hipLaunchKernelGGL(serve_advance, dim3(1), dim3(kServeMaxSlots), 0, stream, d_slots,
                           d_ntiles + kLastFlags);
        ok = hipGetLastError() == hipSuccess;
      }
    }
    return ok;
  }

  // Launch shape (workgroups, tile size index) for the next round of this mix.
  // Shape of the next round; *measuring: the round is timed for the search,
  // *base: it runs the incumbent shape (otherwise the current challenger).
  void choose_shape(uint32_t mask, bool* measuring, bool* base, uint32_t* grid, int* tile) {
    *measuring = false;
    *base = false;
    if (!cfg.autotune) {
      *grid = cfg.grid;
      *tile = cfg_tile;
      return;
    }
    auto it = tune.find(mask);
    if (it == tune.end()) {
      Tune t;
      std::lock_guard<std::mutex> g(g_tune_mu);
      auto c = g_tune_cache.find(((uint64_t)device << 32) | mask);
      if (c != g_tune_cache.end()) {
        t.done = true;
        t.best_grid = c->second.first;
        t.best_tile = c->second.second;
      }
      it = tune.emplace(mask, t).first;
    }
    Tune& t = it->second;
    if (t.done) {
      *grid = t.best_grid;
      *tile = t.best_tile;
      return;
    }
    *measuring = true;
    if (!t.have_base) {
      *base = true;
      *grid = t.best_grid;
      *tile = t.best_tile;
    } else if (t.phase == 0) {
      *grid = kGridCandidates[t.next];
      *tile = t.best_tile;
    } else {
      *grid = t.best_grid;
      *tile = t.next < t.best_tile ? t.next : t.next + 1;  // the other tile sizes
    }
  }

  // seconds_per_byte: the round's time over its algorithmic bytes (2 Q s per
  // node update); version / k: the resident set and steps the round ran.
  void record_tuning(uint32_t mask, bool base, uint32_t grid, int tile, double seconds_per_byte,
                     uint64_t version, int k) {
    Tune& t = tune[mask];
    if (t.done) return;
    if (base) {
      t.have_base = true;
      t.base_score = seconds_per_byte;
      t.base_version = version;
      t.base_k = k;
      return;
    }
    t.have_base = false;
    if (version == t.base_version && k == t.base_k) {
      if (seconds_per_byte < kTuneMargin * t.base_score) {
        t.best_grid = grid;
        t.best_tile = tile;
      }
    } else if (++t.tries < kTuneTries) {
      return;  // the cases changed between the two rounds: pair again
    }
    t.tries = 0;
    const int n = t.phase == 0 ? kNumCandidates : kNumTileSizes - 1;
    if (t.phase == 0 && t.next + 1 < n) {
      t.next++;
      return;
    }
    if (t.phase == 0) {
      t.phase = 1;
      t.next = 0;
      return;
    }
    if (t.next + 1 < n) {
      t.next++;
      return;
    }
    t.done = true;
    {
      std::lock_guard<std::mutex> g(g_tune_mu);
      g_tune_cache[((uint64_t)device << 32) | mask] = {t.best_grid, t.best_tile};
    }
    std::lock_guard<std::mutex> lk(mu);
    tuned_grid = t.best_grid;
    tuned_tile = (uint32_t)kTileSizes[t.best_tile];
  }

  // Detaches a finished (or failed) case from its slot; a helper thread
  // downloads its result and frees it, so the next round does not wait.
  void retire(Request* r, int rc) {
    if (r->slot >= 0) {
      slot_req[r->slot] = nullptr;
      h_slots[r->slot].remaining = 0;
      r->slot = -1;
    }
    dirty = true;
    mix_version++;
    std::lock_guard<std::mutex> lk(mu);
    to_finish.push_back({r, rc});
    cv_help.notify_one();
  }

  void do_finish(std::pair<Request*, int> job) {
    {
      Request* r = job.first;
      int rc = job.second;
      const double tr = now();
      {
        auto hl = enter_shared();
        if (!rc) {
          case_note_external_steps(r->c, r->steps);
          r->rho.resize(r->nodes);
          r->u.resize(r->nodes * r->d);
          rc = lbf_get_fields(r->c, r->rho.data(), r->u.data());
          if (!rc && (r->flags & LBF_SERVE_KEEP_POPS)) {
            r->pops.resize(r->nodes * r->q);
            rc = lbf_get_pops(r->c, 0, r->pops.data());
          }
        }
        if (rc) r->errmsg = lbf_last_error();
        lbf_destroy(r->c);
        r->c = nullptr;
      }
      std::lock_guard<std::mutex> lk(mu);
      retire_seconds += now() - tr;
      resident--;
      finish_locked(r, rc ? LBF_REQ_FAILED : LBF_REQ_DONE, rc);
    }
  }

  // Helper threads: release finished cases first (that frees memory for
  // admission), then set up admitted ones.
  void helper_loop() {
    (void)hipSetDevice(device);
    while (true) {
      Request* pr = nullptr;
      std::pair<Request*, int> fin{nullptr, 0};
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_help.wait(lk, [&] { return stop || !to_finish.empty() || !to_prep.empty(); });
        if (stop) return;
        if (!to_finish.empty()) {
          fin = to_finish.front();
          to_finish.pop_front();
        } else {
          pr = to_prep.front();
          to_prep.pop_front();
        }
      }
      if (fin.first)
        do_finish(fin);
      else
        do_prep(pr);
    }
  }

  void fail_all_resident(int rc) {
    for (int i = 0; i < kServeMaxSlots; i++)
      if (slot_req[i]) retire(slot_req[i], rc);
  }

  void worker_loop() {
    (void)hipSetDevice(device);
    while (true) {
      std::vector<Request*> fresh;
      {
        std::unique_lock<std::mutex> lk(mu);
        admit_locked();
        if (stop) return;
        int in_slots = 0;
        for (int i = 0; i < kServeMaxSlots; i++) in_slots += slot_req[i] != nullptr;
        if (in_slots == 0 && ready.empty()) {
          cv_work.wait_for(lk, std::chrono::milliseconds(5));
          continue;
        }
        fresh.assign(ready.begin(), ready.end());
        ready.clear();
        resident += (int)fresh.size();
      }
      const double t_act = now();
      for (Request* r : fresh) {
        const int rc = activate(r);
        if (rc) {
          retire(r, rc);
          continue;
        }
        {
          std::lock_guard<std::mutex> lk(mu);
          r->state = LBF_REQ_RUNNING;
          r->t_start = t_act;
        }
        if (r->steps == 0) retire(r, 0);
      }
      int n_res = 0;
      long long max_rem = 0;
      for (int i = 0; i < kServeMaxSlots; i++)
        if (slot_req[i]) {
          n_res++;
          max_rem = std::max(max_rem, slot_req[i]->remaining);
        }
      if (!n_res) continue;
      const uint32_t mask = type_mask();
      bool measuring = false, base = false;
      uint32_t grid = 0;
      int tile = kDefaultTileIdx;
      choose_shape(mask, &measuring, &base, &grid, &tile);
      if (tile != cur_tile) {
        cur_tile = tile;  // tile lists are per tile size
        dirty = true;
      }
      int rc = dirty ? upload() : 0;
      if (rc) {
        fail_all_resident(rc);
        continue;
      }
      // K steps per replay; a case that finishes inside the replay skips the
      // rest of it (its remaining count is 0 on the device)
      int K = 1;
      while (2 * K <= (long long)cfg.max_steps_per_round && K < max_rem) K *= 2;
      double work = 0;  // algorithmic bytes this replay moves (2 Q s per node update)
      uint32_t last = 0;  // types with a case whose last step falls in this replay
      for (int i = 0; i < kServeMaxSlots; i++)
        if (slot_req[i]) {
          const Request* r = slot_req[i];
          const int prec = LBF_F32 + (r->type / 2) % 4;
          const double s = (prec == LBF_F32 || prec == LBF_F32_SHIFTED) ? 4.0 : 2.0;
          work += 2.0 * r->q * s * (double)r->nodes * (double)std::min<long long>(K, r->remaining);
          if (r->remaining <= K) last |= 1u << r->type;
        }
      const uint64_t version = mix_version;  // the resident set this round runs
      hipGraphExec_t g = nullptr;
      if (!eager && (rc = get_graph(mask, last, K, grid, tile, &g))) {
        fail_all_resident(rc);
        continue;
      }
      hipError_t e = hipEventRecord(ev0, stream);
      if (e == hipSuccess) {
        if (eager)
          e = enqueue_round(mask, last, K, grid, tile) ? hipSuccess : hipErrorLaunchFailure;
        else
          e = hipGraphLaunch(g, stream);
      }
      if (e == hipSuccess) e = hipEventRecord(ev1, stream);
      if (e == hipSuccess) e = hipEventSynchronize(ev1);
      if (e != hipSuccess) {
        fail_all_resident(err(LBF_EDEVICE, "%s: %s", eager ? "round launches" : "graph replay",
                              hipGetErrorString(e)));
        continue;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev0, ev1);
      if (measuring) record_tuning(mask, base, grid, tile, ms * 1e-3 / work, version, K);
      {
        std::lock_guard<std::mutex> lk(mu);
        rounds++;
        gpu_seconds += ms * 1e-3;
      }
      for (int i = 0; i < kServeMaxSlots; i++) {
        Request* r = slot_req[i];
        if (!r) continue;
        const long long ran = std::min<long long>(K, r->remaining);
        r->remaining -= ran;  // as serve_advance left it
        h_slots[i].remaining = r->remaining;
        h_slots[i].step += (unsigned long long)ran;
        if (r->remaining == 0) retire(r, 0);
      }
      if (K & 1) dirty = true;  // every case changed parity: rebuild the classes
    }
  }

  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    cv_help.notify_all();
    cv_done.notify_all();
    if (worker.joinable()) worker.join();
    for (auto& h : helpers)
      if (h.joinable()) h.join();
    for (auto& job : to_finish)
      if (job.first->c) lbf_destroy(job.first->c), job.first->c = nullptr;
    for (int i = 0; i < kServeMaxSlots; i++)
      if (slot_req[i] && slot_req[i]->c) {
        lbf_destroy(slot_req[i]->c);
        slot_req[i]->c = nullptr;
      }
    for (Request* r : ready)
      if (r->c) lbf_destroy(r->c), r->c = nullptr;
    drop_graphs();
    for (int t = 0; t < kTypes; t++)
      if (d_tiles[t]) (void)hipFree(d_tiles[t]);
    if (d_slots) (void)hipFree(d_slots);
    if (d_ntiles) (void)hipFree(d_ntiles);
    if (h_slots) (void)hipHostFree(h_slots);
    if (h_ntiles) (void)hipHostFree(h_ntiles);
    if (h_tiles) (void)hipHostFree(h_tiles);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// The worker's stream is created with a CU mask that enables every CU. HIP
// gives such a stream a hardware queue of its own instead of one of the
// pooled queues, so the helpers' set-up and release work, which runs on the
// cases' own streams, does not wait behind a graph replay that happens to
// share its queue. Such a stream is blocking: work on the legacy null stream
// (none in this library) would wait for it. If the masked stream cannot be
// created, a plain non-blocking stream serves.
hipError_t create_worker_stream(int device, hipStream_t* out) {
  int cus = 0;
  hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess && cus > 0) {
    std::vector<uint32_t> mask(((size_t)cus + 31) / 32, 0xffffffffu);
    if (hipExtStreamCreateWithCUMask(out, (uint32_t)mask.size(), mask.data()) == hipSuccess)
      return hipSuccess;
    (void)hipGetLastError();  // clear the failed attempt's error
  }
  return hipStreamCreateWithFlags(out, hipStreamNonBlocking);
}

int check_servable(const lbf_case_desc* d, int* type) {
  if (!d) return err(LBF_EINVAL, "descriptor is NULL");
  if (d->abi_version != LBF_ABI_VERSION) return err(LBF_EINVAL, "ABI version %u", d->abi_version);
  *type = type_index(d->velocity_set, d->precision, d->collision.op);
  if (*type < 0)
    return err(LBF_EUNSUPPORTED,
               "serving runs D2Q9/D3Q19, BGK/TRT, F32/F32S/F16S/BF16S storage only");
  if (d->n_scalars) return err(LBF_EUNSUPPORTED, "serving runs momentum-only cases");
  if (d->streaming == LBF_STREAM_TWO_ARRAY) return err(LBF_EUNSUPPORTED, "serving runs the AA scheme");
  const int dim = d->velocity_set == LBF_D2Q9 ? 2 : 3;
  for (int f = 0; f < 2 * dim; f++) {
    const int k = d->faces[f].kind;
    if (k != LBF_BC_NO_SLIP && k != LBF_BC_BOUNCE_BACK && k != LBF_BC_PERIODIC)
      return err(LBF_EUNSUPPORTED, "face %d: serving takes periodic / no-slip / bounce-back", f);
  }
  for (int a = 0; a < dim; a++)
    if ((d->faces[2 * a].kind == LBF_BC_PERIODIC) != (d->faces[2 * a + 1].kind == LBF_BC_PERIODIC))
      return err(LBF_EUNSUPPORTED, "axis %d: opposite faces must both be periodic or not", a);
  return 0;
}

}  // namespace

extern "C" {

int lbf_server_create(const lbf_server_config* cfg, lbf_server** out) {
  if (!cfg || !out) return err(LBF_EINVAL, "NULL argument");
  if (cfg->abi_version != LBF_ABI_VERSION) return err(LBF_EINVAL, "ABI version %u", cfg->abi_version);
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return err(LBF_EDEVICE, "no HIP device");
  auto s = std::make_unique<lbf_server>();
  s->cfg = *cfg;
  if (!s->cfg.max_slots || s->cfg.max_slots > (uint32_t)kServeMaxSlots) s->cfg.max_slots = kServeMaxSlots;
  if (!s->cfg.max_steps_per_round) s->cfg.max_steps_per_round = 64;
  s->cfg.max_steps_per_round = std::min<uint32_t>(s->cfg.max_steps_per_round, 256);
  if (!s->cfg.grid) s->cfg.grid = kGridCandidates[0];
  if (!s->cfg.tile) s->cfg.tile = (uint32_t)kTileSizes[kDefaultTileIdx];
  s->cfg_tile = tile_index_of(s->cfg.tile);
  if (s->cfg_tile < 0) return err(LBF_EINVAL, "tile %u: use 128, 256 or 512", s->cfg.tile);
  s->cur_tile = s->cfg.autotune ? kDefaultTileIdx : s->cfg_tile;
  {
    const char* env = getenv("LBF_SERVE_EAGER");
    s->eager = env && env[0] == '1';
  }
  if (hipGetDevice(&s->device) != hipSuccess) return err(LBF_EDEVICE, "hipGetDevice failed");
  if (!s->cfg.hbm_budget_bytes) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return err(LBF_EDEVICE, "hipMemGetInfo failed");
    s->cfg.hbm_budget_bytes = (uint64_t)(0.8 * (double)free_b);
  }
  s->t0 = std::chrono::steady_clock::now();
  bool ok = create_worker_stream(s->device, &s->stream) == hipSuccess &&
            hipMalloc((void**)&s->d_slots, sizeof(ServeSlot) * kServeMaxSlots) == hipSuccess &&
            hipMalloc((void**)&s->d_ntiles, kCounters * 4) == hipSuccess &&
            hipHostMalloc((void**)&s->h_slots, sizeof(ServeSlot) * kServeMaxSlots) == hipSuccess &&
            hipHostMalloc((void**)&s->h_ntiles, kCounters * 4) == hipSuccess &&
            hipEventCreate(&s->ev0) == hipSuccess && hipEventCreate(&s->ev1) == hipSuccess;
  if (ok) {
    memset(s->h_slots, 0, sizeof(ServeSlot) * kServeMaxSlots);
    memset(s->h_ntiles, 0, kCounters * 4);
    for (int t = 0; t < kTypes && ok; t++) {
      s->tiles_cap[t] = 1024;
      ok = hipMalloc((void**)&s->d_tiles[t], s->tiles_cap[t] * 4) == hipSuccess;
      const int lat = (t / 2) / 4, prec = LBF_F32 + (t / 2) % 4, op = t % 2;
      Launchers L{};
      const bool have = lat ? launchers_d3q19(prec, &L) : launchers_d2q9(prec, &L);
      for (int k = 0; k < kNumTileSizes; k++) {
        ok = ok && have && L.serve[op][k] && L.serve_last[op][k];
        if (ok) {
          s->launch[t][k] = L.serve[op][k];
          s->launch_last[t][k] = L.serve_last[op][k];
        }
      }
    }
    ok = ok &&
         hipMemcpyAsync(s->d_slots, s->h_slots, sizeof(ServeSlot) * kServeMaxSlots,
                        hipMemcpyHostToDevice, s->stream) == hipSuccess &&
         hipMemcpyAsync(s->d_ntiles, s->h_ntiles, kCounters * 4, hipMemcpyHostToDevice, s->stream) ==
             hipSuccess &&
         hipStreamSynchronize(s->stream) == hipSuccess;
  }
  if (!ok) {
    s->shutdown();
    return err(LBF_EDEVICE, "server device set-up failed");
  }
  lbf_server* p = s.release();
  p->worker = std::thread([p] { p->worker_loop(); });
  for (int i = 0; i < kHelpers; i++) p->helpers.emplace_back([p] { p->helper_loop(); });
  *out = p;
  return 0;
}

void lbf_server_destroy(lbf_server* s) {
  if (!s) return;
  s->shutdown();
  delete s;
}

int lbf_server_submit(lbf_server* s, const lbf_case_desc* d, uint64_t steps, uint32_t flags,
                      uint64_t* ticket) {
  if (!s || !ticket) return err(LBF_EINVAL, "NULL argument");
  int type = -1;
  int rc = check_servable(d, &type);
  if (rc) return rc;
  uint64_t bytes = 0;
  if ((rc = lbf_device_bytes(d, &bytes))) return rc;
  if (bytes > s->cfg.hbm_budget_bytes)
    return err(LBF_ENOMEM, "case needs %llu bytes, budget is %llu", (unsigned long long)bytes,
               (unsigned long long)s->cfg.hbm_budget_bytes);
  auto r = std::make_unique<Request>();
  const int dim = d->velocity_set == LBF_D2Q9 ? 2 : 3;
  const uint64_t N = (uint64_t)d->n[0] * d->n[1] * (dim == 3 ? d->n[2] : 1);
  if ((N + kBlock - 1) / kBlock > kServeBlockMask)
    return err(LBF_EUNSUPPORTED, "case too large for a tile index");
  r->desc = *d;
  r->d = dim;
  r->q = d->velocity_set == LBF_D2Q9 ? 9 : 19;
  r->nodes = N;
  r->bytes = bytes;
  r->type = type;
  r->steps = steps;
  r->flags = flags;
  if (d->node_type) r->node_type.assign(d->node_type, d->node_type + N);
  if (d->initial_density) r->rho0.assign(d->initial_density, d->initial_density + N);
  if (d->initial_velocity) r->u0.assign(d->initial_velocity, d->initial_velocity + N * dim);
  if (d->force_kind == LBF_SOURCE_FIELD && d->force_field)
    r->force.assign(d->force_field, d->force_field + N * dim);
  r->desc.node_type = nullptr;
  r->desc.initial_density = nullptr;
  r->desc.initial_velocity = nullptr;
  r->desc.force_field = nullptr;
  r->desc.scalars = nullptr;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return err(LBF_ESTATE, "server is shutting down");
  r->ticket = s->next_ticket++;
  r->t_submit = s->now();
  if (s->first_submit < 0) s->first_submit = r->t_submit;
  *ticket = r->ticket;
  s->queue.push_back(r.get());
  s->reqs[r->ticket] = std::move(r);
  s->submitted++;
  s->cv_work.notify_one();
  return 0;
}

int lbf_server_info(lbf_server* s, uint64_t ticket, lbf_request_info* info) {
  if (!s || !info) return err(LBF_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return err(LBF_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
  const Request& r = *it->second;
  info->state = r.state;
  info->error = r.error;
  info->nodes = r.nodes;
  info->steps = r.steps;
  info->device_bytes = r.bytes;
  info->t_submit = r.t_submit;
  info->t_admit = r.t_admit;
  info->t_start = r.t_start;
  info->t_done = r.t_done;
  return 0;
}

int lbf_server_wait(lbf_server* s, uint64_t ticket, double timeout_s) {
  if (!s) return err(LBF_EINVAL, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return err(LBF_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
  Request* r = it->second.get();
  const bool done = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
    return r->state == LBF_REQ_DONE || r->state == LBF_REQ_FAILED;
  });
  if (!done) return err(LBF_ETIMEOUT, "request %llu not finished", (unsigned long long)ticket);
  if (r->state == LBF_REQ_FAILED)
    return err(r->error ? r->error : LBF_EDEVICE, "request %llu failed: %s",
               (unsigned long long)ticket, r->errmsg.c_str());
  return 0;
}

int lbf_server_result(lbf_server* s, uint64_t ticket, double* rho, double* u, double* pops,
                      int32_t release) {
  if (!s) return err(LBF_EINVAL, "NULL argument");
  std::unique_ptr<Request> owned;  // released request, copied after the lock
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return err(LBF_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    Request& r = *it->second;
    if (r.state == LBF_REQ_FAILED)
      return err(r.error ? r.error : LBF_EDEVICE, "request %llu failed: %s",
                 (unsigned long long)ticket, r.errmsg.c_str());
    if (r.state != LBF_REQ_DONE)
      return err(LBF_ESTATE, "request %llu is not done", (unsigned long long)ticket);
    if (pops && r.pops.empty())
      return err(LBF_ESTATE, "populations were not kept (LBF_SERVE_KEEP_POPS)");
    if (!release) {
      // the request stays visible to other callers: copy under the lock
      if (rho) memcpy(rho, r.rho.data(), r.rho.size() * 8);
      if (u) memcpy(u, r.u.data(), r.u.size() * 8);
      if (pops) memcpy(pops, r.pops.data(), r.pops.size() * 8);
      return 0;
    }
    // a DONE request is no longer touched by the worker or the helpers, so
    // once it leaves the table the copies need not hold the server lock
    owned = std::move(it->second);
    s->reqs.erase(it);
  }
  const Request& r = *owned;
  if (rho) memcpy(rho, r.rho.data(), r.rho.size() * 8);
  if (u) memcpy(u, r.u.data(), r.u.size() * 8);
  if (pops) memcpy(pops, r.pops.data(), r.pops.size() * 8);
  return 0;
}

int lbf_server_drain(lbf_server* s, double timeout_s) {
  if (!s) return err(LBF_EINVAL, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  const bool done = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
    return s->completed + s->failed == s->submitted;
  });
  return done ? 0 : err(LBF_ETIMEOUT, "requests still pending");
}

int lbf_server_get_stats(lbf_server* s, lbf_server_stats* o) {
  if (!s || !o) return err(LBF_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  memset(o, 0, sizeof *o);
  o->submitted = s->submitted;
  o->completed = s->completed;
  o->failed = s->failed;
  o->queued = s->queue.size();
  o->resident = (uint64_t)s->resident;
  o->rounds = s->rounds;
  o->graphs_built = s->graphs_built;
  o->hbm_in_use = s->used;
  o->hbm_peak = s->peak;
  o->hbm_budget = s->cfg.hbm_budget_bytes;
  o->node_updates = s->node_updates;
  o->first_submit = s->first_submit < 0 ? 0 : s->first_submit;
  o->last_done = s->last_done;
  o->gpu_seconds = s->gpu_seconds;
  std::vector<double> l = s->latencies;
  if (!l.empty()) {
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
      const size_t i = (size_t)std::min<double>((double)l.size() - 1, std::ceil(p * l.size()) - 1);
      return l[i];
    };
    o->p50_latency = pct(0.50);
    o->p99_latency = pct(0.99);
    double sum = 0;
    for (double x : l) sum += x;
    o->mean_latency = sum / l.size();
    const double span = s->last_done - o->first_submit;
    if (span > 0) {
      o->cases_per_second = s->completed / span;
      o->mlups = s->node_updates / span / 1e6;
    }
  }
  o->prep_seconds = s->prep_seconds;
  o->retire_seconds = s->retire_seconds;
  o->tuned_grid = s->tuned_grid;
  o->tuned_tile = s->tuned_tile;
  return 0;
}

double lbf_server_now(lbf_server* s) { return s ? s->now() : 0.0; }

}  // extern "C"

// api.hip -- the extern "C" entry points of libpff_hip.so (declared in
// include/pff.h): argument checks, the context and its scratch, and the
// drivers that sequence the kernels (accumulate_constant_Q, reroute_flow).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <algorithm>

#include "pff_internal.hpp"

namespace pff {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return fail(PFF_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

int check_grid(const pff_grid* g, GridD* out) {
  if (!g) return fail(PFF_EINVAL, "grid is NULL");
  if (g->nx < 2 || g->ny < 2) return fail(PFF_EINVAL, "grid needs nx >= 2 and ny >= 2");
  const int64_t N = (int64_t)g->nx * (int64_t)g->ny;
  if (N >= (int64_t)0x7FFFFFFF) return fail(PFF_EINVAL, "grid has 2^31 cells or more");
  if (g->bound_mode < 0 || g->bound_mode > 3) return fail(PFF_EINVAL, "bound_mode must be 0..3");
  if (g->bound_mode == 3 && !g->boundaries) return fail(PFF_EINVAL, "custom boundary mode needs a boundaries array");
  if (!(g->dx > 0.0f)) return fail(PFF_EINVAL, "dx must be positive");
  out->nx = g->nx;
  out->ny = g->ny;
  out->N = (int32_t)N;
  out->dx = g->dx;
  out->mode = g->bound_mode;
  out->B = g->boundaries;
  return 0;
}

int scratch(pff_ctx* ctx, int slot, size_t bytes, void** out, bool zero_new) {
  if (slot < 0 || slot >= S_NSLOTS) return fail(PFF_EINVAL, "scratch slot out of range");
  if (bytes == 0) bytes = 4;
  if (ctx->cap[slot] < bytes && ctx->no_alloc)
    return fail(PFF_EINVAL, "fixed scratch: slot " + std::to_string(slot) + " holds " +
                                std::to_string(ctx->cap[slot]) + " bytes, " + std::to_string(bytes) + " needed");
  if (ctx->cap[slot] < bytes) {
    if (ctx->buf[slot]) {
      // the old buffer may still be read by queued kernels
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(ctx->buf[slot]);
      ctx->buf[slot] = nullptr;
      ctx->cap[slot] = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(PFF_ENOMEM, "device scratch allocation of " + std::to_string(bytes) + " bytes failed");
    }
    ctx->buf[slot] = p;
    ctx->cap[slot] = bytes;
    if (zero_new) PFF_TRY(check_hip(hipMemsetAsync(p, 0, bytes, ctx->stream), "memset"));
  }
  *out = ctx->buf[slot];
  return 0;
}

int wait_stream(pff_ctx* ctx) {
  if (!ctx->sync_ev) PFF_TRY(check_hip(hipEventCreateWithFlags(&ctx->sync_ev, hipEventDisableTiming), "event"));
  PFF_TRY(check_hip(hipEventRecord(ctx->sync_ev, ctx->stream), "event record"));
  for (;;) {
    const hipError_t e = hipEventQuery(ctx->sync_ev);
    if (e == hipSuccess) return 0;
    if (e != hipErrorNotReady) return check_hip(e, "event query");
  }
}

int read_words(pff_ctx* ctx) {
  PFF_TRY(post_words(ctx));
  return wait_words(ctx);
}

int post_words(pff_ctx* ctx) {
  PFF_TRY(check_hip(hipMemcpyAsync(ctx->host_words, ctx->dev_words, W_NWORDS * sizeof(int64_t),
                                   hipMemcpyDeviceToHost, ctx->stream), "d2h"));
  if (!ctx->words_ev) PFF_TRY(check_hip(hipEventCreateWithFlags(&ctx->words_ev, hipEventDisableTiming), "event"));
  return check_hip(hipEventRecord(ctx->words_ev, ctx->stream), "event record");
}

int wait_words(pff_ctx* ctx) {
  for (;;) {
    const hipError_t e = hipEventQuery(ctx->words_ev);
    if (e == hipSuccess) return 0;
    if (e != hipErrorNotReady) return check_hip(e, "event query");
  }
}

unsigned long long* prof_begin(pff_ctx* ctx, const char* name, double bytes_per_cell) {
  ctx->prof_pending = false;
  if (!ctx->prof_on || (int)ctx->prof.size() >= pff_ctx::kProfSlots) return nullptr;
  pff_ctx::ProfRec r;
  r.name = name;
  r.bytes_per_cell = bytes_per_cell;
  r.slot = (int)ctx->prof.size();
  (void)hipEventCreate(&r.start);
  (void)hipEventCreate(&r.stop);
  ctx->prof.push_back(r);
  ctx->prof_pending = true;
  return ctx->prof_cells + r.slot;
}

void prof_events(pff_ctx* ctx, hipEvent_t* start, hipEvent_t* stop) {
  *start = nullptr;
  *stop = nullptr;
  if (!ctx->prof_pending) return;
  ctx->prof_pending = false;
  *start = ctx->prof.back().start;
  *stop = ctx->prof.back().stop;
}

static void prof_clear(pff_ctx* ctx) {
  for (auto& r : ctx->prof) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  ctx->prof.clear();
}

static int begin(pff_ctx* ctx) {
  if (!ctx) return fail(PFF_EINVAL, "context is NULL");
  return check_hip(hipSetDevice(ctx->device), "hipSetDevice");
}

static int need(const void* p, const char* name) {
  if (!p) return fail(PFF_EINVAL, std::string(name) + " is NULL");
  return 0;
}

static int ceil_log2(int64_t n) {
  int r = 0;
  while ((1ll << r) < n) ++r;
  return r;
}

int accumulate_core(pff_ctx* ctx, const GridD& d, const int32_t* rcv, float q0, float* Q, int32_t* converged) {
  int32_t cycle = 0;
  PFF_TRY(tiled_upstream(ctx, d, rcv, 1, nullptr, nullptr, q0, Q, nullptr, nullptr, nullptr, &cycle));
  if (cycle) {
    // not a forest: exact compacted pointer jumping, reported as unconverged
    void* cnt;
    PFF_TRY(scratch(ctx, S_I5, (size_t)d.N * 4, &cnt));
    int32_t cyc2 = 0;
    PFF_TRY(upstream_count(ctx, d, rcv, (uint32_t*)cnt, &cyc2));
    PFF_TRY(launch_scale_counts(ctx, d, (const uint32_t*)cnt, q0, Q));
  }
  if (converged) *converged = cycle ? 0 : 1;
  return 0;
}

// armed lake-outlet log: entries of this trip, cumulative count per trip
static int log_trip(pff_ctx* ctx, const GridD& d, const int32_t* bid, const int64_t* outlet, const int32_t* snode,
                    int trip, bool all_slots) {
  auto& L = ctx->lake_log;
  if (!L.armed || trip >= L.max_trips) return 0;
  PFF_TRY(launch_log_outlets(ctx, d, bid, outlet, snode, L.dev_count, all_slots));
  return check_hip(hipMemcpyAsync(L.trip_counts + trip, L.dev_count, 8, hipMemcpyDeviceToHost, ctx->stream), "d2h");
}

// lakeflow.py:466-544 reroute_flow, the trip loop.  rec_ holds the working
// receivers (rec_ <- rec, :491); with carve, rec ends as the carved receivers.
int reroute_core(pff_ctx* ctx, const GridD& d, int32_t* bid, int32_t* rec, int32_t* rec_, int32_t* rec__,
                 const float* z, float* z_prime, uint8_t* is_border, int64_t* outlet, int64_t* basin_saddle,
                 int32_t* basin_saddlenode, int32_t carve, int32_t* bid_first, pff_lake_status* st, BatchState* bs,
                 bool* fallback) {
  hipStream_t s = ctx->stream;
  const size_t n4 = (size_t)d.N * 4;
  if (st) std::memset(st, 0, sizeof(*st));
  if (fallback) *fallback = false;
  // a canvas runs the tiled path only; anything else is rerun tile by tile
  if (bs && !carve) {
    *fallback = true;
    return 0;
  }
  // rec_ <- rec (lakeflow.py:491)
  PFF_TRY(check_hip(hipMemcpyAsync(rec_, rec, n4, hipMemcpyDeviceToDevice, s), "copy"));
  const int npass = npass_for(d.N);
  // The tiled contraction is used while the receivers form a forest; if a
  // cycle appears (possible only through rerouting, e.g. a basin-graph cycle
  // longer than 2) every later step runs the exact synchronous semantics.
  bool exact = !carve;  // jump mode can leave non-forest receivers: keep exact semantics
  int32_t *levels = nullptr, *fin = nullptr, *rcvp = nullptr;
  uint8_t* mark = nullptr;
  auto ensure_levels = [&]() -> int {
    if (levels) return 0;
    void *lv, *fv, *rp, *mk;
    PFF_TRY(scratch(ctx, S_LEVELS, (size_t)(npass + 1) * n4, &lv));
    PFF_TRY(scratch(ctx, S_I4, n4, &fv));
    PFF_TRY(scratch(ctx, S_RCVP, n4, &rp));
    PFF_TRY(scratch(ctx, S_MARK, (size_t)d.N, &mk));
    levels = (int32_t*)lv;
    fin = (int32_t*)fv;
    rcvp = (int32_t*)rp;
    mark = (uint8_t*)mk;
    return 0;
  };
  // tile activity flags: trip t labels, sorts and carves only the tiles that
  // held a closed basin after trip t-1 (all tiles in the first trip)
  uint32_t *flag_a = nullptr, *flag_b = nullptr;
  PFF_TRY(tile_flags(ctx, d, &flag_a, &flag_b));
  // new receivers: no cached ring walk is valid
  PFF_TRY(ring_cache_reset(ctx, d));
  const uint32_t* tact = nullptr;
  int label_rounds = 0;
  int nouter = 1;
  int64_t nprev = 0;
  bool relabelled = false;  // this trip's labels came from the previous trip's outlets
  int border_cur = -1;      // border list of the previous trip (-1: none yet)
  for (int it = 0; it < nouter; ++it) {
    int64_t ndb = 0;
    relabelled = false;
    if (!exact && it > 0) {
      // labels from the previous labels and outlets (no pass over rec_)
      int32_t cyc = 0;
      PFF_TRY(relabel_from_outlets(ctx, d, bid, outlet, nprev, tact, flag_a, &ndb, &cyc));
      if (cyc) {
        exact = true;
      } else {
        relabelled = true;
        label_rounds = 0;  // the carve finds its own round count
      }
    } else if (!exact) {
      int32_t cyc = 0;
      PFF_TRY(tiled_labels(ctx, d, rec_, tact, &ndb, &cyc, &label_rounds));
      if (cyc) exact = true;
    }
    nprev = ndb;
    if (exact && bs) {
      *fallback = true;
      return 0;
    }
    if (exact) PFF_TRY(depression_count(ctx, d, rec_, &ndb));
    bool frozen_only = false;
    if (bs) PFF_TRY(batch_after_labels(ctx, bs, d, it, ndb, &frozen_only));
    if (it == 0) {
      // lakeflow.py:493 (rec_ == rec here)
      if (st) st->ndep0 = ndb;
      if (ndb == 0) return 0;  // :496-497, bid untouched
      nouter = ceil_log2(ndb) + 1;  // :500
    }
    if (st) {
      if (st->n_loop < 64) st->ndep[st->n_loop] = ndb;  // :501
      st->n_loop++;
    }
    // basin labels on rec_ (:507-511)
    if (relabelled) {
      tact = flag_a;
      std::swap(flag_a, flag_b);
    } else if (!exact) {
      PFF_TRY(tiled_labels_write(ctx, d, bid, tact, flag_a));
      tact = flag_a;
      std::swap(flag_a, flag_b);
    } else if (carve) {
      PFF_TRY(ensure_levels());
      PFF_TRY(labels_split(ctx, d, rec_, bid, rec__, levels, fin, rcvp, mark));
    } else {
      PFF_TRY(basin_labels(ctx, d, rec_, bid, rec__));
    }
    if (it == 0 && bid_first)
      PFF_TRY(check_hip(hipMemcpyAsync(bid_first, bid, n4, hipMemcpyDeviceToDevice, s), "copy"));
    if (ndb == 0 || frozen_only) break;  // :512-514 (per tile on a canvas)
    if (bs) PFF_TRY(batch_freeze(ctx, bs, d, bid));
    // saddles and outlets (:520)
    if (!exact) {
      int32_t* pits = nullptr;
      int64_t* npits = nullptr;
      PFF_TRY(tiled_pits(ctx, d, &pits, &npits));
      if (bs) PFF_TRY(batch_mark42(ctx, bs, d, bid, 1));
      PFF_TRY(saddlesort_active(ctx, d, bid, is_border, z_prime, basin_saddle, basin_saddlenode, outlet, z, tact, pits,
                                npits, &border_cur));
      if (bs) PFF_TRY(batch_mark42(ctx, bs, d, bid, 0));
      PFF_TRY(log_trip(ctx, d, bid, outlet, basin_saddlenode, st ? st->n_iter : it, false));
      if (st) st->n_iter++;
      // carve (:525-526): rec equals rec_ here; rec is edited, then copied to rec_
      int32_t cyc = 0;
      PFF_TRY(carve_tiled(ctx, d, rec, rec_, bid, basin_saddlenode, outlet, rec_, tact, pits, npits, label_rounds,
                          &cyc));
      if (!cyc) continue;
      // not reachable (the labels proved rec_ a forest); fall back regardless
      if (bs) {
        *fallback = true;
        return 0;
      }
      exact = true;
      PFF_TRY(ensure_levels());
      PFF_TRY(labels_split(ctx, d, rec_, bid, rec__, levels, fin, rcvp, mark));
      PFF_TRY(saddlesort_exact(ctx, d, bid, is_border, z_prime, basin_saddle, basin_saddlenode, outlet, z));
    } else {
      PFF_TRY(saddlesort_exact(ctx, d, bid, is_border, z_prime, basin_saddle, basin_saddlenode, outlet, z));
      PFF_TRY(log_trip(ctx, d, bid, outlet, basin_saddlenode, st ? st->n_iter : it, true));
      if (st) st->n_iter++;
    }
    if (carve) {
      PFF_TRY(carve_split(ctx, d, rec, rec_, bid, basin_saddlenode, outlet, rec_, levels, fin, rec__, rcvp, mark,
                          npass));
    } else {
      PFF_TRY(launch_reroute_jump(ctx, d, rec_, outlet));  // :529
    }
  }
  return launch_basin_id_init(ctx, d, bid);  // :537
}

}  // namespace pff

using namespace pff;

extern "C" {

const char* pff_version(void) { return "pyfastflow_amd 0.1.0 (gfx950)"; }
const char* pff_last_error(void) { return g_err.c_str(); }

int pff_ctx_create(pff_ctx** out, int32_t device, void* hip_stream) {
  if (!out) return fail(PFF_EINVAL, "out is NULL");
  *out = nullptr;
  PFF_TRY(check_hip(hipSetDevice(device), "hipSetDevice"));
  pff_ctx* c = new pff_ctx();
  c->device = device;
  c->stream = (hipStream_t)hip_stream;
  if (hipHostMalloc((void**)&c->host_words, W_NWORDS * sizeof(int64_t), hipHostMallocDefault) != hipSuccess) {
    delete c;
    return fail(PFF_ENOMEM, "pinned host words");
  }
  if (hipMalloc((void**)&c->dev_words, W_NWORDS * sizeof(int64_t)) != hipSuccess) {
    (void)hipHostFree(c->host_words);
    delete c;
    return fail(PFF_ENOMEM, "device words");
  }
  (void)hipMemset(c->dev_words, 0, W_NWORDS * sizeof(int64_t));
  *out = c;
  return 0;
}

int pff_ctx_destroy(pff_ctx* ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  prof_clear(ctx);
  if (ctx->prof_cells) (void)hipFree(ctx->prof_cells);
  for (int s = 0; s < pff_ctx::kSlots; ++s)
    if (ctx->buf[s]) (void)hipFree(ctx->buf[s]);
  if (ctx->dev_words) (void)hipFree(ctx->dev_words);
  if (ctx->lake_log.dev_count) (void)hipFree(ctx->lake_log.dev_count);
  if (ctx->sync_ev) (void)hipEventDestroy(ctx->sync_ev);
  if (ctx->words_ev) (void)hipEventDestroy(ctx->words_ev);
  if (ctx->switch_ev) (void)hipEventDestroy(ctx->switch_ev);
  if (ctx->host_words) (void)hipHostFree(ctx->host_words);
  delete ctx;
  return 0;
}

int pff_ctx_set_stream(pff_ctx* ctx, void* hip_stream) {
  if (!ctx) return fail(PFF_EINVAL, "context is NULL");
  const hipStream_t ns = (hipStream_t)hip_stream;
  if (ns == ctx->stream) return 0;
  // work already queued on the old stream may still use the context's
  // scratch and device words: the new stream waits for it before anything
  // the context enqueues there
  PFF_TRY(begin(ctx));
  if (!ctx->switch_ev)
    PFF_TRY(check_hip(hipEventCreateWithFlags(&ctx->switch_ev, hipEventDisableTiming), "event"));
  PFF_TRY(check_hip(hipEventRecord(ctx->switch_ev, ctx->stream), "event record"));
  PFF_TRY(check_hip(hipStreamWaitEvent(ns, ctx->switch_ev, 0), "stream wait"));
  ctx->stream = ns;
  return 0;
}

int pff_ctx_profile(pff_ctx* ctx, int32_t enable) {
  PFF_TRY(begin(ctx));
  PFF_TRY(check_hip(hipStreamSynchronize(ctx->stream), "sync"));
  prof_clear(ctx);
  if (enable && !ctx->prof_cells) {
    if (hipMalloc((void**)&ctx->prof_cells, pff_ctx::kProfSlots * 8) != hipSuccess)
      return fail(PFF_ENOMEM, "profile counters");
  }
  if (enable) PFF_TRY(check_hip(hipMemsetAsync(ctx->prof_cells, 0, pff_ctx::kProfSlots * 8, ctx->stream), "memset"));
  ctx->prof_on = enable != 0;
  return 0;
}

__global__ void k_empty() {}

int pff_ctx_launch_overhead(pff_ctx* ctx, int32_t reps, double* median_us) {
  PFF_TRY(begin(ctx));
  if (!median_us || reps < 1 || reps > 1000) return fail(PFF_EINVAL, "launch_overhead: bad argument");
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  for (auto& e : ev) PFF_TRY(check_hip(hipEventCreate(&e), "hipEventCreate"));
  for (int r = 0; r < reps; ++r)
    hipExtLaunchKernelGGL(k_empty, dim3(1), dim3(64), 0, ctx->stream, ev[2 * r], ev[2 * r + 1], 0);
  int rc = check_hip(hipStreamSynchronize(ctx->stream), "sync");
  std::vector<double> t(reps, 0.0);
  for (int r = 0; r < reps && rc == 0; ++r) {
    float ms = 0.0f;
    rc = check_hip(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]), "hipEventElapsedTime");
    t[r] = 1e3 * ms;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  PFF_TRY(rc);
  std::sort(t.begin(), t.end());
  *median_us = t[reps / 2];
  return 0;
}

int pff_ctx_profile_report(pff_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms,
                           double* total_bytes) {
  PFF_TRY(begin(ctx));
  if (!kernel || !launches || !total_ms || !total_bytes) return fail(PFF_EINVAL, "profile_report: NULL argument");
  PFF_TRY(check_hip(hipStreamSynchronize(ctx->stream), "sync"));
  std::vector<unsigned long long> cells(ctx->prof.size());
  if (!cells.empty())
    PFF_TRY(check_hip(hipMemcpy(cells.data(), ctx->prof_cells, cells.size() * 8, hipMemcpyDeviceToHost), "d2h"));
  int64_t n = 0;
  double ms = 0, bytes = 0;
  for (size_t k = 0; k < ctx->prof.size(); ++k) {
    const auto& r = ctx->prof[k];
    if (r.name != kernel) continue;
    float t = 0;
    PFF_TRY(check_hip(hipEventElapsedTime(&t, r.start, r.stop), "hipEventElapsedTime"));
    ++n;
    ms += t;
    bytes += r.bytes_per_cell * (double)cells[k];
  }
  *launches = n;
  *total_ms = ms;
  *total_bytes = bytes;
  return 0;
}

int pff_ctx_synchronize(pff_ctx* ctx) {
  PFF_TRY(begin(ctx));
  return check_hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

// ---- receivers ------------------------------------------------------------
int pff_compute_receivers(pff_ctx* ctx, const pff_grid* g, const float* z, int32_t* rcv, float* gradient) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(z, "z"));
  PFF_TRY(need(rcv, "receivers"));
  PFF_TRY(need(gradient, "gradient"));
  return launch_receivers(ctx, d, z, rcv, gradient);
}

// ---- accumulation ---------------------------------------------------------
int pff_rcv2donor(pff_ctx* ctx, const pff_grid* g, const int32_t* rcv, int32_t* dnr, int32_t* ndnr) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(rcv, "rcv"));
  PFF_TRY(need(dnr, "dnr"));
  PFF_TRY(need(ndnr, "ndnr"));
  return launch_rcv2donor(ctx, d, rcv, dnr, ndnr);
}

int pff_rake_compress_accum(pff_ctx* ctx, const pff_grid* g, int32_t* dnr, int32_t* ndnr, float* p, int32_t* src,
                            int32_t* dnr_, int32_t* ndnr_, float* p_, int32_t iteration) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!dnr || !ndnr || !p || !src || !dnr_ || !ndnr_ || !p_) return fail(PFF_EINVAL, "rake_compress_accum: NULL array");
  return launch_rake_compress(ctx, d, dnr, ndnr, p, src, dnr_, ndnr_, p_, iteration);
}

int pff_fuse(pff_ctx* ctx, const pff_grid* g, float* A, const int32_t* src, const float* B, int32_t iteration) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!A || !src || !B) return fail(PFF_EINVAL, "fuse: NULL array");
  return launch_fuse(ctx, d, A, src, B, iteration);
}

int pff_upstream_count(pff_ctx* ctx, const pff_grid* g, const int32_t* rcv, uint32_t* count) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(rcv, "rcv"));
  PFF_TRY(need(count, "count"));
  int32_t cycle = 0;
  PFF_TRY(tiled_upstream(ctx, d, rcv, 1, nullptr, nullptr, 0.0f, nullptr, count, nullptr, nullptr, &cycle));
  if (cycle) PFF_TRY(upstream_count(ctx, d, rcv, count, nullptr));
  return 0;
}

int pff_compute_receivers_stochastic(pff_ctx* ctx, const pff_grid* g, const float* z, int32_t* rcv, float* gradient,
                                     uint64_t seed) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!z || !rcv || !gradient) return fail(PFF_EINVAL, "compute_receivers_stochastic: NULL array");
  return launch_receivers_stochastic(ctx, d, z, rcv, gradient, seed);
}

int pff_accumulate_constant_Q(pff_ctx* ctx, const pff_grid* g, const int32_t* rcv, double value, int32_t area,
                              float* Q, int32_t* converged) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(rcv, "rcv"));
  PFF_TRY(need(Q, "Q"));
  // Q.fill(value * (DX*DX if area else 1)) evaluates in double, stored as f32
  // (flowfields.py:279)
  const float q0 = (float)(value * (area ? (double)g->dx * (double)g->dx : 1.0));
  return accumulate_core(ctx, d, rcv, q0, Q, converged);
}

// ---- lakeflow -------------------------------------------------------------
int pff_depression_counter(pff_ctx* ctx, const pff_grid* g, const int32_t* rcv, int64_t* ndep) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(rcv, "rcv"));
  PFF_TRY(need(ndep, "ndep"));
  return depression_count(ctx, d, rcv, ndep);
}

int pff_basin_id_init(pff_ctx* ctx, const pff_grid* g, int32_t* bid) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(bid, "bid"));
  return launch_basin_id_init(ctx, d, bid);
}

int pff_propagate_basin(pff_ctx* ctx, const pff_grid* g, int32_t* bid, int32_t* rec_) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(bid, "bid"));
  PFF_TRY(need(rec_, "rec_"));
  void *b2, *r2;
  PFF_TRY(scratch(ctx, S_I0, (size_t)d.N * 4, &b2));
  PFF_TRY(scratch(ctx, S_I1, (size_t)d.N * 4, &r2));
  PFF_TRY(launch_propagate_basin(ctx, d, bid, rec_, (int32_t*)b2, (int32_t*)r2));
  PFF_TRY(check_hip(hipMemcpyAsync(bid, b2, (size_t)d.N * 4, hipMemcpyDeviceToDevice, ctx->stream), "copy"));
  return check_hip(hipMemcpyAsync(rec_, r2, (size_t)d.N * 4, hipMemcpyDeviceToDevice, ctx->stream), "copy");
}

int pff_basin_identification(pff_ctx* ctx, const pff_grid* g, int32_t* bid, const int32_t* rec, int32_t* rec_) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(bid, "bid"));
  PFF_TRY(need(rec, "rec"));
  PFF_TRY(need(rec_, "rec_"));
  return basin_labels(ctx, d, rec, bid, rec_);
}

int pff_saddlesort(pff_ctx* ctx, const pff_grid* g, const int32_t* bid, uint8_t* is_border, float* z_prime,
                   int64_t* basin_saddle, int32_t* basin_saddlenode, int64_t* outlet, const float* z) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!bid || !is_border || !z_prime || !basin_saddle || !basin_saddlenode || !outlet || !z)
    return fail(PFF_EINVAL, "saddlesort: NULL array");
  return launch_saddlesort(ctx, d, bid, is_border, z_prime, basin_saddle, basin_saddlenode, outlet, z);
}

int pff_reroute_jump(pff_ctx* ctx, const pff_grid* g, int32_t* rec, const int64_t* outlet) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(rec, "rec"));
  PFF_TRY(need(outlet, "outlet"));
  return launch_reroute_jump(ctx, d, rec, outlet);
}

int pff_reroute_carve(pff_ctx* ctx, const pff_grid* g, int32_t* rec, int32_t* rec_, int32_t* rec__, uint8_t* tag,
                      uint8_t* tag_, const int32_t* saddlenode, const int64_t* outlet) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!rec || !rec_ || !rec__ || !saddlenode || !outlet) return fail(PFF_EINVAL, "reroute_carve: NULL array");
  return reroute_carve(ctx, d, rec, rec_, rec__, tag, tag_, saddlenode, outlet, nullptr);
}

int pff_reroute_flow(pff_ctx* ctx, const pff_grid* g, int32_t* bid, int32_t* rec, int32_t* rec_, int32_t* rec__,
                     const float* z, float* z_prime, uint8_t* is_border, int64_t* outlet, int64_t* basin_saddle,
                     int32_t* basin_saddlenode, uint8_t* tag, uint8_t* tag_, int32_t carve, int32_t* bid_first,
                     pff_lake_status* st) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!bid || !rec || !rec_ || !rec__ || !z || !z_prime || !is_border || !outlet || !basin_saddle || !basin_saddlenode)
    return fail(PFF_EINVAL, "reroute_flow: NULL array");
  (void)tag;
  (void)tag_;
  const int rc = reroute_core(ctx, d, bid, rec, rec_, rec__, z, z_prime, is_border, outlet, basin_saddle,
                              basin_saddlenode, carve, bid_first, st);
  if (ctx->lake_log.armed) {
    ctx->lake_log.armed = false;
    PFF_TRY(check_hip(hipStreamSynchronize(ctx->stream), "sync"));
  }
  return rc;
}

int pff_ctx_lake_log(pff_ctx* ctx, int32_t* basin, int64_t* outlet, int32_t* saddlenode, int64_t capacity,
                     int64_t* trip_counts, int32_t max_trips) {
  PFF_TRY(begin(ctx));
  auto& L = ctx->lake_log;
  if (!basin || !outlet || !saddlenode || capacity < 0 || !trip_counts || max_trips < 0)
    return fail(PFF_EINVAL, "ctx_lake_log: bad argument");
  if (!L.dev_count) {
    if (hipMalloc((void**)&L.dev_count, 8) != hipSuccess) return fail(PFF_ENOMEM, "lake log counter");
  }
  PFF_TRY(check_hip(hipMemsetAsync(L.dev_count, 0, 8, ctx->stream), "memset"));
  for (int32_t t = 0; t < max_trips; ++t) trip_counts[t] = -1;
  L.basin = basin;
  L.outlet = outlet;
  L.snode = saddlenode;
  L.capacity = capacity;
  L.trip_counts = trip_counts;
  L.max_trips = max_trips;
  L.armed = true;
  return 0;
}

// ---- ragged batch -----------------------------------------------------------
int pff_batch_route(pff_ctx* ctx, int32_t ntiles, const pff_tile* tiles, float dx, int32_t lakeflow, double value,
                    int32_t area, pff_batch_status* st) {
  PFF_TRY(begin(ctx));
  if (ntiles < 0) return fail(PFF_EINVAL, "batch_route: ntiles < 0");
  if (ntiles > 0 && !tiles) return fail(PFF_EINVAL, "batch_route: tiles is NULL");
  if (!(dx > 0.0f)) return fail(PFF_EINVAL, "batch_route: dx must be positive");
  if (lakeflow < 0 || lakeflow > 2) return fail(PFF_EINVAL, "batch_route: lakeflow must be 0, 1 or 2");
  for (int32_t t = 0; t < ntiles; ++t) {
    const pff_tile& T = tiles[t];
    if (T.nx < 2 || T.ny < 2) return fail(PFF_EINVAL, "batch_route: tile " + std::to_string(t) + " needs nx, ny >= 2");
    if ((int64_t)T.nx * T.ny >= (int64_t)0x7FFFFFFF)
      return fail(PFF_EINVAL, "batch_route: tile " + std::to_string(t) + " has 2^31 cells or more");
    if (!T.z || !T.rcv || !T.Q) return fail(PFF_EINVAL, "batch_route: tile " + std::to_string(t) + " has a NULL array");
  }
  return batch_route(ctx, ntiles, tiles, dx, lakeflow, value, area, st);
}

// ---- fill -------------------------------------------------------------------
int pff_fill_z_add_delta(pff_ctx* ctx, const pff_grid* g, float* zh, float* h, float* z_, const int32_t* rcv,
                         int32_t* rcv_, int32_t* rcv__, double epsilon) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!zh || !h || !z_ || !rcv || !rcv_ || !rcv__) return fail(PFF_EINVAL, "fill_z_add_delta: NULL array");
  const size_t n4 = (size_t)d.N * 4;
  hipStream_t s = ctx->stream;
  PFF_TRY(check_hip(hipMemcpyAsync(z_, zh, n4, hipMemcpyDeviceToDevice, s), "copy"));    // :116
  PFF_TRY(check_hip(hipMemcpyAsync(rcv_, rcv, n4, hipMemcpyDeviceToDevice, s), "copy"));  // :117
  PFF_TRY(fill_passes(ctx, d, zh, z_, rcv_, rcv__, epsilon, -1));                         // :121-122
  return launch_apply_fill(ctx, d, zh, h, z_);                                             // :125
}

int pff_topofill(pff_ctx* ctx, const pff_grid* g, const float* z, float* out, const int32_t* rcv, int32_t* rcv_,
                 int32_t* rcv__, double epsilon) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!z || !out || !rcv || !rcv_ || !rcv__) return fail(PFF_EINVAL, "topofill: NULL array");
  const size_t n4 = (size_t)d.N * 4;
  hipStream_t s = ctx->stream;
  void* zc;
  PFF_TRY(scratch(ctx, S_F0, n4, &zc));
  // the passes compare against the unfilled z; keep a copy when out aliases z
  PFF_TRY(check_hip(hipMemcpyAsync(zc, z, n4, hipMemcpyDeviceToDevice, s), "copy"));
  if (out != z) PFF_TRY(check_hip(hipMemcpyAsync(out, z, n4, hipMemcpyDeviceToDevice, s), "copy"));
  PFF_TRY(check_hip(hipMemcpyAsync(rcv_, rcv, n4, hipMemcpyDeviceToDevice, s), "copy"));
  return fill_passes(ctx, d, (const float*)zc, out, rcv_, rcv__, epsilon, 0);
}

// ---- graphflood -------------------------------------------------------------
int pff_diffuse_Q_constant_prec(pff_ctx* ctx, const pff_grid* g, const pff_hydro* p, const float* zh, float* Q,
                                float* temp) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!p || !zh || !Q || !temp) return fail(PFF_EINVAL, "diffuse_Q_constant_prec: NULL argument");
  PFF_TRY(launch_diffuse(ctx, d, p, zh, Q, temp));
  // Q = temp (gf_hydrodynamics.py:74-75)
  return check_hip(hipMemcpyAsync(Q, temp, (size_t)d.N * 4, hipMemcpyDeviceToDevice, ctx->stream), "copy");
}

int pff_graphflood_core_cte_mannings(pff_ctx* ctx, const pff_grid* g, const pff_hydro* p, float* h, float* zh,
                                     float* dh, const int32_t* rec, const float* Q) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!p || !h || !zh || !dh || !rec || !Q) return fail(PFF_EINVAL, "graphflood_core_cte_mannings: NULL argument");
  return launch_manning(ctx, d, p, h, zh, dh, rec, Q);
}

}  // extern "C"

// ---- GraphFlood drivers ---------------------------------------------------------
namespace {

struct GfFields {
  int32_t *rec_, *rec__, *bid, *snode;
  float *zp, *Qt, *dh, *hprev;
  int64_t *outlet, *saddle;
  uint8_t *isb, *tag, *tag_;
  uint32_t* hist;  // residual histogram (kResidBins counters)
};

// one device block holding every field the iteration needs besides zh, h, Q, rec
int gf_fields(pff_ctx* ctx, const GridD& d, GfFields& f) {
  const size_t N = (size_t)d.N;
  const size_t bytes = N * (4 * 4 + 4 * 4 + 8 * 2) + N * 3 + 8192;
  void* base;
  PFF_TRY(scratch(ctx, S_GF0, bytes, &base, true));
  char* p = (char*)base;
  auto take = [&](size_t n) {
    char* q = p;
    p += (n + 255) / 256 * 256;
    return q;
  };
  f.outlet = (int64_t*)take(N * 8);
  f.saddle = (int64_t*)take(N * 8);
  f.rec_ = (int32_t*)take(N * 4);
  f.rec__ = (int32_t*)take(N * 4);
  f.bid = (int32_t*)take(N * 4);
  f.snode = (int32_t*)take(N * 4);
  f.zp = (float*)take(N * 4);
  f.Qt = (float*)take(N * 4);
  f.dh = (float*)take(N * 4);
  f.hprev = (float*)take(N * 4);
  f.isb = (uint8_t*)take(N);
  f.tag = (uint8_t*)take(N);
  f.tag_ = (uint8_t*)take(N);
  f.hist = (uint32_t*)take(kResidBins * 4);
  return 0;
}

// one Flooder iteration (gf_fields.py:126-141).  Q ping-pongs between the
// caller's array and f.Qt: *qcur holds it on entry and on return.
int gf_iteration(pff_ctx* ctx, const pff_grid* g, const GridD& d, const pff_hydro* p, float* zh, float* h,
                 float** qcur, float* Q, int32_t* rec, int32_t n_diffuse, double eps, GfFields& f) {
  PFF_TRY(launch_receivers(ctx, d, zh, rec, f.zp));  // the gradient output is scratch here
  PFF_TRY(pff_reroute_flow(ctx, g, f.bid, rec, f.rec_, f.rec__, zh, f.zp, f.isb, f.outlet, f.saddle, f.snode, f.tag,
                           f.tag_, 1, nullptr, nullptr));
  float* qin = *qcur;
  float* qout = qin == Q ? f.Qt : Q;
  if (d.mode == 0 && n_diffuse >= 1 && n_diffuse <= 5) {
    // fill_z_add_delta's passes (fill_topo.py:116-122), then its apply, the
    // diffusion and Manning fused in one kernel
    const size_t n4 = (size_t)d.N * 4;
    PFF_TRY(check_hip(hipMemcpyAsync(f.zp, zh, n4, hipMemcpyDeviceToDevice, ctx->stream), "copy"));
    PFF_TRY(check_hip(hipMemcpyAsync(f.rec_, rec, n4, hipMemcpyDeviceToDevice, ctx->stream), "copy"));
    PFF_TRY(fill_passes(ctx, d, zh, f.zp, f.rec_, f.rec__, eps, -1));
    PFF_TRY(launch_hydro_fused(ctx, d, p, n_diffuse, f.zp, zh, h, qin, qout, rec));
    *qcur = qout;
    return 0;
  }
  PFF_TRY(pff_fill_z_add_delta(ctx, g, zh, h, f.zp, rec, f.rec_, f.rec__, eps));
  float* a = qin;
  float* b = qout;
  for (int k = 0; k < n_diffuse; ++k) {
    PFF_TRY(launch_diffuse(ctx, d, p, zh, a, b));
    std::swap(a, b);
  }
  *qcur = a;
  return launch_manning(ctx, d, p, h, zh, f.dh, rec, a);
}

// the caller's Q array receives the iteration's Q if it is in f.Qt
int gf_q_home(pff_ctx* ctx, const GridD& d, float* qcur, float* Q) {
  if (qcur == Q) return 0;
  return check_hip(hipMemcpyAsync(Q, qcur, (size_t)d.N * 4, hipMemcpyDeviceToDevice, ctx->stream), "copy");
}

// upper edge of the histogram bin holding the q-quantile of n values
double resid_quantile(const uint32_t* hist, int64_t n, double q) {
  if (n <= 0) return 0.0;
  const double want = q * (double)n;
  double cum = 0.0;
  for (int k = 0; k < kResidBins; ++k) {
    cum += (double)hist[k];
    if (cum >= want) return k == 0 ? 0.0 : std::pow(10.0, (k - 80) / 8.0);
  }
  return std::pow(10.0, (kResidBins - 80) / 8.0);
}

int gf_check(pff_ctx* ctx, const GridD& d, const float* h, const float* hprev, double tol, uint32_t* hist_dev,
             double* res, int64_t* nans, int64_t* above, double* mean, double* quant /* q50, q90, q99, q999 */) {
  uint32_t* mx = (uint32_t*)(ctx->dev_words + W_GFMAX);
  unsigned long long* nn = (unsigned long long*)(ctx->dev_words + W_GFNAN);
  unsigned long long* ab = (unsigned long long*)(ctx->dev_words + W_GFABOVE);
  double* sm = (double*)(ctx->dev_words + W_GFSUM);
  static_assert(W_GFNAN == W_GFMAX + 1 && W_GFSUM == W_GFABOVE + 1, "residual words are pairs");
  PFF_TRY(launch_fill_i64(ctx, 2, ctx->dev_words + W_GFMAX, 0));
  PFF_TRY(launch_fill_i64(ctx, 2, ctx->dev_words + W_GFABOVE, 0));
  PFF_TRY(launch_fill_u32(ctx, kResidBins, hist_dev, 0u));
  PFF_TRY(launch_maxdiff(ctx, d.N, h, hprev, mx, nn, (float)tol, ab, sm, hist_dev));
  PFF_TRY(read_words(ctx));
  uint32_t hist[kResidBins];
  PFF_TRY(check_hip(hipMemcpy(hist, hist_dev, sizeof(hist), hipMemcpyDeviceToHost), "d2h"));
  const int64_t finite = (int64_t)d.N - ctx->host_words[W_GFNAN];
  quant[0] = resid_quantile(hist, finite, 0.5);
  quant[1] = resid_quantile(hist, finite, 0.9);
  quant[2] = resid_quantile(hist, finite, 0.99);
  quant[3] = resid_quantile(hist, finite, 0.999);
  double sumv;
  std::memcpy(&sumv, &ctx->host_words[W_GFSUM], 8);
  *mean = sumv / (double)d.N;
  uint32_t bits = (uint32_t)(ctx->host_words[W_GFMAX] & 0xFFFFFFFFu);
  float m;
  std::memcpy(&m, &bits, 4);
  *res = (double)m;
  *nans = ctx->host_words[W_GFNAN];
  *above = ctx->host_words[W_GFABOVE];
  return 0;
}

}  // namespace

extern "C" {

int pff_graphflood_run(pff_ctx* ctx, const pff_grid* g, const pff_hydro* p, float* zh, float* h, float* Q,
                       int32_t* rec, int32_t n_iter, int32_t n_diffuse, double fill_eps, pff_gf_status* st) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!p || !zh || !h || !Q || !rec) return fail(PFF_EINVAL, "graphflood_run: NULL argument");
  if (n_iter < 0 || n_diffuse < 0) return fail(PFF_EINVAL, "graphflood_run: negative count");
  GfFields f;
  PFF_TRY(gf_fields(ctx, d, f));
  float* qc = Q;
  for (int it = 0; it < n_iter; ++it) PFF_TRY(gf_iteration(ctx, g, d, p, zh, h, &qc, Q, rec, n_diffuse, fill_eps, f));
  PFF_TRY(gf_q_home(ctx, d, qc, Q));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = n_iter;
  }
  return 0;
}

int pff_graphflood_to_steady(pff_ctx* ctx, const pff_grid* g, const pff_hydro* p, float* zh, float* h, float* Q,
                             int32_t* rec, double tol, double frac, int32_t check_every, int64_t max_iter,
                             int32_t n_diffuse, double fill_eps, pff_gf_status* st) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!p || !zh || !h || !Q || !rec) return fail(PFF_EINVAL, "graphflood_to_steady: NULL argument");
  if (check_every < 1 || max_iter < 0 || n_diffuse < 0) return fail(PFF_EINVAL, "graphflood_to_steady: bad counts");
  if (!(frac >= 0.0 && frac < 1.0)) return fail(PFF_EINVAL, "graphflood_to_steady: frac must be in [0, 1)");
  GfFields f;
  PFF_TRY(gf_fields(ctx, d, f));
  pff_gf_status s;
  std::memset(&s, 0, sizeof(s));
  s.residual = -1.0;
  s.first_residual = -1.0;
  s.above_frac = -1.0;
  s.first_above_frac = -1.0;
  s.mean_change = -1.0;
  s.first_mean_change = -1.0;
  const size_t n4 = (size_t)d.N * 4;
  PFF_TRY(check_hip(hipMemcpyAsync(f.hprev, h, n4, hipMemcpyDeviceToDevice, ctx->stream), "copy"));
  float* qc = Q;
  while (s.iterations < max_iter) {
    const int64_t n = std::min<int64_t>(check_every, max_iter - s.iterations);
    for (int64_t k = 0; k < n; ++k)
      PFF_TRY(gf_iteration(ctx, g, d, p, zh, h, &qc, Q, rec, n_diffuse, fill_eps, f));
    s.iterations += n;
    if (n < check_every) break;  // a partial block cannot be judged
    double res = 0;
    int64_t nans = 0, above = 0;
    double mean = 0, quant[4];
    PFF_TRY(gf_check(ctx, d, h, f.hprev, tol, f.hist, &res, &nans, &above, &mean, quant));
    s.checks++;
    s.residual = res;
    s.nan_cells = nans;
    s.q50 = quant[0];
    s.q90 = quant[1];
    s.q99 = quant[2];
    s.q999 = quant[3];
    if (nans > 0 && s.nan_iteration == 0) s.nan_iteration = s.iterations;
    s.above_frac = (double)above / (double)d.N;
    s.mean_change = mean;
    if (s.first_residual < 0) {
      s.first_residual = res;
      s.first_above_frac = s.above_frac;
      s.first_mean_change = mean;
    }
    if (nans > 0) break;
    // frac = 0: every cell within tol (the max-norm test); otherwise at most
    // frac of the cells may still move by more than tol
    if (frac == 0.0 ? res < tol : (double)above <= frac * (double)d.N) {
      s.converged = 1;
      break;
    }
    PFF_TRY(check_hip(hipMemcpyAsync(f.hprev, h, n4, hipMemcpyDeviceToDevice, ctx->stream), "copy"));
  }
  PFF_TRY(gf_q_home(ctx, d, qc, Q));
  if (st) *st = s;
  return 0;
}

}  // extern "C"

extern "C" {

// ---- packed keys and utilities ------------------------------------------------
int pff_pack_full_float_index(pff_ctx* ctx, int64_t n, int64_t* arr, const float* ft, const int32_t* it) {
  PFF_TRY(begin(ctx));
  if (n < 0 || (n > 0 && (!arr || !ft || !it))) return fail(PFF_EINVAL, "pack: bad arguments");
  if (n == 0) return 0;
  return launch_pack(ctx, n, arr, ft, it);
}

int pff_unpack_full_float_index(pff_ctx* ctx, int64_t n, const int64_t* arr, float* ft, int32_t* it) {
  PFF_TRY(begin(ctx));
  if (n < 0 || (n > 0 && (!arr || !ft || !it))) return fail(PFF_EINVAL, "unpack: bad arguments");
  if (n == 0) return 0;
  return launch_unpack(ctx, n, arr, ft, it);
}

int pff_fill_edges(pff_ctx* ctx, const pff_grid* g, uint8_t* edges) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(edges, "edges"));
  return launch_edge_maps(ctx, d, edges, nullptr);
}

int pff_flow_out_nodes(pff_ctx* ctx, const pff_grid* g, uint8_t* outnodes) {
  PFF_TRY(begin(ctx));
  GridD d;
  PFF_TRY(check_grid(g, &d));
  PFF_TRY(need(outnodes, "outnodes"));
  return launch_edge_maps(ctx, d, nullptr, outnodes);
}

int pff_swap_arrays(pff_ctx* ctx, int64_t n, void* a, void* b) {
  PFF_TRY(begin(ctx));
  if (n < 0 || (n > 0 && (!a || !b))) return fail(PFF_EINVAL, "swap: bad arguments");
  if (n == 0) return 0;
  return launch_swap(ctx, n, (uint32_t*)a, (uint32_t*)b);
}

int pff_init_arange(pff_ctx* ctx, int64_t n, int32_t* a) {
  PFF_TRY(begin(ctx));
  if (n < 0 || (n > 0 && !a)) return fail(PFF_EINVAL, "init_arange: bad arguments");
  if (n == 0) return 0;
  return launch_arange(ctx, n, a);
}

}  // extern "C"

// pff_internal.hpp -- host-side internals of libpff_hip.so: the context, error
// reporting and the launchers each kernel file exports to the C ABI layer.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pff.h"
#include "pff_common.hpp"

struct pff_ctx {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  // growable device scratch, keyed by slot (enum Slot below; checked there)
  static constexpr int kSlots = 48;
  void* buf[kSlots] = {};
  size_t cap[kSlots] = {};
  // pinned host words for counter readback
  int64_t* host_words = nullptr;
  // device words: counters and flags
  int64_t* dev_words = nullptr;
  // reduced-graph rounds the last tiled labelling / upstream count / saddle
  // tagging and basin-chain jump needed (start estimates for the next call)
  int last_label_rounds = 8;
  int last_count_rounds = 8;
  int last_tag_rounds = 8;
  int last_jump_rounds = 4;
  // list lengths of the last exact trip (cells on or draining into a receiver
  // cycle, reduced-graph nodes still live after the first rounds of the
  // cycle-marking labels, tagged cells of the carve's level rounds), read at
  // the next exact trip's depression count; short lists run their rounds in
  // one block (one launch instead of two per round)
  int64_t hint_marked = int64_t(1) << 40, hint_xlive = int64_t(1) << 40, hint_ctag = int64_t(1) << 40;
  // the saddle-node bitmap and the carve tag bytes are all zero between
  // carves (each carve clears what it set); a carve that stopped half-way
  // leaves this set and the next one clears both in full
  bool tag_state_dirty = true;
  // the relabel's next-pit count and change words are known to be zero (the
  // last relabel's first mapping pass zeroed the count and its chase finished)
  bool relabel_words_zero = false;
  // opt-in per-kernel timing (pff_ctx_profile): HIP events around each
  // instrumented launch on the launch stream, plus the cells it processed
  struct ProfRec {
    std::string name;
    double bytes_per_cell;
    hipEvent_t start, stop;
    int slot;
  };
  // armed lake-outlet log for the next reroute_flow call (pff_ctx_lake_log)
  struct LakeLog {
    int32_t* basin = nullptr;
    int64_t* outlet = nullptr;
    int32_t* snode = nullptr;
    int64_t capacity = 0;
    int64_t* trip_counts = nullptr;  // host: entries logged up to the end of each trip
    int32_t max_trips = 0;
    int64_t* dev_count = nullptr;    // device running count (8 bytes, owned by the context)
    bool armed = false;
  } lake_log;
  hipEvent_t sync_ev = nullptr;   // spin-wait event of wait_stream
  hipEvent_t words_ev = nullptr;  // recorded after the device words' copy (post_words)
  hipEvent_t switch_ev = nullptr; // orders a new stream after the old one (pff_ctx_set_stream)
  // scratch is a fixed arena set up by the caller: scratch() never allocates
  // and fails when a slot is too small (the no-allocation entry points)
  bool no_alloc = false;
  bool prof_on = false;
  bool prof_pending = false;  // prof_begin made a record that the next launch takes
  std::vector<ProfRec> prof;
  unsigned long long* prof_cells = nullptr;  // device, kProfSlots counters
  static constexpr int kProfSlots = 4096;
};

namespace pff {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);
int check_grid(const pff_grid* g, GridD* out);
// profiling: returns the device counter for the launch's processed cells
// (null when profiling is off); launch the kernel with PFF_PROF_LAUNCH,
// whose events mark the kernel's own start and end
unsigned long long* prof_begin(pff_ctx* ctx, const char* name, double bytes_per_cell);
void prof_events(pff_ctx* ctx, hipEvent_t* start, hipEvent_t* stop);
#define PFF_PROF_LAUNCH(ctx, kernel, grid, block, shmem, stream, ...)                      \
  do {                                                                                    \
    hipEvent_t e0_, e1_;                                                                  \
    prof_events(ctx, &e0_, &e1_);                                                         \
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, e0_, e1_, 0, __VA_ARGS__); \
  } while (0)
// wait for the context stream by polling an event (a short spin instead of
// the runtime's blocking wait; every lakeflow trip waits once or twice)
int wait_stream(pff_ctx* ctx);
// copy the device words to the pinned host words and wait
int read_words(pff_ctx* ctx);
// the same in two halves: post the copy (the stream can take more work),
// then wait for the copy alone
int post_words(pff_ctx* ctx);
int wait_words(pff_ctx* ctx);
// device scratch slot of at least `bytes`
int scratch(pff_ctx* ctx, int slot, size_t bytes, void** out, bool zero_new = false);

// scratch slot map (each slot is one growable buffer)
enum Slot : int {
  S_I0 = 0, S_I1, S_I2, S_I3, S_I4, S_I5,  // int32 / uint32 N-arrays
  S_F0, S_F1,                              // float N-arrays
  S_WL0, S_WL1,                            // worklists
  S_B0,                                    // byte N-array
  S_LEVELS,                                // pointer-jumping levels
  S_MISC,
  S_TAG,                                   // carve tag words, kept all-zero between calls
  S_B1,                                    // byte N-array (tiled carve tags)
  S_TE, S_TP2, S_TBID, S_TEXIT, S_TS, S_TIN, S_TPIT, S_TACT,  // tiled contraction buffers
  S_GF0,    // GraphFlood driver fields (one block, see graphflood_fields)
  S_BATCH,  // tile-canvas arrays of pff_batch_route (one block)
  S_BATCHT, // tile-canvas tables (per tile)
  S_M0, S_M1, S_TPIT2,  // basin-level relabel: label maps (by basin id) and the next pits list
  S_BL0, S_BL1,         // border cells of closed basins (and marked candidate copies), this trip's and the next trip's; 2N entries
  S_TLIST,              // cells tagged by the tiled carve
  S_RCVP, S_MARK,       // receivers with the cycle cells made roots, cycle-cell flags
  S_WBITS,              // saddle nodes of the trip, one bit per cell in tile order (128 words per tile)
  S_SEGD0, S_SEGD1,     // per-tile list segments (4096 entries per tile), two lists
  S_SEGCO,              // their per-tile counts and offsets
  S_RING,               // ring-slot arrays of the carve tag walks (4 x (256 per tile + 1))
  S_RCACHE,             // cached local ends of the ring walks (u16 per ring slot) + per-tile dirty words
  S_WGPART,             // per-chunk partial weight gradients of a 3x3 convolution
  S_NSLOTS
};
static_assert(S_NSLOTS <= pff_ctx::kSlots, "pff_ctx::kSlots must cover every scratch slot");

// device word indices (dev_words / host_words)
// lists up to this length run their pointer-jumping rounds in one block
constexpr int64_t kShortList = 16384;

enum Word : int { W_COUNT0 = 0, W_COUNT1, W_COUNT2, W_FLAG, W_FLAG2, W_CYCLE, W_NEXIT, W_NDEP, W_NPIT, W_LIVE, W_EMPTY, W_GFMAX, W_GFNAN, W_NPIT2, W_MCHG, W_EMPTY2, W_GFABOVE, W_GFSUM, W_NBL0, W_NBL1, W_NTAG, W_NPREV, W_HMARK, W_HXLIVE, W_HCTAG, W_LAST_WORD, W_NWORDS = 32 };
static_assert(W_LAST_WORD <= W_NWORDS, "device word block too small");

constexpr int kBlock = 256;
inline unsigned grid_for(int64_t n, int block = kBlock, int64_t cap = 1 << 20) {
  int64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}
// fixed grid for grid-stride kernels over a device-side count
inline unsigned stride_grid() { return 2048; }

// ---- launchers (defined in k_*.hip) -----------------------------------------
int launch_receivers(pff_ctx* ctx, const GridD& g, const float* z, int32_t* rcv, float* grad);
int launch_receivers_stochastic(pff_ctx* ctx, const GridD& g, const float* z, int32_t* rcv, float* grad,
                                uint64_t seed);

int launch_rcv2donor(pff_ctx* ctx, const GridD& g, const int32_t* rcv, int32_t* dnr, int32_t* ndnr);
int launch_rake_compress(pff_ctx* ctx, const GridD& g, int32_t* dnr, int32_t* ndnr, float* p, int32_t* src,
                         int32_t* dnr_, int32_t* ndnr_, float* p_, int32_t iteration);
int launch_fuse(pff_ctx* ctx, const GridD& g, float* A, const int32_t* src, const float* B, int32_t iteration);
// exact uint32 upstream counts; *cycle (host) set when a cycle blocks completion
int upstream_count(pff_ctx* ctx, const GridD& g, const int32_t* rcv, uint32_t* cnt, int32_t* cycle);
int launch_scale_counts(pff_ctx* ctx, const GridD& g, const uint32_t* cnt, float q0, float* Q);

int depression_count(pff_ctx* ctx, const GridD& g, const int32_t* rcv, int64_t* out);
int launch_basin_id_init(pff_ctx* ctx, const GridD& g, int32_t* bid);
int launch_propagate_basin(pff_ctx* ctx, const GridD& g, const int32_t* bid_in, const int32_t* rec_in,
                           int32_t* bid_out, int32_t* rec_out);
// bid = init(root(rec)) with ceil(log2 N)+1 synchronous passes' semantics;
// rec_out (nullable) receives the final jumped pointers
int basin_labels(pff_ctx* ctx, const GridD& g, const int32_t* rec, int32_t* bid, int32_t* rec_out);
int launch_saddlesort(pff_ctx* ctx, const GridD& g, const int32_t* bid, uint8_t* is_border, float* zp,
                      int64_t* saddle, int32_t* snode, int64_t* outlet, const float* z);
int saddlesort_exact(pff_ctx* ctx, const GridD& g, const int32_t* bid, uint8_t* isb, float* zp, int64_t* saddle,
                     int32_t* snode, int64_t* outlet, const float* z);
// tiled contraction (forests; *cycle = 1 when the receivers are not a forest)
int tiled_upstream(pff_ctx* ctx, const GridD& g, const int32_t* rcv, int wmode, const int32_t* bid,
                   const int32_t* snode, float q0, float* Q, uint32_t* cnt, uint8_t* tag, const uint32_t* tact,
                   int32_t* cycle, int rounds = 0, int32_t* tag_list = nullptr,
                   int32_t* converged_dev = nullptr);
// scratch bytes of an unweighted tiled_upstream call, and the slots it takes
// (slot, bytes) in the order a fixed arena lays them out
size_t tiled_upstream_slots(const GridD& g, int* slots, size_t* bytes, int* nslots);
// mark_cycles: the receivers may hold cycles; the cells on or draining into
// one get the label -1 in tiled_labels_write (no host read; *ndep, *cycle and
// *rounds_used are not written)
int tiled_labels(pff_ctx* ctx, const GridD& g, const int32_t* rec, const uint32_t* tact_in, int64_t* ndep,
                 int32_t* cycle, int* rounds_used = nullptr, bool mark_cycles = false);
int tiled_labels_write(pff_ctx* ctx, const GridD& g, int32_t* bid, const uint32_t* tact_in, uint32_t* tact_out);
int tiled_pits(pff_ctx* ctx, const GridD& g, int32_t** pits, int64_t** npits);
// Labels of the next lakeflow trip from this trip's labels and outlets,
// without a pass over the receivers: a carved basin's cells all drain through
// its saddle to its outlet, so their new label is the label at the end of the
// basin-level chain b -> bid[outlet(b)] -> ...; uncarved basins keep theirs.
// Rewrites bid on the tiles of tact_in, flags in tact_out the tiles that still
// hold a closed basin, and replaces the pits list by the uncarved basins.
// *ndep (host): the new depression count; *cycle (host): the chain has a
// cycle (the labels are then not written).  n_prev: pits in the current list.
int relabel_from_outlets(pff_ctx* ctx, const GridD& g, int32_t* bid, const int64_t* outlet, int64_t n_prev,
                         const uint32_t* tact_in, uint32_t* tact_out, int64_t* ndep, int32_t* cycle);
int tile_flags(pff_ctx* ctx, const GridD& g, uint32_t** flags_a, uint32_t** flags_b);
// Per-tile list segments (see k_seg_scan): a tile block writes its entries
// to seg + tile * TC and its count, instead of appending with an atomic.
struct SegLists {
  int32_t* seg[2];
  int32_t* cnt[2];  // per tile
  int32_t* off[2];  // per tile
  int32_t* list[2];
  int64_t* count[2];
  int append[2];
};
// segments of up to two lists of a tile pass (slots S_SEGD0/1, S_SEGCO);
// seg_join turns them into the contiguous lists and their counts
int seg_lists(pff_ctx* ctx, const GridD& g, SegLists& sl, int32_t* list0, int64_t* count0, int32_t* list1,
              int64_t* count1);
void seg_join(pff_ctx* ctx, const GridD& g, const SegLists& sl, int nlists);

// saddle-node bitmap (S_WBITS) and carve tag bytes (S_B1), zero between
// carves.  begin: a trip starts setting them (cleared first if the last carve
// did not finish); the carve that clears them resets tag_state_dirty.
int tag_state(pff_ctx* ctx, const GridD& g, uint32_t** wbits, uint8_t** tag, bool begin);
// Carve tags of one trip on a forest of D4 receivers whose saddle nodes are
// set in wbits (the bits are cleared): tag[c] = 1 for every cell with a saddle
// node upstream, each listed once in tag_list (count W_NTAG, zeroed by the
// caller).  Per tile, the paths from the saddle nodes and from the tile's
// ring cells are walked in LDS; only their ends enter the reduced graph.
// rounds > 0: known bound of the reduced rounds (no host read).  Expects the
// W_CYCLE and W_NEXIT words zeroed (saddlesort_active's saddle init does it).
int saddle_tags_walk(pff_ctx* ctx, const GridD& g, const int32_t* rcv, const int32_t* bid, uint32_t* wbits,
                     uint8_t* tag, const uint32_t* tact, int rounds, int32_t* tag_list);
int tiles_count(const GridD& g);
// Cached ring-walk ends of the carve tag walks (k_tag_local) and the per-tile
// dirty words (1 = receivers changed since the tile's walks were cached).
// reset: every tile dirty (the receivers of a new reroute_flow call).
int ring_cache(pff_ctx* ctx, const GridD& g, uint16_t** rcache, uint32_t** tdirty);
int ring_cache_reset(pff_ctx* ctx, const GridD& g);
int carve_tiled(pff_ctx* ctx, const GridD& g, int32_t* rec, const int32_t* rec_in, const int32_t* bid,
                const int32_t* snode, const int64_t* outlet, int32_t* rec_work, const uint32_t* tact,
                const int32_t* pits, const int64_t* npits, int rounds, int32_t* cycle);
int saddlesort_active(pff_ctx* ctx, const GridD& g, const int32_t* bid, uint8_t* isb, float* zp, int64_t* saddle,
                      int32_t* snode, int64_t* outlet, const float* z, const uint32_t* tact, const int32_t* pits,
                      const int64_t* npits, int* border_cur);
int launch_reroute_jump(pff_ctx* ctx, const GridD& g, int32_t* rec, const int64_t* outlet);
// append (b, outlet[b], saddlenode[b]) for every basin b of the current labels
// (cells with bid[i] == i + 1) to the armed lake log; *dev_count counts them
int launch_log_outlets(pff_ctx* ctx, const GridD& g, const int32_t* bid, const int64_t* outlet,
                       const int32_t* snode, int64_t* dev_count, bool all_slots);
// reference-shaped carve (lakeflow.py:405-438)
int reroute_carve(pff_ctx* ctx, const GridD& g, int32_t* rec, int32_t* rec_, int32_t* rec__, uint8_t* tag,
                  uint8_t* tag_, const int32_t* snode, const int64_t* outlet, int32_t* iters);
int npass_for(int32_t N);
int labels_with_levels(pff_ctx* ctx, const GridD& g, const int32_t* rec, int32_t* bid, int32_t* Pfinal,
                       int32_t* levels, int32_t* fin);
int carve_from_levels(pff_ctx* ctx, const GridD& g, int32_t* rec, const int32_t* rec_in, const int32_t* levels,
                      const int32_t* fin, const int32_t* Pfinal, const int32_t* snode, const int64_t* outlet,
                      int32_t* rec_work, uint8_t* tag_bytes, int npass, int32_t* iters);
// Labels and carve of a trip whose receivers may hold cycles, split between
// the cells whose path reaches a root (tiled contraction) and the cells on or
// draining into a cycle (synchronous passes over those cells only).  Same
// results as labels_with_levels + carve_from_levels.  labels_split writes
// rcvp (rec with the cycle cells made roots) and mark (1 on the cycle cells)
// for carve_split; Pfinal, levels and fin only on the cycle cells.
int labels_split(pff_ctx* ctx, const GridD& g, const int32_t* rec, int32_t* bid, int32_t* Pfinal, int32_t* levels,
                 int32_t* fin, int32_t* rcvp, uint8_t* mark);
int carve_split(pff_ctx* ctx, const GridD& g, int32_t* rec, const int32_t* rec_in, const int32_t* bid,
                const int32_t* snode, const int64_t* outlet, int32_t* rec_work, const int32_t* levels,
                const int32_t* fin, const int32_t* Pfinal, const int32_t* rcvp, const uint8_t* mark, int npass);

int fill_passes(pff_ctx* ctx, const GridD& g, const float* zh, float* z_, int32_t* r0, int32_t* r1, double eps,
                int first_exp);
int launch_apply_fill(pff_ctx* ctx, const GridD& g, float* zh, float* h, const float* z_);
int launch_diffuse(pff_ctx* ctx, const GridD& g, const pff_hydro* p, const float* zh, const float* Q, float* temp);
int launch_manning(pff_ctx* ctx, const GridD& g, const pff_hydro* p, float* h, float* zh, float* dh,
                   const int32_t* rec, const float* Q);
// fill apply + n_diffuse diffusion passes + Manning in one kernel ('normal'
// grids, 1..5 passes): zf = filled surface, zh = surface before the fill
// (overwritten with the new surface), Qin -> Qout (distinct buffers)
int launch_hydro_fused(pff_ctx* ctx, const GridD& g, const pff_hydro* p, int n_diffuse, const float* zf, float* zh,
                       float* h, const float* Qin, float* Qout, const int32_t* rec);

// ---- drivers shared by the entry points (api.hip) ----------------------------
struct BatchState;
// reroute_flow on an internal grid.  With a batch state (tile canvas), the
// trip loop applies each tile's own outer cap; *fallback is set (and nothing
// more is done) when the receivers stop being a forest, which the canvas
// does not handle.
int reroute_core(pff_ctx* ctx, const GridD& d, int32_t* bid, int32_t* rec, int32_t* rec_, int32_t* rec__,
                 const float* z, float* z_prime, uint8_t* is_border, int64_t* outlet, int64_t* basin_saddle,
                 int32_t* basin_saddlenode, int32_t carve, int32_t* bid_first, pff_lake_status* st,
                 BatchState* bs = nullptr, bool* fallback = nullptr);
int accumulate_core(pff_ctx* ctx, const GridD& d, const int32_t* rcv, float q0, float* Q, int32_t* converged);

// ---- tile canvas (batch.hip) -------------------------------------------------
// per-trip hook after the labels: counts each tile's depressions; on trip 0
// sets the tiles' outer caps; a tile with depressions left that has used its
// cap is frozen for the rest of the call; *stop when only frozen tiles are left
int batch_after_labels(pff_ctx* ctx, BatchState* bs, const GridD& d, int trip, int64_t ndep, bool* stop);
// bid <- 0 on the frozen tiles, so that no basin of theirs gets a saddle,
// an outlet or a carve in this trip
int batch_freeze(pff_ctx* ctx, BatchState* bs, const GridD& d, int32_t* bid);
// f42[bid[cell 42 of each tile]] = value
int batch_mark42(pff_ctx* ctx, BatchState* bs, const GridD& d, const int32_t* bid, uint8_t value);
int batch_route(pff_ctx* ctx, int32_t ntiles, const pff_tile* tiles, float dx, int32_t lakeflow, double value,
                int32_t area, pff_batch_status* st);

// surrogate (k_conv.hip)
int launch_conv3x3_fwd(pff_ctx* ctx, int n_img, int H, int W, int Cin, int Cout, const uint16_t* x,
                       const uint16_t* wf, const float* bias, int relu, uint16_t* y);
int launch_conv3x3_wgrad(pff_ctx* ctx, int n_img, int H, int W, int Cin, int Cout, const uint16_t* x,
                         const uint16_t* dy, float* dw, float* db);
int launch_pack_weights(pff_ctx* ctx, int Cin, int Cout, const float* w, uint16_t* wf, uint16_t* wt);
int launch_relu_bwd(pff_ctx* ctx, int64_t n, const uint16_t* y, uint16_t* dy);
int launch_avgpool2(pff_ctx* ctx, int n_img, int H, int W, int C, const uint16_t* x, uint16_t* y, int backward);
int launch_upsample2(pff_ctx* ctx, int n_img, int H, int W, int C, const uint16_t* x, uint16_t* y, int backward);
int launch_concat(pff_ctx* ctx, int64_t npix, int Ca, int Cb, const uint16_t* a, const uint16_t* b, uint16_t* y,
                  int backward);
int launch_head(pff_ctx* ctx, int64_t npix, int C, const uint16_t* x, const float* w, const float* b, float* y);
int launch_head_bwd(pff_ctx* ctx, int64_t npix, int C, const uint16_t* x, const float* w, const float* dy,
                    uint16_t* dx, float* dw, float* db);
int launch_pack_input(pff_ctx* ctx, int64_t npix, int nplanes, int C, const float* planes, uint16_t* x);

int launch_edge_maps(pff_ctx* ctx, const GridD& g, uint8_t* edges, uint8_t* out);
// residual histogram of launch_maxdiff: bin 0 = exact zero, bins 1.. eight per
// decade, bin k >= 1 ending at 10^((k - 80) / 8) (1e-10 .. 1e5, ends clamped)
constexpr int kResidBins = 122;
int launch_maxdiff(pff_ctx* ctx, int32_t N, const float* a, const float* b, uint32_t* out_max,
                   unsigned long long* out_nan, float tol = 0.f, unsigned long long* out_above = nullptr,
                   double* out_sum = nullptr, uint32_t* hist = nullptr);
int launch_pack(pff_ctx* ctx, int64_t n, int64_t* arr, const float* ft, const int32_t* it);
int launch_unpack(pff_ctx* ctx, int64_t n, const int64_t* arr, float* ft, int32_t* it);
int launch_swap(pff_ctx* ctx, int64_t n, uint32_t* a, uint32_t* b);
int launch_arange(pff_ctx* ctx, int64_t n, int32_t* a);
int launch_fill_u32(pff_ctx* ctx, int64_t n, uint32_t* a, uint32_t v);
int launch_fill_i64(pff_ctx* ctx, int64_t n, int64_t* a, int64_t v);

}  // namespace pff

#define PFF_TRY(expr)          \
  do {                         \
    int _rc = (expr);          \
    if (_rc != 0) return _rc;  \
  } while (0)

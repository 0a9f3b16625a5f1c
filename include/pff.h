/*
 * pff.h -- C ABI of libpff_hip.so, the MI355X (gfx950) implementation of
 * PyFastFlow's flow-routing and GraphFlood hot path.
 *
 * The reference (bgailleton/pyfastflow) has no FFI: its boundary is a Python
 * call surface over Taichi fields (pyfastflow.flow, pyfastflow.graphflood).
 * Each entry point below replaces one reference kernel or driver; the
 * reference file:line is given at each declaration.  The Python drop-in
 * (pyfastflow_amd) binds these with ctypes.
 *
 * Conventions
 *  - Every array argument is a DEVICE pointer (HBM), row-major 1-D of length
 *    N = nx*ny unless stated, owned by the caller.  The library owns only the
 *    scratch held in a pff_ctx.
 *  - Calls enqueue on the context's HIP stream and return without waiting,
 *    except where a result is returned through a host pointer (stated).
 *  - Every call returns 0 on success or a negative PFF_E* code;
 *    pff_last_error() gives the message (per host thread).
 *  - Grids need nx >= 2, ny >= 2 and nx*ny < 2^31.
 */
#ifndef PFF_H
#define PFF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFF_OK 0
#define PFF_EINVAL -1    /* bad argument (shape, null pointer, mode) */
#define PFF_EHIP -2      /* HIP runtime error */
#define PFF_ENOMEM -3    /* device allocation failed */
#define PFF_ENOTCONV -4  /* an iteration cap was reached before convergence */

/* Grid description; mirrors the globals constants.NX/NY/DX/BOUND_MODE and the
   custom boundary field (constants.py:27-83). */
typedef struct {
  int32_t nx, ny;
  float dx;
  int32_t bound_mode;          /* 0 normal, 1 periodic_EW, 2 periodic_NS, 3 custom */
  const uint8_t* boundaries;   /* device, N codes 0/1/3/7/9, mode 3 only, else NULL */
} pff_grid;

/* Hydraulic parameters; mirror constants.PREC/MANNING/EDGESW/DT_HYDRO
   (constants.py:93-103) as set by Flooder (gf_fields.py:62-108).
   update: 0 = the reference's depth update (h receives dh twice, zh once,
   gf_hydrodynamics.py:121-129); 1 = h and zh receive dh once and h is
   clamped at 0 (no reference counterpart; the reference's update has no
   steady state once lakes are filled). */
typedef struct {
  double prec, manning, edge_slope, dt;
  int32_t update;
} pff_hydro;

typedef struct pff_ctx pff_ctx;

const char* pff_version(void);
const char* pff_last_error(void);

/* Context: device, stream and scratch.  No reference counterpart (the
   reference keeps scratch as FlowRouter fields, flowfields.py:66-159). */
int pff_ctx_create(pff_ctx** out, int32_t device, void* hip_stream);
int pff_ctx_destroy(pff_ctx* ctx);
int pff_ctx_set_stream(pff_ctx* ctx, void* hip_stream);
int pff_ctx_synchronize(pff_ctx* ctx);
/* Opt-in per-kernel timing, no reference counterpart: while enabled, the
   instrumented launches ("k_tile", "k_tag_local", "k_tag_final",
   "k_flag_scatter2", "k_zprime_min_t", "k_topofill", "k_hydro") record HIP
   events at the kernel's
   start and end (hipExtLaunchKernelGGL) and count the algorithmic bytes of
   what they process.  Enabling clears the records; report() syncs and sums
   the launches of one kernel: count, time (ms) and algorithmic bytes.
   launch_overhead() times reps launches of an empty kernel recorded the same
   way and returns the median (us): what such an event pair adds to a
   launch beyond the kernel's own run. */
int pff_ctx_profile(pff_ctx* ctx, int32_t enable);
int pff_ctx_launch_overhead(pff_ctx* ctx, int32_t reps, double* median_us);
int pff_ctx_profile_report(pff_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms,
                           double* total_bytes);

/* ---- flow: receivers -------------------------------------------------- */
/* receivers.py:15-58  compute_receivers(z, receivers, gradient) */
int pff_compute_receivers(pff_ctx* ctx, const pff_grid* g, const float* z, int32_t* rcv, float* gradient);
/* receivers.py:46-47  the same with RAND_RCV (FlowRouter(stochastic_receivers=True)):
   each valid link's slope is multiplied by a uniform draw in [0, 1) before
   the comparison, and gradient holds the weighted slope.  The draws are a
   pure function of (seed, cell, direction): splitmix64 of
   seed + (4*i + k) * 0x9E3779B97F4A7C15, top 24 bits / 2^24.  Taichi's
   ti.random() is another generator, so only the distribution matches the
   reference, not the draws. */
int pff_compute_receivers_stochastic(pff_ctx* ctx, const pff_grid* g, const float* z, int32_t* rcv, float* gradient,
                                     uint64_t seed);

/* ---- flow: accumulation ------------------------------------------------ */
/* downstream_propag.py:77-96  rcv2donor(rcv, dnr, ndnr); ndnr must be zeroed
   by the caller as flowfields.py:275 does; dnr has 4*N slots */
int pff_rcv2donor(pff_ctx* ctx, const pff_grid* g, const int32_t* rcv, int32_t* dnr, int32_t* ndnr);
/* downstream_propag.py:99-179  rake_compress_accum(dnr, ndnr, p, src, dnr_, ndnr_, p_, iteration) */
int pff_rake_compress_accum(pff_ctx* ctx, const pff_grid* g, int32_t* dnr, int32_t* ndnr, float* p, int32_t* src,
                            int32_t* dnr_, int32_t* ndnr_, float* p_, int32_t iteration);
/* downstream_propag.py:59-74  fuse(A, src, B, iteration) */
int pff_fuse(pff_ctx* ctx, const pff_grid* g, float* A, const int32_t* src, const float* B, int32_t iteration);
/* flowfields.py:254-292  FlowRouter.accumulate_constant_Q(value, area):
   Q[i] = f32(value*dx^2 or value) * |upstream(i) + i|, computed as exact
   integer upstream counts on the receiver forest.  *converged (host,
   nullable) is 1 unless the receivers hold a cycle. */
int pff_accumulate_constant_Q(pff_ctx* ctx, const pff_grid* g, const int32_t* rcv, double value, int32_t area,
                              float* Q, int32_t* converged);
/* No counterpart in the reference.  For callers that cannot allocate (code
   under HIP graph capture, tight serving loops): pff_accumulate_constant_Q
   with no context and no allocation.  Its scratch is one fixed device array
   of 64 MiB inside the library (48 bytes per cell + 32 KiB per 64x64 tile,
   e.g. a 1024x1024 grid); larger grids are rejected with PFF_EINVAL.  Work
   is enqueued on hip_stream of the current device and nothing is read back:
   *converged is a DEVICE int32, written 1 on a forest and 0 when the
   receivers hold a cycle (Q is then not meaningful).  All calls on one
   device share the array: calls that overlap on a device race. */
int pff_accumulate_constant_Q_noalloc(void* hip_stream, const pff_grid* g, const int32_t* rcv, double value,
                                      int32_t area, float* Q, int32_t* converged);
/* Upstream cell counts (including self) as uint32; the integer core of
   accumulate_constant_Q. */
int pff_upstream_count(pff_ctx* ctx, const pff_grid* g, const int32_t* rcv, uint32_t* count);

/* ---- flow: lakeflow ---------------------------------------------------- */
/* lakeflow.py:24-44  depression_counter(rec) -> int (host result; syncs) */
int pff_depression_counter(pff_ctx* ctx, const pff_grid* g, const int32_t* rcv, int64_t* ndep);
/* lakeflow.py:46-61  basin_id_init(bid) */
int pff_basin_id_init(pff_ctx* ctx, const pff_grid* g, int32_t* bid);
/* lakeflow.py:63-80  propagate_basin(bid, rec_): one synchronous pointer-jumping pass */
int pff_propagate_basin(pff_ctx* ctx, const pff_grid* g, int32_t* bid, int32_t* rec_);
/* lakeflow.py:82-101  basin_identification(bid, rec, rec_, N): rec_ <- rec,
   init, ceil(log2 N)+1 passes; on return bid = init(root), rec_ = roots */
int pff_basin_identification(pff_ctx* ctx, const pff_grid* g, int32_t* bid, const int32_t* rec, int32_t* rec_);
/* lakeflow.py:105-281  saddlesort(bid, is_border, z_prime, basin_saddle,
   basin_saddlenode, outlet, z); per-basin arrays have N entries */
int pff_saddlesort(pff_ctx* ctx, const pff_grid* g, const int32_t* bid, uint8_t* is_border, float* z_prime,
                   int64_t* basin_saddle, int32_t* basin_saddlenode, int64_t* outlet, const float* z);
/* lakeflow.py:284-304  reroute_jump(rec, outlet) */
int pff_reroute_jump(pff_ctx* ctx, const pff_grid* g, int32_t* rec, const int64_t* outlet);
/* lakeflow.py:405-438  reroute_carve(rec, rec_, rec__, tag, tag_, saddlenode,
   outlet, change): rec <- carved receivers of the working copy rec_ */
int pff_reroute_carve(pff_ctx* ctx, const pff_grid* g, int32_t* rec, int32_t* rec_, int32_t* rec__, uint8_t* tag,
                      uint8_t* tag_, const int32_t* saddlenode, const int64_t* outlet);

/* Per-call statistics of pff_reroute_flow (host struct). */
typedef struct {
  int64_t ndep0;      /* depressions before rerouting (lakeflow.py:493) */
  int32_t n_loop;     /* outer loop trips run */
  int32_t n_iter;     /* trips that rerouted (ran saddlesort) */
  int64_t ndep[64];   /* Ndep_bis of each trip (lakeflow.py:501) */
  int32_t carve_iters[64]; /* tag-propagation rounds of each carve */
} pff_lake_status;

/* lakeflow.py:466-544  reroute_flow(bid, rec, rec_, rec__, z, z_prime,
   is_border, outlet, basin_saddle, basin_saddlenode, tag, tag_, change, carve).
   Arrays as in the reference; on return rec holds the rerouted receivers
   (carve) and rec_ the working copy (the jump-mode result), bid is reset to
   its initial labels (lakeflow.py:537).  Optional: bid_first (device, N)
   receives the first trip's basin labels; st (host) the statistics.  Syncs
   once per outer trip to read the depression count. */
int pff_reroute_flow(pff_ctx* ctx, const pff_grid* g, int32_t* bid, int32_t* rec, int32_t* rec_, int32_t* rec__,
                     const float* z, float* z_prime, uint8_t* is_border, int64_t* outlet, int64_t* basin_saddle,
                     int32_t* basin_saddlenode, uint8_t* tag, uint8_t* tag_, int32_t carve, int32_t* bid_first,
                     pff_lake_status* st);

/* Arms a log of the lake outlets for the next pff_reroute_flow call on this
   context (then disarms; that call syncs).  After each trip's saddle and
   cycle passes (lakeflow.py:520, the state after :255-281), every basin b of
   the trip (the labels held by their own pit; once the receivers contain a
   cycle, every slot with an outlet or a saddle node) appends
   (b, outlet[b], basin_saddlenode[b]) to the device arrays
   basin/outlet/saddlenode (capacity entries; entries past it are counted,
   not written).  trip_counts (host, max_trips) receives the cumulative
   number of entries at the end of each trip, -1 for trips not run.  No
   reference counterpart: the reference keeps only the last trip's arrays. */
int pff_ctx_lake_log(pff_ctx* ctx, int32_t* basin, int64_t* outlet, int32_t* saddlenode, int64_t capacity,
                     int64_t* trip_counts, int32_t max_trips);

/* ---- flow: fill -------------------------------------------------------- */
/* fill_topo.py:99-125  fill_z_add_delta(zh, h, z_, receivers, receivers_, receivers__, epsilon) */
int pff_fill_z_add_delta(pff_ctx* ctx, const pff_grid* g, float* zh, float* h, float* z_, const int32_t* rcv,
                         int32_t* rcv_, int32_t* rcv__, double epsilon);
/* fill_topo.py:48-78  topofill(router, epsilon, custom_z): out = filled z
   (out may alias z: FlowRouter.fill_z, flowfields.py:295) */
int pff_topofill(pff_ctx* ctx, const pff_grid* g, const float* z, float* out, const int32_t* rcv, int32_t* rcv_,
                 int32_t* rcv__, double epsilon);

/* ---- graphflood -------------------------------------------------------- */
/* gf_hydrodynamics.py:23-75  diffuse_Q_constant_prec(zh, Q, temp) */
int pff_diffuse_Q_constant_prec(pff_ctx* ctx, const pff_grid* g, const pff_hydro* p, const float* zh, float* Q,
                                float* temp);
/* gf_hydrodynamics.py:78-129  graphflood_core_cte_mannings(h, zh, dh, rec, Q) */
int pff_graphflood_core_cte_mannings(pff_ctx* ctx, const pff_grid* g, const pff_hydro* p, float* h, float* zh,
                                     float* dh, const int32_t* rec, const float* Q);

/* Per-call result of the GraphFlood drivers (host struct). */
typedef struct {
  int64_t iterations;      /* iterations run by this call */
  double residual;         /* max |h(t) - h(t - check_every)| at the last check, m */
  double first_residual;   /* the same at the first check */
  int32_t converged;       /* 1 when residual < tol at a check */
  int32_t checks;          /* number of residual checks */
  int64_t nan_cells;       /* cells whose h is NaN at the end (h < 0 makes h^(5/3) NaN) */
  double above_frac;       /* fraction of cells with |h(t) - h(t - check_every)| > tol (or NaN), last check */
  double first_above_frac; /* the same at the first check */
  double mean_change;      /* mean |h(t) - h(t - check_every)| over non-NaN cells (/N), last check, m */
  double first_mean_change;/* the same at the first check */
  double q50, q90, q99, q999; /* quantiles of |h(t) - h(t - check_every)| over non-NaN cells, last
                                 check: upper edge of the 1/8-decade histogram bin holding each, m */
  int64_t nan_iteration;   /* iteration count at the first check that saw a NaN h (0: none) */
} pff_gf_status;

/* gf_fields.py:111-141  Flooder.run_graphflood(N): n_iter iterations of
   compute_receivers(zh) -> reroute_flow(carve) -> fill_z_add_delta(fill_eps)
   -> n_diffuse x diffuse_Q_constant_prec -> graphflood_core_cte_mannings.
   zh (router.z), h and Q are updated in place; rec receives the receivers of
   the last iteration.  Scratch fields live in the context.  Syncs once per
   lakeflow trip.  st (host, nullable). */
int pff_graphflood_run(pff_ctx* ctx, const pff_grid* g, const pff_hydro* p, float* zh, float* h, float* Q,
                       int32_t* rec, int32_t n_iter, int32_t n_diffuse, double fill_eps, pff_gf_status* st);
/* The same iteration until it is steady, checked every check_every
   iterations on d = |h(t) - h(t - check_every)|: with frac = 0, when
   max d < tol; with 0 < frac < 1, when at most frac*N cells have d > tol.
   Stops at max_iter iterations or when h has NaN cells.  No reference
   counterpart: the reference runs a fixed count (gf_fields.py:126). */
int pff_graphflood_to_steady(pff_ctx* ctx, const pff_grid* g, const pff_hydro* p, float* zh, float* h, float* Q,
                             int32_t* rec, double tol, double frac, int32_t check_every, int64_t max_iter,
                             int32_t n_diffuse, double fill_eps, pff_gf_status* st);

/* ---- ragged batch of independent grids ---------------------------------- */
/* One grid of a batch: 'normal' boundaries, its own nx, ny; device arrays. */
typedef struct {
  int32_t nx, ny;
  const float* z;   /* in, nx*ny elevations */
  int32_t* rcv;     /* out, nx*ny receivers (tile-local indices) */
  float* Q;         /* out, nx*ny discharge */
} pff_tile;

/* Per-call result of pff_batch_route (host struct). */
typedef struct {
  int64_t cells;          /* sum of the tiles' cells */
  int64_t canvas_cells;   /* cells of the packed canvas, padding included */
  int32_t canvas_nx, canvas_ny;
  int64_t ndep0;          /* initial depressions, all tiles */
  int32_t n_loop;         /* lakeflow trips run over the canvas */
  int32_t per_tile;       /* 1 when the batch was run tile by tile (see below) */
  int32_t converged;      /* 1 when every accumulation converged */
  int64_t capped_tiles;   /* tiles whose lakeflow stopped at their trip cap with depressions left */
} pff_batch_status;

/* No reference counterpart: the reference routes one grid per process
   (environment.py:28-29).  For each tile, the result of FlowRouter(nx, ny,
   dx, 'normal') -> compute_receivers() -> reroute_flow(carve) ->
   accumulate_constant_Q(value, area) (receivers.py:15-58,
   lakeflow.py:466-544, flowfields.py:254-292), for all tiles in one launch
   sequence: the tiles are packed into one canvas separated by padding cells,
   routed together, and unpacked.  lakeflow: 0 no reroute_flow, 1 carve
   (the default of flowfields.py:228), 2 jump (whose result the reference
   leaves in receivers_, so rcv keeps the unrerouted receivers, Q8).
   Results equal per-tile runs bit for bit: each tile keeps its own lakeflow
   trip cap (a tile that reaches it stops there, as its own loop would) and
   the cell 42 that the "no outlet" sentinel names.  When the receivers stop
   being a forest, or lakeflow is 2, the batch runs tile by tile instead
   (st->per_tile = 1).  All tiles share dx.  Syncs (once per lakeflow trip,
   and at the end).  st (host, nullable). */
int pff_batch_route(pff_ctx* ctx, int32_t ntiles, const pff_tile* tiles, float dx, int32_t lakeflow, double value,
                    int32_t area, pff_batch_status* st);

/* ---- GraphFlood warm-start surrogate ------------------------------------
   No reference counterpart (the reference has no learned component).  A
   small U-Net on BF16 activations (x[n][y][x][c], c contiguous, stored as
   uint16_t bit patterns) with FP32 accumulation on the matrix cores.  3x3
   convolutions: stride 1, zero padding 1, channels multiples of 32.  Master
   weights are f32 w[cout][cin][3][3]; pff_conv_pack_weights makes the bf16
   forward (wf[9][cout][cin]) and backward-data (wt[9][cin][cout], taps
   mirrored) copies.  Gradients accumulate (+=) into dw, db. */
int pff_conv_pack_weights(pff_ctx* ctx, int32_t cin, int32_t cout, const float* w, uint16_t* wf, uint16_t* wt);
int pff_conv3x3_fwd(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout, const uint16_t* x,
                    const uint16_t* wf, const float* bias, int32_t relu, uint16_t* y);
/* Inference form of the forward convolution that reads its input from the
   layers in front of it, so that their outputs never go through memory.
   src->mode 0: x [n][h][w][cin] (as pff_conv3x3_fwd); 1: concat(upsample2(x),
   skip) with x [n][h/2][w/2][ca] and skip [n][h][w][cin-ca] (pff_upsample2 +
   pff_concat); 2: the f32 planes [nplanes][n*h*w] rounded to bf16 and zero
   padded to cin = 32 channels (pff_pack_input); 3: x [n][h][w][cin] of which
   only channels 0..7 are read (the others taken as zero; cout = 32, no
   head).  head_out non-NULL (cout = 32): y is not written and head_out[p] =
   head_b[0] + sum_c head_w[c] * act[p][c] (pff_head1x1 of the activated
   output).  Mode 2 with at most 8 planes, cout = 32 and no head sums over
   (tap, channel 0..7) as mode 3 does, and gives the same bits as
   pff_pack_input followed by mode 3; otherwise each mode gives the same bits
   as the unfused sequence.  Needs w a multiple of 32 and h a multiple of
   16. */
typedef struct {
  int32_t mode;
  const uint16_t* x;
  const uint16_t* skip;
  const float* planes;
  int32_t ca, nplanes;
  const float* head_w;
  const float* head_b;
  float* head_out;
} pff_conv_src;
int pff_conv3x3_fwd_fused(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout,
                          const pff_conv_src* src, const uint16_t* wf, const float* bias, int32_t relu, uint16_t* y);
int pff_conv3x3_bwd_data(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout,
                         const uint16_t* dy, const uint16_t* wt, uint16_t* dx);
int pff_conv3x3_bwd_weight(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout,
                           const uint16_t* x, const uint16_t* dy, float* dw, float* db);
/* dy[i] = 0 where y[i] <= 0 (ReLU backward, in place) */
int pff_relu_bwd(pff_ctx* ctx, int64_t n, const uint16_t* y, uint16_t* dy);
/* 2x2 average pooling of an (n, h, w, c) tensor (h, w even) and its backward */
int pff_avgpool2(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t c, const uint16_t* x, uint16_t* y);
int pff_avgpool2_bwd(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t c, const uint16_t* dy, uint16_t* dx);
/* nearest 2x upsampling of an (n, h, w, c) tensor and its backward */
int pff_upsample2(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t c, const uint16_t* x, uint16_t* y);
int pff_upsample2_bwd(pff_ctx* ctx, int32_t n, int32_t h, int32_t w, int32_t c, const uint16_t* dy, uint16_t* dx);
/* channel concatenation y = [a, b] over npix pixels, and its split */
int pff_concat(pff_ctx* ctx, int64_t npix, int32_t ca, int32_t cb, const uint16_t* a, const uint16_t* b,
               uint16_t* y);
int pff_concat_bwd(pff_ctx* ctx, int64_t npix, int32_t ca, int32_t cb, const uint16_t* dy, uint16_t* da,
                   uint16_t* db);
/* 1x1 head to one f32 channel, y = b + x.w (c <= 256), and its backward (dx nullable) */
int pff_head1x1(pff_ctx* ctx, int64_t npix, int32_t c, const uint16_t* x, const float* w, const float* b, float* y);
int pff_head1x1_bwd(pff_ctx* ctx, int64_t npix, int32_t c, const uint16_t* x, const float* w, const float* dy,
                    uint16_t* dx, float* dw, float* db);
/* nplanes f32 planes [nplanes][npix] -> bf16 pixels of c channels (zero padded) */
int pff_pack_input(pff_ctx* ctx, int64_t npix, int32_t nplanes, int32_t c, const float* planes, uint16_t* x);

/* ---- packed keys ------------------------------------------------------- */
/* f32_i32_struct.py:123-135  pack_full_float_index(arr, ft, it) */
int pff_pack_full_float_index(pff_ctx* ctx, int64_t n, int64_t* arr, const float* ft, const int32_t* it);
/* f32_i32_struct.py:105-121  unpack_full_float_index(arr, ft, it) */
int pff_unpack_full_float_index(pff_ctx* ctx, int64_t n, const int64_t* arr, float* ft, int32_t* it);

/* ---- grid debug maps --------------------------------------------------- */
/* neighbourer_flat.py:112-115  fill_edges(edges): edge class 0..8 per cell (u8) */
int pff_fill_edges(pff_ctx* ctx, const pff_grid* g, uint8_t* edges);
/* neighbourer_flat.py:608-611  flow_out_nodes(outnodes): can_leave_domain per cell (u8 0/1) */
int pff_flow_out_nodes(pff_ctx* ctx, const pff_grid* g, uint8_t* outnodes);

/* ---- utilities --------------------------------------------------------- */
/* util_taichi.py:18-35 swap_arrays / :39-44 init_arange (4-byte elements) */
int pff_swap_arrays(pff_ctx* ctx, int64_t n, void* a, void* b);
int pff_init_arange(pff_ctx* ctx, int64_t n, int32_t* a);

#ifdef __cplusplus
}
#endif
#endif

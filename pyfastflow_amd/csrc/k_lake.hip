// k_lake.hip -- lakeflow on gfx950: depression count, basin labels by
// compacted pointer jumping, the five saddlesort passes with packed 64-bit
// atomicMin, and carving by tag doubling over the saved jump levels.
#include "pff_block.hpp"
#include "pff_internal.hpp"

namespace pff {

// ---------------------------------------------------------------------------
// depression_counter (lakeflow.py:24-44): #{i : rcv[i]==i and not can_leave(i)}
// ---------------------------------------------------------------------------
// A fixed grid of 512 blocks strides over the cells; each block sums its
// pits and adds them with one atomic (a returning or not, every atomic on the
// one counter queues at its cache line: one per block, not one per pit).
template <int MODE>
__global__ __launch_bounds__(kBT) void k_depcount(GridD g, const int32_t* __restrict__ rcv, int64_t* out) {
  int mine = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBT + threadIdx.x; i < g.N; i += (int64_t)gridDim.x * kBT) {
    if (rcv[i] == (int32_t)i) {
      int row, col;
      rc_of(g, (int)i, row, col);
      mine += can_leave<MODE>(g, (int)i, row, col) ? 0 : 1;
    }
  }
  block_count(mine, out);
}

int depression_count(pff_ctx* ctx, const GridD& g, const int32_t* rcv, int64_t* out) {
  int64_t* w = ctx->dev_words + W_COUNT2;
  PFF_TRY(check_hip(hipMemsetAsync(w, 0, 8, ctx->stream), "memset"));
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_depcount<M>, dim3(grid_for(g.N, kBT, 512)), dim3(kBT), 0,
                                               ctx->stream, g, rcv, w));
  PFF_TRY(check_hip(hipMemcpyAsync(ctx->host_words + W_COUNT2, w, 8, hipMemcpyDeviceToHost, ctx->stream), "d2h"));
  // the list lengths the last exact trip recorded, read with the same wait
  static_assert(W_HXLIVE == W_HMARK + 1 && W_HCTAG == W_HMARK + 2, "hint words are adjacent");
  PFF_TRY(check_hip(hipMemcpyAsync(ctx->host_words + W_HMARK, ctx->dev_words + W_HMARK, 24, hipMemcpyDeviceToHost,
                                   ctx->stream), "d2h"));
  PFF_TRY(check_hip(hipStreamSynchronize(ctx->stream), "sync"));
  *out = ctx->host_words[W_COUNT2];
  // words never written (no exact trip yet) read 0 from the context's
  // zeroed block: keep the start values then
  if (ctx->host_words[W_HMARK] > 0 || ctx->host_words[W_HXLIVE] > 0 || ctx->host_words[W_HCTAG] > 0) {
    ctx->hint_marked = ctx->host_words[W_HMARK];
    ctx->hint_xlive = ctx->host_words[W_HXLIVE];
    ctx->hint_ctag = ctx->host_words[W_HCTAG];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// basin_id_init (lakeflow.py:46-61) and one synchronous propagate_basin pass
// (lakeflow.py:63-80)
// ---------------------------------------------------------------------------
template <int MODE>
__global__ void k_basin_init(GridD g, int32_t* __restrict__ bid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  int row, col;
  rc_of(g, i, row, col);
  bid[i] = can_leave<MODE>(g, i, row, col) ? 0 : i + 1;
}

int launch_basin_id_init(pff_ctx* ctx, const GridD& g, int32_t* bid) {
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_basin_init<M>, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g,
                                               bid));
  return check_hip(hipGetLastError(), "basin_id_init");
}

__global__ void k_propagate(int32_t N, const int32_t* __restrict__ bid_in, const int32_t* __restrict__ rec_in,
                            int32_t* __restrict__ bid_out, int32_t* __restrict__ rec_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int r = rec_in[i];
  if ((unsigned)r >= (unsigned)N) r = i;
  bid_out[i] = bid_in[r];
  int rr = rec_in[r];
  if ((unsigned)rr >= (unsigned)N) rr = r;
  rec_out[i] = rr;
}

int launch_propagate_basin(pff_ctx* ctx, const GridD& g, const int32_t* bid_in, const int32_t* rec_in, int32_t* bid_out,
                           int32_t* rec_out) {
  hipLaunchKernelGGL(k_propagate, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g.N, bid_in, rec_in, bid_out,
                     rec_out);
  return check_hip(hipGetLastError(), "propagate_basin");
}

// ---------------------------------------------------------------------------
// saddlesort (lakeflow.py:105-281)
// ---------------------------------------------------------------------------
// pass 1 (:132-157): border flag and z'
template <int MODE>
__global__ void k_zprime(GridD g, const int32_t* __restrict__ bid, const float* __restrict__ z,
                         uint8_t* __restrict__ isb, float* __restrict__ zp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  int row, col;
  rc_of(g, i, row, col);
  if (can_leave<MODE>(g, i, row, col)) {
    // the reference never sets is_border here, so its zero-initialised
    // field reads False; the flag is written so that a reused buffer reads
    // the same
    isb[i] = 0;
    zp[i] = z[i];
    return;
  }
  const int bi = bid[i];
  bool border = false;
  float zn = 1e9f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int j = neighbour<MODE>(g, i, row, col, k);
    if (j < 0) continue;
    if (bid[j] != bi) {
      border = true;
      zn = fminf(zn, z[j]);
    }
  }
  isb[i] = border ? 1 : 0;
  zp[i] = border ? fmaxf(z[i], zn) : 1e9f;
}

// pass 2 init (:164-167)
__global__ void k_saddle_init(int32_t N, int64_t* __restrict__ saddle, int64_t* __restrict__ outlet,
                              int32_t* __restrict__ snode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  saddle[i] = INVALID_KEY;
  outlet[i] = INVALID_KEY;
  snode[i] = -1;
}

// pass 2 (:170-194): lexicographic min of (z'_i, bid_j) per basin, packed
// 64-bit signed atomicMin
template <int MODE>
__global__ void k_saddle_min(GridD g, const int32_t* __restrict__ bid, const uint8_t* __restrict__ isb,
                             const float* __restrict__ zp, int64_t* saddle) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N || !isb[i]) return;
  int row, col;
  rc_of(g, i, row, col);
  const int tb = bid[i];
  if (tb >= g.N) return;  // basin N (custom mode, cell N-1 interior) has no per-basin slot
  int64_t res = INVALID_KEY;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int j = neighbour<MODE>(g, i, row, col, k);
    if (j < 0) continue;
    const int bj = bid[j];
    if (bj != tb) {
      const int64_t c = pack_key(zp[i], bj);
      res = c < res ? c : res;
    }
  }
  if (res == INVALID_KEY) return;
  atomicMin((long long*)&saddle[tb], (long long)res);
}

// pass 3 (:199-219): saddle node; the reference's racy store is made
// deterministic as "largest cell index" with atomicMax
template <int MODE>
__global__ void k_saddle_node(GridD g, const int32_t* __restrict__ bid, const uint8_t* __restrict__ isb,
                              const float* __restrict__ zp, const int64_t* __restrict__ saddle, int32_t* snode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N || !isb[i]) return;
  const int bi = bid[i];
  if (bi == 0 || bi >= g.N) return;
  int row, col;
  rc_of(g, i, row, col);
  const int64_t key = saddle[bi];
  const float tz = key_float(key);
  const int tb = key_index(key);
  bool here = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int j = neighbour<MODE>(g, i, row, col, k);
    if (j < 0) continue;
    if (bid[j] == tb && zp[i] == tz) here = true;
  }
  if (here) atomicMax(&snode[bi], i);
}

// pass 4 (:226-250): outlet = lowest other-basin neighbour of the saddle node
// (first direction wins ties, strict < from 1e9)
template <int MODE>
__global__ void k_outlet(GridD g, const int32_t* __restrict__ bid, const float* __restrict__ z,
                         const int64_t* __restrict__ saddle, const int32_t* __restrict__ snode,
                         int64_t* __restrict__ outlet) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < 1 || b >= g.N) return;
  if (saddle[b] == INVALID_KEY) return;
  const int node = snode[b];
  if (node < 0) return;
  int row, col;
  rc_of(g, node, row, col);
  float tz = 1e9f;
  int rr = -1;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int j = neighbour<MODE>(g, node, row, col, k);
    if (j < 0) continue;
    if (bid[j] != b && tz > z[j]) {
      tz = z[j];
      rr = j;
    }
  }
  if (rr > -1) {
    const int64_t c = pack_key(tz, rr);
    if (c < outlet[b]) outlet[b] = c;
  }
}

// pass 5 (:255-281) in two phases.  The reference loop reads outlets other
// basins may invalidate in the same loop; in serial ascending order that
// read only matters for the basin holding cell 42 (the sentinel's index):
// it is also removed when it drains into a lower-id basin that was removed.
//
// holds42(b): b is the basin of the cell the sentinel decodes to, i.e. cell
// 42 of the grid, or of b's own tile on a tile canvas (f42 marks those).
__device__ __forceinline__ bool holds42(int32_t N, const int32_t* __restrict__ bid, const uint8_t* __restrict__ f42,
                                        int b) {
  return f42 ? f42[b] != 0 : (N > 42 && bid[42] == b);
}

// basin of the node an outlet key names; an invalid key decodes to cell 42
// (of the grid, or of the asking basin b's tile on a canvas)
__device__ __forceinline__ int outlet_basin(int32_t N, const int32_t* __restrict__ bid,
                                            const uint8_t* __restrict__ f42, int64_t key, int b) {
  if (f42 && key == INVALID_KEY) return f42[b] ? b : -1;
  const int r = key_index(key);
  return (r >= 0 && r < N) ? bid[r] : -1;
}

__global__ void k_cycle_flag(int32_t N, const int32_t* __restrict__ bid, const int64_t* __restrict__ outlet,
                             uint8_t* __restrict__ inv, const uint8_t* __restrict__ f42) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= N) return;
  uint8_t f = 0;
  if (b >= 1 && outlet[b] != INVALID_KEY) {
    const int r1 = key_index(outlet[b]);
    const int b1 = (r1 >= 0 && r1 < N) ? bid[r1] : -1;
    if (b1 > 0 && b1 < N) {
      const int b2 = outlet_basin(N, bid, f42, outlet[b1], b);
      if (b2 == b && b1 < b) f = 1;
    }
  }
  inv[b] = f;
}

__global__ void k_cycle_apply(int32_t N, const int32_t* __restrict__ bid, int64_t* outlet, int64_t* saddle,
                              int32_t* snode, const uint8_t* __restrict__ inv, const uint8_t* __restrict__ f42) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < 1 || b >= N) return;
  bool kill = inv[b] != 0;
  if (!kill && holds42(N, bid, f42, b) && outlet[b] != INVALID_KEY) {
    const int r1 = key_index(outlet[b]);
    const int b1 = (r1 >= 0 && r1 < N) ? bid[r1] : -1;
    if (b1 > 0 && b1 < b && inv[b1]) kill = true;
  }
  if (kill) {
    outlet[b] = INVALID_KEY;
    saddle[b] = INVALID_KEY;
    snode[b] = -1;
  }
}

int launch_saddlesort(pff_ctx* ctx, const GridD& g, const int32_t* bid, uint8_t* isb, float* zp, int64_t* saddle,
                      int32_t* snode, int64_t* outlet, const float* z) {
  void* inv;
  PFF_TRY(scratch(ctx, S_B0, (size_t)g.N, &inv));
  const dim3 gr(grid_for(g.N)), bl(kBlock);
  hipStream_t s = ctx->stream;
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_zprime<M>, gr, bl, 0, s, g, bid, z, isb, zp));
  hipLaunchKernelGGL(k_saddle_init, gr, bl, 0, s, g.N, saddle, outlet, snode);
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_saddle_min<M>, gr, bl, 0, s, g, bid, isb, zp, saddle));
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_saddle_node<M>, gr, bl, 0, s, g, bid, isb, zp, saddle, snode));
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_outlet<M>, gr, bl, 0, s, g, bid, z, saddle, snode, outlet));
  hipLaunchKernelGGL(k_cycle_flag, gr, bl, 0, s, g.N, bid, outlet, (uint8_t*)inv, g.F42);
  hipLaunchKernelGGL(k_cycle_apply, gr, bl, 0, s, g.N, bid, outlet, saddle, snode, (const uint8_t*)inv, g.F42);
  return check_hip(hipGetLastError(), "saddlesort");
}

// reroute_jump (lakeflow.py:284-304): rec[b-1] = outlet node of basin b
__global__ void k_reroute_jump(int32_t N, int32_t* rec, const int64_t* __restrict__ outlet) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < 1 || b >= N) return;
  const int64_t o = outlet[b];
  if (o == INVALID_KEY) return;
  rec[b - 1] = key_index(o);
}

int launch_reroute_jump(pff_ctx* ctx, const GridD& g, int32_t* rec, const int64_t* outlet) {
  hipLaunchKernelGGL(k_reroute_jump, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g.N, rec, outlet);
  return check_hip(hipGetLastError(), "reroute_jump");
}

// ---------------------------------------------------------------------------
// Carve (lakeflow.py:307-438).  The reference tags, for every saddle node s,
// the set {rec^k(s) : k < 2^t} after t rounds of "tag rec(i) for tagged i,
// then rec <- rec o rec", until a round adds nothing; then it reverses the
// tagged links and joins each saddle node to its outlet.  Here the tagged
// set is a list; round t jumps every listed node by 2^t using the pointer
// levels saved by the labelling, so the work is proportional to the tagged
// path lengths instead of N per round.
// ---------------------------------------------------------------------------

// seed: tag every valid saddle node (init_reroute_carve, :323-330); with
// mark, only the saddle nodes on marked cells
__global__ __launch_bounds__(kBT) void k_carve_seed(int32_t N, const int32_t* __restrict__ snode, uint32_t* tagw,
                                                    int32_t* list, int64_t* count, const uint8_t* __restrict__ mark) {
  constexpr int IT = 16;
  int32_t v[IT];
  bool take[IT];
  const int64_t base = (int64_t)blockIdx.x * IT * kBT;
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int64_t b = base + j * kBT + threadIdx.x;
    take[j] = false;
    v[j] = 0;
    if (b < N) {
      const int s = snode[b];
      v[j] = s;
      if (s >= 0 && (!mark || mark[s])) take[j] = atomicCAS(&tagw[s], 0u, 1u) == 0u;
    }
  }
  block_append<IT>(v, take, list, count);
}

// words: [0] tagged count, [1] count before this round, [2] count before the
// previous round
__global__ void k_carve_prep(int64_t* w) {
  w[2] = w[1];
  w[1] = w[0];
}

// P_t(i): LEVEL[t][i] while i was listed in the labelling, else its final
// pointer
__device__ __forceinline__ int jump_level(const int32_t* levels, const int32_t* fin, const int32_t* Pfinal, int N,
                                          int t, int i) {
  return (t <= fin[i]) ? levels[(size_t)t * N + i] : Pfinal[i];
}

// round t: every node tagged before the round tags its 2^t-th successor
__global__ __launch_bounds__(kBT) void k_carve_round(int32_t N, const int32_t* __restrict__ levels,
                                                     const int32_t* __restrict__ fin, const int32_t* __restrict__ Pfinal,
                                                     uint32_t* tagw, int32_t* list, int64_t* w, int32_t t) {
  constexpr int IT = 16;
  const int64_t n = w[1];
  if (t > 0 && n == w[2]) return;  // the previous round tagged nothing: converged
  for (int64_t base = (int64_t)blockIdx.x * IT * kBT; base < n; base += (int64_t)gridDim.x * IT * kBT) {
    int32_t v[IT];
    bool take[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int64_t e = base + j * kBT + threadIdx.x;
      take[j] = false;
      v[j] = 0;
      if (e < n) {
        const int i = list[e];
        const int jj = jump_level(levels, fin, Pfinal, N, t, i);
        v[j] = jj;
        if (tagw[jj] == 0u) take[j] = atomicCAS(&tagw[jj], 0u, 1u) == 0u;
      }
    }
    block_append<IT>(v, take, list, w);
  }
}

// The rounds t = 0 .. rounds-1 of k_carve_prep + k_carve_round in one block,
// for a short tag list: the same words w (tagged count, count before this
// round, count before the previous round), the same list, stopped at the
// first round that tags nothing (every later round would return at once).
// Any list length is handled.  *record (nullable) receives the final count.
constexpr int kCB = 1024;

__global__ __launch_bounds__(kCB) void k_carve_rounds_block(int32_t N, const int32_t* __restrict__ levels,
                                                            const int32_t* __restrict__ fin,
                                                            const int32_t* __restrict__ Pfinal, uint32_t* tagw,
                                                            int32_t* list, int64_t* w, int rounds, int64_t* record) {
  int64_t c0 = w[0], c1 = w[1], c2 = w[2];
  for (int t = 0; t < rounds; ++t) {
    c2 = c1;
    c1 = c0;
    const int64_t n = c1;
    if (t > 0 && n == c2) break;  // the previous round tagged nothing: converged
    for (int64_t base = 0; base < n; base += kCB) {  // uniform per block
      const int64_t e = base + threadIdx.x;
      bool take = false;
      int jj = 0;
      if (e < n) {
        const int i = list[e];
        jj = jump_level(levels, fin, Pfinal, N, t, i);
        if (tagw[jj] == 0u) take = atomicCAS(&tagw[jj], 0u, 1u) == 0u;
      }
      int tot;
      const int off = block_exclusive_scan<kCB>(take ? 1 : 0, &tot);
      if (take) list[c0 + off] = jj;
      c0 += tot;
    }
    __syncthreads();  // this round's tags and list entries before the next round reads them
  }
  if (threadIdx.x == 0) {
    w[0] = c0;
    w[1] = c0;
    w[2] = c0;
    if (record) *record = c0;
  }
}

// finalise_reroute_carve (:390-397): reverse every tagged link i -> rec_in[i].
// On a receiver graph with a cycle a tagged cell can have two tagged donors;
// the reference's parallel loop then keeps one of them, the rule here is the
// largest donor index (the last writer in index order).  Pass 1 records the
// largest donor + 2 in the tag word (tagged words are nonzero either way),
// pass 2 writes it.
__global__ void k_carve_reverse(int32_t N, const int32_t* __restrict__ list, const int64_t* w, uint32_t* tagw,
                                const int32_t* __restrict__ rec_in) {
  const int64_t n = w[0];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = list[e];
    const int r = rec_in[i];
    if (r != i && (unsigned)r < (unsigned)N && tagw[r]) atomicMax(&tagw[r], (uint32_t)i + 2u);
  }
}

__global__ void k_carve_reverse_apply(const int32_t* __restrict__ list, const int64_t* w,
                                      const uint32_t* __restrict__ tagw, int32_t* rec) {
  const int64_t n = w[0];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = list[e];
    const uint32_t v = tagw[i];
    if (v >= 2u) rec[i] = (int32_t)(v - 2u);
  }
}

// finalise_reroute_carve (:399-403): saddle node -> outlet node
__global__ void k_carve_link(int32_t N, const int64_t* __restrict__ outlet, const int32_t* __restrict__ snode,
                             int32_t* rec) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= N) return;
  const int64_t o = outlet[b];
  const int sn = snode[b];
  if (o != INVALID_KEY && sn >= 0 && sn < N) rec[sn] = key_index(o);
}

// copy the carved receivers of tagged nodes to the working copy and clear tags
__global__ void k_carve_sync(const int32_t* __restrict__ list, const int64_t* w, uint32_t* tagw,
                             const int32_t* __restrict__ rec, int32_t* rec_work, uint8_t* tag_bytes) {
  const int64_t n = w[0];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = list[e];
    if (rec_work) rec_work[i] = rec[i];
    if (tag_bytes) tag_bytes[i] = 1;
    tagw[i] = 0u;
  }
}

__global__ void k_copy_i32(int64_t n, const int32_t* __restrict__ a, int32_t* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

// Full carve for one trip.  rec must equal rec_in on entry (the driver keeps
// them equal); on return rec holds the carved receivers and, if rec_work is
// given, it is brought up to date too.
int carve_from_levels(pff_ctx* ctx, const GridD& g, int32_t* rec, const int32_t* rec_in, const int32_t* levels,
                      const int32_t* fin, const int32_t* Pfinal, const int32_t* snode, const int64_t* outlet,
                      int32_t* rec_work, uint8_t* tag_bytes, int npass, int32_t* iters) {
  void *tagw, *list;
  PFF_TRY(scratch(ctx, S_TAG, (size_t)g.N * 4, &tagw, true));
  PFF_TRY(scratch(ctx, S_MISC, (size_t)g.N * 4, &list));
  int64_t* w = ctx->dev_words + W_COUNT0;  // 3 consecutive words
  PFF_TRY(check_hip(hipMemsetAsync(w, 0, 3 * 8, ctx->stream), "memset"));
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_carve_seed, dim3((g.N + 16 * kBT - 1) / (16 * kBT)), dim3(kBT), 0, s, g.N, snode, (uint32_t*)tagw,
                     (int32_t*)list, w, (const uint8_t*)nullptr);
  if (ctx->hint_ctag <= kShortList) {
    // the last exact trip's carve tagged few cells: all rounds in one block
    hipLaunchKernelGGL(k_carve_rounds_block, dim3(1), dim3(kCB), 0, s, g.N, levels, fin, Pfinal, (uint32_t*)tagw,
                       (int32_t*)list, w, npass + 1, ctx->dev_words + W_HCTAG);
  } else {
    for (int t = 0; t <= npass; ++t) {
      hipLaunchKernelGGL(k_carve_prep, dim3(1), dim3(1), 0, s, w);
      hipLaunchKernelGGL(k_carve_round, dim3(stride_grid()), dim3(kBT), 0, s, g.N, levels, fin, Pfinal,
                         (uint32_t*)tagw, (int32_t*)list, w, t);
    }
    PFF_TRY(check_hip(hipMemcpyAsync(ctx->dev_words + W_HCTAG, w, 8, hipMemcpyDeviceToDevice, s), "copy"));
  }
  hipLaunchKernelGGL(k_carve_reverse, dim3(stride_grid()), dim3(kBlock), 0, s, g.N, (const int32_t*)list, w,
                     (uint32_t*)tagw, rec_in);
  hipLaunchKernelGGL(k_carve_reverse_apply, dim3(stride_grid()), dim3(kBlock), 0, s, (const int32_t*)list, w,
                     (const uint32_t*)tagw, rec);
  hipLaunchKernelGGL(k_carve_link, dim3(grid_for(g.N)), dim3(kBlock), 0, s, g.N, outlet, snode, rec);
  hipLaunchKernelGGL(k_carve_sync, dim3(stride_grid()), dim3(kBlock), 0, s, (const int32_t*)list, w,
                     (uint32_t*)tagw, rec, rec_work, tag_bytes);
  if (iters) *iters = -1;
  return check_hip(hipGetLastError(), "carve");
}

// ---------------------------------------------------------------------------
// Active-region versions used by the reroute_flow driver.  A tile is active
// while it holds a cell of a closed basin (label != 0); cells draining to an
// open edge keep label 0 and their receivers for the rest of the call, so
// their tiles are skipped.  Per-basin passes run over the list of pits
// (basin id = pit cell + 1) instead of all N basin slots.
// ---------------------------------------------------------------------------
constexpr int TSL = 64;  // tile side, as in the tiled contraction

// every basin entry (lakeflow.py:164-167 initialises all of them): in the
// first trip the pits touch almost every cache line of the three arrays, so
// one streaming pass of whole lines is cheaper than the scattered per-pit
// stores; also zeroes the words of this trip (clear0, clear1, and the two
// words at clear2)
__global__ __launch_bounds__(kBT) void k_saddle_init_all(int32_t n, int64_t* __restrict__ saddle,
                                                         int64_t* __restrict__ outlet, int32_t* __restrict__ snode,
                                                         int64_t* clear0, int64_t* clear1, int64_t* clear2) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *clear0 = 0;
    *clear1 = 0;
    clear2[0] = 0;
    clear2[1] = 0;
  }
  for (int64_t b = (int64_t)blockIdx.x * kBT + threadIdx.x; b < n; b += (int64_t)gridDim.x * kBT) {
    saddle[b] = INVALID_KEY;
    outlet[b] = INVALID_KEY;
    snode[b] = -1;
  }
}

// also zeroes the words of this trip (clear0, clear1, and the two words at
// clear2; nullable).  Basin N (custom mode, cell N-1 an interior pit) has no
// per-basin slot and is skipped.
__global__ __launch_bounds__(kBT) void k_saddle_init_list(int32_t N, const int32_t* __restrict__ pits,
                                                          const int64_t* np, int64_t* saddle, int64_t* outlet,
                                                          int32_t* snode, int64_t* clear0, int64_t* clear1,
                                                          int64_t* clear2) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (clear0) *clear0 = 0;
    if (clear1) *clear1 = 0;
    if (clear2) {
      clear2[0] = 0;
      clear2[1] = 0;
    }
  }
  const int64_t n = *np;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int b = pits[w];
    if (b <= 0 || b >= N) continue;
    saddle[b] = INVALID_KEY;
    outlet[b] = INVALID_KEY;
    snode[b] = -1;
  }
}

// Per-basin minima of packed saddle keys.  Each lane offers at most one
// (basin, key) pair; a segmented minimum over runs of lanes of the same basin
// leaves one offer per run, which goes to a small LDS table per block (one
// global atomic per basin and block at the flush) or, when the table has no
// room for the basin, straight to a global atomicMin after a plain read shows
// that it can lower the stored minimum.
constexpr int kSaddleSlots = 1024;  // a 64x64 tile holds ~550 basins in the first trip at 4096^2

struct SaddleTable {
  int32_t b[kSaddleSlots];
  long long k[kSaddleSlots];
};

__device__ __forceinline__ void saddle_table_init(SaddleTable& t) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  for (int q = tid; q < kSaddleSlots; q += kBT) {
    t.b[q] = -1;
    t.k[q] = (long long)INVALID_KEY;
  }
  __syncthreads();
}

// keyb < 0: no offer from this lane (every lane of the wave must call)
__device__ __forceinline__ void saddle_offer(SaddleTable& t, int keyb, int64_t res, int64_t* saddle) {
  const int lane = (threadIdx.x + threadIdx.y * blockDim.x) & 63;
  if (keyb < 0) keyb = -1 - lane;  // lanes without an offer never match another lane
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int ob = __shfl_down(keyb, d);
    const int64_t ov = __shfl_down(res, d);
    if (lane + d < 64 && ob == keyb && ov < res) res = ov;
  }
  const int prevb = __shfl_up(keyb, 1);
  if (keyb >= 0 && (lane == 0 || prevb != keyb)) {
    bool done = false;
    int slot = (int)(((uint32_t)keyb * 2654435761u) >> 22);  // 10-bit hash
    for (int probe = 0; probe < 8 && !done; ++probe, slot = (slot + 1) & (kSaddleSlots - 1)) {
      const int prev = atomicCAS(&t.b[slot], -1, keyb);
      if (prev == -1 || prev == keyb) {
        atomicMin(&t.k[slot], (long long)res);
        done = true;
      }
    }
    if (!done && res < saddle[keyb]) atomicMin((long long*)&saddle[keyb], (long long)res);
  }
}

// every stored minimum is read before the first atomic: a read issued after
// an atomic waits for that atomic's round trip as well
__device__ __forceinline__ void saddle_table_flush(SaddleTable& t, int64_t* saddle) {
  __syncthreads();
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  constexpr int Q = kSaddleSlots / kBT;
  int b[Q];
  long long v[Q], cur[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    b[k] = t.b[tid + k * kBT];
    v[k] = t.k[tid + k * kBT];
  }
#pragma unroll
  for (int k = 0; k < Q; ++k) cur[k] = b[k] >= 0 ? (long long)saddle[b[k]] : 0;
#pragma unroll
  for (int k = 0; k < Q; ++k)
    if (b[k] >= 0 && v[k] < cur[k]) atomicMin((long long*)&saddle[b[k]], v[k]);
}

// slot of basin b in the table, or -1 when its offers went to the global
// minimum (the probe sequence of saddle_offer: a basin that found a slot
// keeps it, one that found none finds none later, as slots are never freed)
__device__ __forceinline__ int saddle_table_find(const SaddleTable& t, int b) {
  int slot = (int)(((uint32_t)b * 2654435761u) >> 22);
  for (int probe = 0; probe < 8; ++probe, slot = (slot + 1) & (kSaddleSlots - 1)) {
    const int v = t.b[slot];
    if (v == b) return slot;
    if (v == -1) return -1;
  }
  return -1;
}

// A tiled z' pass lists every border cell of a closed basin (cell index,
// below 2^28) and after them, with bit 31 set (kSnCand), the cells that can
// be saddle nodes; the list pass of the next trip skips the marked entries,
// the saddle-node pass of this trip takes only them
constexpr uint32_t kSnCand = 0x80000000u;

// z', border flag and the offer of cell i (not an outlet cell); returns true
// when i is a border cell of a closed basin (the cells the saddle passes use)
template <int MODE>
__device__ __forceinline__ bool zprime_cell(const GridD& g, int i, int row, int col, float zi,
                                            const int32_t* __restrict__ bid, const float* __restrict__ z,
                                            uint8_t* __restrict__ isb, float* __restrict__ zp, int& keyb,
                                            int64_t& res) {
  const int bi = bid[i];
  bool border = false;
  float zn = 1e9f;
  int nbj[4];
  // the neighbours' elevations are loaded with their labels (one round trip
  // instead of a second one that waits for the label comparison)
  int jn[4];
  float zj[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    jn[k] = neighbour<MODE>(g, i, row, col, k);
    zj[k] = jn[k] >= 0 ? z[jn[k]] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    nbj[k] = -1;
    if (jn[k] < 0) continue;
    const int bj = bid[jn[k]];
    if (bj != bi) {
      border = true;
      zn = fminf(zn, zj[k]);
      nbj[k] = bj;
    }
  }
  isb[i] = border ? 1 : 0;
  const float zpi = border ? fmaxf(zi, zn) : 1e9f;
  zp[i] = zpi;
  if (!border || bi == 0 || bi >= g.N) return false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (nbj[k] < 0) continue;
    const int64_t c = pack_key(zpi, nbj[k]);
    res = c < res ? c : res;
  }
  if (res != INVALID_KEY) keyb = bi;
  return true;
}

// z', border flag and offer of a cell whose neighbours' elevations, labels
// and link validity are already in registers (d = 0 up, 1 left, 2 right,
// 3 down, the order of neighbour()): the arithmetic of zprime_cell
__device__ __forceinline__ bool zprime_cell_held(int N, int i, float zi, int bi, const float (&zj)[4],
                                                 const int (&bj)[4], const bool (&ok)[4], uint8_t* __restrict__ isb,
                                                 float* __restrict__ zp, int& keyb, int64_t& res) {
  bool border = false;
  float zn = 1e9f;
  int nbj[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    nbj[k] = -1;
    if (ok[k] && bj[k] != bi) {
      border = true;
      zn = fminf(zn, zj[k]);
      nbj[k] = bj[k];
    }
  }
  isb[i] = border ? 1 : 0;
  const float zpi = border ? fmaxf(zi, zn) : 1e9f;
  zp[i] = zpi;
  if (!border || bi == 0 || bi >= N) return false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (nbj[k] < 0) continue;
    const int64_t c = pack_key(zpi, nbj[k]);
    res = c < res ? c : res;
  }
  if (res != INVALID_KEY) keyb = bi;
  return true;
}

// passes 1+2 fused (lakeflow.py:132-194) over the active tiles: z', border
// flag and the packed minimum of (z'_i, bid_j) per basin; the border cells of
// closed basins are appended to blist in row-major order (blist may be null).
// One block per 64x64 tile; wave w takes rows 16w..16w+15 of the tile.  A wave
// whose rows and columns, with a one-cell frame, lie inside the grid loads
// every elevation, label (and code) it needs before its first store (on gfx9
// a wait for a load also waits for every store issued before it) and takes
// the left and right neighbours from the neighbour lanes (the frame columns
// by readlane); there every link is the plain D4 one (modes 0-2) or valid
// when both codes are nonzero (modes 3, 4), as neighbour() gives away from
// the border.  Other waves go cell by cell through neighbour().
constexpr int kZRows = TSL / 4;  // rows per wave
template <int MODE>
__global__ __launch_bounds__(kBT) void k_zprime_min_t(GridD g, int tiles_x, const uint32_t* __restrict__ tact,
                                                      const int32_t* __restrict__ bid, const float* __restrict__ z,
                                                      uint8_t* __restrict__ isb, float* __restrict__ zp,
                                                      int64_t* saddle, int32_t* blist, int64_t* nblist,
                                                      unsigned long long* prof) {
  static_assert(kZRows <= 32, "frame columns are held in lanes 0..kZRows-1 and 32..31+kZRows");
  constexpr bool kCodes = MODE == 3 || MODE == 4;
  __shared__ SaddleTable t;
  int bx_, by_;
  xcd_block(bx_, by_);
  if (tact && tact[by_ * tiles_x + bx_] == 0u) return;  // uniform per block
  saddle_table_init(t);
  constexpr int R = kZRows;
  int32_t lv[R];
  bool lt[R];
  const int lane = threadIdx.x;
  const int col0 = bx_ * 64, col = col0 + lane;
  const int r0 = by_ * TSL + kZRows * threadIdx.y;
  const int nx = g.nx;
  if (col0 > 0 && col0 + 64 < nx && r0 > 0 && r0 + R < g.ny) {  // uniform per wave
    float zz[R + 2];
    int bb[R + 2];
    uint32_t cc[R + 2];
#pragma unroll
    for (int k = 0; k < R + 2; ++k) {
      const size_t c = (size_t)(r0 - 1 + k) * nx + col;
      zz[k] = z[c];
      bb[k] = bid[c];
      if constexpr (kCodes) cc[k] = g.B[c];
    }
    // lanes 0..R-1: column col0-1 of row r0+lane; lanes 32..: column col0+64
    const int hk = lane & 31;
    const size_t hcell = (size_t)(r0 + hk) * nx + (lane < 32 ? col0 - 1 : col0 + 64);
    float hz = 0.f;
    int hb = 0;
    uint32_t hc = 0u;
    if (hk < R) {
      hz = z[hcell];
      hb = bid[hcell];
      if constexpr (kCodes) hc = g.B[hcell];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int i = (r0 + k) * nx + col;
      const float zi = zz[k + 1];
      const int bi = bb[k + 1];
      float zl = __shfl_up(zi, 1), zr = __shfl_down(zi, 1);
      int bl = __shfl_up(bi, 1), br = __shfl_down(bi, 1);
      zl = lane == 0 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hz), k)) : zl;
      zr = lane == 63 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hz), 32 + k)) : zr;
      bl = lane == 0 ? __builtin_amdgcn_readlane(hb, k) : bl;
      br = lane == 63 ? __builtin_amdgcn_readlane(hb, 32 + k) : br;
      bool ok[4] = {true, true, true, true};
      bool leave = false;
      if constexpr (kCodes) {
        const uint32_t ci = cc[k + 1];
        uint32_t cl = __shfl_up(ci, 1), cr = __shfl_down(ci, 1);
        cl = lane == 0 ? (uint32_t)__builtin_amdgcn_readlane((int)hc, k) : cl;
        cr = lane == 63 ? (uint32_t)__builtin_amdgcn_readlane((int)hc, 32 + k) : cr;
        ok[0] = ci != 0 && cc[k] != 0;
        ok[1] = ci != 0 && cl != 0;
        ok[2] = ci != 0 && cr != 0;
        ok[3] = ci != 0 && cc[k + 2] != 0;
        leave = MODE == 4 ? ci != 1 : ci == 3;  // can_leave()
      }
      int keyb = -1;
      int64_t res = INVALID_KEY;
      lv[k] = i;
      lt[k] = false;
      if (leave) {
        isb[i] = 0;
        zp[i] = zi;
      } else {
        const float zj[4] = {zz[k], zl, zr, zz[k + 2]};
        const int bj[4] = {bb[k], bl, br, bb[k + 2]};
        lt[k] = zprime_cell_held(g.N, i, zi, bi, zj, bj, ok, isb, zp, keyb, res);
      }
      saddle_offer(t, keyb, res, saddle);
    }
  } else {
    // no early exit inside the row loop: the whole wave (one row of 64
    // cells) takes part in the segmented minimum
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = r0 + r;
      lv[r] = 0;
      lt[r] = false;
      if (row >= g.ny) continue;  // uniform per wave
      const bool inside = col < nx;
      const int i = inside ? row * nx + col : 0;
      int keyb = -1;
      int64_t res = INVALID_KEY;
      if (inside) {
        const float zi = z[i];
        if (can_leave<MODE>(g, i, row, col)) {
          isb[i] = 0;
          zp[i] = zi;
        } else {
          lt[r] = zprime_cell<MODE>(g, i, row, col, zi, bid, z, isb, zp, keyb, res);
          lv[r] = i;
        }
      }
      saddle_offer(t, keyb, res, saddle);
    }
  }
  if (blist) {
    // saddle-node candidates: a saddle node's key (z'_i, basin across) is
    // its basin's global minimum, which is at most the tile's minimum and at
    // least the node's own key, so the tile's minimum has the node's z'.
    // Listed cells whose z' is the z' of their basin's tile minimum (and
    // those whose basin went to the global minimum directly, or that made no
    // offer) are listed again, marked.  Their labels and z' are read back
    // (this block wrote the z'), which keeps them out of registers.
    __syncthreads();  // every offer is in the table
    int cb[R];
    float cz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cb[r] = lt[r] ? bid[lv[r]] : 0;
      cz[r] = lt[r] ? zp[lv[r]] : 0.f;
    }
    bool cand[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cand[r] = lt[r];
      if (lt[r]) {
        const int slot = saddle_table_find(t, cb[r]);
        if (slot >= 0) cand[r] = cz[r] == key_float((int64_t)t.k[slot]);
      }
    }
    block_append_wave_major2<R>(lv, lt, cand, kSnCand, blist, nblist);
  }
  if (prof) {
    // profiling: algorithmic bytes of the tile (z, bid in; z', border flag
    // out; 4 per listed border cell)
    int mine = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) mine += lt[r] ? 1 : 0;
    int total;
    (void)block_exclusive_scan<kBT>(mine, &total);
    const int tw = min(TSL, g.nx - bx_ * TSL), th = min(TSL, g.ny - by_ * TSL);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomicAdd(prof, (unsigned long long)(13 * tw * th + 4 * total));
  }
  saddle_table_flush(t, saddle);
}

// the same over the previous trip's border list: a border cell of this trip
// was a border cell of the previous one (labels only merge), so the cells
// that are still border cells of closed basins are the new list
// R list entries per thread and chunk, one append per chunk: long lists take
// large chunks (few appends on the shared counter), short ones small chunks
// (enough blocks to fill the chip)
template <int MODE, int R>
__global__ __launch_bounds__(kBT) void k_zprime_list(GridD g, const int32_t* __restrict__ lin, const int64_t* nin,
                                                     const int32_t* __restrict__ bid, const float* __restrict__ z,
                                                     uint8_t* __restrict__ isb, float* __restrict__ zp,
                                                     int64_t* saddle, int32_t* lout, int64_t* nout) {
  __shared__ SaddleTable t;
  saddle_table_init(t);
  const int64_t n = *nin;
  constexpr int64_t chunk = (int64_t)kBT * R;
  for (int64_t base = (int64_t)blockIdx.x * chunk; base < n; base += (int64_t)gridDim.x * chunk) {
    int32_t lv[R];
    bool lt[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t e = base + r * kBT + threadIdx.x;
      lv[r] = 0;
      lt[r] = false;
      int keyb = -1;
      int64_t res = INVALID_KEY;
      if (e < n) {
        const int i = lin[e];
        if (i >= 0) {  // marked saddle-node candidates (bit 31) are second copies
          int row, col;
          rc_of(g, i, row, col);
          lt[r] = zprime_cell<MODE>(g, i, row, col, z[i], bid, z, isb, zp, keyb, res);
          lv[r] = i;
        }
      }
      saddle_offer(t, keyb, res, saddle);
    }
    block_append_ordered<R>(lv, lt, lout, nout);
  }
  saddle_table_flush(t, saddle);
}

// Pass 3 with pass 4's neighbour scan folded in, for the tiled trips: a
// saddle-node candidate also finds its lowest neighbour outside the basin
// (first direction wins ties, strict < from 1e9: the outlet the reference
// computes for the final saddle node, :226-250) and offers (i << 3) | k,
// k that neighbour's direction or 4 for none (also when the basin has no
// saddle key).  The largest offer is the largest candidate index with its
// own direction; k_outlet_list decodes it (cell indices below 2^28).
template <int MODE>
__device__ __forceinline__ void saddle_node_cell_packed(const GridD& g, int i, int row, int col, int bi,
                                                        const int32_t* __restrict__ bid, const float* __restrict__ z,
                                                        const float* __restrict__ zp,
                                                        const int64_t* __restrict__ saddle, int32_t* snode) {
  int jn[4], bj[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    jn[k] = neighbour<MODE>(g, i, row, col, k);
    bj[k] = jn[k] >= 0 ? bid[jn[k]] : -1;
  }
  const int64_t key = saddle[bi];
  if (zp[i] != key_float(key)) return;
  const int tb = key_index(key);
  bool here = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) here |= bj[k] >= 0 && bj[k] == tb;
  if (!here) return;
  int code = 4;
  if (key != INVALID_KEY) {
    float tz = 1e9f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (jn[k] < 0 || bj[k] == bi) continue;
      const float zj = z[jn[k]];
      if (tz > zj) {
        tz = zj;
        code = k;
      }
    }
  }
  atomicMax(&snode[bi], (i << 3) | code);
}

// over this trip's border list
// (cand_only: only the entries a tiled z' pass marked kSnCand, else the
// unmarked ones)
template <int MODE>
__global__ __launch_bounds__(kBT) void k_saddle_node_list(GridD g, const int32_t* __restrict__ list,
                                                          const int64_t* nlist, const int32_t* __restrict__ bid,
                                                          const float* __restrict__ z, const float* __restrict__ zp,
                                                          const int64_t* __restrict__ saddle, int32_t* snode,
                                                          int cand_only) {
  const int64_t n = *nlist;
  for (int64_t e = (int64_t)blockIdx.x * kBT + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBT) {
    const uint32_t v = (uint32_t)list[e];
    if (((v & kSnCand) != 0) != (cand_only != 0)) continue;
    const int i = (int)(v & ~kSnCand);
    int row, col;
    rc_of(g, i, row, col);
    saddle_node_cell_packed<MODE>(g, i, row, col, bid[i], bid, z, zp, saddle, snode);
  }
}

// pass 4 for pit b from its packed saddle-node offer: the saddle node, and
// the outlet key of the neighbour its direction names
template <int MODE>
__device__ __forceinline__ void outlet_of_packed(const GridD& g, int b, const float* __restrict__ z, int32_t* snode,
                                                 int64_t* __restrict__ outlet) {
  if (b <= 0 || b >= g.N) return;
  const int sp = snode[b];
  if (sp < 0) return;
  const int node = sp >> 3, code = sp & 7;
  snode[b] = node;
  if (code >= 4) return;
  int row, col;
  rc_of(g, node, row, col);
  const int rr = neighbour<MODE>(g, node, row, col, code);
  const int64_t c = pack_key(z[rr], rr);
  if (c < outlet[b]) outlet[b] = c;
}

// pass 5 (:255-281), first phase for pit b: 1 when b and the basin of its
// outlet drain into each other and that basin has the lower id
__device__ __forceinline__ uint8_t cycle_flag_of(int32_t N, int b, const int32_t* __restrict__ bid,
                                                 const int64_t* __restrict__ outlet, const uint8_t* __restrict__ f42) {
  if (outlet[b] == INVALID_KEY) return 0;
  const int r1 = key_index(outlet[b]);
  const int b1 = (r1 >= 0 && r1 < N) ? bid[r1] : -1;
  if (b1 <= 0 || b1 >= N) return 0;
  const int b2 = outlet_basin(N, bid, f42, outlet[b1], b);
  return (b2 == b && b1 < b) ? 1 : 0;
}

// second phase for pit b: remove the flagged outlets (and the basin holding
// cell 42 when it drains into a lower flagged basin); set the bit of every
// remaining saddle node in the tile-ordered bitmap wbits (128 words per
// 64x64 tile) for the carve's tag walks
__device__ __forceinline__ void cycle_apply_of(int32_t N, int b, const int32_t* __restrict__ bid, int64_t* outlet,
                                               int64_t* saddle, int32_t* snode, const uint8_t* __restrict__ inv,
                                               const uint8_t* __restrict__ f42, int32_t nx, int tiles_x,
                                               uint32_t* wbits) {
  bool kill = inv[b] != 0;
  if (!kill && holds42(N, bid, f42, b) && outlet[b] != INVALID_KEY) {
    const int r1 = key_index(outlet[b]);
    const int b1 = (r1 >= 0 && r1 < N) ? bid[r1] : -1;
    if (b1 > 0 && b1 < b && inv[b1]) kill = true;
  }
  if (kill) {
    outlet[b] = INVALID_KEY;
    saddle[b] = INVALID_KEY;
    snode[b] = -1;
  } else {
    const int sn = snode[b];
    if (sn >= 0 && sn < N) {
      const int row = sn / nx, col = sn - row * nx;
      const int li = (row % TSL) * TSL + col % TSL;
      const size_t word = (size_t)((row / TSL) * tiles_x + col / TSL) * (TSL * TSL / 32) + (li >> 5);
      atomicOr(&wbits[word], 1u << (li & 31));
    }
  }
}

// pass 4 over the pits, from the packed saddle-node offers
template <int MODE>
__global__ __launch_bounds__(kBT) void k_outlet_list(GridD g, const int32_t* __restrict__ pits, const int64_t* np,
                                                     const float* __restrict__ z, int32_t* snode,
                                                     int64_t* __restrict__ outlet) {
  const int64_t n = *np;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT)
    outlet_of_packed<MODE>(g, pits[w], z, snode, outlet);
}

// pass 5 over the pits, in two kernels (every flag is read after all are set)
__global__ __launch_bounds__(kBT) void k_cycle_flag_list(int32_t N, const int32_t* __restrict__ pits, const int64_t* np,
                                                         const int32_t* __restrict__ bid,
                                                         const int64_t* __restrict__ outlet, uint8_t* __restrict__ inv,
                                                         const uint8_t* __restrict__ f42) {
  const int64_t n = *np;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int b = pits[w];
    if (b < N) inv[b] = cycle_flag_of(N, b, bid, outlet, f42);
  }
}

__global__ __launch_bounds__(kBT) void k_cycle_apply_list(int32_t N, const int32_t* __restrict__ pits,
                                                          const int64_t* np, const int32_t* __restrict__ bid,
                                                          int64_t* outlet, int64_t* saddle, int32_t* snode,
                                                          const uint8_t* __restrict__ inv,
                                                          const uint8_t* __restrict__ f42, int32_t nx, int tiles_x,
                                                          uint32_t* wbits) {
  const int64_t n = *np;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int b = pits[w];
    if (b < N) cycle_apply_of(N, b, bid, outlet, saddle, snode, inv, f42, nx, tiles_x, wbits);
  }
}

// lake-outlet log: one entry per basin of the trip.  On a forest every basin
// is the label of its pit, which holds its own id; with all_slots (receivers
// with a cycle, where labels of cycle cells need not be held by their own
// cell, and every basin slot was reset this trip) every slot with an outlet
// or a saddle node is a basin.
__global__ __launch_bounds__(kBT) void k_log_outlets(int32_t N, const int32_t* __restrict__ bid,
                                                     const int64_t* __restrict__ outlet,
                                                     const int32_t* __restrict__ snode, int32_t* __restrict__ lb,
                                                     int64_t* __restrict__ lo, int32_t* __restrict__ ls,
                                                     int64_t capacity, int64_t* count, int all_slots) {
  for (int64_t base = (int64_t)blockIdx.x * kBT; base < N; base += (int64_t)gridDim.x * kBT) {
    const int64_t i = base + threadIdx.x;
    int32_t v[1] = {0};
    bool take[1] = {false};
    if (i < N) {
      if (all_slots) {
        if (outlet[i] != INVALID_KEY || snode[i] != -1) {
          v[0] = (int32_t)i;
          take[0] = true;
        }
      } else if (bid[i] == (int32_t)i + 1) {
        v[0] = (int32_t)i + 1;
        take[0] = true;
      }
    }
    // positions past the capacity are counted but not written
    int total;
    const int off = block_exclusive_scan<kBT>(take[0] ? 1 : 0, &total);
    __shared__ int64_t base_s;
    if (threadIdx.x == 0) base_s = total ? (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)total) : 0;
    __syncthreads();
    const int64_t p = base_s + off;
    if (take[0] && p < capacity) {
      const int b = v[0];
      lb[p] = b;
      lo[p] = b < N ? outlet[b] : INVALID_KEY;
      ls[p] = b < N ? snode[b] : -1;
    }
    __syncthreads();
  }
}

int launch_log_outlets(pff_ctx* ctx, const GridD& g, const int32_t* bid, const int64_t* outlet,
                       const int32_t* snode, int64_t* dev_count, bool all_slots) {
  const auto& L = ctx->lake_log;
  hipLaunchKernelGGL(k_log_outlets, dim3(grid_for(g.N, kBT, 8192)), dim3(kBT), 0, ctx->stream, g.N, bid, outlet, snode,
                     L.basin, L.outlet, L.snode, L.capacity, dev_count, all_slots ? 1 : 0);
  return check_hip(hipGetLastError(), "log_outlets");
}

// Saddles and outlets of one trip on the tiled path.  *border_cur (host) is
// the border list of the previous trip (0 or 1), or -1 for a full pass over
// the active tiles (the first trip); on return it names this trip's list.
// A list longer than half of the grid is also passed over in full (the
// coalesced pass is cheaper than the gathers then); its length is the one the
// last read of the device words saw.  Also zeroes the tag-list counter that
// carve_tiled appends to.
int saddlesort_active(pff_ctx* ctx, const GridD& g, const int32_t* bid, uint8_t* isb, float* zp, int64_t* saddle,
                      int32_t* snode, int64_t* outlet, const float* z, const uint32_t* tact, const int32_t* pits,
                      const int64_t* npits, int* border_cur) {
  void *inv, *bl0, *bl1;
  PFF_TRY(scratch(ctx, S_B0, (size_t)g.N, &inv));
  // a tiled pass lists every border cell of a closed basin (at most N) and a
  // marked copy of each saddle-node candidate among them: up to 2N entries
  PFF_TRY(scratch(ctx, S_BL0, (size_t)g.N * 8, &bl0));
  PFF_TRY(scratch(ctx, S_BL1, (size_t)g.N * 8, &bl1));
  int32_t* BL[2] = {(int32_t*)bl0, (int32_t*)bl1};
  static_assert(W_NBL1 == W_NBL0 + 1, "border list counters are adjacent words");
  int64_t* NB[2] = {ctx->dev_words + W_NBL0, ctx->dev_words + W_NBL1};
  const int in = *border_cur;
  const int out = in < 0 ? 0 : (in ^ 1);
  hipStream_t s = ctx->stream;
  const int tiles_x = (g.nx + TSL - 1) / TSL;
  const dim3 lg(1024), bl(kBT), cb(64, 4);
  // also zeroes the cycle flag and the exit count that this trip's carve
  // tags (saddle_tags_walk, the next use of them) count into
  static_assert(W_NEXIT == W_CYCLE + 1, "cycle flag and exit count are adjacent words");
  int64_t* cyc_words = ctx->dev_words + W_CYCLE;
  if (in < 0)  // first trip of the call: every tile is active
    hipLaunchKernelGGL(k_saddle_init_all, dim3(8192), bl, 0, s, g.N, saddle, outlet, snode, NB[out],
                       ctx->dev_words + W_NTAG, cyc_words);
  else
    hipLaunchKernelGGL(k_saddle_init_list, lg, bl, 0, s, g.N, pits, npits, saddle, outlet, snode, NB[out],
                       ctx->dev_words + W_NTAG, cyc_words);
  const bool full = in < 0 || ctx->host_words[W_NBL0 + in] * 2 > (int64_t)g.N;
  const dim3 lgz(2048);
  if (full) {
    const dim3 tt((g.nx + TSL - 1) / TSL, (g.ny + TSL - 1) / TSL);
    unsigned long long* prof = prof_begin(ctx, "k_zprime_min_t", 1.0);
    PFF_MODE_DISPATCH(g.mode, PFF_PROF_LAUNCH(ctx, k_zprime_min_t<M>, tt, cb, 0, s, g, tiles_x, tact, bid, z, isb, zp,
                                              saddle, BL[out], NB[out], prof));
  } else {
    if (ctx->host_words[W_NBL0 + in] > (int64_t)1 << 21)
      PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL((k_zprime_list<M, 16>), lgz, bl, 0, s, g, (const int32_t*)BL[in],
                                                   (const int64_t*)NB[in], bid, z, isb, zp, saddle, BL[out], NB[out]));
    else
      PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL((k_zprime_list<M, 2>), lgz, bl, 0, s, g, (const int32_t*)BL[in],
                                                   (const int64_t*)NB[in], bid, z, isb, zp, saddle, BL[out], NB[out]));
  }
  if (g.N > (1 << 28)) return fail(PFF_EINVAL, "saddlesort_active: grids above 2^28 cells are not supported");
  // after the tiled pass only the entries it marked can be saddle nodes
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_saddle_node_list<M>, lgz, bl, 0, s, g, (const int32_t*)BL[out],
                                               (const int64_t*)NB[out], bid, z, (const float*)zp,
                                               (const int64_t*)saddle, snode, full ? 1 : 0));
  *border_cur = out;
  uint32_t* wbits;
  uint8_t* tag;
  PFF_TRY(tag_state(ctx, g, &wbits, &tag, true));
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_outlet_list<M>, lg, bl, 0, s, g, pits, npits, z, snode, outlet));
  hipLaunchKernelGGL(k_cycle_flag_list, lg, bl, 0, s, g.N, pits, npits, bid, (const int64_t*)outlet, (uint8_t*)inv,
                     g.F42);
  hipLaunchKernelGGL(k_cycle_apply_list, lg, bl, 0, s, g.N, pits, npits, bid, outlet, saddle, snode,
                     (const uint8_t*)inv, g.F42, g.nx, tiles_x, wbits);
  return check_hip(hipGetLastError(), "saddlesort_active");
}

// The five saddle passes over the whole grid for the driver's exact trips
// (labels then need not be pit ids): as launch_saddlesort, with passes 1 and
// 2 fused into the tiled z' pass and its per-tile minimum table.  Basin 0
// (draining out) gets no saddle entry here; the driver never reads it.
int saddlesort_exact(pff_ctx* ctx, const GridD& g, const int32_t* bid, uint8_t* isb, float* zp, int64_t* saddle,
                     int32_t* snode, int64_t* outlet, const float* z) {
  void* inv;
  PFF_TRY(scratch(ctx, S_B0, (size_t)g.N, &inv));
  const dim3 gr(grid_for(g.N)), bl(kBlock);
  hipStream_t s = ctx->stream;
  const int tiles_x = (g.nx + TSL - 1) / TSL;
  const dim3 tt(tiles_x, (g.ny + TSL - 1) / TSL), cb(64, 4);
  hipLaunchKernelGGL(k_saddle_init, gr, bl, 0, s, g.N, saddle, outlet, snode);
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_zprime_min_t<M>, tt, cb, 0, s, g, tiles_x, (const uint32_t*)nullptr,
                                               bid, z, isb, zp, saddle, (int32_t*)nullptr, (int64_t*)nullptr,
                                               (unsigned long long*)nullptr));
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_saddle_node<M>, gr, bl, 0, s, g, bid, isb, zp, saddle, snode));
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_outlet<M>, gr, bl, 0, s, g, bid, z, saddle, snode, outlet));
  hipLaunchKernelGGL(k_cycle_flag, gr, bl, 0, s, g.N, bid, outlet, (uint8_t*)inv, g.F42);
  hipLaunchKernelGGL(k_cycle_apply, gr, bl, 0, s, g.N, bid, outlet, saddle, snode, (const uint8_t*)inv, g.F42);
  return check_hip(hipGetLastError(), "saddlesort_exact");
}

// tiled carve finalise (lakeflow.py:390-397): reverse the tagged links, over
// the tagged cells
__global__ __launch_bounds__(kBT) void k_carve_reverse_list(int32_t N, const int32_t* __restrict__ list,
                                                            const int64_t* nlist, const uint8_t* __restrict__ tag,
                                                            const int32_t* __restrict__ rec_in, int32_t* rec) {
  const int64_t n = *nlist;
  for (int64_t e = (int64_t)blockIdx.x * kBT + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBT) {
    const int i = list[e];
    const int r = rec_in[i];
    if (r != i && (unsigned)r < (unsigned)N && tag[r]) rec[r] = i;
  }
}

// saddle node -> outlet (lakeflow.py:399-403) over the pits.  Basin N (custom
// mode, cell N-1 an interior pit) has no per-basin slot: it has no outlet.
__global__ __launch_bounds__(kBT) void k_carve_link_list(int32_t N, const int32_t* __restrict__ pits,
                                                         const int64_t* np, const int64_t* __restrict__ outlet,
                                                         const int32_t* __restrict__ snode, int32_t* rec) {
  const int64_t n = *np;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int b = pits[w];
    if (b <= 0 || b >= N) continue;
    const int64_t o = outlet[b];
    const int sn = snode[b];
    if (o != INVALID_KEY && sn >= 0 && sn < N) rec[sn] = key_index(o);
  }
}

// every changed receiver belongs to a tagged cell: copy those to the working copy
// (and clears the tag bytes of the listed cells: zero between carves)
// (tdirty, nullable: the tiles whose receivers change are marked dirty for
// the ring-walk cache of the next trip's tag walks)
__global__ __launch_bounds__(kBT) void k_carve_sync_list(const int32_t* __restrict__ list, const int64_t* nlist,
                                                         const int32_t* __restrict__ rec, int32_t* __restrict__ rec_work,
                                                         uint8_t* __restrict__ tag, uint32_t* tdirty, int nx,
                                                         int tiles_x) {
  const int64_t n = *nlist;
  for (int64_t e = (int64_t)blockIdx.x * kBT + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBT) {
    const int i = list[e];
    if (rec_work) rec_work[i] = rec[i];
    tag[i] = 0;
    if (tdirty) {
      const int row = i / nx;
      const int t = (row / TSL) * tiles_x + (i - row * nx) / TSL;
      if (tdirty[t] == 0u) tdirty[t] = 1u;
    }
  }
}

// Carve for one trip on a forest of D4 receivers: the cells the reference
// tags (every cell on the path from a saddle node to its pit) are exactly the
// cells with a saddle node upstream, found by the tag walks (saddle_tags_walk)
// from the saddle-node bits that this trip's saddlesort_active set.  rec
// equals rec_in on entry; on return rec holds the carved receivers and
// rec_work (if given) the same; the tag bytes are cleared again.  *cycle is
// always 0 (rec_in was proved a forest by this trip's labels).
int carve_tiled(pff_ctx* ctx, const GridD& g, int32_t* rec, const int32_t* rec_in, const int32_t* bid,
                const int32_t* snode, const int64_t* outlet, int32_t* rec_work, const uint32_t* tact,
                const int32_t* pits, const int64_t* npits, int rounds, int32_t* cycle) {
  void* tl;
  uint32_t* wbits;
  uint8_t* tag;
  PFF_TRY(tag_state(ctx, g, &wbits, &tag, false));
  PFF_TRY(scratch(ctx, S_TLIST, (size_t)g.N * 4, &tl));
  // rec_in was proved a forest by this trip's labels: no cycle check needed.
  // The tag walks also list the tagged cells (the counter was zeroed by
  // this trip's saddlesort_active, which also set the saddle-node bits).
  const int64_t* ntag = ctx->dev_words + W_NTAG;
  PFF_TRY(saddle_tags_walk(ctx, g, rec_in, bid, wbits, tag, tact, rounds, (int32_t*)tl));
  *cycle = 0;
  const dim3 lg(2048), bl(kBT);
  hipLaunchKernelGGL(k_carve_reverse_list, lg, bl, 0, ctx->stream, g.N, (const int32_t*)tl, ntag,
                     (const uint8_t*)tag, rec_in, rec);
  hipLaunchKernelGGL(k_carve_link_list, dim3(1024), bl, 0, ctx->stream, g.N, pits, npits, outlet, snode, rec);
  uint16_t* rcache;
  uint32_t* tdirty;
  PFF_TRY(ring_cache(ctx, g, &rcache, &tdirty));
  hipLaunchKernelGGL(k_carve_sync_list, lg, bl, 0, ctx->stream, (const int32_t*)tl, ntag, (const int32_t*)rec,
                     rec_work, tag, tdirty, g.nx, (g.nx + TSL - 1) / TSL);
  ctx->tag_state_dirty = false;  // k_tag_local cleared the saddle bits, k_carve_sync_list the tags
  return check_hip(hipGetLastError(), "carve_tiled");
}

// Carve of a trip whose receivers may hold cycles (labels from
// labels_split).  The reference's tag loop runs until the tags stop changing
// (:431), so the tagged cells are every cell reachable from a saddle node.
// From a saddle node whose path reaches a root that is the path (rcvp keeps
// those links): the tiled saddle tags.  From a saddle node on a marked cell
// it stays among the marked cells: the jumping rounds of carve_from_levels
// over those cells alone.  A cell of the first kind has one tagged donor at
// most; the reversal of the second takes the largest donor, as
// carve_from_levels does.  Then every saddle node links to its outlet.
int carve_split(pff_ctx* ctx, const GridD& g, int32_t* rec, const int32_t* rec_in, const int32_t* bid,
                const int32_t* snode, const int64_t* outlet, int32_t* rec_work, const int32_t* levels,
                const int32_t* fin, const int32_t* Pfinal, const int32_t* rcvp, const uint8_t* mark, int npass) {
  void *tl, *tagw, *list;
  uint32_t* wbits;
  uint8_t* tag;
  PFF_TRY(tag_state(ctx, g, &wbits, &tag, true));
  PFF_TRY(scratch(ctx, S_TLIST, (size_t)g.N * 4, &tl));
  PFF_TRY(scratch(ctx, S_TAG, (size_t)g.N * 4, &tagw, true));
  PFF_TRY(scratch(ctx, S_MISC, (size_t)g.N * 4, &list));
  hipStream_t s = ctx->stream;
  int64_t* ntag = ctx->dev_words + W_NTAG;
  PFF_TRY(check_hip(hipMemsetAsync(ntag, 0, 8, s), "memset"));
  // paths that reach a root (a forest once the marked cells are roots)
  PFF_TRY(tiled_upstream(ctx, g, rcvp, 2, bid, snode, 0.0f, nullptr, nullptr, (uint8_t*)tag, nullptr, nullptr, 0,
                         (int32_t*)tl));
  const dim3 lg(2048), bl(kBT);
  hipLaunchKernelGGL(k_carve_reverse_list, lg, bl, 0, s, g.N, (const int32_t*)tl, (const int64_t*)ntag,
                     (const uint8_t*)tag, rcvp, rec);
  // the marked cells
  int64_t* w = ctx->dev_words + W_COUNT0;  // 3 consecutive words
  PFF_TRY(check_hip(hipMemsetAsync(w, 0, 3 * 8, s), "memset"));
  hipLaunchKernelGGL(k_carve_seed, dim3((g.N + 16 * kBT - 1) / (16 * kBT)), dim3(kBT), 0, s, g.N, snode,
                     (uint32_t*)tagw, (int32_t*)list, w, mark);
  if (ctx->hint_ctag <= kShortList) {
    // the last exact trip's carve tagged few cells: all rounds in one block
    hipLaunchKernelGGL(k_carve_rounds_block, dim3(1), dim3(kCB), 0, s, g.N, levels, fin, Pfinal, (uint32_t*)tagw,
                       (int32_t*)list, w, npass + 1, ctx->dev_words + W_HCTAG);
  } else {
    for (int t = 0; t <= npass; ++t) {
      hipLaunchKernelGGL(k_carve_prep, dim3(1), dim3(1), 0, s, w);
      hipLaunchKernelGGL(k_carve_round, dim3(stride_grid()), dim3(kBT), 0, s, g.N, levels, fin, Pfinal,
                         (uint32_t*)tagw, (int32_t*)list, w, t);
    }
    PFF_TRY(check_hip(hipMemcpyAsync(ctx->dev_words + W_HCTAG, w, 8, hipMemcpyDeviceToDevice, s), "copy"));
  }
  hipLaunchKernelGGL(k_carve_reverse, dim3(stride_grid()), dim3(kBlock), 0, s, g.N, (const int32_t*)list, w,
                     (uint32_t*)tagw, rec_in);
  hipLaunchKernelGGL(k_carve_reverse_apply, dim3(stride_grid()), dim3(kBlock), 0, s, (const int32_t*)list, w,
                     (const uint32_t*)tagw, rec);
  hipLaunchKernelGGL(k_carve_link, dim3(grid_for(g.N)), dim3(kBlock), 0, s, g.N, outlet, snode, rec);
  hipLaunchKernelGGL(k_carve_sync, dim3(stride_grid()), dim3(kBlock), 0, s, (const int32_t*)list, w,
                     (uint32_t*)tagw, rec, rec_work, (uint8_t*)nullptr);
  hipLaunchKernelGGL(k_carve_sync_list, lg, bl, 0, s, (const int32_t*)tl, (const int64_t*)ntag, (const int32_t*)rec,
                     rec_work, tag, (uint32_t*)nullptr, g.nx, (g.nx + TSL - 1) / TSL);
  ctx->tag_state_dirty = false;
  return check_hip(hipGetLastError(), "carve_split");
}

// Reference-shaped carve (lakeflow.py:405-438) for the module-level API.
// rec_ is the working copy on entry and is left holding the jumped pointers
// (the reference clobbers it too); rec__ receives the pre-carve copy; rec
// receives the carved receivers; tag (nullable) the tagged cells.
int reroute_carve(pff_ctx* ctx, const GridD& g, int32_t* rec, int32_t* rec_, int32_t* rec__, uint8_t* tag,
                  uint8_t* tag_, const int32_t* snode, const int64_t* outlet, int32_t* iters) {
  const int npass = npass_for(g.N);
  void *lv, *fin, *bid;
  PFF_TRY(scratch(ctx, S_LEVELS, (size_t)(npass + 1) * g.N * 4, &lv));
  PFF_TRY(scratch(ctx, S_I4, (size_t)g.N * 4, &fin));
  PFF_TRY(scratch(ctx, S_F0, (size_t)g.N * 4, &bid));
  const dim3 gr(grid_for(g.N)), bl(kBlock);
  hipLaunchKernelGGL(k_copy_i32, gr, bl, 0, ctx->stream, (int64_t)g.N, rec_, rec__);  // :428
  hipLaunchKernelGGL(k_copy_i32, gr, bl, 0, ctx->stream, (int64_t)g.N, rec_, rec);    // :427
  PFF_TRY(labels_with_levels(ctx, g, rec__, (int32_t*)bid, rec_, (int32_t*)lv, (int32_t*)fin));
  if (tag) PFF_TRY(check_hip(hipMemsetAsync(tag, 0, (size_t)g.N, ctx->stream), "memset"));
  PFF_TRY(carve_from_levels(ctx, g, rec, rec__, (const int32_t*)lv, (const int32_t*)fin, rec_, snode, outlet,
                            nullptr, tag, npass, iters));
  if (tag_ && tag)
    PFF_TRY(check_hip(hipMemcpyAsync(tag_, tag, (size_t)g.N, hipMemcpyDeviceToDevice, ctx->stream), "copy"));
  return 0;
}

}  // namespace pff

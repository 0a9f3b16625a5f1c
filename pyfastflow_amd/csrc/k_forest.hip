// k_forest.hip -- receiver-forest primitives on gfx950: roots (basin labels)
// and upstream sums (drainage area, carve path tags).
//
// Two implementations share the same semantics on forests:
//
//  * global compacted pointer jumping (exact for any receiver graph, including
//    cycles, where it reproduces ceil(log2 N)+1 synchronous passes);
//  * tiled contraction (forests only, detects cycles): 64x64 tiles are loaded
//    into LDS and contracted there by pointer jumping -- every cell learns the
//    last cell of its path inside the tile ("local end") and its in-tile
//    upstream sum.  Only local ends whose receiver leaves the tile ("exits",
//    a few percent of the cells) are then jumped globally, and a second tile
//    pass folds the inflow of each tile's entry cells back in.  Each cell is
//    read and written a small constant number of times, instead of once per
//    global pointer-jumping round.
#include "pff_block.hpp"
#include "pff_internal.hpp"

#include <algorithm>

namespace pff {

// ---------------------------------------------------------------------------
// division by the grid width with a multiply-high (Granlund & Montgomery 1994)
// ---------------------------------------------------------------------------
struct FastDiv {
  uint32_t m, s;  // s >= 1 (divisor >= 2)
};

static FastDiv make_fastdiv(uint32_t d) {
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  const uint64_t m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
  return FastDiv{(uint32_t)m, s};
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
  const uint32_t t = __umulhi(n, f.m);
  return (t + ((n - t) >> 1)) >> (f.s - 1);
}

// ---------------------------------------------------------------------------
// Global compacted pointer jumping.  Lists are processed in chunks of
// kChunk entries per block iteration; appends cost one atomic per chunk.
// ---------------------------------------------------------------------------
constexpr int kItems = 16;
constexpr int kChunk = kItems * kBT;  // 4096
// the scatter halves of the list rounds: 4 nodes per thread, so that the
// lists of later rounds (often ~1e5 nodes) still spread over ~100 blocks,
// with one append per 1024 nodes
constexpr int kRItems = 4;
constexpr int kRChunk = kRItems * kBT;

// label of the cells that tiled labels with cycle marking finds on or
// draining into a receiver cycle
constexpr int32_t kCycleLabel = -1;

// counts ---------------------------------------------------------------------
// init from all cells: S = 1, P = rcv (or -1 at roots), list = non-roots
__global__ __launch_bounds__(kBT) void k_count_init_all(int32_t N, const int32_t* __restrict__ rcv,
                                                        uint32_t* __restrict__ S, int32_t* __restrict__ P,
                                                        int32_t* __restrict__ list, int64_t* count) {
  int32_t val[kItems];
  bool take[kItems];
  const int64_t base = (int64_t)blockIdx.x * kChunk;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t t = base + j * kBT + threadIdx.x;
    take[j] = false;
    val[j] = (int32_t)t;
    if (t < N) {
      const int r = rcv[t];
      const bool root = (r == (int)t) || (unsigned)r >= (unsigned)N;
      S[t] = 1u;
      P[t] = root ? -1 : r;
      take[j] = !root;
    }
  }
  block_append<kItems>(val, take, list, count);
}

// phase 1 of a round: read pre-round values of each listed node
__global__ __launch_bounds__(kBT) void k_count_gather2(const int32_t* __restrict__ list, const int64_t* count,
                                                       const uint32_t* __restrict__ S, const int32_t* __restrict__ P,
                                                       int32_t* __restrict__ tgt, uint32_t* __restrict__ val,
                                                       int32_t* __restrict__ nxt, int64_t* next_count,
                                                       int64_t* first_empty, int32_t round) {
  const int64_t n = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *next_count = 0;
    if (n == 0 && first_empty) atomicMin((long long*)first_empty, (long long)round);
  }
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int m = list[w];
    const int pm = P[m];
    tgt[w] = pm;
    val[w] = S[m];
    nxt[w] = P[pm];
  }
}

// phase 2: scatter-add into the 2^k-th successor, advance, compact.  As
// the pointers converge, many nodes of a block add into the same few
// successors; their adds are summed per block in an LDS table first (one
// global atomic per distinct successor and block; a successor that finds no
// slot adds directly).  The sums are integers, so the order does not matter.
constexpr int kAddSlots = 1024;

__global__ __launch_bounds__(kBT) void k_count_scatter2(const int32_t* __restrict__ list, const int64_t* count,
                                                        uint32_t* S, int32_t* P, const int32_t* __restrict__ tgt,
                                                        const uint32_t* __restrict__ val,
                                                        const int32_t* __restrict__ nxt, int32_t* __restrict__ out,
                                                        int64_t* next_count) {
  __shared__ int32_t tk[kAddSlots];
  __shared__ uint32_t tv[kAddSlots];
  for (int q = threadIdx.x; q < kAddSlots; q += kBT) {
    tk[q] = -1;
    tv[q] = 0u;
  }
  __syncthreads();
  const int64_t n = *count;
  for (int64_t base = (int64_t)blockIdx.x * kRChunk; base < n; base += (int64_t)gridDim.x * kRChunk) {
    int32_t v[kRItems];
    bool take[kRItems];
#pragma unroll
    for (int j = 0; j < kRItems; ++j) {
      const int64_t w = base + j * kBT + threadIdx.x;
      take[j] = false;
      v[j] = 0;
      if (w < n) {
        const int m = list[w];
        const int t = tgt[w];
        const uint32_t a = val[w];
        if (a != 0u) {
          bool done = false;
          int slot = (int)(((uint32_t)t * 2654435761u) >> 22);  // 10-bit hash
          for (int probe = 0; probe < 4 && !done; ++probe, slot = (slot + 1) & (kAddSlots - 1)) {
            const int prev = atomicCAS(&tk[slot], -1, t);
            if (prev == -1 || prev == t) {
              atomicAdd(&tv[slot], a);
              done = true;
            }
          }
          if (!done) atomicAdd(&S[t], a);
        }
        const int q = nxt[w];
        P[m] = q;
        v[j] = m;
        take[j] = q >= 0;
      }
    }
    block_append<kRItems>(v, take, out, next_count);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < kAddSlots; q += kBT)
    if (tk[q] >= 0 && tv[q] != 0u) atomicAdd(&S[tk[q]], tv[q]);
}

// phase 2 for flags ("is anything upstream marked"): OR into the 2^k-th
// successor.  OR is idempotent, so a plain store of 1 into a zero target is
// enough and no atomics are needed; the doubling schedule of the sums
// delivers every marked descendant to every ancestor at least once.
__global__ __launch_bounds__(kBT) void k_flag_scatter2(const int32_t* __restrict__ list, const int64_t* count,
                                                       uint32_t* S, int32_t* P, const int32_t* __restrict__ tgt,
                                                       const uint32_t* __restrict__ val,
                                                       const int32_t* __restrict__ nxt, int32_t* __restrict__ out,
                                                       int64_t* next_count, unsigned long long* prof) {
  const int64_t n = *count;
  // profiling: algorithmic bytes per listed node (list, tgt, val, nxt, S[t]
  // in; P, list out)
  if (prof && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(prof, 28ull * (unsigned long long)n);
  for (int64_t base = (int64_t)blockIdx.x * kRChunk; base < n; base += (int64_t)gridDim.x * kRChunk) {
    int32_t v[kRItems];
    bool take[kRItems];
#pragma unroll
    for (int j = 0; j < kRItems; ++j) {
      const int64_t w = base + j * kBT + threadIdx.x;
      take[j] = false;
      v[j] = 0;
      if (w < n) {
        const int m = list[w];
        if (val[w] != 0u) {
          const int t = tgt[w];
          if (S[t] == 0u) S[t] = 1u;
        }
        const int q = nxt[w];
        P[m] = q;
        v[j] = m;
        take[j] = q >= 0;
      }
    }
    block_append<kRItems>(v, take, out, next_count);
  }
}

// labels ---------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(kBT) void k_label_init_all(GridD g, const int32_t* __restrict__ rec,
                                                        int32_t* __restrict__ bid, int32_t* __restrict__ P,
                                                        int32_t* __restrict__ fin, int32_t* __restrict__ level0,
                                                        int32_t* __restrict__ list, int64_t* count) {
  int32_t val[kItems];
  bool take[kItems];
  const int64_t base = (int64_t)blockIdx.x * kChunk;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t t = base + j * kBT + threadIdx.x;
    const int i = (int)t;
    take[j] = false;
    val[j] = i;
    if (t < g.N) {
      int row, col;
      rc_of(g, i, row, col);
      bid[i] = can_leave<MODE>(g, i, row, col) ? 0 : i + 1;
      int r = rec[i];
      if ((unsigned)r >= (unsigned)g.N) r = i;  // out-of-range receivers act as roots
      P[i] = r;
      if (level0) level0[i] = r;
      take[j] = r != i;
      if (fin) fin[i] = take[j] ? 0x7fffffff : 0;
    }
  }
  block_append<kItems>(val, take, list, count);
}

__global__ __launch_bounds__(kBT) void k_label_gather2(const int32_t* __restrict__ list, const int64_t* count,
                                                       const int32_t* __restrict__ bid, const int32_t* __restrict__ P,
                                                       int32_t* __restrict__ nb, int32_t* __restrict__ np,
                                                       int64_t* next_count, int64_t* first_empty, int32_t round) {
  const int64_t n = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *next_count = 0;
    if (n == 0 && first_empty) atomicMin((long long*)first_empty, (long long)round);
  }
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int m = list[w];
    const int pm = P[m];
    nb[w] = bid[pm];
    const int q = P[pm];
    np[w] = (q == pm) ? ~q : q;  // pm is a fixed point: m finishes this pass
  }
}

__global__ __launch_bounds__(kBT) void k_label_scatter2(const int32_t* __restrict__ list, const int64_t* count,
                                                        int32_t* bid, int32_t* P, const int32_t* __restrict__ nb,
                                                        const int32_t* __restrict__ np, int32_t* __restrict__ out,
                                                        int64_t* next_count, int32_t* __restrict__ level_next,
                                                        int32_t* __restrict__ fin, int32_t round, int64_t* self_loop) {
  const int64_t n = *count;
  for (int64_t base = (int64_t)blockIdx.x * kRChunk; base < n; base += (int64_t)gridDim.x * kRChunk) {
    int32_t v[kRItems];
    bool take[kRItems];
#pragma unroll
    for (int j = 0; j < kRItems; ++j) {
      const int64_t w = base + j * kBT + threadIdx.x;
      take[j] = false;
      v[j] = 0;
      if (w < n) {
        const int m = list[w];
        int q = np[w];
        const bool done = q < 0;
        if (done) q = ~q;
        bid[m] = nb[w];
        P[m] = q;
        if (level_next) level_next[m] = q;
        if (done && fin) fin[m] = round + 1;
        // a node that jumps onto itself lies on a cycle (on a forest only
        // roots point to themselves); the lists keep such a node live
        if (!done && q == m && self_loop) *self_loop = 1;
        v[j] = m;
        take[j] = !done;
      }
    }
    block_append<kRItems>(v, take, out, next_count);
  }
}

// round drivers: lists ping-pong between la/lb with counts in ca/cb
struct Lists {
  int32_t* l[2];
  int64_t* c[2];
  int cur = 0;
};

static unsigned list_grid() { return 1024; }

static int count_rounds(pff_ctx* ctx, Lists& L, uint32_t* S, int32_t* P, int32_t* tgt, uint32_t* val, int32_t* nxt,
                        int rounds, bool flags = false, int first_round = 0, int64_t* first_empty = nullptr) {
  for (int k = first_round; k < first_round + rounds; ++k) {
    const int a = L.cur, b = 1 - L.cur;
    hipLaunchKernelGGL(k_count_gather2, dim3(list_grid()), dim3(kBT), 0, ctx->stream, L.l[a], L.c[a], S, P, tgt, val,
                       nxt, L.c[b], first_empty, k);
    if (flags) {
      unsigned long long* prof = prof_begin(ctx, "k_flag_scatter2", 1.0);
      PFF_PROF_LAUNCH(ctx, k_flag_scatter2, dim3(list_grid()), dim3(kBT), 0, ctx->stream, L.l[a], L.c[a], S, P, tgt,
                      val, nxt, L.l[b], L.c[b], prof);
    }
    else
      hipLaunchKernelGGL(k_count_scatter2, dim3(list_grid()), dim3(kBT), 0, ctx->stream, L.l[a], L.c[a], S, P, tgt,
                         val, nxt, L.l[b], L.c[b]);
    L.cur = b;
  }
  return check_hip(hipGetLastError(), "count rounds");
}

// All rounds [first_round, first_round + rounds) of k_label_gather2 +
// k_label_scatter2 in one block, for short lists: per round the gather over
// the whole list, a barrier, the scatter with the compaction of the live
// nodes into the other list, a barrier.  The lists ping-pong between l0
// (current on entry, count *c0) and l1; afterwards the current list is l0 if
// rounds is even, else l1, and its count is in the matching word, as after
// the two-kernel rounds.  Any list length is handled.  *record (nullable)
// receives the length of the list on entry.
constexpr int kLB = 1024;

__global__ __launch_bounds__(kLB) void k_label_rounds_block(int32_t* l0, int32_t* l1, int64_t* c0, int64_t* c1,
                                                            int32_t* bid, int32_t* P, int32_t* __restrict__ nb,
                                                            int32_t* __restrict__ np, int32_t* levels, int32_t N,
                                                            int32_t* fin, int first_round, int rounds,
                                                            int64_t* first_empty, int64_t* self_loop,
                                                            int64_t* record) {
  int32_t* la = l0;
  int32_t* lb = l1;
  int64_t* cb = c1;
  int64_t n = *c0;
  if (record && threadIdx.x == 0) *record = n;
  for (int k = first_round; k < first_round + rounds; ++k) {
    if (n == 0 && first_empty && threadIdx.x == 0) atomicMin((long long*)first_empty, (long long)k);
    for (int64_t w = threadIdx.x; w < n; w += kLB) {
      const int m = la[w];
      const int pm = P[m];
      nb[w] = bid[pm];
      const int q = P[pm];
      np[w] = (q == pm) ? ~q : q;
    }
    __syncthreads();
    int32_t* lvl = levels ? levels + (size_t)(k + 1) * N : nullptr;
    int64_t out = 0;
    for (int64_t base = 0; base < n; base += kLB) {  // uniform per block
      const int64_t w = base + threadIdx.x;
      bool take = false;
      int m = 0;
      if (w < n) {
        m = la[w];
        int q = np[w];
        const bool done = q < 0;
        if (done) q = ~q;
        bid[m] = nb[w];
        P[m] = q;
        if (lvl) lvl[m] = q;
        if (done && fin) fin[m] = k + 1;
        if (!done && q == m && self_loop) *self_loop = 1;
        take = !done;
      }
      int tot;
      const int off = block_exclusive_scan<kLB>(take ? 1 : 0, &tot);
      if (take) lb[out + off] = m;
      out += tot;
    }
    if (threadIdx.x == 0) *cb = out;
    int32_t* tl = la;
    la = lb;
    lb = tl;
    cb = (cb == c1) ? c0 : c1;
    n = out;
    __syncthreads();
  }
}

__global__ void k_copy_word(const int64_t* src, int64_t* dst) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = *src;
}

static int label_rounds_run(pff_ctx* ctx, Lists& L, int32_t* bid, int32_t* P, int32_t* nb, int32_t* np,
                            int32_t* levels, int32_t N, int32_t* fin, int first_round, int rounds,
                            int64_t* self_loop = nullptr, bool one_block = false, int64_t* record = nullptr) {
  if (one_block) {
    const int a = L.cur;
    hipLaunchKernelGGL(k_label_rounds_block, dim3(1), dim3(kLB), 0, ctx->stream, L.l[a], L.l[1 - a], L.c[a],
                       L.c[1 - a], bid, P, nb, np, levels, N, fin, first_round, rounds, ctx->dev_words + W_EMPTY,
                       self_loop, record);
    if (rounds & 1) L.cur = 1 - a;
    return check_hip(hipGetLastError(), "label rounds (one block)");
  }
  if (record) hipLaunchKernelGGL(k_copy_word, dim3(1), dim3(64), 0, ctx->stream, (const int64_t*)L.c[L.cur], record);
  for (int k = first_round; k < first_round + rounds; ++k) {
    const int a = L.cur, b = 1 - L.cur;
    hipLaunchKernelGGL(k_label_gather2, dim3(list_grid()), dim3(kBT), 0, ctx->stream, L.l[a], L.c[a], bid, P, nb, np,
                       L.c[b], ctx->dev_words + W_EMPTY, k);
    int32_t* lvl = levels ? levels + (size_t)(k + 1) * N : nullptr;
    hipLaunchKernelGGL(k_label_scatter2, dim3(list_grid()), dim3(kBT), 0, ctx->stream, L.l[a], L.c[a], bid, P, nb,
                       np, L.l[b], L.c[b], lvl, fin, k, self_loop);
    L.cur = b;
  }
  return check_hip(hipGetLastError(), "label rounds");
}

static int ceil_log2i(int64_t n) {
  int r = 0;
  while ((1ll << r) < n) ++r;
  return r;
}

static int read_word(pff_ctx* ctx, const int64_t* dev, int64_t* out) {
  PFF_TRY(check_hip(hipMemcpyAsync(ctx->host_words + W_FLAG2, dev, 8, hipMemcpyDeviceToHost, ctx->stream), "d2h"));
  PFF_TRY(check_hip(hipStreamSynchronize(ctx->stream), "sync"));
  *out = ctx->host_words[W_FLAG2];
  return 0;
}

// exact upstream counts, any receiver graph; *cycle set when a cycle blocks completion
int upstream_count(pff_ctx* ctx, const GridD& g, const int32_t* rcv, uint32_t* cnt, int32_t* cycle) {
  const size_t n4 = (size_t)g.N * 4;
  void *P, *L0, *L1, *tgt, *val, *nxt;
  PFF_TRY(scratch(ctx, S_I0, n4, &P));
  PFF_TRY(scratch(ctx, S_WL0, n4, &L0));
  PFF_TRY(scratch(ctx, S_WL1, n4, &L1));
  PFF_TRY(scratch(ctx, S_I1, n4, &tgt));
  PFF_TRY(scratch(ctx, S_I2, n4, &val));
  PFF_TRY(scratch(ctx, S_I3, n4, &nxt));
  Lists L;
  L.l[0] = (int32_t*)L0;
  L.l[1] = (int32_t*)L1;
  L.c[0] = ctx->dev_words + W_COUNT0;
  L.c[1] = ctx->dev_words + W_COUNT1;
  PFF_TRY(check_hip(hipMemsetAsync(L.c[0], 0, 8, ctx->stream), "memset"));
  hipLaunchKernelGGL(k_count_init_all, dim3((g.N + kChunk - 1) / kChunk), dim3(kBT), 0, ctx->stream, g.N, rcv, cnt,
                     (int32_t*)P, L.l[0], L.c[0]);
  PFF_TRY(count_rounds(ctx, L, cnt, (int32_t*)P, (int32_t*)tgt, (uint32_t*)val, (int32_t*)nxt,
                       ceil_log2i(g.N) + 1));
  if (cycle) {
    int64_t live = 0;
    PFF_TRY(read_word(ctx, L.c[L.cur], &live));
    *cycle = live != 0;
  }
  return 0;
}

// exact labels (ceil(log2 N)+1 synchronous passes' result), optional levels
int labels_global(pff_ctx* ctx, const GridD& g, const int32_t* rec, int32_t* bid, int32_t* P, int32_t* levels,
                  int32_t* fin) {
  const size_t n4 = (size_t)g.N * 4;
  void *L0, *L1, *nbv, *npv;
  PFF_TRY(scratch(ctx, S_WL0, n4, &L0));
  PFF_TRY(scratch(ctx, S_WL1, n4, &L1));
  PFF_TRY(scratch(ctx, S_I2, n4, &nbv));
  PFF_TRY(scratch(ctx, S_I3, n4, &npv));
  Lists L;
  L.l[0] = (int32_t*)L0;
  L.l[1] = (int32_t*)L1;
  L.c[0] = ctx->dev_words + W_COUNT0;
  L.c[1] = ctx->dev_words + W_COUNT1;
  PFF_TRY(check_hip(hipMemsetAsync(L.c[0], 0, 8, ctx->stream), "memset"));
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_label_init_all<M>, dim3((g.N + kChunk - 1) / kChunk), dim3(kBT), 0,
                                               ctx->stream, g, rec, bid, P, fin, levels, L.l[0], L.c[0]));
  return label_rounds_run(ctx, L, bid, P, (int32_t*)nbv, (int32_t*)npv, levels, g.N, fin, 0, ceil_log2i(g.N) + 1);
}

int basin_labels(pff_ctx* ctx, const GridD& g, const int32_t* rec, int32_t* bid, int32_t* rec_out) {
  void* P;
  if (rec_out) {
    P = rec_out;
  } else {
    PFF_TRY(scratch(ctx, S_I4, (size_t)g.N * 4, &P));
  }
  return labels_global(ctx, g, rec, bid, (int32_t*)P, nullptr, nullptr);
}

int labels_with_levels(pff_ctx* ctx, const GridD& g, const int32_t* rec, int32_t* bid, int32_t* Pfinal,
                       int32_t* levels, int32_t* fin) {
  return labels_global(ctx, g, rec, bid, Pfinal, levels, fin);
}

int npass_for(int32_t N) { return ceil_log2i(N) + 1; }

// After tiled labels with cycle marking: the cells labelled kCycleLabel (on
// or draining into a receiver cycle) get the set-up of the synchronous
// passes (init label, pointer, level 0, not finished) and are listed; every
// other cell keeps its root's label.  rcvp = rec with the listed cells made
// roots (-1), mark = 1 on the listed cells.
template <int MODE>
__global__ __launch_bounds__(kBT) void k_cycle_split(GridD g, const int32_t* __restrict__ rec, int32_t* __restrict__ bid,
                                                     int32_t* __restrict__ P, int32_t* __restrict__ level0,
                                                     int32_t* __restrict__ fin, int32_t* __restrict__ rcvp,
                                                     uint8_t* __restrict__ mark, int32_t* __restrict__ list,
                                                     int64_t* count) {
  int32_t val[kItems];
  bool take[kItems];
  const int64_t base = (int64_t)blockIdx.x * kChunk;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t t = base + j * kBT + threadIdx.x;
    const int i = (int)t;
    take[j] = false;
    val[j] = i;
    if (t < g.N) {
      int r = rec[i];
      if ((unsigned)r >= (unsigned)g.N) r = i;
      const bool m = bid[i] == kCycleLabel;
      rcvp[i] = m ? -1 : r;
      mark[i] = m ? 1 : 0;
      if (m) {
        int row, col;
        rc_of(g, i, row, col);
        bid[i] = can_leave<MODE>(g, i, row, col) ? 0 : i + 1;
        P[i] = r;
        level0[i] = r;
        fin[i] = 0x7fffffff;
        take[j] = true;
      }
    }
  }
  block_append<kItems>(val, take, list, count);
}

// Labels with the semantics of ceil(log2 N)+1 synchronous passes on
// receivers that may hold cycles.  A cell whose path reaches a root gets its
// root's label from the tiled contraction (the passes give the same); only
// the cells on or draining into a cycle run the passes, and their successors
// are such cells too, so the passes over them alone give the labels, levels
// and final pointers that the passes over every cell would.  Pfinal, levels
// and fin are written for those cells only; rcvp and mark as in
// k_cycle_split.
int labels_split(pff_ctx* ctx, const GridD& g, const int32_t* rec, int32_t* bid, int32_t* Pfinal, int32_t* levels,
                 int32_t* fin, int32_t* rcvp, uint8_t* mark) {
  int64_t ndep_unused = 0;
  int32_t cyc_unused = 0;
  PFF_TRY(tiled_labels(ctx, g, rec, nullptr, &ndep_unused, &cyc_unused, nullptr, true));
  PFF_TRY(tiled_labels_write(ctx, g, bid, nullptr, nullptr));
  const size_t n4 = (size_t)g.N * 4;
  void *L0, *L1, *nbv, *npv;
  PFF_TRY(scratch(ctx, S_WL0, n4, &L0));
  PFF_TRY(scratch(ctx, S_WL1, n4, &L1));
  PFF_TRY(scratch(ctx, S_I2, n4, &nbv));
  PFF_TRY(scratch(ctx, S_I3, n4, &npv));
  Lists L;
  L.l[0] = (int32_t*)L0;
  L.l[1] = (int32_t*)L1;
  L.c[0] = ctx->dev_words + W_COUNT0;
  L.c[1] = ctx->dev_words + W_COUNT1;
  PFF_TRY(check_hip(hipMemsetAsync(L.c[0], 0, 8, ctx->stream), "memset"));
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_cycle_split<M>, dim3((g.N + kChunk - 1) / kChunk), dim3(kBT), 0,
                                               ctx->stream, g, rec, bid, Pfinal, levels, fin, rcvp, mark, L.l[0],
                                               L.c[0]));
  return label_rounds_run(ctx, L, bid, Pfinal, (int32_t*)nbv, (int32_t*)npv, levels, g.N, fin, 0,
                          ceil_log2i(g.N) + 1, nullptr, ctx->hint_marked <= kShortList, ctx->dev_words + W_HMARK);
}

// ---------------------------------------------------------------------------
// Tiled contraction
// ---------------------------------------------------------------------------
constexpr int TS = 64;               // tile side (cells)
constexpr int TC = TS * TS;          // 4096 cells per tile
constexpr int kTB = 512;             // threads per tile block
constexpr int CPT = TC / kTB;        // 8 cells per thread
constexpr uint16_t NONE = 0xFFFF;
constexpr int kTileRounds = 12;      // 2^12 = TC: every acyclic in-tile path is contracted

enum WMode : int { W_ROOTS = 0, W_UNIT = 1, W_SADDLE = 2 };

struct TileArgs {
  GridD g;
  FastDiv fd;
  int tiles_x;
  const int32_t* rcv;
  const uint32_t* tact;  // per-tile "active" flags; null = every tile active
  // local pass outputs
  int32_t* E;        // global local-end of every cell
  uint32_t* S;       // in-tile sums (later: totals at local ends)
  uint32_t* inflow;  // inflow at entry cells (zeroed by the local pass)
  int32_t* P2;       // reduced-graph pointer at local ends
  int32_t* bid2;     // labels at local ends
  int32_t* exits;
  int64_t* n_exits;
  int32_t* pits;     // basin ids (cell+1) of interior roots
  int64_t* n_pits;
  int64_t* cycle;
  // saddle weights
  const int32_t* bid;
  const int32_t* snode;
  // final pass outputs
  float q0;
  float* Q;
  uint32_t* cnt;
  uint8_t* tag;
  // saddle tags: tiles that receive a marked inflow (the final pass skips
  // the others, whose tags the local pass already wrote)
  uint32_t* tin;
  // saddle tags: list of the tagged cells (nullable), each cell once
  // carve tag walks: the reduced graph lives on ring slots (256 per tile,
  // slot = tile * 256 + ring index); ringE[slot] = slot of the ring cell's
  // local end if that is an exit, else the terminal slot; ringIn = tagged
  // inflow of the ring cell; ringS = saddle flag of an exit slot
  int32_t* ringE;
  uint32_t* ringIn;
  uint32_t* ringS;
  int ring_term;  // terminal slot (ntiles * 256): its pointer is -1
  int32_t* tlist;
  int64_t* n_tlist;
  // labels on receivers that may hold cycles: a cell on or draining into a
  // cycle gets E = -1 (in-tile cycle) or the label kCycleLabel (through the
  // reduced graph) instead of failing the pass
  int mark_cycles;
  unsigned long long* prof_cells;  // profiling: cells processed by this launch
  // carve tag walks across the trips of one reroute_flow call: per ring slot
  // the local end of its walk (12 bits) and whether that end is a root
  // (bit 15), valid for the tiles whose tdirty word is 0 (their receivers
  // have not changed since the walk)
  uint16_t* rcache;
  uint32_t* tdirty;
};

template <int W>
__device__ __forceinline__ uint32_t tile_weight(const TileArgs& a, int c) {
  if constexpr (W == W_UNIT) {
    return 1u;
  } else if constexpr (W == W_SADDLE) {
    const int b = a.bid[c];
    return (b > 0 && b < a.g.N && a.snode[b] == c) ? 1u : 0u;
  } else {
    return 0u;
  }
}

__device__ __forceinline__ bool tile_active(const uint32_t* tact, int tiles_x, int bx, int by) {
  return !tact || tact[by * tiles_x + bx] != 0u;
}

// PH 1: local pass; PH 3: final pass (weights += inflow, emit results)
template <int MODE, int PH, int W>
__global__ __launch_bounds__(kTB) void k_tile(TileArgs a, SegLists sl) {
  __shared__ uint16_t P[TC];
  __shared__ uint16_t E[(PH == 1) ? TC : 1];
  __shared__ uint32_t S[(W != W_ROOTS) ? TC : 1];
  const int tile_id = blockIdx.y * a.tiles_x + blockIdx.x;
  if (!tile_active(a.tact, a.tiles_x, blockIdx.x, blockIdx.y)) {  // uniform per block
    if constexpr (PH == 1) {
      if (threadIdx.x == 0) {
        sl.cnt[0][tile_id] = 0;
        if constexpr (W == W_ROOTS) sl.cnt[1][tile_id] = 0;
      }
    }
    return;
  }
  if constexpr (PH == 3 && W == W_SADDLE) {
    if (a.tin && a.tin[blockIdx.y * a.tiles_x + blockIdx.x] == 0u) return;  // uniform per block
  }
  const GridD& g = a.g;
  const int x0 = blockIdx.x * TS, y0 = blockIdx.y * TS;
  const int tw = min(TS, g.nx - x0), th = min(TS, g.ny - y0);
  if (a.prof_cells && threadIdx.x == 0) atomicAdd(a.prof_cells, (unsigned long long)(tw * th));
  int32_t cell[CPT];
  uint32_t rootmask = 0;
  int wany = 0;
  int farlink = 0;  // a receiver that is not a D4 neighbour (labels pass only)
#pragma unroll
  for (int j = 0; j < CPT; ++j) {
    const int li = threadIdx.x + j * kTB;
    const int ly = li >> 6, lx = li & 63;
    uint16_t lp = NONE;
    cell[j] = -1;
    uint32_t w = 0;
    if (lx < tw && ly < th) {
      const int c = (y0 + ly) * g.nx + x0 + lx;
      cell[j] = c;
      int r = a.rcv[c];
      if ((unsigned)r >= (unsigned)g.N) r = c;
      if (r == c) {
        rootmask |= 1u << j;
      } else {
        const int rr = (int)fdiv((uint32_t)r, a.fd);
        const int dy = rr - y0, dxx = r - rr * g.nx - x0;
        if ((unsigned)dy < (unsigned)th && (unsigned)dxx < (unsigned)tw) lp = (uint16_t)(dy * TS + dxx);
        if constexpr (PH == 1 && W == W_ROOTS) {
          // D4 links, wrapping ones included: a receiver in another tile
          // then lies on that tile's ring, which the carve's tag walks use
          const int ady = abs(dy - ly), adx = abs(dxx - lx);
          if (!(ady + adx == 1 || (ady == 0 && adx == g.nx - 1) || (adx == 0 && ady == g.ny - 1))) farlink = 1;
        }
      }
      if constexpr (W != W_ROOTS) {
        w = tile_weight<W>(a, c);
        if constexpr (PH == 3) w += a.inflow[c];
      }
    }
    P[li] = lp;
    if constexpr (PH == 1) E[li] = (lp == NONE) ? (uint16_t)li : lp;
    if constexpr (W != W_ROOTS) {
      S[li] = w;
      wany |= w != 0u;
    }
  }
  // saddle weights are sparse: a tile without any nonzero weight keeps S = 0
  // and skips the sum part of the rounds (uniform per block)
  bool sums = W == W_UNIT;
  if constexpr (W == W_SADDLE) {
    sums = __syncthreads_or(wany) != 0;
  } else if constexpr (W == W_ROOTS && PH == 1) {
    // receivers that are not D4 links take the exact path (reported as a
    // cycle; uniform per block)
    if (__syncthreads_or(farlink) && threadIdx.x == 0) atomicOr((unsigned long long*)a.cycle, 1ull);
  } else {
    __syncthreads();
  }
  // Saddle tags run on receivers already known to be a forest; a tile with
  // no weight there only needs its local ends, jumped on E alone until they
  // stop changing (uniform per block)
  bool eonly = false;
  if constexpr (PH == 1 && W == W_SADDLE) eonly = !sums;
  bool cyc = true;
  for (int k = 0; k < kTileRounds; ++k) {
    uint16_t pm[CPT], pp[CPT], ee[CPT];
    uint32_t sv[CPT];
    int any = 0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int li = threadIdx.x + j * kTB;
      pm[j] = NONE;
      pp[j] = NONE;
      sv[j] = 0;
      if (!eonly) {
        pm[j] = P[li];
        if (pm[j] != NONE) {
          pp[j] = P[pm[j]];
          if constexpr (W != W_ROOTS) {
            if (sums) sv[j] = S[li];
          }
        }
      }
      if constexpr (PH == 1) {
        const uint16_t e0 = E[li];
        ee[j] = E[e0];
        if (eonly) any |= ee[j] != e0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int li = threadIdx.x + j * kTB;
      if (pm[j] != NONE) {
        if constexpr (W != W_ROOTS) {
          if (sums) atomicAdd(&S[pm[j]], sv[j]);
        }
        P[li] = pp[j];
        any |= (pp[j] != NONE);
      }
      if constexpr (PH == 1) E[li] = ee[j];
    }
    if (!__syncthreads_or(any)) {
      cyc = false;
      break;
    }
  }
  if (cyc && threadIdx.x == 0) atomicOr((unsigned long long*)a.cycle, 1ull);
  if constexpr (PH == 1) {
    int32_t ex[CPT], pv[CPT];
    bool take[CPT], tpit[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int li = threadIdx.x + j * kTB;
      take[j] = false;
      tpit[j] = false;
      ex[j] = cell[j];
      pv[j] = cell[j] + 1;
      const int c = cell[j];
      if (c < 0) continue;
      bool end;
      {
        if constexpr (W == W_ROOTS) {
          // the in-tile pointer of a cell whose path ends in the tile has
          // reached NONE after kTileRounds rounds; one that has not is on or
          // drains into a cycle inside the tile
          if (a.mark_cycles && P[li] != NONE) {
            a.E[c] = -1;
            continue;
          }
        }
        const int e = E[li];
        a.E[c] = (y0 + (e >> 6)) * g.nx + x0 + (e & 63);
        end = e == li;
      }
      if constexpr (W != W_ROOTS) {
        a.S[c] = S[li];
        a.inflow[c] = 0u;
      }
      if constexpr (W == W_SADDLE) a.tag[c] = S[li] ? 1 : 0;  // final unless the tile gets a marked inflow
      // (the tagged cells are listed below)
      if (end) {  // local end
        if (rootmask & (1u << j)) {
          if constexpr (W == W_ROOTS) {
            const int row = y0 + (li >> 6), col = x0 + (li & 63);
            const bool out = can_leave<MODE>(g, c, row, col);
            a.P2[c] = c;
            a.bid2[c] = out ? 0 : c + 1;
            tpit[j] = !out;
          } else {
            a.P2[c] = -1;
          }
        } else if constexpr (W == W_SADDLE) {
          // only paths inside closed basins can carry a saddle tag
          const int b = a.bid[c];
          if (b != 0) take[j] = true;
          else a.P2[c] = -1;
        } else {
          take[j] = true;  // exit: its receiver is in another tile
        }
      }
    }
    if constexpr (W == W_ROOTS) seg_append<CPT, kTB>(pv, tpit, sl.seg[1] + (size_t)tile_id * TC, &sl.cnt[1][tile_id]);
    seg_append<CPT, kTB>(ex, take, sl.seg[0] + (size_t)tile_id * TC, &sl.cnt[0][tile_id]);
    if constexpr (W == W_SADDLE) {
      if (sums && a.tlist) {  // uniform per block; a tile without weights has no tag yet
#pragma unroll
        for (int j = 0; j < CPT; ++j) tpit[j] = cell[j] >= 0 && S[threadIdx.x + j * kTB] != 0u;
        block_append<CPT, kTB>(cell, tpit, a.tlist, a.n_tlist);
      }
    }
  } else {
    bool fresh[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int li = threadIdx.x + j * kTB;
      const int c = cell[j];
      fresh[j] = false;
      if (c < 0) continue;
      if constexpr (W == W_UNIT) {
        if (a.cnt) a.cnt[c] = S[li];
        if (a.Q) a.Q[c] = a.q0 * (float)S[li];
      } else if constexpr (W == W_SADDLE) {
        // the local pass listed the cells it tagged: list the new ones only
        const bool t = S[li] != 0u;
        fresh[j] = t && a.tag[c] == 0;
        a.tag[c] = t ? 1 : 0;
      }
    }
    if constexpr (W == W_SADDLE) {
      if (a.tlist) block_append<CPT, kTB>(cell, fresh, a.tlist, a.n_tlist);
    }
  }
}

// reduced graph: exit x points to the local end of the cell it drains into.
// For labels, a target in an inactive tile drains to an open edge (label 0):
// the exit becomes a terminal labelled 0.
__global__ __launch_bounds__(kBT) void k_exit_link(TileArgs a, const int32_t* __restrict__ exits,
                                                   const int64_t* count, int labels, int32_t* __restrict__ list_out,
                                                   int64_t* list_count, int64_t* first_empty, uint32_t* tin,
                                                   int ntiles) {
  const int64_t n = *count;
  // set-up of the rounds that follow: list length, first-empty marker, and
  // the per-tile inflow flags that k_exit_inflow sets afterwards
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (list_count) *list_count = n;
    if (first_empty) *first_empty = (int64_t)1 << 40;
  }
  if (tin)
    for (int t = blockIdx.x * kBT + threadIdx.x; t < ntiles; t += gridDim.x * kBT) tin[t] = 0u;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int x = exits[w];
    const int r = a.rcv[x];
    if (labels && a.tact) {
      const int rr = (int)fdiv((uint32_t)r, a.fd);
      const int rc = r - rr * a.g.nx;
      if (!tile_active(a.tact, a.tiles_x, rc / TS, rr / TS)) {
        a.P2[x] = x;
        a.bid2[x] = 0;
        list_out[w] = x;
        continue;
      }
    }
    const int er = a.E[r];
    if (labels && er < 0) {  // drains into an in-tile cycle (cycle marking only)
      a.P2[x] = x;
      a.bid2[x] = kCycleLabel;
      list_out[w] = x;
      continue;
    }
    a.P2[x] = er;
    // an exit keeps its initial label only if its chain never reaches a root
    if (labels) a.bid2[x] = a.mark_cycles ? kCycleLabel : 0;
    list_out[w] = x;
  }
}

// inflow of entry cells: sum of the totals of the exits draining into them;
// with tin, flags the tiles that receive a nonzero inflow
__global__ __launch_bounds__(kBT) void k_exit_inflow(TileArgs a, const int32_t* __restrict__ exits,
                                                     const int64_t* count, const uint32_t* __restrict__ S,
                                                     uint32_t* inflow) {
  const int64_t n = *count;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int x = exits[w];
    const uint32_t v = S[x];
    if (v == 0u) continue;
    const int r = a.rcv[x];
    atomicAdd(&inflow[r], v);
    if (a.tin) {
      const int rr = (int)fdiv((uint32_t)r, a.fd);
      const int rc = r - rr * a.g.nx;
      a.tin[(rr / TS) * a.tiles_x + rc / TS] = 1u;
    }
  }
}

// final labels on active tiles: bid = label of the local end; the tile stays
// active for the next trip if any of its cells is in a closed basin
// (tact_out zeroed beforehand; one lane per cell, 16 blocks per tile)
__global__ __launch_bounds__(kBT) void k_label_tiles(TileArgs a, int32_t* __restrict__ bid,
                                                     uint32_t* __restrict__ tact_out) {
  int bx, by;
  xcd_block(bx, by);
  const int t = (by >> 4) * a.tiles_x + bx;
  if (a.tact && a.tact[t] == 0u) return;  // uniform per block
  const int col = bx * 64 + threadIdx.x;
  const int row = by * 4 + threadIdx.y;
  int any = 0;
  if (col < a.g.nx && row < a.g.ny) {
    const int c = row * a.g.nx + col;
    const int e = a.E[c];
    const int b = e >= 0 ? a.bid2[e] : kCycleLabel;
    bid[c] = b;
    any = b != 0;
  }
  any = __syncthreads_or(any);
  if (tact_out && any && threadIdx.x == 0 && threadIdx.y == 0) atomicOr(&tact_out[t], 1u);
}

// ---------------------------------------------------------------------------
// Per-tile list segments: a tile block writes its entries to its own 4096
// slots (seg + tile * TC) and its count; k_seg_scan turns the counts into
// offsets (after the list's current length when appending) and the new
// length, k_seg_copy moves the segments into the contiguous list.  One
// block of k_seg_scan per list (blockIdx.x), one block of k_seg_copy per
// tile and list.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_seg_scan(SegLists s, int ntiles) {
  const int l = blockIdx.x;
  const int per = (ntiles + 1023) / 1024;
  const int t0 = min((int)threadIdx.x * per, ntiles), t1 = min(t0 + per, ntiles);
  const int64_t base = s.append[l] ? *s.count[l] : 0;
  int sum = 0;
  for (int t = t0; t < t1; ++t) sum += s.cnt[l][t];
  int total;
  int acc = (int)base + block_exclusive_scan<1024>(sum, &total);
  for (int t = t0; t < t1; ++t) {
    s.off[l][t] = acc;
    acc += s.cnt[l][t];
  }
  if (threadIdx.x == 0) *s.count[l] = base + total;  // every thread read base before the scan's barriers
}

__global__ __launch_bounds__(kBT) void k_seg_copy(SegLists s) {
  const int t = blockIdx.x, l = blockIdx.y;
  const int n = s.cnt[l][t], o = s.off[l][t];
  const int32_t* src = s.seg[l] + (size_t)t * TC;
  for (int k = threadIdx.x; k < n; k += kBT) s.list[l][o + k] = src[k];
}

// ---------------------------------------------------------------------------
// Carve tags by walks.  A tile's carve tags only come from its saddle nodes
// (the paths from them down to the tile's local ends) and from the tagged
// flow that enters it through ring cells (a D4 receiver in another tile lies
// on that tile's ring).  So per tile only those paths are followed, one lane
// per start cell over LDS local pointers, instead of contracting every cell:
// the local pass walks from the saddle nodes (tagging) and from every ring
// cell (finding its local end, which the reduced graph needs), the final
// pass walks from the ring cells that received a tag (tagging).  A walk
// stops at a cell already tagged: that cell's walker carries on downstream.
// ---------------------------------------------------------------------------
constexpr int TW = TC / 32;  // bitmap words per tile
constexpr int kWB = 256;     // threads per tile block of the walk passes (more tiles in flight per CU)
constexpr int CPW = TC / kWB;

__device__ __forceinline__ bool bit_test_set(uint32_t* bm, int li) {
  const uint32_t b = 1u << (li & 31);
  return (atomicOr(&bm[li >> 5], b) & b) != 0u;
}
__device__ __forceinline__ bool bit_test(const uint32_t* bm, int li) { return (bm[li >> 5] >> (li & 31)) & 1u; }

// slot of this lane in an LDS staging list: one atomic per wave for the
// lanes active at the call
__device__ __forceinline__ int wave_slot(int* count) {
  const uint64_t m = __ballot(1);
  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(count, __popcll(m));
  return __shfl(base, leader) + (int)below;
}

// store a 64-cell row flag (one wave, lanes in cell order) into a bitmap
__device__ __forceinline__ void row_bits(uint32_t* bm, int li_lane0, bool flag) {
  const uint64_t m = __ballot(flag);
  const int lane = threadIdx.x & 63;
  if (lane < 2) bm[(li_lane0 >> 5) + lane] = (uint32_t)(m >> (32 * lane));
}

// k-th ring cell (local index) of a tw x th tile; ring_count cells in all
__device__ __forceinline__ int ring_count(int tw, int th) {
  if (th == 1) return tw;
  if (tw == 1) return th;
  return 2 * tw + 2 * (th - 2);
}
__device__ __forceinline__ int ring_cell(int k, int tw, int th) {
  if (th == 1) return k;
  if (tw == 1) return k * TS;
  if (k < tw) return k;
  k -= tw;
  if (k < tw) return (th - 1) * TS + k;
  k -= tw;
  if (k < th - 2) return (k + 1) * TS;
  k -= th - 2;
  return (k + 1) * TS + tw - 1;
}

// A tile's receivers, loaded before any of them is used: on gfx9 a wait for
// a load also waits for every earlier store, so each pass of the walk
// kernels is ordered loads, then decisions, then stores.
__device__ __forceinline__ void tile_load_rcv(const TileArgs& a, int x0, int y0, int tw, int th, int (&rv)[CPW]) {
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    const int li = threadIdx.x + j * kWB;
    const int ly = li >> 6, lx = li & 63;
    rv[j] = -2;
    if (lx < tw && ly < th) rv[j] = a.rcv[(y0 + ly) * a.g.nx + x0 + lx];
  }
}

// inverse of ring_cell for a ring cell (ly, lx)
__device__ __forceinline__ int ring_index(int ly, int lx, int tw, int th) {
  if (th == 1) return lx;
  if (tw == 1) return ly;
  if (ly == 0) return lx;
  if (ly == th - 1) return tw + lx;
  if (lx == 0) return 2 * tw + ly - 1;
  return 2 * tw + (th - 2) + ly - 1;
}
constexpr int kRing = 256;  // ring slots per tile (2*64 + 2*62 = 252 used)

__device__ __forceinline__ void tile_dims(const GridD& g, int tx, int ty, int& tw, int& th) {
  tw = min(TS, g.nx - tx * TS);
  th = min(TS, g.ny - ty * TS);
}

// ring slot of a cell that lies on its tile's ring
__device__ __forceinline__ int slot_of(const TileArgs& a, int c) {
  const int row = (int)fdiv((uint32_t)c, a.fd), col = c - row * a.g.nx;
  const int tx = col / TS, ty = row / TS;
  int tw, th;
  tile_dims(a.g, tx, ty, tw, th);
  return (ty * a.tiles_x + tx) * kRing + ring_index(row - ty * TS, col - tx * TS, tw, th);
}

// cell of a ring slot
__device__ __forceinline__ int slot_cell(const TileArgs& a, int x) {
  const int tile = x / kRing, k = x - tile * kRing;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  int tw, th;
  tile_dims(a.g, tx, ty, tw, th);
  const int li = ring_cell(k, tw, th);
  return (ty * TS + (li >> 6)) * a.g.nx + tx * TS + (li & 63);
}

// local pointers of the tile into P (NONE: root or leaves the tile), roots
// into the bitmap R (each wave writes the words of its row)
__device__ __forceinline__ void tile_build_pointers(const TileArgs& a, int x0, int y0, int tw, int th,
                                                    const int (&rv)[CPW], uint16_t* P, uint32_t* R) {
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    const int li = threadIdx.x + j * kWB;
    const int ly = li >> 6, lx = li & 63;
    uint16_t lp = NONE;
    bool root = false;
    if (rv[j] != -2) {
      const int c = (y0 + ly) * a.g.nx + x0 + lx;
      int r = rv[j];
      if ((unsigned)r >= (unsigned)a.g.N) r = c;
      if (r == c) {
        root = true;
      } else {
        const int rr = (int)fdiv((uint32_t)r, a.fd);
        const int dy = rr - y0, dxx = r - rr * a.g.nx - x0;
        if ((unsigned)dy < (unsigned)th && (unsigned)dxx < (unsigned)tw) lp = (uint16_t)(dy * TS + dxx);
      }
    }
    P[li] = lp;
    row_bits(R, li - lx, root);
  }
}

// tag the path from li down to the first tagged cell or the local end.
// Returns the local end if this walk tagged it, else -1.
__device__ __forceinline__ int walk_tag(int li, const uint16_t* P, uint32_t* T, int64_t* cycle) {
  int c = li;
  for (int s = 0; s < TC; ++s) {
    // the pointer read is issued before the tag atomic returns: both LDS
    // round trips of a step overlap
    const uint16_t n = P[c];
    if (bit_test_set(T, c)) return -1;
    if (n == NONE) return c;
    c = n;
  }
  atomicOr((unsigned long long*)cycle, 1ull);  // not a forest
  return -1;
}

__device__ __forceinline__ int walk_end(int li, const uint16_t* P, int64_t* cycle) {
  int c = li;
  for (int s = 0; s < TC; ++s) {
    const uint16_t n = P[c];
    if (n == NONE) return c;
    c = n;
  }
  atomicOr((unsigned long long*)cycle, 1ull);
  return li;
}

// the same walk cut after `cap` steps: -1 if it has not reached the end
__device__ __forceinline__ int walk_end_capped(int li, const uint16_t* P, int cap) {
  int c = li;
  for (int s = 0; s < cap; ++s) {
    const uint16_t n = P[c];
    if (n == NONE) return c;
    c = n;
  }
  return -1;
}

// Local ends of a whole tile by pointer jumping in place in LDS: P[c] <-
// P[P[c]] while that is not the end marker.  Concurrent updates only move a
// pointer further along its own chain, so any interleaving converges; after
// it, the end of c is c if P[c] is NONE, else P[c].  13 rounds cover chains
// of 4096 cells; a pointer still moving then lies on a cycle.
template <int NT, int PER>
__device__ __forceinline__ void tile_jump_ends(uint16_t* P, int64_t* cycle) {
  for (int r = 0; r < 14; ++r) {
    bool moved = false;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int li = threadIdx.x + j * NT;
      const uint16_t n = P[li];
      if (n != NONE) {
        const uint16_t nn = P[n];
        if (nn != NONE) {
          P[li] = nn;
          moved = true;
        }
      }
    }
    if (!__syncthreads_or(moved)) return;
  }
  if (threadIdx.x == 0) atomicOr((unsigned long long*)cycle, 1ull);
}

constexpr int kWalkCap = 384;  // ring walks longer than this switch the tile to pointer jumping

// local pass: tags below the tile's saddle nodes (wbits, cleared here), the
// local end E of every ring cell, the reduced-graph nodes (the ends those
// walks reach: listed exits with S = 1 if a saddle path reaches them, or
// terminals P2 = -1), inflow = 0 on the ring
__global__ __launch_bounds__(kWB) void k_tag_local(TileArgs a, uint32_t* __restrict__ wbits, SegLists sl) {
  __shared__ uint16_t P[TC];
  __shared__ uint16_t wl[TC];  // saddle nodes of the tile
  __shared__ uint32_t T[TW], Lb[TW], F[TW], R[TW], Wb[TW];
  __shared__ int nw;
  const int tile = blockIdx.y * a.tiles_x + blockIdx.x;
  if (!tile_active(a.tact, a.tiles_x, blockIdx.x, blockIdx.y)) {  // uniform per block
    if (threadIdx.x == 0) {
      sl.cnt[0][tile] = 0;
      sl.cnt[1][tile] = 0;
    }
    return;
  }
  const GridD& g = a.g;
  const int x0 = blockIdx.x * TS, y0 = blockIdx.y * TS;
  const int tw = min(TS, g.nx - x0), th = min(TS, g.ny - y0);
  const uint32_t w = threadIdx.x < TW ? wbits[(size_t)tile * TW + threadIdx.x] : 0u;
  const int base = tile * kRing;
  // A tile whose receivers have not changed since its ring walks were cached
  // and that holds no saddle node this trip adds no tag: only its ring-slot
  // graph is needed, and the cached ends give it without loading the tile
  // (the exit test reads the end's current label: labels only merge).
  const bool clean = a.rcache && a.tdirty[tile] == 0u;  // uniform per block
  if (clean && !__syncthreads_or(w != 0u)) {
    const int nr = ring_count(tw, th);
    int32_t xs[1] = {0};
    bool xt[1] = {false};
    if ((int)threadIdx.x < nr) {
      const int k = threadIdx.x;
      const uint32_t cv = a.rcache[base + k];
      const int e = (int)(cv & 0x0FFFu);
      const bool eroot = (cv >> 15) != 0u;
      const bool ex = !eroot && a.bid[(y0 + (e >> 6)) * g.nx + x0 + (e & 63)] != 0;
      a.ringE[base + k] = ex ? base + ring_index(e >> 6, e & 63, tw, th) : a.ring_term;
      a.ringIn[base + k] = 0u;
      a.ringS[base + k] = 0u;
      xt[0] = ex && e == ring_cell(k, tw, th);  // the ring cell is its own end: an exit
      xs[0] = base + k;
    }
    if (threadIdx.x == 0) sl.cnt[0][tile] = 0;
    seg_append<1, kWB>(xs, xt, sl.seg[1] + (size_t)tile * TC, &sl.cnt[1][tile]);
    if (a.prof_cells && threadIdx.x == 0)
      atomicAdd(a.prof_cells, (unsigned long long)(16 * nr + 4 * sl.cnt[1][tile]));
    return;
  }
  int rv[CPW];
  tile_load_rcv(a, x0, y0, tw, th, rv);
  if (threadIdx.x < TW) {
    T[threadIdx.x] = 0u;
    Lb[threadIdx.x] = 0u;
    F[threadIdx.x] = 0u;
    Wb[threadIdx.x] = w;
  }
  if (threadIdx.x == 0) nw = 0;
  tile_build_pointers(a, x0, y0, tw, th, rv, P, R);
  __syncthreads();
  if (w) wbits[(size_t)tile * TW + threadIdx.x] = 0u;
  // saddle nodes into a list, so that their walks spread over all lanes
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    const int li = threadIdx.x + j * kWB;
    if (bit_test(Wb, li)) wl[wave_slot(&nw)] = (uint16_t)li;
  }
  __syncthreads();
  const int nwl = nw;
  for (int k = threadIdx.x; k < nwl; k += kWB) {
    const int e = walk_tag(wl[k], P, T, a.cycle);
    if (e >= 0) {
      atomicOr(&F[e >> 5], 1u << (e & 31));
      atomicOr(&Lb[e >> 5], 1u << (e & 31));
    }
  }
  // ring cells: their local ends.  Walks are cut at kWalkCap steps; if one
  // is longer (long paths through the tile, common once basins have been
  // carved together) the tile finds every local end by pointer jumping,
  // which takes log2 of the path length in LDS rounds.  The tags are done,
  // so the pointers may be overwritten.
  __shared__ uint16_t rend[kRing];
  __shared__ uint32_t Xb[TW];
  const int nr = ring_count(tw, th);
  bool unfinished = false;
  for (int k = threadIdx.x; k < nr; k += kWB) {
    const int e = walk_end_capped(ring_cell(k, tw, th), P, kWalkCap);
    rend[k] = e < 0 ? NONE : (uint16_t)e;
    if (e >= 0) atomicOr(&Lb[e >> 5], 1u << (e & 31));
    else unfinished = true;
  }
  if (__syncthreads_or(unfinished)) {
    tile_jump_ends<kWB, CPW>(P, a.cycle);
    for (int k = threadIdx.x; k < nr; k += kWB) {
      if (rend[k] != NONE) continue;
      const int c = ring_cell(k, tw, th);
      const int e = P[c] == NONE ? c : P[c];
      rend[k] = (uint16_t)e;
      atomicOr(&Lb[e >> 5], 1u << (e & 31));
    }
  }
  __syncthreads();
  // reached ends: roots and ends outside closed basins are terminals, the
  // others exits (on the ring: the receiver is in another tile)
  auto cell_of = [&](int j) {
    const int li = threadIdx.x + j * kWB;
    return (y0 + (li >> 6)) * g.nx + x0 + (li & 63);
  };
  int bv[CPW];
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    const int li = threadIdx.x + j * kWB;
    bv[j] = 1;
    if (bit_test(Lb, li) && !bit_test(R, li)) bv[j] = a.bid[cell_of(j)];
  }
  uint32_t tmask = 0u;
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    const int li = threadIdx.x + j * kWB;
    if (bit_test(T, li)) tmask |= 1u << j;
    row_bits(Xb, li - (li & 63), bit_test(Lb, li) && !bit_test(R, li) && bv[j] != 0);
  }
  __syncthreads();
  // one ring slot per thread (nr <= 252 < kWB): its end's slot, its inflow,
  // and, if the ring cell is itself an exit, its saddle flag
  int32_t xs[1] = {0};
  bool xt[1] = {false};
  if ((int)threadIdx.x < nr) {
    const int k = threadIdx.x;
    const int e = rend[k];
    a.ringE[base + k] = bit_test(Xb, e) ? base + ring_index(e >> 6, e & 63, tw, th) : a.ring_term;
    a.ringIn[base + k] = 0u;
    if (a.rcache) a.rcache[base + k] = (uint16_t)(e | (bit_test(R, e) ? 0x8000 : 0));
    const int li = ring_cell(k, tw, th);
    xt[0] = bit_test(Xb, li);
    xs[0] = base + k;
    a.ringS[base + k] = (xt[0] && bit_test(F, li)) ? 1u : 0u;
  }
#pragma unroll
  for (int j = 0; j < CPW; ++j)
    if ((tmask >> j) & 1u) a.tag[cell_of(j)] = 1;
  seg_append_mask<CPW, kWB>(tmask, cell_of, sl.seg[0] + (size_t)tile * TC, &sl.cnt[0][tile]);
  seg_append<1, kWB>(xs, xt, sl.seg[1] + (size_t)tile * TC, &sl.cnt[1][tile]);
  if (a.tdirty && threadIdx.x == 0) a.tdirty[tile] = 0u;  // the cache now holds this tile's walks
  // profiling: algorithmic bytes of the tile (rcv in; ring slots, tags and
  // list entries out)
  if (a.prof_cells && threadIdx.x == 0)
    atomicAdd(a.prof_cells, (unsigned long long)(4 * tw * th + 12 * nr + 5 * sl.cnt[0][tile] + 4 * sl.cnt[1][tile]));
}

// final pass on the tiles that receive a tag: walk from the ring cells with
// a tagged inflow, tagging until a tagged cell; list the new tags
__global__ __launch_bounds__(kWB, 8) void k_tag_final(TileArgs a, SegLists sl) {
  __shared__ uint16_t P[TC];
  __shared__ uint32_t T[TW], T0[TW], R[TW];
  const int tile = blockIdx.y * a.tiles_x + blockIdx.x;
  if (!tile_active(a.tact, a.tiles_x, blockIdx.x, blockIdx.y) || a.tin[tile] == 0u) {  // uniform per block
    if (threadIdx.x == 0) sl.cnt[0][tile] = 0;
    return;
  }
  const GridD& g = a.g;
  const int x0 = blockIdx.x * TS, y0 = blockIdx.y * TS;
  const int tw = min(TS, g.nx - x0), th = min(TS, g.ny - y0);
  if (a.prof_cells && threadIdx.x == 0) atomicAdd(a.prof_cells, (unsigned long long)(tw * th));
  int rv[CPW];
  tile_load_rcv(a, x0, y0, tw, th, rv);
  uint8_t tb[CPW];
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    const int li = threadIdx.x + j * kWB;
    const int ly = li >> 6, lx = li & 63;
    tb[j] = (lx < tw && ly < th) ? a.tag[(y0 + ly) * g.nx + x0 + lx] : (uint8_t)0;
  }
  tile_build_pointers(a, x0, y0, tw, th, rv, P, R);
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    const int li = threadIdx.x + j * kWB;
    row_bits(T, li - (li & 63), tb[j] != 0);
    row_bits(T0, li - (li & 63), tb[j] != 0);
  }
  __syncthreads();
  const int nr = ring_count(tw, th);
  for (int k = threadIdx.x; k < nr; k += kWB)
    if (a.ringIn[tile * kRing + k] != 0u) (void)walk_tag(ring_cell(k, tw, th), P, T, a.cycle);
  __syncthreads();
  auto cell_of = [&](int j) {
    const int li = threadIdx.x + j * kWB;
    return (y0 + (li >> 6)) * g.nx + x0 + (li & 63);
  };
  uint32_t tmask = 0u;
#pragma unroll
  for (int j = 0; j < CPW; ++j) {
    const int li = threadIdx.x + j * kWB;
    if (bit_test(T, li) && !bit_test(T0, li)) {
      tmask |= 1u << j;
      a.tag[cell_of(j)] = 1;
    }
  }
  seg_append_mask<CPW, kWB>(tmask, cell_of, sl.seg[0] + (size_t)tile * TC, &sl.cnt[0][tile]);
}

// reduced graph of the tag walks: exit slot x -> the slot of the local end
// its receiver's ring cell reaches (or the terminal slot), list set-up
__global__ __launch_bounds__(kBT) void k_ring_link(TileArgs a, const int32_t* __restrict__ exits, const int64_t* count,
                                                   int32_t* __restrict__ P2, int32_t* __restrict__ list_out,
                                                   int64_t* list_count, int64_t* first_empty, uint32_t* tin,
                                                   int ntiles) {
  const int64_t n = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *list_count = n;
    *first_empty = (int64_t)1 << 40;
    P2[a.ring_term] = -1;
  }
  for (int t = blockIdx.x * kBT + threadIdx.x; t < ntiles; t += gridDim.x * kBT) tin[t] = 0u;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int x = exits[w];
    P2[x] = a.ringE[slot_of(a, a.rcv[slot_cell(a, x)])];
    list_out[w] = x;
  }
}

// flagged exits pass their flag to the ring cell they drain into
__global__ __launch_bounds__(kBT) void k_ring_inflow(TileArgs a, const int32_t* __restrict__ exits, const int64_t* count,
                                                     uint32_t* tin) {
  const int64_t n = *count;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int x = exits[w];
    const uint32_t v = a.ringS[x];
    if (v == 0u) continue;
    const int slot = slot_of(a, a.rcv[slot_cell(a, x)]);
    atomicAdd(&a.ringIn[slot], v);
    tin[slot / kRing] = 1u;
  }
}

// converged = 1 when the reduced rounds emptied their list and no tile saw
// an in-tile cycle (the receivers are a forest), else 0
__global__ void k_forest_status(const int64_t* live, const int64_t* cycle, int32_t* converged) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *converged = (*live == 0 && *cycle == 0) ? 1 : 0;
}

struct TileBufs {
  int32_t *E, *P2, *bid2, *exits, *L0, *L1, *t0, *t1, *t2, *pits;
  uint32_t *S, *inflow, *tact, *tin;
};

static int tile_bufs(pff_ctx* ctx, const GridD& g, TileBufs& b) {
  const size_t n4 = (size_t)g.N * 4;
  void* p;
  PFF_TRY(scratch(ctx, S_TE, n4, &p)); b.E = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_TP2, n4, &p)); b.P2 = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_TBID, n4, &p)); b.bid2 = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_TEXIT, n4, &p)); b.exits = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_WL0, n4, &p)); b.L0 = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_WL1, n4, &p)); b.L1 = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_I1, n4, &p)); b.t0 = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_I2, n4, &p)); b.t1 = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_I3, n4, &p)); b.t2 = (int32_t*)p;
  PFF_TRY(scratch(ctx, S_TS, n4, &p)); b.S = (uint32_t*)p;
  PFF_TRY(scratch(ctx, S_TIN, n4, &p)); b.inflow = (uint32_t*)p;
  PFF_TRY(scratch(ctx, S_TPIT, n4, &p)); b.pits = (int32_t*)p;
  const size_t nt = (size_t)((g.nx + TS - 1) / TS) * ((g.ny + TS - 1) / TS);
  PFF_TRY(scratch(ctx, S_TACT, nt * 4 * 4, &p)); b.tact = (uint32_t*)p;
  b.tin = b.tact + 2 * nt;
  return 0;
}

size_t tiled_upstream_slots(const GridD& g, int* slots, size_t* bytes, int* nslots) {
  const size_t n4 = (size_t)g.N * 4;
  const size_t nt = (size_t)tiles_count(g);
  const int s_n[] = {S_TE, S_TP2, S_TBID, S_TEXIT, S_WL0, S_WL1, S_I1, S_I2, S_I3, S_TS, S_TIN, S_TPIT};
  int k = 0;
  for (int sl : s_n) {
    slots[k] = sl;
    bytes[k++] = n4;
  }
  slots[k] = S_TACT;
  bytes[k++] = nt * 4 * 4;
  slots[k] = S_SEGD0;
  bytes[k++] = nt * TC * 4;
  slots[k] = S_SEGD1;
  bytes[k++] = nt * TC * 4;
  slots[k] = S_SEGCO;
  bytes[k++] = nt * 4 * 4;
  *nslots = k;
  size_t total = 0;
  for (int j = 0; j < k; ++j) total += (bytes[j] + 255) & ~(size_t)255;
  return total;
}

static TileArgs tile_args(const GridD& g, const TileBufs& b, pff_ctx* ctx, const int32_t* rcv) {
  TileArgs a{};
  a.g = g;
  a.fd = make_fastdiv((uint32_t)g.nx);
  a.tiles_x = (g.nx + TS - 1) / TS;
  a.rcv = rcv;
  a.E = b.E;
  a.S = b.S;
  a.inflow = b.inflow;
  a.P2 = b.P2;
  a.bid2 = b.bid2;
  a.exits = b.exits;
  a.n_exits = ctx->dev_words + W_NEXIT;
  a.pits = b.pits;
  a.n_pits = ctx->dev_words + W_NPIT;
  a.cycle = ctx->dev_words + W_CYCLE;
  return a;
}

static dim3 tile_grid(const GridD& g) { return dim3((g.nx + TS - 1) / TS, (g.ny + TS - 1) / TS); }

int tiles_count(const GridD& g) { return ((g.nx + TS - 1) / TS) * ((g.ny + TS - 1) / TS); }

// per-tile segments for up to two lists of a tile pass (S_SEGD0/1, S_SEGCO)
int seg_lists(pff_ctx* ctx, const GridD& g, SegLists& sl, int32_t* list0, int64_t* count0, int32_t* list1,
                     int64_t* count1) {
  const int nt = tiles_count(g);
  void *d0, *d1, *co;
  PFF_TRY(scratch(ctx, S_SEGD0, (size_t)nt * TC * 4, &d0));
  PFF_TRY(scratch(ctx, S_SEGD1, (size_t)nt * TC * 4, &d1));
  PFF_TRY(scratch(ctx, S_SEGCO, (size_t)nt * 4 * 4, &co));
  sl = SegLists{};
  sl.seg[0] = (int32_t*)d0;
  sl.seg[1] = (int32_t*)d1;
  sl.cnt[0] = (int32_t*)co;
  sl.cnt[1] = sl.cnt[0] + nt;
  sl.off[0] = sl.cnt[1] + nt;
  sl.off[1] = sl.off[0] + nt;
  sl.list[0] = list0;
  sl.count[0] = count0;
  sl.list[1] = list1;
  sl.count[1] = count1;
  return 0;
}

// join the segments of nlists lists into their contiguous lists
void seg_join(pff_ctx* ctx, const GridD& g, const SegLists& sl, int nlists) {
  const int nt = tiles_count(g);
  hipLaunchKernelGGL(k_seg_scan, dim3(nlists), dim3(1024), 0, ctx->stream, sl, nt);
  hipLaunchKernelGGL(k_seg_copy, dim3(nt, nlists), dim3(kBT), 0, ctx->stream, sl);
}

// Upstream sums with the tiled contraction.  W_UNIT: counts -> cnt and/or
// Q = q0*count; W_SADDLE: tag[c] = 1 where a saddle node lies upstream
// (tiles flagged inactive in tact are skipped: they hold no closed basin).
// rounds > 0 with cycle null: the graph is known to be a forest whose reduced
// depth needs at most `rounds` pointer-jumping rounds (e.g. the label rounds
// of the same graph); nothing syncs.  Otherwise the rounds run until the list
// empties, checked on the host every few rounds, and *cycle (host) reports a
// graph that is not a forest (outputs are then not meaningful).
int tiled_upstream(pff_ctx* ctx, const GridD& g, const int32_t* rcv, int wmode, const int32_t* bid,
                   const int32_t* snode, float q0, float* Q, uint32_t* cnt, uint8_t* tag, const uint32_t* tact,
                   int32_t* cycle, int rounds, int32_t* tag_list, int32_t* converged_dev) {
  TileBufs b;
  PFF_TRY(tile_bufs(ctx, g, b));
  TileArgs a = tile_args(g, b, ctx, rcv);
  if (tag_list) {
    a.tlist = tag_list;
    a.n_tlist = ctx->dev_words + W_NTAG;
  }
  a.bid = bid;
  a.snode = snode;
  a.q0 = q0;
  a.Q = Q;
  a.cnt = cnt;
  a.tag = tag;
  a.tact = tact;
  hipStream_t s = ctx->stream;
  static_assert(W_NEXIT == W_CYCLE + 1, "cycle flag and exit count are adjacent words");
  PFF_TRY(launch_fill_i64(ctx, 2, ctx->dev_words + W_CYCLE, 0));  // a kernel, not a runtime fill: capturable
  const dim3 tg = tile_grid(g);
  // algorithmic bytes per cell: rcv in; E, S, inflow out (+ bid, snode for saddle weights)
  a.prof_cells = prof_begin(ctx, "k_tile", wmode == W_UNIT ? 16.0 : 24.0);
  SegLists sl;
  PFF_TRY(seg_lists(ctx, g, sl, b.exits, a.n_exits, nullptr, nullptr));
  if (wmode == W_UNIT) PFF_PROF_LAUNCH(ctx, (k_tile<0, 1, W_UNIT>), tg, dim3(kTB), 0, s, a, sl);
  else PFF_PROF_LAUNCH(ctx, (k_tile<0, 1, W_SADDLE>), tg, dim3(kTB), 0, s, a, sl);
  a.prof_cells = nullptr;
  seg_join(ctx, g, sl, 1);
  // reduced graph over the exits: P2 (roots = -1), S holds the in-tile sums
  Lists L;
  L.l[0] = b.L0;
  L.l[1] = b.L1;
  L.c[0] = ctx->dev_words + W_COUNT0;
  L.c[1] = ctx->dev_words + W_COUNT1;
  // saddle tags only ask whether anything upstream is marked: OR rounds
  const bool flags = wmode == W_SADDLE;
  int64_t* fe = ctx->dev_words + W_EMPTY2;
  hipLaunchKernelGGL(k_exit_link, dim3(list_grid()), dim3(kBT), 0, s, a, b.exits, a.n_exits, 0, b.L0, L.c[0], fe,
                     flags ? b.tin : nullptr, tiles_count(g));
  if (rounds > 0 && !cycle) {
    // the caller knows the depth bound (the labels of the same graph), or
    // runs the full ceil(log2 N)+1 rounds (rounds on an empty list return at
    // once) and learns on the device whether the list emptied
    PFF_TRY(count_rounds(ctx, L, b.S, b.P2, b.t0, (uint32_t*)b.t1, b.t2, rounds, flags));
    if (converged_dev)
      hipLaunchKernelGGL(k_forest_status, dim3(1), dim3(64), 0, s, L.c[L.cur], ctx->dev_words + W_CYCLE,
                         converged_dev);
  } else {
    // start from the round count the previous call of this kind needed,
    // then confirm on the host that the list has emptied
    const int cap = ceil_log2i(g.N) + 1;
    int& last = flags ? ctx->last_tag_rounds : ctx->last_count_rounds;
    int done_rounds = 0;
    int64_t live = 1;
    int step = rounds > 0 ? rounds : std::max(2, std::min(last + 1, cap));
    while (done_rounds < cap) {
      const int r = std::min(step, cap - done_rounds);
      PFF_TRY(count_rounds(ctx, L, b.S, b.P2, b.t0, (uint32_t*)b.t1, b.t2, r, flags, done_rounds, fe));
      done_rounds += r;
      step = 3;
      PFF_TRY(read_words(ctx));
      live = ctx->host_words[L.cur == 0 ? W_COUNT0 : W_COUNT1];
      if (live == 0 || ctx->host_words[W_CYCLE] != 0) break;
    }
    const int64_t f = ctx->host_words[W_EMPTY2];
    last = (live == 0 && f < done_rounds) ? std::max<int>((int)f, 1) : done_rounds;
    const int cyc = (live != 0 || ctx->host_words[W_CYCLE] != 0) ? 1 : 0;
    if (cycle) *cycle = cyc;
    if (cyc) return cycle ? 0 : fail(PFF_ENOTCONV, "tiled_upstream: receivers are not a forest");
  }
  if (flags) a.tin = b.tin;  // zeroed by k_exit_link
  hipLaunchKernelGGL(k_exit_inflow, dim3(list_grid()), dim3(kBT), 0, s, a, b.exits, a.n_exits, b.S, b.inflow);
  // rcv, inflow in; Q (or count) out / rcv, bid, snode, inflow in; tag out
  a.prof_cells = prof_begin(ctx, "k_tile", wmode == W_UNIT ? 12.0 : 17.0);
  if (wmode == W_UNIT) PFF_PROF_LAUNCH(ctx, (k_tile<0, 3, W_UNIT>), tg, dim3(kTB), 0, s, a, sl);
  else PFF_PROF_LAUNCH(ctx, (k_tile<0, 3, W_SADDLE>), tg, dim3(kTB), 0, s, a, sl);
  a.prof_cells = nullptr;
  return check_hip(hipGetLastError(), "tiled_upstream");
}

int tag_state(pff_ctx* ctx, const GridD& g, uint32_t** wbits, uint8_t** tag, bool begin) {
  void *w, *t;
  const size_t wb = (size_t)tiles_count(g) * TW * 4;
  PFF_TRY(scratch(ctx, S_WBITS, wb, &w, true));
  PFF_TRY(scratch(ctx, S_B1, (size_t)g.N, &t, true));
  if (begin) {
    if (ctx->tag_state_dirty) {
      PFF_TRY(check_hip(hipMemsetAsync(w, 0, wb, ctx->stream), "memset"));
      PFF_TRY(check_hip(hipMemsetAsync(t, 0, (size_t)g.N, ctx->stream), "memset"));
    }
    ctx->tag_state_dirty = true;  // until the carve has cleared what it sets
  }
  *wbits = (uint32_t*)w;
  *tag = (uint8_t*)t;
  return 0;
}

int ring_cache(pff_ctx* ctx, const GridD& g, uint16_t** rcache, uint32_t** tdirty) {
  const size_t nt = (size_t)tiles_count(g);
  const size_t cb = (nt * kRing * 2 + 255) & ~(size_t)255;
  void* p;
  PFF_TRY(scratch(ctx, S_RCACHE, cb + nt * 4, &p));
  *rcache = (uint16_t*)p;
  *tdirty = (uint32_t*)((char*)p + cb);
  return 0;
}

int ring_cache_reset(pff_ctx* ctx, const GridD& g) {
  uint16_t* rc;
  uint32_t* td;
  PFF_TRY(ring_cache(ctx, g, &rc, &td));
  return launch_fill_u32(ctx, tiles_count(g), td, 1u);
}

int saddle_tags_walk(pff_ctx* ctx, const GridD& g, const int32_t* rcv, const int32_t* bid, uint32_t* wbits,
                     uint8_t* tag, const uint32_t* tact, int rounds, int32_t* tag_list) {
  TileBufs b;
  PFF_TRY(tile_bufs(ctx, g, b));
  TileArgs a = tile_args(g, b, ctx, rcv);
  a.tlist = tag_list;
  a.n_tlist = ctx->dev_words + W_NTAG;
  a.bid = bid;
  a.tag = tag;
  a.tact = tact;
  hipStream_t s = ctx->stream;
  // the cycle flag and the exit count were zeroed by this trip's saddle
  // initialisation (saddlesort_active, the only caller's predecessor)
  const dim3 tg = tile_grid(g);
  const int nt = tiles_count(g);
  const int nslots = nt * kRing;
  int32_t* ringP2;
  {
    void* rg;
    PFF_TRY(scratch(ctx, S_RING, (size_t)(nslots + 1) * 4 * 4, &rg));
    a.ringE = (int32_t*)rg;
    a.ringIn = (uint32_t*)(a.ringE + (nslots + 1));
    a.ringS = a.ringIn + (nslots + 1);
    ringP2 = (int32_t*)(a.ringS + (nslots + 1));
    a.ring_term = nslots;
  }
  PFF_TRY(ring_cache(ctx, g, &a.rcache, &a.tdirty));
  SegLists sl{};
  {
    void *d0, *d1, *co;
    PFF_TRY(scratch(ctx, S_SEGD0, (size_t)nt * TC * 4, &d0));
    PFF_TRY(scratch(ctx, S_SEGD1, (size_t)nt * TC * 4, &d1));
    PFF_TRY(scratch(ctx, S_SEGCO, (size_t)nt * 4 * 4, &co));
    sl.seg[0] = (int32_t*)d0;
    sl.seg[1] = (int32_t*)d1;
    sl.cnt[0] = (int32_t*)co;
    sl.cnt[1] = sl.cnt[0] + nt;
    sl.off[0] = sl.cnt[1] + nt;
    sl.off[1] = sl.off[0] + nt;
    sl.list[0] = tag_list;
    sl.count[0] = a.n_tlist;
    sl.append[0] = 1;
    sl.list[1] = b.exits;
    sl.count[1] = a.n_exits;
    sl.append[1] = 0;
  }
  // rcv in; tags and the ring's E and inflow out (the ring is ~6% of a tile)
  a.prof_cells = prof_begin(ctx, "k_tag_local", 1.0);  // the kernel adds its bytes
  PFF_PROF_LAUNCH(ctx, k_tag_local, tg, dim3(kWB), 0, s, a, wbits, sl);
  a.prof_cells = nullptr;
  hipLaunchKernelGGL(k_seg_scan, dim3(2), dim3(1024), 0, s, sl, nt);
  hipLaunchKernelGGL(k_seg_copy, dim3(nt, 2), dim3(kBT), 0, s, sl);
  Lists L;
  L.l[0] = b.L0;
  L.l[1] = b.L1;
  L.c[0] = ctx->dev_words + W_COUNT0;
  L.c[1] = ctx->dev_words + W_COUNT1;
  int64_t* fe = ctx->dev_words + W_EMPTY2;
  hipLaunchKernelGGL(k_ring_link, dim3(list_grid()), dim3(kBT), 0, s, a, (const int32_t*)b.exits,
                     (const int64_t*)a.n_exits, ringP2, b.L0, L.c[0], fe, b.tin, nt);
  a.tin = b.tin;
  // flagged exits mark the ring cells they drain into, the final pass tags
  // below those, the new tags join the list.  All three only add tags: run
  // again after more rounds, they add the tags the earlier flags missed.
  auto finish = [&]() -> int {
    hipLaunchKernelGGL(k_ring_inflow, dim3(list_grid()), dim3(kBT), 0, s, a, (const int32_t*)b.exits,
                       (const int64_t*)a.n_exits, b.tin);
    // rcv and the tile's tags in; new tags out
    a.prof_cells = prof_begin(ctx, "k_tag_final", 5.0);
    PFF_PROF_LAUNCH(ctx, k_tag_final, tg, dim3(kWB), 0, s, a, sl);
    a.prof_cells = nullptr;
    hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(1024), 0, s, sl, nt);
    hipLaunchKernelGGL(k_seg_copy, dim3(nt, 1), dim3(kBT), 0, s, sl);
    return check_hip(hipGetLastError(), "saddle_tags_walk");
  };
  if (rounds > 0) {
    PFF_TRY(count_rounds(ctx, L, a.ringS, ringP2, b.t0, (uint32_t*)b.t1, b.t2, rounds, true));
    return finish();
  }
  const int cap = ceil_log2i(g.N) + 1;
  int done_rounds = 0;
  int64_t live = 1;
  int step = std::max(2, std::min(ctx->last_tag_rounds + 1, cap));
  while (done_rounds < cap) {
    const int r = std::min(step, cap - done_rounds);
    PFF_TRY(count_rounds(ctx, L, a.ringS, ringP2, b.t0, (uint32_t*)b.t1, b.t2, r, true, done_rounds, fe));
    done_rounds += r;
    step = 3;
    // the finishing passes run on the device while the host waits for the
    // list count of the rounds
    PFF_TRY(post_words(ctx));
    PFF_TRY(finish());
    PFF_TRY(wait_words(ctx));
    live = ctx->host_words[L.cur == 0 ? W_COUNT0 : W_COUNT1];
    if (live == 0 || ctx->host_words[W_CYCLE] != 0) break;
  }
  const int64_t f = ctx->host_words[W_EMPTY2];
  ctx->last_tag_rounds = (live == 0 && f < done_rounds) ? std::max<int>((int)f, 1) : done_rounds;
  if (live != 0 || ctx->host_words[W_CYCLE] != 0) return fail(PFF_ENOTCONV, "saddle tags: receivers are not a forest");
  return 0;
}

// Labels bid = init(root) with the tiled contraction, plus the depression
// count and the list of pits (basin ids of interior roots).  tact_in: tiles
// to label (null = all); cells of other tiles keep their labels, which must
// be 0 (they drain to an open edge).  The labels are written by
// tiled_labels_write, which also flags the tiles that still hold a closed
// basin.  *ndep, *npits, *cycle are host outputs; syncs once (per six
// reduced-graph rounds).
int tiled_labels(pff_ctx* ctx, const GridD& g, const int32_t* rec, const uint32_t* tact_in, int64_t* ndep,
                 int32_t* cycle, int* rounds_used, bool mark_cycles) {
  TileBufs b;
  PFF_TRY(tile_bufs(ctx, g, b));
  TileArgs a = tile_args(g, b, ctx, rec);
  a.tact = tact_in;
  a.mark_cycles = mark_cycles ? 1 : 0;
  hipStream_t s = ctx->stream;
  // cycle flag, exit count and pit count in one fill (the word between them
  // is unused)
  static_assert(W_NEXIT == W_CYCLE + 1 && W_NDEP == W_CYCLE + 2 && W_NPIT == W_CYCLE + 3,
                "cycle flag, exit count and pit count lie in four adjacent words");
  PFF_TRY(check_hip(hipMemsetAsync(ctx->dev_words + W_CYCLE, 0, 32, s), "memset"));
  const dim3 tg = tile_grid(g);
  a.prof_cells = prof_begin(ctx, "k_tile", 8.0);  // rcv in, E out
  SegLists sl;
  PFF_TRY(seg_lists(ctx, g, sl, b.exits, a.n_exits, b.pits, a.n_pits));
  PFF_MODE_DISPATCH(g.mode, PFF_PROF_LAUNCH(ctx, (k_tile<M, 1, W_ROOTS>), tg, dim3(kTB), 0, s, a, sl));
  a.prof_cells = nullptr;
  seg_join(ctx, g, sl, 2);
  Lists L;
  L.l[0] = b.L0;
  L.l[1] = b.L1;
  L.c[0] = ctx->dev_words + W_COUNT0;
  L.c[1] = ctx->dev_words + W_COUNT1;
  hipLaunchKernelGGL(k_exit_link, dim3(list_grid()), dim3(kBT), 0, s, a, b.exits, a.n_exits, 1, b.L0, L.c[0],
                     ctx->dev_words + W_EMPTY, nullptr, 0);
  const int cap = ceil_log2i(g.N) + 1;
  if (mark_cycles) {
    // chains that reach a root do so within cap rounds; the others keep
    // kCycleLabel.  Nothing is read back.  After the first rounds only the
    // nodes on or draining into a cycle (and long chains) are left: when the
    // last exact trip left few, the remaining rounds run in one block.
    constexpr int kFirst = 6;
    const int r0 = std::min(kFirst, cap);
    PFF_TRY(label_rounds_run(ctx, L, b.bid2, b.P2, b.t0, b.t1, nullptr, g.N, nullptr, 0, r0));
    const bool short_tail = ctx->hint_xlive <= kShortList;
    PFF_TRY(label_rounds_run(ctx, L, b.bid2, b.P2, b.t0, b.t1, nullptr, g.N, nullptr, r0, cap - r0, nullptr,
                             short_tail, ctx->dev_words + W_HXLIVE));
    return 0;
  }
  // start from the previous call's round count, then confirm on the host
  int done_rounds = 0;
  int64_t live = 1;
  int step = std::max(2, std::min(ctx->last_label_rounds + 1, cap));
  while (done_rounds < cap) {
    const int r = std::min(step, cap - done_rounds);
    PFF_TRY(label_rounds_run(ctx, L, b.bid2, b.P2, b.t0, b.t1, nullptr, g.N, nullptr, done_rounds, r,
                             ctx->dev_words + W_CYCLE));
    done_rounds += r;
    step = 3;
    // one readback: live count (a cycle across tiles), in-tile cycle flag, pits
    PFF_TRY(read_words(ctx));
    live = ctx->host_words[L.cur == 0 ? W_COUNT0 : W_COUNT1];
    if (live == 0 || ctx->host_words[W_CYCLE] != 0) break;
  }
  // rounds that actually had work (the first empty round index, or all)
  const int64_t fe = ctx->host_words[W_EMPTY];
  const int needed = (live == 0 && fe < done_rounds) ? (int)fe : done_rounds;
  ctx->last_label_rounds = std::max(needed, 1);
  done_rounds = ctx->last_label_rounds + 1;
  *ndep = ctx->host_words[W_NPIT];
  *cycle = (live != 0 || ctx->host_words[W_CYCLE] != 0) ? 1 : 0;
  if (rounds_used) *rounds_used = done_rounds;
  return 0;
}

int tiled_labels_write(pff_ctx* ctx, const GridD& g, int32_t* bid, const uint32_t* tact_in, uint32_t* tact_out) {
  TileBufs b;
  PFF_TRY(tile_bufs(ctx, g, b));
  TileArgs a = tile_args(g, b, ctx, nullptr);
  a.tact = tact_in;
  if (tact_out)
    PFF_TRY(check_hip(hipMemsetAsync(tact_out, 0, (size_t)tiles_count(g) * 4, ctx->stream), "memset"));
  hipLaunchKernelGGL(k_label_tiles, dim3((g.nx + 63) / 64, (g.ny + 3) / 4), dim3(64, 4), 0, ctx->stream, a, bid,
                     tact_out);
  return check_hip(hipGetLastError(), "tiled_labels_write");
}

// device list of pits from the last tiled_labels call
int tiled_pits(pff_ctx* ctx, const GridD& g, int32_t** pits, int64_t** npits) {
  TileBufs b;
  PFF_TRY(tile_bufs(ctx, g, b));
  *pits = b.pits;
  *npits = ctx->dev_words + W_NPIT;
  return 0;
}

// ---- basin-level relabel between lakeflow trips ------------------------------
// After a trip, the receivers change only on the carved paths: every cell of a
// carved basin b now drains through its saddle node to its outlet cell, which
// lies in basin bid[outlet(b)] (0 if it drains out), and every other cell
// keeps its path.  The next trip's labels init(root) (lakeflow.py:507-511)
// are therefore the old labels mapped through the basin-level chain
// b -> bid[outlet(b)] -> ... to its end: an uncarved basin (its pit stays a
// root) or 0.  The pits of the next trip are the uncarved basins.  Basin N
// (custom mode, cell N-1 an interior pit) has no per-basin slot, never gets
// an outlet and so stays a pit; a chain that reaches it ends there, as at
// every other uncarved basin.

// M[b] = bid[outlet node of b], or b when b has no outlet (then b is kept in
// the next pits list); M[0] = 0; *np_saved = *np (the rounds after the
// relabel read the old pit count there once the relabel has replaced *np)
__global__ __launch_bounds__(kBT) void k_basin_next(const int32_t* __restrict__ pits, const int64_t* np,
                                                    const int64_t* __restrict__ outlet,
                                                    const int32_t* __restrict__ bid, int32_t N, int32_t* __restrict__ M0,
                                                    int32_t* __restrict__ M1, int32_t* __restrict__ next_pits,
                                                    int64_t* n_next, int64_t* np_saved) {
  const int64_t n = *np;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    M0[0] = 0;
    M1[0] = 0;
    *np_saved = n;
  }
  // the loop bound is uniform per block: block_append needs every thread
  for (int64_t base = (int64_t)blockIdx.x * kBT; base < n; base += (int64_t)gridDim.x * kBT) {
    const int64_t w = base + threadIdx.x;
    int32_t v[1] = {0};
    bool take[1] = {false};
    if (w < n) {
      const int b = pits[w];
      if (b > 0 && b < N) {
        const int64_t o = outlet[b];
        int m = b;
        if (o != INVALID_KEY) {
          const int r = key_index(o);
          m = (r >= 0 && r < N) ? bid[r] : 0;
        } else {
          v[0] = b;
          take[0] = true;
        }
        M0[b] = m;  // M1 gets every pit's entry from the first jump round, if one runs
      } else if (b == N) {
        v[0] = b;
        take[0] = true;
      }
    }
    block_append<1, kBT>(v, take, next_pits, n_next);
  }
}

// Chain ends by chasing: every basin follows M until a fixed point (an
// uncarved basin keeps M[b] = b; 0 is the open edge, and an id out of range
// leads there too), storing each step into its own entry so that later
// chasers take longer strides.  An entry only ever moves further along its
// own chain, so a chaser that reads an entry another thread is updating
// stays on its chain.  A chaser never stores its own id into its entry (a
// chain that comes back to it is a cycle), so every fixed point it meets is
// a true one.  A basin on or feeding a cycle, or on a chain longer than
// max_steps, sets *unfinished; the synchronous rounds then finish from here
// (and report the cycle).
__global__ __launch_bounds__(kBT) void k_basin_chase(const int32_t* __restrict__ pits, const int64_t* np, int32_t N,
                                                     int32_t* M, int max_steps, int64_t* unfinished) {
  const int64_t n = *np;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int b = pits[w];
    if (b <= 0 || b >= N) continue;
    int m = M[b];
    if (m == b) continue;  // uncarved: a fixed point
    bool done = false;
    for (int k = 0; k < max_steps; ++k) {
      if (m == N) {  // basin N: an uncarved end without a slot
        done = true;
        break;
      }
      if (m <= 0 || m > N) {
        M[b] = 0;
        done = true;
        break;
      }
      const int mm = M[m];
      if (mm == m) {
        done = true;
        break;
      }
      if (mm == b) break;  // the chain comes back to b: a cycle
      m = mm;
      M[b] = m;
    }
    if (!done) *unfinished = 1;
  }
}

// one synchronous pointer-jumping round Mout[b] = Min[Min[b]] over the pits;
// *chg = the latest round (1-based) in which a label changed
__global__ __launch_bounds__(kBT) void k_basin_jump(const int32_t* __restrict__ pits, const int64_t* np, int32_t N,
                                                    const int32_t* __restrict__ Min, int32_t* __restrict__ Mout,
                                                    int round, int64_t* chg) {
  const int64_t n = *np;
  for (int64_t base = (int64_t)blockIdx.x * kBT; base < n; base += (int64_t)gridDim.x * kBT) {
    const int64_t w = base + threadIdx.x;
    bool changed = false;
    if (w < n) {
      const int b = pits[w];
      if (b > 0 && b < N) {
        const int m = Min[b];
        const int mm = (m >= 0 && m < N) ? Min[m] : (m == N ? N : 0);
        Mout[b] = mm;
        changed = mm != m;
      }
    }
    // rounds run in order, so a plain store of this round's number is the
    // latest change; one store per block that saw a change
    if (__syncthreads_or(changed) && threadIdx.x == 0) *chg = round;
  }
}

// Synchronous jumping on a cycle can settle on an entry that points to
// itself (a cycle of length 2^k becomes the identity after k rounds, and the
// chase leaves mixed strides that can do the same sooner): a true chain end
// is the open edge or a basin without an outlet.  *bad = 1 if some basin
// ends elsewhere.
__global__ __launch_bounds__(kBT) void k_basin_check_ends(const int32_t* __restrict__ pits, const int64_t* np,
                                                          int32_t N, const int32_t* __restrict__ M,
                                                          const int64_t* __restrict__ outlet, int64_t* bad) {
  const int64_t n = *np;
  for (int64_t w = (int64_t)blockIdx.x * kBT + threadIdx.x; w < n; w += (int64_t)gridDim.x * kBT) {
    const int b = pits[w];
    if (b <= 0 || b >= N) continue;
    const int e = M[b];
    if (e > 0 && e < N && outlet[e] != INVALID_KEY) *bad = 1;
  }
}

// bid <- M[bid] on the tiles of tact_in; tact_out[t] = 1 for the tiles that
// still hold a closed basin, 0 for the others (one block per 64x64 tile, so
// every flag is written and none needs clearing first).  With npit:
// *npit = *n_next, and *n_next is zeroed for the next relabel.
__global__ __launch_bounds__(kBT) void k_relabel_tiles(GridD g, int tiles_x, const uint32_t* __restrict__ tact_in,
                                                       const int32_t* __restrict__ M, int32_t* __restrict__ bid,
                                                       uint32_t* __restrict__ tact_out, int64_t* npit,
                                                       int64_t* n_next) {
  if (npit && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    *npit = *n_next;
    *n_next = 0;
  }
  int bx, by;
  xcd_block(bx, by);
  const int t = by * tiles_x + bx;
  if (tact_in && tact_in[t] == 0u) {  // uniform per block
    if (threadIdx.x == 0 && threadIdx.y == 0) tact_out[t] = 0u;
    return;
  }
  const int col = bx * TS + threadIdx.x;
  constexpr int RR = TS / 4;  // rows per thread
  // all loads, then all gathers, then the stores (a wait for a load also
  // waits for every earlier store)
  int b[RR], nb[RR];
#pragma unroll
  for (int k = 0; k < RR; ++k) {
    const int row = by * TS + threadIdx.y + 4 * k;
    b[k] = (col < g.nx && row < g.ny) ? bid[row * g.nx + col] : 0;
  }
#pragma unroll
  for (int k = 0; k < RR; ++k) nb[k] = (b[k] > 0 && b[k] < g.N) ? M[b[k]] : (b[k] == g.N ? g.N : 0);
  int any = 0;
#pragma unroll
  for (int k = 0; k < RR; ++k) {
    if (b[k] == 0) continue;
    const int row = by * TS + threadIdx.y + 4 * k;
    if (nb[k] != b[k]) bid[row * g.nx + col] = nb[k];
    any |= nb[k] != 0;
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0 && threadIdx.y == 0) tact_out[t] = any ? 1u : 0u;
}

int relabel_from_outlets(pff_ctx* ctx, const GridD& g, int32_t* bid, const int64_t* outlet, int64_t n_prev,
                         const uint32_t* tact_in, uint32_t* tact_out, int64_t* ndep, int32_t* cycle) {
  TileBufs b;
  PFF_TRY(tile_bufs(ctx, g, b));
  const size_t n4 = (size_t)g.N * 4;
  void *m0, *m1, *np2;
  PFF_TRY(scratch(ctx, S_M0, n4, &m0));
  PFF_TRY(scratch(ctx, S_M1, n4, &m1));
  PFF_TRY(scratch(ctx, S_TPIT2, n4, &np2));
  int32_t* M[2] = {(int32_t*)m0, (int32_t*)m1};
  int32_t* next_pits = (int32_t*)np2;
  hipStream_t s = ctx->stream;
  int64_t* npit = ctx->dev_words + W_NPIT;
  int64_t* n_next = ctx->dev_words + W_NPIT2;
  int64_t* chg = ctx->dev_words + W_MCHG;
  const int64_t* nprev_d = ctx->dev_words + W_NPREV;  // the old pit count, kept by k_basin_next
  *cycle = 0;
  static_assert(W_MCHG == W_NPIT2 + 1, "next pit count and change round are adjacent words");
  if (!ctx->relabel_words_zero) PFF_TRY(check_hip(hipMemsetAsync(n_next, 0, 16, s), "memset"));
  ctx->relabel_words_zero = false;
  const unsigned lg = grid_for(std::max<int64_t>(n_prev, 1), kBT, 4096);
  hipLaunchKernelGGL(k_basin_next, dim3(lg), dim3(kBT), 0, s, (const int32_t*)b.pits, (const int64_t*)npit, outlet,
                     (const int32_t*)bid, g.N, M[0], M[1], next_pits, n_next, ctx->dev_words + W_NPREV);
  // chase the chains to their ends in one pass; the synchronous rounds below
  // finish what a cycle or a very long chain leaves unfinished
  hipLaunchKernelGGL(k_basin_chase, dim3(lg), dim3(kBT), 0, s, (const int32_t*)b.pits, (const int64_t*)npit, g.N,
                     M[0], 64, chg);
  PFF_TRY(check_hip(hipGetLastError(), "basin_chase"));
  // The labels are mapped with the chased map while the host reads whether
  // the chase finished.  If it did not, the rounds below finish the map and a
  // second pass maps the labels again: each label is a basin on its own
  // chain, so the final map sends it to the chain's end.  This pass also
  // replaces the pit count; the rounds read the old one from nprev_d.
  PFF_TRY(post_words(ctx));
  hipLaunchKernelGGL(k_relabel_tiles, tile_grid(g), dim3(64, 4), 0, s, g, (g.nx + TS - 1) / TS, tact_in,
                     (const int32_t*)M[0], bid, tact_out, npit, n_next);
  PFF_TRY(wait_words(ctx));
  const bool chased = ctx->host_words[W_MCHG] == 0;
  // read before the rounds' read-backs: the mapping pass zeroes the device word
  const int64_t n2 = ctx->host_words[W_NPIT2];
  if (!chased) PFF_TRY(check_hip(hipMemsetAsync(chg, 0, 8, s), "memset"));
  // a chain of basins is at most n_prev long: ceil(log2 n_prev) + 1 rounds
  // reach every end; a label still changing then means a cycle
  const int cap = ceil_log2i(std::max<int64_t>(n_prev, 2)) + 1;
  int cur = 0, done = 0;
  int step = std::max(2, std::min(cap + 1, ctx->last_jump_rounds));
  while (!chased) {
    for (int k = 0; k < step; ++k) {
      hipLaunchKernelGGL(k_basin_jump, dim3(lg), dim3(kBT), 0, s, (const int32_t*)b.pits, nprev_d, g.N,
                         (const int32_t*)M[cur], M[cur ^ 1], done + k + 1, chg);
      cur ^= 1;
    }
    done += step;
    PFF_TRY(check_hip(hipGetLastError(), "basin_jump"));
    PFF_TRY(read_words(ctx));
    // converged when the last round changed nothing; the ends must then be
    // true ends, or the chains run round a cycle
    if (ctx->host_words[W_MCHG] < done) {
      ctx->last_jump_rounds = std::max<int>((int)ctx->host_words[W_MCHG] + 1, 1);
      PFF_TRY(check_hip(hipMemsetAsync(chg, 0, 8, s), "memset"));
      hipLaunchKernelGGL(k_basin_check_ends, dim3(lg), dim3(kBT), 0, s, (const int32_t*)b.pits, nprev_d,
                         g.N, (const int32_t*)M[cur], outlet, chg);
      PFF_TRY(check_hip(hipGetLastError(), "basin_check_ends"));
      PFF_TRY(read_words(ctx));
      if (ctx->host_words[W_MCHG] != 0) {
        *cycle = 1;
        return 0;
      }
      break;
    }
    if (done >= cap + 1) {
      *cycle = 1;
      return 0;
    }
    step = std::min(4, cap + 1 - done);
  }
  if (!chased)
    hipLaunchKernelGGL(k_relabel_tiles, tile_grid(g), dim3(64, 4), 0, s, g, (g.nx + TS - 1) / TS, tact_in,
                       (const int32_t*)M[cur], bid, tact_out, (int64_t*)nullptr, n_next);
  else
    ctx->relabel_words_zero = true;  // count zeroed by the mapping pass, no change recorded
  PFF_TRY(check_hip(hipGetLastError(), "relabel_tiles"));
  // the uncarved basins become the pits list: swap the two list buffers
  std::swap(ctx->buf[S_TPIT], ctx->buf[S_TPIT2]);
  std::swap(ctx->cap[S_TPIT], ctx->cap[S_TPIT2]);
  *ndep = n2;
  return 0;
}

int tile_flags(pff_ctx* ctx, const GridD& g, uint32_t** flags_a, uint32_t** flags_b) {
  TileBufs b;
  PFF_TRY(tile_bufs(ctx, g, b));
  *flags_a = b.tact;
  *flags_b = b.tact + tiles_count(g);
  return 0;
}

}  // namespace pff

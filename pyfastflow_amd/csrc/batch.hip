// batch.hip -- ragged batch of independent 'normal' grids routed in one launch
// sequence (pff_batch_route).
//
// The tiles are packed into one canvas (grid mode 4): each tile keeps its
// row-major shape, tiles are placed on shelves with one padding row and
// column between them, and padding cells carry code 0, tile edges code 3 and
// tile interiors code 1.  A padding cell has no valid link and counts as an
// outlet, so it is never a pit and never receives flow.  Inside a tile every
// link, receiver, basin label and packed key is what the tile's own 'normal'
// grid gives, and canvas indices grow with tile indices (a row-major tile maps
// to increasing canvas rows at a fixed column offset), so every comparison of
// basin ids or node indices in the lakeflow rules orders the same way.  Two
// things are per tile in the reference and are kept per tile here: the outer
// trip cap ceil(log2 Ndep)+1 (lakeflow.py:500) of each tile's own depression
// count, and the cell that the "no outlet" sentinel decodes to (index 42,
// lakeflow.py:127, read by the cycle pass :255-281), which is cell 42 of the
// tile.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "pff_block.hpp"
#include "pff_internal.hpp"

namespace pff {

struct TileD {
  const float* z;
  int32_t* rcv;
  float* Q;
  int32_t nx, ny;
  int32_t x0, y0;     // canvas position of the tile's cell 0
  int32_t bx;         // 64-wide block columns of the tile
  int32_t blk0;       // first pack block of the tile
};

struct BatchState {
  int32_t ntiles = 0;
  int32_t W = 0;                   // canvas row length
  int32_t nblk = 0;                // pack blocks over all tiles
  const TileD* td = nullptr;       // device tile table
  const int32_t* blk = nullptr;    // device first pack block per tile (+ total)
  const int32_t* tmap = nullptr;   // canvas cell -> tile (tile cells only)
  const int32_t* c42 = nullptr;    // per tile: canvas index of its cell 42, or -1
  uint8_t* f42 = nullptr;          // per basin id
  int32_t* cnt_d = nullptr;        // per tile: pits of the current trip
  uint8_t* frozen_d = nullptr;     // per tile: 1 once its trip cap is used up
  std::vector<int32_t> cnt_h, cap;
  std::vector<uint8_t> frozen_h;
  bool any_frozen = false;
  int64_t frozen_tiles = 0;
};

namespace {

constexpr int kPW = 64, kPH = 4;  // pack block: 64 x 4 cells

__device__ __forceinline__ int tile_of_block(const int32_t* __restrict__ blk_start, int ntiles, int b) {
  int lo = 0, hi = ntiles - 1;  // last t with blk_start[t] <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk_start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// one 64x4 block of one tile: copy z in, write the canvas codes and tile map
__global__ __launch_bounds__(kPW* kPH) void k_batch_pack(const TileD* __restrict__ tiles,
                                                          const int32_t* __restrict__ blk_start, int ntiles,
                                                          int32_t W, float* __restrict__ zc, uint8_t* __restrict__ code,
                                                          int32_t* __restrict__ tmap) {
  const int t = tile_of_block(blk_start, ntiles, blockIdx.x);
  const TileD T = tiles[t];
  const int lb = blockIdx.x - T.blk0;
  const int col = (lb % T.bx) * kPW + threadIdx.x;
  const int row = (lb / T.bx) * kPH + threadIdx.y;
  if (col >= T.nx || row >= T.ny) return;
  const int64_t c = (int64_t)(T.y0 + row) * W + T.x0 + col;
  zc[c] = T.z[(int64_t)row * T.nx + col];
  code[c] = (row == 0 || row == T.ny - 1 || col == 0 || col == T.nx - 1) ? 3 : 1;
  tmap[c] = t;
}

// receivers back to tile-local indices, and Q
__global__ __launch_bounds__(kPW* kPH) void k_batch_unpack(const TileD* __restrict__ tiles,
                                                            const int32_t* __restrict__ blk_start, int ntiles,
                                                            int32_t W, const int32_t* __restrict__ rc,
                                                            const float* __restrict__ Qc) {
  const int t = tile_of_block(blk_start, ntiles, blockIdx.x);
  const TileD T = tiles[t];
  const int lb = blockIdx.x - T.blk0;
  const int col = (lb % T.bx) * kPW + threadIdx.x;
  const int row = (lb / T.bx) * kPH + threadIdx.y;
  if (col >= T.nx || row >= T.ny) return;
  const int64_t c = (int64_t)(T.y0 + row) * W + T.x0 + col;
  const int32_t r = rc[c];
  const int32_t rr = r / W - T.y0;
  const int32_t rcol = r - (r / W) * W - T.x0;
  const int64_t local = (int64_t)row * T.nx + col;
  T.rcv[local] = rr * T.nx + rcol;
  T.Q[local] = Qc[c];
}

// pits (basin ids) of the trip -> per-tile counts.  The pit list comes in
// canvas tile order, so neighbouring entries mostly belong to one batch tile:
// each wave adds one count per run of equal tiles (the run's first lane adds
// the run length), not one atomic per pit on a few hundred shared counters.
__global__ __launch_bounds__(kBT) void k_batch_count(const int32_t* __restrict__ pits, const int64_t* np,
                                                     const int32_t* __restrict__ tmap, int32_t* __restrict__ cnt) {
  const int64_t n = *np;
  const int lane = threadIdx.x & 63;
  for (int64_t base = (int64_t)blockIdx.x * kBT; base < n; base += (int64_t)gridDim.x * kBT) {  // uniform per block
    const int64_t w = base + threadIdx.x;
    const int t = w < n ? tmap[pits[w] - 1] : -1;
    const int tp = __shfl_up(t, 1);
    const uint64_t heads = __ballot(lane == 0 || t != tp);
    if (t >= 0 && ((heads >> lane) & 1ull)) {
      const uint64_t later = lane < 63 ? heads >> (lane + 1) : 0ull;
      const int len = later ? __builtin_ctzll(later) + 1 : 64 - lane;
      atomicAdd(cnt + t, len);
    }
  }
}

// bid <- 0 on the cells of frozen tiles
__global__ __launch_bounds__(kPW* kPH) void k_batch_freeze(const TileD* __restrict__ tiles,
                                                            const int32_t* __restrict__ blk_start, int ntiles,
                                                            int32_t W, const uint8_t* __restrict__ frozen,
                                                            int32_t* __restrict__ bid) {
  const int t = tile_of_block(blk_start, ntiles, blockIdx.x);
  if (!frozen[t]) return;
  const TileD T = tiles[t];
  const int lb = blockIdx.x - T.blk0;
  const int col = (lb % T.bx) * kPW + threadIdx.x;
  const int row = (lb / T.bx) * kPH + threadIdx.y;
  if (col >= T.nx || row >= T.ny) return;
  bid[(int64_t)(T.y0 + row) * W + T.x0 + col] = 0;
}

__global__ void k_batch_mark42(int ntiles, const int32_t* __restrict__ c42, const int32_t* __restrict__ bid, int32_t N,
                               uint8_t* __restrict__ f42, uint8_t v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles || c42[t] < 0) return;
  const int b = bid[c42[t]];
  if (b > 0 && b < N) f42[b] = v;
}

int ceil_log2_i64(int64_t n) {
  int r = 0;
  while ((1ll << r) < n) ++r;
  return r;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int batch_after_labels(pff_ctx* ctx, BatchState* bs, const GridD& d, int trip, int64_t ndep, bool* stop) {
  *stop = false;
  hipStream_t s = ctx->stream;
  const int nt = bs->ntiles;
  if (ndep == 0) {
    std::fill(bs->cnt_h.begin(), bs->cnt_h.end(), 0);
  } else {
    int32_t* pits;
    int64_t* npits;
    PFF_TRY(tiled_pits(ctx, d, &pits, &npits));
    PFF_TRY(check_hip(hipMemsetAsync(bs->cnt_d, 0, (size_t)nt * 4, s), "memset"));
    hipLaunchKernelGGL(k_batch_count, dim3(grid_for(ndep, kBT, 4096)), dim3(kBT), 0, s, (const int32_t*)pits,
                       (const int64_t*)npits, bs->tmap, bs->cnt_d);
    PFF_TRY(check_hip(hipGetLastError(), "batch_count"));
    PFF_TRY(check_hip(hipMemcpyAsync(bs->cnt_h.data(), bs->cnt_d, (size_t)nt * 4, hipMemcpyDeviceToHost, s), "d2h"));
    PFF_TRY(wait_stream(ctx));
  }
  if (trip == 0) {
    // each tile's own cap, lakeflow.py:500 (a tile without depressions
    // returns before the loop, :496-497)
    for (int t = 0; t < nt; ++t) bs->cap[t] = bs->cnt_h[t] > 0 ? ceil_log2_i64(bs->cnt_h[t]) + 1 : 0;
    return 0;
  }
  // a tile whose loop has run its cap trips stops there with its depressions
  // (lakeflow.py:500-514): freeze it
  bool changed = false, live = false;
  for (int t = 0; t < nt; ++t) {
    if (bs->cnt_h[t] == 0) continue;
    if (trip >= bs->cap[t]) {
      if (!bs->frozen_h[t]) {
        bs->frozen_h[t] = 1;
        bs->frozen_tiles++;
        changed = true;
      }
    } else {
      live = true;
    }
  }
  if (changed) {
    bs->any_frozen = true;
    PFF_TRY(check_hip(hipMemcpyAsync(bs->frozen_d, bs->frozen_h.data(), (size_t)nt, hipMemcpyHostToDevice, s), "h2d"));
    // the host copy must stay unchanged until the copy ran
    PFF_TRY(check_hip(hipStreamSynchronize(s), "sync"));
  }
  *stop = !live;
  return 0;
}

int batch_freeze(pff_ctx* ctx, BatchState* bs, const GridD& d, int32_t* bid) {
  (void)d;
  if (!bs->any_frozen) return 0;
  hipLaunchKernelGGL(k_batch_freeze, dim3(bs->nblk), dim3(kPW, kPH), 0, ctx->stream, bs->td, bs->blk, bs->ntiles,
                     bs->W, (const uint8_t*)bs->frozen_d, bid);
  return check_hip(hipGetLastError(), "batch_freeze");
}

int batch_mark42(pff_ctx* ctx, BatchState* bs, const GridD& d, const int32_t* bid, uint8_t value) {
  if (bs->ntiles == 0) return 0;
  hipLaunchKernelGGL(k_batch_mark42, dim3((bs->ntiles + 255) / 256), dim3(256), 0, ctx->stream, bs->ntiles, bs->c42,
                     bid, d.N, bs->f42, value);
  return check_hip(hipGetLastError(), "batch_mark42");
}

// shelf packing: tiles by decreasing height, left to right on shelves of a
// canvas about as wide as it is tall, one padding row and column per tile
static int batch_layout(int32_t ntiles, const pff_tile* tiles, std::vector<TileD>& td, int32_t& W, int32_t& H,
                        std::vector<int32_t>& blk_start) {
  int64_t area = 0;
  int32_t wmax = 0;
  for (int t = 0; t < ntiles; ++t) {
    area += (int64_t)(tiles[t].nx + 1) * (tiles[t].ny + 1);
    wmax = std::max(wmax, tiles[t].nx + 1);
  }
  int64_t w = std::max<int64_t>(wmax, (int64_t)std::ceil(std::sqrt((double)area) * 1.02));
  w = (w + 63) / 64 * 64;
  std::vector<int32_t> order(ntiles);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tiles[a].ny > tiles[b].ny; });
  td.assign(ntiles, TileD{});
  int64_t x = 0, y = 0, shelf = 0;
  for (int t : order) {
    const pff_tile& T = tiles[t];
    if (x + T.nx + 1 > w) {
      y += shelf;
      x = 0;
      shelf = 0;
    }
    td[t].x0 = (int32_t)x;
    td[t].y0 = (int32_t)y;
    x += T.nx + 1;
    shelf = std::max<int64_t>(shelf, T.ny + 1);
  }
  const int64_t h = y + shelf;
  if (w * h >= (int64_t)0x7FFFFFFF) return fail(PFF_EINVAL, "batch_route: the packed canvas has 2^31 cells or more");
  W = (int32_t)w;
  H = (int32_t)std::max<int64_t>(h, 2);
  blk_start.assign(ntiles + 1, 0);
  int64_t nb = 0;
  for (int t = 0; t < ntiles; ++t) {
    TileD& D = td[t];
    D.z = tiles[t].z;
    D.rcv = tiles[t].rcv;
    D.Q = tiles[t].Q;
    D.nx = tiles[t].nx;
    D.ny = tiles[t].ny;
    D.bx = (D.nx + kPW - 1) / kPW;
    D.blk0 = (int32_t)nb;
    blk_start[t] = (int32_t)nb;
    nb += (int64_t)D.bx * ((D.ny + kPH - 1) / kPH);
    if (nb >= (int64_t)0x7FFFFFFF) return fail(PFF_EINVAL, "batch_route: too many cells");
  }
  blk_start[ntiles] = (int32_t)nb;
  return 0;
}

int batch_route(pff_ctx* ctx, int32_t ntiles, const pff_tile* tiles, float dx, int32_t lakeflow, double value,
                int32_t area, pff_batch_status* st) {
  if (st) std::memset(st, 0, sizeof(*st));
  if (ntiles == 0) {
    if (st) st->converged = 1;
    return 0;
  }
  hipStream_t s = ctx->stream;
  std::vector<TileD> td;
  std::vector<int32_t> blk_start;
  int32_t W = 0, H = 0;
  PFF_TRY(batch_layout(ntiles, tiles, td, W, H, blk_start));
  GridD d;
  d.nx = W;
  d.ny = H;
  d.N = W * H;
  d.dx = dx;
  d.mode = 4;
  const size_t N = (size_t)d.N;
  int64_t cells = 0;
  for (int t = 0; t < ntiles; ++t) cells += (int64_t)tiles[t].nx * tiles[t].ny;

  // canvas arrays, one block: f32 z, grad, zp, Q; i32 tmap, rcv, bid, rec_,
  // rec__, snode; i64 outlet, saddle; u8 code, isb, f42
  const size_t a4 = align256(N * 4), a8 = align256(N * 8), a1 = align256(N);
  void* blk;
  PFF_TRY(scratch(ctx, S_BATCH, 4 * a4 + 6 * a4 + 2 * a8 + 3 * a1, &blk));
  char* p = (char*)blk;
  auto take = [&](size_t b) {
    char* q = p;
    p += b;
    return (void*)q;
  };
  float* zc = (float*)take(a4);
  float* grad = (float*)take(a4);
  float* zp = (float*)take(a4);
  float* Qc = (float*)take(a4);
  int32_t* tmap = (int32_t*)take(a4);
  int32_t* rc = (int32_t*)take(a4);
  int32_t* bid = (int32_t*)take(a4);
  int32_t* rec_ = (int32_t*)take(a4);
  int32_t* rec__ = (int32_t*)take(a4);
  int32_t* snode = (int32_t*)take(a4);
  int64_t* outlet = (int64_t*)take(a8);
  int64_t* saddle = (int64_t*)take(a8);
  uint8_t* code = (uint8_t*)take(a1);
  uint8_t* isb = (uint8_t*)take(a1);
  uint8_t* f42 = (uint8_t*)take(a1);

  // per-tile tables: TileD[], blk_start[], c42[], cnt[]
  std::vector<int32_t> c42(ntiles);
  for (int t = 0; t < ntiles; ++t) {
    const TileD& D = td[t];
    c42[t] = (int64_t)D.nx * D.ny > 42 ? (D.y0 + 42 / D.nx) * W + D.x0 + 42 % D.nx : -1;
  }
  const size_t tb = align256(sizeof(TileD) * ntiles), bb = align256(4 * (size_t)(ntiles + 1)),
               cb = align256(4 * (size_t)ntiles);
  void* tbl;
  PFF_TRY(scratch(ctx, S_BATCHT, tb + bb + 3 * cb, &tbl));
  TileD* td_d = (TileD*)tbl;
  int32_t* blk_d = (int32_t*)((char*)tbl + tb);
  int32_t* c42_d = (int32_t*)((char*)tbl + tb + bb);
  int32_t* cnt_d = (int32_t*)((char*)tbl + tb + bb + cb);
  uint8_t* frozen_d = (uint8_t*)((char*)tbl + tb + bb + 2 * cb);
  PFF_TRY(check_hip(hipMemcpyAsync(td_d, td.data(), sizeof(TileD) * ntiles, hipMemcpyHostToDevice, s), "h2d"));
  PFF_TRY(check_hip(hipMemcpyAsync(blk_d, blk_start.data(), 4 * (size_t)(ntiles + 1), hipMemcpyHostToDevice, s),
                    "h2d"));
  PFF_TRY(check_hip(hipMemcpyAsync(c42_d, c42.data(), 4 * (size_t)ntiles, hipMemcpyHostToDevice, s), "h2d"));

  // padding codes 0 everywhere first; z of padding is never compared but is
  // kept finite
  PFF_TRY(check_hip(hipMemsetAsync(code, 0, N, s), "memset"));
  PFF_TRY(check_hip(hipMemsetAsync(zc, 0, N * 4, s), "memset"));
  PFF_TRY(check_hip(hipMemsetAsync(f42, 0, N, s), "memset"));
  const int nblk = blk_start[ntiles];
  hipLaunchKernelGGL(k_batch_pack, dim3(nblk), dim3(kPW, kPH), 0, s, (const TileD*)td_d, (const int32_t*)blk_d,
                     ntiles, W, zc, code, tmap);
  PFF_TRY(check_hip(hipGetLastError(), "batch_pack"));
  d.B = code;
  d.F42 = f42;

  BatchState bs;
  bs.ntiles = ntiles;
  bs.W = W;
  bs.nblk = nblk;
  bs.td = td_d;
  bs.blk = blk_d;
  bs.frozen_d = frozen_d;
  bs.frozen_h.assign(ntiles, 0);
  bs.tmap = tmap;
  bs.c42 = c42_d;
  bs.f42 = f42;
  bs.cnt_d = cnt_d;
  bs.cnt_h.assign(ntiles, 0);
  bs.cap.assign(ntiles, 0);

  const float q0 = (float)(value * (area ? (double)dx * (double)dx : 1.0));
  PFF_TRY(launch_receivers(ctx, d, zc, rc, grad));
  const int32_t carve = lakeflow == 1 ? 1 : 0;
  bool fb = false;
  pff_lake_status ls;
  std::memset(&ls, 0, sizeof(ls));
  if (lakeflow)
    PFF_TRY(reroute_core(ctx, d, bid, rc, rec_, rec__, zc, zp, isb, outlet, saddle, snode, carve, nullptr, &ls, &bs,
                         &fb));
  int32_t conv = 1;
  if (!fb) {
    PFF_TRY(accumulate_core(ctx, d, rc, q0, Qc, &conv));
    hipLaunchKernelGGL(k_batch_unpack, dim3(nblk), dim3(kPW, kPH), 0, s, (const TileD*)td_d, (const int32_t*)blk_d,
                       ntiles, W, (const int32_t*)rc, (const float*)Qc);
    PFF_TRY(check_hip(hipGetLastError(), "batch_unpack"));
  } else {
    // tile by tile with the single-grid drivers, on the canvas arrays as
    // scratch (each tile is smaller than the canvas)
    conv = 1;
    for (int t = 0; t < ntiles; ++t) {
      GridD g1;
      g1.nx = tiles[t].nx;
      g1.ny = tiles[t].ny;
      g1.N = g1.nx * g1.ny;
      g1.dx = dx;
      g1.mode = 0;
      g1.B = nullptr;
      PFF_TRY(launch_receivers(ctx, g1, tiles[t].z, tiles[t].rcv, grad));
      if (lakeflow)
        PFF_TRY(reroute_core(ctx, g1, bid, tiles[t].rcv, rec_, rec__, tiles[t].z, zp, isb, outlet, saddle, snode,
                             carve, nullptr, nullptr));
      int32_t c1 = 1;
      PFF_TRY(accumulate_core(ctx, g1, tiles[t].rcv, q0, tiles[t].Q, &c1));
      conv &= c1;
    }
  }
  PFF_TRY(check_hip(hipStreamSynchronize(s), "sync"));
  if (st) {
    st->cells = cells;
    st->canvas_cells = (int64_t)d.N;
    st->canvas_nx = W;
    st->canvas_ny = H;
    st->ndep0 = ls.ndep0;
    st->n_loop = ls.n_loop;
    st->per_tile = fb ? 1 : 0;
    st->capped_tiles = fb ? 0 : bs.frozen_tiles;
    st->converged = conv;
  }
  return 0;
}

}  // namespace pff

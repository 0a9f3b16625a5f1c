// k_flow.hip -- receivers, donors, rake-and-compress and exact upstream
// counts, packed-key and utility kernels for libpff_hip.so (gfx950).
#include "pff_internal.hpp"

namespace pff {

// ---------------------------------------------------------------------------
// D4 steepest-descent receivers (receivers.py:15-58).
// Each wavefront owns a strip of 64 columns and kRcvRows rows.  Ties keep the
// earliest direction (strict >), invalid links never win, and the slope is
// the IEEE f32 quotient (z_i - z_j) / dx, so receivers match the reference
// bit for bit.
//
// Inside the grid (the strip and its rows away from the border, the common
// case) a wave loads its kRcvRows + 2 rows (and their codes in the custom and
// canvas modes) and, with one load, the columns either side of the strip, all
// before its first store (on gfx9 a wait for a load also waits for every
// store issued before it); left and right elevations come from the neighbour
// lanes, the strip-edge ones by readlane.  A cell whose four links are valid
// needs one division (steepest4).  Elsewhere (grid border, stochastic draws)
// each lane keeps its column's rows above, at and below in registers,
// marching down, and loads the neighbours outside them.  The kernel is bound by its
// vector instructions (a wave64 instruction takes four cycles), which is
// what the interior path cuts.
// ---------------------------------------------------------------------------
constexpr int kRcvRows = 4;
// Uniform draw in [0, 1) for cell i, direction k of one stochastic call
// (RAND_RCV, receivers.py:46-47): splitmix64 (Steele, Lea and Flood 2014) of
// seed + (4i+k) times the golden-ratio constant, top 24 bits times 2^-24
// (exact in f32).  The oracle's ora_uniform01 is the same function.
__device__ __forceinline__ float uniform01(uint64_t seed, int i, int k) {
  uint64_t x = seed + (uint64_t)(4 * (int64_t)i + k) * 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (float)(x >> 40) * 0x1p-24f;
}

// Slope and receiver of a cell whose four links are valid (neighbours i-nx,
// i-1, i+1, i+nx in this order), dx > 0.  fl(d/dx) is non-decreasing in d,
// so the largest slope is q = fl(dmax/dx) and the first direction whose
// difference equals dmax has it.  A direction with d < dmax rounds to q only
// when d is within 2^-22 of dmax relative to it (q normal); such near ties,
// and q tiny, subnormal or infinite, take the four divisions of the
// reference.  NaN differences never win, as in the reference (NaN > s is
// false): fmaxf drops them, and a cell whose differences are all NaN keeps
// itself with slope 0.
__device__ __forceinline__ void steepest4(float zc, float zu, float zl, float zr, float zd, float dx, int i, int nx,
                                          int& rcv_out, float& grad_out) {
  const float d0 = zc - zu, d1 = zc - zl, d2 = zc - zr, d3 = zc - zd;
  const float dm = fmaxf(fmaxf(d0, d1), fmaxf(d2, d3));
  const float q = dm / dx;
  const bool pos = q > 0.0f;
  int r = d2 == dm ? i + 1 : i + nx;
  r = d1 == dm ? i - 1 : r;
  r = d0 == dm ? i - nx : r;
  const float thr = dm - dm * 0x1p-20f;
  const bool near = ((d0 < dm) & (d0 >= thr)) | ((d1 < dm) & (d1 >= thr)) | ((d2 < dm) & (d2 >= thr)) |
                    ((d3 < dm) & (d3 >= thr)) | !((q >= 0x1p-100f) & (q <= 3.402823466e38f) & (dm >= 0x1p-100f));
  rcv_out = pos ? r : i;
  grad_out = pos ? q : 0.0f;
  if (__builtin_expect(pos & near, 0)) {
    const float dd[4] = {d0, d1, d2, d3};
    const int jj[4] = {i - nx, i - 1, i + 1, i + nx};
    int rr = i;
    float sr = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float t = dd[k] / dx;
      if (t > sr) {
        sr = t;
        rr = jj[k];
      }
    }
    rcv_out = rr;
    grad_out = sr;
  }
}

template <int MODE, bool STOCH = false>
__global__ __launch_bounds__(256) void k_receivers(GridD g, const float* __restrict__ z, int32_t* __restrict__ rcv,
                                                   float* __restrict__ grad, uint64_t seed = 0) {
  static_assert(kRcvRows <= 32, "strip-edge columns are held in lanes 0..kRcvRows-1 and 32..31+kRcvRows");
  int bx, by;
  xcd_block(bx, by);
  const int lane = threadIdx.x;
  const int col0 = bx * 64, col = col0 + lane;
  const int r0 = (by * 4 + threadIdx.y) * kRcvRows;
  if (r0 >= g.ny) return;  // uniform per wavefront
  const int nx = g.nx;
  if (!STOCH && col0 > 0 && col0 + 64 < nx && r0 > 0 && r0 + kRcvRows < g.ny && g.dx > 0.0f) {
    // every cell of the wave is inside: rows r0-1 .. r0+kRcvRows and columns
    // col0-1 .. col0+64 exist.  Away from the border a link is the plain D4
    // one in modes 0-2, and in modes 3 and 4 it is valid when both of its
    // cells have a nonzero code (neighbourer_flat.py:360-462 for an interior
    // cell; the canvas rule).
    constexpr bool kCodes = MODE == 3 || MODE == 4;
    float zz[kRcvRows + 2];
    uint32_t bb[kRcvRows + 2];
#pragma unroll
    for (int k = 0; k < kRcvRows + 2; ++k) zz[k] = z[(size_t)(r0 - 1 + k) * nx + col];
    if constexpr (kCodes) {
#pragma unroll
      for (int k = 0; k < kRcvRows + 2; ++k) bb[k] = g.B[(size_t)(r0 - 1 + k) * nx + col];
    }
    // lanes 0..kRcvRows-1: column col0-1 of row r0+lane; lanes 32..: column col0+64
    const int hk = lane & 31;
    const size_t hcell = (size_t)(r0 + hk) * nx + (lane < 32 ? col0 - 1 : col0 + 64);
    float halo = 0.0f;
    uint32_t hcode = 0;
    if (hk < kRcvRows) {
      halo = z[hcell];
      if constexpr (kCodes) hcode = g.B[hcell];
    }
#pragma unroll
    for (int k = 0; k < kRcvRows; ++k) {
      const float zc = zz[k + 1];
      float zl = __shfl_up(zc, 1);
      float zr = __shfl_down(zc, 1);
      zl = lane == 0 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(halo), k)) : zl;
      zr = lane == 63 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(halo), 32 + k)) : zr;
      const int i = (r0 + k) * nx + col;
      int r;
      float s;
      if constexpr (!kCodes) {
        steepest4(zc, zz[k], zl, zr, zz[k + 2], g.dx, i, nx, r, s);
      } else {
        const uint32_t bc = bb[k + 1];
        uint32_t bl = __shfl_up(bc, 1), br = __shfl_down(bc, 1);
        bl = lane == 0 ? __builtin_amdgcn_readlane(hcode, k) : bl;
        br = lane == 63 ? __builtin_amdgcn_readlane(hcode, 32 + k) : br;
        const bool ok[4] = {bb[k] != 0, bl != 0, br != 0, bb[k + 2] != 0};
        if (bc != 0 && ok[0] && ok[1] && ok[2] && ok[3]) {
          steepest4(zc, zz[k], zl, zr, zz[k + 2], g.dx, i, nx, r, s);
        } else {
          // some links invalid: the reference loop over the valid ones
          const float zn[4] = {zz[k], zl, zr, zz[k + 2]};
          const int jj[4] = {i - nx, i - 1, i + 1, i + nx};
          r = i;
          s = 0.0f;
          if (bc != 0) {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
              if (!ok[d]) continue;
              const float t = (zc - zn[d]) / g.dx;
              if (t > s) {
                s = t;
                r = jj[d];
              }
            }
          }
        }
      }
      rcv[i] = r;
      grad[i] = s;
    }
    return;
  }
  const int r1 = min(r0 + kRcvRows, g.ny);
  const bool cv = col < g.nx;
  float zu = (cv && r0 > 0) ? z[(r0 - 1) * g.nx + col] : 0.0f;  // row above
  float zc = cv ? z[r0 * g.nx + col] : 0.0f;                     // current row
  for (int row = r0; row < r1; ++row) {
    const int i = row * g.nx + col;
    const float zd = (cv && row + 1 < g.ny) ? z[i + g.nx] : 0.0f;  // row below
    const float zl = __shfl_up(zc, 1);    // column - 1 (lanes >= 1)
    const float zr = __shfl_down(zc, 1);  // column + 1 (lanes <= 62)
    if (cv) {
      int jn[4];
      float zn[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int j = neighbour<MODE>(g, i, row, col, k);
        jn[k] = j;
        // neighbours outside the registers are loaded, all before any use
        const bool held = j == i - g.nx || j == i + g.nx || (j == i - 1 && lane > 0) ||
                          (j == i + 1 && lane < 63 && col + 1 < g.nx);
        zn[k] = (j >= 0 && !held) ? z[j] : 0.0f;
      }
      int r = i;
      float sr = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int j = jn[k];
        if (j >= 0) {
          const float zj = j == i - g.nx ? zu
                         : j == i + g.nx ? zd
                         : (j == i - 1 && lane > 0) ? zl
                         : (j == i + 1 && lane < 63 && col + 1 < g.nx) ? zr : zn[k];
          float t = (zc - zj) / g.dx;
          if constexpr (STOCH) t = t * uniform01(seed, i, k);  // the weighted slope is also the stored gradient
          if (t > sr) {
            sr = t;
            r = j;
          }
        }
      }
      rcv[i] = r;
      grad[i] = sr;
    }
    zu = zc;
    zc = zd;
  }
}

static dim3 receivers_grid(const GridD& g) {
  return dim3((g.nx + 63) / 64, (g.ny + 4 * kRcvRows - 1) / (4 * kRcvRows));
}

int launch_receivers(pff_ctx* ctx, const GridD& g, const float* z, int32_t* rcv, float* grad) {
  dim3 block(64, 4);
  dim3 grid = receivers_grid(g);
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_receivers<M>, grid, block, 0, ctx->stream, g, z, rcv, grad, 0ull));
  return check_hip(hipGetLastError(), "receivers");
}

int launch_receivers_stochastic(pff_ctx* ctx, const GridD& g, const float* z, int32_t* rcv, float* grad,
                                uint64_t seed) {
  dim3 block(64, 4);
  dim3 grid = receivers_grid(g);
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL((k_receivers<M, true>), grid, block, 0, ctx->stream, g, z, rcv, grad,
                                               seed));
  return check_hip(hipGetLastError(), "receivers (stochastic)");
}

// ---------------------------------------------------------------------------
// rcv2donor (downstream_propag.py:77-96): 4 donor slots per node, slot taken
// with an atomic on the receiver's donor count.
// ---------------------------------------------------------------------------
__global__ void k_rcv2donor(int32_t N, const int32_t* __restrict__ rcv, int32_t* __restrict__ dnr,
                            int32_t* __restrict__ ndnr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const int r = rcv[t];
  if (r != t && (unsigned)r < (unsigned)N) {
    const int old = atomicAdd(&ndnr[r], 1);
    const int64_t did = (int64_t)r * 4 + old;
    if (did < (int64_t)N * 4) dnr[did] = t;
  }
}

int launch_rcv2donor(pff_ctx* ctx, const GridD& g, const int32_t* rcv, int32_t* dnr, int32_t* ndnr) {
  hipLaunchKernelGGL(k_rcv2donor, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g.N, rcv, dnr, ndnr);
  return check_hip(hipGetLastError(), "rcv2donor");
}

// getSrc / updateSrc ping-pong selector (downstream_propag.py:20-56): a node
// that already wrote in round `it` is read from its pre-round buffer.
__device__ __forceinline__ bool get_src(const int32_t* src, int t, int it) {
  const int e = src[t];
  bool flip = e < 0;
  if ((e < 0 ? -e : e) == it + 1) flip = !flip;
  return flip;
}

// One rake-and-compress round (downstream_propag.py:99-179).
__global__ void k_rake_compress(int32_t N, int32_t* __restrict__ dnr, int32_t* __restrict__ ndnr, float* p,
                                int32_t* src, int32_t* __restrict__ dnr_, int32_t* __restrict__ ndnr_, float* p_,
                                int32_t it) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const bool flip = get_src(src, t, it);
  bool worked = false;
  int donors[4] = {-1, -1, -1, -1};
  int todo = flip ? ndnr_[t] : ndnr[t];
  const int64_t base = (int64_t)t * 4;
  const int64_t lim = (int64_t)N * 4;
  float p_added = 0.0f;
  int i = 0;
  while (i < todo && i < 4) {
    if (donors[i] == -1) donors[i] = flip ? dnr_[base + i] : dnr[base + i];
    const int did = donors[i];
    const bool fd = get_src(src, did, it);
    const int nd = fd ? ndnr_[did] : ndnr[did];
    if (nd <= 1) {
      if (!worked) p_added = flip ? p_[t] : p[t];
      worked = true;
      p_added += fd ? p_[did] : p[did];
      if (nd == 0) {
        todo -= 1;
        if (todo > i && base + todo < lim) donors[i] = flip ? dnr_[base + todo] : dnr[base + todo];
        i -= 1;
      } else {
        donors[i] = fd ? dnr_[(int64_t)did * 4] : dnr[(int64_t)did * 4];
      }
    }
    i += 1;
  }
  if (worked) {
    const int m = todo < 4 ? todo : 4;
    if (flip) {
      ndnr[t] = todo;
      p[t] = p_added;
      for (int j = 0; j < m; ++j) dnr[base + j] = donors[j];
    } else {
      ndnr_[t] = todo;
      p_[t] = p_added;
      for (int j = 0; j < m; ++j) dnr_[base + j] = donors[j];
    }
    src[t] = (flip ? 1 : -1) * (it + 1);
  }
}

int launch_rake_compress(pff_ctx* ctx, const GridD& g, int32_t* dnr, int32_t* ndnr, float* p, int32_t* src,
                         int32_t* dnr_, int32_t* ndnr_, float* p_, int32_t iteration) {
  hipLaunchKernelGGL(k_rake_compress, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g.N, dnr, ndnr, p, src,
                     dnr_, ndnr_, p_, iteration);
  return check_hip(hipGetLastError(), "rake_compress_accum");
}

// fuse (downstream_propag.py:59-74)
__global__ void k_fuse(int32_t N, float* A, const int32_t* src, const float* B, int32_t it) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  if (get_src(src, t, it)) A[t] = B[t];
}

int launch_fuse(pff_ctx* ctx, const GridD& g, float* A, const int32_t* src, const float* B, int32_t iteration) {
  hipLaunchKernelGGL(k_fuse, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g.N, A, src, B, iteration);
  return check_hip(hipGetLastError(), "fuse");
}

__global__ void k_scale_counts(int32_t N, const uint32_t* __restrict__ cnt, float q0, float* __restrict__ Q) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N) Q[t] = q0 * (float)cnt[t];
}

int launch_scale_counts(pff_ctx* ctx, const GridD& g, const uint32_t* cnt, float q0, float* Q) {
  hipLaunchKernelGGL(k_scale_counts, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g.N, cnt, q0, Q);
  return check_hip(hipGetLastError(), "scale_counts");
}

// ---------------------------------------------------------------------------
// packed keys and utilities
// ---------------------------------------------------------------------------
__global__ void k_pack(int64_t n, int64_t* arr, const float* ft, const int32_t* it) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) arr[t] = pack_key(ft[t], it[t]);
}
__global__ void k_unpack(int64_t n, const int64_t* arr, float* ft, int32_t* it) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    ft[t] = key_float(arr[t]);
    it[t] = key_index(arr[t]);
  }
}
__global__ void k_swap(int64_t n, uint32_t* a, uint32_t* b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const uint32_t x = a[t];
    a[t] = b[t];
    b[t] = x;
  }
}
__global__ void k_arange(int64_t n, int32_t* a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = (int32_t)t;
}
__global__ void k_fill_u32(int64_t n, uint32_t* a, uint32_t v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}
__global__ void k_fill_i64(int64_t n, int64_t* a, int64_t v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// fill_edges / flow_out_nodes debug maps (neighbourer_flat.py:112-115, 608-611)
template <int MODE>
__global__ void k_edge_maps(GridD g, uint8_t* __restrict__ edges, uint8_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  int row, col;
  rc_of(g, i, row, col);
  if (edges) edges[i] = (uint8_t)edge_class(g, row, col);
  if (out) out[i] = can_leave<MODE>(g, i, row, col) ? 1 : 0;
}

int launch_edge_maps(pff_ctx* ctx, const GridD& g, uint8_t* edges, uint8_t* out) {
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_edge_maps<M>, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g,
                                               edges, out));
  return check_hip(hipGetLastError(), "edge_maps");
}

int launch_pack(pff_ctx* ctx, int64_t n, int64_t* arr, const float* ft, const int32_t* it) {
  hipLaunchKernelGGL(k_pack, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, arr, ft, it);
  return check_hip(hipGetLastError(), "pack");
}
int launch_unpack(pff_ctx* ctx, int64_t n, const int64_t* arr, float* ft, int32_t* it) {
  hipLaunchKernelGGL(k_unpack, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, arr, ft, it);
  return check_hip(hipGetLastError(), "unpack");
}
int launch_swap(pff_ctx* ctx, int64_t n, uint32_t* a, uint32_t* b) {
  hipLaunchKernelGGL(k_swap, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, a, b);
  return check_hip(hipGetLastError(), "swap");
}
int launch_arange(pff_ctx* ctx, int64_t n, int32_t* a) {
  hipLaunchKernelGGL(k_arange, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, a);
  return check_hip(hipGetLastError(), "arange");
}
int launch_fill_u32(pff_ctx* ctx, int64_t n, uint32_t* a, uint32_t v) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, a, v);
  return check_hip(hipGetLastError(), "fill_u32");
}
int launch_fill_i64(pff_ctx* ctx, int64_t n, int64_t* a, int64_t v) {
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, a, v);
  return check_hip(hipGetLastError(), "fill_i64");
}

}  // namespace pff

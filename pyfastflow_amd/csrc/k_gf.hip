// k_gf.hip -- GraphFlood on gfx950: epsilon-fill by pointer jumping,
// gather-form (atomic-free) discharge diffusion and the Manning depth update.
#include "pff_internal.hpp"

// No fused multiply-add contraction in this file: every f32 operation rounds
// as written, as in the oracle's C restatement (built with contraction off),
// so a GraphFlood iteration gives the same bits on both.
#pragma clang fp contract(off)

namespace pff {

// h^(5/3) rounded to f32 (gf_hydrodynamics.py:115, ti.pow(h, 5/3)), from
// IEEE-exact double operations only so that the oracle's ora_pow53 gives the
// same bits: cube root y of x by Newton's method y <- y - (y^3 - x)/(3y^2)
// from the seed bits(x)/3 + 0x2A9F7893782DA1CE (the exponent divided by
// three, within 4% of the root), four steps (quadratic convergence: 4e-2,
// 1e-3, 1e-6, 1e-12, then the last bit), then x*y*y rounded once.  Negative or NaN depth gives NaN, as
// powf does for a non-integer exponent.
__device__ __forceinline__ float pow53(float h) {
  if (!(h >= 0.0f)) return __builtin_nanf("");
  if (h == 0.0f) return 0.0f;
  if (h == __builtin_inff()) return h;
  const double x = (double)h;
  const unsigned long long b = (unsigned long long)__double_as_longlong(x) / 3ull + 0x2A9F7893782DA1CEull;
  double y = __longlong_as_double((long long)b);
#pragma unroll
  for (int k = 0; k < 4; ++k) y = y - (y * y * y - x) / (3.0 * y * y);
  return (float)(x * y * y);
}

// ---------------------------------------------------------------------------
// _topofill passes (fill_topo.py:16-45), synchronous form: every read sees the
// pass's input (ping-pong buffers), one legal interleaving of the in-place
// reference pass.  Pass `it` uses the increment f32(2^(it+first_exp) * eps)
// with first_exp = -1 for fill_z_add_delta (fill_topo.py:121-122) and 0 for
// topofill (fill_topo.py:70-72).
// ---------------------------------------------------------------------------
// Early exit: a pass that changes neither an elevation nor a pointer leaves
// every later pass without work too (its pointers are all at their roots;
// a cell not skipped by the rule then lies above its root's elevation +
// the previous increment, so the next increment cannot raise it either).
// Pass it marks one of the kFillFlags words of its row chg[it][.] (a plain
// store by each block that changed something, spread over separate cache
// lines so that the blocks do not queue on one address) and records
// last[0] = it; a pass whose predecessor's row is all zero returns at once,
// and k_fill_final picks the buffer the last pass wrote.  Fixed grid,
// grid-stride over the cells.
constexpr int kFillFlags = 64, kFillStride = 16;  // words per pass row; 64 B apart
constexpr unsigned kFillBlocks = 2048;
__global__ __launch_bounds__(256) void k_topofill(int32_t N, const float* __restrict__ zh,
                                                  const float* __restrict__ zin, float* __restrict__ zout,
                                                  const int32_t* __restrict__ r0, int32_t* __restrict__ r1,
                                                  float inc, int it, int32_t* chg, int32_t* last,
                                                  unsigned long long* prof) {
  if (it > 0) {
    __shared__ int busy;
    if (threadIdx.x < 64) {
      const bool c = chg[((it - 1) * kFillFlags + threadIdx.x) * kFillStride] != 0;
      const unsigned long long b = __ballot(c);
      if (threadIdx.x == 0) busy = b != 0ull;
    }
    __syncthreads();
    if (!busy) return;  // the previous pass was a no-op
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    last[0] = it;
    if (prof) *prof = (unsigned long long)N;  // cells processed (profiling)
  }
  bool moved = false;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    int r = r0[i];
    if ((unsigned)r >= (unsigned)N) r = i;
    int rr = r0[r];
    if ((unsigned)rr >= (unsigned)N) rr = r;
    r1[i] = rr;
    moved |= rr != r;
    const float zi = zin[i];
    float zo = zi;
    if (!(i == r || (zi > zh[r] && rr == r))) {
      const float c = zin[r] + inc;
      zo = zi > c ? zi : c;
    }
    zout[i] = zo;
    moved |= !(zo == zi);
  }
  if (__syncthreads_or(moved) && threadIdx.x == 0)
    chg[(it * kFillFlags + (blockIdx.x % kFillFlags)) * kFillStride] = 1;
}

// the last pass that ran wrote (zb, rb) if its index is even, else (za, ra):
// leave the elevations in za (= z_) and the pointers in both ra and rb
__global__ void k_fill_final(int32_t N, const int32_t* __restrict__ last, float* __restrict__ za,
                             const float* __restrict__ zb, int32_t* __restrict__ ra, int32_t* __restrict__ rb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if ((last[0] & 1) == 0) {
    za[i] = zb[i];
    ra[i] = rb[i];
  } else {
    rb[i] = ra[i];
  }
}

__global__ void k_copy_f32(int64_t n, const float* __restrict__ a, float* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}
__global__ void k_copy_i32b(int64_t n, const int32_t* __restrict__ a, int32_t* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

// Runs ceil(log2 N) passes.  z_ holds the start elevations and receives the
// filled ones; r0 holds the receivers and receives the final jumped pointers
// (r1 too).
int fill_passes(pff_ctx* ctx, const GridD& g, const float* zh, float* z_, int32_t* r0, int32_t* r1, double eps,
                int first_exp) {
  // ping-pong elevations (N floats) + the passes' change flags (one row of
  // kFillFlags words, kFillStride apart, per pass) and the index of the last
  // pass that ran
  void* zt;
  const size_t zbytes = ((size_t)g.N * 4 + 255) & ~(size_t)255;
  const size_t nflag = (size_t)32 * kFillFlags * kFillStride;
  PFF_TRY(scratch(ctx, S_F1, zbytes + (nflag + 64) * 4, &zt));
  int32_t* chg = (int32_t*)((char*)zt + zbytes);
  int32_t* last = chg + nflag;
  int nit = 0;
  while ((1ll << nit) < (int64_t)g.N) ++nit;  // ceil(log2 N) <= 31
  const dim3 gr(grid_for(g.N)), bl(kBlock);
  const dim3 gp(std::min<unsigned>(kFillBlocks, grid_for(g.N)));
  PFF_TRY(launch_fill_u32(ctx, (int64_t)nit * kFillFlags * kFillStride, (uint32_t*)chg, 0u));
  float* za = z_;
  float* zb = (float*)zt;
  int32_t* ra = r0;
  int32_t* rb = r1;
  for (int it = 0; it < nit; ++it) {
    const float inc = (float)(ldexp(1.0, it + first_exp) * eps);
    // algorithmic bytes per cell: z_, zh[r], r, r[r], z_[r] in; z_, r out
    unsigned long long* prof = prof_begin(ctx, "k_topofill", 28.0);
    PFF_PROF_LAUNCH(ctx, k_topofill, gp, bl, 0, ctx->stream, g.N, zh, za, zb, ra, rb, inc, it, chg, last, prof);
    float* tz = za; za = zb; zb = tz;
    int32_t* tr = ra; ra = rb; rb = tr;
  }
  // reference end state: z_ holds the filled elevations, rec_ and rec__ both
  // the final jumped pointers
  hipLaunchKernelGGL(k_fill_final, gr, bl, 0, ctx->stream, g.N, last, z_, (const float*)zt, r0, r1);
  return check_hip(hipGetLastError(), "topofill");
}

// _apply_fill_z_add_delta (fill_topo.py:81-97)
__global__ void k_apply_fill(int32_t N, float* __restrict__ zh, float* __restrict__ h, const float* __restrict__ z_) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float zf = z_[i];
  h[i] += zf - zh[i];
  zh[i] = zf;
}

int launch_apply_fill(pff_ctx* ctx, const GridD& g, float* zh, float* h, const float* z_) {
  hipLaunchKernelGGL(k_apply_fill, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g.N, zh, h, z_);
  return check_hip(hipGetLastError(), "apply_fill");
}

// ---------------------------------------------------------------------------
// diffuse_Q_constant_prec (gf_hydrodynamics.py:23-75) in gather form.
// Cell j (not can_leave) sends S_jk / sum_k S_jk * Q_j through each link k
// with S_jk = max(0, (zh_j - zh_nk)/dx); can_leave cells send nothing (their
// Q is zeroed first, :46-49).  temp_i = P*dx^2 + sum of what reaches i.
// Instead of the reference's float atomics, each cell gathers from the cells
// that link to it, so the sum order is fixed and the result deterministic.
// ---------------------------------------------------------------------------
template <int MODE>
__device__ __forceinline__ float share_to(const GridD& g, const float* zh, const float* Q, int j, int target) {
  int row, col;
  rc_of(g, j, row, col);
  if (can_leave<MODE>(g, j, row, col)) return 0.0f;
  const float zj = zh[j];
  float s[4];
  int nk[4];
  float sums = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int n = neighbour<MODE>(g, j, row, col, k);
    nk[k] = n;
    const float t = (n >= 0) ? (zj - zh[n]) / g.dx : 0.0f;
    s[k] = t > 0.0f ? t : 0.0f;
    sums += s[k];
  }
  if (sums == 0.0f) return 0.0f;
  const float q = Q[j];
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (nk[k] == target) acc += s[k] / sums * q;
  return acc;
}

template <int MODE>
__global__ void k_diffuse(GridD g, float pin, const float* __restrict__ zh, const float* __restrict__ Q,
                          float* __restrict__ temp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.N) return;
  int row, col;
  rc_of(g, i, row, col);
  float acc = pin;
  if constexpr (MODE != 3) {
    // modes 0-2 have symmetric links: the sources of i are its neighbours
    int prev[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int j = neighbour<MODE>(g, i, row, col, k);
      prev[k] = j;
      bool dup = false;
#pragma unroll
      for (int m = 0; m < k; ++m) dup |= (prev[m] == j);
      if (j >= 0 && !dup) acc += share_to<MODE>(g, zh, Q, j, i);
    }
  } else {
    // custom codes can make links one-way: test every geometric neighbour
    // (both axes wrapped) for a link into i
    int cand[4];
    cand[0] = row > 0 ? i - g.nx : i + g.nx * (g.ny - 1);
    cand[1] = col > 0 ? i - 1 : i + g.nx - 1;
    cand[2] = col < g.nx - 1 ? i + 1 : i - g.nx + 1;
    cand[3] = row < g.ny - 1 ? i + g.nx : i - g.nx * (g.ny - 1);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      bool dup = cand[k] == i;
      for (int m = 0; m < k; ++m) dup |= (cand[m] == cand[k]);
      if (!dup) acc += share_to<MODE>(g, zh, Q, cand[k], i);
    }
  }
  temp[i] = acc;
}

int launch_diffuse(pff_ctx* ctx, const GridD& g, const pff_hydro* p, const float* zh, const float* Q, float* temp) {
  const float pin = (float)(p->prec * (double)g.dx * (double)g.dx);
  PFF_MODE_DISPATCH(g.mode, hipLaunchKernelGGL(k_diffuse<M>, dim3(grid_for(g.N)), dim3(kBlock), 0, ctx->stream, g,
                                               pin, zh, Q, temp));
  return check_hip(hipGetLastError(), "diffuse");
}

// ---------------------------------------------------------------------------
// graphflood_core_cte_mannings (gf_hydrodynamics.py:78-129).  h receives tdh
// twice (:121, :128) and zh once (:129), as in the reference; the second
// kernel applies dh after every slope has been read.
// ---------------------------------------------------------------------------
struct ManningC {
  float dx, edgesw, manning, dt, dx2;
  int update;
};

__global__ void k_manning1(int32_t N, ManningC c, float* __restrict__ h, const float* __restrict__ zh,
                           float* __restrict__ dh, const int32_t* __restrict__ rec, const float* __restrict__ Q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float tS = c.edgesw;
  const int r = rec[i];
  if (r != i && (unsigned)r < (unsigned)N) {
    const float s = (zh[i] - zh[r]) / c.dx;
    tS = s > 1e-4f ? s : 1e-4f;
  }
  const float hi = h[i];
  const float Qo = c.dx * pow53(hi) / c.manning * sqrtf(tS);
  float tdh = (Q[i] - Qo) / c.dx2 * c.dt;
  if (c.update == 0) {
    const float hn = hi + tdh;
    if (hn < 0.0f) tdh += hn;
    h[i] = hn;
    dh[i] = tdh;
  } else {
    // single update, depth clamped at zero: the second kernel adds dh to h and zh
    if (hi + tdh < 0.0f) tdh = -hi;
    dh[i] = tdh;
  }
}

__global__ void k_manning2(int32_t N, float* __restrict__ h, float* __restrict__ zh, const float* __restrict__ dh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float d = dh[i];
  h[i] += d;
  zh[i] += d;
}

// ---------------------------------------------------------------------------
// Fused hydro step of one GraphFlood iteration on a 'normal' grid: the fill's
// apply (fill_topo.py:94-97: h += z_ - zh, zh = z_), ND diffusion passes
// (gf_hydrodynamics.py:23-75) and the Manning update (:78-129), in one pass
// over 64 x 32 tiles.  A block stages the filled surface of its tile with a
// halo of ND + 1 cells in LDS (one independent load per cell, so the loads
// overlap), derives the share fractions f_jk = s_k / sum_k s_k of every
// source in the ND halo once (zh does not change during the diffusion),
// stages Q, runs the ND Jacobi passes in LDS (the trusted region shrinks by
// one cell per pass, so after ND passes the interior is exact), then applies
// Manning to the interior.  Per cell and pass the arithmetic is that of
// k_diffuse (pin, then the shares of the top, left, right and bottom
// sources, each f * Q) and of k_manning1/2, so the result is bit-exact with
// the unfused kernels.  zf is the filled surface (read only), zh the surface
// before the fill (each cell reads and then overwrites its own entry),
// Qin/Qout distinct.
// ---------------------------------------------------------------------------
constexpr int kHX = 64, kHY = 32, kHT = 512;

template <int ND>
__global__ __launch_bounds__(kHT) void k_hydro(GridD g, ManningC c, float pin, const float* __restrict__ zf,
                                               float* zh, float* __restrict__ h, const float* __restrict__ Qin,
                                               float* __restrict__ Qout, const int32_t* __restrict__ rec,
                                               unsigned long long* prof) {
  constexpr int RW = kHX + 2 * ND, RH = kHY + 2 * ND, RN = RW * RH;  // diffusion region (halo ND)
  if (prof && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *prof = (unsigned long long)g.N;
  constexpr int ZW = RW + 2, ZH = RH + 2, ZN = ZW * ZH;               // filled surface (halo ND + 1)
  constexpr int PER = (RN + kHT - 1) / kHT, PERZ = (ZN + kHT - 1) / kHT;
  __shared__ float F[4][RN];
  __shared__ float Qs[RN];
  __shared__ float Z[ZN];
  __shared__ uint8_t snd[RN];
  int bx, by;
  xcd_block(bx, by);
  const int x0 = bx * kHX - ND, y0 = by * kHY - ND;  // region origin (grid coordinates)
  // stage the filled surface (halo ND + 1) and Q (halo ND)
  {
    float zv[PERZ];
#pragma unroll
    for (int m = 0; m < PERZ; ++m) {
      const int l = threadIdx.x + m * kHT;
      const int row = y0 - 1 + l / ZW, col = x0 - 1 + l % ZW;
      zv[m] = (l < ZN && row >= 0 && row < g.ny && col >= 0 && col < g.nx) ? zf[row * g.nx + col] : 0.0f;
    }
    float qv[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int l = threadIdx.x + m * kHT;
      const int row = y0 + l / RW, col = x0 + l % RW;
      qv[m] = (l < RN && row >= 0 && row < g.ny && col >= 0 && col < g.nx) ? Qin[row * g.nx + col] : 0.0f;
    }
#pragma unroll
    for (int m = 0; m < PERZ; ++m) {
      const int l = threadIdx.x + m * kHT;
      if (l < ZN) Z[l] = zv[m];
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int l = threadIdx.x + m * kHT;
      if (l < RN) Qs[l] = qv[m];
    }
  }
  __syncthreads();
  // share fractions of every region cell (0 and not sending outside the grid,
  // on can_leave cells and where no neighbour is lower)
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int l = threadIdx.x + m * kHT;
    if (l >= RN) continue;
    const int lr = l / RW, lc = l - lr * RW;
    const int row = y0 + lr, col = x0 + lc;
    uint8_t sends = 0;
    float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (row >= 0 && row < g.ny && col >= 0 && col < g.nx && !can_leave<0>(g, row * g.nx + col, row, col)) {
      const int zl = (lr + 1) * ZW + lc + 1;
      const float zj = Z[zl];
      // neighbours top, left, right, bottom; on a 'normal' grid a cell that
      // cannot leave is interior, so all four exist
      const float zn[4] = {Z[zl - ZW], Z[zl - 1], Z[zl + 1], Z[zl + ZW]};
      float sk[4];
      float sums = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float t = (zj - zn[k]) / g.dx;
        sk[k] = t > 0.0f ? t : 0.0f;
        sums += sk[k];
      }
      if (!(sums == 0.0f)) {
        sends = 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) f[k] = sk[k] / sums;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k][l] = f[k];
    snd[l] = sends;
  }
  __syncthreads();
  // ND Jacobi passes in LDS: temp_i = pin + sum over the sources j linked to
  // i of f_j(j -> i) * Q_j, sources in the order top, left, right, bottom
  for (int p = 0; p < ND; ++p) {
    float acc[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int l = threadIdx.x + m * kHT;
      acc[m] = 0.0f;
      if (l >= RN) continue;
      const int lr = l / RW, lc = l - lr * RW;
      const int row = y0 + lr, col = x0 + lc;
      if (lr == 0 || lr == RH - 1 || lc == 0 || lc == RW - 1) continue;  // region edge: not trusted
      if (row < 0 || row >= g.ny || col < 0 || col >= g.nx) continue;
      float a = pin;
      // top source (its bottom link, k=3), left (right link), right (left link), bottom (top link)
      if (row > 0 && snd[l - RW]) a += F[3][l - RW] * Qs[l - RW];
      if (col > 0 && snd[l - 1]) a += F[2][l - 1] * Qs[l - 1];
      if (col < g.nx - 1 && snd[l + 1]) a += F[1][l + 1] * Qs[l + 1];
      if (row < g.ny - 1 && snd[l + RW]) a += F[0][l + RW] * Qs[l + RW];
      acc[m] = a;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int l = threadIdx.x + m * kHT;
      if (l >= RN) continue;
      const int lr = l / RW, lc = l - lr * RW;
      if (lr == 0 || lr == RH - 1 || lc == 0 || lc == RW - 1) continue;
      Qs[l] = acc[m];
    }
    __syncthreads();
  }
  // apply of the fill, then Manning, on the interior
  constexpr int PERI = kHX * kHY / kHT;
  float hv[PERI], zo[PERI];
  int rv[PERI];
#pragma unroll
  for (int m = 0; m < PERI; ++m) {
    const int l = threadIdx.x + m * kHT;
    const int row = by * kHY + l / kHX, col = bx * kHX + l % kHX;
    const bool in = row < g.ny && col < g.nx;
    const int i = row * g.nx + col;
    hv[m] = in ? h[i] : 0.0f;
    zo[m] = in ? zh[i] : 0.0f;
    rv[m] = in ? rec[i] : 0;
  }
#pragma unroll
  for (int m = 0; m < PERI; ++m) {
    const int l = threadIdx.x + m * kHT;
    const int row = by * kHY + l / kHX, col = bx * kHX + l % kHX;
    if (row >= g.ny || col >= g.nx) continue;
    const int i = row * g.nx + col;
    const int zl = (row - y0 + 1) * ZW + (col - x0 + 1);
    const float zfi = Z[zl];
    const float h1 = hv[m] + (zfi - zo[m]);
    const float q = Qs[(row - y0) * RW + (col - x0)];
    float tS = c.edgesw;
    const int r = rv[m];
    if (r != i && (unsigned)r < (unsigned)g.N) {
      // the receiver is a D4 neighbour after the carve; read its surface
      // from LDS then, from memory otherwise
      float zr;
      if (r == i - g.nx) zr = Z[zl - ZW];
      else if (r == i - 1 && col > 0) zr = Z[zl - 1];
      else if (r == i + 1 && col < g.nx - 1) zr = Z[zl + 1];
      else if (r == i + g.nx) zr = Z[zl + ZW];
      else zr = zf[r];
      const float sl = (zfi - zr) / c.dx;
      tS = sl > 1e-4f ? sl : 1e-4f;
    }
    const float Qo = c.dx * pow53(h1) / c.manning * sqrtf(tS);
    float tdh = (q - Qo) / c.dx2 * c.dt;
    float hn;
    if (c.update == 0) {
      hn = h1 + tdh;
      if (hn < 0.0f) tdh += hn;
    } else {
      if (h1 + tdh < 0.0f) tdh = -h1;
      hn = h1;
    }
    h[i] = hn + tdh;
    zh[i] = zfi + tdh;
    Qout[i] = q;
  }
}

int launch_hydro_fused(pff_ctx* ctx, const GridD& g, const pff_hydro* p, int n_diffuse, const float* zf, float* zh,
                       float* h, const float* Qin, float* Qout, const int32_t* rec) {
  if (g.mode != 0 || n_diffuse < 1 || n_diffuse > 5) return fail(PFF_EINVAL, "fused hydro step: mode 0, 1-5 passes");
  ManningC c;
  c.dx = g.dx;
  c.edgesw = (float)p->edge_slope;
  c.manning = (float)p->manning;
  c.dt = (float)p->dt;
  c.dx2 = (float)((double)g.dx * (double)g.dx);
  c.update = p->update;
  const float pin = (float)(p->prec * (double)g.dx * (double)g.dx);
  const dim3 gr((g.nx + kHX - 1) / kHX, (g.ny + kHY - 1) / kHY), bl(kHT);
  // algorithmic bytes per cell: z_, zh, h, Q, rec in; h, zh, Q out
  unsigned long long* prof = prof_begin(ctx, "k_hydro", 32.0);
  switch (n_diffuse) {
    case 1: PFF_PROF_LAUNCH(ctx, k_hydro<1>, gr, bl, 0, ctx->stream, g, c, pin, zf, zh, h, Qin, Qout, rec, prof); break;
    case 2: PFF_PROF_LAUNCH(ctx, k_hydro<2>, gr, bl, 0, ctx->stream, g, c, pin, zf, zh, h, Qin, Qout, rec, prof); break;
    case 3: PFF_PROF_LAUNCH(ctx, k_hydro<3>, gr, bl, 0, ctx->stream, g, c, pin, zf, zh, h, Qin, Qout, rec, prof); break;
    case 4: PFF_PROF_LAUNCH(ctx, k_hydro<4>, gr, bl, 0, ctx->stream, g, c, pin, zf, zh, h, Qin, Qout, rec, prof); break;
    default: PFF_PROF_LAUNCH(ctx, k_hydro<5>, gr, bl, 0, ctx->stream, g, c, pin, zf, zh, h, Qin, Qout, rec, prof); break;
  }
  return check_hip(hipGetLastError(), "hydro (fused)");
}

int launch_manning(pff_ctx* ctx, const GridD& g, const pff_hydro* p, float* h, float* zh, float* dh,
                   const int32_t* rec, const float* Q) {
  ManningC c;
  c.dx = g.dx;
  c.edgesw = (float)p->edge_slope;
  c.manning = (float)p->manning;
  c.dt = (float)p->dt;
  c.dx2 = (float)((double)g.dx * (double)g.dx);
  c.update = p->update;
  const dim3 gr(grid_for(g.N)), bl(kBlock);
  hipLaunchKernelGGL(k_manning1, gr, bl, 0, ctx->stream, g.N, c, h, zh, dh, rec, Q);
  hipLaunchKernelGGL(k_manning2, gr, bl, 0, ctx->stream, g.N, h, zh, dh);
  return check_hip(hipGetLastError(), "manning");
}

// bin of |a - b| in the residual histogram: 0 for an exact zero, then eight
// bins per decade from 1e-10 (bin 1 holds everything below) to 1e5 (the last
// bin holds everything above): bin k >= 1 ends at 10^((k - 80) / 8)
__device__ __forceinline__ int resid_bin(float d) {
  if (d == 0.0f) return 0;
  const float l = floorf(8.0f * log10f(d)) + 81.0f;
  return l < 1.0f ? 1 : (l > (float)(kResidBins - 1) ? kResidBins - 1 : (int)l);
}

// max |a - b| over non-NaN cells (as float bits: non-negative floats order as
// integers), NaN count of a, (with out_above) the cells with |a - b| > tol or
// NaN, the sum of |a - b|, and (with hist) its histogram (resid_bin)
__global__ __launch_bounds__(256) void k_maxdiff(int32_t N, const float* __restrict__ a, const float* __restrict__ b,
                                                 uint32_t* out_max, unsigned long long* out_nan, float tol,
                                                 unsigned long long* out_above, double* out_sum, uint32_t* hist) {
  __shared__ uint32_t lh[kResidBins];
  for (int k = threadIdx.x; k < kResidBins; k += blockDim.x) lh[k] = 0u;
  __syncthreads();
  float m = 0.0f;
  double sum = 0.0;
  unsigned nan = 0, above = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const float x = a[i];
    if (x != x) {
      ++nan;
      continue;
    }
    const float d = fabsf(x - b[i]);
    m = d > m ? d : m;
    above += d > tol ? 1u : 0u;
    sum += (double)d;
    if (hist) atomicAdd(&lh[resid_bin(d)], 1u);
  }
  __syncthreads();
  if (hist)
    for (int k = threadIdx.x; k < kResidBins; k += blockDim.x)
      if (lh[k]) atomicAdd(&hist[k], lh[k]);
  for (int o = 32; o > 0; o >>= 1) {
    m = fmaxf(m, __shfl_down(m, o));
    nan += __shfl_down(nan, o);
    above += __shfl_down(above, o);
    sum += __shfl_down(sum, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out_max, __float_as_uint(m));
    if (nan) atomicAdd(out_nan, (unsigned long long)nan);
    if (out_above && (above + nan)) atomicAdd(out_above, (unsigned long long)(above + nan));
    if (out_sum) atomicAdd(out_sum, sum);
  }
}

int launch_maxdiff(pff_ctx* ctx, int32_t N, const float* a, const float* b, uint32_t* out_max,
                   unsigned long long* out_nan, float tol, unsigned long long* out_above, double* out_sum,
                   uint32_t* hist) {
  hipLaunchKernelGGL(k_maxdiff, dim3(1024), dim3(256), 0, ctx->stream, N, a, b, out_max, out_nan, tol, out_above,
                     out_sum, hist);
  return check_hip(hipGetLastError(), "maxdiff");
}

}  // namespace pff

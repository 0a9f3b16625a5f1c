/*
 * pff_oracle.c -- CPU restatement of PyFastFlow's flow-routing and GraphFlood
 * path (reference: bgailleton/pyfastflow, Taichi).  TEST INFRASTRUCTURE ONLY:
 * the parity checker and the CPU baseline, never part of the product.
 *
 * Layout: 1-D row-major arrays, i = row*nx + col, row 0 = top/north, exactly
 * as pyfastflow/flow/neighbourer_flat.py:22-50.
 *
 * Determinism rules where the reference races (all equal a single-threaded
 * ascending-order run unless stated):
 *   - saddle node (lakeflow.py:218-219, multi-writer store): the largest cell
 *     index wins.
 *   - pass 5 cycle removal (lakeflow.py:255-281) reads entries other basins
 *     may invalidate in the same loop: ascending serial order.
 *   - propagate_basin (lakeflow.py:63-80) and _topofill (fill_topo.py:16-45)
 *     update in place: every read sees the values from before the pass (one
 *     legal interleaving of the racy reference pass).
 *   - rcv2donor slot order (downstream_propag.py:77-96): ascending donor id.
 *   - diffuse_Q_constant_prec float atomics (gf_hydrodynamics.py:71):
 *     ascending source cell order, directions 0..3.
 *
 * OpenMP is used only on passes whose result is independent of the order,
 * so OMP_NUM_THREADS changes the speed, never the result (except the f32
 * sum order of ora_accumulate_* via rcv2donor slots, which is run serial).
 *
 * How each part is pinned (tests/test_oracle_*.py, tests/test_golden_routing.py,
 * never by re-typing the formula):
 *   - grid topology: numpy-roll closed forms of the neighbour tables, the
 *     custom codes' rules;
 *   - receivers: a vectorised numpy argmax; stochastic receivers: splitmix64's
 *     published outputs, the draws' moments, the closed-form probability that
 *     the steeper of two links wins;
 *   - accumulation: brute-force path walks, the outlet-flux invariant, the
 *     reference's rake-and-compress schedule converging to the counts;
 *   - labels: init(root) by path walking; bid0[f^(2^P-1)(i)] on receiver
 *     graphs with cycles;
 *   - saddlesort: the priority-flood spill elevation (Barnes et al. 2014) on a
 *     bowl; every saddle node passes the tiled pass's candidate rule;
 *   - carve and jump: the explicit reversal of the pit path; on receiver
 *     graphs with loops, the tags are the forward closure of the saddle
 *     nodes; DEMs on which the sentinel rule and the trip cap fire;
 *   - packed keys: the reference's own vectors (archive/test_pack_unpack.py)
 *     and an independent closed form;
 *   - fill: the closed form on chains, strict descent of the filled surface;
 *     diffusion: mass balance; Manning: its steady state, the double update,
 *     f32 against f64; pow53 against the correctly rounded double power.
 * Parity unpinned (only running Taichi could pin them): whether Taichi's
 * fast math turns x/DX into x*(1/DX) (ora_receivers_div_ambiguity and
 * ora_gradient_div_ambiguity count the cells it would change), Taichi's
 * powf/sqrtf to the ulp, the stream of ti.random().
 */
#include "pff_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NN(g) ((int64_t)(g)->nx * (int64_t)(g)->ny)

int ora_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Grid topology: neighbourer_flat.py                                          */
/* ------------------------------------------------------------------------- */

/* which_edge: neighbourer_flat.py:75-110 (0 interior, 1..8 edge/corner class) */
static inline int which_edge(int32_t nx, int32_t ny, int32_t i) {
  const int32_t N = nx * ny;
  if (i == 0) return 1;
  if (i < nx - 1) return 2;
  if (i == nx - 1) return 3;
  if (i < N - nx) {
    const int32_t tx = i % nx;
    if (tx == 0) return 4;
    if (tx == nx - 1) return 5;
    return 0;
  }
  {
    const int32_t ty = i / nx, tx = i % nx;
    if (ty == ny - 1) {
      if (tx == 0) return 6;
      if (tx == nx - 1) return 8;
      return 7;
    }
  }
  return 0;
}

/* validate_link_n: neighbourer_flat.py:146-159 */
static inline int valid_n(int e, int k) {
  if (e > 0) {
    if (k == 0 && e <= 3) return 0;
    if (k == 1 && (e == 1 || e == 4 || e == 6)) return 0;
    if (k == 2 && (e == 3 || e == 5 || e == 8)) return 0;
    if (k == 3 && e >= 6) return 0;
  }
  return 1;
}
/* validate_link_pew: neighbourer_flat.py:228-237 */
static inline int valid_pew(int e, int k) {
  if (e > 0) {
    if (k == 0 && e <= 3) return 0;
    if (k == 3 && e >= 6) return 0;
  }
  return 1;
}
/* validate_link_pns: neighbourer_flat.py:305-314 */
static inline int valid_pns(int e, int k) {
  if (e > 0) {
    if (k == 1 && (e == 1 || e == 4 || e == 6)) return 0;
    if (k == 2 && (e == 3 || e == 5 || e == 8)) return 0;
  }
  return 1;
}

/* raw neighbours: normal :123-141, periodic EW :193-223, periodic NS :270-300 */
static inline int32_t raw_n(int32_t nx, int32_t i, int k) {
  switch (k) {
    case 0: return i - nx;
    case 1: return i - 1;
    case 2: return i + 1;
    default: return i + nx;
  }
}
static inline int32_t raw_pew(int32_t nx, int32_t i, int k) {
  const int32_t col = i % nx;
  switch (k) {
    case 0: return i - nx;
    case 1: return col > 0 ? i - 1 : i + nx - 1;
    case 2: return col < nx - 1 ? i + 1 : i - nx + 1;
    default: return i + nx;
  }
}
static inline int32_t raw_pns(int32_t nx, int32_t ny, int32_t i, int k) {
  const int32_t row = i / nx;
  switch (k) {
    case 0: return row > 0 ? i - nx : i + nx * (ny - 1);
    case 1: return i - 1;
    case 2: return i + 1;
    default: return row < ny - 1 ? i + nx : i - nx * (ny - 1);
  }
}

/* neighbour(i,k): neighbourer_flat.py:542-569, custom mode :360-462 */
static inline int32_t nb(const ora_grid* g, int32_t i, int k) {
  const int32_t nx = g->nx, ny = g->ny;
  const int e = which_edge(nx, ny, i);
  switch (g->bound_mode) {
    case 1: return valid_pew(e, k) ? raw_pew(nx, i, k) : -1;
    case 2: return valid_pns(e, k) ? raw_pns(nx, ny, i, k) : -1;
    case 3: {
      const uint8_t* B = g->boundaries;
      const int32_t N = nx * ny;
      const uint8_t tb = B[i];
      int ok;
      if (tb == 0) ok = 0;
      else if (tb != 9) ok = valid_n(e, k);
      else if (e == 0) ok = valid_n(e, k);
      else if (e <= 3 || e >= 6) ok = valid_pns(e, k);
      else ok = valid_pew(e, k);
      if (!ok) return -1;
      {
        int32_t node = -1;
        if (tb == 9) node = (k == 0 || k == 3) ? raw_pns(nx, ny, i, k) : raw_pew(nx, i, k);
        else if (tb > 0) node = raw_n(nx, i, k);
        if (node > -1 && node < N && B[node] == 0) node = -1;
        /* neighbour_custom re-checks boundaries[j]==0 (:458); a -1 stays -1 */
        if (node < 0 || node >= N) return -1;
        return node;
      }
    }
    default: return valid_n(e, k) ? raw_n(nx, i, k) : -1;
  }
}

/* can_leave_domain: normal :183-185, pew :259-262, pns :337-340, custom :465-467 */
static inline int can_leave(const ora_grid* g, int32_t i) {
  const int e = which_edge(g->nx, g->ny, i);
  switch (g->bound_mode) {
    case 1: return (e <= 3 || e >= 6) && e != 0;
    case 2: return e == 1 || e == 3 || e == 4 || e == 5 || e == 6 || e == 8;
    case 3: return g->boundaries[i] == 3;
    default: return e > 0;
  }
}

int32_t ora_neighbour(const ora_grid* g, int32_t i, int32_t k) { return nb(g, i, k); }
int32_t ora_can_leave(const ora_grid* g, int32_t i) { return can_leave(g, i); }
int32_t ora_which_edge(const ora_grid* g, int32_t i) { return which_edge(g->nx, g->ny, i); }

/* ------------------------------------------------------------------------- */
/* Packed (f32, i32) key: f32_i32_struct.py:14-103                              */
/* ------------------------------------------------------------------------- */

int64_t ora_pack_key(float f, int32_t idx) {
  uint32_t u;
  memcpy(&u, &f, 4);
  /* flip_float_bits (:27-30): negatives clear the sign, positives invert */
  const uint32_t enc = (u & 0x80000000u) ? (u ^ 0x80000000u) : ~u;
  const uint64_t packed = ((uint64_t)enc << 32) | (uint64_t)(uint32_t)idx;
  /* pack_float_index (:64-75): invert the upper word, keep the lower */
  const uint64_t r = (~packed & 0xFFFFFFFF00000000ull) | (packed & 0xFFFFFFFFull);
  return (int64_t)r;
}

void ora_unpack_key(int64_t key, float* f, int32_t* idx) {
  const uint64_t k = (uint64_t)key;
  /* unpack_float_index (:90-103) */
  const uint64_t un = (~k & 0xFFFFFFFF00000000ull) | (k & 0xFFFFFFFFull);
  const uint32_t enc = (uint32_t)(un >> 32);
  const uint32_t lo = (uint32_t)(un & 0xFFFFFFFFull);
  /* unflip_float_bits (:47-48) */
  const uint32_t u = (enc & 0x80000000u) ? ~enc : (enc ^ 0x80000000u);
  memcpy(f, &u, 4);
  *idx = (int32_t)lo;
}

void ora_pack_array(int64_t n, const float* f, const int32_t* idx, int64_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = ora_pack_key(f[i], idx[i]);
}
void ora_unpack_array(int64_t n, const int64_t* key, float* f, int32_t* idx) {
  for (int64_t i = 0; i < n; ++i) ora_unpack_key(key[i], &f[i], &idx[i]);
}

/* ------------------------------------------------------------------------- */
/* Receivers: receivers.py:15-58                                                */
/* ------------------------------------------------------------------------- */

void ora_compute_receivers(const ora_grid* g, const float* z, int32_t* rcv, float* grad) {
  const int64_t N = NN(g);
  const float dx = g->dx;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    int32_t r = i;
    float sr = 0.0f;
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      /* (z[i]-z[j])/DX in f32 with IEEE division; -1 for invalid links (:43) */
      const float t = (j != -1) ? (z[i] - z[j]) / dx : -1.0f;
      if (j != -1 && t > sr) { /* strict > : first direction wins ties (:50-52) */
        sr = t;
        r = j;
      }
    }
    rcv[i] = r;
    grad[i] = sr;
  }
}

/* Uniform draw in [0, 1) for cell i, direction k of one stochastic call:
 * splitmix64 (Steele, Lea and Flood 2014) of seed + (4i+k) times the golden
 * ratio constant, top 24 bits scaled by 2^-24 (exact in f32).  Stands in for
 * ti.random() of receivers.py:46-47, whose generator is Taichi's own: the
 * draws cannot match the reference's, only their distribution. */
float ora_uniform01(uint64_t seed, int64_t i, int k) {
  uint64_t x = seed + (uint64_t)(4 * i + k) * 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (float)(x >> 40) * 0x1p-24f;
}

/* receivers.py:15-58 with RAND_RCV (:46-47): each valid link's slope is
 * multiplied by a uniform draw before the strict > comparison; the stored
 * gradient is the weighted slope, as in the reference. */
void ora_compute_receivers_stochastic(const ora_grid* g, const float* z, int32_t* rcv, float* grad, uint64_t seed) {
  const int64_t N = NN(g);
  const float dx = g->dx;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    int32_t r = i;
    float sr = 0.0f;
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      if (j == -1) continue; /* -1 * random() <= 0 never beats sr >= 0 */
      const float t = ((z[i] - z[j]) / dx) * ora_uniform01(seed, i, k);
      if (t > sr) {
        sr = t;
        r = j;
      }
    }
    rcv[i] = r;
    grad[i] = sr;
  }
}

int64_t ora_receivers_div_ambiguity(const ora_grid* g, const float* z) {
  const int64_t N = NN(g);
  const float dx = g->dx;
  const float inv = 1.0f / dx;
  int64_t cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    int32_t r1 = i, r2 = i;
    float s1 = 0.0f, s2 = 0.0f;
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      if (j == -1) continue;
      const float d = z[i] - z[j];
      const float t1 = d / dx, t2 = d * inv;
      if (t1 > s1) { s1 = t1; r1 = j; }
      if (t2 > s2) { s2 = t2; r2 = j; }
    }
    if (r1 != r2) cnt++;
  }
  return cnt;
}

/* Cells whose stored gradient (receivers.py:43,55-58) would differ in its bits
   if the division by DX were a multiplication by f32(1/DX) (what a fast-math
   compile may emit); receivers that differ are counted by
   ora_receivers_div_ambiguity. */
int64_t ora_gradient_div_ambiguity(const ora_grid* g, const float* z) {
  const int64_t N = NN(g);
  const float dx = g->dx;
  const float inv = 1.0f / dx;
  int64_t cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    float s1 = 0.0f, s2 = 0.0f;
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      if (j == -1) continue;
      const float d = z[i] - z[j];
      const float t1 = d / dx, t2 = d * inv;
      if (t1 > s1) s1 = t1;
      if (t2 > s2) s2 = t2;
    }
    uint32_t b1, b2;
    memcpy(&b1, &s1, 4);
    memcpy(&b2, &s2, 4);
    if (b1 != b2) cnt++;
  }
  return cnt;
}

/* ------------------------------------------------------------------------- */
/* Accumulation: downstream_propag.py + flowfields.py:254-292                  */
/* ------------------------------------------------------------------------- */

/* rcv2donor: downstream_propag.py:77-96.  Serial ascending order fixes the
   slot order (the reference's atomic order is unspecified). */
void ora_rcv2donor(const ora_grid* g, const int32_t* rcv, int32_t* dnr, int32_t* ndnr) {
  const int64_t N = NN(g);
  for (int64_t t = 0; t < N; ++t) {
    const int32_t r = rcv[t];
    if (r != (int32_t)t) {
      const int32_t old = ndnr[r]++;
      const int64_t did = (int64_t)r * 4 + old;
      if (did < N * 4) dnr[did] = (int32_t)t;
    }
  }
}

/* One rake_compress_accum round (downstream_propag.py:99-179) in its
   synchronous form: every read of another node sees that node's state before
   the round (the src ping-pong guarantees this in the reference, :20-38), and
   the new state goes to the *_n buffers.  Returns 1 if any node worked. */
static int rake_round(int64_t N, const int32_t* dnr, const int32_t* ndnr, const float* p, int32_t* dnr_n,
                      int32_t* ndnr_n, float* p_n) {
  int any = 0;
#pragma omp parallel for reduction(| : any) schedule(static)
  for (int64_t t = 0; t < N; ++t) {
    int32_t donors[4] = {-1, -1, -1, -1};
    int32_t todo = ndnr[t];
    const int64_t base = t * 4;
    int worked = 0;
    float p_added = 0.0f;
    int32_t i = 0;
    while (i < todo && i < 4) {
      if (donors[i] == -1) donors[i] = dnr[base + i];
      const int32_t did = donors[i];
      const int32_t nd = ndnr[did];
      if (nd <= 1) { /* rake (:142-150) */
        if (!worked) p_added = p[t];
        worked = 1;
        p_added += p[did];
        if (nd == 0) { /* drop: swap the last slot in (:153-158) */
          todo -= 1;
          if (todo > i && base + todo < N * 4) donors[i] = dnr[base + todo];
          i -= 1;
        } else { /* compress: splice in the donor's single donor (:159-161) */
          donors[i] = dnr[(int64_t)did * 4];
        }
      }
      i += 1;
    }
    if (worked) {
      ndnr_n[t] = todo;
      p_n[t] = p_added;
      for (int32_t j = 0; j < (todo < 4 ? todo : 4); ++j) dnr_n[base + j] = donors[j];
      for (int32_t j = (todo < 4 ? todo : 4); j < 4; ++j) dnr_n[base + j] = dnr[base + j];
      any = 1;
    } else {
      ndnr_n[t] = ndnr[t];
      p_n[t] = p[t];
      for (int j = 0; j < 4; ++j) dnr_n[base + j] = dnr[base + j];
    }
  }
  return any;
}

int ora_accumulate_rake_compress(const ora_grid* g, const int32_t* rcv, const float* w, int32_t nrounds,
                                 float* Q) {
  const int64_t N = NN(g);
  int32_t* dnr = (int32_t*)calloc((size_t)N * 4, 4);
  int32_t* dnr2 = (int32_t*)calloc((size_t)N * 4, 4);
  int32_t* nd = (int32_t*)calloc((size_t)N, 4);
  int32_t* nd2 = (int32_t*)calloc((size_t)N, 4);
  float* p = (float*)malloc((size_t)N * 4);
  float* p2 = (float*)malloc((size_t)N * 4);
  memcpy(p, w, (size_t)N * 4);
  ora_rcv2donor(g, rcv, dnr, nd);
  for (int32_t r = 0; r < nrounds; ++r) {
    rake_round(N, dnr, nd, p, dnr2, nd2, p2);
    int32_t* ti = dnr; dnr = dnr2; dnr2 = ti;
    ti = nd; nd = nd2; nd2 = ti;
    float* tf = p; p = p2; p2 = tf;
  }
  int done = 1;
  for (int64_t i = 0; i < N; ++i) {
    Q[i] = p[i];
    if (nd[i] != 0) done = 0;
  }
  free(dnr); free(dnr2); free(nd); free(nd2); free(p); free(p2);
  return done;
}

/* accumulate_constant_Q: flowfields.py:254-292.  logn = ceil(log2 N)+1; the
   reference runs rounds 0..logn and fuses at it=logn, which discards the work
   of round logn (the src inversion in getSrc), so logn rounds take effect. */
int ora_accumulate_constant_Q(const ora_grid* g, const int32_t* rcv, double value, int32_t area, float* Q) {
  const int64_t N = NN(g);
  const int32_t logn = (int32_t)ceil(log2((double)N)) + 1;
  const float q0 = (float)(value * (area ? (double)g->dx * (double)g->dx : 1.0));
  float* w = (float*)malloc((size_t)N * 4);
  for (int64_t i = 0; i < N; ++i) w[i] = q0;
  const int done = ora_accumulate_rake_compress(g, rcv, w, logn, Q);
  free(w);
  return done;
}

/* Exact upstream counts by topological order (Kahn).  Returns -1 on a cycle. */
int ora_upstream_count(const ora_grid* g, const int32_t* rcv, uint32_t* cnt) {
  const int64_t N = NN(g);
  int32_t* indeg = (int32_t*)calloc((size_t)N, 4);
  int32_t* stack = (int32_t*)malloc((size_t)N * 4);
  for (int64_t i = 0; i < N; ++i) {
    cnt[i] = 1;
    if (rcv[i] != (int32_t)i) indeg[rcv[i]]++;
  }
  int64_t top = 0, done = 0;
  for (int64_t i = 0; i < N; ++i)
    if (indeg[i] == 0) stack[top++] = (int32_t)i;
  while (top > 0) {
    const int32_t u = stack[--top];
    done++;
    const int32_t r = rcv[u];
    if (r != u) {
      cnt[r] += cnt[u];
      if (--indeg[r] == 0) stack[top++] = r;
    }
  }
  free(indeg);
  free(stack);
  return done == N ? 0 : -1;
}

int ora_accumulate_f64(const ora_grid* g, const int32_t* rcv, const double* w, double* Q) {
  const int64_t N = NN(g);
  int32_t* indeg = (int32_t*)calloc((size_t)N, 4);
  int32_t* stack = (int32_t*)malloc((size_t)N * 4);
  for (int64_t i = 0; i < N; ++i) {
    Q[i] = w[i];
    if (rcv[i] != (int32_t)i) indeg[rcv[i]]++;
  }
  int64_t top = 0, done = 0;
  for (int64_t i = 0; i < N; ++i)
    if (indeg[i] == 0) stack[top++] = (int32_t)i;
  while (top > 0) {
    const int32_t u = stack[--top];
    done++;
    const int32_t r = rcv[u];
    if (r != u) {
      Q[r] += Q[u];
      if (--indeg[r] == 0) stack[top++] = r;
    }
  }
  free(indeg);
  free(stack);
  return done == N ? 0 : -1;
}

/* ------------------------------------------------------------------------- */
/* Lakeflow: lakeflow.py                                                        */
/* ------------------------------------------------------------------------- */

#define INVALID_KEY ((int64_t)0x4CBEBC200000002Aull) /* pack_float_index(1e8, 42), lakeflow.py:127 */

/* depression_counter: lakeflow.py:24-44 */
int64_t ora_depression_counter(const ora_grid* g, const int32_t* rcv) {
  const int64_t N = NN(g);
  int64_t n = 0;
#pragma omp parallel for reduction(+ : n) schedule(static)
  for (int64_t i = 0; i < N; ++i)
    if (rcv[i] == (int32_t)i && !can_leave(g, (int32_t)i)) n++;
  return n;
}

/* basin_id_init: lakeflow.py:46-61 */
void ora_basin_id_init(const ora_grid* g, int32_t* bid) {
  const int64_t N = NN(g);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) bid[i] = can_leave(g, (int32_t)i) ? 0 : (int32_t)(i + 1);
}

/* basin_id_init + npass propagate_basin passes (lakeflow.py:63-80, driver
   :507-511), each pass synchronous. */
void ora_basin_identification(const ora_grid* g, const int32_t* rec, int32_t npass, int32_t* bid) {
  const int64_t N = NN(g);
  int32_t* r0 = (int32_t*)malloc((size_t)N * 4);
  int32_t* r1 = (int32_t*)malloc((size_t)N * 4);
  int32_t* b1 = (int32_t*)malloc((size_t)N * 4);
  memcpy(r0, rec, (size_t)N * 4);
  ora_basin_id_init(g, bid);
  for (int32_t p = 0; p < npass; ++p) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      b1[i] = bid[r0[i]];
      r1[i] = r0[r0[i]];
    }
    memcpy(bid, b1, (size_t)N * 4);
    int32_t* t = r0; r0 = r1; r1 = t;
  }
  free(r0); free(r1); free(b1);
}

/* saddlesort: lakeflow.py:105-281 (five passes). */
static int64_t g_sentinel_hits = 0;

void ora_saddlesort(const ora_grid* g, const int32_t* bid, const float* z, uint8_t* is_border, float* zp,
                    int64_t* basin_saddle, int32_t* saddlenode, int64_t* outlet) {
  const int64_t N = NN(g);
  /* pass 1: border flag and z' (:132-157).  can_leave cells keep is_border. */
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    if (can_leave(g, i)) {
      zp[i] = z[i];
      continue;
    }
    int border = 0;
    float zn = 1e9f;
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      if (j == -1) continue;
      if (bid[j] != bid[i]) {
        border = 1;
        zn = fminf(zn, z[j]);
      }
    }
    is_border[i] = (uint8_t)border;
    zp[i] = border ? fmaxf(z[i], zn) : 1e9f;
  }
  /* pass 2 init (:164-167) */
  for (int64_t i = 0; i < N; ++i) {
    basin_saddle[i] = INVALID_KEY;
    outlet[i] = INVALID_KEY;
    saddlenode[i] = -1;
  }
  /* pass 2: minimum saddle key per basin (:170-194) */
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    if (!is_border[i]) continue;
    const int32_t tb = bid[i];
    if (tb >= N) continue; /* basin id N (custom mode) has no per-basin slot */
    int64_t res = INVALID_KEY;
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      if (j == -1) continue;
      if (bid[j] != tb) {
        const int64_t c = ora_pack_key(zp[i], bid[j]);
        if (c < res) res = c;
      }
    }
    if (res == INVALID_KEY) continue;
    if (res < basin_saddle[tb]) basin_saddle[tb] = res;
  }
  /* pass 3: saddle node (:199-219); ascending order => largest index wins */
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    if (!is_border[i] || bid[i] == 0 || bid[i] >= N) continue;
    float tz;
    int32_t tb;
    ora_unpack_key(basin_saddle[bid[i]], &tz, &tb);
    int here = 0;
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      if (j == -1) continue;
      if (bid[j] == tb && zp[i] == tz) here = 1;
    }
    if (here) saddlenode[bid[i]] = i;
  }
  /* pass 4: outlet of each basin (:226-250) */
  for (int64_t bb = 1; bb < N; ++bb) {
    const int32_t b = (int32_t)bb;
    if (basin_saddle[b] == INVALID_KEY) continue;
    const int32_t node = saddlenode[b];
    if (node < 0) continue; /* cannot happen for a valid saddle; guards an OOB read */
    float tz = 1e9f;
    int32_t rr = -1;
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, node, k);
      if (j == -1) continue;
      if (bid[j] != b && tz > z[j]) {
        tz = z[j];
        rr = j;
      }
    }
    if (rr > -1) {
      const int64_t c = ora_pack_key(tz, rr);
      if (c < outlet[b]) outlet[b] = c;
    }
  }
  /* pass 5: 2-cycle removal (:255-281), serial ascending.  Decoding the
     sentinel yields node 42; for grids with N <= 42 that read is out of
     bounds in the reference and is taken here as "no basin". */
  for (int64_t bb = 1; bb < N; ++bb) {
    const int32_t b = (int32_t)bb;
    if (outlet[b] == INVALID_KEY) continue;
    float tf;
    int32_t r1, r2;
    ora_unpack_key(outlet[b], &tf, &r1);
    const int32_t b1 = (r1 >= 0 && r1 < N) ? bid[r1] : -1;
    if (b1 <= 0 || b1 >= N) continue;
    ora_unpack_key(outlet[b1], &tf, &r2);
    const int32_t b2 = (r2 >= 0 && r2 < N) ? bid[r2] : -1;
    if (b2 == b && b1 < b) {
      if (outlet[b1] == INVALID_KEY) g_sentinel_hits++;
      outlet[b] = INVALID_KEY;
      basin_saddle[b] = INVALID_KEY;
      saddlenode[b] = -1;
    }
  }
}

/* number of pass-5 removals decided by reading an already removed outlet
   (the sentinel, which decodes to node 42) since the last reset */
int64_t ora_sentinel_hits(int32_t reset) {
  const int64_t h = g_sentinel_hits;
  if (reset) g_sentinel_hits = 0;
  return h;
}

/* reroute_jump: lakeflow.py:284-304 */
void ora_reroute_jump(const ora_grid* g, int32_t* rec, const int64_t* outlet) {
  const int64_t N = NN(g);
  for (int64_t i = 0; i < N; ++i) {
    if (outlet[i] == INVALID_KEY) continue;
    float tf;
    int32_t node;
    ora_unpack_key(outlet[i], &tf, &node);
    if (i >= 1) rec[i - 1] = node;
  }
}

/* reroute_carve (lakeflow.py:405-438) with init (:307-334), iteration
   (:336-368) and finalise (:371-403).  rec_in is the working receiver copy
   (router.receivers_); rec receives the carved receivers. */
int ora_reroute_carve(const ora_grid* g, int32_t* rec, const int32_t* rec_in, const int32_t* saddlenode,
                      const int64_t* outlet, uint8_t* tag_out) {
  const int64_t N = NN(g);
  uint8_t* tag = (uint8_t*)calloc((size_t)N, 1);
  uint8_t* tag_ = (uint8_t*)calloc((size_t)N, 1);
  int32_t* r = (int32_t*)malloc((size_t)N * 4);  /* rec during the jump loop */
  int32_t* r_ = (int32_t*)malloc((size_t)N * 4); /* rec_ during the jump loop */
  for (int64_t i = 0; i < N; ++i)
    if (saddlenode[i] != -1) tag[saddlenode[i]] = 1;
  memcpy(tag_, tag, (size_t)N);
  memcpy(r, rec_in, (size_t)N * 4);
  int change = 1, it = 0;
  while (change) {
    it++;
    change = 0;
    for (int64_t i = 0; i < N; ++i) {
      if (tag[i]) tag_[r[i]] = 1;
      r_[i] = r[i];
    }
    for (int64_t i = 0; i < N; ++i) {
      r[i] = r_[r_[i]];
      if (tag[i] != tag_[i]) change = 1;
      tag[i] = tag_[i];
    }
  }
  /* finalise: rec <- pre-carve receivers; reverse tagged links; saddle -> outlet */
  memcpy(rec, rec_in, (size_t)N * 4);
  for (int64_t i = 0; i < N; ++i) {
    const int32_t ri = rec_in[i];
    if (tag[ri] && tag[i] && (int32_t)i != ri) rec[ri] = (int32_t)i;
  }
  for (int64_t i = 0; i < N; ++i) {
    if (outlet[i] != INVALID_KEY) {
      float tf;
      int32_t node;
      ora_unpack_key(outlet[i], &tf, &node);
      rec[saddlenode[i]] = node;
    }
  }
  if (tag_out) memcpy(tag_out, tag, (size_t)N);
  free(tag); free(tag_); free(r); free(r_);
  return it;
}

/* reroute_flow driver: lakeflow.py:466-544 (the code after the `return` at
   :538 is unreachable and is not restated). */
void ora_reroute_flow(const ora_grid* g, const float* z, int32_t* rec, int32_t* rec_out, int32_t carve,
                      ora_lake_log* log) {
  const int64_t N = NN(g);
  int32_t* rec_ = (int32_t*)malloc((size_t)N * 4);
  int32_t* bid = (int32_t*)malloc((size_t)N * 4);
  uint8_t* isb = (uint8_t*)calloc((size_t)N, 1);
  float* zp = (float*)malloc((size_t)N * 4);
  int64_t* bs = (int64_t*)malloc((size_t)N * 8);
  int64_t* ol = (int64_t*)malloc((size_t)N * 8);
  int32_t* sn = (int32_t*)malloc((size_t)N * 4);
  memcpy(rec_, rec, (size_t)N * 4);
  const int64_t ndep = ora_depression_counter(g, rec);
  if (log) { log->n_iter = 0; log->n_loop = 0; log->ndep0 = ndep; }
  if (ndep > 0) {
    const int32_t nouter = (int32_t)ceil(log2((double)ndep)) + 1;
    const int32_t npass = (int32_t)ceil(log2((double)N)) + 1;
    for (int32_t it = 0; it < nouter; ++it) {
      const int64_t ndb = ora_depression_counter(g, rec_);
      if (log) {
        if (log->n_loop < 64) log->ndep[log->n_loop] = ndb;
        log->n_loop++;
      }
      ora_basin_identification(g, rec_, npass, bid);
      if (log && it == 0 && log->bid_first) memcpy(log->bid_first, bid, (size_t)N * 4);
      if (ndb == 0) break;
      ora_saddlesort(g, bid, z, isb, zp, bs, sn, ol);
      if (log) {
        if (log->n_iter < log->max_log_iter) {
          if (log->outlet_log) memcpy(log->outlet_log + (int64_t)log->n_iter * N, ol, (size_t)N * 8);
          if (log->saddle_log) memcpy(log->saddle_log + (int64_t)log->n_iter * N, bs, (size_t)N * 8);
          if (log->saddlenode_log) memcpy(log->saddlenode_log + (int64_t)log->n_iter * N, sn, (size_t)N * 4);
        }
        log->n_iter++;
      }
      if (carve) {
        ora_reroute_carve(g, rec, rec_, sn, ol, NULL);
        memcpy(rec_, rec, (size_t)N * 4);
      } else {
        ora_reroute_jump(g, rec_, ol);
      }
    }
  }
  if (rec_out) memcpy(rec_out, rec_, (size_t)N * 4);
  free(rec_); free(bid); free(isb); free(zp); free(bs); free(ol); free(sn);
}

/* ------------------------------------------------------------------------- */
/* Fill: fill_topo.py                                                           */
/* ------------------------------------------------------------------------- */

/* _topofill passes (fill_topo.py:16-45) in synchronous form.  `first_exp` is
   the exponent of the first increment: -1 for fill_z_add_delta (iteration
   starts at 0, :121-122), 0 for topofill (iteration starts at 1, :70-72).
   The increment is the f32 rounding of 2^e * eps computed in double. */
static void topofill_passes_f32(const ora_grid* g, const float* zh, float* z_, const int32_t* rec, double eps,
                                int first_exp) {
  const int64_t N = NN(g);
  const int32_t nit = (int32_t)ceil(log2((double)N));
  int32_t* r0 = (int32_t*)malloc((size_t)N * 4);
  int32_t* r1 = (int32_t*)malloc((size_t)N * 4);
  float* zn = (float*)malloc((size_t)N * 4);
  memcpy(r0, rec, (size_t)N * 4);
  for (int32_t it = 0; it < nit; ++it) {
    const float inc = (float)(ldexp(1.0, it + first_exp) * eps);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      const int32_t r = r0[i];
      r1[i] = r0[r];
      if ((int32_t)i == r || (z_[i] > zh[r] && r0[r] == r)) {
        zn[i] = z_[i];
        continue;
      }
      const float c = z_[r] + inc;
      zn[i] = z_[i] > c ? z_[i] : c; /* ti.max */
    }
    memcpy(z_, zn, (size_t)N * 4);
    int32_t* t = r0; r0 = r1; r1 = t;
  }
  free(r0); free(r1); free(zn);
}

static void topofill_passes_f64(const ora_grid* g, const double* zh, double* z_, const int32_t* rec, double eps,
                                int first_exp) {
  const int64_t N = NN(g);
  const int32_t nit = (int32_t)ceil(log2((double)N));
  int32_t* r0 = (int32_t*)malloc((size_t)N * 4);
  int32_t* r1 = (int32_t*)malloc((size_t)N * 4);
  double* zn = (double*)malloc((size_t)N * 8);
  memcpy(r0, rec, (size_t)N * 4);
  for (int32_t it = 0; it < nit; ++it) {
    const double inc = ldexp(1.0, it + first_exp) * eps;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      const int32_t r = r0[i];
      r1[i] = r0[r];
      if ((int32_t)i == r || (z_[i] > zh[r] && r0[r] == r)) {
        zn[i] = z_[i];
        continue;
      }
      const double c = z_[r] + inc;
      zn[i] = z_[i] > c ? z_[i] : c;
    }
    memcpy(z_, zn, (size_t)N * 8);
    int32_t* t = r0; r0 = r1; r1 = t;
  }
  free(r0); free(r1); free(zn);
}

/* fill_z_add_delta: fill_topo.py:99-125 (+ _apply_fill_z_add_delta :81-97) */
void ora_fill_z_add_delta(const ora_grid* g, float* zh, float* h, const int32_t* rec, double eps) {
  const int64_t N = NN(g);
  float* z_ = (float*)malloc((size_t)N * 4);
  memcpy(z_, zh, (size_t)N * 4);
  topofill_passes_f32(g, zh, z_, rec, eps, -1);
  for (int64_t i = 0; i < N; ++i) {
    const float d = z_[i] - zh[i];
    h[i] += d;
    zh[i] = z_[i];
  }
  free(z_);
}

void ora_fill_z_add_delta_f64(const ora_grid* g, double* zh, double* h, const int32_t* rec, double eps) {
  const int64_t N = NN(g);
  double* z_ = (double*)malloc((size_t)N * 8);
  memcpy(z_, zh, (size_t)N * 8);
  topofill_passes_f64(g, zh, z_, rec, eps, -1);
  for (int64_t i = 0; i < N; ++i) {
    h[i] += z_[i] - zh[i];
    zh[i] = z_[i];
  }
  free(z_);
}

/* topofill with custom_z=None: fill_topo.py:48-78 (z <- filled z) */
void ora_topofill(const ora_grid* g, float* z, const int32_t* rec, double eps) {
  const int64_t N = NN(g);
  float* z_ = (float*)malloc((size_t)N * 4);
  memcpy(z_, z, (size_t)N * 4);
  topofill_passes_f32(g, z, z_, rec, eps, 0);
  memcpy(z, z_, (size_t)N * 4);
  free(z_);
}

/* ------------------------------------------------------------------------- */
/* GraphFlood: gf_hydrodynamics.py                                              */
/* ------------------------------------------------------------------------- */

/* diffuse_Q_constant_prec: gf_hydrodynamics.py:23-75 (f32, serial scatter) */
void ora_diffuse_Q_constant_prec(const ora_grid* g, const ora_hydro* p, const float* zh, float* Q) {
  const int64_t N = NN(g);
  const float dx = g->dx;
  const float pin = (float)(p->prec * (double)g->dx * (double)g->dx);
  float* temp = (float*)malloc((size_t)N * 4);
  for (int64_t i = 0; i < N; ++i) {
    temp[i] = pin;
    if (can_leave(g, (int32_t)i)) Q[i] = 0.0f;
  }
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    if (can_leave(g, i)) continue;
    float sums = 0.0f;
    float s[4];
    int32_t js[4];
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      js[k] = j;
      const float t = (j != -1) ? (zh[i] - zh[j]) / dx : 0.0f;
      s[k] = t > 0.0f ? t : 0.0f; /* ti.max(0., .) */
      sums += s[k];
    }
    if (sums == 0.0f) continue;
    for (int k = 0; k < 4; ++k) {
      if (js[k] == -1) continue; /* the reference adds 0 to temp[-1] here */
      temp[js[k]] += s[k] / sums * Q[i];
    }
  }
  memcpy(Q, temp, (size_t)N * 4);
  free(temp);
}

void ora_diffuse_Q_constant_prec_f64(const ora_grid* g, const ora_hydro* p, const double* zh, double* Q) {
  const int64_t N = NN(g);
  const double dx = (double)g->dx;
  const double pin = p->prec * dx * dx;
  double* temp = (double*)malloc((size_t)N * 8);
  for (int64_t i = 0; i < N; ++i) {
    temp[i] = pin;
    if (can_leave(g, (int32_t)i)) Q[i] = 0.0;
  }
  for (int64_t ii = 0; ii < N; ++ii) {
    const int32_t i = (int32_t)ii;
    if (can_leave(g, i)) continue;
    double sums = 0.0, s[4];
    int32_t js[4];
    for (int k = 0; k < 4; ++k) {
      const int32_t j = nb(g, i, k);
      js[k] = j;
      const double t = (j != -1) ? (zh[i] - zh[j]) / dx : 0.0;
      s[k] = t > 0.0 ? t : 0.0;
      sums += s[k];
    }
    if (sums == 0.0) continue;
    for (int k = 0; k < 4; ++k) {
      if (js[k] == -1) continue;
      temp[js[k]] += s[k] / sums * Q[i];
    }
  }
  memcpy(Q, temp, (size_t)N * 8);
  free(temp);
}

/* graphflood_core_cte_mannings: gf_hydrodynamics.py:78-129.  h receives tdh
   twice (:121 and :128) and zh once (:129); the negative clamp at :122-123
   only changes dh. */
/* h^(5/3) rounded to f32 (gf_hydrodynamics.py:115 computes ti.pow(h, 5/3)
 * in f32; Taichi's powf is not pinned here).  Built from IEEE-exact double
 * operations only, so that the HIP kernel's copy of this function gives the
 * same bits: the cube root y of x by Newton's method y <- y - (y^3 - x)/(3y^2)
 * from the bit-level seed bits(x)/3 + 0x2A9F7893782DA1CE (a division of the
 * exponent by three, within 4% of the root), four steps (quadratic
 * convergence: 4e-2, 1e-3, 1e-6, 1e-12, then the last bit), then x*y*y
 * rounded once to f32.  Negative
 * and NaN depths give NaN, as powf does for a non-integer exponent.  Equal to
 * f32(pow((double)h, 5/3)) on every value tested (tests/test_oracle_fill_gf.py). */
float ora_pow53(float h) {
  if (h != h || h < 0.0f) return NAN;
  if (h == 0.0f) return 0.0f;
  if (isinf(h)) return INFINITY;
  const double x = (double)h;
  uint64_t b;
  memcpy(&b, &x, 8);
  b = b / 3u + 0x2A9F7893782DA1CEull;
  double y;
  memcpy(&y, &b, 8);
  for (int k = 0; k < 4; ++k) y = y - (y * y * y - x) / (3.0 * y * y);
  return (float)(x * y * y);
}

void ora_pow53_array(int64_t n, const float* in, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = ora_pow53(in[i]);
}

void ora_graphflood_core_cte_mannings(const ora_grid* g, const ora_hydro* p, float* h, float* zh,
                                      const int32_t* rec, const float* Q) {
  const int64_t N = NN(g);
  const float dx = g->dx;
  const float edgesw = (float)p->edge_slope;
  const float man = (float)p->manning;
  const float dt = (float)p->dt;
  const float dx2 = (float)((double)g->dx * (double)g->dx);
  float* dh = (float*)malloc((size_t)N * 4);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    float tS = edgesw;
    if (rec[i] != (int32_t)i) {
      const float s = (zh[i] - zh[rec[i]]) / dx;
      tS = s > 1e-4f ? s : 1e-4f;
    }
    const float Qo = dx * ora_pow53(h[i]) / man * sqrtf(tS);
    float tdh = (Q[i] - Qo) / dx2 * dt;
    if (p->update == 0) {
      h[i] += tdh;
      if (h[i] < 0.0f) tdh += h[i];
    } else if (h[i] + tdh < 0.0f) {
      tdh = -h[i]; /* single update with h >= 0 (not the reference's update) */
    }
    dh[i] = tdh;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    h[i] += dh[i];
    zh[i] += dh[i];
  }
  free(dh);
}

void ora_graphflood_core_cte_mannings_f64(const ora_grid* g, const ora_hydro* p, double* h, double* zh,
                                          const int32_t* rec, const double* Q) {
  const int64_t N = NN(g);
  const double dx = (double)g->dx;
  double* dh = (double*)malloc((size_t)N * 8);
  for (int64_t i = 0; i < N; ++i) {
    double tS = p->edge_slope;
    if (rec[i] != (int32_t)i) {
      const double s = (zh[i] - zh[rec[i]]) / dx;
      tS = s > 1e-4 ? s : 1e-4;
    }
    const double Qo = dx * pow(h[i], 5.0 / 3.0) / p->manning * sqrt(tS);
    double tdh = (Q[i] - Qo) / (dx * dx) * p->dt;
    if (p->update == 0) {
      h[i] += tdh;
      if (h[i] < 0.0) tdh += h[i];
    } else if (h[i] + tdh < 0.0) {
      tdh = -h[i];
    }
    dh[i] = tdh;
  }
  for (int64_t i = 0; i < N; ++i) {
    h[i] += dh[i];
    zh[i] += dh[i];
  }
  free(dh);
}

/* One iteration of Flooder.run_graphflood: gf_fields.py:126-141 */
void ora_graphflood_iteration(const ora_grid* g, const ora_hydro* p, float* zh, float* h, float* Q,
                              int32_t* rec_out) {
  const int64_t N = NN(g);
  int32_t* rec = (int32_t*)malloc((size_t)N * 4);
  float* grad = (float*)malloc((size_t)N * 4);
  ora_compute_receivers(g, zh, rec, grad);
  ora_reroute_flow(g, zh, rec, NULL, 1, NULL);
  ora_fill_z_add_delta(g, zh, h, rec, 1e-3);
  for (int k = 0; k < 5; ++k) ora_diffuse_Q_constant_prec(g, p, zh, Q);
  ora_graphflood_core_cte_mannings(g, p, h, zh, rec, Q);
  if (rec_out) memcpy(rec_out, rec, (size_t)N * 4);
  free(rec);
  free(grad);
}

void ora_graphflood_iteration_f64(const ora_grid* g, const ora_hydro* p, double* zh, double* h, double* Q,
                                  int32_t* rec_out) {
  const int64_t N = NN(g);
  int32_t* rec = (int32_t*)malloc((size_t)N * 4);
  float* grad = (float*)malloc((size_t)N * 4);
  float* z32 = (float*)malloc((size_t)N * 4);
  for (int64_t i = 0; i < N; ++i) z32[i] = (float)zh[i];
  ora_compute_receivers(g, z32, rec, grad);
  ora_reroute_flow(g, z32, rec, NULL, 1, NULL);
  ora_fill_z_add_delta_f64(g, zh, h, rec, 1e-3);
  for (int k = 0; k < 5; ++k) ora_diffuse_Q_constant_prec_f64(g, p, zh, Q);
  ora_graphflood_core_cte_mannings_f64(g, p, h, zh, rec, Q);
  if (rec_out) memcpy(rec_out, rec, (size_t)N * 4);
  free(rec);
  free(grad);
  free(z32);
}

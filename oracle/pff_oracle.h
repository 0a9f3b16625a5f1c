/*
 * pff_oracle.h -- CPU restatement of PyFastFlow's flow-routing and GraphFlood path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product (pyfastflow_amd) never links or calls it.
 *
 * Every function restates one reference kernel or driver pass for pass; the
 * reference file:line is cited at each definition in pff_oracle.c.  Where the
 * reference races (in-place reads, multi-writer stores, float atomics) the
 * oracle fixes one deterministic rule, written at the definition.
 *
 * Parity status of each function (what pins it is in tests/):
 *   neighbour / can_leave ............ pinned (brute-force tables, tests/test_oracle_grid.py)
 *   compute_receivers ................ pinned (numpy closed form + hand cases)
 *   rcv2donor / rake_compress / fuse . pinned (exact upstream counts, flux invariant)
 *   upstream_count / accumulate_f64 .. pinned (brute-force path walk)
 *   pack / unpack key ................ pinned (reference test vectors, archive/test_pack_unpack.py)
 *   basin labels / depression count .. pinned (closed form bid = init(root))
 *   saddlesort / carve / reroute ..... pinned (hand-built DEMs, invariants, priority-flood spill)
 *   fill (topofill, fill_z_add_delta)  pinned (closed-form fixed point on chains)
 *   diffuse / manning ................ pinned (per-cell formulas in FP64, mass balance)
 */
#ifndef PFF_ORACLE_H
#define PFF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t nx, ny;
  float dx;
  int32_t bound_mode;          /* 0 normal, 1 periodic_EW, 2 periodic_NS, 3 custom */
  const uint8_t* boundaries;   /* custom codes (mode 3) else NULL */
} ora_grid;

/* grid topology */
int32_t ora_neighbour(const ora_grid* g, int32_t i, int32_t k);
int32_t ora_can_leave(const ora_grid* g, int32_t i);
int32_t ora_which_edge(const ora_grid* g, int32_t i);

/* packed (f32, i32) key */
int64_t ora_pack_key(float f, int32_t idx);
void    ora_unpack_key(int64_t key, float* f, int32_t* idx);
void    ora_pack_array(int64_t n, const float* f, const int32_t* idx, int64_t* out);
void    ora_unpack_array(int64_t n, const int64_t* key, float* f, int32_t* idx);

/* receivers */
/* stochastic receivers (RAND_RCV): seeded uniform draws, see pff_oracle.c */
float ora_uniform01(uint64_t seed, int64_t i, int k);
void ora_compute_receivers_stochastic(const ora_grid* g, const float* z, int32_t* rcv, float* grad, uint64_t seed);
void ora_compute_receivers(const ora_grid* g, const float* z, int32_t* rcv, float* grad);
/* number of cells whose receiver choice would differ if (z_i - z_j)/dx were
   computed as (z_i - z_j) * fl(1/dx) */
int64_t ora_receivers_div_ambiguity(const ora_grid* g, const float* z);
int64_t ora_gradient_div_ambiguity(const ora_grid* g, const float* z);

/* accumulation: reference schedule (f32) and exact forms */
void ora_rcv2donor(const ora_grid* g, const int32_t* rcv, int32_t* dnr, int32_t* ndnr);
/* one full accumulate_constant_Q with the reference schedule; returns 1 if every
   node finished within the reference's fixed round cap, 0 if the cap truncated */
int  ora_accumulate_constant_Q(const ora_grid* g, const int32_t* rcv, double value, int32_t area,
                               float* Q);
/* generic rake-compress on a user weight field (reference schedule, f32) */
int  ora_accumulate_rake_compress(const ora_grid* g, const int32_t* rcv, const float* w, int32_t nrounds,
                                  float* Q);
/* exact upstream cell counts (including self); returns -1 if rcv has a cycle */
int  ora_upstream_count(const ora_grid* g, const int32_t* rcv, uint32_t* cnt);
/* FP64 upstream sum of arbitrary weights; returns -1 on cycle */
int  ora_accumulate_f64(const ora_grid* g, const int32_t* rcv, const double* w, double* Q);

/* lakeflow pieces */
int64_t ora_depression_counter(const ora_grid* g, const int32_t* rcv);
void ora_basin_id_init(const ora_grid* g, int32_t* bid);
/* bid = init, then npass synchronous pointer-jumping passes on a copy of rec */
void ora_basin_identification(const ora_grid* g, const int32_t* rec, int32_t npass, int32_t* bid);
void ora_saddlesort(const ora_grid* g, const int32_t* bid, const float* z, uint8_t* is_border,
                    float* z_prime, int64_t* basin_saddle, int32_t* saddlenode, int64_t* outlet);
void ora_reroute_jump(const ora_grid* g, int32_t* rec, const int64_t* outlet);
/* pass-5 removals that read the sentinel (node 42) since the last reset */
int64_t ora_sentinel_hits(int32_t reset);
/* returns the number of tag iterations run */
int  ora_reroute_carve(const ora_grid* g, int32_t* rec, const int32_t* rec_in, const int32_t* saddlenode,
                       const int64_t* outlet, uint8_t* tag_out);

typedef struct {
  int32_t max_log_iter;      /* capacity of the per-iteration logs below */
  int32_t n_iter;            /* outer iterations that ran saddlesort */
  int32_t n_loop;            /* outer loop trips (including the one that broke) */
  int64_t ndep0;             /* initial depression count */
  int64_t ndep[64];          /* Ndep_bis per loop trip */
  int32_t* bid_first;        /* N, nullable: labels of the first loop trip */
  int64_t* outlet_log;       /* max_log_iter*N, nullable */
  int64_t* saddle_log;       /* max_log_iter*N, nullable */
  int32_t* saddlenode_log;   /* max_log_iter*N, nullable */
} ora_lake_log;

/* lakeflow.py:reroute_flow.  rec in/out = router.receivers, rec_out = router.receivers_ */
void ora_reroute_flow(const ora_grid* g, const float* z, int32_t* rec, int32_t* rec_out, int32_t carve,
                      ora_lake_log* log);

/* fill (synchronous semantics, see pff_oracle.c) */
void ora_fill_z_add_delta(const ora_grid* g, float* zh, float* h, const int32_t* rec, double eps);
void ora_fill_z_add_delta_f64(const ora_grid* g, double* zh, double* h, const int32_t* rec, double eps);
void ora_topofill(const ora_grid* g, float* z, const int32_t* rec, double eps);

/* GraphFlood kernels */
typedef struct { double prec, manning, edge_slope, dt; int32_t update; } ora_hydro;
void ora_diffuse_Q_constant_prec(const ora_grid* g, const ora_hydro* p, const float* zh, float* Q);
void ora_diffuse_Q_constant_prec_f64(const ora_grid* g, const ora_hydro* p, const double* zh, double* Q);
/* h^(5/3) rounded to f32 from exact double operations (the HIP kernel holds the same function) */
float ora_pow53(float h);
void ora_pow53_array(int64_t n, const float* in, float* out);
void ora_graphflood_core_cte_mannings(const ora_grid* g, const ora_hydro* p, float* h, float* zh,
                                      const int32_t* rec, const float* Q);
void ora_graphflood_core_cte_mannings_f64(const ora_grid* g, const ora_hydro* p, double* h, double* zh,
                                          const int32_t* rec, const double* Q);
/* one run_graphflood iteration in f32 (gf_fields.py:126-141) */
void ora_graphflood_iteration(const ora_grid* g, const ora_hydro* p, float* zh, float* h, float* Q,
                              int32_t* rec_out);
/* the same iteration with the continuous parts in FP64, the discrete routing
   decided on the f32 rounding of zh (so integer decisions match the f32 path) */
void ora_graphflood_iteration_f64(const ora_grid* g, const ora_hydro* p, double* zh, double* h, double* Q,
                                  int32_t* rec_out);

int  ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif

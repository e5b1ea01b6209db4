/*
 * pdhg_hip.h - C-ABI of the MI355X-native restarted-PDHG engine (gfx950).
 *
 * This is the drop-in boundary for the reference solver's PDHG engine
 * (src/pdhg/pdhg.h, src/pdhg/pdhg.cpp and src/gpu/cuda_driver.h of the
 * reference).  Plain pointers and sizes only; no exception crosses it.
 * Functions returning int give 0 on success and -1 on error, with the message
 * in pdhg_last_error() (thread-local), as the reference C API does with
 * pramana_last_error (include/pramana/pramana.h:44, src/api/c_api.cpp:19).
 * Functions returning a handle give NULL on error.
 *
 * Problem form (src/core/model.h:1-10): min c'x + offset  s.t.
 * row_lower <= A x <= row_upper, col_lower <= x <= col_upper, A in CSC.
 * A bound with |v| >= 1e20 is infinite (src/util/common.h:15).
 *
 * Threading: one caller per handle at a time; engine entries may be called
 * from any thread.  All GPU work runs on the engine's own HIP stream.
 */
#ifndef PDHG_HIP_H
#define PDHG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status values keep the numbering of the reference Status enum
 * (src/util/common.h:34-46). */
enum {
  PDHG_NOT_SOLVED = 0,
  PDHG_OPTIMAL = 1,
  PDHG_TIME_LIMIT = 5,
  PDHG_ITERATION_LIMIT = 6,
  PDHG_NUMERICAL_FAILURE = 8,
  PDHG_ERROR = 10
};

/* Backend vector ids, same order as the reference `enum V`
 * (src/pdhg/pdhg.cpp:45).  Column vectors (X..UP) have n*k entries, row
 * vectors (Y..RU) m*k, both k-interleaved: element (i, b) at i*k + b. */
enum {
  PDHG_V_X = 0, PDHG_V_XP, PDHG_V_X0, PDHG_V_ATY, PDHG_V_ATYP, PDHG_V_C, PDHG_V_LO, PDHG_V_UP,
  PDHG_V_Y, PDHG_V_YP, PDHG_V_Y0, PDHG_V_AX, PDHG_V_AXP, PDHG_V_AX0, PDHG_V_RL, PDHG_V_RU,
  PDHG_NV
};

/* Iterate precision. */
enum { PDHG_FP32 = 0, /* FP32 iterates + SpMV, FP64 norms/restarts/KKT */
       PDHG_FP64 = 1, /* FP64 everywhere (the reference's arithmetic)   */
       PDHG_FP32_AUTO = 2 /* FP32 iterates; the run moves to FP64 iterates when
                             half its live instances stop improving (the FP32
                             resolution floor) */ };

/* Read-only LP description (minimization form), CSC matrix. */
typedef struct pdhg_lp {
  int m, n;
  const int* col_start;    /* n+1 */
  const int* row_index;    /* nnz */
  const double* value;     /* nnz */
  const double* cost;      /* n */
  const double* col_lower; /* n */
  const double* col_upper; /* n */
  const double* row_lower; /* m */
  const double* row_upper; /* m */
  double obj_offset;
} pdhg_lp;

/* One case of a family sharing A: LpCase (src/pdhg/pdhg.h:68-70).  NULL
 * means "same as the base model".  x0 / y0 (original space) are a warm
 * start, which the reference does not have (it hard-wires x0 = clamp(0),
 * y0 = 0, src/pdhg/pdhg.cpp:509,519); NULL keeps that cold start.
 * primal_weight > 0 replaces the initial primal weight ||c_hat|| / ||b_hat||
 * (src/pdhg/pdhg.cpp:513-529), e.g. the weight a solve of a neighbouring
 * case ended with; 0 keeps the reference's. */
typedef struct pdhg_case {
  const double* cost;
  const double* col_lower;
  const double* col_upper;
  const double* row_lower;
  const double* row_upper;
  const double* x0;
  const double* y0;
  double primal_weight;
} pdhg_case;

/* PdhgOptions (src/pdhg/pdhg.h:32-42) plus the engine's own fields. */
typedef struct pdhg_options {
  double tolerance;            /* relative KKT tolerance (1e-6) */
  long long max_iterations;    /* 1000000 */
  int check_every;             /* iterations between checks (64); >= 2 */
  int use_gpu;                 /* must be 1: there is no CPU fallback */
  int ruiz_iterations;         /* 10 */
  int pock_chambolle;          /* 1 */
  int rescale;                 /* objective/bound rescaling beta, gamma (1) */
  double reflection;           /* rho (1.0) */
  int log_interval;            /* log every N checks to stderr (0 = off) */
  int precision;               /* PDHG_FP64 (default), PDHG_FP32 or PDHG_FP32_AUTO */
  int use_graph;               /* capture check_every iterations in a hipGraph (1) */
  int lanes_per_row;           /* 0 = autotune; else nnz-per-lane target (1..64) */
  int row_map;                 /* 0 = auto, 1 = row-length buckets, 2 = one mapping,
                                  3, 4, 5 = LDS-staged row tiles of 16, 8, 4 rows
                                  (width 64; else as 1) */
  int xcd_remap;               /* -1 = auto, 0 = off, 1 = contiguous rows per XCD */
  double time_limit;           /* seconds, <= 0: none (Deadline, src/util/timer.h:22) */
  int device;                  /* HIP device ordinal (0) */
  int kernel_timing_reps;      /* > 0: after the solve, time this many launches of each
                                  iteration kernel, then this many iterations, with HIP
                                  events (stats *_kernel_ms) */
  int compact;                 /* 1: once half the lanes that may hold live instances are
                                  live, move the live ones into the first lanes of the run
                                  width measured cheapest (default 1) */
  double progress_seconds;     /* > 0: a progress line on stderr at this interval */
  long long warm_fallback;     /* cold-start fallback of warm-started cases (x0, y0 or
                                  primal_weight given): a case not converged after this
                                  many iterations of its own restarts in its lane from the
                                  reference's cold start (x0 = clamp(0), y0 = 0, initial
                                  primal weight), with a full max_iterations of its own;
                                  its result counts the iterations of both parts.  -1 (the
                                  default) = max_iterations / 4, rounded up to check_every;
                                  0 = never */
  /* Optional stop request, polled on the host at every check (the reference's
   * Deadline, src/util/timer.h:22-40: a time limit or a cancellation flag);
   * a nonzero return ends the solve as a time limit does.  NULL = none. */
  int (*should_stop)(void* user);
  void* stop_user;
} pdhg_options;

/* PdhgStats (src/pdhg/pdhg.h:44-56) plus engine counters. */
typedef struct pdhg_stats {
  long long iterations;
  int restarts;
  double primal_residual, dual_residual, gap, primal_weight;
  double setup_seconds, solve_seconds, kernel_seconds, transfer_seconds;
  long long bytes_uploaded, bytes_downloaded, kernel_launches, graph_launches;
  double spectral_norm;
  int batch;                   /* instances in the run */
  int precision;
  long long fp64_switch_iteration; /* PDHG_FP32_AUTO: iteration of the move to FP64 (-1: none) */
  double spmv_bytes_per_iter;  /* algorithmic bytes of the 2 SpMVs per iteration */
  double iter_bytes;           /* algorithmic bytes of one fused iteration */
  double col_kernel_ms, row_kernel_ms;        /* mean launch time (kernel_timing_reps > 0) */
  double col_kernel_bytes, row_kernel_bytes;  /* algorithmic bytes per launch */
  /* Optional per-check trace of instance 0 (caller-owned, may be NULL). */
  double* trace;               /* records of 8 doubles: it, relP, relD, gap, pobj, dobj, omega, restarted */
  int trace_capacity;
  int trace_length;
  char backend[32];
  char device_name[64];
  /* Row mapping of the iteration kernels of the first run, per direction
   * (column kernel over A^T, row kernel over A): nnz-per-lane target (0 =
   * one row per B-lane group, rows' entries read through the scalar cache),
   * 1 = row-length buckets (grouped form: rows ordered by length) / 2 = one
   * mapping / 3, 4, 5 = LDS-staged tiles of 16, 8, 4 rows, XCD remap on/off.  autotuned = 1
   * when the options left all three on auto and the autotuner chose them
   * (timed on the device at the run's width, cached per matrix and width). */
  int autotuned;
  int col_nnz_per_lane, col_row_map, col_xcd;
  int row_nnz_per_lane, row_row_map, row_xcd;
  double tune_seconds;
  int col_rows_per_wave, row_rows_per_wave; /* B = 64: 1, or 2 rows per wavefront */
  int warm_fallbacks;          /* warm-started cases restarted from the cold start */
  double iter_kernel_ms;       /* mean iteration, column and row kernel in turn
                                  (kernel_timing_reps > 0) */
} pdhg_stats;

/* PdhgResult (src/pdhg/pdhg.h:58-65).  x (n) and row_dual (m) are
 * caller-allocated and may be NULL. */
typedef struct pdhg_result {
  int status;
  int restarts;
  long long iterations;
  double objective, dual_objective, relative_kkt;
  double rel_primal, rel_dual, rel_gap, primal_weight;
  double* x;
  double* row_dual;
} pdhg_result;

/* ---- errors / device --------------------------------------------------- */
const char* pdhg_last_error(void);
void pdhg_default_options(pdhg_options* opts);
/* gpuAvailable (src/pdhg/pdhg.h:79). */
int pdhg_gpu_available(void);
/* gpuInfoJson (src/pdhg/pdhg.h:80, src/pdhg/pdhg.cpp:732-746): writes a JSON
 * object with name, gfx arch, CUs, memory, peak bandwidth into buf. */
int pdhg_gpu_info_json(char* buf, size_t cap);

/* ---- engine entries ---------------------------------------------------- */
/* solvePdhg (src/pdhg/pdhg.h:73).  stats may be NULL. */
int pdhg_solve(const pdhg_lp* lp, const pdhg_options* opts, pdhg_result* result, pdhg_stats* stats);
/* solvePdhgBatch (src/pdhg/pdhg.h:76-77): k cases sharing A in one batched
 * run; results[k].  Beta/gamma/scaling come from the base model as in the
 * reference (src/pdhg/pdhg.cpp:441-450). */
int pdhg_solve_batch(const pdhg_lp* base, const pdhg_case* cases, int k, const pdhg_options* opts,
                     pdhg_result* results, pdhg_stats* batch_stats);
/* Packed uneven batch: `count` unrelated LPs solved together by one
 * persistent kernel with per-instance restart state and a device work
 * queue.  No reference counterpart (the reference solves such LPs one by
 * one through solvePdhg, src/pdhg/pdhg.h:73). */
int pdhg_solve_packed(const pdhg_lp* lps, int count, const pdhg_options* opts, pdhg_result* results,
                      pdhg_stats* stats);

/* prepare (src/pdhg/pdhg.cpp:424-459): Ruiz + Pock-Chambolle factors R[m],
 * C[n] (powers of two), scalars[3] = {beta, gamma, 1.01 ||A_scaled||_2} and
 * the CSR of the scaled matrix (a_ptr[m+1], a_idx[nnz], a_val[nnz]).  Any
 * output may be NULL.  Host FP64, bit-identical to the reference. */
int pdhg_prepare(const pdhg_lp* lp, int ruiz_iterations, int pock_chambolle, int rescale, double* R, double* C,
                 double* scalars, int* a_ptr, int* a_idx, double* a_val);

/* ---- op-level backend (the Backend vtable, src/pdhg/pdhg.cpp:51-70) ----- */
typedef struct pdhg_backend pdhg_backend;
/* init(A, AT, rfac, cfac, k): CSR of the scaled A (m rows) and of its
 * transpose (n rows), both FP64 on the host. */
pdhg_backend* pdhg_backend_create(int m, int n, const int* a_ptr, const int* a_idx, const double* a_val,
                                  const int* t_ptr, const int* t_idx, const double* t_val, const double* rfac,
                                  const double* cfac, int k, int precision, int device);
void pdhg_backend_destroy(pdhg_backend* be);
int pdhg_backend_set(pdhg_backend* be, int vec, const double* data);   /* set (pdhg.cpp:56) */
int pdhg_backend_get(pdhg_backend* be, int vec, double* out);          /* get (pdhg.cpp:57) */
int pdhg_backend_spmv(pdhg_backend* be, int in, int out);              /* spmv (pdhg.cpp:58) */
int pdhg_backend_spmvT(pdhg_backend* be, int in, int out);             /* spmvT (pdhg.cpp:59) */
int pdhg_backend_set_steps(pdhg_backend* be, const double* tau, const double* sigma); /* pdhg.cpp:60 */
int pdhg_backend_primal_step(pdhg_backend* be);                        /* pdhg.cpp:61 */
int pdhg_backend_dual_step(pdhg_backend* be);                          /* pdhg.cpp:62 */
int pdhg_backend_halpern(pdhg_backend* be, double rho);                /* pdhg.cpp:63 */
int pdhg_backend_restart(pdhg_backend* be, const int* mask);           /* pdhg.cpp:64 */
int pdhg_backend_kkt(pdhg_backend* be, double* rows3k, double* cols3k); /* pdhg.cpp:65 */
int pdhg_backend_diff2(pdhg_backend* be, int a, int b, double* out);   /* pdhg.cpp:66 */
/* Times `reps` back-to-back SpMVs (transpose: 0 = A x, 1 = A^T y) with HIP
 * events on the backend stream; writes the mean kernel time in ms. */
int pdhg_backend_time_spmv(pdhg_backend* be, int transpose, int reps, double* ms);
/* The storage the backend's SpMV (transpose: 0 = A x, 1 = A^T y) runs on,
 * the fastest the backend timed: 0 = CSR rows in the tuned mapping, 1 = sliced
 * ELLPACK (width 1), 2 = CSR rows in LDS-staged tiles (width 33..64).
 * Engine-side; the reference's backend has one form. */
int pdhg_backend_spmv_layout(pdhg_backend* be, int transpose, int* layout);
/* Forces a storage with its default plan: 0 = one row per wavefront (rows
 * bucketed by length below width 64), 1 = sliced ELLPACK (width 1 only),
 * 2 = 8-row LDS tiles (width 33..64 only; -1 if a row exceeds their budget). */
int pdhg_backend_set_spmv_layout(pdhg_backend* be, int transpose, int layout);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The KKT sums of pdhg_backend_kkt over caller-owned
 * DEVICE arrays, k-interleaved like the backend's vectors, in FP64 (double
 * arrays) or FP32 (float arrays, precision PDHG_FP32); rfac[m] and cfac[n] are
 * FP64.  rows3k[3k] and cols3k[3k] are device FP64 outputs, bit-identical to
 * pdhg_backend_kkt's on the same data.  The work is enqueued on `stream` (a
 * hipStream_t of the current device; NULL = the null stream) and not waited
 * for.  It calls no allocating routine: the block partials live in one fixed
 * device array of the library (2 x 256 blocks x 64 instances x 3 doubles =
 * 768 KiB per GPU, held until the process ends), so k is limited to 1..64 and
 * calls that overlap on one GPU race.  Runs on the current device. */
int pdhg_kkt_noalloc(int m, int n, int k, int precision, const void* axp, const void* rl, const void* ru,
                     const void* yp, const double* rfac, const void* xp, const void* c, const void* atyp,
                     const void* lo, const void* up, const double* cfac, double* rows3k, double* cols3k,
                     void* stream);

/* ---- certifier handoff (src/cert/certifier.h:198-201) ------------------ */
/* Rigorous Neumaier-Shcherbina lower bound on min c'x valid for any y,
 * outward-rounded FP64 interval arithmetic on the host. */
int pdhg_safe_dual_bound(const pdhg_lp* lp, const double* y, double* bound, int* infinite_terms,
                         double* cost_perturbation);
/* Family acceptance of the reference (src/parametric/parametric.cpp:509-524):
 * implied bounds by activity propagation, safe bound, |obj - LB| <=
 * 1e-4 (1 + |obj|).  out[4] = {safe bound, |obj-LB|/(1+|obj|), max primal
 * violation (relative), accepted (0/1)}. */
int pdhg_certify(const pdhg_lp* lp, const double* x, const double* y, double objective, double* out);
/* The same acceptance for k cases of a family (the per-case loop of
 * src/parametric/parametric.cpp:505-525) on `threads` host threads; cases
 * that share bounds share one propagation.  xs (k x n, may be NULL), ys
 * (k x m), objectives[k]; out[4k] as pdhg_certify. */
int pdhg_certify_batch(const pdhg_lp* base, const pdhg_case* cases, int k, const double* xs, const double* ys,
                       const double* objectives, int threads, double* out);

/* ---- learned warm start (no reference counterpart: the reference hard-wires
 * x0 = clamp(0), y0 = 0, src/pdhg/pdhg.cpp:509,519) ----------------------- */
/* POD basis of k solution snapshots Z (k x d, row-major), host FP64 with a
 * cyclic Jacobi eigensolver on the snapshot Gram matrix.  Writes mean[d],
 * basis[d x rank] (row-major), coeffs[k x rank], energy[rank]; returns the
 * number of modes kept (>= 1) or -1. */
int pdhg_pod_basis(const double* Z, int k, int d, int rank, double* mean, double* basis, double* coeffs,
                   double* energy);
/* out (N x d) = mean + coeffs (N x rank) basis^T on BF16 MFMA tiles with FP32
 * accumulation (device 0). */
int pdhg_decode_bf16(const double* basis, const double* mean, int d, int rank, const double* coeffs, int N,
                     double* out);
/* MLP in -> hidden (GELU) -> hidden (GELU) -> out; forward, backward and Adam
 * are HIP kernels on BF16 MFMA tiles with FP32 accumulation and FP32 master
 * weights.  Parameters: W1[hidden x in], b1, W2[hidden x hidden], b2,
 * W3[out x hidden], b3 in one vector.  Loss: mean squared error. */
typedef struct pdhg_mlp pdhg_mlp;
pdhg_mlp* pdhg_mlp_create(int in, int hidden, int out, uint64_t seed, int device);
void pdhg_mlp_destroy(pdhg_mlp* mlp);
int pdhg_mlp_param_count(const pdhg_mlp* mlp);
int pdhg_mlp_get_params(pdhg_mlp* mlp, double* params);
int pdhg_mlp_set_params(pdhg_mlp* mlp, const double* params);
int pdhg_mlp_forward(pdhg_mlp* mlp, const double* X, int N, double* Y);
int pdhg_mlp_loss_grad(pdhg_mlp* mlp, const double* X, const double* T, int N, double* loss, double* grad);
/* `steps` full-batch Adam steps (beta 0.9 / 0.999); losses[steps] may be NULL. */
int pdhg_mlp_train(pdhg_mlp* mlp, const double* X, const double* T, int N, int steps, double lr, double* losses);

/* ---- synthetic LP generators (owned models) ---------------------------- */
typedef struct pdhg_model pdhg_model;
/* Random LP with a known optimum: x* half zeros / half U(0,1), s*
 * complementary U(0,1), y* ~ N(0,1), b = A x* (equality rows), c = A'y* + s*,
 * x >= 0.  row_sigma == 0: every column gets nnz_per_col entries at uniform
 * rows (duplicates merged).  row_sigma > 0: row lengths are log-normal(sigma)
 * scaled to mean nnz_per_col * n / m, columns uniform within each row. */
pdhg_model* pdhg_gen_random_lp(uint64_t seed, int m, int n, double nnz_per_col, double row_sigma);
/* Refinery-style multi-period blending LP (block staircase, inventory
 * linking rows), every column bounded, rows and columns scaled by
 * 10^U(-scale_decades, scale_decades). */
pdhg_model* pdhg_gen_refinery(uint64_t seed, int periods, int streams, int crudes, int products, int qualities,
                              double scale_decades);
/* Price family of a refinery model: cost_k = c0 + sum_q theta_kq d_q, where
 * d_q is crude q's base purchase price on its purchase columns (q < 4) and
 * theta_kq ~ U[-range, range].  thetas (k x 4) and costs (k x n) are written. */
int pdhg_gen_family(const pdhg_model* base, uint64_t seed, int k, double range, double* thetas, double* costs);
/* Same family map for given thetas (k x 4). */
int pdhg_family_costs(const pdhg_model* base, int k, const double* thetas, double* costs);
/* The reference's `calibrate` LP (src/parametric/parametric.cpp:543-563):
 * m = s rows >= 1, n = 2s columns in [0, 10], 4 entries 0.5 + U per column,
 * cost U - 0.3.  Used to fit the CPU/GPU crossover. */
pdhg_model* pdhg_gen_calibration(uint64_t seed, int s);
/* Reads an MPS file (free format, or fixed columns once names contain
 * blanks; plain or gzip) into a model in minimisation form: the LP part of
 * the reference's readMps (src/io/mps.h:15-19) followed by toMinimization
 * (src/core/model.cpp:92-101).  *sense (may be NULL) is -1 for an OBJSENSE
 * MAX file, whose costs and offset come back negated, else 1.  Quadratic
 * sections and semi-continuous bounds are rejected (NULL). */
pdhg_model* pdhg_read_mps(const char* path, int* sense);
/* Borrowed view of a model; x_star / y_star may be NULL (refinery). */
int pdhg_model_view(const pdhg_model* mdl, pdhg_lp* lp, const double** x_star, const double** y_star);
void pdhg_model_free(pdhg_model* mdl);

#ifdef __cplusplus
}
#endif

#endif /* PDHG_HIP_H */

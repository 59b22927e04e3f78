/*
 * cosmicblup_amd — C-ABI of the MI355X (gfx950) GBLUP/ssGBLUP hot path.
 *
 * Plain pointers and sizes only. Every call is synchronous with respect to
 * the host unless stated otherwise, returns CB_OK (0) or a negative status,
 * and leaves a message for cb_last_error() (thread-local) on failure.
 * Host arrays are caller-owned; device memory is owned by the handles.
 * Dense matrices crossing the boundary are row-major with a leading
 * dimension; the symmetric ones (G, G^-1) are identical in Eigen's
 * column-major order.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   PlinkReader::load / openBed           src/core/plink_reader.cpp:27-91
 *   GrmBuilder::compute(PlinkReader)      src/core/grm_builder.cpp:63-405
 *   GrmBuilder getters                    src/core/grm_builder.h:55-64
 *   MMELHSBuilder / PCGSolver / MatrixFreePCG / AI_REML (see each entry)
 */
#ifndef COSMICBLUP_AMD_H
#define COSMICBLUP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CB_OK 0
#define CB_ERR_INVALID -1       /* bad argument / malformed input          */
#define CB_ERR_OOM -2           /* device or host allocation failed         */
#define CB_ERR_DEVICE -3        /* HIP runtime error                        */
#define CB_ERR_NOT_CONVERGED -4 /* iterative solver hit its iteration cap   */
#define CB_ERR_INTERNAL -5
#define CB_ERR_IO -6            /* a file could not be opened, read or written */

typedef struct cb_context cb_context;
typedef struct cb_genotypes cb_genotypes;
typedef struct cb_grm cb_grm;

/* ---- context ----------------------------------------------------------- */
/* One context per host thread; owns a HIP stream (default priority) on
 * `device`. No reference counterpart: the reference is single-process CPU
 * code with global OpenMP state (src/app/runtime.cpp:105-117). */
int cb_context_create(int device, cb_context** out);
void cb_context_destroy(cb_context* ctx);
/* Message of the last failed call on this host thread ("" if none). */
const char* cb_last_error(void);
/* Use a caller-owned hipStream_t (e.g. torch's current stream) instead of
 * the context's own; NULL restores the context's stream. */
int cb_context_set_stream(cb_context* ctx, void* hip_stream);
void* cb_context_stream(cb_context* ctx);
int cb_context_synchronize(cb_context* ctx);
/* Per-kernel timing with HIP events recorded on the launch stream. */
int cb_context_set_profiling(cb_context* ctx, int enabled);
/* Replay the fixed launch sequence of PCG iterations as HIP graphs (used
 * only on a non-null stream that the caller is not capturing). Results are
 * identical to individual launches. Default off: on MI355X the per-kernel
 * cost of these iterations is device-side, so a replay saves nothing and
 * the per-solve capture costs about 0.5 ms. No reference counterpart
 * (launch mechanics). */
int cb_context_set_graphs(cb_context* ctx, int enabled);
int cb_context_reset_profile(cb_context* ctx);
/* Total milliseconds and launch count booked under `kernel` since the last
 * reset (0 / 0 when the name is unknown). */
int cb_context_kernel_time(cb_context* ctx, const char* kernel, double* total_ms,
                           int64_t* launches);

/* ---- genotypes (PLINK BED) -------------------------------------------- */
/* Replaces PlinkReader::openBed (src/core/plink_reader.cpp:66-91): `bed` is
 * the whole .bed file (3-byte magic 6C 1B 01, SNP-major, ceil(n/4) bytes per
 * SNP, sample i at byte i>>2, bits 2(i&3)). n and m come from the .fam and
 * .bim line counts (plink_reader.cpp:33-64). Errors: wrong magic,
 * individual-major mode, empty dims (reference: runtime_error), and a file
 * shorter than 3 + m*ceil(n/4) bytes (the reference only warns, then reads
 * past the mapping). The bytes are copied to device memory. */
int cb_genotypes_from_bed(cb_context* ctx, const uint8_t* bed, size_t bed_size,
                          int64_t n_samples, int64_t n_snps, cb_genotypes** out);
/* Same, from SNP rows already resident on `ctx`'s device (no magic):
 * row j starts at d_rows + j*row_pitch; row_pitch >= ceil(n/4). Copies on
 * the context's stream and returns when the copy is done; work on other
 * streams that writes d_rows must be finished before the call. */
int cb_genotypes_from_device_rows(cb_context* ctx, const void* d_rows, size_t row_pitch,
                                  int64_t n_samples, int64_t n_snps, cb_genotypes** out);
void cb_genotypes_destroy(cb_genotypes* g);
int cb_genotypes_dims(const cb_genotypes* g, int64_t* n_samples, int64_t* n_snps);

/* ---- GRM (VanRaden-1) -------------------------------------------------- */
/* Mirrors GrmBuilder::Options (src/core/grm_builder.h:14-29) for the fields
 * that reach the PLINK VanRaden path. algorithm must be 1, min_hwe 0
 * (HWE filtering is off in the reference default). min_maf below 2^-7 is
 * rejected (the integer centring needs 2p >= 2^-6; the reference GRM path
 * always uses 0.01f, relationship_builder.cpp:206-212). block_size only
 * shapes the order in which the float 2pq values are summed, exactly as
 * the reference's SNP blocks do. */
typedef struct {
    int64_t block_size; /* default 2048 */
    float min_maf;      /* default 0.01f */
    float min_hwe;      /* must be 0 */
    int algorithm;      /* must be 1 (VanRaden) */
} cb_grm_options;

void cb_grm_options_default(cb_grm_options* opts);

/* Output selection for cb_grm_compute. */
#define CB_GRM_OUT_F64 1   /* FP64 G, n x n                                    */
#define CB_GRM_OUT_F32 2   /* FP32 copy of G, n x n                            */
#define CB_GRM_OUT_INT 4   /* raw integer Gram X X^T (int32), for exactness    */

int cb_grm_create(cb_context* ctx, const cb_grm_options* opts, cb_grm** out);
void cb_grm_destroy(cb_grm* grm);
/* GrmBuilder::compute(const PlinkReader&) (grm_builder.cpp:63-405): per-SNP
 * statistics and MAF filter, the integer Gram on INT8 MFMA with INT32
 * accumulation, FP64 centring/missing-call corrections, division by
 * sum(2pq). Results stay on the device until copied out. */
int cb_grm_compute(cb_grm* grm, const cb_genotypes* g, int outputs);
/* getGrm() (grm_builder.h:55): copy G (n x n, leading dimension ld). */
int cb_grm_copy_grm(const cb_grm* grm, double* out, int64_t ld);
int cb_grm_copy_grm_f32(const cb_grm* grm, float* out, int64_t ld);
int cb_grm_copy_int_gram(const cb_grm* grm, int32_t* out, int64_t ld);
/* FP64 G into caller-owned device memory on the same device (e.g. to form
 * G^-1 there); leading dimension ld. */
int cb_grm_copy_grm_device(const cb_grm* grm, void* dst, int64_t ld);
/* Device pointer to an output (CB_GRM_OUT_*) and its leading dimension;
 * valid until the next compute or destroy. */
int cb_grm_device_ptr(const cb_grm* grm, int which, void** ptr, int64_t* ld);
/* The matrix part of write_named_matrix (src/io/relationship_builder.cpp:
 * 43-53) for the FP64 G, streamed from device memory: format
 * CB_MATRIX_TXT writes write_matrix_txt's "i<TAB>j<TAB>G(i,j)" lines for
 * j <= i, 1-based, 6 significant digits (src/core/matrix_io.cpp:464-472);
 * CB_MATRIX_BIN writes COSMIC_UPPER_PACKED (write_upper_packed_binary,
 * src/core/upper_packed_binary.cpp:54-69). The caller writes the .id file.
 * CB_ERR_IO when the file cannot be written. */
#define CB_MATRIX_TXT 0
#define CB_MATRIX_BIN 1
int cb_grm_write(const cb_grm* grm, const char* path, int format);
/* getSum2pq(), getTotalSnps(), getAlleleFreqs() (grm_builder.h:56-61):
 * allele frequencies are (float)p for SNPs that passed the filter and -1
 * otherwise (grm_builder.cpp:111,242). */
int cb_grm_sum2pq(const cb_grm* grm, double* out);
int cb_grm_total_snps(const cb_grm* grm, int64_t* out);
int cb_grm_num_samples(const cb_grm* grm, int64_t* out);
int cb_grm_copy_allele_freqs(const cb_grm* grm, float* out, int64_t m);
/* Per-SNP statistics of grm_builder.cpp:197-246 for parity checks:
 * counts (n_obs, sum_x, n_hom1, n_het, n_hom2), pass flag, (float)(2p),
 * (float)(2p(1-p)); any pointer may be NULL. */
int cb_grm_copy_snp_stats(const cb_grm* grm, int64_t m, int32_t* n_obs, int32_t* sum_x,
                          int32_t* n_hom1, int32_t* n_het, int32_t* n_hom2, uint8_t* valid,
                          float* mean_f, float* twopq_f);


/* ---- mixed-model equations -------------------------------------------- */
typedef struct cb_mme cb_mme;

/* MMELHSBuilder(recs, fd, comps) (src/core/mme_builder.cpp:72-257) for one
 * random component: n records; record r has the fixed-effect row
 * X[r,:] as CSR (x_ptr[n+1], x_col, x_val over p columns — FixedDesignG.rows,
 * src/core/design.cpp:93-108), random level aid[r] in [0,q) or -1 for none
 * (GenRecord.aid - 1, mme_builder.cpp:137) and phenotype y[r]. The LHS
 * W'W (W = [X Z]) and RHS W'y are assembled once on the host and uploaded. */
int cb_mme_create(cb_context* ctx, int64_t n, int64_t p, int64_t q, const int64_t* x_ptr,
                  const int32_t* x_col, const double* x_val, const int32_t* aid, const double* y,
                  cb_mme** out);
void cb_mme_destroy(cb_mme* mme);
int cb_mme_dims(const cb_mme* mme, int64_t* n, int64_t* p, int64_t* q);
/* The relationship inverse Q of RandomComponent (mme_builder.h:10-21):
 * dense q x q row-major (DenseMatrixAdapter, src/core/matrix_adapter.h:116-153),
 * from host or device memory, or CSR with full symmetric storage
 * (SparseMatrixAdapter, matrix_adapter.h:56-113). Kept in FP64 and FP32. */
int cb_mme_set_qinv_dense(cb_mme* mme, const double* Q, int64_t ld);
int cb_mme_set_qinv_dense_device(cb_mme* mme, const double* dQ, int64_t ld);
int cb_mme_set_qinv_csr(cb_mme* mme, const int64_t* ptr, const int32_t* col, const double* val);
/* ssGBLUP H^-1 as SplitMatrixAdapter (matrix_adapter.h:232-373) holds it:
 * Q x = A^-1 x + scatter(D gather(x)) with A^-1 CSR (as above) and
 * D = tau G^-1 - omega A22^-1 one dense ng x ng block in device memory
 * (leading dimension ldd, e.g. from cb_hinv_blocks); gmap[i] is the row of
 * Q of genotyped row i (-1: not in the pedigree), entries distinct. */
int cb_mme_set_qinv_split(cb_mme* mme, const int64_t* ptr, const int32_t* col, const double* val, int64_t ng,
                          const int32_t* gmap, const double* dD, int64_t ldd);
/* build_rhs (mme_builder.cpp:321-328): [X'y; Z'y], length p+q. */
int cb_mme_rhs(const cb_mme* mme, double* rhs);
/* Y = C(lambda) X, C = W'W + lambda blockdiag(0, Q), FP64; X, Y column-major (p+q) x r. */
int cb_mme_apply(cb_mme* mme, double lambda, int64_t r, const double* X, double* Y);

#define CB_PREC_FP64 0   /* FP64 operator and vectors ("plain") */
#define CB_PREC_MIXED 1  /* FP32 block PCG inside FP64 residual refinement */
/* As MIXED, with the dense Q X of the FP32 recursion (r > 2 columns) on FP16
   matrix cores: Q and X as hi + lo FP16 halves of power-of-two scaled values
   (22 significant bits), three products with FP32 accumulation. No
   counterpart in the reference (its operator is FP64, implicit_mme.h). */
#define CB_PREC_MIXED_F16 2
#define CB_PRECOND_JACOBI 0        /* JacobiPreconditioner, pcg_solver.h:35-57 */
#define CB_PRECOND_BLOCK_JACOBI 1  /* BlockJacobiPreconditioner, pcg_solver.h:60-128 */
#define CB_STOP_RELATIVE 0  /* ||r|| / ||b|| < tol, PCGSolver (pcg_solver.h:501) */
#define CB_STOP_ABSOLUTE 1  /* ||r|| < tol, MatrixFreePCG (matrix_free_pcg.h:51) */

typedef struct {
    int precision;      /* CB_PREC_*; default MIXED */
    int precond;        /* CB_PRECOND_*; default BLOCK_JACOBI */
    int stop_rule;      /* CB_STOP_*; default RELATIVE */
    int batch_rhs;      /* FP64 only: 1 = all columns as one block, 0 = one at a time */
    double tol;         /* default 1e-10 (config.h:41) */
    int64_t max_iter;   /* default 10000 (config.h:42) */
    double stall_tol;   /* MIXED: the stop test is on the true FP64 residual b - C x,
                           which has a rounding floor near cond(C) * eps64. A column
                           whose best true residual has not halved over eight
                           residual replacements has stalled there: it is accepted
                           when its residual is below stall_tol (the call then
                           returns CB_OK with *rel_residual above tol), else it
                           iterates on and the solve returns CB_ERR_NOT_CONVERGED
                           at max_iter. Default 0 (no stall accepted);
                           cb_reml_options_default sets 1e-8 for the REML solves.
                           (The reference's tol tests the recursive residual,
                           pcg_solver.h:501.) */
} cb_solve_options;

void cb_solve_options_default(cb_solve_options* opts);
/* Solve C(lambda) X = B for r right-hand sides (column-major (p+q) x r host
 * arrays), each column its own CG recurrence and stop test. Replaces
 * PCGSolver::solve / MatrixFreePCG::solve as called by AI_REML::solve_lhs
 * (vce.cpp:1124-1236) and GSSolve (mixed_model_solver.cpp:243-255).
 * Returns CB_ERR_NOT_CONVERGED (X then holds the last iterate) when the
 * iteration cap is hit. */
int cb_mme_solve(cb_mme* mme, double lambda, const cb_solve_options* opts, int64_t r,
                 const double* B, double* X, int64_t* iterations, double* rel_residual);
/* cb_mme_solve with B and X in device memory on this context's device
 * ((p+q) x r column blocks, leading dimensions ldb, ldx >= p+q): for callers
 * that keep right-hand sides and solutions on the GPU (no host transfer).
 * The work that produced B must be finished before the call. */
int cb_mme_solve_device(cb_mme* mme, double lambda, const cb_solve_options* opts, int64_t r,
                        const double* dB, int64_t ldb, double* dX, int64_t ldx, int64_t* iterations,
                        double* rel_residual);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cb_mme_solve_device with its scratch (PCG vectors,
 * right-hand-side and solution blocks, dot and split-K partials) taken from
 * a fixed 64 MiB array on the device instead of the model's growing
 * buffers; it calls no hipMalloc or hipFree. Dense or CSR Qinv,
 * CB_PREC_FP64 or CB_PREC_MIXED; sizes whose scratch would exceed the array
 * (about r * (p + q) * 96 bytes plus partials) return CB_ERR_INVALID. The
 * array is shared by every call on one GPU and keeps the last call's
 * contents: calls that overlap on one GPU (several contexts or threads)
 * race. Results equal cb_mme_solve_device's. */
int cb_mme_solve_device_noalloc(cb_mme* mme, double lambda, const cb_solve_options* opts, int64_t r,
                                const double* dB, int64_t ldb, double* dX, int64_t ldx, int64_t* iterations,
                                double* rel_residual);
/* GSSolve's fixed-effect covariance and standard errors (calc_se, on by
 * default: config.h:45; mixed_model_solver.cpp:346-417): C(lambda) x = e_i
 * for each fixed effect i < p, solved as blocks of unit right-hand sides;
 * cinv_beta (optional, p x p, column i = rows 0..p of C^-1 e_i, i.e.
 * Cinv_beta) and beta_se (optional, p) = sqrt(max(0, var_e (C^-1)_ii)).
 * iterations (optional): PCG iterations of the slowest block. */
int cb_mme_beta_cov(cb_mme* mme, double lambda, double var_e, const cb_solve_options* opts, double* cinv_beta,
                    double* beta_se, int64_t* iterations);

/* ---- REML -------------------------------------------------------------- */
#define CB_VCE_AI 0  /* data-driven AI step, vce.cpp:2414-2857 */
#define CB_VCE_EM 1  /* EM + Hutchinson step, vce.cpp:1490-1574 */
#define CB_VCE_AI_EXACT 2  /* dense-exact AI step (run_dense_exact_step, vce.cpp:2859-3195):
                              exact traces from dense P = (I - W C^-1 W')/ve on the device,
                              <= 6 logL halvings; n, p + q <= 40000 */
/* The reference's own choice per iteration (AI_REML::run_ai_iteration,
 * vce.cpp:1334-1574, with the solver set-up of build_lhs, vce.cpp:1041-1110,
 * and estimate_trace_Qinv_Cuu, vce.cpp:1266-1315), from dim = p + q, the
 * kind of Q and q (cb_reml_branch reports it):
 *   dim < 5000                         dense-exact AI step (as CB_VCE_AI_EXACT);
 *   direct-solver sizes (dense or split Q with dim < 15000, CSR Q with
 *   q <= 100000 and dim < 200000)      AI step (CSR: from the first iteration;
 *                                      dense/split: EM first, AI after), the
 *                                      Newton step accepted only when logL
 *                                      improves (<= 15 halvings, vce.cpp:
 *                                      2797-2848; p + q <= 65536) else EM;
 *                                      traces exact when q <= 5000, else
 *                                      adaptive with ns = max(50, mc_samples);
 *   otherwise                          EM with adaptive traces, ns =
 *                                      max(5, mc_samples / 2).
 * Adaptive traces: Rademacher probes in batches of 16 until at least
 * max(30, ns/3) and SE/|mean| < 0.05, at most max(100, ns)
 * (AdaptiveTraceEstimator::estimate, adaptive_trace_estimator.cpp:11-83,
 * with the batch size of one thread). */
#define CB_VCE_AUTO 3

typedef struct {
    int algorithm;      /* CB_VCE_*; default AI */
    int64_t max_iter;   /* ai_maxit, default 20 (config.h:52) */
    double tol;         /* vce_tol on sum |d sigma^2|, default 1e-4 (config.h:57) */
    int64_t n_probes;   /* Rademacher probes for tr(Q C^uu), fixed for the run; default 64 */
    uint64_t seed;      /* probe stream: sign of splitmix64(seed + g*(c*q + j + 1)), g = 0x9E3779B97F4A7C15 */
    int trace_exact;    /* 1: probes are the q unit vectors (exact trace) */
    cb_solve_options solve;
    int64_t block_cols; /* right-hand sides per block solve (y + probes); 0 (default): as many as
                           fit in half the free device memory. A smaller block takes the probes in
                           chunks whose vectors and Q z are regenerated at each evaluation rather
                           than stored for the run; the estimator is unchanged (the reference
                           solves probes in batches too, adaptive_trace_estimator.cpp:11-83) */
    int64_t mc_samples; /* VCEConfig.mc_samples (config.h:56), default 100: sets the adaptive
                           probe counts (CB_VCE_AUTO, or adaptive = 1) */
    int adaptive;       /* 1: the adaptive trace rule (see CB_VCE_AUTO) with ns = max(5,
                           mc_samples / 2) for CB_VCE_AI / CB_VCE_EM instead of n_probes fixed
                           probes; default 0. CB_VCE_AUTO sets its own. */
} cb_reml_options;

void cb_reml_options_default(cb_reml_options* opts);
/* VCESolver::solve for the single-component GBLUP/PBLUP model
 * (VCESolver::create + AI_REML::solve, vce.cpp:741-837,3592-3611):
 * iterate from (var_u0, var_e0) until sum |d sigma^2| < tol or max_iter.
 * hist_vu / hist_ve (optional, length max_iter) receive each iterate;
 * solution (optional, p+q) receives the MME solution at the estimates
 * (getSolution()). ai_last (optional, 2x2 row-major) receives the projected
 * AI matrix of the last AI step (last_AI_mat; for an EM run, the AI matrix
 * at the estimates) and se (optional, 2) the standard errors of
 * (var_u, var_e) from it: sqrt(diag(AI^-1)) (calculate_SE, vce.cpp:3197-3368). */
int cb_reml_run(cb_mme* mme, const cb_reml_options* opts, double var_u0, double var_e0,
                double* var_u, double* var_e, int64_t* iterations, int* converged,
                double* hist_vu, double* hist_ve, double* solution, double* ai_last, double* se);
/* REML score dlogL/d(var_u, var_e) at the given variances from one block
 * solve [W'y | probes] (the data-driven gradient of vce.cpp:2647-2675),
 * optionally the AI matrix (2x2 row-major, vce.cpp:2596-2645), the trace
 * estimate of tr(Q C^uu) and the MME solution. */
int cb_reml_score(cb_mme* mme, const cb_reml_options* opts, double var_u, double var_e,
                  double* grad, double* ai, double* trace, double* solution);
/* REML log-likelihood calc_logL_internal (vce.cpp:3533-3588):
 * -1/2 (log|C| + (n-p-q) log var_e + q log var_u + (y'y - sol'rhs)/var_e)
 * with C = C(var_e/var_u) assembled densely on the device and factorised by
 * Cholesky (p + q <= 65536). A C that is not positive definite gives
 * -1e20, the reference's value for a failed determinant. logdet_c may be NULL. */
int cb_reml_logl(cb_mme* mme, double var_u, double var_e, double* logl, double* logdet_c);
/* The branch CB_VCE_AUTO takes for this model (VCESolver::create's size-based
 * choice): *first_step and *later_step are CB_VCE_AI_EXACT, CB_VCE_AI or
 * CB_VCE_EM for iteration 0 and the iterations after it; *trace_rule is 0
 * (fixed n_probes), 1 (adaptive) or 2 (exact: q unit probes); *logl_checked
 * is 1 when AI steps are accepted on a logL increase; *min_probes and
 * *max_probes are the adaptive rule's bounds. Any output may be NULL. */
int cb_reml_branch(const cb_mme* mme, const cb_reml_options* opts, int* first_step, int* later_step,
                   int* trace_rule, int* logl_checked, int64_t* min_probes, int64_t* max_probes);
/* The probe matrix the run uses (q x k, column-major, leading dimension ld). */
int cb_reml_copy_probes(cb_mme* mme, const cb_reml_options* opts, double* out, int64_t ld);

/* ---- relationship-matrix inverse --------------------------------------- */
/* invert_psd_matrix(M, ridge) (src/core/matrix_io.cpp:488-501): eigen-bent
 * inverse of a symmetric PSD n x n matrix, M + ridge I = U diag(l) U^T,
 * eps = max(1e-12 max(l), 1e-6), l <- max(l, eps), out = U diag(1/l) U^T
 * (G^-1 at relationship_builder.cpp:471). Host arrays (M, out: leading
 * dimensions ldm, ldo). min_eig / max_eig (may be NULL) receive the extreme
 * eigenvalues of M + ridge I before the floor. With both NULL the library
 * avoids the eigensolve where the floor allows: no eigenvalue below 1e-6
 * (inverse by Cholesky) or exactly one (that eigenpair by inverse iteration,
 * floored by a rank-1 update, then Cholesky), each certified by a Cholesky
 * factorisation, with the row sums bounding max(l) by 1e6; otherwise, and
 * whenever an eigenvalue is asked for, the symmetric eigensolve. */
int cb_invert_psd(cb_context* ctx, const double* M, int64_t ldm, int64_t n, double ridge, double* out,
                  int64_t ldo, double* min_eig, double* max_eig);
/* Same with M and out in device memory of ctx's device (they may alias). */
int cb_invert_psd_device(cb_context* ctx, const double* dM, int64_t ldm, int64_t n, double ridge, double* dout,
                         int64_t ldo, double* min_eig, double* max_eig);

/* ---- ssGBLUP H^-1 ------------------------------------------------------ */
typedef struct {
    int tune_g;     /* align G to A22 (config.h:120), default 1 */
    double alpha;   /* G <- (1 - alpha) G + alpha A22 (config.h:121), default 0.05 */
    double tau;     /* G^-1 scale (config.h:88), default 1 */
    double omega;   /* A22^-1 scale (config.h:89), default 1 */
    double ridge;   /* added to G and A22 before inversion (config.h:118), default 0 */
} cb_hinv_options;
void cb_hinv_options_default(cb_hinv_options* opts);
/* The split-mode H^-1 set-up of src/io/relationship_builder.cpp:481-633:
 * A22 of the genotyped animals (row gmap[i] of the pedigree for genotyped
 * row i, -1 if absent: identity row/column), tuneG a + b G matching A22's
 * diagonal and off-diagonal means, the blend, G^-1 by invert_psd_matrix,
 * A22^-1 by Cholesky, and D = tau G^-1 - omega A22^-1 (ng x ng, device,
 * leading dimension ldd). dam/sire as for cb_ainv_henderson (host arrays);
 * dG: the ng x ng GRM of the genotyped rows on the device (unchanged).
 * dA22 (optional, device) receives A22. stats (optional, 6) receives
 * a, b, mean diag A22, mean off-diag A22, mean diag G, mean off-diag G. */
int cb_hinv_blocks(cb_context* ctx, int64_t n, const int32_t* dam, const int32_t* sire, int64_t ng,
                   const int32_t* gmap, const double* dG, int64_t ldg, const cb_hinv_options* opts, double* dD,
                   int64_t ldd, double* dA22, int64_t lda22, double* stats);

/* ---- multi-trait ------------------------------------------------------- */
typedef struct {
    int64_t max_iter;  /* default 500 = max(500, 25 * ai_maxit) (variance_estimator.cpp:291) */
    double tol;        /* relative Frobenius change of Vg and Ve, default 1e-4 (vce_tol) */
} cb_mv_options;
void cb_mv_options_default(cb_mv_options* opts);
/* MvEigenVce::prepare + runNullModel with covariate projection
 * (src/core/eigen_vce.cpp:106-262, as variance_estimator.cpp:285-294 runs
 * it): complete cases only. Y: n x t (t <= 4) column-major host array, X:
 * n x c dense column-major host array (c < n - 1), dG: the n x n G on the
 * device (leading dimension ldg). Vg, Ve: t x t outputs (row-major). */
int cb_mv_vce(cb_context* ctx, int64_t n, int64_t t, int64_t c, const double* Y, const double* X, const double* dG,
              int64_t ldg, const cb_mv_options* opts, double* Vg, double* Ve, int64_t* iterations, int* converged);
/* GSSolve::solveMultiTrait (src/solve/mixed_model_solver.cpp:425-524): the
 * solution of the Kronecker MME [Ve^-1 (x) X'X, Ve^-1 (x) X'; .,
 * Ve^-1 (x) I + Vg^-1 (x) G^-1] (the reference factorises it by dense
 * LDLT). Here: canonical transformation of the t traits (complete records,
 * one design), t single-trait FP64 PCG solves with lambda_a = 1/d_a on the
 * device to relative residual 1e-12, transformed back. Vg and Ve (t x t,
 * t <= 4) must be positive definite. dGinv: n x n G^-1 on the device;
 * beta (c x t) and u (n x t) column-major. */
int cb_mv_solve(cb_context* ctx, int64_t n, int64_t t, int64_t c, const double* Y, const double* X,
                const double* dGinv, int64_t ldg, const double* Vg, const double* Ve, double* beta, double* u);

/* ---- matrix files ------------------------------------------------------ */
/* A host symmetric matrix (row-major, leading dimension lda) as
 * write_matrix_txt (CB_MATRIX_TXT, src/core/matrix_io.cpp:464-472) or
 * write_upper_packed_binary (CB_MATRIX_BIN, src/core/upper_packed_binary.cpp:
 * 54-69) writes it. */
int cb_matrix_write(const char* path, int64_t n, const double* A, int64_t lda, int format);
/* read_upper_packed_binary (upper_packed_binary.cpp:87-138): *n from the
 * header; when out is non-NULL the matrix is expanded symmetrically into it
 * (n x n, leading dimension ldo). The reference's checks and messages:
 * magic, metadata, dimension, truncation, trailing bytes (CB_ERR_IO). */
int cb_upper_packed_read(const char* path, int64_t* n, double* out, int64_t ldo);

/* ---- pedigree relationship inverse ----------------------------------- */
typedef struct cb_ainv cb_ainv;
/* genmatrix::invertA_henderson(dam, sire) (src/core/a_matrix.cpp:392-439)
 * with the inbreeding of ml() (a_matrix.cpp:127-186): Henderson's rules with
 * d_ii from the parents' F, the reference's duplicate-summing order,
 * symmetrisation, diagonal floor and prune (a_matrix.cpp:58-80). Parents are
 * 0-based row indices; an index outside [0, n) is an unknown parent
 * (complete_missing_parents_add_phantoms, a_matrix.cpp:43-52). When every
 * parent precedes its offspring the result is bit-identical to the
 * reference's. Otherwise (e.g. the first-appearance numbering of
 * relationship_builder.cpp:83-125 puts parents without a row of their own
 * after their first offspring) the pedigree is renumbered parents-first,
 * solved and mapped back: exact F where the reference's ml() scan, which
 * starts at max(s, d), misses ancestors numbered above it. A cycle or a
 * self-parent is CB_ERR_INVALID. Host code. */
int cb_ainv_henderson(int64_t n, const int32_t* dam, const int32_t* sire, cb_ainv** out);
int cb_ainv_nnz(const cb_ainv* a, int64_t* nnz);
/* CSR rows (ptr: n+1, col/val: nnz), columns ascending within a row. */
int cb_ainv_copy_csr(const cb_ainv* a, int64_t* ptr, int32_t* col, double* val);
/* Inbreeding coefficients F and Mendelian-sampling variances d_ii (either may be NULL). */
int cb_ainv_copy_inbreeding(const cb_ainv* a, double* F, double* dii);
void cb_ainv_destroy(cb_ainv* a);

/* ---- operator tuning -------------------------------------------------- */
/* No counterpart in the reference: the autotuner's interface
 * (cosmicblup_amd/tune.py). The dense operator Q X (k_dense_mfma) splits the
 * k range of Q over blocks; the split is looked up per (element size, q, r)
 * in a table the tuner generates (cb_dense_tuned.h).
 * cb_tune_dense_time: average ms of one Q X on synthetic q x q Q and r
 * columns (elem_bytes 2: the FP16 halves of CB_PREC_MIXED_F16, 4: FP32,
 * 8: FP64; q in [64, 200000], r in [3, 80])
 * with about target_blocks blocks and splits of at least min_kspan k, over
 * reps launches after one warm-up.
 * cb_tune_dense_lookup: the split the library uses for that shape. */
int cb_tune_dense_time(cb_context* ctx, int elem_bytes, int64_t q, int r, int target_blocks, int min_kspan, int reps,
                       double* ms);
int cb_tune_dense_lookup(int elem_bytes, int64_t q, int r, int* target_blocks, int* min_kspan);

#ifdef __cplusplus
}
#endif
#endif /* COSMICBLUP_AMD_H */

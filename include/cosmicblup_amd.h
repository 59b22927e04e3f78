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
 * row j starts at d_rows + j*row_pitch; row_pitch >= ceil(n/4). Copies. */
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
/* Device pointer to an output (CB_GRM_OUT_*) and its leading dimension;
 * valid until the next compute or destroy. */
int cb_grm_device_ptr(const cb_grm* grm, int which, void** ptr, int64_t* ld);
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

#ifdef __cplusplus
}
#endif
#endif /* COSMICBLUP_AMD_H */

/*
 * CPU oracle (test infrastructure, NOT product code) for the GRM part of the
 * GBLUP hot path: PLINK BED decoding, per-SNP statistics and filtering, the
 * integer Gram matrix, and the VanRaden-1 GRM.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as a checker / CPU baseline.
 *
 * Every function restates a reference routine; the citation is on each one.
 * Pinning (how this restatement is itself checked) is listed per function:
 *   orc_snp_stats      pinned: PLINK 1 code table on the decoded toy fixture,
 *                      closed-form counts of the synthetic writer's dosages,
 *                      brute force in numpy (tests/test_oracle_grm.py).
 *   orc_int_gram       pinned: brute-force numpy int64 X X^T.
 *   orc_grm_vanraden   pinned: brute-force numpy Z Z^T / sum(2pq); invariant
 *                      G 1 ~ 0 without missing calls; the identity
 *                      G = (XX^T - r1^T - 1r^T + c11^T)/S.
 *   orc_grm_block_raw,
 *   orc_grm_rows_raw   pinned: sub-blocks and row lists of S * G from
 *                      orc_grm_vanraden, whose pins are above, and a brute-
 *                      force numpy Z Z^T with missing calls, bit for bit on
 *                      the same block order (tests/test_oracle_grm.py).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifdef _OPENMP
#include <omp.h>
#endif


/* Code table of src/core/plink_reader.cpp:17-20: 00 -> 2, 01 -> missing,
 * 10 -> 1, 11 -> 0 (dosage of the .bim column-5 allele). */
static inline int code_at(const uint8_t *row, int64_t i) {
    return (row[i >> 2] >> (2 * (i & 3))) & 3;
}

/*
 * Per-SNP statistics and filter of src/core/grm_builder.cpp:175-263
 * (algorithm 1, one population, no SNP weights, min_hwe = 0).
 *   bed       : m rows of bytes_per_snp bytes (file content after the 3-byte
 *               magic), sample i at byte i>>2, bits 2(i&3).
 * Outputs (length m): counts, valid flag, and the float-rounded quantities
 * the reference keeps: freq_f = (float)p, mean_f = (float)(2p),
 * twopq_f = (float)(2p(1-p)).  Invalid SNPs get freq_f = -1
 * (allele_freqs.assign(m, -1.0f), grm_builder.cpp:111).
 */
void orc_snp_stats(const uint8_t *bed, int64_t n, int64_t m, int64_t bytes_per_snp,
                   float min_maf,
                   int64_t *n_obs, int64_t *sum_x, int64_t *n_hom1, int64_t *n_het,
                   int64_t *n_hom2, uint8_t *valid, float *freq_f, float *mean_f,
                   float *twopq_f) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < m; ++j) {
        const uint8_t *row = bed + j * bytes_per_snp;
        int64_t obs = 0, sx = 0, h1 = 0, het = 0, h2 = 0;
        for (int64_t i = 0; i < n; ++i) {
            int c = code_at(row, i);
            if (c == 1) continue;
            obs++;
            if (c == 0) { sx += 2; h1++; }
            else if (c == 2) { sx += 1; het++; }
            else { h2++; }
        }
        n_obs[j] = obs; sum_x[j] = sx; n_hom1[j] = h1; n_het[j] = het; n_hom2[j] = h2;
        valid[j] = 0; freq_f[j] = -1.0f; mean_f[j] = 0.0f; twopq_f[j] = 0.0f;
        if (obs == 0) continue;
        double freq = (double)sx / (2.0 * (double)obs);
        double maf = freq < 1.0 - freq ? freq : 1.0 - freq;
        if (maf < (double)min_maf || fabs(maf) < 1e-8) continue;
        /* weight (1.0f) * 2.0 * p * (1 - p), evaluated left to right */
        double w = 1.0;
        twopq_f[j] = (float)(((w * 2.0) * freq) * (1.0 - freq));
        freq_f[j] = (float)freq;
        mean_f[j] = (float)(freq * 2.0);
        valid[j] = 1;
    }
}

/* Dosage used by the integer Gram: missing calls count as 0 (they are
 * centred to 0 in the reference, grm_builder.cpp:309, which the FP64
 * corrections restore). Invalid SNPs contribute nothing. */
static inline int dosage(int c) { return c == 0 ? 2 : (c == 2 ? 1 : 0); }

/* Integer Gram X X^T over valid SNPs, full n x n, row-major int64.
 * The reference never forms it; it is the exact integer part of
 * Z Z^T (grm_builder.cpp:351) once centring is pulled out. */
void orc_int_gram(const uint8_t *bed, int64_t n, int64_t m, int64_t bytes_per_snp,
                  const uint8_t *valid, int64_t *out) {
    int8_t *x = (int8_t *)malloc((size_t)n * (size_t)m);
    for (int64_t j = 0; j < m; ++j) {
        const uint8_t *row = bed + j * bytes_per_snp;
        for (int64_t i = 0; i < n; ++i)
            x[i * m + j] = valid[j] ? (int8_t)dosage(code_at(row, i)) : 0;
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t a = 0; a < n; ++a) {
        for (int64_t b = 0; b <= a; ++b) {
            const int8_t *xa = x + a * m, *xb = x + b * m;
            int64_t s = 0;
            for (int64_t j = 0; j < m; ++j) s += (int64_t)xa[j] * xb[j];
            out[a * n + b] = s;
            out[b * n + a] = s;
        }
    }
    free(x);
}

/*
 * VanRaden-1 GRM exactly as src/core/grm_builder.cpp:63-404 forms it:
 * SNP blocks of block_size; per block the centred FP64 Z with
 * Z[i,c] = {2-mu, 0 (missing), 1-mu, -mu}[code], mu = (double)(float)(2p)
 * (:299-311); G += Z Z^T (:351); block_sum_2pq is a double sum of the float
 * 2pq values in SNP order (:272-278); total_sum_2pq += block_sum_2pq (:353);
 * finally G /= total_sum_2pq (:398).
 * G: full n x n row-major double (symmetric). Returns sum_2pq and the number
 * of valid SNPs through the pointers.
 */
void orc_grm_vanraden(const uint8_t *bed, int64_t n, int64_t m, int64_t bytes_per_snp,
                      int64_t block_size, float min_maf, double *G,
                      double *sum_2pq_out, int64_t *total_valid_out) {
    int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * 5 * (size_t)m);
    uint8_t *valid = (uint8_t *)malloc((size_t)m);
    float *freq_f = (float *)malloc(sizeof(float) * (size_t)m);
    float *mean_f = (float *)malloc(sizeof(float) * (size_t)m);
    float *twopq_f = (float *)malloc(sizeof(float) * (size_t)m);
    orc_snp_stats(bed, n, m, bytes_per_snp, min_maf, cnt, cnt + m, cnt + 2 * m,
                  cnt + 3 * m, cnt + 4 * m, valid, freq_f, mean_f, twopq_f);
    memset(G, 0, sizeof(double) * (size_t)n * (size_t)n);
    double total_sum_2pq = 0.0;
    int64_t total_valid = 0;
    double *Z = (double *)malloc(sizeof(double) * (size_t)n * (size_t)block_size);
    for (int64_t start = 0; start < m; start += block_size) {
        int64_t end = start + block_size < m ? start + block_size : m;
        int64_t nv = 0;
        double block_sum_2pq = 0.0;
        for (int64_t j = start; j < end; ++j) {
            if (!valid[j]) continue;
            block_sum_2pq += (double)twopq_f[j];
            double mu = (double)mean_f[j];
            double map[4] = {2.0 - mu, 0.0, 1.0 - mu, 0.0 - mu};
            const uint8_t *row = bed + j * bytes_per_snp;
            for (int64_t i = 0; i < n; ++i) Z[i * block_size + nv] = map[code_at(row, i)];
            nv++;
        }
        if (nv == 0) continue;
        /* G += Z Z^T (Z stored sample-major per block), lower triangle then mirrored */
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t a = 0; a < n; ++a) {
            const double *za = Z + a * block_size;
            for (int64_t b = 0; b <= a; ++b) {
                const double *zb = Z + b * block_size;
                double s = 0.0;
                for (int64_t c = 0; c < nv; ++c) s += za[c] * zb[c];
                G[a * n + b] += s;
            }
        }
        total_sum_2pq += block_sum_2pq;
        total_valid += nv;
    }
    for (int64_t a = 0; a < n; ++a)
        for (int64_t b = 0; b <= a; ++b) {
            double v = total_sum_2pq > 0 ? G[a * n + b] / total_sum_2pq : G[a * n + b];
            G[a * n + b] = v;
            G[b * n + a] = v;
        }
    *sum_2pq_out = total_sum_2pq;
    *total_valid_out = total_valid;
    free(Z); free(cnt); free(valid); free(freq_f); free(mean_f); free(twopq_f);
}

/*
 * Same GRM restricted to a rows x cols sub-block (rows [r0,r0+nr),
 * cols [c0,c0+nc)), used to time a bounded sample of a large GRM on the
 * host cores. Statistics use all n samples like the full build.
 * Gsub: nr x nc row-major, unnormalised Z Z^T sum (caller divides).
 */
void orc_grm_block_raw(const uint8_t *bed, int64_t n, int64_t m, int64_t bytes_per_snp,
                       const uint8_t *valid, const float *mean_f,
                       int64_t r0, int64_t nr, int64_t c0, int64_t nc,
                       int64_t block_size, double *Gsub) {
    memset(Gsub, 0, sizeof(double) * (size_t)nr * (size_t)nc);
    double *Zr = (double *)malloc(sizeof(double) * (size_t)nr * (size_t)block_size);
    double *Zc = (double *)malloc(sizeof(double) * (size_t)nc * (size_t)block_size);
    for (int64_t start = 0; start < m; start += block_size) {
        int64_t end = start + block_size < m ? start + block_size : m;
        int64_t nv = 0;
        for (int64_t j = start; j < end; ++j) {
            if (!valid[j]) continue;
            double mu = (double)mean_f[j];
            double map[4] = {2.0 - mu, 0.0, 1.0 - mu, 0.0 - mu};
            const uint8_t *row = bed + j * bytes_per_snp;
            for (int64_t i = 0; i < nr; ++i) Zr[i * block_size + nv] = map[code_at(row, r0 + i)];
            for (int64_t i = 0; i < nc; ++i) Zc[i * block_size + nv] = map[code_at(row, c0 + i)];
            nv++;
        }
#pragma omp parallel for schedule(static)
        for (int64_t a = 0; a < nr; ++a) {
            const double *za = Zr + a * block_size;
            for (int64_t b = 0; b < nc; ++b) {
                const double *zb = Zc + b * block_size;
                double s = 0.0;
                for (int64_t c = 0; c < nv; ++c) s += za[c] * zb[c];
                Gsub[a * nc + b] += s;
            }
        }
    }
    free(Zr); free(Zc);
}

/*
 * The same unnormalised Z Z^T sum for an explicit list of rows against the
 * column range [c0, c0+nc): out[a*nc + b] = sum_j Z[rows[a], j] Z[c0+b, j],
 * SNP blocks in the reference's order. The column panel of each SNP block
 * is filled in parallel, so a few rows of a 50,000-sample G take seconds.
 */
void orc_grm_rows_raw(const uint8_t *bed, int64_t n, int64_t m, int64_t bytes_per_snp,
                      const uint8_t *valid, const float *mean_f, const int64_t *rows, int64_t nr,
                      int64_t c0, int64_t nc, int64_t block_size, double *out) {
    (void)n;
    memset(out, 0, sizeof(double) * (size_t)nr * (size_t)nc);
    double *Zr = (double *)malloc(sizeof(double) * (size_t)nr * (size_t)block_size);
    double *Zc = (double *)malloc(sizeof(double) * (size_t)nc * (size_t)block_size);
    int64_t *cols = (int64_t *)malloc(sizeof(int64_t) * (size_t)block_size);
    for (int64_t start = 0; start < m; start += block_size) {
        int64_t end = start + block_size < m ? start + block_size : m;
        int64_t nv = 0;
        for (int64_t j = start; j < end; ++j)
            if (valid[j]) cols[nv++] = j;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nc; ++i)
            for (int64_t c = 0; c < nv; ++c) {
                const int64_t j = cols[c];
                const double mu = (double)mean_f[j];
                const double map[4] = {2.0 - mu, 0.0, 1.0 - mu, 0.0 - mu};
                Zc[i * block_size + c] = map[code_at(bed + j * bytes_per_snp, c0 + i)];
            }
        for (int64_t a = 0; a < nr; ++a)
            for (int64_t c = 0; c < nv; ++c) {
                const int64_t j = cols[c];
                const double mu = (double)mean_f[j];
                const double map[4] = {2.0 - mu, 0.0, 1.0 - mu, 0.0 - mu};
                Zr[a * block_size + c] = map[code_at(bed + j * bytes_per_snp, rows[a])];
            }
#pragma omp parallel for schedule(static)
        for (int64_t b = 0; b < nc; ++b) {
            const double *zb = Zc + b * block_size;
            for (int64_t a = 0; a < nr; ++a) {
                const double *za = Zr + a * block_size;
                double s = 0.0;
                for (int64_t c = 0; c < nv; ++c) s += za[c] * zb[c];
                out[a * nc + b] += s;
            }
        }
    }
    free(Zr); free(Zc); free(cols);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// VanRaden-1 GRM on gfx950 — the MI355X counterpart of
// GrmBuilder::compute(const PlinkReader&) (src/core/grm_builder.cpp:63-405).
//
// Algebra. With X the 0/1/2 dosage (missing calls and filtered SNPs set to
// 0), O the observed mask and mu_j = (double)(float)(2 p_j), the reference
// forms Z = O .* (X - 1 mu^T) and G = Z Z^T / S. Expanding:
//   Z Z^T(i,k) = X X^T(i,k) - u_i - u_k + C + E(i,k) + E(k,i)
//   u_i   = sum_j mu_j X_ij + sum_{j miss i} mu_j^2        (per sample)
//   C     = sum_j mu_j^2                                   (scalar)
//   E(k,i)= sum_{j miss k} mu_j X_ij + F(k,i)/2,  F(k,i) = sum_{j miss i and k} mu_j^2
// X X^T is an exact integer product (INT8 MFMA, INT32 accumulation). Every
// mu_j is a multiple of 2^-29 (2p >= 2^-6 after the MAF filter), so the
// mu-weighted dosage sums are computed exactly as integers
// K_j = mu_j * 2^29 (u_i's first term, and E via dot4 byte chunks of K_j).
// Without missing calls E = 0 and the correction is rank-1.
//
// Kernels:
//   k_snp_stats   per-SNP counts + filter (grm_builder.cpp:175-263), HBM-bound
//   k_pack        SNP-major BED -> sample-major 2-bit dosage tiles, HBM-bound
//   k_row_sums    exact integer sum_j K_j X_ij per sample
//   k_miss_count, k_miss_scan, k_miss_emit  per-sample missing-call lists,
//                 no atomics; k_miss_mfma  E with F folded in (INT8 MFMA)
//   k_gram        INT8 MFMA Gram over lower-triangle 256x256 tiles with the
//                 FP64 centring/scaling fused into the epilogue, MFMA-bound
#include <algorithm>
#include <cmath>
#include <numeric>

#include "cb_genotypes.h"
#include "cb_internal.h"

using namespace cbamd;

namespace {

constexpr int TILE = 256;        // samples per Gram tile side
constexpr int STAGE_SNPS = 256;  // SNPs per LDS stage (16 packed words)
constexpr int LDS_PITCH = 20;    // words per LDS row (16 + 4 pad: conflict-free b128 reads)
constexpr double TWO_M29 = 1.0 / 536870912.0;  // 2^-29

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// 2-bit-field mask (low bit of each field) for the samples of word w that are < n.
__device__ __forceinline__ uint32_t field_mask(int64_t n, int64_t w) {
    const int64_t rem = n - w * 16;
    if (rem >= 16) return 0x55555555u;
    if (rem <= 0) return 0u;
    return 0x55555555u & ((1u << (2 * rem)) - 1u);
}

// PLINK codes (plink_reader.cpp:17-20) -> dosage in each 2-bit field:
// 00 -> 2, 01 (missing) -> 0, 10 -> 1, 11 -> 0.
__device__ __forceinline__ uint32_t dosage_swar(uint32_t x) {
    const uint32_t lo = x & 0x55555555u;
    const uint32_t hi = (x >> 1) & 0x55555555u;
    return (hi & ~lo) | ((~(hi | lo) & 0x55555555u) << 1);
}

// ---------------------------------------------------------------------------
// Per-SNP statistics: one wave per SNP, popcounts over 16-sample words.
__global__ __launch_bounds__(256) void k_snp_stats(
    const uint8_t* __restrict__ bed, int64_t pitch, int64_t n, int64_t m, double min_maf,
    int32_t* __restrict__ o_nobs, int32_t* __restrict__ o_sumx, int32_t* __restrict__ o_h1,
    int32_t* __restrict__ o_het, int32_t* __restrict__ o_h2, uint8_t* __restrict__ o_valid,
    float* __restrict__ o_freq, float* __restrict__ o_mean, float* __restrict__ o_2pq,
    uint32_t* __restrict__ o_kfix) {
#pragma clang fp contract(off)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + wave;
    if (j >= m) return;
    const uint32_t* row = reinterpret_cast<const uint32_t*>(bed + j * pitch);
    const int64_t nwords = (n + 15) / 16;
    uint32_t c_miss = 0, c_h1 = 0, c_het = 0, c_h2 = 0;
    for (int64_t w = lane; w < nwords; w += 64) {
        const uint32_t x = row[w];
        const uint32_t lo = x & 0x55555555u, hi = (x >> 1) & 0x55555555u;
        const uint32_t vm = field_mask(n, w);
        c_h1 += __popc(~hi & ~lo & vm);
        c_miss += __popc(~hi & lo & vm);
        c_het += __popc(hi & ~lo & vm);
        c_h2 += __popc(hi & lo & vm);
    }
    for (int off = 32; off > 0; off >>= 1) {
        c_miss += __shfl_down(c_miss, off);
        c_h1 += __shfl_down(c_h1, off);
        c_het += __shfl_down(c_het, off);
        c_h2 += __shfl_down(c_h2, off);
    }
    if (lane != 0) return;
    const int64_t obs = n - (int64_t)c_miss;
    const int64_t sx = 2 * (int64_t)c_h1 + (int64_t)c_het;
    o_nobs[j] = (int32_t)obs;
    o_sumx[j] = (int32_t)sx;
    o_h1[j] = (int32_t)c_h1;
    o_het[j] = (int32_t)c_het;
    o_h2[j] = (int32_t)c_h2;
    uint8_t valid = 0;
    float freq_f = -1.0f, mean_f = 0.0f, twopq_f = 0.0f;
    uint32_t kfix = 0;
    if (obs > 0) {
        // grm_builder.cpp:219-246, same operations in the same order.
        const double freq = (double)sx / (2.0 * (double)obs);
        const double maf = freq < 1.0 - freq ? freq : 1.0 - freq;
        if (!(maf < min_maf || fabs(maf) < 1e-8)) {
            twopq_f = (float)((2.0 * freq) * (1.0 - freq));
            freq_f = (float)freq;
            mean_f = (float)(freq * 2.0);
            kfix = (uint32_t)((double)mean_f * 536870912.0);
            valid = 1;
        }
    }
    o_valid[j] = valid;
    o_freq[j] = freq_f;
    o_mean[j] = mean_f;
    o_2pq[j] = twopq_f;
    o_kfix[j] = kfix;
}

// ---------------------------------------------------------------------------
// SNP-major BED -> sample-major 2-bit dosage, tiled [stage][sample][16 words]:
// word w of stage s for sample i holds SNPs 256s+16w+t at bits 2t.
// Filtered SNPs, missing calls, padded SNPs and padded samples are 0.
__global__ __launch_bounds__(256) void k_pack(const uint8_t* __restrict__ bed, int64_t pitch,
                                              int64_t n, int64_t m, int64_t n_pad,
                                              const uint8_t* __restrict__ valid,
                                              uint32_t* __restrict__ X2) {
    __shared__ uint32_t lds[256 * 17];
    const int t = threadIdx.x;
    const int64_t sb = blockIdx.x;  // 256-sample block
    const int64_t st = blockIdx.y;  // 256-SNP stage
    const int64_t j = st * STAGE_SNPS + t;
    const bool ok = j < m && valid[j] != 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t byte0 = sb * 64 + q * 16;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (ok && byte0 < pitch) v = *reinterpret_cast<const uint4*>(bed + j * pitch + byte0);
        const int64_t w0 = sb * 16 + q * 4;
        uint32_t f[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t vm = field_mask(n, w0 + e);
            // (a zero byte is code 00 = dosage 2, so rows that are not kept are zeroed here)
            lds[t * 17 + q * 4 + e] = ok ? dosage_swar(f[e]) & (vm | (vm << 1)) : 0u;
        }
    }
    __syncthreads();
    const int sw = t >> 4, sf = t & 15;
    uint32_t outw[16];
#pragma unroll
    for (int w = 0; w < 16; ++w) {
        uint32_t acc = 0;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const uint32_t x = lds[(16 * w + s) * 17 + sw];
            acc |= ((x >> (2 * sf)) & 3u) << (2 * s);
        }
        outw[w] = acc;
    }
    uint4* dst = reinterpret_cast<uint4*>(X2 + ((st * n_pad) + sb * TILE + t) * 16);
    dst[0] = make_uint4(outw[0], outw[1], outw[2], outw[3]);
    dst[1] = make_uint4(outw[4], outw[5], outw[6], outw[7]);
    dst[2] = make_uint4(outw[8], outw[9], outw[10], outw[11]);
    dst[3] = make_uint4(outw[12], outw[13], outw[14], outw[15]);
}

// ---------------------------------------------------------------------------
// r_i = sum_j mu_j X_ij, exact: integer sum of K_j X_ij times 2^-29.
__global__ __launch_bounds__(256) void k_row_sums(const uint32_t* __restrict__ X2, int64_t n,
                                                  int64_t n_pad, int64_t n_stages,
                                                  const uint32_t* __restrict__ kfix,
                                                  double* __restrict__ r_out) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + wave;
    if (i >= n) return;
    uint64_t acc = 0;
    const int64_t nq = n_stages * 16;
    for (int64_t q = lane; q < nq; q += 64) {
        const int64_t s = q >> 4, w = q & 15;
        const uint32_t x = X2[(s * n_pad + i) * 16 + w];
        if (x == 0) continue;
        const uint32_t* kk = kfix + s * STAGE_SNPS + w * 16;
#pragma unroll
        for (int t = 0; t < 16; ++t) acc += (uint64_t)(kk[t] * ((x >> (2 * t)) & 3u));
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if (lane == 0) r_out[i] = (double)acc * TWO_M29;
}

// ---------------------------------------------------------------------------
// Missing-call lists (per sample: the SNP j and K_j of every missing call at
// a kept SNP), built without atomics. The SNPs are cut into chunks of
// MC_ROWS; thread (word column w, chunk c) reads word w of each BED row of
// the chunk, the 2-bit codes of samples 16 w .. 16 w + 15, and keeps its 16
// per-sample counters in LDS (private to the thread). Pass 1 writes the
// counts of each (chunk, sample); a scan over the chunks turns them into
// offsets; pass 2 repeats the reads and writes each entry at its sample's
// offset. A sample's list comes out in SNP order. Pads (to a multiple of 8)
// keep j = 0 and K = 0.
constexpr int MC_ROWS = 128;

// the missing-call (code 01) bit of each 2-bit field of word w of SNP j,
// zero for a dropped SNP and for samples >= n
__device__ __forceinline__ uint32_t missing_bits(const uint8_t* __restrict__ bed, int64_t pitch,
                                                 const uint8_t* __restrict__ valid, int64_t j, int64_t w,
                                                 uint32_t fmask) {
    const uint32_t x = valid[j] ? reinterpret_cast<const uint32_t*>(bed + j * pitch)[w] : 0u;
    return ~(x >> 1) & x & fmask;
}

// Pass 1: cnt[c n + i] = missing calls of sample i at the kept SNPs of chunk c.
__global__ __launch_bounds__(256) void k_miss_count(const uint8_t* __restrict__ bed, int64_t pitch, int64_t n,
                                                    int64_t m, const uint8_t* __restrict__ valid,
                                                    int32_t* __restrict__ cnt) {
    __shared__ int32_t lc[256 * 17];  // 16 counters per thread (pitch 17)
    const int tid = threadIdx.x;
    const int64_t nw = (n + 15) / 16, w = (int64_t)blockIdx.x * 256 + tid, c = blockIdx.y;
    if (w >= nw) return;  // the counters are private: no barrier follows
    int32_t* my = lc + tid * 17;
#pragma unroll
    for (int f = 0; f < 16; ++f) my[f] = 0;
    const uint32_t fm = field_mask(n, w);
    const int64_t j0 = c * MC_ROWS, j1 = min(m, j0 + MC_ROWS);
    for (int64_t jb = j0; jb < j1; jb += 8) {
        uint32_t mb[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) mb[t] = jb + t < j1 ? missing_bits(bed, pitch, valid, jb + t, w, fm) : 0u;
#pragma unroll
        for (int t = 0; t < 8; ++t)
            for (uint32_t b = mb[t]; b; b &= b - 1) ++my[(__ffs(b) - 1) >> 1];
    }
#pragma unroll
    for (int f = 0; f < 16; ++f)
        if (w * 16 + f < n) cnt[c * n + w * 16 + f] = my[f];
}

// Per sample (one thread): the counts of the nc chunks become exclusive
// offsets, in place, and the total goes to smp_cnt.
__global__ __launch_bounds__(256) void k_miss_scan(int32_t* __restrict__ cnt, int64_t n, int64_t nc,
                                                   int32_t* __restrict__ smp_cnt) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int32_t run = 0;
    for (int64_t c = 0; c < nc; ++c) {
        const int32_t t = cnt[c * n + i];
        cnt[c * n + i] = run;
        run += t;
    }
    smp_cnt[i] = run;
}

// Pass 2: the entries (j, K_j) of chunk c at smp_off[i] + cnt[c n + i] + ...
__global__ __launch_bounds__(256) void k_miss_emit(const uint8_t* __restrict__ bed, int64_t pitch, int64_t n,
                                                   int64_t m, const uint8_t* __restrict__ valid,
                                                   const uint32_t* __restrict__ kfix,
                                                   const int64_t* __restrict__ smp_off,
                                                   const int32_t* __restrict__ cnt,
                                                   unsigned long long* __restrict__ ljk) {
    __shared__ int64_t lp[256 * 17];  // 16 write positions per thread (pitch 17)
    const int tid = threadIdx.x;
    const int64_t nw = (n + 15) / 16, w = (int64_t)blockIdx.x * 256 + tid, c = blockIdx.y;
    if (w >= nw) return;  // the positions are private: no barrier follows
    int64_t* my = lp + tid * 17;
#pragma unroll
    for (int f = 0; f < 16; ++f) {
        const int64_t i = w * 16 + f;
        my[f] = i < n ? smp_off[i] + cnt[c * n + i] : 0;
    }
    const uint32_t fm = field_mask(n, w);
    const int64_t j0 = c * MC_ROWS, j1 = min(m, j0 + MC_ROWS);
    for (int64_t jb = j0; jb < j1; jb += 8) {
        uint32_t mb[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) mb[t] = jb + t < j1 ? missing_bits(bed, pitch, valid, jb + t, w, fm) : 0u;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (!mb[t]) continue;
            const int32_t j = (int32_t)(jb + t);
            const uint32_t K = kfix[j];
            for (uint32_t b = mb[t]; b; b &= b - 1) {
                const int64_t pos = my[(__ffs(b) - 1) >> 1]++;
                ljk[pos] = ((unsigned long long)K << 32) | (uint32_t)j;  // one 8-byte write per entry
            }
        }
    }
}

// s_i = sum over sample i's list of K^2 >> 6 (mu^2 in 2^-52 units) and
// k_i = sum of K, one wave per sample. A long list overflows 64 bits, so the
// low and high words of the K^2 terms are summed apart (exact for lists
// below 2^31 entries).
__global__ __launch_bounds__(256) void k_sample_sq(const int64_t* __restrict__ smp_off,
                                                   const unsigned long long* __restrict__ ljk, int64_t n,
                                                   unsigned long long* __restrict__ sq_lo,
                                                   unsigned long long* __restrict__ sq_hi,
                                                   unsigned long long* __restrict__ ksum) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + wave;
    if (i >= n) return;
    unsigned long long lo = 0, hi = 0, ks = 0;
    for (int64_t p = smp_off[i] + lane; p < smp_off[i + 1]; p += 64) {
        const unsigned long long K = ljk[p] >> 32;
        const unsigned long long k2 = (K * K) >> 6;
        lo += k2 & 0xffffffffull;
        hi += k2 >> 32;
        ks += K;
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo += __shfl_down(lo, off);
        hi += __shfl_down(hi, off);
        ks += __shfl_down(ks, off);
    }
    if (lane == 0) {
        sq_lo[i] = lo;
        sq_hi[i] = hi;
        ksum[i] = ks;
    }
}

// ---------------------------------------------------------------------------
// Missing-call correction on INT8 MFMA (k_miss_mfma).
//
// E(k, i) = e(k, i) + F(k, i) / 2 with
//   e(k, i) = sum_{j in miss(k)} mu_j X_ij     (X = 0 where i is missing)
//   F(k, i) = sum_{j in miss(k) and miss(i)} mu_j^2.
// With mu_j = K_j 2^-29 (K_j < 2^30 an integer), e is the exact integer
// sum_j K_j X_ij and F the exact integer sum_j K_j^2 (units 2^-58).
// The kernel forms, over a group of four samples k, the int8 matrix product
// e'(k, i) = sum_{j in miss(k)} K_j c_ij, where c_ij is the 2-bit code of the
// I-block rows: the dosage, or 3 for a missing call. So e' = e + 3 F1 with
// F1(k, i) = sum_{j in miss(k) and miss(i)} K_j, which is symmetric in
// (k, i). Only E(k, i) + E(i, k) is consumed, and it needs
//   e(k, i) + e(i, k) + F = e'(k, i) + e'(i, k) + D,
//   D(k, i) = sum_{j in miss(k) and miss(i)} (K_j^2 - 3 2^30 K_j)   (units 2^-58),
// so D, again symmetric, is added once per pair: in full to E(k, i) when i's
// I-block follows k's, halved to both when they share one, and not at all
// when it precedes. D is the int8 product of the digits of K^2 - 3 2^30 K
// with the missing-call indicators, formed only in the workgroups that add
// it.
// K is written in signed base-256 digits d_c in [-128, 127] (K = sum_c d_c
// 256^c: the bytes of (K + 0x80808080) ^ 0x80808080), K^2 - 3 2^30 K (a
// negative 64-bit integer) likewise with 8 digits. A has 16 rows = 4 samples
// x 4 digits; row (kl, c) holds d_c of the entries of sample kl and 0
// elsewhere. Each K-step of 64 list entries holds 16 entries of each sample
// of the group (k-block kl = sample kl), so a lane's A row is non-zero only in
// its own k-block, fixed for the whole kernel. B is 64 entries x 256 samples
// i from the entries' SNP rows; 64 entries per v_mfma_i32_16x16x64_i8. Row
// (kl, c) of the 16x16 output is element c of lane 16 kl + i's accumulator,
// so the digits combine in the lane: sum_c acc_c 256^c, exactly, in int64.
constexpr int MB_SAMPLES = 256;  // samples per I-block
constexpr int MB_STEP = 64;      // list entries per K-step
constexpr int MB_GROUP = 4;      // samples k per group (4 x 4 digit rows)
constexpr int MB_KBLK = 16;      // entries per k-block; a sample's list is padded to whole k-blocks
constexpr int MB_PF = 2;         // K-steps of row gathers in flight; group steps are padded to a multiple
constexpr int MB_CHUNK = 24;     // K-steps of one group staged in LDS at once (a multiple of MB_PF)
// LDS layout of the missing-call kernel, chosen so that its reads and
// writes hit distinct banks: a digit plane of one K-step takes 20 dwords
// (16 used), a K-step's 4 K planes 96 dwords and its 8 D planes 160; a
// gathered row takes 64 bytes, its 16-byte quarters swizzled by row
constexpr int MB_PLN = 20;
constexpr int MB_SPK = 96;
constexpr int MB_SPD = 160;
constexpr int MB_XCD = 8;        // XCDs; workgroups are dealt to them round-robin
// chunks between folds of the digit sums into E: an accumulator row (kl, c)
// sees 16 entries per K-step, N = 85 x 24 x 16 = 32 640 between folds, each
// adding at most 128 x 128, so the int32 digit sums stay below 2^30; and the
// values they encode, at most N 2^30 x 128 < 2^52 for the codes and
// N 2^31 x 64 < 2^52 for either half of D, are integers that the FP64
// recombination of their digits forms exactly
constexpr int MB_FOLD = 85;

// SNP rows in I-block-major order: block b (samples 256 b .. 256 b + 255)
// of SNP j at dwords (b m + j) 16 .. + 15, dosage 0/1/2 per 2-bit field,
// 3 for a missing call (PLINK 01), 0 for samples past n.
__global__ void k_encode_iblock(const uint8_t* __restrict__ bed, int64_t pitch, int64_t n, int64_t m, int64_t nb,
                                uint32_t* __restrict__ out) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // dword of the row
    if (w >= nb * 16) return;
    for (int64_t j = blockIdx.y; j < m; j += gridDim.y) {
        uint32_t v = 0;
        if (w * 4 < pitch) {
            const uint32_t x = reinterpret_cast<const uint32_t*>(bed + j * pitch)[w];
            const uint32_t vm = field_mask(n, w);
            const uint32_t miss = x & ~(x >> 1) & 0x55555555u;
            v = (dosage_swar(x) | miss | (miss << 1)) & (vm | (vm << 1));
        }
        out[((w >> 4) * m + j) * 16 + (w & 15)] = v;
    }
}

// The group lists: K-step t of group g holds entries 16 t .. 16 t + 15 of
// samples 4g .. 4g + 3 in its k-blocks 0 .. 3, pads (j = 0, K = 0) past the
// end of a sample's list; the group has as many K-steps as its longest
// list needs, rounded up to a multiple of MB_PF. gj = the 64 SNPs of each
// K-step, gK = their K. One block per group.
__global__ __launch_bounds__(256) void k_group_entries(const int64_t* __restrict__ smp_off,
                                                       const int32_t* __restrict__ smp_cnt,
                                                       const unsigned long long* __restrict__ ljk, int64_t n,
                                                       const int64_t* __restrict__ gstep, int32_t* __restrict__ gj,
                                                       uint32_t* __restrict__ gK) {
    const int64_t g = blockIdx.x;
    const int64_t base = gstep[g] * MB_STEP, len = (gstep[g + 1] - gstep[g]) * MB_STEP;
    for (int64_t e = threadIdx.x; e < len; e += blockDim.x) {
        const int kl = (int)((e >> 4) & 3);
        const int64_t r = (e >> 6) * MB_KBLK + (e & 15);  // entry r of sample kl's list
        const int64_t k = g * MB_GROUP + kl;
        int32_t j = 0;
        uint32_t K = 0;
        if (k < n && r < smp_cnt[k]) {
            const unsigned long long v = ljk[smp_off[k] + r];
            j = (int32_t)(uint32_t)v;
            K = (uint32_t)(v >> 32);
        }
        gj[base + e] = j;
        gK[base + e] = K;
    }
}

// 4 x 4 byte transpose: byte e of o[c] = byte c of d[e].
__device__ __forceinline__ void transpose4x4(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3, uint32_t (&o)[4]) {
    const uint32_t ta = __builtin_amdgcn_perm(d1, d0, 0x05010400u);
    const uint32_t tb = __builtin_amdgcn_perm(d1, d0, 0x07030602u);
    const uint32_t tc = __builtin_amdgcn_perm(d3, d2, 0x05010400u);
    const uint32_t td = __builtin_amdgcn_perm(d3, d2, 0x07030602u);
    o[0] = __builtin_amdgcn_perm(tc, ta, 0x05040100u);
    o[1] = __builtin_amdgcn_perm(tc, ta, 0x07060302u);
    o[2] = __builtin_amdgcn_perm(td, tb, 0x05040100u);
    o[3] = __builtin_amdgcn_perm(td, tb, 0x07060302u);
}

// Orders one wave's LDS accesses across its lanes (hardware keeps a wave's
// LDS operations in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Transposed 8-byte LDS read (ds_read_b64_tr_b8): within each group of 16
// lanes, lane i receives byte (i & 7) of the 8-byte rows addressed by lanes
// 2t + (i >> 3), t = 0 .. 7, as its byte t.
__device__ __forceinline__ uint2 lds_read_tr8(const void* p) {
    typedef int v2i_t __attribute__((ext_vector_type(2)));
    const v2i_t v = __builtin_amdgcn_ds_read_tr8_b64_v2i32(
        (__attribute__((address_space(3))) v2i_t*)(const_cast<void*>(p)));
    return make_uint2((uint32_t)v.x, (uint32_t)v.y);
}

// B fragments of one K-step, one per sample of the quad, from the quad's 16
// entry bytes x: the 2-bit codes c (dosage, 3 for a missing call) as c, 4c
// and 16c for samples 0-2 and, for sample 3, its field in place with the top
// bit flipped, which as a signed byte is 64 (c - 2) (one bitwise operation;
// the -128 per entry is put back from the sum of the entries' K); where D is
// added, the missing-call indicators (scaled by 1, 4, 16, 64).
__device__ __forceinline__ void code_fragments(const uint32_t (&x)[4], v4i (&bd)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bd[0][q] = (int)(x[q] & 0x03030303u);
        bd[1][q] = (int)(x[q] & 0x0c0c0c0cu);
        bd[2][q] = (int)(x[q] & 0x30303030u);
        bd[3][q] = (int)((x[q] ^ 0x80808080u) & 0xc0c0c0c0u);
    }
}
__device__ __forceinline__ void missing_fragments(const uint32_t (&x)[4], v4i (&bm)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t y = x[q] & (x[q] >> 1);
        bm[0][q] = (int)(y & 0x01010101u);
        bm[1][q] = (int)(y & 0x04040404u);
        bm[2][q] = (int)(y & 0x10101010u);
        bm[3][q] = (int)(y & 0x40404040u);  // an indicator at bit 6 (64) is still a positive int8
    }
}

// blockIdx.x = b P + p: I-block b (samples 256 b .. + 255) and group range p
// (groups gp[p] .. gp[p + 1] - 1, about equal in K-steps). With P a multiple
// of MB_XCD every XCD holds P / MB_XCD consecutive workgroups of each
// I-block, so the workgroups resident on an XCD gather from the same
// m x 64 bytes of row segments, which stay in its L2. 256 threads; per group:
//   stage    per chunk of MB_CHUNK K-steps: the SNPs, and the digit planes
//            of K (4) and, if D is added here, of K^2 - 3 2^30 K (8), formed
//            from gK (plane c, byte t = digit c of entry t); the lists stream
//            past L2 (non-temporal loads), which keeps the gathered rows;
//   gather   wave w loads the 64-byte rows of entries 16 w .. 16 w + 15, a
//            quarter row (16 bytes) per lane (MB_PF steps in flight), and
//            stores them untransposed in LDS (double-buffered, one barrier
//            per step; the quarter q of row r sits at q ^ ((r >> 1) & 3));
//   product  lane l of wave w takes quad Q = 16 w + l % 16 and k-block
//            kb = l / 16: two transposed reads (ds_read_b64_tr_b8) give its
//            16 entries' bytes (lane pairs 2t, 2t + 1 of a lane group address
//            the two 8-byte halves of the wave's quads in entry row
//            16 kb + 8h + t), from which the 4 samples' codes (and
//            missing-call indicators) are masked out, exact scales removed
//            at the end; 4 MFMAs for e' and 8 for D. The A fragment of row
//            (kl, c) is plane c of k-block kb when kb = kl, else a zero block;
//   write    E(k, i) for the group's samples k < n and the I-block's i < n;
//            E(k, k) = s_k / 2 from sqd. A group longer than MB_FOLD chunks
//            adds its sums into E every MB_FOLD chunks and starts them again,
//            so that the int32 digit sums never wrap, whatever the list
//            lengths.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_miss_mfma(const uint32_t* __restrict__ bedi, int64_t m, int64_t n,
                                                   int64_t P, const int32_t* __restrict__ gp,
                                                   const int64_t* __restrict__ gstep, const int32_t* __restrict__ gj,
                                                   const uint32_t* __restrict__ gK, const double* __restrict__ sqd,
                                                   const unsigned long long* __restrict__ sks,
                                                   double* __restrict__ E, int64_t lde) {
    __shared__ __attribute__((aligned(128))) uint32_t sP[MB_CHUNK * MB_SPK];           // K planes, 9 KB
    __shared__ __attribute__((aligned(128))) uint32_t sV[MB_CHUNK * MB_SPD];           // D planes, 15 KB
    __shared__ __attribute__((aligned(16))) int32_t sJ[MB_CHUNK * MB_STEP];            // 6 KB
    __shared__ __attribute__((aligned(128))) uint8_t sR[2][MB_STEP * 64];              // rows, 8 KB
    // the A fragments of rows off their k-block: zero blocks at dwords 0, 8,
    // 16 and 24, each in banks that no on-block read of the same pass uses
    __shared__ __attribute__((aligned(128))) uint32_t sZero[32];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid < 32) sZero[tid] = 0u;  // ordered before its readers by the first chunk's barriers
    const int64_t b = blockIdx.x / P, p = blockIdx.x % P;
    const int64_t I0 = b * MB_SAMPLES;
    const int kl_out = lane >> 4;                // output rows (kl_out, c) of this lane
    const int Qf = wave * 16 + (lane & 15);      // sample quad of this lane's B fragments
    const int kb = lane >> 4;                    // k-block of this lane's fragments
    const int kl_a = (lane & 15) >> 2, c_a = lane & 3;  // A row (kl, c) of this lane
    const bool on = kb == kl_a;                  // A row (kl, c) is non-zero only in k-block kl
    const uint32_t* const pa = on ? sP + c_a * MB_PLN + kb * 4 : sZero + (kb < 2 ? 16 : 24);
    const uint32_t* const pv0 = on ? sV + c_a * MB_PLN + kb * 4 : sZero + (kb < 2 ? 16 : 24);
    const uint32_t* const pv1 = on ? sV + (4 + c_a) * MB_PLN + kb * 4 : sZero + (kb < 2 ? 0 : 8);
    const int sa = on ? MB_SPK : 0, sv = on ? MB_SPD : 0;  // dwords per step
    const int er = 16 * wave + (lane >> 2), qr = lane & 3;  // gather: entry er, quarter qr
    const int wst = er * 64 + 16 * (qr ^ ((er >> 1) & 3));
    // rows 16 kb + 8 h + (lane & 15) / 2 have (r >> 1) & 3 = (lane >> 2) & 3 for both h
    const int rtr = (16 * kb + ((lane & 15) >> 1)) * 64 + 16 * (wave ^ ((lane >> 2) & 3)) + 8 * (lane & 1);
    // rows of I-block b through a buffer resource (m * 64 bytes < 2^31)
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(bedi + b * m * 16), (short)0, (int)(m * 64), 0x00020000);
    const int64_t g1 = gp[p + 1];
    for (int64_t g = gp[p]; g < g1; ++g) {
        const int64_t step0 = gstep[g], nsteps = gstep[g + 1] - step0;
        const int64_t bk = (g * MB_GROUP) / MB_SAMPLES;
        const int fm = b > bk ? 2 : b == bk ? 1 : 0;  // D in full, halved, or not at all
        v4i acc[4], af0[4], af1[4];
#pragma unroll
        for (int s_ = 0; s_ < 4; ++s_) acc[s_] = af0[s_] = af1[s_] = (v4i){0, 0, 0, 0};
        // sum_c a_c 256^c of the 4 digit rows in FP64: every partial sum is
        // an integer multiple of 2^8 or 2^16 below 2^53 in magnitude (the fold
        // bound), so the value is exact
        auto digits = [](const v4i& a) {
            double t = fma((double)a[3], 0x1p24, (double)a[2] * 0x1p16);
            t = fma((double)a[1], 0x1p8, t);
            return t + (double)a[0];
        };
        // E(k, i) from the digit sums, scales removed: written, or added to a
        // folded part; the last call also sets E(k, k) and puts back sample 3's
        // code offset, 2 K for every entry of sample k's list (2 sks[k]).
        auto emit = [&](bool add, bool last) {
            const int64_t k_out = g * MB_GROUP + kl_out;
            if (k_out >= n) return;
            double* row = E + k_out * lde + I0;
            const double fsc = fm == 2 ? 0x1p-58 : 0x1p-59;
            const double off3 = last ? 2.0 * (double)sks[k_out] : 0.0;  // exact below 2^22 entries
#pragma unroll
            for (int s_ = 0; s_ < 4; ++s_) {
                const int64_t i_loc = 4 * Qf + s_;
                if (I0 + i_loc >= n) continue;
                // the code scales 1, 4, 16, and 64 (c - 2) for sample 3
                const double ev = s_ == 3 ? digits(acc[3]) * 0x1p-6 + off3
                                          : digits(acc[s_]) * (s_ == 0 ? 1.0 : s_ == 1 ? 0.25 : 0.0625);
                double v = ev * TWO_M29;
                if (fm != 0) {
                    const double isc = s_ == 0 ? 1.0 : s_ == 1 ? 0.25 : s_ == 2 ? 0.0625 : 0.015625;  // indicators
                    const double flo = digits(af0[s_]) * isc, fhi = digits(af1[s_]) * isc;
                    v += (fhi * 4294967296.0 + flo) * fsc;
                }
                if (add) v += row[i_loc];
                if (last && I0 + i_loc == k_out) v = sqd[k_out] * 0.5;  // E(k, k) = s_k / 2
                row[i_loc] = v;
            }
        };

        for (int64_t c0 = 0; c0 < nsteps; c0 += MB_CHUNK) {
            const int csteps = (int)(nsteps - c0 < MB_CHUNK ? nsteps - c0 : MB_CHUNK);
            __syncthreads();  // the previous chunk's readers are done with the staging
            {
                const int64_t s0 = step0 + c0;
                const v4i* srcJ = reinterpret_cast<const v4i*>(gj + s0 * MB_STEP);
                const v4i* srcK = reinterpret_cast<const v4i*>(gK + s0 * MB_STEP);
                for (int x = tid; x < csteps * 16; x += 256) {
                    reinterpret_cast<v4i*>(sJ)[x] = __builtin_nontemporal_load(srcJ + x);
                    const v4i kv = __builtin_nontemporal_load(srcK + x);
                    const uint4 k4 = make_uint4((uint32_t)kv[0], (uint32_t)kv[1], (uint32_t)kv[2], (uint32_t)kv[3]);
                    const int st = x >> 4, t4 = x & 15;  // entries 4 t4 .. + 3 of step st
                    uint32_t o[4];
                    transpose4x4((k4.x + 0x80808080u) ^ 0x80808080u, (k4.y + 0x80808080u) ^ 0x80808080u,
                                 (k4.z + 0x80808080u) ^ 0x80808080u, (k4.w + 0x80808080u) ^ 0x80808080u, o);
#pragma unroll
                    for (int c = 0; c < 4; ++c) sP[st * MB_SPK + c * MB_PLN + t4] = o[c];
                    if (fm != 0) {
                        // K^2 - 3 2^30 K (mod 2^64, a negative integer), in signed digits
                        const unsigned long long B8 = 0x8080808080808080ull;
                        auto dval = [&](uint32_t K) {
                            const unsigned long long k = K;
                            return (k * k - 3ull * (k << 30) + B8) ^ B8;
                        };
                        const unsigned long long v0 = dval(k4.x), v1 = dval(k4.y), v2 = dval(k4.z), v3 = dval(k4.w);
                        transpose4x4((uint32_t)v0, (uint32_t)v1, (uint32_t)v2, (uint32_t)v3, o);
#pragma unroll
                        for (int c = 0; c < 4; ++c) sV[st * MB_SPD + c * MB_PLN + t4] = o[c];
                        transpose4x4((uint32_t)(v0 >> 32), (uint32_t)(v1 >> 32), (uint32_t)(v2 >> 32),
                                     (uint32_t)(v3 >> 32), o);
#pragma unroll
                        for (int c = 0; c < 4; ++c) sV[st * MB_SPD + (4 + c) * MB_PLN + t4] = o[c];
                    }
                }
            }
            __syncthreads();  // staged
            // Quarter rows of MB_PF steps in flight: step st's in ring[st % MB_PF].
            // Every step issues the gather of step st + MB_PF (a step past the
            // chunk reloads its last one, unused), so the wait for a step's
            // rows never covers the gathers issued after them.
            uint4 ring[MB_PF];
            auto fetch = [&](int st) {
                const int sc = st < csteps ? st : csteps - 1;
                const int32_t j = sJ[sc * MB_STEP + er];
                const v4i v = __builtin_amdgcn_raw_buffer_load_b128(rs, (uint32_t)(j * 64 + 16 * qr), 0, 0);
                return make_uint4((uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2], (uint32_t)v[3]);
            };
#pragma unroll
            for (int u = 0; u < MB_PF; ++u) {
                ring[u] = fetch(u);
                __builtin_amdgcn_sched_barrier(0);  // issued in ring order
            }
            for (int st0 = 0; st0 < csteps; st0 += MB_PF) {
#pragma unroll
                for (int u = 0; u < MB_PF; ++u) {
                    const int st = st0 + u;  // csteps is a multiple of MB_PF
                    uint8_t* const rb = sR[st & 1];
                    *reinterpret_cast<uint4*>(rb + wst) = ring[u];
                    ring[u] = fetch(st + MB_PF);
                    __syncthreads();
                    const uint2 lo = lds_read_tr8(rb + rtr);
                    const uint2 hi = lds_read_tr8(rb + rtr + 8 * 64);
                    const uint32_t x[4] = {lo.x, lo.y, hi.x, hi.y};
                    const v4i ak = *reinterpret_cast<const v4i*>(pa + st * sa);
                    v4i bd[4];
                    code_fragments(x, bd);
#pragma unroll
                    for (int s_ = 0; s_ < 4; ++s_) acc[s_] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ak, bd[s_], acc[s_], 0, 0, 0);
                    if (fm != 0) {
                        v4i bm[4];
                        missing_fragments(x, bm);
                        const v4i a0 = *reinterpret_cast<const v4i*>(pv0 + st * sv);
                        const v4i a1 = *reinterpret_cast<const v4i*>(pv1 + st * sv);
#pragma unroll
                        for (int s_ = 0; s_ < 4; ++s_) {
                            af0[s_] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, bm[s_], af0[s_], 0, 0, 0);
                            af1[s_] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, bm[s_], af1[s_], 0, 0, 0);
                        }
                    }
                }
            }
            // a long list: fold the sums into E and start them again
            const int64_t cdone = c0 / MB_CHUNK + 1;
            if (cdone % MB_FOLD == 0 && c0 + MB_CHUNK < nsteps) {
                emit(cdone > MB_FOLD, false);
#pragma unroll
                for (int s_ = 0; s_ < 4; ++s_) acc[s_] = af0[s_] = af1[s_] = (v4i){0, 0, 0, 0};
            }
        }
        emit(nsteps > (int64_t)MB_FOLD * MB_CHUNK, true);
    }
}

// ---------------------------------------------------------------------------

// One 32x32x32 int8 fragment (16 bytes per lane) from one packed word:
// byte t of dword q is the 2-bit field at bits 8t+2q.
__device__ __forceinline__ v4i decode16(uint32_t w) {
    return (v4i){(int)(w & 0x03030303u), (int)((w >> 2) & 0x03030303u),
                 (int)((w >> 4) & 0x03030303u), (int)((w >> 6) & 0x03030303u)};
}

// INT8-MFMA Gram over the lower triangle of 256x256 tiles.
// 4 waves in 2x2, each 128x128 = 4x4 blocks of v_mfma_i32_32x32x32_i8.
// A 32x32x32 fragment takes 16 int8 per lane = one packed word decoded with
// (w >> 2q) & 0x03030303; A and B fragments share that k-layout, so the
// product is the Gram regardless of how the hardware orders k.
// MODE 0: FP64/FP32 G with fused correction including the missing-call
// term E; MODE 2: the same without missing calls; MODE 1: raw int32 Gram.
// W2 (n even, so every row of G and E starts 16-byte aligned): the FP64/FP32
// epilogue moves two adjacent entries per lane and access.
template <int MODE, bool W2 = false>
__global__ __launch_bounds__(256) void k_gram(const uint32_t* __restrict__ X2, int64_t n,
                                              int64_t n_pad, int64_t n_stages,
                                              const double* __restrict__ u, double Cc,
                                              double S, const double* __restrict__ E,
                                              int64_t lde, double* __restrict__ G64,
                                              float* __restrict__ G32, int32_t* __restrict__ Gi,
                                              int64_t ld) {
    constexpr bool HAS_E = MODE == 0;
    // One LDS block: three buffers of A and B panels in the main loop (120 KB);
    // four 32x33 FP64 slabs plus four 32x33 int32 slabs (50.7 KB) in the
    // epilogue.
    constexpr int PANELS_WORDS = 2 * TILE * LDS_PITCH;  // A and B of one buffer
    constexpr int LDS_WORDS = 3 * PANELS_WORDS;
    static_assert(LDS_WORDS >= 4 * 32 * 33 * 3, "epilogue slabs exceed LDS block");
    __shared__ __attribute__((aligned(16))) uint32_t lds[LDS_WORDS];
    uint32_t* ldsA = lds;  // buffer b: lds + b * PANELS_WORDS
    uint32_t* ldsB = lds + TILE * LDS_PITCH;

    // XCD-aware order: hardware deals workgroups round-robin over 8 XCDs;
    // give each XCD a contiguous run of tiles so neighbours share L2 panels.
    const int64_t nb = gridDim.x, bid = blockIdx.x;
    const int64_t xcd = bid % 8, q8 = nb / 8, r8 = nb % 8;
    const int64_t L = xcd * q8 + (xcd < r8 ? xcd : r8) + bid / 8;
    int64_t ti = (int64_t)((sqrt(8.0 * (double)L + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > L) --ti;
    while ((ti + 1) * (ti + 2) / 2 <= L) ++ti;
    const int64_t tj = L - ti * (ti + 1) / 2;

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wr = wave >> 1, wc = wave & 1;
    const int h = lane >> 5, r = lane & 31;

    v16i acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (v16i){0};

    const u32x4* gA = reinterpret_cast<const u32x4*>(X2 + (ti * TILE + tid) * 16);
    const u32x4* gB = reinterpret_cast<const u32x4*>(X2 + (tj * TILE + tid) * 16);
    const int64_t stage_stride = n_pad * 4;  // in 16-byte units
    // Three LDS buffers: stage s is computed from buffer s%3 while the rows
    // of stage s+2 (in registers) go to buffer (s+2)%3 and stage s+3 is
    // fetched. A lane's words of a stage are read in two halves (k-steps
    // 0-3, 4-7), each half one stage-half ahead of its use, and the decode
    // of the next k-step runs under the MFMAs of the current one, across
    // stage boundaries, so the MFMA pipe never waits for LDS. One barrier
    // per stage, at its start.
    u32x4 pa0, pa1, pa2, pa3, pb0, pb1, pb2, pb3;
    auto gload = [&](int64_t st_) {
        const int64_t o = st_ * stage_stride;
        pa0 = gA[o]; pa1 = gA[o + 1]; pa2 = gA[o + 2]; pa3 = gA[o + 3];
        pb0 = gB[o]; pb1 = gB[o + 1]; pb2 = gB[o + 2]; pb3 = gB[o + 3];
    };
    auto lstore = [&](int buf) {
        u32x4* la = reinterpret_cast<u32x4*>(ldsA + buf * PANELS_WORDS + tid * LDS_PITCH);
        u32x4* lb = reinterpret_cast<u32x4*>(ldsB + buf * PANELS_WORDS + tid * LDS_PITCH);
        la[0] = pa0; la[1] = pa1; la[2] = pa2; la[3] = pa3;
        lb[0] = pb0; lb[1] = pb1; lb[2] = pb2; lb[3] = pb3;
    };
    // half hf (0: k-steps 0-3, 1: 4-7) of buffer buf: word kk of row block a is .[kk&3]
    u32x4 H0a[4], H0b[4], H1a[4], H1b[4];
    auto lread = [&](int buf, int hf, u32x4 (&xa)[4], u32x4 (&xb)[4]) {
        const uint32_t* cA = ldsA + buf * PANELS_WORDS;
        const uint32_t* cB = ldsB + buf * PANELS_WORDS;
#pragma unroll
        for (int a = 0; a < 4; ++a)
            xa[a] = *reinterpret_cast<const u32x4*>(cA + (wr * 128 + a * 32 + r) * LDS_PITCH + 8 * h + 4 * hf);
#pragma unroll
        for (int b = 0; b < 4; ++b)
            xb[b] = *reinterpret_cast<const u32x4*>(cB + (wc * 128 + b * 32 + r) * LDS_PITCH + 8 * h + 4 * hf);
    };
    auto wsel = [](const u32x4& x, int k) { return k == 0 ? x.x : k == 1 ? x.y : k == 2 ? x.z : x.w; };

    gload(0);
    lstore(0);
    if (n_stages > 1) {
        gload(1);
        lstore(1);
    }
    gload(n_stages > 2 ? 2 : n_stages - 1);
    __syncthreads();
    lread(0, 0, H0a, H0b);
    v4i fa[4], fb[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) fa[a] = decode16(H0a[a].x);
#pragma unroll
    for (int b = 0; b < 4; ++b) fb[b] = decode16(H0b[b].x);
    int cur = 0;
    for (int64_t s = 0; s < n_stages; ++s) {
        if (s > 0) __syncthreads();
        const int nxt = cur == 2 ? 0 : cur + 1, far = nxt == 2 ? 0 : nxt + 1;
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            // branch-free stage body: past the last stage the copies are
            // redundant (clamped stage index, buffers no longer read)
            if (kk == 0) lread(cur, 1, H1a, H1b);
            if (kk == 1) lstore(far);
            if (kk == 2) gload(s + 3 < n_stages ? s + 3 : n_stages - 1);
            if (kk == 4) lread(nxt, 0, H0a, H0b);
            v4i na[4], nb[4];
            {
#pragma unroll
                for (int a = 0; a < 4; ++a)
                    na[a] = decode16(kk < 3 ? wsel(H0a[a], kk + 1) : kk < 7 ? wsel(H1a[a], kk - 3) : H0a[a].x);
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    nb[b] = decode16(kk < 3 ? wsel(H0b[b], kk + 1) : kk < 7 ? wsel(H1b[b], kk - 3) : H0b[b].x);
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[a], fb[b], acc[a][b], 0, 0, 0);
#pragma unroll
            for (int g = 0; g < 16; ++g) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // four VALU
                if (g < 8 && (kk == 0 || kk == 1 || kk == 2 || kk == 4))
                    __builtin_amdgcn_sched_group_barrier(0x320, 1, 0);  // one LDS or global access
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int a = 0; a < 4; ++a) fa[a] = na[a];
#pragma unroll
            for (int b = 0; b < 4; ++b) fb[b] = nb[b];
        }
        cur = nxt;
    }

    // Epilogue, one 32x32 block at a time through wave-private LDS slabs
    // (the main-loop panels are dead by now) so that both the own-position
    // rows and the mirrored rows are written with coalesced stores. The
    // value is formed symmetrically in (i,j) — x_ij - (u_i + u_j) + C +
    // (E(i,j) + E(j,i)) — so diagonal tiles come out bitwise symmetric.
    // 32x32 accumulator layout: register q of lane l holds
    // row 8(q/4) + 4(l/32) + q%4, column l%32.
    // The E values a block needs (16 per lane in each pattern) are loaded
    // one block ahead, so their latency hides under the previous block's
    // stores; u of the tile's rows and columns is staged in LDS once. The
    // slabs are wave-private, so a wave orders its own LDS traffic and
    // never waits for the other waves. E is read once and G written once:
    // both go non-temporal, so that they do not evict the sample panels the
    // other tiles of the XCD are still reading from L2.
    __syncthreads();
    double* slabD = reinterpret_cast<double*>(lds) + wave * (32 * 33);
    int32_t* slabX = reinterpret_cast<int32_t*>(reinterpret_cast<double*>(lds) + 4 * 32 * 33) + wave * (32 * 33);
    double* uI = reinterpret_cast<double*>(lds) + 4 * 32 * 33 + 2 * 32 * 33;  // after the int32 slabs
    double* uJ = uI + TILE;
    static_assert((4 * 32 * 33 + 2 * 32 * 33 + 2 * TILE) * 8 <= LDS_WORDS * 4, "epilogue LDS");
    const bool diag = ti == tj;
    const double Sinv = 1.0 / (S > 0.0 ? S : 1.0);
    if (MODE != 1) {
        const int64_t ri = ti * TILE + tid, rj = tj * TILE + tid;
        uI[tid] = ri < n ? u[ri] : 0.0;
        uJ[tid] = rj < n ? u[rj] : 0.0;
        __syncthreads();
    }
    auto block_to_slab = [&](int ab) {
        // registers cannot be indexed at run time: block ab's accumulator is
        // picked by a (wave-uniform) branch, the rest of the body stays rolled
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c == ab) {
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    slabX[(8 * (q >> 2) + 4 * h + (q & 3)) * 33 + r] = acc[c >> 2][c & 3][q];
            }
        wave_sync();
    };
    if (MODE == 1) {
#pragma unroll 1
        for (int ab = 0; ab < 16; ++ab) {
            block_to_slab(ab);
            const int64_t I0 = ti * TILE + wr * 128 + (ab >> 2) * 32;  // block rows
            const int64_t J0 = tj * TILE + wc * 128 + (ab & 3) * 32;   // block columns
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) {
                const int row = h * 16 + rr;
                const int64_t i = I0 + row, j = J0 + r;
                if (i < n && j < n) Gi[i * ld + j] = slabX[row * 33 + r];
                if (!diag) {
                    const int64_t mi = J0 + row, mj = I0 + r;  // G(mi, mj) = G(mj, mi)
                    if (mi < n && mj < n) Gi[mi * ld + mj] = slabX[r * 33 + row];
                }
            }
            wave_sync();
        }
        return;
    }
    if constexpr (W2) {
        // Lane (q, p) = (lane / 16, lane % 16) holds the pair of entries
        // 2p, 2p+1 of a 32-wide run and takes four runs (4k + q) per access:
        // the E loads and the G64 stores move 16 bytes per lane, the G32
        // stores 8, half the accesses of the one-entry path below. The values
        // are formed exactly as there. With n even a pair is either wholly
        // inside the matrix or wholly outside, and n - 2 is a valid clamp.
        const int p = lane & 15, q = lane >> 4;
        auto load_e2 = [&](int ab, f64x2 (&ev)[8], f64x2 (&et)[8]) {
            const int64_t I0 = ti * TILE + wr * 128 + (ab >> 2) * 32;
            const int64_t J0 = tj * TILE + wc * 128 + (ab & 3) * 32;
            const int64_t jc = J0 + 2 * p < n ? J0 + 2 * p : n - 2, ic = I0 + 2 * p < n ? I0 + 2 * p : n - 2;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int64_t i = I0 + 4 * k + q < n ? I0 + 4 * k + q : n - 1;
                ev[k] = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(E + i * lde + jc));
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int64_t j = J0 + 4 * k + q < n ? J0 + 4 * k + q : n - 1;
                et[k] = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(E + j * lde + ic));
            }
        };
        auto process2 = [&](int ab, const f64x2 (&ev)[8], const f64x2 (&et)[8]) {
            const int a = ab >> 2, b = ab & 3;
            block_to_slab(ab);
            const int64_t I0 = ti * TILE + wr * 128 + a * 32;
            const int64_t J0 = tj * TILE + wc * 128 + b * 32;
            if (HAS_E) {  // stage E(i,j), rows 4k + q, columns 2p, 2p+1
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    slabD[(4 * k + q) * 33 + 2 * p] = ev[k].x;
                    slabD[(4 * k + q) * 33 + 2 * p + 1] = ev[k].y;
                }
                wave_sync();
            }
            // rows i, i+1 = I0 + 2p, +1, columns j = J0 + 4k + q: form g, write the mirror
            const int64_t i = I0 + 2 * p;
            const double ui0 = uI[wr * 128 + a * 32 + 2 * p], ui1 = uI[wr * 128 + a * 32 + 2 * p + 1];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int col = 4 * k + q;
                const int64_t j = J0 + col;
                const double uj = uJ[wc * 128 + b * 32 + col];
                const double e0 = HAS_E ? slabD[(2 * p) * 33 + col] + et[k].x : 0.0;
                const double e1 = HAS_E ? slabD[(2 * p + 1) * 33 + col] + et[k].y : 0.0;
                const double v0 = (((double)slabX[(2 * p) * 33 + col] - (ui0 + uj)) + Cc) + e0;
                const double v1 = (((double)slabX[(2 * p + 1) * 33 + col] - (ui1 + uj)) + Cc) + e1;
                const double g0 = v0 * Sinv, g1 = v1 * Sinv;
                slabD[(2 * p) * 33 + col] = g0;
                slabD[(2 * p + 1) * 33 + col] = g1;
                if (!diag && i < n && j < n) {
                    if (G64) __builtin_nontemporal_store((f64x2){g0, g1}, reinterpret_cast<f64x2*>(G64 + j * ld + i));
                    if (G32) __builtin_nontemporal_store((f32x2){(float)g0, (float)g1}, reinterpret_cast<f32x2*>(G32 + j * ld + i));
                }
            }
            wave_sync();
            // own positions: rows I0 + 4k + q, columns J0 + 2p, +1
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int row = 4 * k + q;
                const int64_t ii = I0 + row, j = J0 + 2 * p;
                if (ii < n && j < n) {
                    const double g0 = slabD[row * 33 + 2 * p], g1 = slabD[row * 33 + 2 * p + 1];
                    if (G64) __builtin_nontemporal_store((f64x2){g0, g1}, reinterpret_cast<f64x2*>(G64 + ii * ld + j));
                    if (G32) __builtin_nontemporal_store((f32x2){(float)g0, (float)g1}, reinterpret_cast<f32x2*>(G32 + ii * ld + j));
                }
            }
            wave_sync();
        };
        f64x2 evA[8], etA[8], evB[8], etB[8];
        if (HAS_E) load_e2(0, evA, etA);
#pragma unroll 1
        for (int ab = 0; ab < 16; ab += 2) {
            if (HAS_E) load_e2(ab + 1, evB, etB);
            process2(ab, evA, etA);
            if (HAS_E && ab + 2 < 16) load_e2(ab + 2, evA, etA);
            process2(ab + 1, evB, etB);
        }
        return;
    }
    // E(i,j) for rows I0 + 16h + rr, column J0 + r (row pattern) and
    // E(j,i) for rows J0 + 16h + rr, column I0 + r (transposed pattern)
    // Indices are clamped into the matrix instead of predicating the loads
    // (values outside it are never used), which keeps the loads unbranched
    // so that waits cover only the block they belong to.
    auto load_e = [&](int ab, double (&ev)[16], double (&et)[16]) {
        const int64_t I0 = ti * TILE + wr * 128 + (ab >> 2) * 32;
        const int64_t J0 = tj * TILE + wc * 128 + (ab & 3) * 32;
        const int64_t jr = J0 + r < n ? J0 + r : n - 1, ir = I0 + r < n ? I0 + r : n - 1;
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            const int64_t i = I0 + h * 16 + rr < n ? I0 + h * 16 + rr : n - 1;
            ev[rr] = __builtin_nontemporal_load(E + i * lde + jr);
        }
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            const int64_t j = J0 + h * 16 + rr < n ? J0 + h * 16 + rr : n - 1;
            et[rr] = __builtin_nontemporal_load(E + j * lde + ir);
        }
    };
    auto process = [&](int ab, const double (&ev)[16], const double (&et)[16]) {
        const int a = ab >> 2, b = ab & 3;
        block_to_slab(ab);
        const int64_t I0 = ti * TILE + wr * 128 + a * 32;  // block rows
        const int64_t J0 = tj * TILE + wc * 128 + b * 32;  // block columns
        if (HAS_E) {  // pass 1 (row pattern): stage E(i,j)
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) slabD[(h * 16 + rr) * 33 + r] = ev[rr];
            wave_sync();
        }
        // pass 2 (transposed pattern): form g, write the mirror
        const int64_t i = I0 + r;
        const double ui = uI[wr * 128 + a * 32 + r];
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            const int col = h * 16 + rr;
            const int64_t j = J0 + col;
            if (i < n && j < n) {
                const double e = HAS_E ? slabD[r * 33 + col] + et[rr] : 0.0;
                const double v = (((double)slabX[r * 33 + col] - (ui + uJ[wc * 128 + b * 32 + col])) + Cc) + e;
                const double g = v * Sinv;
                slabD[r * 33 + col] = g;
                if (!diag) {
                    if (G64) __builtin_nontemporal_store(g, G64 + j * ld + i);
                    if (G32) __builtin_nontemporal_store((float)g, G32 + j * ld + i);
                }
            }
        }
        wave_sync();
        // pass 3 (row pattern): own positions
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
            const int row = h * 16 + rr;
            const int64_t ii = I0 + row, j = J0 + r;
            if (ii < n && j < n) {
                const double g = slabD[row * 33 + r];
                if (G64) __builtin_nontemporal_store(g, G64 + ii * ld + j);
                if (G32) __builtin_nontemporal_store((float)g, G32 + ii * ld + j);
            }
        }
        wave_sync();
    };
    double evA[16], etA[16], evB[16], etB[16];
    if (HAS_E) load_e(0, evA, etA);
#pragma unroll 1
    for (int ab = 0; ab < 16; ab += 2) {
        if (HAS_E) load_e(ab + 1, evB, etB);
        process(ab, evA, etA);
        if (HAS_E && ab + 2 < 16) load_e(ab + 2, evA, etA);
        process(ab + 1, evB, etB);
    }
}

}  // namespace
struct cb_grm {
    cb_context* ctx = nullptr;
    cb_grm_options opts{};
    int64_t n = 0, m = 0;
    int outputs = 0;
    double sum2pq = 0.0;
    int64_t total_valid = 0;
    // per-SNP stats (device)
    DevBuf<int32_t> nobs, sumx, h1, het, h2;
    DevBuf<uint8_t> valid;
    DevBuf<float> freq, mean, twopq;
    DevBuf<uint32_t> kfix;
    // outputs
    DevBuf<double> G64;
    DevBuf<float> G32;
    DevBuf<int32_t> Gi;
    // work buffers of compute, kept from one call to the next and only grown,
    // so a repeated compute at a fixed shape allocates and frees nothing
    DevBuf<uint32_t> X2;              // sample-major packed dosages
    DevBuf<double> u;                 // per-sample centring terms
    DevBuf<double> E;                 // missing-call correction, n x n
    DevBuf<int64_t> smp_off;
    DevBuf<int32_t> mcnt, smp_cnt;
    DevBuf<unsigned long long> ljk;  // list entries: K_j << 32 | j
    DevBuf<unsigned long long> sq;
    DevBuf<double> sqd;
    // the missing-call kernel's inputs: rows in I-block-major order, the
    // I-block segments of the SNP lists, the group lists and A fragments
    DevBuf<uint32_t> bedi, gK;
    DevBuf<int32_t> gj, gp;
    DevBuf<int64_t> gstep;
};

namespace {
// grow-only allocation of a work buffer (contents undefined)
template <class T>
void grow(DevBuf<T>& b, size_t count) {
    if (b.n < count) b.alloc(count);
}
}  // namespace

static void compute_grm(cb_grm* grm, const cb_genotypes* g, int outputs) {
    cb_context* ctx = grm->ctx;
    hipStream_t st = ctx->stream;
    const int64_t n = g->n, m = g->m;
    const int64_t n_pad = round_up(n, TILE);
    const int64_t m_pad = round_up(m, STAGE_SNPS);
    const int64_t n_stages = m_pad / STAGE_SNPS;
    grm->n = n;
    grm->m = m;
    grm->outputs = outputs;

    // 1. per-SNP statistics
    grm->nobs.alloc(m); grm->sumx.alloc(m); grm->h1.alloc(m); grm->het.alloc(m); grm->h2.alloc(m);
    grm->valid.alloc(m); grm->freq.alloc(m); grm->mean.alloc(m); grm->twopq.alloc(m);
    grm->kfix.alloc(m_pad);
    CB_HIP(hipMemsetAsync(grm->kfix.p, 0, grm->kfix.bytes(), st));
    const double min_maf = (double)grm->opts.min_maf;
    timed(ctx, "k_snp_stats", [&] {
        k_snp_stats<<<dim3((unsigned)ceil_div(m, 4)), dim3(256), 0, st>>>(
            g->bed.p, g->pitch, n, m, min_maf, grm->nobs.p, grm->sumx.p, grm->h1.p, grm->het.p,
            grm->h2.p, grm->valid.p, grm->freq.p, grm->mean.p, grm->twopq.p, grm->kfix.p);
    });
    std::vector<uint8_t> hvalid(m);
    std::vector<float> hmean(m), h2pq(m);
    std::vector<int32_t> hnobs(m);
    copy_host(st, hvalid.data(), grm->valid.p, m, hipMemcpyDeviceToHost);
    copy_host(st, hmean.data(), grm->mean.p, m * 4, hipMemcpyDeviceToHost);
    copy_host(st, h2pq.data(), grm->twopq.p, m * 4, hipMemcpyDeviceToHost);
    copy_host(st, hnobs.data(), grm->nobs.p, m * 4, hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));

    // sum(2pq) in the reference's order: double sum of float 2pq per SNP
    // block, then across blocks (grm_builder.cpp:272-278,353); C = sum mu^2.
    double total = 0.0, Cc = 0.0;
    int64_t nvalid = 0, nmiss_total = 0;
    const int64_t bs = grm->opts.block_size;
    for (int64_t b0 = 0; b0 < m; b0 += bs) {
        double blk = 0.0;
        const int64_t b1 = std::min(m, b0 + bs);
        for (int64_t j = b0; j < b1; ++j) {
            if (!hvalid[j]) continue;
            blk += (double)h2pq[j];
            const double mu = (double)hmean[j];
            Cc += mu * mu;
            ++nvalid;
            nmiss_total += n - hnobs[j];
        }
        total += blk;
    }
    grm->sum2pq = total;
    grm->total_valid = nvalid;

    // 2. sample-major packed dosages
    DevBuf<uint32_t>& X2 = grm->X2;
    grow(X2, (size_t)n_stages * n_pad * 16);
    timed(ctx, "k_pack", [&] {
        k_pack<<<dim3((unsigned)(n_pad / TILE), (unsigned)n_stages), dim3(256), 0, st>>>(
            g->bed.p, g->pitch, n, m, n_pad, grm->valid.p, X2.p);
    });

    // 3. per-sample centring terms
    DevBuf<double>& u = grm->u;
    grow(u, n_pad);
    CB_HIP(hipMemsetAsync(u.p, 0, n_pad * 8, st));
    timed(ctx, "k_row_sums", [&] {
        k_row_sums<<<dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, st>>>(X2.p, n, n_pad, n_stages,
                                                                        grm->kfix.p, u.p);
    });

    // 4. missing calls among the kept SNPs
    const bool has_e = nmiss_total > 0 && (outputs & (CB_GRM_OUT_F64 | CB_GRM_OUT_F32));
    DevBuf<double>& E = grm->E;
    if (has_e) {
        const int64_t nw = ceil_div(n, 16), nc = ceil_div(m, MC_ROWS);
        CB_REQUIRE(nc < 65536, "GRM: too many SNPs for the missing-call list kernels' grid");
        DevBuf<int32_t>& mcnt = grm->mcnt;
        DevBuf<int32_t>& smp_cnt = grm->smp_cnt;
        grow(mcnt, nc * n);
        grow(smp_cnt, n);
        const dim3 lgrid((unsigned)ceil_div(nw, 256), (unsigned)nc);
        timed(ctx, "k_miss_count", [&] {
            k_miss_count<<<lgrid, 256, 0, st>>>(g->bed.p, g->pitch, n, m, grm->valid.p, mcnt.p);
        });
        timed(ctx, "k_miss_scan", [&] {
            k_miss_scan<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(mcnt.p, n, nc, smp_cnt.p);
        });
        std::vector<int32_t> hcnt(n);
        copy_host(st, hcnt.data(), smp_cnt.p, n * 4, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        std::vector<int64_t> smp_off(n + 1, 0);
        for (int64_t i = 0; i < n; ++i) smp_off[i + 1] = smp_off[i] + round_up(hcnt[i], 8);
        DevBuf<int64_t>& d_smp_off = grm->smp_off;
        grow(d_smp_off, n + 1);
        copy_host(st, d_smp_off.p, smp_off.data(), (n + 1) * 8, hipMemcpyHostToDevice);
        const int64_t nlist = std::max<int64_t>(8, smp_off[n]);
        DevBuf<unsigned long long>& ljk = grm->ljk;
        DevBuf<unsigned long long>& sq = grm->sq;  // per sample: K^2 low words, high words, then the sums of K
        grow(ljk, nlist);
        grow(sq, 3 * n);
        CB_HIP(hipMemsetAsync(ljk.p, 0, nlist * 8, st));
        timed(ctx, "k_miss_emit", [&] {
            k_miss_emit<<<lgrid, 256, 0, st>>>(g->bed.p, g->pitch, n, m, grm->valid.p, grm->kfix.p, d_smp_off.p,
                                               mcnt.p, ljk.p);
        });
        timed(ctx, "k_sample_sq", [&] {
            k_sample_sq<<<dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, st>>>(d_smp_off.p, ljk.p, n, sq.p, sq.p + n,
                                                                            sq.p + 2 * n);
        });
        // u_i += sum_{j miss i} mu_j^2
        std::vector<unsigned long long> hsq(2 * n);
        std::vector<double> hu(n);
        copy_host(st, hsq.data(), sq.p, 2 * n * 8, hipMemcpyDeviceToHost);
        copy_host(st, hu.data(), u.p, n * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        // s_i = sum_{j miss i} mu_j^2, an exact integer in 2^-52 units, converted once
        std::vector<double> hsqd(n);
        for (int64_t i = 0; i < n; ++i) {
            const unsigned __int128 s_i = ((unsigned __int128)hsq[n + i] << 32) + hsq[i];
            hsqd[i] = (double)s_i * (1.0 / 4503599627370496.0);
            hu[i] += hsqd[i];
        }
        DevBuf<double>& sqd = grm->sqd;
        grow(sqd, n);
        copy_host(st, sqd.p, hsqd.data(), n * 8, hipMemcpyHostToDevice);
        copy_host(st, u.p, hu.data(), n * 8, hipMemcpyHostToDevice);

        grow(E, (size_t)n * n);
        // 4b. the correction on INT8 MFMA (k_miss_mfma): rows in I-block-major
        // order, the group lists and their A fragments
        const int64_t nb = ceil_div(n, MB_SAMPLES);
        CB_REQUIRE(m * 64 < ((int64_t)1 << 31), "GRM: too many SNPs for the missing-call kernel's row buffer");
        DevBuf<uint32_t>& bedi = grm->bedi;
        grow(bedi, (size_t)nb * m * 16);
        const unsigned gy = (unsigned)std::min<int64_t>(m, 65535);
        timed(ctx, "k_encode_iblock", [&] {
            k_encode_iblock<<<dim3((unsigned)ceil_div(nb * 16, 256), gy), 256, 0, st>>>(g->bed.p, g->pitch, n, m,
                                                                                      nb, bedi.p);
        });
        // groups of 4 samples; a group has as many K-steps as its longest list
        // needs (16 entries of each sample per K-step), rounded up to a
        // multiple of MB_PF
        const int64_t ngrp = ceil_div(n, MB_GROUP);
        std::vector<int64_t> gstep(ngrp + 1, 0);
        for (int64_t gi = 0; gi < ngrp; ++gi) {
            int64_t longest = 0;
            for (int64_t k = gi * MB_GROUP; k < std::min(n, (gi + 1) * MB_GROUP); ++k)
                longest = std::max<int64_t>(longest, hcnt[k]);
            gstep[gi + 1] = gstep[gi] + round_up(ceil_div(longest, MB_KBLK), MB_PF);
        }
        const int64_t tsteps = gstep[ngrp];
        DevBuf<int64_t>& dgstep = grm->gstep;
        grow(dgstep, ngrp + 1);
        copy_host(st, dgstep.p, gstep.data(), (ngrp + 1) * 8, hipMemcpyHostToDevice);
        DevBuf<int32_t>& gj = grm->gj;
        DevBuf<uint32_t>& gK = grm->gK;
        grow(gj, std::max<int64_t>(1, tsteps * MB_STEP));
        grow(gK, std::max<int64_t>(1, tsteps * MB_STEP));
        timed(ctx, "k_group_entries", [&] {
            k_group_entries<<<(unsigned)ngrp, 256, 0, st>>>(d_smp_off.p, smp_cnt.p, ljk.p, n, dgstep.p, gj.p,
                                                           gK.p);
        });
        // group ranges of about equal K-steps, P of them per I-block
        const int64_t P = std::min<int64_t>(4096, round_up(ngrp, MB_XCD));
        std::vector<int32_t> gp(P + 1);
        for (int64_t q = 0; q < P; ++q)
            gp[q] = (int32_t)(std::lower_bound(gstep.begin(), gstep.end(), q * tsteps / P) - gstep.begin());
        gp[P] = (int32_t)ngrp;
        DevBuf<int32_t>& dgp = grm->gp;
        grow(dgp, P + 1);
        copy_host(st, dgp.p, gp.data(), (P + 1) * 4, hipMemcpyHostToDevice);
        CB_REQUIRE(nb * P < ((int64_t)1 << 31), "GRM: too many samples for the missing-call kernel's grid");
        timed(ctx, "k_miss_cross", [&] {
            k_miss_mfma<<<(unsigned)(nb * P), 256, 0, st>>>(bedi.p, m, n, P, dgp.p, dgstep.p, gj.p, gK.p, sqd.p,
                                                           sq.p + 2 * n, E.p, n);
        });
    }

    // 5. Gram + fused correction
    const int64_t T = n_pad / TILE;
    const int64_t ntiles = T * (T + 1) / 2;
    CB_REQUIRE(ntiles < (int64_t(1) << 31), "GRM too large for one launch");
    if (outputs & CB_GRM_OUT_F64) grm->G64.alloc((size_t)n * n); else grm->G64.release();
    if (outputs & CB_GRM_OUT_F32) grm->G32.alloc((size_t)n * n); else grm->G32.release();
    if (outputs & CB_GRM_OUT_INT) grm->Gi.alloc((size_t)n * n); else grm->Gi.release();
    if (outputs & (CB_GRM_OUT_F64 | CB_GRM_OUT_F32)) {
        // G and E are n x n with leading dimension n: for even n every row
        // starts 16-byte aligned (device allocations are 256-byte aligned)
        const bool w2 = n % 2 == 0;
        timed(ctx, "k_gram", [&] {
            if (has_e && w2)
                k_gram<0, true><<<dim3((unsigned)ntiles), dim3(256), 0, st>>>(
                    X2.p, n, n_pad, n_stages, u.p, Cc, total, E.p, n, grm->G64.p, grm->G32.p, nullptr, n);
            else if (has_e)
                k_gram<0><<<dim3((unsigned)ntiles), dim3(256), 0, st>>>(
                    X2.p, n, n_pad, n_stages, u.p, Cc, total, E.p, n, grm->G64.p, grm->G32.p, nullptr, n);
            else if (w2)
                k_gram<2, true><<<dim3((unsigned)ntiles), dim3(256), 0, st>>>(
                    X2.p, n, n_pad, n_stages, u.p, Cc, total, nullptr, n, grm->G64.p, grm->G32.p, nullptr, n);
            else
                k_gram<2><<<dim3((unsigned)ntiles), dim3(256), 0, st>>>(
                    X2.p, n, n_pad, n_stages, u.p, Cc, total, nullptr, n, grm->G64.p, grm->G32.p, nullptr, n);
        });
    }
    if (outputs & CB_GRM_OUT_INT) {
        timed(ctx, "k_gram_int", [&] {
            k_gram<1><<<dim3((unsigned)ntiles), dim3(256), 0, st>>>(
                X2.p, n, n_pad, n_stages, u.p, Cc, total, nullptr, n, nullptr, nullptr, grm->Gi.p, n);
        });
    }
    CB_HIP(hipStreamSynchronize(st));
}

extern "C" {

void cb_grm_options_default(cb_grm_options* o) {
    if (!o) return;
    o->block_size = 2048;
    o->min_maf = 0.01f;
    o->min_hwe = 0.0f;
    o->algorithm = 1;
}

int cb_grm_create(cb_context* ctx, const cb_grm_options* opts, cb_grm** out) {
    return guarded([&] {
        CB_REQUIRE(ctx && out, "cb_grm_create: NULL argument");
        cb_grm_options o;
        cb_grm_options_default(&o);
        if (opts) o = *opts;
        CB_REQUIRE(o.algorithm == 1, "cb_grm_create: only algorithm 1 (VanRaden) is implemented");
        CB_REQUIRE(o.min_hwe == 0.0f, "cb_grm_create: HWE filtering is not implemented (min_hwe must be 0)");
        CB_REQUIRE(o.block_size > 0, "cb_grm_create: block_size must be positive");
        CB_REQUIRE(o.min_maf >= 0.0078125f,
                   "cb_grm_create: min_maf below 2^-7 is not supported by the integer centring");
        auto* g = new cb_grm();
        g->ctx = ctx;
        g->opts = o;
        *out = g;
    });
}

void cb_grm_destroy(cb_grm* grm) {
    if (!grm) return;
    (void)hipSetDevice(grm->ctx->device);
    tls_stream = grm->ctx->stream;  // its buffers are released behind the context's queued work
    delete grm;
}

int cb_grm_compute(cb_grm* grm, const cb_genotypes* g, int outputs) {
    return guarded([&] {
        CB_REQUIRE(grm && g, "cb_grm_compute: NULL argument");
        CB_REQUIRE(g->ctx->device == grm->ctx->device, "cb_grm_compute: genotypes live on another device");
        CB_REQUIRE(outputs != 0 && (outputs & ~7) == 0, "cb_grm_compute: bad output mask");
        set_device(grm->ctx);
        compute_grm(grm, g, outputs);
    });
}

static void copy_out(const cb_grm* grm, const void* src, size_t esz, void* dst, int64_t ld) {
    CB_REQUIRE(src, "requested GRM output was not computed");
    CB_REQUIRE(dst && ld >= grm->n, "bad destination / leading dimension");
    set_device(grm->ctx);
    copy_host_2d(grm->ctx->stream, dst, ld * esz, src, grm->n * esz, grm->n * esz, grm->n, hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(grm->ctx->stream));
}

int cb_grm_copy_grm(const cb_grm* grm, double* out, int64_t ld) {
    return guarded([&] { CB_REQUIRE(grm, "NULL grm"); copy_out(grm, grm->G64.p, 8, out, ld); });
}
int cb_grm_copy_grm_f32(const cb_grm* grm, float* out, int64_t ld) {
    return guarded([&] { CB_REQUIRE(grm, "NULL grm"); copy_out(grm, grm->G32.p, 4, out, ld); });
}
int cb_grm_copy_int_gram(const cb_grm* grm, int32_t* out, int64_t ld) {
    return guarded([&] { CB_REQUIRE(grm, "NULL grm"); copy_out(grm, grm->Gi.p, 4, out, ld); });
}

int cb_grm_write(const cb_grm* grm, const char* path, int format) {
    return guarded([&] {
        CB_REQUIRE(grm && path, "cb_grm_write: NULL argument");
        CB_REQUIRE(grm->G64.p, "cb_grm_write: FP64 G not computed");
        set_device(grm->ctx);
        CB_HIP(hipStreamSynchronize(grm->ctx->stream));
        write_symmetric(grm->ctx, grm->G64.p, grm->n, grm->n, true, path, format);
    });
}

int cb_grm_copy_grm_device(const cb_grm* grm, void* dst, int64_t ld) {
    return guarded([&] {
        CB_REQUIRE(grm && grm->G64.p, "cb_grm_copy_grm_device: FP64 G not computed");
        CB_REQUIRE(dst && ld >= grm->n, "cb_grm_copy_grm_device: bad destination");
        set_device(grm->ctx);
        CB_HIP(hipMemcpy2DAsync(dst, ld * 8, grm->G64.p, grm->n * 8, grm->n * 8, grm->n,
                                hipMemcpyDeviceToDevice, grm->ctx->stream));
        CB_HIP(hipStreamSynchronize(grm->ctx->stream));
    });
}

int cb_grm_device_ptr(const cb_grm* grm, int which, void** ptr, int64_t* ld) {
    return guarded([&] {
        CB_REQUIRE(grm && ptr && ld, "cb_grm_device_ptr: NULL argument");
        void* p = which == CB_GRM_OUT_F64 ? (void*)grm->G64.p
                  : which == CB_GRM_OUT_F32 ? (void*)grm->G32.p
                  : which == CB_GRM_OUT_INT ? (void*)grm->Gi.p
                                            : nullptr;
        CB_REQUIRE(p, "cb_grm_device_ptr: output not computed");
        *ptr = p;
        *ld = grm->n;
    });
}

int cb_grm_sum2pq(const cb_grm* grm, double* out) {
    return guarded([&] { CB_REQUIRE(grm && out, "NULL argument"); *out = grm->sum2pq; });
}
int cb_grm_total_snps(const cb_grm* grm, int64_t* out) {
    return guarded([&] { CB_REQUIRE(grm && out, "NULL argument"); *out = grm->total_valid; });
}
int cb_grm_num_samples(const cb_grm* grm, int64_t* out) {
    return guarded([&] { CB_REQUIRE(grm && out, "NULL argument"); *out = grm->n; });
}

int cb_grm_copy_allele_freqs(const cb_grm* grm, float* out, int64_t m) {
    return guarded([&] {
        CB_REQUIRE(grm && out && m == grm->m && grm->freq.p, "cb_grm_copy_allele_freqs: bad argument");
        set_device(grm->ctx);
        copy_host(grm->ctx->stream, out, grm->freq.p, m * 4, hipMemcpyDeviceToHost);
    });
}

int cb_grm_copy_snp_stats(const cb_grm* grm, int64_t m, int32_t* n_obs, int32_t* sum_x,
                          int32_t* n_hom1, int32_t* n_het, int32_t* n_hom2, uint8_t* valid,
                          float* mean_f, float* twopq_f) {
    return guarded([&] {
        CB_REQUIRE(grm && m == grm->m && grm->nobs.p, "cb_grm_copy_snp_stats: bad argument");
        set_device(grm->ctx);
        if (n_obs) copy_host(grm->ctx->stream, n_obs, grm->nobs.p, m * 4, hipMemcpyDeviceToHost);
        if (sum_x) copy_host(grm->ctx->stream, sum_x, grm->sumx.p, m * 4, hipMemcpyDeviceToHost);
        if (n_hom1) copy_host(grm->ctx->stream, n_hom1, grm->h1.p, m * 4, hipMemcpyDeviceToHost);
        if (n_het) copy_host(grm->ctx->stream, n_het, grm->het.p, m * 4, hipMemcpyDeviceToHost);
        if (n_hom2) copy_host(grm->ctx->stream, n_hom2, grm->h2.p, m * 4, hipMemcpyDeviceToHost);
        if (valid) copy_host(grm->ctx->stream, valid, grm->valid.p, m, hipMemcpyDeviceToHost);
        if (mean_f) copy_host(grm->ctx->stream, mean_f, grm->mean.p, m * 4, hipMemcpyDeviceToHost);
        if (twopq_f) copy_host(grm->ctx->stream, twopq_f, grm->twopq.p, m * 4, hipMemcpyDeviceToHost);
    });
}

}  // extern "C"

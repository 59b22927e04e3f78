// Mixed-model equations on gfx950: the MME operator, block PCG in FP64 and
// in FP32 with FP64 iterative refinement, and the AI-REML / EM-REML driver
// with Hutchinson trace probes.
//
// Reference counterparts (paths relative to the reference tree):
//   MMELHSBuilder (LHS = [X'X X'Z; Z'X Z'Z + lambda Q], RHS = W'y)
//                                   src/core/mme_builder.cpp:72-423
//   PCGSolver (relative stop, block-Jacobi)  src/core/pcg_solver.h:35-128,436-536
//   MatrixFreePCG (absolute stop)    src/core/matrix_free_pcg.h:20-66
//   AI_REML data-driven AI step      src/core/vce.cpp:2414-2857
//   AI_REML EM + Hutchinson step     src/core/vce.cpp:1490-1574
//   estimate_trace_Qinv_Cuu          src/core/vce.cpp:1266-1315
//   project_spd_2x2                  src/core/vce.cpp:267-283
//
// Layout. dim = p + q unknowns (fixed effects first). Right-hand-side blocks
// are column-major [ldv x r] with ldv = dim rounded up to 64. The
// coefficient matrix C(lambda) = W'W + lambda * blockdiag(0, Q) is applied
// as a CSR SpMM of W'W (assembled once on the host, like the reference's
// LHS with its lambda update map) plus lambda * Q X_u, Q dense (G^-1) or
// CSR (A^-1). Q and W'W are held in FP64 and FP32.
//
// Kernels and their bounds (per operator application on an r-column block):
//   k_dense_mfma  Q X_u on MFMA 16x16x4 (FP32 or FP64), split-K; q^2*s bytes
//                 of Q read once (s = 4 or 8), 2 q^2 r flops: HBM-bound for
//                 small r, matrix-core bound above ~20 (FP32) / ~10 (FP64)
//                 columns
//   k_gemv        r <= 2 dense product, HBM-bound
//   k_csr_spmm    W'W and CSR Q, HBM/L2-bound (8-12 bytes per nonzero)
//   k_col_dots    per-column FP64 dot products, two-stage deterministic
//   k_pcg_*       fused CG updates, HBM-bound
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

#include "cb_blas.h"
#include "cb_dense_tuned.h"
#include "cb_internal.h"

using namespace cbamd;

namespace {

constexpr int DOT_CHUNKS = 128;   // row chunks of the two-stage column dot
constexpr int CG_RPT = 4;         // rows per thread of the CG vector kernels (k_cg_xr_diag, k_cg_p)
constexpr int PB_LDS = 4096;      // block rows whose r k_precond_block_dots stages in LDS
constexpr int PB_U = 16;          // Binv entries a lane of k_precond_block_dots loads at once
constexpr double REPLACE_DROP2 = 1e-6;  // squared residual drop that triggers a replacement
constexpr int64_t LONG_CHUNK = 4096;    // nonzeros per block of a long W'W row
constexpr int64_t MULTI_FIN_MAX = 16;   // long rows of several chunks a dot chunk may finish
constexpr int MF_ROWS = 128;      // rows per k_dense_mfma tile (4 waves x 32)
constexpr int MF_KC = 32;         // k per chunk
constexpr int MF_KCP = MF_KC + 4; // LDS pitch of a staged X column
constexpr int MF_MAXNT = 5;       // 16-column subtiles per launch (80 columns)

// ---------------------------------------------------------------------------
// Generic device kernels

// The few long rows of W'W (fixed effects; the mean row touches every
// unknown) in two deterministic stages: block per (chunk of <= LONG_CHUNK
// nonzeros, column) -> partial, then a wave per (row, column) summing its
// chunks in order. FP64 accumulation. Thread t adds nonzeros t, t + 256, ...
// of the chunk in that order; four terms' loads are issued before their
// gathers and their adds. When row1 is given, a chunk that is the whole of
// its row also writes the row's result y_c[row] (the value the finishing
// sum of that one partial gives, 0 + part), so only rows of several chunks
// are left to finish.
template <class TM, class TV>
__device__ __forceinline__ void long_chunk_part(int64_t ch, int64_t nchunks, const int64_t* __restrict__ cbeg,
                                                const int64_t* __restrict__ cend, const int32_t* __restrict__ col,
                                                const TM* __restrict__ val, const TV* __restrict__ x,
                                                double* __restrict__ part_c, const int32_t* __restrict__ row1,
                                                TV* __restrict__ y_c) {
    __shared__ double red[4];
    double acc = 0;
    const int64_t e = cend[ch];
    int64_t k = cbeg[ch] + threadIdx.x;
    for (; k + 3 * 256 < e; k += 4 * 256) {
        const int32_t c0 = col[k], c1 = col[k + 256], c2 = col[k + 512], c3 = col[k + 768];
        const TM v0 = val[k], v1 = val[k + 256], v2 = val[k + 512], v3 = val[k + 768];
        const TV x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        acc += (double)v0 * (double)x0;
        acc += (double)v1 * (double)x1;
        acc += (double)v2 * (double)x2;
        acc += (double)v3 * (double)x3;
    }
    for (; k < e; k += 256) acc += (double)val[k] * (double)x[col[k]];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double v = (red[0] + red[1]) + (red[2] + red[3]);
        part_c[ch] = v;
        if (row1 && row1[ch] >= 0) y_c[row1[ch]] = (TV)(0.0 + v);
    }
}

// One lane's share of a CSR row sum: val[k] x[col[k]] for k = b, b + s, ...
// < e, added in that order. The first two terms' indices and values are
// loaded together, then their two gathers, so that a row of up to 2 s
// nonzeros waits for one round trip of each kind instead of one per term;
// the loads of terms past the end read entry 0 (and x[0]) and are not
// added (the arrays and x hold at least one entry).
template <class TM, class TV>
struct CsrLane {
    int64_t k0, k1;
    bool t0, t1;
    int32_t c0, c1;
    TM v0, v1;
    TV x0, x1;
    __device__ __forceinline__ void indices(const int32_t* __restrict__ col, const TM* __restrict__ val,
                                            int64_t b, int64_t e, int s, bool lane_on) {
        t0 = lane_on && b < e;
        t1 = lane_on && b + s < e;
        k0 = t0 ? b : 0;
        k1 = t1 ? b + s : 0;
        c0 = col[k0];
        c1 = col[k1];
        v0 = val[k0];
        v1 = val[k1];
    }
    __device__ __forceinline__ void gather(const TV* __restrict__ x) {
        x0 = x[t0 ? c0 : 0];
        x1 = x[t1 ? c1 : 0];
    }
    // the two terms, then the rest of the row one term at a time
    __device__ __forceinline__ TV sum(const int32_t* __restrict__ col, const TM* __restrict__ val,
                                      const TV* __restrict__ x, int64_t b, int64_t e, int s, TV acc) const {
        if (t0) acc += (TV)v0 * x0;
        if (t1) acc += (TV)v1 * x1;
        if (t1)
            for (int64_t k = b + 2 * s; k < e; k += s) acc += (TV)val[k] * x[col[k]];
        return acc;
    }
};

// Y[row, c] = alpha * sum_k val(row,k) X[col(k), c] (+ Y[row, c] when
// accumulating) for rows in [row0, row1): S lanes per (row, c), lane j
// taking nonzeros j, j + S, ... of the row (consecutive lanes read
// consecutive values and column indices), then a butterfly over the S
// lanes. S follows the matrix's mean row length (csr_lanes); S = 1 is a
// thread per row.
template <class TM, class TV, int S>
__global__ __launch_bounds__(256) void k_csr_spmm(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                                  const TM* __restrict__ val, int64_t row0, int64_t row1,
                                                  int64_t col_off, const TV* __restrict__ X, int64_t ldx,
                                                  TV* __restrict__ Y, int64_t ldy, int64_t y_off, TV alpha,
                                                  int accumulate) {
    const int sub = (int)threadIdx.x & (S - 1);
    const int64_t row = row0 + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / S;
    const int64_t c = blockIdx.y;
    TV acc = 0;
    if (row < row1) {  // no return before the butterfly: every lane of a group takes part
        const TV* x = X + c * ldx + col_off;
        const int64_t b = ptr[row] + sub, e = ptr[row + 1];
        CsrLane<TM, TV> L;
        L.indices(col, val, b, e, S, true);
        L.gather(x);
        acc = L.sum(col, val, x, b, e, S, acc);
    }
#pragma unroll
    for (int off = S / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, S);
    if (row < row1 && sub == 0) {
        TV* y = Y + c * ldy + y_off + row;
        *y = accumulate ? *y + alpha * acc : alpha * acc;
    }
}

// Y[row, c] = W[row] x_c + alpha Q[row - p] x_{c, p..} for rows in [row0,
// row1) (the Q term for rows >= p only): W'W's rows and a CSR Q's in one pass
// over the rows, S lanes per (row, c). Each matrix keeps its own lane count
// (s1 <= S for W, S for Q) and butterfly, so the two sums and the final
// y = w + alpha q are the ones the separate kernels form, bit for bit. Both
// matrices' first terms are loaded before either's gathers.
//
// The first nlc workgroups instead form the chunk partials of W'W's long rows
// [0, row0) (k_csr_spmm_chunks' work, same sums), so that those rows' gathers
// run beside the short rows'; with lc_row1 the rows of one chunk are written
// here, the others are finished later from the partials.
template <class TM, class TV, int S, bool HAS_Q = true>
__global__ __launch_bounds__(256) void k_csr2_spmm(const int64_t* __restrict__ ptr1, const int32_t* __restrict__ col1,
                                                   const TM* __restrict__ val1, int s1,
                                                   const int64_t* __restrict__ ptr2, const int32_t* __restrict__ col2,
                                                   const TM* __restrict__ val2, int64_t p, int64_t row0,
                                                   int64_t row1, const TV* __restrict__ X, int64_t ldx,
                                                   TV* __restrict__ Y, int64_t ldy, TV alpha, int64_t nlc,
                                                   const int64_t* __restrict__ lc_beg,
                                                   const int64_t* __restrict__ lc_end, double* __restrict__ lpart,
                                                   const int32_t* __restrict__ lc_row1) {
    const int64_t c = blockIdx.y;
    const TV* x = X + c * ldx;
    if ((int64_t)blockIdx.x < nlc) {
        long_chunk_part<TM, TV>(blockIdx.x, nlc, lc_beg, lc_end, col1, val1, x, lpart + c * nlc, lc_row1,
                                Y + c * ldy);
        return;
    }
    const int sub = (int)threadIdx.x & (S - 1);
    const int64_t row = row0 + (((int64_t)blockIdx.x - nlc) * blockDim.x + threadIdx.x) / S;
    TV a1 = 0, a2 = 0;
    if (row < row1) {  // no return before the butterflies: every lane of a group takes part
        const int64_t b1 = ptr1[row] + sub, e1 = ptr1[row + 1];
        CsrLane<TM, TV> L1;
        L1.indices(col1, val1, b1, e1, s1, sub < s1);
        if (HAS_Q) {
            const bool has2 = row >= p;
            const int64_t r2 = has2 ? row - p : 0;
            const int64_t b2 = ptr2[r2] + sub, e2 = has2 ? ptr2[r2 + 1] : 0;
            CsrLane<TM, TV> L2;
            L2.indices(col2, val2, b2, e2, S, has2);
            L1.gather(x);
            L2.gather(x + p);
            a1 = L1.sum(col1, val1, x, b1, e1, s1, a1);
            a2 = L2.sum(col2, val2, x + p, b2, e2, S, a2);
        } else {
            L1.gather(x);
            a1 = L1.sum(col1, val1, x, b1, e1, s1, a1);
        }
    }
    for (int off = s1 / 2; off > 0; off >>= 1) a1 += __shfl_xor(a1, off, S);
    if (HAS_Q) {
#pragma unroll
        for (int off = S / 2; off > 0; off >>= 1) a2 += __shfl_xor(a2, off, S);
    }
    if (row < row1 && sub == 0) {
        TV* y = Y + c * ldy + row;
        const TV w = (TV)1 * a1;
        *y = HAS_Q && row >= p ? w + alpha * a2 : w;
    }
}

// lanes per row for a CSR block: the mean row length rounded down to a
// power of two, 2 when it lies in (1, 2), at most 16 (the 1M-animal A^-1,
// 6.0 nonzeros per row: 4 lanes 7.1 ms per PBLUP solve, 8 lanes 7.6 ms,
// 16 lanes 8.4 ms)
inline int csr_lanes(int64_t nnz, int64_t rows) {
    if (nnz <= rows) return 1;
    int S = 2;
    while (S < 16 && (int64_t)(2 * S) * rows <= nnz) S *= 2;
    return S;
}

template <class TM, class TV>
void csr_spmm(hipStream_t st, int S, int r, const int64_t* ptr, const int32_t* col, const TM* val, int64_t row0,
              int64_t row1, int64_t col_off, const TV* X, int64_t ldx, TV* Y, int64_t ldy, int64_t y_off, TV alpha,
              int accumulate) {
    const int64_t rows = row1 - row0;
    if (rows <= 0) return;
    const dim3 grid((unsigned)std::max<int64_t>(1, (rows * S + 255) / 256), (unsigned)r);
#define CB_CSR(S_) \
    k_csr_spmm<TM, TV, S_><<<grid, 256, 0, st>>>(ptr, col, val, row0, row1, col_off, X, ldx, Y, ldy, y_off, alpha, accumulate)
    switch (S) {
        case 1: CB_CSR(1); break;
        case 2: CB_CSR(2); break;
        case 4: CB_CSR(4); break;
        case 8: CB_CSR(8); break;
        default: CB_CSR(16); break;
    }
#undef CB_CSR
    CB_HIP(hipGetLastError());
}

template <class TM, class TV>
__global__ __launch_bounds__(256) void k_csr_spmm_chunks(const int64_t* __restrict__ cbeg,
                                                         const int64_t* __restrict__ cend,
                                                         int64_t nchunks, const int32_t* __restrict__ col,
                                                         const TM* __restrict__ val, const TV* __restrict__ X,
                                                         int64_t ldx, double* __restrict__ part) {
    const int64_t c = blockIdx.y;
    long_chunk_part<TM, TV>(blockIdx.x, nchunks, cbeg, cend, col, val, X + c * ldx, part + c * nchunks, nullptr,
                            nullptr);
}

// a wave per (row, column): lane l sums the row's chunks l, l + 64, ... in
// order, then a fixed butterfly over the wave (the mean row of W'W has
// hundreds of chunks, which one thread would add one dependent load at a time)
template <class TV>
__global__ __launch_bounds__(256) void k_csr_chunks_finish(const int64_t* __restrict__ rstart, int64_t rows,
                                                           int64_t nchunks, const double* __restrict__ part,
                                                           TV* __restrict__ Y, int64_t ldy) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t c = blockIdx.y;
    if (row >= rows) return;  // whole waves
    double s = 0;
    for (int64_t k = rstart[row] + lane; k < rstart[row + 1]; k += 64) s += part[c * nchunks + k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) Y[c * ldy + row] = (TV)s;
}

// MFMA 16x16x4 (FP32 or FP64). Both take lane l's operands as row (l & 15)
// of A and column (l & 15) of B at k-slot l >> 4. The accumulators differ:
// register j of lane l is element (4 * (l >> 4) + j, l & 15) of the 16 x 16
// tile for FP32 and ((l >> 4) + 4 j, l & 15) for FP64 (checked on gfx950).
template <class T>
struct Mfma16;
template <>
struct Mfma16<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int j) { return 4 * g + j; }
};
template <>
struct Mfma16<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int j) { return g + 4 * j; }
};

// Split-K partial of Q X on the matrix cores:
//   part[split][c][row] = sum_{k in split} Q[row,k] X[k,c]
// Q row-major q x q with leading dimension ldq (a multiple of 32, columns
// [q, ldq) zero); X column-major (ldx), columns [0, r), r <= 16 NT.
// 1-D grid: block b takes split b % nsplit of row tile b / nsplit, so
// consecutive blocks (which the dispatcher deals to different XCDs) share a
// row tile and each XCD keeps to a few splits of X in its L2.
// Wave w owns rows [32w, 32w + 32) of the 128-row tile as two 16-row
// subtiles. In each 32-k chunk, lane l holds row (l & 15) and the 8
// consecutive k of group g = l >> 4: its A operand is one contiguous run of a
// Q row loaded straight from global memory (Q is read exactly once), and MFMA
// step t multiplies k = kc + 8g + t. The X chunk is staged in LDS as
// [column][k] (double-buffered) and read in the same k order, which is all a
// product needs. EX further columns (r = 16 NT + EX) are done by VALU FMAs on
// the same A registers, each lane group summing its k-slice, the four slices
// added by lane shuffles at the end: cheaper than a 16-column MFMA tile for
// one useful column.
// A load of Q: streaming (non-temporal) when Q is too large to stay in the
// 256 MB Infinity Cache between launches, cacheable otherwise (a Q of up to
// Q_CACHE_BYTES is then re-read from the cache by the next solve iteration:
// at q = 5000 with FP16 halves, 0.044 -> 0.036 ms per product)
constexpr double Q_CACHE_BYTES = 128e6;
template <bool STREAM, class V>
__device__ __forceinline__ V load_q(const V* p) {
    if constexpr (STREAM) return __builtin_nontemporal_load(p);
    else return *p;
}

template <class T, int NT, int EX, bool STREAM>
__global__ __launch_bounds__(256) void k_dense_mfma(const T* __restrict__ Q, int64_t ldq, int64_t q,
                                                    const T* __restrict__ X, int64_t ldx, int r,
                                                    int64_t kspan, int nsplit, T* __restrict__ part,
                                                    int64_t ldp) {
    using V4 = typename Mfma16<T>::v4;
    constexpr int NC = NT * 16 + EX;                      // staged X columns
    constexpr int XE = (NC * MF_KC + 255) / 256;          // staged X elements per thread
    __shared__ T xs[2][NC][MF_KCP];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int split = (int)(blockIdx.x % (unsigned)nsplit);
    const int64_t row0 = (int64_t)(blockIdx.x / (unsigned)nsplit) * MF_ROWS;
    const int64_t k0 = (int64_t)split * kspan;
    const int64_t k1 = min(q, k0 + kspan);
    const int nchunks = (int)((k1 - k0 + MF_KC - 1) / MF_KC);
    const T* qa[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int64_t rr = min(row0 + w * 32 + s * 16 + li, q - 1);  // rows >= q: computed, not stored
        qa[s] = Q + rr * ldq + k0 + 8 * g;
    }
    T xr[XE], aa[2][8], ab[2][8];
    V4 acc[2][NT];
    T ex[2][EX > 0 ? EX : 1];  // the EX columns past the MFMA tiles, by VALU
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[s][n] = V4{0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < (EX > 0 ? EX : 1); ++e) ex[s][e] = 0;
    }
    // Unbranched loads: X indices clamped into the block (values outside
    // zeroed after the load), Q chunks clamped to the last one. A chunk past
    // the end therefore stages zeros and adds exactly nothing, so the loop
    // runs two chunks per trip with fixed register sets and no branches,
    // which keeps the compiler's wait counts exact.
    auto load_x = [&](int64_t kc) {
#pragma unroll
        for (int i = 0; i < XE; ++i) {
            const int e = tid + 256 * i, c = e >> 5;
            const int64_t k = kc + (e & 31);
            xr[i] = X[(int64_t)(c < r ? c : r - 1) * ldx + (k < k1 ? k : k1 - 1)];
        }
    };
    auto store_x = [&](int buf, int64_t kc) {
#pragma unroll
        for (int i = 0; i < XE; ++i) {
            const int e = tid + 256 * i, c = e >> 5;
            const int64_t k = kc + (e & 31);
            if ((NC * MF_KC) % 256 == 0 || c < NC) xs[buf][c][e & 31] = (c < r && k < k1) ? xr[i] : (T)0;
        }
    };
    // k beyond k1 reads the zero padding of Q (last split) and meets zero X
    auto load_a = [&](int ch, T (&dst)[2][8]) {
        const int off = (ch < nchunks ? ch : nchunks - 1) * MF_KC;
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 8; ++t) dst[s][t] = load_q<STREAM>(qa[s] + off + t);
    };
    // one chunk from xs[buf] and register set cur; the next chunk's loads
    // (into nxt) issue first, its X is staged into the other buffer after
    auto half = [&](int ch, int buf, const T (&cur)[2][8], T (&nxt)[2][8]) {
        load_x(k0 + (int64_t)(ch + 1) * MF_KC);
        load_a(ch + 1, nxt);
        T bx[NT][8];
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int t = 0; t < 8; ++t) bx[n][t] = xs[buf][n * 16 + li][8 * g + t];
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int n = 0; n < NT; ++n) acc[s][n] = Mfma16<T>::step(cur[s][t], bx[n][t], acc[s][n]);
#pragma unroll
        for (int e = 0; e < EX; ++e) {
            T xe[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) xe[t] = xs[buf][NT * 16 + e][8 * g + t];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 8; ++t) ex[s][e] = fma(cur[s][t], xe[t], ex[s][e]);
        }
        __builtin_amdgcn_sched_group_barrier(0x020, XE + 4, 0);  // the loads first,
        __builtin_amdgcn_sched_group_barrier(0x008, 16 * NT, 0); // then the MFMAs;
        __builtin_amdgcn_sched_barrier(0);  // the staging (which waits for X) after them
        store_x(buf ^ 1, k0 + (int64_t)(ch + 1) * MF_KC);
        __syncthreads();
    };
    load_x(k0);
    load_a(0, aa);
    store_x(0, k0);
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch += 2) {
        half(ch, 0, aa, ab);
        half(ch + 1, 1, ab, aa);
    }
    T* pp = part + (int64_t)split * r * ldp;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int c = n * 16 + li;
            if (c >= r) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t row = row0 + w * 32 + s * 16 + Mfma16<T>::row(g, j);
                if (row < q) pp[(int64_t)c * ldp + row] = acc[s][n][j];
            }
        }
    // extra columns: the four lane groups hold k-slices of the same rows
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int e = 0; e < EX; ++e) {
            T v = ex[s][e];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            const int c = NT * 16 + e;
            const int64_t row = row0 + w * 32 + s * 16 + li;
            if (g == 0 && c < r && row < q) pp[(int64_t)c * ldp + row] = v;
        }
}

// Split-FP16 variant of k_dense_mfma for the FP32 recursion of the mixed
// PCG. Q and X are each held as two FP16 halves, hi = fp16(v) and lo =
// fp16(v - hi), of power-of-two scaled values (Q by one scale, each X
// column by its own, so that the largest magnitude sits near 2^14): hi + lo
// carries 22 significant bits. Q X ~ Qh Xh + Qh Xl + Ql Xh on
// v_mfma_f32_16x16x32_f16 with FP32 accumulation (the dropped Ql Xl term is
// 2^-22 relative). Q's halves take the same 4 bytes per element as FP32, so
// the launch reads q^2 * 4 bytes, and three FP16 MFMAs do the work of 32
// FP32 ones: HBM-bound instead of MFMA-bound. Q's halves are stored in
// MFMA fragment order: fragment (strip, kb) = rows [16 strip, 16 strip + 16)
// x k [32 kb, 32 kb + 32) is one contiguous 2 KB run, the hi halves (1 KB)
// then the lo halves, each with lane l = li + 16 g's 8 k of group g of row
// li at 16 l bytes, and a strip's fragments follow each other along k
// (ldq / 32 per strip). A wave thus streams each of its subtiles
// sequentially, and each of its 16-byte loads covers whole 128-byte lines. The partials stay scaled; k_dense_reduce applies each column's
// power of two in FP64.
// Tiling as k_dense_mfma: 128-row tiles, split-K over a 1-D grid, wave w
// owns rows [32w, 32w + 32) as two 16-row subtiles; in each 32-k chunk lane
// l holds row (l & 15) and the 8 consecutive k of group g = l >> 4 (one
// 16-byte load per half), and its B operand is the same 8 k of column
// (l & 15) of the staged X chunk, so A and B agree on k slot by slot.
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int MH_KCP = MF_KC + 8;  // LDS pitch (halves) of a staged X column: 16-byte aligned runs
template <int NT, bool STREAM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_dense_h(const _Float16* __restrict__ Qhl, int64_t ldq, int64_t q,
                                                 const _Float16* __restrict__ Xh,
                                                 const _Float16* __restrict__ Xl, int64_t ldh, int r, int64_t kspan,
                                                 int nsplit, float* __restrict__ part, int64_t ldp) {
    constexpr int NC = NT * 16;
    constexpr int FRAG = 64 * 16;             // halves per fragment
    constexpr int XP = NC * MF_KC / 2;        // staged half pairs per buffer and half
    constexpr int XE = (XP + 255) / 256;      // pairs per thread
    __shared__ __attribute__((aligned(16))) _Float16 xs[2][2][NC][MH_KCP];  // [buf][hi, lo][column][k]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int split = (int)(blockIdx.x % (unsigned)nsplit);
    const int64_t row0 = (int64_t)(blockIdx.x / (unsigned)nsplit) * MF_ROWS;
    const int64_t k0 = (int64_t)split * kspan;
    const int64_t k1 = min(q, k0 + kspan);
    const int nchunks = (int)((k1 - k0 + MF_KC - 1) / MF_KC);
    int64_t qoff[2];
    const int64_t kb_row = ldq / MF_KC, last_strip = (q - 1) / 16;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        // strips past the last: computed from the last one, not stored
        const int64_t strip = min((row0 + w * 32 + s * 16) / 16, last_strip);
        qoff[s] = (strip * kb_row + k0 / MF_KC) * FRAG + (int64_t)lane * 8;
    }
    uint32_t xr[2][2][XE];                       // [register set][hi, lo][pair]
    h8v ah[3][2], al[3][2];                      // [register set][subtile]
    f4v acc[2][NT];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[s][n] = f4v{0, 0, 0, 0};
    // X half pairs (k, k+1) of column c; pairs at or past k1 stage zeros
    // (the padding rows of Xh/Xl past q are zero, so a pair that straddles
    // q = k1 reads zero for its second half)
    auto load_x = [&](int64_t kc, int set) {
#pragma unroll
        for (int i = 0; i < XE; ++i) {
            const int e = tid + 256 * i, c = e >> 4;
            const int64_t k = kc + 2 * (e & 15);
            const int64_t off = (int64_t)(c < r ? c : r - 1) * ldh + (k < k1 ? k : k0);
            xr[set][0][i] = *reinterpret_cast<const uint32_t*>(Xh + off);
            xr[set][1][i] = *reinterpret_cast<const uint32_t*>(Xl + off);
        }
    };
    auto store_x = [&](int buf, int64_t kc, int set) {
#pragma unroll
        for (int i = 0; i < XE; ++i) {
            const int e = tid + 256 * i, c = e >> 4;
            const int64_t k = kc + 2 * (e & 15);
            if (XP % 256 == 0 || c < NC) {
                const bool ok = c < r && k < k1;
#pragma unroll
                for (int hl = 0; hl < 2; ++hl)
                    *reinterpret_cast<uint32_t*>(&xs[buf][hl][c][2 * (e & 15)]) = ok ? xr[set][hl][i] : 0u;
            }
        }
    };
    // k beyond k1 reads the zero padding of Q (last split) and meets zero X
    auto load_a = [&](int ch, int set) {
        const int off = (ch < nchunks ? ch : nchunks - 1) * FRAG;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            ah[set][s] = load_q<STREAM>(reinterpret_cast<const h8v*>(Qhl + qoff[s] + off));
            al[set][s] = load_q<STREAM>(reinterpret_cast<const h8v*>(Qhl + qoff[s] + off + FRAG / 2));
        }
    };
    // chunk ch: its Q registers (set qs = ch % 3) were loaded two chunks
    // ahead, its X staged in xs[buf]; X of chunk ch + 1 (register set
    // xset ^ 1, loaded one chunk ago) is staged after the products, and X of
    // chunk ch + 2 loads into set xset: a Q load has two chunks to arrive, an
    // X load (from L2, X being shared by all row tiles) one
    auto half = [&](int ch, int buf, int qs, int xset) {
        load_a(ch + 2, (qs + 2) % 3);
        load_x(k0 + (int64_t)(ch + 2) * MF_KC, xset);
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const h8v xh = *reinterpret_cast<const h8v*>(&xs[buf][0][n * 16 + li][8 * g]);
            const h8v xl = *reinterpret_cast<const h8v*>(&xs[buf][1][n * 16 + li][8 * g]);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[qs][s], xh, acc[s][n], 0, 0, 0);
                acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[qs][s], xl, acc[s][n], 0, 0, 0);
                acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[qs][s], xh, acc[s][n], 0, 0, 0);
            }
        }
        // xs[buf ^ 1] was last read before the previous chunk's barrier
        store_x(buf ^ 1, k0 + (int64_t)(ch + 1) * MF_KC, xset ^ 1);
        __syncthreads();
    };
    load_x(k0, 0);
    load_a(0, 0);
    load_a(1, 1);
    store_x(0, k0, 0);
    load_x(k0 + MF_KC, 1);
    __syncthreads();
    // six chunks per trip keep every register set index static; chunks past
    // the split's end stage zero X (kspan is a multiple of 6 chunks except in
    // the last split)
    for (int ch = 0; ch < nchunks; ch += 6) {
        half(ch, 0, 0, 0);
        half(ch + 1, 1, 1, 1);
        half(ch + 2, 0, 2, 0);
        half(ch + 3, 1, 0, 1);
        half(ch + 4, 0, 1, 0);
        half(ch + 5, 1, 2, 1);
    }
    float* pp = part + (int64_t)split * r * ldp;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int c = n * 16 + li;
            if (c >= r) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t row = row0 + w * 32 + s * 16 + 4 * g + j;
                if (row < q) pp[(int64_t)c * ldp + row] = acc[s][n][j];
            }
        }
}

// Largest |Q| of n elements as float bits into *out (non-negative floats
// order as their bit patterns); *out starts at 0. A NaN counts as +inf, so
// it reaches the result (fmaxf alone would drop it).
__global__ __launch_bounds__(256) void k_absmax_bits(const float* __restrict__ a, int64_t n, unsigned* out) {
    __shared__ float red[4];
    float m = 0.0f;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const float v = fabsf(a[e]);
        m = fmaxf(m, v == v ? v : __builtin_inff());
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}

__device__ __forceinline__ void split_h(float v, _Float16* hi, _Float16* lo) {
    const _Float16 h = (_Float16)v;
    *hi = h;
    *lo = (_Float16)(v - (float)h);  // v - h is exact in FP32
}

// hi + lo = Q32 * 2^sq to 22 bits, written in k_dense_h's fragment order
// (ldq a multiple of 32; Q32 padding zeros stay zero, and rows past the
// last up to a multiple of 16 are zeroed by the caller)
__global__ void k_split_q(const float* __restrict__ Q, int64_t n, int64_t ldq, int sq, _Float16* __restrict__ Qhl) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int64_t i = e / ldq, k = e - i * ldq;
    const int64_t frag = (i >> 4) * (ldq / MF_KC) + (k >> 5);
    const int lane = (int)(i & 15) + 16 * (int)((k & 31) >> 3);
    _Float16* d = Qhl + frag * (64 * 16) + lane * 8 + (k & 7);
    split_h(ldexpf(Q[e], sq), d, d + 64 * 8);
}

// One block per column c of X (q rows): its largest |x| sets sx so that the
// column's values times 2^sx stay below 2^14; Xh + Xl = x * 2^sx, rows
// [q, ldh) zero; cscale[c] = 2^-(sq + sx) undoes both scales of Q X.
constexpr int SPLIT_X_THREADS = 1024;
__global__ __launch_bounds__(SPLIT_X_THREADS) void k_split_x(const float* __restrict__ X, int64_t ldx, int64_t q,
                                                             int64_t ldh, int sq, _Float16* __restrict__ Xh,
                                                             _Float16* __restrict__ Xl, double* __restrict__ cscale) {
    constexpr int NW = SPLIT_X_THREADS / 64;
    __shared__ float red[NW];
    const int c = blockIdx.x;
    const float* x = X + (int64_t)c * ldx;
    float m = 0.0f;
    for (int64_t i = threadIdx.x; i < q; i += SPLIT_X_THREADS) m = fmaxf(m, fabsf(x[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    m = red[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) m = fmaxf(m, red[w]);
    int e = 0;
    if (m > 0.0f && m <= 3.4e38f) (void)frexpf(m, &e);  // m < 2^e
    const int sx = m > 0.0f && m <= 3.4e38f ? 14 - e : 0;
    for (int64_t i = threadIdx.x; i < ldh; i += SPLIT_X_THREADS)
        split_h(i < q ? ldexpf(x[i], sx) : 0.0f, Xh + (int64_t)c * ldh + i, Xl + (int64_t)c * ldh + i);
    if (threadIdx.x == 0) cscale[c] = ldexp(1.0, -(sq + sx));
}

// Y[row, c] += alpha * sum_split part[split][c][row]   (fixed split order);
// with cscale set, the split sum is first scaled by cscale[c] (the split-FP16
// path's power of two); with sv set, then Y[row, c] += scoef * (sum of spart[c0 + c]'s chunks) *
// sv[row]: the stiff rank-1 term of apply_C, rounded exactly as its own
// kernel (k_stiff_axpy) rounds it.
template <class T>
__global__ void k_dense_reduce(const T* __restrict__ part, int64_t ldp, int nsplit, int64_t q, int r,
                               T* __restrict__ Y, int64_t ldy, int64_t y_off, T alpha,
                               const double* __restrict__ sv, const double* __restrict__ spart, int c0,
                               double scoef, const double* __restrict__ cscale = nullptr) {
    __shared__ double t;
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (sv) {
        if (threadIdx.x == 0) {
            double a = 0;
            for (int k = 0; k < DOT_CHUNKS; ++k) a += spart[(int64_t)(c0 + c) * DOT_CHUNKS + k];
            t = a;
        }
        __syncthreads();
    }
    if (row >= q) return;
    double s = 0;
    for (int sp = 0; sp < nsplit; ++sp) s += (double)part[((int64_t)sp * r + c) * ldp + row];
    if (cscale) s *= cscale[c];  // a power of two: exact
    T* y = Y + (int64_t)c * ldy + y_off + row;
    T v = (T)((double)*y + (double)alpha * s);
    if (sv) v = (T)((double)v + scoef * t * sv[row]);
    *y = v;
}

// 16-byte vector of T
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int n = 2;
};

// y_c[row] += alpha * Q[row,:] x_c for r <= 2 columns. Wave per GEMV_ROWS
// rows: each lane streams 16-byte runs of those rows (Q rows start 128-byte
// aligned, leading dimension a multiple of 32) and loads x once for all of
// them; FP64 accumulation. HBM-bound: q^2 * sizeof(T) bytes per launch.
// GEMV_PF runs of every row, with their x, are loaded before the first is
// used, so that a lane has GEMV_PF * GEMV_ROWS loads in flight (FP32 r = 2 at
// q = 20 000: 0.46 -> 0.24 ms; one run per trip left the FP64 conversions
// and FMAs of each run waiting on its own load).
constexpr int GEMV_ROWS = 4, GEMV_PF = 4;
template <class T>
__global__ __launch_bounds__(256) void k_gemv(const T* __restrict__ Q, int64_t ldq, int64_t q,
                                              const T* __restrict__ X, int64_t ldx, int r,
                                              T* __restrict__ Y, int64_t ldy, int64_t y_off, T alpha) {
    using VT = typename Vec16<T>::type;
    constexpr int VW = Vec16<T>::n;
    const int lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * GEMV_ROWS;
    if (row0 >= q) return;
    const VT* qr[GEMV_ROWS];
#pragma unroll
    for (int a = 0; a < GEMV_ROWS; ++a) qr[a] = (const VT*)(Q + min(row0 + a, q - 1) * ldq);
    double acc0[GEMV_ROWS], acc1[GEMV_ROWS];
#pragma unroll
    for (int a = 0; a < GEMV_ROWS; ++a) acc0[a] = acc1[a] = 0;
    const int64_t nv = q / VW;
    const bool two = r > 1;
    for (int64_t v0 = lane; v0 < nv; v0 += 64 * GEMV_PF) {
        VT w[GEMV_PF][GEMV_ROWS];
        VT x0[GEMV_PF], x1[GEMV_PF];
#pragma unroll
        for (int t = 0; t < GEMV_PF; ++t) {
            const int64_t v = min(v0 + 64 * t, nv - 1);  // a run past the end is loaded (clamped), not used
            // x runs as 16-byte loads: X sits at the random-effect offset p of
            // a block, so only element alignment is known (one unaligned
            // dwordx4 rather than VW dword loads per run and column)
            __builtin_memcpy(&x0[t], X + v * VW, sizeof(VT));
            if (two)
                __builtin_memcpy(&x1[t], X + ldx + v * VW, sizeof(VT));
            else
                x1[t] = VT{};
#pragma unroll
            for (int a = 0; a < GEMV_ROWS; ++a) w[t][a] = __builtin_nontemporal_load(qr[a] + v);
        }
#pragma unroll
        for (int t = 0; t < GEMV_PF; ++t) {
            if (v0 + 64 * t < nv) {
#pragma unroll
                for (int a = 0; a < GEMV_ROWS; ++a)
#pragma unroll
                    for (int e = 0; e < VW; ++e) {
                        acc0[a] += (double)w[t][a][e] * (double)x0[t][e];
                        acc1[a] += (double)w[t][a][e] * (double)x1[t][e];
                    }
            }
        }
    }
    for (int64_t k = nv * VW + lane; k < q; k += 64) {  // tail of rows whose length is not a multiple of VW
        const double x0 = (double)X[k], x1 = two ? (double)X[ldx + k] : 0.0;
#pragma unroll
        for (int a = 0; a < GEMV_ROWS; ++a) {
            const double w = (double)((const T*)qr[a])[k];
            acc0[a] += w * x0;
            acc1[a] += w * x1;
        }
    }
#pragma unroll
    for (int a = 0; a < GEMV_ROWS; ++a)
        for (int off = 32; off > 0; off >>= 1) {
            acc0[a] += __shfl_down(acc0[a], off);
            acc1[a] += __shfl_down(acc1[a], off);
        }
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < GEMV_ROWS; ++a) {
            const int64_t row = row0 + a;
            if (row >= q) break;
            Y[y_off + row] += alpha * (T)acc0[a];
            if (two) Y[ldy + y_off + row] += alpha * (T)acc1[a];
        }
    }
}

// part[c][chunk] = sum over the chunk's rows of A[row,c] * B[row,c] (FP64),
// NTH threads per chunk (1024 for long vectors, so that 128 chunks still
// fill the chip)
template <class T, int NTH = 256>
__global__ __launch_bounds__(NTH) void k_col_dots(const T* __restrict__ A, const T* __restrict__ B,
                                                  int64_t ld, int64_t rows,
                                                  double* __restrict__ part) {
    __shared__ double red[NTH];
    const int c = blockIdx.y;
    const int64_t per = (rows + DOT_CHUNKS - 1) / DOT_CHUNKS;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(rows, r0 + per);
    double s = 0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += NTH)
        s += (double)A[(int64_t)c * ld + i] * (double)B[(int64_t)c * ld + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)c * DOT_CHUNKS + blockIdx.x] = red[0];
}

// k_col_dots with the long rows of B = C P that span several chunks finished
// first: each such row (listed in multi) is summed from its chunk partials (a
// wave per row, the order of k_csr_chunks_finish) by the dot chunk whose row
// range holds it and written to B before that chunk's dot reads it. Rows of
// one chunk were written by the chunk pass.
template <class T, int NTH = 256>
__global__ __launch_bounds__(NTH) void k_col_dots_fin(const T* __restrict__ A, T* __restrict__ B, int64_t ld,
                                                      int64_t rows, double* __restrict__ part,
                                                      const int64_t* __restrict__ rstart, int64_t long_rows,
                                                      int64_t nchunks, const double* __restrict__ lpart,
                                                      const int64_t* __restrict__ multi, int64_t n_multi) {
    __shared__ double red[NTH];
    const int c = blockIdx.y;
    const int64_t per = (rows + DOT_CHUNKS - 1) / DOT_CHUNKS;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(rows, r0 + per);
    if (r0 < min(r1, long_rows)) {  // the same for the whole workgroup
        const int lane = threadIdx.x & 63;
        for (int64_t j = threadIdx.x >> 6; j < n_multi; j += NTH / 64) {
            const int64_t row = multi[j];
            if (row < r0 || row >= r1) continue;
            double s = 0;
            for (int64_t k = rstart[row] + lane; k < rstart[row + 1]; k += 64) s += lpart[c * nchunks + k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0) B[(int64_t)c * ld + row] = (T)s;
        }
        __syncthreads();
    }
    double s = 0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += NTH)
        s += (double)A[(int64_t)c * ld + i] * (double)B[(int64_t)c * ld + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)c * DOT_CHUNKS + blockIdx.x] = red[0];
}

// vectors at least this long take 1024 threads per dot chunk
constexpr int64_t DOT_LONG_ROWS = int64_t(1) << 18;

__global__ void k_dots_finish(const double* __restrict__ part, int r, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= r) return;
    double s = 0;
    for (int k = 0; k < DOT_CHUNKS; ++k) s += part[(int64_t)c * DOT_CHUNKS + k];
    out[c] = s;
}

// Jacobi part of the preconditioner: z = r * dinv for rows >= p0
template <class T>
__global__ void k_precond_diag(const T* __restrict__ R, T* __restrict__ Z, const double* __restrict__ dinv,
                               int64_t p0, int64_t rows, int64_t ld) {
    const int64_t i = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= rows) return;
    Z[(int64_t)c * ld + i] = (T)((double)R[(int64_t)c * ld + i] * dinv[i]);
}

// block part: z[0:p, c] = Binv r[0:p, c]  (Binv p x p row-major, FP64);
// wave per row, coalesced along the row
template <class T>
__global__ __launch_bounds__(256) void k_precond_block(const T* __restrict__ R, T* __restrict__ Z,
                                                       const double* __restrict__ Binv, int64_t p, int64_t ld) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int c = blockIdx.y;
    if (i >= p) return;
    const T* rc = R + (int64_t)c * ld;
    double s = 0;
    for (int64_t k = lane; k < p; k += 64) s += Binv[i * p + k] * (double)rc[k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) Z[(int64_t)c * ld + i] = (T)s;
}

// partial dots of each column's u-block with the stiff vector
template <class T>
__global__ __launch_bounds__(256) void k_stiff_dot(const T* __restrict__ Zu, const double* __restrict__ v,
                                                   int64_t ld, int64_t q, double* __restrict__ part) {
    __shared__ double red[256];
    const int c = blockIdx.y;
    const int64_t per = (q + DOT_CHUNKS - 1) / DOT_CHUNKS;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(q, r0 + per);
    double s = 0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) s += (double)Zu[(int64_t)c * ld + i] * v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)c * DOT_CHUNKS + blockIdx.x] = red[0];
}

// Zu[:, c] += coef * (sum of the column's partial dots) * y
template <class T>
__global__ void k_stiff_axpy(T* __restrict__ Zu, const double* __restrict__ y, const double* __restrict__ part,
                             double coef, int64_t ld, int64_t q) {
    __shared__ double t;
    __shared__ double sp[DOT_CHUNKS];
    const int c = blockIdx.y;
    if (threadIdx.x < DOT_CHUNKS) sp[threadIdx.x] = part[(int64_t)c * DOT_CHUNKS + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int k = 0; k < DOT_CHUNKS; ++k) s += sp[k];
        t = s;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q) return;
    Zu[(int64_t)c * ld + i] = (T)((double)Zu[(int64_t)c * ld + i] + coef * t * y[i]);
}

// Q32[i][j] = (float)(Q64[i][j] - mu v_i v_j)
__global__ void k_q32_bulk(const double* __restrict__ Q, float* __restrict__ Q32, const double* __restrict__ v,
                           double mu, int64_t q, int64_t ldq) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= q * ldq) return;
    const int64_t i = e / ldq, j = e - i * ldq;
    Q32[e] = j < q ? (float)(Q[e] - mu * v[i] * v[j]) : 0.0f;  // padding stays zero
}

// CG scalar state per column (kept on the device so the host need not sync
// every iteration). rho_new is the r.z of the current direction; rho the one
// before it, moved up by the next update of x and r.
struct ColState {
    double rho, pq, rho_new, rr, bb, alpha, beta;
    int active, restarted;
    int xr_act;  // x and r moved by alpha in this iteration (k_cg_xr_block, read by k_cg_p3)
};

// The scalar steps run inside the vector kernels that use them: thread 0 of
// every block sums the dot products' chunk partials in chunk order (as
// k_dots_finish does), so every block derives the same scalars, and block 0
// of each column writes them back. A field the kernel writes is never one
// its blocks read for their scalars, so the order of blocks does not matter.
//   k_col_dots2      the partials of A.B and A.C in one pass over A
//   k_cg_xr_diag     alpha = rho_new / p.q (a breakdown freezes the column),
//                    x += alpha p, r -= alpha q, then the Jacobi rows of z
//   k_cg_p           rho_new = r.z, rr = r.r, convergence, beta, p = z + beta p
template <class T, int NTH = 256>
__global__ __launch_bounds__(NTH) void k_col_dots2(const T* __restrict__ A, const T* __restrict__ B,
                                                   const T* __restrict__ C, int64_t ld, int64_t rows,
                                                   double* __restrict__ part_ab, double* __restrict__ part_ac) {
    __shared__ double red[2][NTH];
    const int c = blockIdx.y;
    const int64_t per = (rows + DOT_CHUNKS - 1) / DOT_CHUNKS;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(rows, r0 + per);
    double s = 0, u = 0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += NTH) {
        const double a = (double)A[(int64_t)c * ld + i];
        s += a * (double)B[(int64_t)c * ld + i];
        u += a * (double)C[(int64_t)c * ld + i];
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = u;
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_ab[(int64_t)c * DOT_CHUNKS + blockIdx.x] = red[0][0];
        part_ac[(int64_t)c * DOT_CHUNKS + blockIdx.x] = red[1][0];
    }
}

// the partials of column c into LDS by DOT_CHUNKS threads at once (one round
// trip instead of a chain of dependent loads); thread 0 then sums them in
// chunk order
__device__ __forceinline__ void stage_chunks(const double* __restrict__ part, int c, double* __restrict__ sp,
                                             int t) {
    if (t < DOT_CHUNKS) sp[t] = part[(int64_t)c * DOT_CHUNKS + t];
}
__device__ __forceinline__ double sum_staged(const double* __restrict__ sp) {
    double s = 0;
    for (int k = 0; k < DOT_CHUNKS; ++k) s += sp[k];
    return s;
}

// rho_new = r.z, rr = r.r, convergence (or NaN) and beta from the partials;
// active columns: p = z + beta p
template <class T>
__global__ __launch_bounds__(256) void k_cg_p(T* __restrict__ P, const T* __restrict__ Z,
                                              ColState* __restrict__ st, const double* __restrict__ part_rz,
                                              const double* __restrict__ part_rr, double tol2, int absolute,
                                              int64_t rows, int64_t ld) {
    __shared__ double s_beta;
    __shared__ int s_act;
    __shared__ double sp[2][DOT_CHUNKS];
    const int c = blockIdx.y;
    // CG_RPT rows per thread (rows base + j 256 + t); their loads go out
    // before the scalar step, which they do not need
    const int64_t base = (int64_t)blockIdx.x * (256 * CG_RPT) + threadIdx.x;
    T zv[CG_RPT], pv[CG_RPT];
#pragma unroll
    for (int j = 0; j < CG_RPT; ++j) {
        const int64_t i = base + j * 256;
        const int64_t o = (int64_t)c * ld + (i < rows ? i : 0);
        zv[j] = Z[o];
        pv[j] = P[o];
    }
    static_assert(2 * DOT_CHUNKS <= 256, "one partial per thread");
    if (threadIdx.x < DOT_CHUNKS) stage_chunks(part_rz, c, sp[0], threadIdx.x);
    else stage_chunks(part_rr, c, sp[1], threadIdx.x - DOT_CHUNKS);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double rho = st[c].rho, bb = st[c].bb;
        int act = st[c].active;
        double beta = 0.0;
        if (act) {
            const double rho_new = sum_staged(sp[0]), rr = sum_staged(sp[1]);
            const double thr = absolute ? tol2 : tol2 * bb;
            if (!(rr >= thr)) act = 0;  // converged (or NaN)
            beta = rho != 0.0 ? rho_new / rho : 0.0;
            if (blockIdx.x == 0) {
                st[c].rho_new = rho_new;
                st[c].rr = rr;
                st[c].beta = beta;
                st[c].active = act;
            }
        }
        s_beta = beta;
        s_act = act;
    }
    __syncthreads();
    if (!s_act) return;
#pragma unroll
    for (int j = 0; j < CG_RPT; ++j) {
        const int64_t i = base + j * 256;
        if (i < rows) P[(int64_t)c * ld + i] = (T)((double)zv[j] + s_beta * (double)pv[j]);
    }
}

// alpha = rho_new / p.q from the partials (a breakdown freezes the column);
// active columns: x += alpha p, r -= alpha q; every column: z = r dinv on
// the rows >= p0 (the block-Jacobi rows < p0 follow in their own kernel).
// A workgroup covers 256 CG_RPT rows. With px, each workgroup also writes its
// sums of r.z over its rows >= p0 and of r.r over all its rows (each
// thread's rows in order, then a fixed tree over the workgroup) to px[c * 2 nx +
// blockIdx.x] and px[c * 2 nx + nx + blockIdx.x] (nx = gridDim.x), from the
// values stored in R and Z.
template <class TX, class T>
__global__ __launch_bounds__(256) void k_cg_xr_diag(TX* __restrict__ X, T* __restrict__ R, const T* __restrict__ P,
                                                    const T* __restrict__ Qp, T* __restrict__ Z,
                                                    const double* __restrict__ dinv, ColState* __restrict__ st,
                                                    const double* __restrict__ part, int64_t p0, int64_t rows,
                                                    int64_t ld, double* __restrict__ px) {
    __shared__ double s_alpha;
    __shared__ int s_act;
    __shared__ double sp[DOT_CHUNKS];
    __shared__ double red[2][256];
    const int c = blockIdx.y;
    // CG_RPT rows per thread (rows base + j 256 + t); their loads go out
    // before the scalar step, which they do not need
    const int64_t base = (int64_t)blockIdx.x * (256 * CG_RPT) + threadIdx.x;
    TX xv[CG_RPT];
    T rv0[CG_RPT], pv[CG_RPT], qv[CG_RPT];
    double dv[CG_RPT];
#pragma unroll
    for (int j = 0; j < CG_RPT; ++j) {
        const int64_t i = base + j * 256, ii = i < rows ? i : 0;
        const int64_t o = (int64_t)c * ld + ii;
        xv[j] = X[o];
        rv0[j] = R[o];
        pv[j] = P[o];
        qv[j] = Qp[o];
        dv[j] = dinv[ii];
    }
    stage_chunks(part, c, sp, threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double rho = st[c].rho_new;
        int act = st[c].active;
        const double pq = sum_staged(sp);
        if (act && fabs(pq) < 1e-300) act = 0;  // breakdown: freeze the column
        const double alpha = act ? rho / pq : 0.0;
        if (blockIdx.x == 0) {
            st[c].pq = pq;
            st[c].alpha = alpha;
            st[c].active = act;
            st[c].rho = rho;
        }
        s_alpha = alpha;
        s_act = act;
    }
    __syncthreads();
    double rz = 0, rr = 0;
#pragma unroll
    for (int j = 0; j < CG_RPT; ++j) {
        const int64_t i = base + j * 256;
        if (i >= rows) break;
        const int64_t o = (int64_t)c * ld + i;
        T rv = rv0[j];
        if (s_act) {
            const double a = s_alpha;
            X[o] = (TX)((double)xv[j] + a * (double)pv[j]);
            rv = (T)((double)rv - a * (double)qv[j]);
            R[o] = rv;
        }
        rr += (double)rv * (double)rv;
        if (i >= p0) {
            const T zv = (T)((double)rv * dv[j]);
            Z[o] = zv;
            rz += (double)rv * (double)zv;
        }
    }
    if (!px) return;  // the same for every workgroup
    red[0][threadIdx.x] = rz;
    red[1][threadIdx.x] = rr;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        px[(int64_t)c * 2 * gridDim.x + blockIdx.x] = red[0][0];
        px[(int64_t)c * 2 * gridDim.x + gridDim.x + blockIdx.x] = red[1][0];
    }
}

// z[0:p, c] = Binv r[0:p, c] for the block rows and the column's r.z and r.r
// chunk partials for k_cg_p, in DOT_CHUNKS workgroups of NTH threads: a
// wave per block row (r[0:p] staged in LDS when p <= PB_LDS; a lane loads
// PB_U of its row's entries at once, into four running sums), and
// workgroup b also adds the b-th of DOT_CHUNKS contiguous runs of
// k_cg_xr_diag's nx workgroup partials (px: r.z of rows >= p, then r.r).
// part_rz[c][b] = that run's r.z partials + r_i z_i of the workgroup's block
// rows, part_rr[c][b] = the run's r.r partials; fixed order throughout.
template <class T, int NTH>
__global__ __launch_bounds__(NTH) void k_precond_block_dots(const T* __restrict__ R, T* __restrict__ Z,
                                                            const double* __restrict__ Binv, int64_t p, int64_t ld,
                                                            const double* __restrict__ px, int64_t nx,
                                                            double* __restrict__ part_rz,
                                                            double* __restrict__ part_rr) {
    constexpr int W = NTH / 64;
    __shared__ double red[2][NTH];
    __shared__ double rs[PB_LDS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = blockIdx.y;
    const T* rc = R + (int64_t)c * ld;
    const bool staged = p <= PB_LDS && (int64_t)blockIdx.x * W < p;  // the same for the whole workgroup
    if (staged) {
        for (int64_t k = threadIdx.x; k < p; k += NTH) rs[k] = (double)rc[k];
        __syncthreads();
    }
    double a = 0, u = 0;
    for (int64_t i = (int64_t)blockIdx.x * W + wave; i < p; i += (int64_t)DOT_CHUNKS * W) {
        const double* b = Binv + i * p;
        double sum[4] = {0, 0, 0, 0};
        for (int64_t k0 = lane; k0 < p; k0 += 64 * PB_U) {
            double bv[PB_U], rv[PB_U];
#pragma unroll
            for (int j = 0; j < PB_U; ++j) {
                const int64_t k = k0 + 64 * j;
                bv[j] = k < p ? b[k] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < PB_U; ++j) {
                const int64_t k = k0 + 64 * j;
                rv[j] = k < p ? (staged ? rs[k] : (double)rc[k]) : 0.0;
            }
#pragma unroll
            for (int j = 0; j < PB_U; ++j) sum[j & 3] += bv[j] * rv[j];
        }
        double s = (sum[0] + sum[1]) + (sum[2] + sum[3]);
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
        if (lane == 0) {
            const T z = (T)s;
            Z[(int64_t)c * ld + i] = z;
            a += (double)rc[i] * (double)z;
        }
    }
    const int64_t per = (nx + DOT_CHUNKS - 1) / DOT_CHUNKS;
    const int64_t k0 = (int64_t)blockIdx.x * per, k1 = min(nx, k0 + per);
    const double* pc = px + (int64_t)c * 2 * nx;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += NTH) {
        a += pc[k];
        u += pc[nx + k];
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = u;
    __syncthreads();
    for (int w = NTH / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_rz[(int64_t)c * DOT_CHUNKS + blockIdx.x] = red[0][0];
        part_rr[(int64_t)c * DOT_CHUNKS + blockIdx.x] = red[1][0];
    }
}

// The CG step after the p.q dot when the block-Jacobi block (if any) has p0
// <= PB_LDS rows: k_cg_xr_diag, the preconditioner's block and stiff parts
// and the r.z / r.r dots in one launch. Grid (nbw + nx, r), 256 threads,
// dynamic LDS of p0 doubles. Every workgroup forms alpha from the dot's
// chunk partials as k_cg_xr_diag does.
//  - workgroups nbw.. (CG_RPT rows per thread, loads before the scalar
//    step): x += alpha p on every row; on the rows >= p0 also r -= alpha q,
//    z = r dinv, and the partials of r.z and r.r; with STIFF also of v.z
//    and r.y over those rows (v the stiff direction, y = D'^-1 v);
//  - workgroups 0..nbw-1: the block rows' new r (r - alpha q) staged in LDS,
//    then z = Binv r for the rows i = 4 b + wave, + 4 nbw, ..., summed as
//    k_precond_block_dots sums them, and those rows' r.z and r.r. The block
//    rows' r stays unwritten here (every block workgroup reads the old one);
//    k_cg_p3 writes it, with the same alpha.
// px[c][NS ntot] (NS = 2, or 4 with STIFF): r.z partials of the ntot =
// gridDim.x workgroups, then r.r, then v.z and r.y.
template <class TX, class T, bool STIFF>
__global__ __launch_bounds__(256) void k_cg_xr_block(TX* __restrict__ X, T* __restrict__ R, const T* __restrict__ P,
                                                     const T* __restrict__ Qp, T* __restrict__ Z,
                                                     const double* __restrict__ dinv, const double* __restrict__ Binv,
                                                     const double* __restrict__ sv, const double* __restrict__ sy,
                                                     ColState* __restrict__ st, const double* __restrict__ part,
                                                     int64_t p0, int64_t rows, int64_t ld, int64_t nbw,
                                                     double* __restrict__ px) {
    constexpr int NS = STIFF ? 4 : 2;
    extern __shared__ double rs[];  // p0 doubles
    __shared__ double red[NS][256];
    __shared__ double sp[DOT_CHUNKS];
    __shared__ double s_alpha;
    __shared__ int s_act;
    const int c = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool blk = (int64_t)blockIdx.x < nbw;  // the same for the whole workgroup
    const int64_t ntot = gridDim.x;
    const int64_t base = ((int64_t)blockIdx.x - nbw) * (256 * CG_RPT) + tid;
    TX xv[CG_RPT];
    T rv0[CG_RPT], pv[CG_RPT], qv[CG_RPT];
    double dv[CG_RPT], vv[CG_RPT], yv[CG_RPT];
    if (!blk) {
#pragma unroll
        for (int j = 0; j < CG_RPT; ++j) {
            const int64_t i = base + j * 256, ii = i < rows ? i : 0;
            const int64_t o = (int64_t)c * ld + ii;
            xv[j] = X[o];
            pv[j] = P[o];
            rv0[j] = R[o];
            qv[j] = Qp[o];
            dv[j] = dinv[ii];
            if (STIFF) {
                const int64_t u = ii >= p0 ? ii - p0 : 0;
                vv[j] = sv[u];
                yv[j] = sy[u];
            }
        }
    }
    stage_chunks(part, c, sp, tid);
    __syncthreads();
    if (tid == 0) {
        const double rho = st[c].rho_new;
        int act = st[c].active;
        const double pq = sum_staged(sp);
        if (act && fabs(pq) < 1e-300) act = 0;  // breakdown: freeze the column
        const double alpha = act ? rho / pq : 0.0;
        if (blockIdx.x == 0) {
            st[c].pq = pq;
            st[c].alpha = alpha;
            st[c].active = act;
            st[c].rho = rho;
            st[c].xr_act = act;
        }
        s_alpha = alpha;
        s_act = act;
    }
    __syncthreads();
    const double a = s_alpha;
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0;  // r.z, r.r (, v.z, r.y)
    if (!blk) {
#pragma unroll
        for (int j = 0; j < CG_RPT; ++j) {
            const int64_t i = base + j * 256;
            if (i >= rows) break;
            const int64_t o = (int64_t)c * ld + i;
            if (s_act) X[o] = (TX)((double)xv[j] + a * (double)pv[j]);
            if (i >= p0) {
                T rv = rv0[j];
                if (s_act) {
                    rv = (T)((double)rv - a * (double)qv[j]);
                    R[o] = rv;
                }
                acc[1] += (double)rv * (double)rv;
                const T zv = (T)((double)rv * dv[j]);
                Z[o] = zv;
                acc[0] += (double)rv * (double)zv;
                if (STIFF) {
                    acc[NS - 2] += (double)zv * vv[j];
                    acc[NS - 1] += (double)rv * yv[j];
                }
            }
        }
    } else {
        const T* rc = R + (int64_t)c * ld;
        const T* qc = Qp + (int64_t)c * ld;
        for (int64_t k = tid; k < p0; k += 256) {
            const double r0 = (double)rc[k];
            rs[k] = s_act ? (double)(T)(r0 - a * (double)qc[k]) : r0;
        }
        __syncthreads();
        for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < p0; i += nbw * 4) {
            const double* b = Binv + i * p0;
            double sum[4] = {0, 0, 0, 0};
            for (int64_t k0 = lane; k0 < p0; k0 += 64 * PB_U) {
                double bv[PB_U], rv[PB_U];
#pragma unroll
                for (int j = 0; j < PB_U; ++j) {
                    const int64_t k = k0 + 64 * j;
                    bv[j] = k < p0 ? b[k] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < PB_U; ++j) {
                    const int64_t k = k0 + 64 * j;
                    rv[j] = k < p0 ? rs[k] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < PB_U; ++j) sum[j & 3] += bv[j] * rv[j];
            }
            double t = (sum[0] + sum[1]) + (sum[2] + sum[3]);
            for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off);
            if (lane == 0) {
                const T z = (T)t;
                Z[(int64_t)c * ld + i] = z;
                acc[0] += rs[i] * (double)z;
                acc[1] += rs[i] * rs[i];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) red[k][tid] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
#pragma unroll
            for (int k = 0; k < NS; ++k) red[k][tid] += red[k][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NS; ++k) px[((int64_t)c * NS + k) * ntot + blockIdx.x] = red[k][0];
    }
}

// After k_cg_xr_block: rho_new = r.z and rr = r.r from its ntot partials
// (each thread its strided run in order, then a fixed tree), convergence and
// beta as k_cg_p; the block rows' r moved by that kernel's alpha (r - alpha
// q, the value it used); active columns: p = z + beta p. With STIFF the
// preconditioner's rank-1 part applies here: z_u -= gamma (v.z_u) y on the
// rows >= p0 (written back to Z), and r.z is r.z_u - gamma (v.z_u) (r.y)
// plus the block rows' r.z.
template <class T, bool STIFF>
__global__ __launch_bounds__(256) void k_cg_p3(T* __restrict__ P, T* __restrict__ Z, T* __restrict__ R,
                                               const T* __restrict__ Qp, const double* __restrict__ sy,
                                               double gamma, ColState* __restrict__ st,
                                               const double* __restrict__ px, int64_t ntot, double tol2,
                                               int absolute, int64_t p0, int64_t rows, int64_t ld) {
    constexpr int NS = STIFF ? 4 : 2;
    __shared__ double red[NS][256];
    __shared__ double s_beta, s_alpha, s_t;
    __shared__ int s_act, s_xr;
    const int c = blockIdx.y;
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * (256 * CG_RPT) + tid;
    const bool has_block = (int64_t)blockIdx.x * (256 * CG_RPT) < p0;  // the same for the whole workgroup
    T zv[CG_RPT], pv[CG_RPT], rb[CG_RPT], qb[CG_RPT];
    double yv[CG_RPT];
#pragma unroll
    for (int j = 0; j < CG_RPT; ++j) {
        const int64_t i = base + j * 256, ii = i < rows ? i : 0;
        const int64_t o = (int64_t)c * ld + ii;
        zv[j] = Z[o];
        pv[j] = P[o];
        if (STIFF) yv[j] = sy[ii >= p0 ? ii - p0 : 0];
        if (has_block) {
            rb[j] = R[o];
            qb[j] = Qp[o];
        }
    }
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0;
    for (int64_t k = tid; k < ntot; k += 256) {
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[q] += px[((int64_t)c * NS + q) * ntot + k];
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) red[k][tid] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
#pragma unroll
            for (int k = 0; k < NS; ++k) red[k][tid] += red[k][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double rho = st[c].rho, bb = st[c].bb;
        int act = st[c].active;
        double beta = 0.0;
        const double t = STIFF ? red[NS - 2][0] : 0.0;  // v.z_u
        if (act) {
            const double rho_new = STIFF ? red[0][0] - gamma * t * red[NS - 1][0] : red[0][0];
            const double rr = red[1][0];
            const double thr = absolute ? tol2 : tol2 * bb;
            if (!(rr >= thr)) act = 0;  // converged (or NaN)
            beta = rho != 0.0 ? rho_new / rho : 0.0;
            if (blockIdx.x == 0) {
                st[c].rho_new = rho_new;
                st[c].rr = rr;
                st[c].beta = beta;
                st[c].active = act;
            }
        }
        s_beta = beta;
        s_act = act;
        s_t = t;
        s_alpha = st[c].alpha;  // this kernel writes neither alpha nor xr_act
        s_xr = st[c].xr_act;
    }
    __syncthreads();
    if (has_block && s_xr) {
#pragma unroll
        for (int j = 0; j < CG_RPT; ++j) {
            const int64_t i = base + j * 256;
            if (i < p0) R[(int64_t)c * ld + i] = (T)((double)rb[j] - s_alpha * (double)qb[j]);
        }
    }
    if (!s_act) return;
#pragma unroll
    for (int j = 0; j < CG_RPT; ++j) {
        const int64_t i = base + j * 256;
        if (i >= rows) continue;
        T z = zv[j];
        if (STIFF && i >= p0) {
            z = (T)((double)z - gamma * s_t * yv[j]);
            Z[(int64_t)c * ld + i] = z;
        }
        P[(int64_t)c * ld + i] = (T)((double)z + s_beta * (double)pv[j]);
    }
}

// R = B - A   (column-major, same ld)
template <class TB, class T>
__global__ void k_sub(const TB* __restrict__ B, const T* __restrict__ A, T* __restrict__ R, int64_t rows,
                      int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= rows) return;
    const int64_t o = (int64_t)c * ld + i;
    R[o] = (T)((double)B[o] - (double)A[o]);
}

template <class TD, class TS>
__global__ void k_convert(const TS* __restrict__ s, TD* __restrict__ d, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (TD)s[i];
}

// Rademacher probes: z[c][j] = +-1 from splitmix64(seed + golden*(c*q + j + 1));
// the sign is the top bit of the mixed word. Written at rows [off, off+q).
__device__ __host__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void k_probes(double* __restrict__ Z, int64_t ld, int64_t off, int64_t q, int k, uint64_t seed,
                         int exact, int64_t first) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (j >= q) return;
    double v;
    if (exact) {
        v = (first + c == j) ? 1.0 : 0.0;
    } else {
        const uint64_t h = splitmix64(seed + 0x9E3779B97F4A7C15ull * ((uint64_t)(first + c) * (uint64_t)q + (uint64_t)j + 1ull));
        v = (h >> 63) ? -1.0 : 1.0;
    }
    Z[(int64_t)c * ld + off + j] = v;
}

// B[:, c] = e_{first + c} (B zeroed by the caller)
__global__ void k_unit_cols(double* __restrict__ B, int64_t ld, int64_t first, int r) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < r) B[(int64_t)c * ld + first + c] = 1.0;
}

// synthetic operands for the tuner: uniform in [-0.5, 0.5) from a hash of the index
template <class T>
__global__ void k_tune_fill(T* __restrict__ a, int64_t n, uint64_t seed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (T)((double)(splitmix64(seed + (uint64_t)i) >> 11) * 0x1.0p-53 - 0.5);
}

// t = W s (records): t[r] = sum x_val * s[x_col] + (aid>=0 ? s[p+aid] : 0)
__global__ void k_w_apply(const int64_t* __restrict__ xp, const int32_t* __restrict__ xc,
                          const double* __restrict__ xv, const int32_t* __restrict__ aid, int64_t n,
                          int64_t p, const double* __restrict__ S, int64_t lds, double* __restrict__ T,
                          int64_t ldt) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= n) return;
    const double* s = S + (int64_t)c * lds;
    double t = 0;
    for (int64_t k = xp[r]; k < xp[r + 1]; ++k) t += xv[k] * s[xc[k]];
    if (aid[r] >= 0) t += s[p + aid[r]];
    T[(int64_t)c * ldt + r] = t;
}

// out = W' v via the transposed design (CSR dim x n), a wave per row: the
// fixed-effect rows hold a record each of a level, the animal rows a few
template <int LANES>
__global__ __launch_bounds__(256) void k_wt_apply(const int64_t* __restrict__ tp, const int32_t* __restrict__ tc,
                                                  const double* __restrict__ tv, int64_t dim,
                                                  const double* __restrict__ V, int64_t ldv, double* __restrict__ O,
                                                  int64_t ldo) {
    const int lane = threadIdx.x % LANES;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
    const int c = blockIdx.y;
    if (i >= dim) return;
    const double* v = V + (int64_t)c * ldv;
    double s = 0;
    for (int64_t k = tp[i] + lane; k < tp[i + 1]; k += LANES) s += tv[k] * v[tc[k]];
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LANES);
    if (lane == 0) O[(int64_t)c * ldo + i] = s;
}

// AI phase vectors: vu[r] = u_hat[aid[r]] / var_u ; ve[r] = e[r] / var_e
__global__ void k_ai_vectors(const int32_t* __restrict__ aid, int64_t n, int64_t p,
                             const double* __restrict__ sol, const double* __restrict__ e,
                             double inv_vu, double inv_ve, double* __restrict__ V, int64_t ldv) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    V[r] = aid[r] >= 0 ? sol[p + aid[r]] * inv_vu : 0.0;
    V[ldv + r] = e[r] * inv_ve;
}

// The REML score's data terms from one block of MME solutions, in one pass
// (vce.cpp:2655-2675 uses u'Qu, e'e and tr(Q C^uu)). Chunk partials in the
// layout of k_col_dots, block (chunk, column):
//   with_y, column 0   e = y - W s_0 over the n records (stored), e'e
//   with_y, column 1   u'Qu with u = s_0[p:], Qu given
//   trace column j     z_j' Q C^uu z_j = (Q z_j)' s_{t0+j}[p:]
// (t0 = 1 with the y column, else 0). Each partial sums its chunk's rows in
// thread-strided order and a fixed tree, so the terms do not depend on the
// launch.
__global__ __launch_bounds__(256) void k_reml_score_terms(
    const int64_t* __restrict__ xp, const int32_t* __restrict__ xc, const double* __restrict__ xv,
    const int32_t* __restrict__ aid, const double* __restrict__ y, int64_t n, int64_t p, int64_t q,
    const double* __restrict__ S, int64_t ld, const double* __restrict__ Qu, const double* __restrict__ Qz,
    int with_y, double* __restrict__ e, double* __restrict__ part) {
    __shared__ double red[256];
    const int c = blockIdx.y;
    const bool resid = with_y && c == 0;
    const int64_t rows = resid ? n : q;
    const int64_t per = (rows + DOT_CHUNKS - 1) / DOT_CHUNKS;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(rows, r0 + per);
    double acc = 0;
    if (resid) {
        for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
            double t = 0;
            for (int64_t k = xp[r]; k < xp[r + 1]; ++k) t += xv[k] * S[xc[k]];
            if (aid[r] >= 0) t += S[p + aid[r]];
            const double ev = y[r] - t;
            e[r] = ev;
            acc += ev * ev;
        }
    } else {
        const double* a;
        const double* b;
        if (with_y && c == 1) {
            a = S + p;
            b = Qu + p;
        } else {
            const int j = c - (with_y ? 2 : 0);
            a = S + (int64_t)(j + (with_y ? 1 : 0)) * ld + p;
            b = Qz + (int64_t)j * ld + p;
        }
        for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) acc += a[i] * b[i];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)c * DOT_CHUNKS + blockIdx.x] = red[0];
}

inline unsigned blocks(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, ceil_div(n, t)); }

}  // namespace

// ---------------------------------------------------------------------------
// The device scratch of cb_mme_solve_device_noalloc: one array of
// NOALLOC_SCRATCH_BYTES on every GPU the library runs on, held until the
// process ends. Host code never uses the array's name as an address: each
// device's address comes from hipGetSymbolAddress, looked up once per
// device with that device current (at cb_mme_create) and kept in
// g_scratch_ptr.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;  // device -> address of its scratch

void init_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw cbamd::Error(CB_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
}

unsigned char* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    return it == g_scratch_ptr.end() ? nullptr : it->second;
}
}  // namespace

// Scratch of an allocation-free call: sub-buffers carved from one fixed
// device region, one per key, 256-byte aligned and reused by the same key
// within the call. Running out is an error, never an allocation.
struct ScratchArena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    struct Slot {
        int key;
        char* p;
        size_t bytes;
    };
    std::vector<Slot> slots;
    char* get(int key, size_t bytes) {
        for (Slot& s : slots)
            if (s.key == key && s.bytes >= bytes) return s.p;
        const size_t at = (size_t)round_up((int64_t)used, 256);
        if (at + bytes > cap)
            throw cbamd::Error(CB_ERR_INVALID, "allocation-free call: its scratch (" + std::to_string(cap) +
                                                   " bytes) is too small for this size");
        for (Slot& s : slots)
            if (s.key == key) {
                s.p = base + at;
                s.bytes = bytes;
                used = at + bytes;
                return base + at;
            }
        slots.push_back({key, base + at, bytes});
        used = at + bytes;
        return base + at;
    }
};

struct cb_mme {
    ScratchArena* arena = nullptr;  // set for the duration of an allocation-free call
    cb_context* ctx = nullptr;
    int64_t n = 0, p = 0, q = 0, dim = 0, ldv = 0;
    // design (device): W rows (records), W^T rows (unknowns)
    DevBuf<int64_t> xp, tp;
    DevBuf<int32_t> xc, tc, aid;
    DevBuf<double> xv, tv, y;
    // W'W CSR
    DevBuf<int64_t> wp;
    DevBuf<int32_t> wc;
    DevBuf<double> wv64;
    DevBuf<float> wv32;
    int64_t long_rows = 0;  // rows [0, long_rows) use the chunked two-stage SpMM
    int ww_lanes = 1, q_lanes = 1;  // lanes per row of the W'W rows past long_rows and of a CSR Q
    // chunks of at most LONG_CHUNK nonzeros of the long rows, in row order
    int64_t n_lchunks = 0;
    DevBuf<int64_t> lc_beg, lc_end, lr_start;
    DevBuf<int32_t> lc_row1;   // per chunk: its row when that row is one chunk, else -1
    DevBuf<int64_t> lr_multi;  // the long rows of more than one chunk
    int64_t n_multi = 0;
    DevBuf<double> lc_part;
    std::vector<double> wtw_diag;  // diagonal of W'W
    std::vector<double> xtx;       // p x p block (constant in lambda)
    std::vector<double> rhs;       // W'y
    double yty = 0.0;
    // Q
    int qkind = -1;  // 0 dense, 1 CSR, 2 split (CSR A^-1 + dense block on the genotyped rows)
    int64_t ng = 0;                // split: genotyped rows
    DevBuf<int32_t> gmap;          // split: row of Q for each genotyped row (-1: none)
    DevBuf<double> xg64, yg64;     // split: gathered blocks
    DevBuf<float> xg32, yg32;
    DevBuf<double> Qd64;  // q rows, leading dimension ldq (zero-padded)
    DevBuf<float> Qd32;
    // FP16 halves of Qd32 * 2^q16_scale for CB_PREC_MIXED_F16 solves, built
    // before such a solve when Qd32 has changed (q16_rows == 0: not built)
    DevBuf<_Float16> Qhl16;        // hi and lo halves in k_dense_h's fragment order
    int64_t q16_rows = 0;
    int q16_scale = 0;
    bool use_f16 = false;          // set for the duration of a CB_PREC_MIXED_F16 solve
    DevBuf<_Float16> xh16, xl16;   // the operand block's halves
    DevBuf<double> xscale;         // per column: 2^-(q16_scale + its own scale)
    int64_t ldq = 0;
    DevBuf<int64_t> qp;
    DevBuf<int32_t> qc;
    DevBuf<double> qv64;
    DevBuf<float> qv32;
    std::vector<double> q_diag;
    // Stiff direction of a dense Q (the eigenvector of the floored centring
    // null direction of an eigen-bent G^-1): unit v and its eigenvalue mu.
    // Deflated in the preconditioner when mu >> the mean diagonal of Q.
    bool stiff = false;
    double stiff_mu = 0.0;
    DevBuf<double> stiff_v;   // q
    DevBuf<double> stiff_y;   // q: D'^-1 v for the current lambda
    DevBuf<double> stiff_part;  // DOT_CHUNKS x r partial dots
    double stiff_gamma = 0.0; // 1 / (1/(lambda mu) + v' D'^-1 v)
    // preconditioner
    DevBuf<double> Binv;   // p x p
    DevBuf<double> dinv;   // dim
    double dinv_lambda = std::nan("");
    int dinv_block = -1;  // preconditioner kind the cached inverse was built for
    // scratch for dense products
    DevBuf<double> part64;
    DevBuf<float> part32;
    // work vectors of the PCG solves and of cb_mme_solve's blocks, kept from
    // one solve to the next (grown, never shrunk): repeated solves allocate
    // and free nothing, and so never wait on the device for a free
    DevBuf<uint8_t> ws[16];
    // pinned host page, allocated with the model, for the solvers' small
    // exchanges of device state (the column states and dots of each check):
    // a kernel stores them there (or loads them from there), so a check
    // waits on the stream once and uses no copy engine. Lower half: device
    // to host; upper half: host to device (h2d_pending until the stream has
    // been waited on since).
    void* hpage = nullptr;
    bool h2d_pending = false;
    ~cb_mme() {
        if (hpage) (void)hipHostFree(hpage);
    }
};

namespace {

template <class T>
T* qdense(cb_mme* M);
template <>
double* qdense<double>(cb_mme* M) { return M->Qd64.p; }
template <>
float* qdense<float>(cb_mme* M) { return M->Qd32.p; }
template <class T>
const T* wvals(cb_mme* M);
template <>
const double* wvals<double>(cb_mme* M) { return M->wv64.p; }
template <>
const float* wvals<float>(cb_mme* M) { return M->wv32.p; }
template <class T>
const T* qvals(cb_mme* M);
template <>
const double* qvals<double>(cb_mme* M) { return M->qv64.p; }
template <>
const float* qvals<float>(cb_mme* M) { return M->qv32.p; }
// A typed view of one of M's work buffers, grown to hold count elements.
enum WsSlot { WS_R, WS_Z, WS_P, WS_QP, WS_R64, WS_CX, WS_PART, WS_PART2, WS_D1, WS_D2, WS_CS, WS_TMPX, WS_XT,
              WS_IO_B, WS_IO_X, WS_COUNT };
template <class T>
struct WsPtr {
    T* p;
    size_t n;
    size_t bytes() const { return n * sizeof(T); }
};
// count elements of one of M's grow-only buffers, or, during an
// allocation-free call, of the arena's region for `key`
enum ScratchKey { SK_PART64 = 20, SK_PART32, SK_LCPART, SK_STIFF };
template <class T>
T* scratch_or(cb_mme* M, int key, DevBuf<T>& own, size_t count) {
    if (M->arena) return reinterpret_cast<T*>(M->arena->get(key, count * sizeof(T)));
    if (own.n < count) own.alloc(count);
    return own.p;
}

template <class T>
WsPtr<T> ws(cb_mme* M, WsSlot slot, size_t count) {
    static_assert(WS_COUNT <= 16, "work buffer slots");
    uint8_t* p = scratch_or<uint8_t>(M, (int)slot, M->ws[slot], count * sizeof(T));
    return WsPtr<T>{reinterpret_cast<T*>(p), count};
}

constexpr size_t HPAGE_HALF = size_t(32) << 10;  // bytes of each half of a model's pinned page

// Host polls the stream until its queued work is done (no sleep, so the
// solver's checks see the device finish at once).
void wait_idle(cb_mme* M) {
    hipStream_t st = M->ctx->stream;
    for (;;) {
        const hipError_t e = hipStreamQuery(st);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) CB_HIP(e);
        for (int k = 0; k < 64; ++k) __builtin_ia32_pause();
    }
    M->h2d_pending = false;
}

// h[...] = the words of up to three device regions, one after the other
__global__ void k_to_host(const uint32_t* __restrict__ a, int64_t na, const uint32_t* __restrict__ b, int64_t nb,
                          const uint32_t* __restrict__ c, int64_t nc, uint32_t* __restrict__ h) {
    const int64_t n = na + nb + nc;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        h[i] = i < na ? a[i] : i < na + nb ? b[i - na] : c[i - na - nb];
}
__global__ void k_from_host(const uint32_t* __restrict__ h, int64_t n, uint32_t* __restrict__ d) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) d[i] = h[i];
}

struct HostPart {
    void* host;
    const void* dev;
    size_t bytes;
};
// Reads up to three small device regions (whole words) behind the stream's
// queued work: one kernel stores them into the pinned page and one wait
// follows; without the page, or past its half, through copy_host.
void read_small(cb_mme* M, std::initializer_list<HostPart> parts) {
    hipStream_t st = M->ctx->stream;
    size_t tot = 0;
    bool words = parts.size() <= 3;
    for (const HostPart& q : parts) {
        tot += q.bytes;
        words = words && q.bytes % 4 == 0;
    }
    if (!M->hpage || tot > HPAGE_HALF || !words) {
        for (const HostPart& q : parts) copy_host(st, q.host, q.dev, q.bytes, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        M->h2d_pending = false;
        return;
    }
    const uint32_t* src[3] = {nullptr, nullptr, nullptr};
    int64_t nw[3] = {0, 0, 0};
    int k = 0;
    for (const HostPart& q : parts) {
        src[k] = static_cast<const uint32_t*>(q.dev);
        nw[k] = (int64_t)(q.bytes / 4);
        ++k;
    }
    k_to_host<<<1, 256, 0, st>>>(src[0], nw[0], src[1], nw[1], src[2], nw[2], static_cast<uint32_t*>(M->hpage));
    CB_HIP(hipGetLastError());
    wait_idle(M);
    size_t off = 0;
    for (const HostPart& q : parts) {
        std::memcpy(q.host, static_cast<const char*>(M->hpage) + off, q.bytes);
        off += q.bytes;
    }
}
// Writes a small host block (whole words) to the device in stream order:
// staged in the page's upper half and loaded by a kernel, with no wait
// (a second write first waits for the stream, which has then read the first).
void write_small(cb_mme* M, void* dev, const void* host, size_t bytes) {
    hipStream_t st = M->ctx->stream;
    if (!M->hpage || bytes > HPAGE_HALF || bytes % 4 != 0) {
        copy_host(st, dev, host, bytes, hipMemcpyHostToDevice);
        return;
    }
    if (M->h2d_pending) wait_idle(M);
    char* up = static_cast<char*>(M->hpage) + HPAGE_HALF;
    std::memcpy(up, host, bytes);
    k_from_host<<<1, 256, 0, st>>>(reinterpret_cast<const uint32_t*>(up), (int64_t)(bytes / 4),
                                   static_cast<uint32_t*>(dev));
    CB_HIP(hipGetLastError());
    M->h2d_pending = true;
}

template <class T>
T* partbuf(cb_mme* M, size_t need);
template <>
double* partbuf<double>(cb_mme* M, size_t need) {
    return scratch_or<double>(M, SK_PART64, M->part64, need);
}
template <>
float* partbuf<float>(cb_mme* M, size_t need) {
    return scratch_or<float>(M, SK_PART32, M->part32, need);
}

// xg[i, c] = X[gmap[i], c] (0 for gmap < 0); yg[i, c] = 0
template <class T>
__global__ void k_gather_rows(const T* __restrict__ X, int64_t ldx, const int32_t* __restrict__ gmap, int64_t ng,
                              T* __restrict__ xg, int64_t ldg, T* __restrict__ yg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (i >= ng) return;
    const int32_t g = gmap[i];
    xg[c * ldg + i] = g >= 0 ? X[c * ldx + g] : (T)0;
    yg[c * ldg + i] = (T)0;
}

// Y[gmap[i], c] += alpha * yg[i, c] (gmap entries distinct)
template <class T>
__global__ void k_scatter_add_rows(const T* __restrict__ yg, int64_t ldg, const int32_t* __restrict__ gmap,
                                   int64_t ng, T alpha, T* __restrict__ Y, int64_t ldy) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (i >= ng) return;
    const int32_t g = gmap[i];
    if (g >= 0) Y[c * ldy + g] += alpha * yg[c * ldg + i];
}

// Split-K shape of one k_dense_mfma launch: from the tuned table, or a
// shape being timed by the tuner on this thread
struct DenseSplit {
    int target_blocks, min_kspan;
};
thread_local const DenseSplit* g_split_probe = nullptr;
DenseSplit dense_split(int elem, int64_t qn, int rc) {
    if (g_split_probe) return *g_split_probe;
    for (const DenseTunedEntry& e : kDenseTuned)
        if (e.elem == elem && qn <= e.q_max && rc <= e.r_max) return {e.target_blocks, e.min_kspan};
    return {2048, 256};
}

// Y[:, c] += coef (v' x_c) v, a rank-1 term: v, and the per-column partial
// dots of v' x_c (DOT_CHUNKS per column) already on the device
struct StiffTerm {
    const double* v;
    const double* part;
    double coef;
};

// Y[0:qn, c] += alpha * Qd X[0:qn, c] for r columns: the dense block Qd
// (qn rows, leading dimension ldq, zero-padded) of this MME. X and Y point
// at row 0 of the block's index space. With sk, the rank-1 term follows in
// the same pass.
// Split shape of one k_dense_mfma / k_dense_h launch over qn rows
struct DenseShape {
    int64_t row_tiles, kspan, ldp;
    int nsplit;
};
DenseShape dense_shape(int elem, int64_t qn, int r) {
    // k_dense_h runs six 32-k chunks per loop trip
    const int64_t kq = elem == 2 ? 6 * MF_KC : MF_KC;
    DenseShape d;
    d.row_tiles = ceil_div(qn, MF_ROWS);
    // about target_blocks blocks, splits of at least min_kspan k (tuned table)
    const DenseSplit sp = dense_split(elem, qn, std::min(r, MF_MAXNT * 16));
    d.nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(sp.target_blocks, d.row_tiles), qn / sp.min_kspan));
    d.kspan = round_up(ceil_div(qn, d.nsplit), kq);
    d.nsplit = (int)ceil_div(qn, d.kspan);
    d.ldp = round_up(qn, 64);
    return d;
}

// Y[0:qn, c] += alpha * Qd X[0:qn, c] through the FP16 halves of Qd32
// (CB_PREC_MIXED_F16): X split per column (k_split_x), the product on FP16
// MFMA (k_dense_h), then the FP64 split sum, scale and stiff term as in the
// FP32 path.
void dense_apply_h(cb_mme* M, int64_t qn, const float* X, int64_t ldx, float* Y, int64_t ldy, int r, float alpha,
                   const StiffTerm* sk) {
    cb_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    CB_REQUIRE(M->q16_rows == qn && M->Qhl16.p, "FP16 halves of Q missing");
    const int64_t ldh = round_up(qn, 64);
    if (M->xh16.n < (size_t)ldh * r) {
        M->xh16.alloc((size_t)ldh * r);
        M->xl16.alloc((size_t)ldh * r);
    }
    if (M->xscale.n < (size_t)r) M->xscale.alloc(r);
    k_split_x<<<(unsigned)r, SPLIT_X_THREADS, 0, st>>>(X, ldx, qn, ldh, M->q16_scale, M->xh16.p, M->xl16.p,
                                                       M->xscale.p);
    const DenseShape d = dense_shape(2, qn, r);
    const bool stream = (double)qn * (double)M->ldq * 4.0 > Q_CACHE_BYTES;
    for (int c0 = 0; c0 < r; c0 += MF_MAXNT * 16) {
        const int rc = std::min(MF_MAXNT * 16, r - c0);
        float* part = partbuf<float>(M, (size_t)d.nsplit * rc * d.ldp);
        const unsigned grid = (unsigned)(d.row_tiles * d.nsplit);
        const _Float16* xh = M->xh16.p + (int64_t)c0 * ldh;
        const _Float16* xl = M->xl16.p + (int64_t)c0 * ldh;
        timed(ctx, "k_dense_h", [&] {
#define CB_DH(NT_)                                                                                              \
    if (stream)                                                                                                 \
        k_dense_h<NT_, true><<<grid, 256, 0, st>>>(M->Qhl16.p, M->ldq, qn, xh, xl, ldh, rc, d.kspan, d.nsplit,  \
                                                   part, d.ldp);                                                \
    else                                                                                                        \
        k_dense_h<NT_, false><<<grid, 256, 0, st>>>(M->Qhl16.p, M->ldq, qn, xh, xl, ldh, rc, d.kspan, d.nsplit, \
                                                    part, d.ldp)
            switch ((rc + 15) / 16) {
                case 1: CB_DH(1); break;
                case 2: CB_DH(2); break;
                case 3: CB_DH(3); break;
                case 4: CB_DH(4); break;
                default: CB_DH(5); break;
            }
#undef CB_DH
        });
        k_dense_reduce<float><<<dim3(blocks(qn), (unsigned)rc), 256, 0, st>>>(
            part, d.ldp, d.nsplit, qn, rc, Y + (int64_t)c0 * ldy, ldy, 0, alpha, sk ? sk->v : nullptr,
            sk ? sk->part : nullptr, c0, sk ? sk->coef : 0.0, M->xscale.p + c0);
        CB_HIP(hipGetLastError());
    }
}

// Builds the FP16 halves of Qd32 (rows x ldq) for CB_PREC_MIXED_F16, unless
// they are current: one scale for the whole block puts its largest
// magnitude just below 2^14.
void ensure_q16(cb_mme* M, int64_t rows) {
    if (M->q16_rows == rows && M->Qhl16.p) return;
    hipStream_t st = M->ctx->stream;
    const int64_t n = rows * M->ldq;
    CB_REQUIRE(M->ldq % MF_KC == 0, "FP16 halves of Q: leading dimension not a multiple of 32");
    M->Qhl16.alloc((size_t)2 * round_up(rows, 16) * M->ldq);
    CB_HIP(hipMemsetAsync(M->Qhl16.p, 0, M->Qhl16.bytes(), st));
    DevBuf<unsigned> mx;
    mx.alloc(1);
    CB_HIP(hipMemsetAsync(mx.p, 0, sizeof(unsigned), st));
    k_absmax_bits<<<1024, 256, 0, st>>>(M->Qd32.p, n, mx.p);
    CB_HIP(hipGetLastError());
    unsigned bits = 0;
    copy_host(st, &bits, mx.p, sizeof(unsigned), hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    float m;
    std::memcpy(&m, &bits, sizeof m);
    CB_REQUIRE(std::isfinite(m), "FP16 halves of Q: non-finite entries");
    int e = 0;
    if (m > 0) (void)std::frexp(m, &e);
    M->q16_scale = m > 0 ? 14 - e : 0;
    k_split_q<<<blocks(n), 256, 0, st>>>(M->Qd32.p, n, M->ldq, M->q16_scale, M->Qhl16.p);
    CB_HIP(hipGetLastError());
    M->q16_rows = rows;
}

template <class T>
void dense_apply(cb_mme* M, int64_t qn, const T* X, int64_t ldx, T* Y, int64_t ldy, int r, T alpha,
                 const StiffTerm* sk = nullptr) {
    cb_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const T* Q = qdense<T>(M);
    const int64_t ldq = M->ldq;
    if (r <= 2) {
        timed(ctx, sizeof(T) == 8 ? "k_gemv_f64" : "k_gemv_f32", [&] {
            k_gemv<T><<<(unsigned)ceil_div(qn, 4 * GEMV_ROWS), 256, 0, st>>>(Q, ldq, qn, X, ldx, r, Y, ldy, 0,
                                                                         alpha);
        });
        if (sk)
            k_stiff_axpy<T><<<dim3(blocks(qn), (unsigned)r), 256, 0, st>>>(Y, sk->v, sk->part, sk->coef, ldy, qn);
        CB_HIP(hipGetLastError());
        return;
    }
    if constexpr (sizeof(T) == 4) {
        if (M->use_f16) {
            dense_apply_h(M, qn, X, ldx, Y, ldy, r, alpha, sk);
            return;
        }
    }
    const DenseShape d = dense_shape((int)sizeof(T), qn, r);
    const bool stream = (double)qn * (double)ldq * (double)sizeof(T) > Q_CACHE_BYTES;
    const int64_t row_tiles = d.row_tiles, kspan = d.kspan, ldp = d.ldp;
    const int nsplit = d.nsplit;
    for (int c0 = 0; c0 < r; c0 += MF_MAXNT * 16) {
        const int rc = std::min(MF_MAXNT * 16, r - c0);
        // FP32: whole 16-column MFMA tiles and one or two columns past them
        // by VALU (r = 16k + 1 is the probes-plus-y block); FP64 measured
        // faster with the padded tile
        int nt = (rc + 15) / 16, ex = 0;
        if (sizeof(T) == 4 && rc > 16 && rc % 16 != 0 && rc % 16 <= 2) {
            nt = rc / 16;
            ex = rc % 16;
        }
        T* part = partbuf<T>(M, (size_t)nsplit * rc * ldp);
        const T* Xc = X + (int64_t)c0 * ldx;
        const unsigned grid = (unsigned)(row_tiles * nsplit);
        timed(ctx, sizeof(T) == 8 ? "k_dense_mfma_f64" : "k_dense_mfma_f32", [&] {
#define CB_DM(NT_, EX_)                                                                                    \
    if (stream)                                                                                            \
        k_dense_mfma<T, NT_, EX_, true><<<grid, 256, 0, st>>>(Q, ldq, qn, Xc, ldx, rc, kspan, nsplit, part, ldp); \
    else                                                                                                   \
        k_dense_mfma<T, NT_, EX_, false><<<grid, 256, 0, st>>>(Q, ldq, qn, Xc, ldx, rc, kspan, nsplit, part, ldp)
            switch (nt * 4 + ex) {
                case 1 * 4: CB_DM(1, 0); break;
                case 1 * 4 + 1: CB_DM(1, 1); break;
                case 1 * 4 + 2: CB_DM(1, 2); break;
                case 2 * 4: CB_DM(2, 0); break;
                case 2 * 4 + 1: CB_DM(2, 1); break;
                case 2 * 4 + 2: CB_DM(2, 2); break;
                case 3 * 4: CB_DM(3, 0); break;
                case 3 * 4 + 1: CB_DM(3, 1); break;
                case 3 * 4 + 2: CB_DM(3, 2); break;
                case 4 * 4: CB_DM(4, 0); break;
                case 4 * 4 + 1: CB_DM(4, 1); break;
                case 4 * 4 + 2: CB_DM(4, 2); break;
                default: CB_DM(5, 0); break;
            }
#undef CB_DM
        });
        k_dense_reduce<T><<<dim3(blocks(qn), (unsigned)rc), 256, 0, st>>>(
            part, ldp, nsplit, qn, rc, Y + (int64_t)c0 * ldy, ldy, 0, alpha, sk ? sk->v : nullptr,
            sk ? sk->part : nullptr, c0, sk ? sk->coef : 0.0);
        CB_HIP(hipGetLastError());
    }
}

template <class T>
DevBuf<T>& gbuf(cb_mme* M, int which);
template <>
DevBuf<double>& gbuf<double>(cb_mme* M, int which) { return which ? M->yg64 : M->xg64; }
template <>
DevBuf<float>& gbuf<float>(cb_mme* M, int which) { return which ? M->yg32 : M->xg32; }

// Y_u += alpha * Q X_u on an r-column block, every kind of Q:
//   dense   Qd X_u
//   CSR     Q X_u
//   split   A^-1 X_u + scatter(D gather(X_u))  (SplitMatrixAdapter, matrix_adapter.h:253-278,
//           with D = tau G^-1 - omega A22^-1 held as one dense block)
// The rank-1 stiff term of a dense Q's FP32 copy is added by apply_C.
template <class T>
void q_apply(cb_mme* M, const T* X, int64_t ldx, T* Y, int64_t ldy, int r, T alpha, const StiffTerm* sk = nullptr,
             bool csr_done = false) {
    hipStream_t st = M->ctx->stream;
    const int64_t p = M->p, q = M->q;
    if (M->qkind == 0) {
        dense_apply<T>(M, q, X + p, ldx, Y + p, ldy, r, alpha, sk);
        return;
    }
    if (!csr_done) csr_spmm<T, T>(st, M->q_lanes, r, M->qp.p, M->qc.p, qvals<T>(M), 0, q, p, X, ldx, Y, ldy, p, alpha, 1);
    CB_HIP(hipGetLastError());
    if (M->qkind == 2) {
        const int64_t ng = M->ng, ldg = round_up(ng, 64);
        DevBuf<T>& xg = gbuf<T>(M, 0);
        DevBuf<T>& yg = gbuf<T>(M, 1);
        if (xg.n < (size_t)ldg * r) xg.alloc((size_t)ldg * r);
        if (yg.n < (size_t)ldg * r) yg.alloc((size_t)ldg * r);
        k_gather_rows<T><<<dim3(blocks(ng), (unsigned)r), 256, 0, st>>>(X + p, ldx, M->gmap.p, ng, xg.p, ldg, yg.p);
        CB_HIP(hipGetLastError());
        dense_apply<T>(M, ng, xg.p, ldg, yg.p, ldg, r, (T)1);
        k_scatter_add_rows<T><<<dim3(blocks(ng), (unsigned)r), 256, 0, st>>>(yg.p, ldg, M->gmap.p, ng, alpha, Y + p,
                                                                            ldy);
        CB_HIP(hipGetLastError());
    }
}

// a CSR Q rides in the pass over W'W's short rows when those hold every
// random-effect row and W'W's rows need no more lanes than Q's
inline bool csr_fused(const cb_mme* M) {
    return M->qkind != 0 && M->long_rows <= M->p && M->dim > M->long_rows && M->q_lanes >= 2 &&
           M->ww_lanes <= M->q_lanes;
}

// Y = C(lambda) X for an r-column block, arithmetic in T. With lcp_out the
// long rows of W'W that span several chunks may be left unfinished when the
// pass over the short rows carries the chunk partials (its condition puts
// every long row among the fixed effects, where no Q term lands) and there
// are at most MULTI_FIN_MAX of them: *lcp_out then points at the partials,
// which the caller must finish (k_col_dots_fin) before Y's rows [0,
// long_rows) are read; otherwise *lcp_out is null and Y is complete.
template <class T>
void apply_C(cb_mme* M, double lambda, const T* X, T* Y, int r, double** lcp_out = nullptr) {
    cb_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const int64_t ld = M->ldv;
    const bool fused = csr_fused(M);
    // a dense Q: W'W's short rows and its long rows' chunk partials in one
    // pass of their own (the long rows are fixed-effect rows, where Q adds
    // nothing)
    const bool wonly = M->qkind == 0 && M->long_rows > 0 && M->long_rows <= M->p && M->dim > M->long_rows;
    const bool defer =
        lcp_out != nullptr && (fused || wonly) && M->long_rows > 0 && M->n_multi <= MULTI_FIN_MAX;
    if (lcp_out) *lcp_out = nullptr;
    double* lcp = nullptr;
    // W'W part (overwrites Y); the long rows' chunk partials ride in the
    // fused pass when there is one
    if (M->long_rows > 0) {
        lcp = scratch_or<double>(M, SK_LCPART, M->lc_part, (size_t)std::max<int64_t>(1, M->n_lchunks) * r);
        if (M->n_lchunks > 0 && !fused && !wonly)
            k_csr_spmm_chunks<T, T><<<dim3((unsigned)M->n_lchunks, (unsigned)r), 256, 0, st>>>(
            M->lc_beg.p, M->lc_end.p, M->n_lchunks, M->wc.p, wvals<T>(M), X, ld, lcp);
    }
    if (wonly) {
        const int S = M->ww_lanes;
        const int64_t nlc = M->n_lchunks;
        const dim3 grid((unsigned)(nlc + std::max<int64_t>(1, ((M->dim - M->long_rows) * S + 255) / 256)),
                        (unsigned)r);
        timed(ctx, sizeof(T) == 8 ? "k_csr_spmm_f64" : "k_csr_spmm_f32", [&] {
#define CB_CSRW(S_)                                                                                                  \
    k_csr2_spmm<T, T, S_, false><<<grid, 256, 0, st>>>(M->wp.p, M->wc.p, wvals<T>(M), S_, nullptr, nullptr, nullptr,  \
                                                       M->dim, M->long_rows, M->dim, X, ld, Y, ld, (T)0, nlc,         \
                                                       M->lc_beg.p, M->lc_end.p, lcp, M->lc_row1.p)
            switch (S) {
                case 1: CB_CSRW(1); break;
                case 2: CB_CSRW(2); break;
                case 4: CB_CSRW(4); break;
                case 8: CB_CSRW(8); break;
                default: CB_CSRW(16); break;
            }
#undef CB_CSRW
        });
    }
    if (fused) {
        const int S = M->q_lanes;
        const int64_t nlc = M->long_rows > 0 ? M->n_lchunks : 0;
        const dim3 grid((unsigned)(nlc + std::max<int64_t>(1, ((M->dim - M->long_rows) * S + 255) / 256)),
                        (unsigned)r);
        timed(ctx, sizeof(T) == 8 ? "k_csr_spmm_f64" : "k_csr_spmm_f32", [&] {
#define CB_CSR2(S_)                                                                                                  \
    k_csr2_spmm<T, T, S_><<<grid, 256, 0, st>>>(M->wp.p, M->wc.p, wvals<T>(M), M->ww_lanes, M->qp.p, M->qc.p,        \
                                                qvals<T>(M), M->p, M->long_rows, M->dim, X, ld, Y, ld, (T)lambda, nlc, \
                                                M->lc_beg.p, M->lc_end.p, lcp, M->lc_row1.p)
            switch (S) {
                case 2: CB_CSR2(2); break;
                case 4: CB_CSR2(4); break;
                case 8: CB_CSR2(8); break;
                default: CB_CSR2(16); break;
            }
#undef CB_CSR2
        });
    } else if (!wonly && M->dim > M->long_rows) {
        timed(ctx, sizeof(T) == 8 ? "k_csr_spmm_f64" : "k_csr_spmm_f32", [&] {
            csr_spmm<T, T>(st, M->ww_lanes, r, M->wp.p, M->wc.p, wvals<T>(M), M->long_rows, M->dim, 0, X, ld, Y, ld, 0,
                           (T)1, 0);
        });
    }
    CB_HIP(hipGetLastError());
    if (M->long_rows > 0) {
        if (defer) {
            *lcp_out = lcp;
        } else {
            k_csr_chunks_finish<T><<<dim3(blocks(M->long_rows, 4), (unsigned)r), 256, 0, st>>>(
                M->lr_start.p, M->long_rows, M->n_lchunks, lcp, Y, ld);
            CB_HIP(hipGetLastError());
        }
    }
    // lambda Q part; the FP32 copy of a dense Q holds Q - mu v v', and
    // lambda mu v (v' x_u) is added in the same pass that adds lambda Q x_u
    if (M->qkind == 0 && M->stiff && sizeof(T) == 4) {
        double* sp = scratch_or<double>(M, SK_STIFF, M->stiff_part, (size_t)DOT_CHUNKS * r);
        k_stiff_dot<T><<<dim3(DOT_CHUNKS, (unsigned)r), 256, 0, st>>>(X + M->p, M->stiff_v.p, ld, M->q, sp);
        CB_HIP(hipGetLastError());
        const StiffTerm sk{M->stiff_v.p, sp, lambda * M->stiff_mu};
        q_apply<T>(M, X, ld, Y, ld, r, (T)lambda, &sk);
    } else {
        q_apply<T>(M, X, ld, Y, ld, r, (T)lambda, nullptr, fused);
    }
}

// Preconditioner inverse for this lambda (block-Jacobi: Cholesky of the p x p
// block with the reference's ridge, then 1/diag for the random effects).
void build_precond(cb_mme* M, double lambda, bool block) {
    if (M->dinv_lambda == lambda && M->dinv_block == (int)block && M->Binv.p) return;
    const int64_t p = M->p, dim = M->dim;
    std::vector<double> dinv(dim);
    std::vector<double> hv;
    if (M->stiff) {
        hv.resize(M->q);
        copy_host(M->ctx->stream, hv.data(), M->stiff_v.p, M->q * 8, hipMemcpyDeviceToHost);
    }
    for (int64_t i = 0; i < dim; ++i) {
        double d = M->wtw_diag[i] + (i >= p ? lambda * M->q_diag[i - p] : 0.0);
        // with deflation the stiff rank-1 part lambda mu v v' leaves the diagonal
        if (M->stiff && i >= p) d -= lambda * M->stiff_mu * hv[i - p] * hv[i - p];
        dinv[i] = std::fabs(d) > 1e-12 ? 1.0 / d : 1.0;
    }
    if (M->stiff) {
        // Sherman-Morrison for (D' + s v v')^-1, s = lambda mu
        std::vector<double> y(M->q);
        double vy = 0;
        for (int64_t i = 0; i < M->q; ++i) {
            y[i] = dinv[p + i] * hv[i];
            vy += hv[i] * y[i];
        }
        M->stiff_gamma = 1.0 / (1.0 / (lambda * M->stiff_mu) + vy);
        M->stiff_y.alloc(M->q);
        copy_host(M->ctx->stream, M->stiff_y.p, y.data(), M->q * 8, hipMemcpyHostToDevice);
    }
    std::vector<double> binv((size_t)std::max<int64_t>(1, p * p), 0.0);
    if (p > 0) {
        bool ok = block;
        if (block) {
            // pcg_solver.h:77-91: ridge = max(1e-8 * max|diag|, 1e-6), then factorise
            std::vector<double> A = M->xtx;
            double maxd = 0;
            for (int64_t i = 0; i < p; ++i) maxd = std::max(maxd, std::fabs(A[i * p + i]));
            const double ridge = std::max(1e-8 * maxd, 1e-6);
            for (int64_t i = 0; i < p; ++i) A[i * p + i] += ridge;
            // Cholesky A = L L^T (row-major, lower)
            for (int64_t j = 0; j < p && ok; ++j) {
                double s = A[j * p + j];
                for (int64_t k = 0; k < j; ++k) s -= A[j * p + k] * A[j * p + k];
                if (!(s > 0)) { ok = false; break; }
                const double ljj = std::sqrt(s);
                A[j * p + j] = ljj;
                for (int64_t i = j + 1; i < p; ++i) {
                    double t = A[i * p + j];
                    for (int64_t k = 0; k < j; ++k) t -= A[i * p + k] * A[j * p + k];
                    A[i * p + j] = t / ljj;
                }
            }
            if (ok) {
                // inverse via solves with unit vectors
                std::vector<double> col(p);
                for (int64_t c = 0; c < p; ++c) {
                    std::fill(col.begin(), col.end(), 0.0);
                    col[c] = 1.0;
                    for (int64_t i = 0; i < p; ++i) {
                        double t = col[i];
                        for (int64_t k = 0; k < i; ++k) t -= A[i * p + k] * col[k];
                        col[i] = t / A[i * p + i];
                    }
                    for (int64_t i = p - 1; i >= 0; --i) {
                        double t = col[i];
                        for (int64_t k = i + 1; k < p; ++k) t -= A[k * p + i] * col[k];
                        col[i] = t / A[i * p + i];
                    }
                    for (int64_t i = 0; i < p; ++i) binv[i * p + c] = col[i];
                }
            }
        }
        if (!ok) {  // diagonal fallback for the fixed effects
            for (int64_t i = 0; i < p; ++i) {
                const double d = M->wtw_diag[i];
                binv[i * p + i] = std::fabs(d) > 1e-14 ? 1.0 / d : 1.0;
            }
        }
    }
    M->Binv.alloc(binv.size());
    M->dinv.alloc(dim);
    copy_host(M->ctx->stream, M->Binv.p, binv.data(), binv.size() * 8, hipMemcpyHostToDevice);
    copy_host(M->ctx->stream, M->dinv.p, dinv.data(), dim * 8, hipMemcpyHostToDevice);
    CB_HIP(hipStreamSynchronize(M->ctx->stream));
    M->dinv_lambda = lambda;
    M->dinv_block = (int)block;
}

// the preconditioner's block-Jacobi rows and stiff term; the Jacobi rows
// (>= p) of Z must hold R * dinv already
template <class T>
void apply_precond_rest(cb_mme* M, const T* R, T* Z, int r) {
    hipStream_t st = M->ctx->stream;
    if (M->p > 0)
        k_precond_block<T><<<dim3(blocks(M->p, 4), (unsigned)r), 256, 0, st>>>(R, Z, M->Binv.p, M->p, M->ldv);
    if (M->stiff) {
        // z_u -= gamma (v' z_u) D'^-1 v
        double* sp = scratch_or<double>(M, SK_STIFF, M->stiff_part, (size_t)DOT_CHUNKS * r);
        k_stiff_dot<T><<<dim3(DOT_CHUNKS, (unsigned)r), 256, 0, st>>>(Z + M->p, M->stiff_v.p, M->ldv, M->q, sp);
        k_stiff_axpy<T><<<dim3(blocks(M->q), (unsigned)r), 256, 0, st>>>(Z + M->p, M->stiff_y.p, sp,
                                                                        -M->stiff_gamma, M->ldv, M->q);
    }
    CB_HIP(hipGetLastError());
}

template <class T>
void apply_precond(cb_mme* M, const T* R, T* Z, int r) {
    hipStream_t st = M->ctx->stream;
    k_precond_diag<T><<<dim3(blocks(M->dim - M->p), (unsigned)r), 256, 0, st>>>(R, Z, M->dinv.p, M->p, M->dim, M->ldv);
    apply_precond_rest<T>(M, R, Z, r);
}

// The iteration's r.z and r.r are formed inside the x/r update and the
// block preconditioner when there is a block (p > 0) and no stiff term
// (which changes z after the block); part2 then holds the DOT_CHUNKS r.r
// chunk partials per column followed by the x/r update's two sets of
// workgroup partials per column.
inline bool fused_dots(const cb_mme* M) { return M->p > 0 && !M->stiff; }

// The x/r update and the block preconditioner in one launch (k_cg_xr_block,
// then k_cg_p3): one block row per wave, a block of at most PB_LDS rows
inline bool xr_block_path(const cb_mme* M) { return M->p <= PB_LDS; }
inline int64_t xr_block_wgs(const cb_mme* M) { return ceil_div(M->p, 4); }
// doubles per column of the r.z / r.r partials (the workspace part2)
inline size_t part2_count(const cb_mme* M, int r) {
    return ((size_t)DOT_CHUNKS + 4 * ((size_t)ceil_div(M->dim, 256 * CG_RPT) + (size_t)ceil_div(M->p, 4))) * r;
}

// One CG iteration on an r-column block after P holds the search
// directions: Qp = C P, alpha, x and r, z = M^-1 r, beta, convergence and P.
template <class TX, class T>
void cg_iteration(cb_mme* M, double lambda, TX* X, T* R, T* Z, T* P, T* Qp, ColState* cs, double* part,
                  double* part2, int r, double tol2, bool absolute) {
    hipStream_t st = M->ctx->stream;
    const int64_t ld = M->ldv, dim = M->dim;
    const dim3 g2(blocks(dim, 256 * CG_RPT), (unsigned)r);  // the CG vector kernels' grid
    const dim3 gd(DOT_CHUNKS, (unsigned)r);
    double* lcp = nullptr;
    apply_C<T>(M, lambda, P, Qp, r, &lcp);
    const bool long_dots = dim >= DOT_LONG_ROWS;
    if (lcp) {  // Qp's long rows are finished by the dot
        if (long_dots)
            k_col_dots_fin<T, 1024><<<gd, 1024, 0, st>>>(P, Qp, ld, dim, part, M->lr_start.p, M->long_rows,
                                                         M->n_lchunks, lcp, M->lr_multi.p, M->n_multi);
        else
            k_col_dots_fin<T><<<gd, 256, 0, st>>>(P, Qp, ld, dim, part, M->lr_start.p, M->long_rows, M->n_lchunks,
                                                  lcp, M->lr_multi.p, M->n_multi);
    } else if (long_dots) {
        k_col_dots<T, 1024><<<gd, 1024, 0, st>>>(P, Qp, ld, dim, part);
    } else {
        k_col_dots<T><<<gd, 256, 0, st>>>(P, Qp, ld, dim, part);
    }
    if (xr_block_path(M)) {
        // alpha, x, r and z of the Jacobi and block rows with the dot
        // partials in one launch, then the stiff part, beta, p and the block
        // rows' r
        const int64_t nbw = xr_block_wgs(M), ntot = nbw + g2.x;
        const dim3 g3((unsigned)ntot, (unsigned)r);
        const size_t lds = (size_t)M->p * sizeof(double);
        if (M->stiff) {
            k_cg_xr_block<TX, T, true><<<g3, 256, lds, st>>>(X, R, P, Qp, Z, M->dinv.p, M->Binv.p, M->stiff_v.p,
                                                             M->stiff_y.p, cs, part, M->p, dim, ld, nbw, part2);
            k_cg_p3<T, true><<<g2, 256, 0, st>>>(P, Z, R, Qp, M->stiff_y.p, M->stiff_gamma, cs, part2, ntot, tol2,
                                                 absolute ? 1 : 0, M->p, dim, ld);
        } else {
            k_cg_xr_block<TX, T, false><<<g3, 256, lds, st>>>(X, R, P, Qp, Z, M->dinv.p, M->Binv.p, nullptr,
                                                              nullptr, cs, part, M->p, dim, ld, nbw, part2);
            k_cg_p3<T, false><<<g2, 256, 0, st>>>(P, Z, R, Qp, nullptr, 0.0, cs, part2, ntot, tol2,
                                                  absolute ? 1 : 0, M->p, dim, ld);
        }
        CB_HIP(hipGetLastError());
        return;
    }
    if (fused_dots(M)) {
        // r.z and r.r ride in the x/r update (workgroup partials after the
        // DOT_CHUNKS r.r chunk partials in part2) and the block
        // preconditioner, which folds them into the chunk partials k_cg_p reads
        const int64_t nx = g2.x;
        double* px = part2 + (size_t)DOT_CHUNKS * r;
        k_cg_xr_diag<TX, T><<<g2, 256, 0, st>>>(X, R, P, Qp, Z, M->dinv.p, cs, part, M->p, dim, ld, px);
        k_precond_block_dots<T, 1024><<<gd, 1024, 0, st>>>(R, Z, M->Binv.p, M->p, ld, px, nx, part, part2);
    } else {
        k_cg_xr_diag<TX, T><<<g2, 256, 0, st>>>(X, R, P, Qp, Z, M->dinv.p, cs, part, M->p, dim, ld, nullptr);
        apply_precond_rest<T>(M, R, Z, r);
        if (long_dots)
            k_col_dots2<T, 1024><<<gd, 1024, 0, st>>>(R, Z, R, ld, dim, part, part2);
        else
            k_col_dots2<T><<<gd, 256, 0, st>>>(R, Z, R, ld, dim, part, part2);
    }
    k_cg_p<T><<<g2, 256, 0, st>>>(P, Z, cs, part, part2, tol2, absolute ? 1 : 0, dim, ld);
    CB_HIP(hipGetLastError());
}

template <class T>
void col_dots(cb_mme* M, const T* A, const T* B, int r, double* part, double* out) {
    hipStream_t st = M->ctx->stream;
    k_col_dots<T><<<dim3(DOT_CHUNKS, (unsigned)r), 256, 0, st>>>(A, B, M->ldv, M->dim, part);
    k_dots_finish<<<blocks(r, 64), 64, 0, st>>>(part, r, out);
    CB_HIP(hipGetLastError());
}

// C[gmap[i], gmap[j]] += alpha * D[i, j] (D row-major, ldd)
__global__ void k_dense_add_block(const double* __restrict__ D, int64_t ldd, int64_t ng,
                                  const int32_t* __restrict__ gmap, double alpha, double* __restrict__ C, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= ng) return;
    const int32_t gi = gmap[i], gj = gmap[j];
    if (gi >= 0 && gj >= 0) C[(int64_t)gj * ld + gi] += alpha * D[j * ldd + i];
}

// C[row, col - col_off] (column-major, ld) += alpha * CSR(row, col) for rows [0, rows)
__global__ void k_dense_from_csr(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                 const double* __restrict__ val, int64_t rows, int64_t col_off, double alpha,
                                 double* __restrict__ C, int64_t ld) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k) C[((int64_t)col[k] - col_off) * ld + r] += alpha * val[k];
}

// C[i, j] += alpha * Q[i, j] (Q row-major with leading dimension ldq, symmetric)
__global__ void k_dense_add_q(const double* __restrict__ Q, int64_t ldq, int64_t q, double alpha,
                              double* __restrict__ C, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= q) return;
    C[j * ld + i] += alpha * Q[j * ldq + i];
}

// partial sums of log(L_ii) over DOT_CHUNKS row chunks (fixed order)
__global__ __launch_bounds__(256) void k_logdiag_part(const double* __restrict__ L, int64_t n,
                                                      double* __restrict__ part) {
    __shared__ double red[256];
    const int64_t per = (n + DOT_CHUNKS - 1) / DOT_CHUNKS;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(n, r0 + per);
    double s = 0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) s += log(L[i * n + i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

struct SolveStats {
    int64_t iterations = 0;
    double max_rel = 0.0;
};

// Block PCG (one CG recurrence per column) on C(lambda) X = B, arithmetic T,
// solution accumulated in TX. X holds the start (zero or warm). The stop
// test is per column: ||r|| < tol ||b|| (relative) or ||r|| < tol (absolute);
// converged columns are frozen. Returns iterations of the slowest column.
template <class T, class TX>
int64_t pcg_block(cb_mme* M, double lambda, const T* B, TX* X, int r, double tol, bool absolute,
                  int64_t max_iter, bool block_precond, std::vector<double>* rel_out, bool x_zero) {
    cb_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const int64_t ld = M->ldv, dim = M->dim;
    const size_t nv = (size_t)ld * r;
    build_precond(M, lambda, block_precond);
    const WsPtr<T> R = ws<T>(M, WS_R, nv), Z = ws<T>(M, WS_Z, nv), P = ws<T>(M, WS_P, nv), Qp = ws<T>(M, WS_QP, nv);
    const WsPtr<double> part = ws<double>(M, WS_PART, (size_t)DOT_CHUNKS * r), d1 = ws<double>(M, WS_D1, 2 * (size_t)r),
                        d2 = ws<double>(M, WS_D2, r);
    const WsPtr<ColState> cs = ws<ColState>(M, WS_CS, r);
    CB_HIP(hipMemsetAsync(R.p, 0, R.bytes(), st));
    CB_HIP(hipMemsetAsync(Z.p, 0, Z.bytes(), st));
    CB_HIP(hipMemsetAsync(P.p, 0, P.bytes(), st));
    CB_HIP(hipMemsetAsync(Qp.p, 0, Qp.bytes(), st));
    const dim3 g2(blocks(dim), (unsigned)r);
    // r = b - C x
    if (x_zero) {
        CB_HIP(hipMemcpyAsync(R.p, B, nv * sizeof(T), hipMemcpyDeviceToDevice, st));
    } else {
        const WsPtr<TX> tmpX = ws<TX>(M, WS_TMPX, nv);
        CB_HIP(hipMemcpyAsync(tmpX.p, X, nv * sizeof(TX), hipMemcpyDeviceToDevice, st));
        const WsPtr<T> xt = ws<T>(M, WS_XT, nv);
        k_convert<T, TX><<<blocks((int64_t)nv), 256, 0, st>>>(tmpX.p, xt.p, (int64_t)nv);
        apply_C<T>(M, lambda, xt.p, Qp.p, r);
        k_sub<T, T><<<g2, 256, 0, st>>>(B, Qp.p, R.p, dim, ld);
        CB_HIP(hipStreamSynchronize(st));
    }
    apply_precond<T>(M, R.p, Z.p, r);
    CB_HIP(hipMemcpyAsync(P.p, Z.p, nv * sizeof(T), hipMemcpyDeviceToDevice, st));
    col_dots<T>(M, R.p, Z.p, r, part.p, d1.p);
    col_dots<T>(M, B, B, r, part.p, d2.p);
    col_dots<T>(M, R.p, R.p, r, part.p, d1.p + r);  // the initial residual check
    std::vector<double> rz(r), bb(r), rr0(r);
    read_small(M, {{rz.data(), d1.p, (size_t)r * 8},
                   {bb.data(), d2.p, (size_t)r * 8},
                   {rr0.data(), d1.p + r, (size_t)r * 8}});
    std::vector<ColState> hs(r);
    for (int c = 0; c < r; ++c) {
        hs[c] = ColState{};
        hs[c].rho = hs[c].rho_new = rz[c];
        hs[c].bb = bb[c] == 0.0 ? 1.0 : bb[c];  // pcg_solver.h:463: bnorm == 0 -> 1
        hs[c].active = 1;
    }
    // (a zero RHS column is solved by x = 0)
    const double tol2 = tol * tol;
    for (int c = 0; c < r; ++c) {
        const double thr = absolute ? tol2 : tol2 * hs[c].bb;
        hs[c].rr = rr0[c];
        if (!(rr0[c] >= thr)) hs[c].active = 0;
    }
    write_small(M, cs.p, hs.data(), r * sizeof(ColState));
    int64_t it = 0;
    const int check_every = 8;
    bool any_active = std::any_of(hs.begin(), hs.end(), [](const ColState& s) { return s.active; });
    const WsPtr<double> part2 = ws<double>(M, WS_PART2, part2_count(M, r));
    auto iteration = [&] {
        cg_iteration<TX, T>(M, lambda, X, R.p, Z.p, P.p, Qp.p, cs.p, part.p, part2.p, r, tol2, absolute);
    };
    // The first check_every iterations run as plain launches (they size the
    // operator's scratch); later blocks of check_every iterations replay one
    // captured graph. Converged columns are frozen by their ColState, so a
    // block runs the same launches whatever converged inside it.
    LaunchGraph graph;
    const bool use_graph = LaunchGraph::usable(ctx);
    while (any_active && it < max_iter) {
        if (use_graph && it >= check_every && it % check_every == 0 && it + check_every <= max_iter) {
            if (!graph.exec)
                graph.capture(ctx, [&] {
                    for (int k = 0; k < check_every; ++k) iteration();
                });
            graph.launch(ctx);
            it += check_every;
        } else {
            iteration();
            ++it;
        }
        if (it % check_every == 0 || it == max_iter) {
            read_small(M, {{hs.data(), cs.p, r * sizeof(ColState)}});
            any_active = std::any_of(hs.begin(), hs.end(), [](const ColState& s) { return s.active; });
        }
    }
    read_small(M, {{hs.data(), cs.p, r * sizeof(ColState)}});
    if (rel_out) {
        rel_out->resize(r);
        for (int c = 0; c < r; ++c)
            (*rel_out)[c] = absolute ? std::sqrt(hs[c].rr) : std::sqrt(hs[c].rr / hs[c].bb);
    }
    return it;
}

}  // namespace

namespace {

__global__ void k_to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t rows, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= rows) return;
    out[(int64_t)c * ld + i] = (float)in[(int64_t)c * ld + i];
}

// Block PCG in FP32 with FP64 iterate and FP64 residual replacement.
// Returns iterations of the slowest column; *max_rel = worst true relative
// (or absolute) residual at exit.
int64_t pcg_reliable(cb_mme* M, double lambda, const double* B64, double* X, int r,
                     const cb_solve_options& o, double* max_rel, bool warm) {
    cb_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const int64_t ld = M->ldv, dim = M->dim;
    const size_t nv = (size_t)ld * r;
    const bool absolute = o.stop_rule == CB_STOP_ABSOLUTE;
    build_precond(M, lambda, o.precond == CB_PRECOND_BLOCK_JACOBI);
    const WsPtr<float> R = ws<float>(M, WS_R, nv), Z = ws<float>(M, WS_Z, nv), P = ws<float>(M, WS_P, nv),
                       Qp = ws<float>(M, WS_QP, nv);
    const WsPtr<double> R64 = ws<double>(M, WS_R64, nv), CX = ws<double>(M, WS_CX, nv),
                        part = ws<double>(M, WS_PART, (size_t)DOT_CHUNKS * r), d1 = ws<double>(M, WS_D1, r),
                        d2 = ws<double>(M, WS_D2, r);
    const WsPtr<ColState> cs = ws<ColState>(M, WS_CS, r);
    CB_HIP(hipMemsetAsync(R.p, 0, R.bytes(), st));
    CB_HIP(hipMemsetAsync(Z.p, 0, Z.bytes(), st));
    CB_HIP(hipMemsetAsync(P.p, 0, P.bytes(), st));
    CB_HIP(hipMemsetAsync(Qp.p, 0, Qp.bytes(), st));
    CB_HIP(hipMemsetAsync(CX.p, 0, CX.bytes(), st));
    const dim3 g2(blocks(dim), (unsigned)r);
    const double tol2 = o.tol * o.tol;
    std::vector<double> bb(r), rr(r), rz(r), rr_ref(r);
    std::vector<std::vector<double>> hist(r);  // true residual at each replacement
    std::vector<ColState> hs(r);
    col_dots<double>(M, B64, B64, r, part.p, d1.p);
    read_small(M, {{bb.data(), d1.p, (size_t)r * 8}});
    for (int c = 0; c < r; ++c) bb[c] = bb[c] == 0.0 ? 1.0 : bb[c];
    std::vector<int> done(r, 0);  // true-residual convergence per column
    // (re)start from the true residual: R64 = B - C x, R = (float)R64,
    // Z = M^-1 R, rho = R.Z; returns true when every column has converged
    auto replace = [&](bool first) -> bool {
        if (first && !warm) {
            CB_HIP(hipMemcpyAsync(R64.p, B64, nv * 8, hipMemcpyDeviceToDevice, st));
        } else {
            apply_C<double>(M, lambda, X, CX.p, r);
            k_sub<double, double><<<g2, 256, 0, st>>>(B64, CX.p, R64.p, dim, ld);
        }
        k_to_f32<<<g2, 256, 0, st>>>(R64.p, R.p, dim, ld);
        apply_precond<float>(M, R.p, Z.p, r);
        col_dots<double>(M, R64.p, R64.p, r, part.p, d1.p);
        col_dots<float>(M, R.p, Z.p, r, part.p, d2.p);
        if (first) read_small(M, {{rr.data(), d1.p, (size_t)r * 8}, {rz.data(), d2.p, (size_t)r * 8}});
        else
            read_small(M, {{rr.data(), d1.p, (size_t)r * 8},
                           {rz.data(), d2.p, (size_t)r * 8},
                           {hs.data(), cs.p, r * sizeof(ColState)}});
        bool all = true;
        for (int c = 0; c < r; ++c) {
            if (!std::isfinite(rr[c]))
                throw Error(CB_ERR_NOT_CONVERGED, "mixed-precision PCG: non-finite residual");
            const double thr = absolute ? tol2 : tol2 * bb[c];
            if (first) { hs[c] = ColState{}; hs[c].bb = bb[c]; }
            if (!done[c]) {
                if (!(rr[c] >= thr)) {
                    done[c] = 1;
                } else if (hist[c].size() >= 8) {
                    // the best true residual of the last eight replacements is
                    // not half the best before them: the FP64 rounding floor,
                    // around which the true residual only fluctuates
                    const size_t h = hist[c].size();
                    double before = hist[c][0], recent = rr[c];
                    for (size_t k = 1; k + 7 < h; ++k) before = std::min(before, hist[c][k]);
                    for (size_t k = h - 7; k < h; ++k) recent = std::min(recent, hist[c][k]);
                    const double rel = absolute ? std::sqrt(rr[c]) : std::sqrt(rr[c] / bb[c]);
                    // accepted only on request (stall_tol > 0); otherwise the
                    // column iterates on and a true floor above tol ends in
                    // CB_ERR_NOT_CONVERGED at max_iter
                    if (recent > 0.25 * before && rel < o.stall_tol) done[c] = 1;
                }
            }
            // a column paused by the recursive test (or one whose recursion
            // blew up) holds a direction that predates its last residual:
            // restart it from the new preconditioned residual
            const bool restart = first || !hs[c].active || !(hs[c].rr <= 100.0 * rr_ref[c]);
            hist[c].push_back(rr[c]);
            rr_ref[c] = rr[c];
            hs[c].rho = hs[c].rho_new = rz[c];
            hs[c].rr = rr[c];
            hs[c].active = done[c] ? 0 : 1;
            if (restart && !done[c])
                CB_HIP(hipMemcpyAsync(P.p + (size_t)c * ld, Z.p + (size_t)c * ld, ld * 4,
                                      hipMemcpyDeviceToDevice, st));
            all = all && done[c];
        }
        write_small(M, cs.p, hs.data(), r * sizeof(ColState));
        return all;
    };
    bool all_done = replace(true);
    int64_t it = 0;
    const int check_every = 4;
    const WsPtr<double> part2 = ws<double>(M, WS_PART2, part2_count(M, r));
    // the recursive test in the iteration only pauses a column until the
    // next replacement
    auto iteration = [&] {
        cg_iteration<double, float>(M, lambda, X, R.p, Z.p, P.p, Qp.p, cs.p, part.p, part2.p, r, tol2, absolute);
    };
    // as in pcg_block: plain launches for the first block, then a captured
    // block of check_every iterations replayed between the host checks (the
    // replacements in between are ordinary launches and copies)
    LaunchGraph graph;
    const bool use_graph = LaunchGraph::usable(ctx);
    while (!all_done && it < o.max_iter) {
        if (use_graph && it >= check_every && it % check_every == 0 && it + check_every <= o.max_iter) {
            if (!graph.exec)
                graph.capture(ctx, [&] {
                    for (int k = 0; k < check_every; ++k) iteration();
                });
            graph.launch(ctx);
            it += check_every;
        } else {
            iteration();
            ++it;
        }
        if (it % check_every == 0 || it == o.max_iter) {
            read_small(M, {{hs.data(), cs.p, r * sizeof(ColState)}});
            bool need = false;
            for (int c = 0; c < r; ++c) {
                if (done[c]) continue;
                // paused by the recursive test, fallen a thousandfold since the
                // last replacement (the FP32 recursion drifts from the true
                // residual by about eps32 times the largest residual since
                // then), or grown tenfold (lost conjugacy)
                if (!hs[c].active || hs[c].rr < REPLACE_DROP2 * rr_ref[c] || !(hs[c].rr <= 100.0 * rr_ref[c]))
                    need = true;
            }
            if (need) {
                all_done = replace(false);
            }
        }
    }
    if (!all_done) all_done = replace(false);
    double worst = 0;
    for (int c = 0; c < r; ++c)
        worst = std::max(worst, absolute ? std::sqrt(rr[c]) : std::sqrt(rr[c] / bb[c]));
    *max_rel = worst;
    if (!all_done) throw Error(CB_ERR_NOT_CONVERGED, "mixed-precision PCG did not reach the tolerance");
    return it;
}

}  // namespace

// Solve driver shared by cb_mme_solve and the REML loop. B, X device FP64
// blocks [ldv x r]; X is the output (zero start).
// warm: X holds the initial guess (else the solve starts from 0)
static SolveStats solve_block(cb_mme* M, double lambda, const cb_solve_options& o, const double* B,
                              double* X, int r, bool warm = false) {
    cb_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const int64_t ld = M->ldv;
    const size_t nv = (size_t)ld * r;
    const bool absolute = o.stop_rule == CB_STOP_ABSOLUTE;
    const bool block = o.precond == CB_PRECOND_BLOCK_JACOBI;
    SolveStats S;
    if (!warm) CB_HIP(hipMemsetAsync(X, 0, nv * 8, st));
    std::vector<double> rel;
    CB_REQUIRE(o.precision >= CB_PREC_FP64 && o.precision <= CB_PREC_MIXED_F16, "unknown precision");
    if (o.precision == CB_PREC_FP64) {
        if (o.batch_rhs || r == 1) {
            S.iterations = pcg_block<double, double>(M, lambda, B, X, r, o.tol, absolute, o.max_iter, block, &rel, !warm);
            for (double v : rel) S.max_rel = std::max(S.max_rel, v);
        } else {
            // the plain path: one right-hand side at a time
            for (int c = 0; c < r; ++c) {
                const int64_t itc = pcg_block<double, double>(M, lambda, B + (size_t)c * ld, X + (size_t)c * ld, 1,
                                                              o.tol, absolute, o.max_iter, block, &rel, !warm);
                S.iterations += itc;
                S.max_rel = std::max(S.max_rel, rel[0]);
            }
        }
        if (S.max_rel > o.tol) throw Error(CB_ERR_NOT_CONVERGED, "PCG did not reach the tolerance");
        return S;
    }
    // Mixed: one CG recursion per column with the operator, preconditioner
    // and direction vectors in FP32, the iterate accumulated in FP64, and the
    // residual replaced by the true FP64 residual b - C x whenever the
    // recursive one has dropped a thousandfold since the last replacement
    // (and before any column is declared converged). The search directions are
    // kept across replacements.
    CB_REQUIRE(M->qkind == 1 || M->Qd32.p, "FP32 copy of Q missing");
    struct F16Scope {
        cb_mme* M;
        ~F16Scope() { M->use_f16 = false; }
    } f16{M};
    if (o.precision == CB_PREC_MIXED_F16 && M->qkind != 1) {
        ensure_q16(M, M->qkind == 0 ? M->q : M->ng);
        M->use_f16 = true;
    }
    S.iterations = pcg_reliable(M, lambda, B, X, r, o, &S.max_rel, warm);
    return S;
}

// ---------------------------------------------------------------------------
// MME assembly (host) — MMELHSBuilder::initialize (mme_builder.cpp:72-257)
static void assemble(cb_mme* M, const int64_t* x_ptr, const int32_t* x_col, const double* x_val,
                     const int32_t* aid, const double* y) {
    const int64_t n = M->n, p = M->p, q = M->q, dim = M->dim;
    struct Trip { int64_t r, c; double v; };
    std::vector<Trip> wt;  // W^T entries (unknown, record, value)
    std::vector<Trip> tr;  // W'W entries
    wt.reserve((size_t)(x_ptr[n] + n));
    for (int64_t r = 0; r < n; ++r) {
        CB_REQUIRE(aid[r] < q && aid[r] >= -1, "record random-effect index out of range");
        std::vector<std::pair<int64_t, double>> row;
        for (int64_t k = x_ptr[r]; k < x_ptr[r + 1]; ++k) {
            CB_REQUIRE(x_col[k] >= 0 && x_col[k] < p, "fixed-effect column out of range");
            row.emplace_back(x_col[k], x_val[k]);
        }
        if (aid[r] >= 0) row.emplace_back(p + aid[r], 1.0);
        for (auto& a : row) {
            wt.push_back({a.first, r, a.second});
            for (auto& b : row) tr.push_back({a.first, b.first, a.second * b.second});
        }
    }
    auto to_csr = [](std::vector<Trip>& t, int64_t rows, std::vector<int64_t>& ptr, std::vector<int32_t>& col,
                     std::vector<double>& val) {
        std::sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
        ptr.assign(rows + 1, 0);
        col.clear();
        val.clear();
        for (size_t i = 0; i < t.size();) {
            size_t j = i;
            double s = 0;
            while (j < t.size() && t[j].r == t[i].r && t[j].c == t[i].c) s += t[j++].v;
            col.push_back((int32_t)t[i].c);
            val.push_back(s);
            ptr[t[i].r + 1]++;
            i = j;
        }
        for (int64_t r = 0; r < rows; ++r) ptr[r + 1] += ptr[r];
    };
    std::vector<int64_t> tp, wp;
    std::vector<int32_t> tc, wc;
    std::vector<double> tv, wv;
    to_csr(wt, dim, tp, tc, tv);
    to_csr(tr, dim, wp, wc, wv);
    M->wtw_diag.assign(dim, 0.0);
    M->xtx.assign((size_t)p * p, 0.0);
    int64_t long_rows = 0;
    for (int64_t r = 0; r < dim; ++r)
        for (int64_t k = wp[r]; k < wp[r + 1]; ++k) {
            if (wc[k] == r) M->wtw_diag[r] = wv[k];
            if (r < p && wc[k] < p) M->xtx[r * p + wc[k]] = wv[k];
        }
    for (int64_t r = 0; r < p; ++r)
        if (wp[r + 1] - wp[r] > 32) long_rows = r + 1;
    M->long_rows = long_rows;
    M->ww_lanes = csr_lanes(wp[dim] - wp[long_rows], dim - long_rows);
    std::vector<int64_t> lcb, lce, lrs(long_rows + 1, 0);
    for (int64_t r = 0; r < long_rows; ++r) {
        for (int64_t k = wp[r]; k < wp[r + 1]; k += LONG_CHUNK) {
            lcb.push_back(k);
            lce.push_back(std::min<int64_t>(wp[r + 1], k + LONG_CHUNK));
        }
        lrs[r + 1] = (int64_t)lcb.size();
    }
    M->n_lchunks = (int64_t)lcb.size();
    std::vector<int32_t> lrow1(lcb.size(), -1);
    std::vector<int64_t> lmulti;
    for (int64_t r = 0; r < long_rows; ++r) {
        if (lrs[r + 1] - lrs[r] == 1) lrow1[lrs[r]] = (int32_t)r;
        else lmulti.push_back(r);
    }
    M->n_multi = (int64_t)lmulti.size();
    // RHS = W'y (mme_builder.cpp:321-328), y'y
    M->rhs.assign(dim, 0.0);
    for (int64_t i = 0; i < dim; ++i)
        for (int64_t k = tp[i]; k < tp[i + 1]; ++k) M->rhs[i] += tv[k] * y[tc[k]];
    M->yty = 0;
    for (int64_t r = 0; r < n; ++r) M->yty += y[r] * y[r];
    // upload
    hipStream_t st = M->ctx->stream;
    auto up = [&](auto& buf, const auto& vec) {
        using E = std::remove_reference_t<decltype(*buf.p)>;
        buf.alloc(std::max<size_t>(1, vec.size()));
        if (!vec.empty()) copy_host(st, buf.p, vec.data(), vec.size() * sizeof(E), hipMemcpyHostToDevice);
    };
    std::vector<int64_t> xp(x_ptr, x_ptr + n + 1);
    std::vector<int32_t> xc(x_col, x_col + x_ptr[n]), av(aid, aid + n);
    std::vector<double> xv(x_val, x_val + x_ptr[n]), yv(y, y + n);
    up(M->xp, xp); up(M->xc, xc); up(M->xv, xv); up(M->aid, av); up(M->y, yv);
    up(M->tp, tp); up(M->tc, tc); up(M->tv, tv);
    up(M->wp, wp); up(M->wc, wc); up(M->wv64, wv);
    up(M->lc_beg, lcb); up(M->lc_end, lce); up(M->lr_start, lrs);
    up(M->lc_row1, lrow1); up(M->lr_multi, lmulti);
    std::vector<float> wv32(wv.begin(), wv.end());
    up(M->wv32, wv32);
    CB_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// REML
namespace {

struct RemlState {
    double vu, ve;
    double grad[2], ai[4];
    double tr, uQu, esse;
};

// 2x2 SPD projection of vce.cpp:267-283 (closed-form symmetric eigen)
void project_spd_2x2(double A[4]) {
    const double a = A[0], b = 0.5 * (A[1] + A[2]), d = A[3];
    const double tr = a + d, det = a * d - b * b;
    const double disc = std::sqrt(std::max(0.0, 0.25 * (a - d) * (a - d) + b * b));
    double l1 = 0.5 * tr + disc, l2 = 0.5 * tr - disc;  // l1 >= l2
    double v1x, v1y;
    if (std::fabs(b) > 0) { v1x = l1 - d; v1y = b; }
    else if (a >= d) { v1x = 1; v1y = 0; }
    else { v1x = 0; v1y = 1; }
    const double nv = std::sqrt(v1x * v1x + v1y * v1y);
    v1x /= nv; v1y /= nv;
    const double v2x = -v1y, v2y = v1x;
    const double maxe = std::max(std::max(l1, l2), 0.0);
    const double floor = std::max(1e-12, 1e-8 * maxe);
    l1 = std::max(l1, floor);
    l2 = std::max(l2, floor);
    (void)det;
    A[0] = l1 * v1x * v1x + l2 * v2x * v2x;
    A[1] = A[2] = l1 * v1x * v1y + l2 * v2x * v2y;
    A[3] = l1 * v1y * v1y + l2 * v2y * v2y;
}

}  // namespace

// Device state of one REML run. The block solve takes y's right-hand side
// W'y and the k probe columns [0; z] together, bw columns at a time: bw =
// 1 + k when the block fits (probes and Q z generated once for the run),
// else the probes go through in chunks of bw columns whose vectors and Q z are
// regenerated at every evaluation instead of being stored, y's column riding
// with the last chunk.
// The solutions of one evaluation start the next one's solves (the probe
// block when it is one block, and the AI pair; y's column, which the last
// chunk leaves in S, starts the final solve), as AI_REML::solve_lhs starts
// PCG from last_solution (vce.cpp:1225-1234).
struct cb_reml_work {
    DevBuf<double> B, S, Qz, B2, S2, V, T, part, dots, e, rhs, Qu;
    int k = 0, bw = 0;
    bool chunked = false;
    bool have_s = false, have_s2 = false;  // S (last block) and S2 hold solutions
    // adaptive traces: the k = max_k probe columns are solved in sub-blocks
    // (y + probes up to the first check point, then batches of 16) until the
    // stopping rule holds; sub_warm[b]: sub-block b holds solutions
    bool adaptive = false;
    bool exact = false;  // probes are the q unit vectors
    int64_t min_k = 0, max_k = 0;
    std::vector<char> sub_warm;
    int64_t used_k = 0;  // probes the last evaluation used
};

// The branch of one REML run: the step of iteration 0 and of later
// iterations (CB_VCE_AI_EXACT / AI / EM), the trace rule (0 fixed n_probes,
// 1 adaptive, 2 exact by unit probes), whether AI steps need a logL increase,
// and the adaptive rule's probe bounds.
struct RemlPlan {
    int first = CB_VCE_AI, later = CB_VCE_AI;
    int trace = 0;
    bool logl = false;
    int64_t min_k = 0, max_k = 0;
};

namespace {

// device bytes per block column of a REML solve: B, S and Q z (FP64), the PCG
// vectors (four FP32 and two FP64, or four FP64) and the split-K partials
constexpr int64_t REML_BYTES_PER_COL_ROW = 128;

int reml_block_width(cb_mme* M, const cb_reml_options& o, int r) {
    if (o.block_cols > 0) return (int)std::min<int64_t>(r, std::max<int64_t>(2, o.block_cols));
    size_t free_b = 0, total_b = 0;
    CB_HIP(hipMemGetInfo(&free_b, &total_b));
    const int64_t fit = (int64_t)(free_b / 2) / (REML_BYTES_PER_COL_ROW * M->ldv);
    return (int)std::min<int64_t>(r, std::max<int64_t>(2, fit));
}

// probe columns [first, first + count) into B's columns dst.. (rows p.. ; the
// fixed-effect rows stay zero) and, with Qz, Q z into Qz's columns 0..count
void make_probes(cb_mme* M, const cb_reml_options& o, cb_reml_work& w, int64_t first, int count, int dst,
                 bool with_qz) {
    hipStream_t st = M->ctx->stream;
    const int64_t ld = M->ldv, p = M->p, q = M->q;
    CB_HIP(hipMemsetAsync(w.B.p + (size_t)dst * ld, 0, (size_t)count * ld * 8, st));
    k_probes<<<dim3(blocks(q), (unsigned)count), 256, 0, st>>>(w.B.p + (size_t)dst * ld, ld, p, q, count, o.seed,
                                                              w.exact ? 1 : 0, first);
    CB_HIP(hipGetLastError());
    if (with_qz) {
        CB_HIP(hipMemsetAsync(w.Qz.p, 0, (size_t)count * ld * 8, st));
        q_apply<double>(M, w.B.p + (size_t)dst * ld, ld, w.Qz.p, ld, count, 1.0);
    }
}

}  // namespace

// AdaptiveTraceEstimator bounds for n_samples ns (vce.cpp:1312-1314)
static void adaptive_bounds(int64_t ns, RemlPlan& pl) {
    pl.min_k = std::max<int64_t>(30, ns / 3);
    pl.max_k = std::max<int64_t>(100, ns);
}

// CB_VCE_AUTO: the branch AI_REML takes for a single random component
// (run_ai_iteration, vce.cpp:1334-1574), from the solvers it would have:
// Ainv factorised for an in-memory sparse Q with q <= 100000 (and not
// q >= 50000 with density > 0.001, vce.cpp:877-1001); build_lhs's dense
// direct solver for dim < 10000 or LHS density > 0.05 with dim < 15000, else
// its sparse direct solver for dim < 200000 and density < 0.05
// (vce.cpp:1041-1110). The LHS density is estimated from nnz(W'W) plus the
// nonzeros of Q (q^2 dense; CSR; CSR + ng^2 split).
static RemlPlan reml_plan(const cb_mme* M, const cb_reml_options& o) {
    RemlPlan pl;
    const int64_t mc = o.mc_samples > 0 ? o.mc_samples : 100;
    if (o.algorithm != CB_VCE_AUTO) {
        pl.first = pl.later = o.algorithm;
        pl.trace = o.trace_exact ? 2 : (o.adaptive ? 1 : 0);
        adaptive_bounds(std::max<int64_t>(5, mc / 2), pl);
        return pl;
    }
    const int64_t dim = M->dim, q = M->q;
    if (dim < 5000) {  // run_dense_exact_step (vce.cpp:1397-1422)
        pl.first = pl.later = CB_VCE_AI_EXACT;
        return pl;
    }
    const double ww_nnz = M->wc.n ? (double)M->wc.n : 0.0;
    double q_nnz = 0;
    if (M->qkind == 0) q_nnz = (double)q * (double)q;
    else q_nnz = (double)M->qc.n + (M->qkind == 2 ? (double)M->ng * (double)M->ng : 0.0);
    const double density = (ww_nnz + q_nnz) / ((double)dim * (double)dim);
    const bool dense_direct = dim < 10000 || (density > 0.05 && dim < 15000);
    const bool sparse_direct = !dense_direct && dim < 200000 && density < 0.05;
    const double q_density = M->qkind == 1 ? (double)M->qc.n / ((double)q * (double)q) : 1.0;
    const bool ainv_fact = M->qkind == 1 && q <= 100000 && !(q >= 50000 && q_density > 0.001);
    if (dense_direct || sparse_direct) {
        // iteration 0: AI needs Ainv factorised or a solver, which build_lhs
        // creates only inside the step; EM otherwise. Later: the AI step
        // with its logL monotonicity check.
        pl.first = ainv_fact ? CB_VCE_AI : CB_VCE_EM;
        pl.later = CB_VCE_AI;
        pl.logl = dim <= 65536;
        if (q <= 5000) {
            pl.trace = 2;  // estimate_trace_Qinv_Cuu's direct branch (vce.cpp:1278-1302)
        } else {
            pl.trace = 1;
            adaptive_bounds(std::max<int64_t>(50, mc), pl);
        }
        return pl;
    }
    // no direct solver: the AI step cannot check logL and falls back to EM
    // every iteration (vce.cpp:2788-2795, 1441-1454)
    pl.first = pl.later = CB_VCE_EM;
    pl.trace = 1;
    adaptive_bounds(std::max<int64_t>(5, mc / 2), pl);
    return pl;
}

// The adaptive rule of AdaptiveTraceEstimator::estimate
// (adaptive_trace_estimator.cpp:11-83) over per-probe values t[0..done):
// checked at each batch end once done >= min_k.
static bool adaptive_stop(const std::vector<double>& t, int64_t done, int64_t min_k) {
    if (done < min_k || done < 2) return false;
    double sum = 0, sq = 0;
    for (int64_t i = 0; i < done; ++i) {
        sum += t[i];
        sq += t[i] * t[i];
    }
    const double mean = sum / (double)done;
    const double var = (sq - (sum * sum) / (double)done) / ((double)done - 1.0);
    const double se = std::sqrt(std::max(0.0, var) / (double)done);
    return std::fabs(mean) > 1e-12 && se / std::fabs(mean) < 0.05;
}

// One evaluation at (vu, ve): solves, score terms, optionally the AI matrix.
static void reml_eval(cb_mme* M, const cb_reml_options& o, cb_reml_work& w, double vu, double ve,
                      bool want_ai, RemlState& out, double* sol_out) {
    cb_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const int64_t ld = M->ldv, p = M->p, q = M->q, n = M->n, dim = M->dim;
    const int k = w.k, bw = w.bw;
    const double lambda = ve / std::max(vu, 1e-9);
    // probe chunks: [first, first + count), y's column with the last one
    std::vector<std::pair<int64_t, int>> chunks;
    if (!w.chunked) {
        chunks.push_back({0, k});
    } else {
        const int64_t last = std::min<int64_t>(k, bw - 1);
        for (int64_t f = 0; f < k - last; f += bw) chunks.push_back({f, (int)std::min<int64_t>(bw, k - last - f)});
        chunks.push_back({k - last, (int)last});
    }
    std::vector<double> traces(k), hd(2 + bw);
    DevBuf<double>& Qu = w.Qu;
    if (w.adaptive) {
        // sub-block 0: y and the probes up to the first check point (the rule
        // is not checked before min_k); then batches of 16
        int64_t done = 0;
        int sub = 0;
        bool stop = false;
        while (!stop && done < w.max_k) {
            int64_t end = done == 0 ? std::min<int64_t>(w.max_k, round_up(w.min_k, 16))
                                    : std::min<int64_t>(w.max_k, done + 16);
            const bool with_y = done == 0;
            const int64_t off = with_y ? 0 : 1 + done;  // first column of the sub-block
            const int cnt = (int)(end - done) + (with_y ? 1 : 0);
            if ((int)w.sub_warm.size() <= sub) w.sub_warm.resize(sub + 1, 0);
            solve_block(M, lambda, o.solve, w.B.p + (size_t)off * ld, w.S.p + (size_t)off * ld, cnt,
                        w.sub_warm[sub] != 0);
            w.sub_warm[sub] = 1;
            if (with_y) {
                CB_HIP(hipMemsetAsync(Qu.p, 0, ld * 8, st));
                q_apply<double>(M, w.S.p, ld, Qu.p, ld, 1, 1.0);
            }
            const int ncol = (with_y ? 2 : 0) + (int)(end - done);
            k_reml_score_terms<<<dim3(DOT_CHUNKS, (unsigned)ncol), 256, 0, st>>>(
                M->xp.p, M->xc.p, M->xv.p, M->aid.p, M->y.p, n, p, q, w.S.p + (size_t)off * ld, ld, Qu.p,
                w.Qz.p + (size_t)done * ld, with_y ? 1 : 0, w.e.p, w.part.p);
            k_dots_finish<<<blocks(ncol, 64), 64, 0, st>>>(w.part.p, ncol, w.dots.p);
            CB_HIP(hipGetLastError());
            copy_host(st, hd.data(), w.dots.p, ncol * 8, hipMemcpyDeviceToHost);
            CB_HIP(hipStreamSynchronize(st));
            const double* th = hd.data() + (with_y ? 2 : 0);
            for (int64_t j = done; j < end; ++j) traces[j] = th[j - done];
            if (with_y) {
                out.esse = hd[0];
                out.uQu = hd[1];
            }
            done = end;
            ++sub;
            stop = adaptive_stop(traces, done, w.min_k);
        }
        w.used_k = done;
        w.have_s = true;
        double tr = 0;
        for (int64_t i = 0; i < done; ++i) tr += traces[i];
        tr /= (double)done;
        out.vu = vu;
        out.ve = ve;
        out.tr = tr;
    }
    for (size_t ci = 0; ci < (w.adaptive ? 0 : chunks.size()); ++ci) {
        const bool with_y = ci + 1 == chunks.size();
        const int64_t first = chunks[ci].first;
        const int count = chunks[ci].second;
        const int t0 = with_y ? 1 : 0;
        if (w.chunked) {
            if (with_y) CB_HIP(hipMemcpyAsync(w.B.p, w.rhs.p, ld * 8, hipMemcpyDeviceToDevice, st));
            if (count > 0) make_probes(M, o, w, first, count, t0, true);
        }
        // (chunks share S, so only one block keeps its solutions to the next evaluation)
        solve_block(M, lambda, o.solve, w.B.p, w.S.p, t0 + count, !w.chunked && w.have_s);
        // the score terms of this block ("adjoint" pass over its solutions)
        if (with_y) {
            CB_HIP(hipMemsetAsync(Qu.p, 0, ld * 8, st));
            q_apply<double>(M, w.S.p, ld, Qu.p, ld, 1, 1.0);
        }
        const int ncol = (with_y ? 2 : 0) + count;
        k_reml_score_terms<<<dim3(DOT_CHUNKS, (unsigned)ncol), 256, 0, st>>>(
            M->xp.p, M->xc.p, M->xv.p, M->aid.p, M->y.p, n, p, q, w.S.p, ld, Qu.p, w.Qz.p, with_y ? 1 : 0, w.e.p,
            w.part.p);
        k_dots_finish<<<blocks(ncol, 64), 64, 0, st>>>(w.part.p, ncol, w.dots.p);
        CB_HIP(hipGetLastError());
        copy_host(st, hd.data(), w.dots.p, ncol * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        const double* th = hd.data() + (with_y ? 2 : 0);
        for (int j = 0; j < count; ++j) traces[first + j] = th[j];
        if (with_y) {
            out.esse = hd[0];
            out.uQu = hd[1];
        }
    }
    if (!w.adaptive) {
        w.have_s = true;
        double tr = 0;
        for (int i = 0; i < k; ++i) tr += traces[i];
        if (!w.exact && k > 0) tr /= k;
        out.vu = vu;
        out.ve = ve;
        out.tr = tr;
        w.used_k = k;
    }
    const double tr = out.tr;
    // score, vce.cpp:2655 and :2661-2675 (lambda*tr capped below q)
    double lambda_tr = lambda * tr;
    if (lambda_tr > q - 1e-6) lambda_tr = q - 1e-6;
    out.grad[0] = 0.5 * (out.uQu + tr * ve - q * vu) / (vu * vu);
    out.grad[1] = 0.5 * (out.esse - ve * ((double)n - p - q + lambda_tr)) / (ve * ve);
    if (sol_out) copy_host(st, sol_out, w.S.p, dim * 8, hipMemcpyDeviceToHost);
    if (!want_ai) return;
    // AI matrix, vce.cpp:2596-2645: v_u = Z u / vu, v_e = e / ve; P v = (v - W C^-1 W'v) / ve
    const int64_t ldn = round_up(n, 64);
    k_ai_vectors<<<blocks(n), 256, 0, st>>>(M->aid.p, n, p, w.S.p, w.e.p, 1.0 / vu, 1.0 / ve, w.V.p, ldn);
    k_wt_apply<64><<<dim3((unsigned)ceil_div(dim, 4), 2), 256, 0, st>>>(M->tp.p, M->tc.p, M->tv.p, dim, w.V.p, ldn,
                                                                       w.B2.p, ld);
    CB_HIP(hipGetLastError());
    solve_block(M, lambda, o.solve, w.B2.p, w.S2.p, 2, w.have_s2);
    w.have_s2 = true;
    k_w_apply<<<dim3(blocks(n), 2), 256, 0, st>>>(M->xp.p, M->xc.p, M->xv.p, M->aid.p, n, p, w.S2.p, ld, w.T.p, ldn);
    std::vector<double> hV((size_t)ldn * 2), hT((size_t)ldn * 2);
    copy_host(st, hV.data(), w.V.p, ldn * 2 * 8, hipMemcpyDeviceToHost);
    copy_host(st, hT.data(), w.T.p, ldn * 2 * 8, hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    double a_uu = 0, a_ue = 0, a_ee = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double vu_i = hV[i], ve_i = hV[ldn + i];
        const double pvu = (vu_i - hT[i]) / ve, pve = (ve_i - hT[ldn + i]) / ve;
        a_uu += vu_i * pvu;
        a_ue += vu_i * pve;
        a_ee += ve_i * pve;
    }
    out.ai[0] = 0.5 * a_uu;
    out.ai[1] = out.ai[2] = 0.5 * a_ue;
    out.ai[3] = 0.5 * a_ee;
}

static void reml_setup(cb_mme* M, const cb_reml_options& o, cb_reml_work& w, const RemlPlan& pl) {
    cb_context* ctx = M->ctx;
    hipStream_t st = ctx->stream;
    const int64_t ld = M->ldv, n = M->n;
    CB_REQUIRE(o.block_cols >= 0, "REML: block_cols must be >= 0");
    w.adaptive = pl.trace == 1;
    w.exact = pl.trace == 2;
    w.min_k = pl.min_k;
    w.max_k = pl.max_k;
    w.k = (int)(pl.trace == 2 ? M->q : pl.trace == 1 ? pl.max_k : o.n_probes);
    const int r = 1 + w.k;
    w.bw = reml_block_width(M, o, r);
    w.chunked = w.bw < r;
    CB_REQUIRE(!(w.adaptive && w.chunked),
               "REML: adaptive traces need y and max(100, ns) probe columns in one block (raise block_cols)");
    const int bw = w.bw;
    w.B.alloc((size_t)ld * bw);
    w.S.alloc((size_t)ld * bw);
    w.Qz.alloc((size_t)ld * std::max(1, std::min(bw, w.k)));
    w.rhs.alloc(ld);
    w.Qu.alloc(ld);
    w.B2.alloc((size_t)ld * 2);
    w.S2.alloc((size_t)ld * 2);
    const int64_t ldn = round_up(n, 64);
    w.V.alloc((size_t)ldn * 2);
    w.T.alloc((size_t)ldn * 2);
    w.e.alloc(ldn);
    w.part.alloc((size_t)DOT_CHUNKS * (bw + 1));
    w.dots.alloc(bw + 1);
    CB_HIP(hipMemsetAsync(w.B.p, 0, w.B.bytes(), st));
    CB_HIP(hipMemsetAsync(w.Qz.p, 0, w.Qz.bytes(), st));
    CB_HIP(hipMemsetAsync(w.B2.p, 0, w.B2.bytes(), st));
    CB_HIP(hipMemsetAsync(w.rhs.p, 0, w.rhs.bytes(), st));
    copy_host(st, w.rhs.p, M->rhs.data(), M->dim * 8, hipMemcpyHostToDevice);
    CB_HIP(hipMemcpyAsync(w.B.p, w.rhs.p, ld * 8, hipMemcpyDeviceToDevice, st));
    // one block: the probes and Q z (fixed for the run) are made once
    if (!w.chunked && w.k > 0) make_probes(M, o, w, 0, w.k, 1, true);
    CB_HIP(hipStreamSynchronize(st));
}

extern "C" {

void cb_solve_options_default(cb_solve_options* o) {
    if (!o) return;
    o->precision = CB_PREC_MIXED;
    o->precond = CB_PRECOND_BLOCK_JACOBI;
    o->stop_rule = CB_STOP_RELATIVE;
    o->batch_rhs = 1;
    o->tol = 1e-10;
    o->max_iter = 10000;
    o->stall_tol = 0.0;  // strict: a stall above tol is CB_ERR_NOT_CONVERGED
}

void cb_reml_options_default(cb_reml_options* o) {
    if (!o) return;
    o->algorithm = CB_VCE_AI;
    o->max_iter = 20;
    o->tol = 1e-4;
    o->n_probes = 64;
    o->seed = 42;
    o->trace_exact = 0;
    cb_solve_options_default(&o->solve);
    o->mc_samples = 100;  // vce_mc_samples (config.h:56)
    o->adaptive = 0;
    // REML solves accept the FP64 rounding floor up to 1e-8: the probes'
    // sampling noise in the trace is orders of magnitude larger
    o->solve.stall_tol = 1e-8;
    o->block_cols = 0;
}

int cb_mme_create(cb_context* ctx, int64_t n, int64_t p, int64_t q, const int64_t* x_ptr,
                  const int32_t* x_col, const double* x_val, const int32_t* aid, const double* y,
                  cb_mme** out) {
    return guarded([&] {
        CB_REQUIRE(ctx && out && x_ptr && aid && y, "cb_mme_create: NULL argument");
        CB_REQUIRE(n > 0 && p >= 0 && q > 0, "cb_mme_create: bad dimensions");
        CB_REQUIRE(x_ptr[0] == 0 && x_ptr[n] >= 0 && (x_ptr[n] == 0 || (x_col && x_val)),
                   "cb_mme_create: bad fixed-effect CSR");
        set_device(ctx);
        init_scratch(ctx->device);  // this device's scratch address (the allocation-free solve)
        auto* M = new cb_mme();
        M->ctx = ctx;
        M->n = n; M->p = p; M->q = q; M->dim = p + q;
        M->ldv = round_up(M->dim, 64);
        try {
            assemble(M, x_ptr, x_col, x_val, aid, y);
        } catch (...) {
            delete M;
            throw;
        }
        // the pinned page of the solvers' state exchanges (without it they
        // copy through the staging buffer)
        if (hipHostMalloc(&M->hpage, 2 * HPAGE_HALF, hipHostMallocCoherent) != hipSuccess) {
            M->hpage = nullptr;
            (void)hipGetLastError();
        }
        *out = M;
    });
}

void cb_mme_destroy(cb_mme* M) {
    if (!M) return;
    (void)hipSetDevice(M->ctx->device);
    tls_stream = M->ctx->stream;  // its buffers are released behind the context's queued work
    delete M;
}

int cb_mme_dims(const cb_mme* M, int64_t* n, int64_t* p, int64_t* q) {
    return guarded([&] {
        CB_REQUIRE(M, "cb_mme_dims: NULL");
        if (n) *n = M->n;
        if (p) *p = M->p;
        if (q) *q = M->q;
    });
}

// Dominant eigenpair of a dense Q by power iteration from the constant
// vector (the centring null direction of a VanRaden G). It is deflated in
// the preconditioner when it stands far above the rest of the spectrum:
// mu > 1e3 * (trace(Q) - mu) / q.
static void find_stiff_direction(cb_mme* M) {
    const int64_t q = M->q;
    hipStream_t st = M->ctx->stream;
    M->stiff = false;
    if (q < 2) return;
    std::vector<double> v(q, 1.0 / std::sqrt((double)q)), y(q);
    DevBuf<double> dv, dy;
    dv.alloc(q);
    dy.alloc(q);
    double mu = 0, mu_prev = 0;
    for (int it = 0; it < 100; ++it) {
        copy_host(st, dv.p, v.data(), q * 8, hipMemcpyHostToDevice);
        CB_HIP(hipMemsetAsync(dy.p, 0, q * 8, st));
        k_gemv<double><<<(unsigned)ceil_div(q, 4 * GEMV_ROWS), 256, 0, st>>>(M->Qd64.p, M->ldq, q, dv.p, q, 1, dy.p, q,
                                                                             0, 1.0);
        CB_HIP(hipGetLastError());
        copy_host(st, y.data(), dy.p, q * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        double vy = 0, yy = 0;
        for (int64_t i = 0; i < q; ++i) { vy += v[i] * y[i]; yy += y[i] * y[i]; }
        mu = vy;
        const double ny = std::sqrt(yy);
        if (!(ny > 0)) return;
        for (int64_t i = 0; i < q; ++i) v[i] = y[i] / ny;
        if (it > 2 && std::fabs(mu - mu_prev) <= 1e-13 * std::fabs(mu)) break;
        mu_prev = mu;
    }
    double tr = 0;
    for (double d : M->q_diag) tr += d;
    const double bulk = (tr - mu) / (double)q;
    if (!(mu > 1e3 * std::fabs(bulk))) return;
    M->stiff = true;
    M->stiff_mu = mu;
    M->stiff_v.alloc(q);
    copy_host(M->ctx->stream, M->stiff_v.p, v.data(), q * 8, hipMemcpyHostToDevice);
    // FP32 copy of the bulk only: Q32 = fp32(Q - mu v v'); the rank-1 part is
    // applied separately with FP64 accumulation (apply_C<float>)
    k_q32_bulk<<<blocks(q * M->ldq), 256, 0, st>>>(M->Qd64.p, M->Qd32.p, M->stiff_v.p, mu, q, M->ldq);
    M->q16_rows = 0;
    CB_HIP(hipGetLastError());
    CB_HIP(hipStreamSynchronize(st));
}

// The preconditioner's buffers at their fixed sizes (p x p, dim, q), so
// that building it for a new lambda inside a solve allocates nothing.
static void presize_precond(cb_mme* M) {
    M->Binv.alloc((size_t)std::max<int64_t>(1, M->p * M->p));
    M->dinv.alloc(M->dim);
    if (M->stiff) M->stiff_y.alloc(M->q);
}

static void set_q_csr_common(cb_mme* M, const int64_t* ptr, const int32_t* col, const double* val) {
    set_device(M->ctx);
    const int64_t q = M->q, nnz = ptr[q];
    CB_REQUIRE(ptr[0] == 0 && nnz >= 0, "Qinv CSR: bad row pointers");
    M->q_diag.assign(q, 0.0);
    for (int64_t i = 0; i < q; ++i)
        for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            CB_REQUIRE(col[k] >= 0 && col[k] < q, "Qinv CSR: column out of range");
            if (col[k] == i) M->q_diag[i] += val[k];
        }
    hipStream_t st = M->ctx->stream;
    M->q_lanes = csr_lanes(nnz, q);
    M->qp.alloc(q + 1);
    M->qc.alloc(std::max<int64_t>(1, nnz));
    M->qv64.alloc(std::max<int64_t>(1, nnz));
    M->qv32.alloc(std::max<int64_t>(1, nnz));
    copy_host(st, M->qp.p, ptr, (q + 1) * 8, hipMemcpyHostToDevice);
    copy_host(st, M->qc.p, col, nnz * 4, hipMemcpyHostToDevice);
    copy_host(st, M->qv64.p, val, nnz * 8, hipMemcpyHostToDevice);
    k_convert<float, double><<<blocks(nnz), 256, 0, st>>>(M->qv64.p, M->qv32.p, nnz);
    CB_HIP(hipGetLastError());
    CB_HIP(hipStreamSynchronize(st));
    M->stiff = false;
    M->Qd64.release();
    M->Qd32.release();
    M->ng = 0;
    M->dinv_lambda = std::nan("");
    presize_precond(M);
}

static void set_q_dense_common(cb_mme* M, const double* Q, int64_t ld, hipMemcpyKind kind) {
    const int64_t q = M->q;
    CB_REQUIRE(Q && ld >= q, "cb_mme_set_qinv_dense: bad matrix");
    set_device(M->ctx);
    hipStream_t st = M->ctx->stream;
    // rows padded to a multiple of 32 elements with zeros (k_dense_mfma reads
    // whole 32-k chunks)
    const int64_t ldq = round_up(q, MF_KC);
    M->ldq = ldq;
    M->Qd64.alloc((size_t)q * ldq);
    M->Qd32.alloc((size_t)q * ldq);
    if (ldq > q) CB_HIP(hipMemsetAsync(M->Qd64.p, 0, M->Qd64.bytes(), st));
    copy_host_2d(st, M->Qd64.p, ldq * 8, Q, ld * 8, q * 8, q, kind);
    k_convert<float, double><<<blocks(q * ldq), 256, 0, st>>>(M->Qd64.p, M->Qd32.p, q * ldq);
    M->q16_rows = 0;
    CB_HIP(hipGetLastError());
    M->q_diag.resize(q);
    DevBuf<double> dg;
    dg.alloc(q);
    CB_HIP(hipMemcpy2DAsync(dg.p, 8, M->Qd64.p, (ldq + 1) * 8, 8, q, hipMemcpyDeviceToDevice, st));
    copy_host(st, M->q_diag.data(), dg.p, q * 8, hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    M->qkind = 0;
    M->qp.release(); M->qc.release(); M->qv64.release(); M->qv32.release();
    M->dinv_lambda = std::nan("");
    find_stiff_direction(M);
    presize_precond(M);
}

int cb_mme_set_qinv_dense(cb_mme* M, const double* Q, int64_t ld) {
    return guarded([&] {
        CB_REQUIRE(M, "NULL mme");
        set_q_dense_common(M, Q, ld, hipMemcpyHostToDevice);
    });
}

int cb_mme_set_qinv_dense_device(cb_mme* M, const double* dQ, int64_t ld) {
    return guarded([&] {
        CB_REQUIRE(M, "NULL mme");
        set_q_dense_common(M, dQ, ld, hipMemcpyDeviceToDevice);
    });
}

int cb_mme_set_qinv_csr(cb_mme* M, const int64_t* ptr, const int32_t* col, const double* val) {
    return guarded([&] {
        CB_REQUIRE(M && ptr && col && val, "cb_mme_set_qinv_csr: NULL argument");
        set_q_csr_common(M, ptr, col, val);
        M->qkind = 1;
    });
}

int cb_mme_set_qinv_split(cb_mme* M, const int64_t* ptr, const int32_t* col, const double* val, int64_t ng,
                          const int32_t* gmap, const double* dD, int64_t ldd) {
    return guarded([&] {
        CB_REQUIRE(M && ptr && col && val && gmap && dD, "cb_mme_set_qinv_split: NULL argument");
        CB_REQUIRE(ng > 0 && ldd >= ng, "cb_mme_set_qinv_split: bad block dimensions");
        std::vector<char> seen(M->q, 0);
        for (int64_t i = 0; i < ng; ++i) {
            CB_REQUIRE(gmap[i] >= -1 && gmap[i] < M->q, "cb_mme_set_qinv_split: map entry out of range");
            if (gmap[i] >= 0) {
                CB_REQUIRE(!seen[gmap[i]], "cb_mme_set_qinv_split: map entries must be distinct");
                seen[gmap[i]] = 1;
            }
        }
        set_q_csr_common(M, ptr, col, val);
        hipStream_t st = M->ctx->stream;
        const int64_t ldq = round_up(ng, MF_KC);
        M->ng = ng;
        M->ldq = ldq;
        M->Qd64.alloc((size_t)ng * ldq);
        M->Qd32.alloc((size_t)ng * ldq);
        if (ldq > ng) CB_HIP(hipMemsetAsync(M->Qd64.p, 0, M->Qd64.bytes(), st));
        CB_HIP(hipMemcpy2DAsync(M->Qd64.p, ldq * 8, dD, ldd * 8, ng * 8, ng, hipMemcpyDeviceToDevice, st));
        k_convert<float, double><<<blocks(ng * ldq), 256, 0, st>>>(M->Qd64.p, M->Qd32.p, ng * ldq);
        M->q16_rows = 0;
        CB_HIP(hipGetLastError());
        M->gmap.alloc(ng);
        copy_host(st, M->gmap.p, gmap, ng * 4, hipMemcpyHostToDevice);
        std::vector<double> dg(ng);
        copy_host_2d(st, dg.data(), 8, M->Qd64.p, (ldq + 1) * 8, 8, ng, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < ng; ++i)
            if (gmap[i] >= 0) M->q_diag[gmap[i]] += dg[i];  // add_diagonal_to (matrix_adapter.h:280-300)
        M->qkind = 2;
    });
}

int cb_mme_rhs(const cb_mme* M, double* rhs) {
    return guarded([&] {
        CB_REQUIRE(M && rhs, "cb_mme_rhs: NULL argument");
        std::copy(M->rhs.begin(), M->rhs.end(), rhs);
    });
}

int cb_mme_apply(cb_mme* M, double lambda, int64_t r, const double* X, double* Y) {
    return guarded([&] {
        CB_REQUIRE(M && X && Y && r > 0, "cb_mme_apply: bad argument");
        CB_REQUIRE(M->qkind >= 0, "cb_mme_apply: Qinv not set");
        set_device(M->ctx);
        hipStream_t st = M->ctx->stream;
        const int64_t ld = M->ldv;
        DevBuf<double> dX, dY;
        dX.alloc((size_t)ld * r);
        dY.alloc((size_t)ld * r);
        CB_HIP(hipMemsetAsync(dX.p, 0, dX.bytes(), st));
        copy_host_2d(st, dX.p, ld * 8, X, M->dim * 8, M->dim * 8, r, hipMemcpyHostToDevice);
        apply_C<double>(M, lambda, dX.p, dY.p, (int)r);
        copy_host_2d(st, Y, M->dim * 8, dY.p, ld * 8, M->dim * 8, r, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"

// The tuner's probe: one dense operator shape with a given split.
template <class T>
static double tune_dense_time(cb_context* ctx, int64_t q, int r, const DenseSplit& sp, int reps, bool f16 = false) {
    hipStream_t st = ctx->stream;
    cb_mme M;
    M.ctx = ctx;
    M.q = q;
    M.ldq = round_up(q, 32);
    M.qkind = 0;
    const int64_t ldx = round_up(q, 64);
    DevBuf<T> X, Y;
    X.alloc((size_t)ldx * r);
    Y.alloc((size_t)ldx * r);
    DevBuf<T>& Qb = [&]() -> DevBuf<T>& {
        if constexpr (sizeof(T) == 8) return M.Qd64;
        else return M.Qd32;
    }();
    Qb.alloc((size_t)q * M.ldq);
    k_tune_fill<T><<<blocks((int64_t)Qb.n), 256, 0, st>>>(Qb.p, (int64_t)Qb.n, 1);
    k_tune_fill<T><<<blocks((int64_t)X.n), 256, 0, st>>>(X.p, (int64_t)X.n, 2);
    CB_HIP(hipMemsetAsync(Y.p, 0, Y.bytes(), st));
    CB_HIP(hipGetLastError());
    if constexpr (sizeof(T) == 4) {
        if (f16) {  // the FP16-halves product of CB_PREC_MIXED_F16 solves
            ensure_q16(&M, q);
            M.use_f16 = true;
        }
    }
    struct Probe {
        Probe(const DenseSplit* s) { g_split_probe = s; }
        ~Probe() { g_split_probe = nullptr; }
    } probe(&sp);
    dense_apply<T>(&M, q, X.p, ldx, Y.p, ldx, r, (T)1);  // warm-up (sizes the partial-sum buffer)
    hipEvent_t e0, e1;
    CB_HIP(hipEventCreate(&e0));
    CB_HIP(hipEventCreate(&e1));
    float ms = 0.f;
    CB_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < reps; ++i) dense_apply<T>(&M, q, X.p, ldx, Y.p, ldx, r, (T)1);
    CB_HIP(hipEventRecord(e1, st));
    const hipError_t e = hipEventSynchronize(e1);
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    CB_HIP(e);
    return (double)ms / reps;
}

extern "C" {

int cb_tune_dense_time(cb_context* ctx, int elem_bytes, int64_t q, int r, int target_blocks, int min_kspan, int reps,
                       double* ms) {
    return guarded([&] {
        CB_REQUIRE(ctx && ms, "cb_tune_dense_time: NULL argument");
        CB_REQUIRE(elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8,
                   "cb_tune_dense_time: elem_bytes must be 2 (FP16 halves), 4 or 8");
        CB_REQUIRE(q >= 64 && q <= 200000 && r >= 3 && r <= MF_MAXNT * 16 && reps >= 1,
                   "cb_tune_dense_time: q in [64, 200000], r in [3, 80], reps >= 1");
        CB_REQUIRE(target_blocks >= 1 && min_kspan >= MF_KC, "cb_tune_dense_time: target_blocks >= 1, min_kspan >= 32");
        set_device(ctx);
        const DenseSplit sp{target_blocks, min_kspan};
        *ms = elem_bytes == 8 ? tune_dense_time<double>(ctx, q, r, sp, reps)
                              : tune_dense_time<float>(ctx, q, r, sp, reps, elem_bytes == 2);
    });
}

int cb_tune_dense_lookup(int elem_bytes, int64_t q, int r, int* target_blocks, int* min_kspan) {
    return guarded([&] {
        CB_REQUIRE(target_blocks && min_kspan, "cb_tune_dense_lookup: NULL argument");
        const DenseSplit sp = dense_split(elem_bytes, q, std::min(r, MF_MAXNT * 16));
        *target_blocks = sp.target_blocks;
        *min_kspan = sp.min_kspan;
    });
}

}  // extern "C"

// cb_mme_solve / cb_mme_solve_device: B and X are dim x r column blocks with
// leading dimensions ldb, ldx, in host memory (kind H2D / D2H) or in device
// memory (D2D). The work blocks pad each column to ldv rows with zeros.
static int mme_solve_blocks(cb_mme* M, double lambda, const cb_solve_options* opts, int64_t r, const double* B,
                            int64_t ldb, double* X, int64_t ldx, bool on_device, int64_t* iterations,
                            double* rel_residual, const char* who) {
    cb_solve_options o;
    cb_solve_options_default(&o);
    if (opts) o = *opts;
    int64_t its = 0;
    double rel = 0;
    const int st = guarded([&] {
        CB_REQUIRE(M && B && X && r > 0, std::string(who) + ": bad argument");
        CB_REQUIRE(ldb >= M->dim && ldx >= M->dim, std::string(who) + ": leading dimension < p + q");
        CB_REQUIRE(M->qkind >= 0, std::string(who) + ": Qinv not set");
        CB_REQUIRE(lambda >= 0 && std::isfinite(lambda), std::string(who) + ": lambda must be finite and >= 0");
        set_device(M->ctx);
        hipStream_t s = M->ctx->stream;
        const int64_t ld = M->ldv;
        const hipMemcpyKind in = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const hipMemcpyKind out = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        const WsPtr<double> dB = ws<double>(M, WS_IO_B, (size_t)ld * r), dX = ws<double>(M, WS_IO_X, (size_t)ld * r);
        CB_HIP(hipMemsetAsync(dB.p, 0, dB.bytes(), s));
        copy_host_2d(s, dB.p, ld * 8, B, ldb * 8, M->dim * 8, r, in);
        try {
            SolveStats S = solve_block(M, lambda, o, dB.p, dX.p, (int)r);
            its = S.iterations;
            rel = S.max_rel;
        } catch (const Error& e) {
            // hand back the unconverged iterate (matrix_free_pcg.h:64-65 returns it silently)
            copy_host_2d(s, X, ldx * 8, dX.p, ld * 8, M->dim * 8, r, out);
            CB_HIP(hipStreamSynchronize(s));
            throw;
        }
        copy_host_2d(s, X, ldx * 8, dX.p, ld * 8, M->dim * 8, r, out);
        CB_HIP(hipStreamSynchronize(s));
    });
    if (iterations) *iterations = its;
    if (rel_residual) *rel_residual = rel;
    return st;
}

extern "C" {

int cb_mme_solve(cb_mme* M, double lambda, const cb_solve_options* opts, int64_t r, const double* B,
                 double* X, int64_t* iterations, double* rel_residual) {
    const int64_t dim = M ? M->dim : 0;
    return mme_solve_blocks(M, lambda, opts, r, B, dim, X, dim, false, iterations, rel_residual, "cb_mme_solve");
}

int cb_mme_solve_device(cb_mme* M, double lambda, const cb_solve_options* opts, int64_t r, const double* dB,
                        int64_t ldb, double* dX, int64_t ldx, int64_t* iterations, double* rel_residual) {
    return mme_solve_blocks(M, lambda, opts, r, dB, ldb, dX, ldx, true, iterations, rel_residual,
                            "cb_mme_solve_device");
}

// Upper bound of the scratch bytes one solve of r columns takes (the
// buffers solve_block and apply_C ask for, 256-byte aligned each).
static size_t noalloc_need(const cb_mme* M, int64_t r, int precision) {
    const size_t ld = (size_t)M->ldv, R = (size_t)r, a = 256;
    const size_t vec = ld * R;
    size_t need = 2 * (vec * 8 + a);  // the right-hand-side and solution blocks
    need += (precision == CB_PREC_FP64 ? 4 * (vec * 8 + a) : 4 * (vec * 4 + a) + 2 * (vec * 8 + a));
    need += ((size_t)DOT_CHUNKS * R * 8 + a) + (part2_count(M, (int)r) * 8 + a) + (3 * R * 8 + 2 * a) + R * sizeof(ColState) + a;
    if (M->long_rows > 0) need += (size_t)std::max<int64_t>(1, M->n_lchunks) * R * 8 + a;
    if (M->qkind == 0) {
        if (M->stiff) need += (size_t)DOT_CHUNKS * R * 8 + a;
        if (r > 2) {
            const int rc = (int)std::min<int64_t>(r, MF_MAXNT * 16);
            const DenseShape d8 = dense_shape(8, M->q, rc), d4 = dense_shape(4, M->q, rc);
            need += (size_t)d8.nsplit * rc * d8.ldp * 8 + a;
            if (precision != CB_PREC_FP64) need += (size_t)d4.nsplit * rc * d4.ldp * 4 + a;
        }
    }
    return need;
}

// cb_mme_solve_device with its scratch in this device's fixed scratch array
// instead of the model's growing buffers: no device allocation or release.
int cb_mme_solve_device_noalloc(cb_mme* M, double lambda, const cb_solve_options* opts, int64_t r,
                                const double* dB, int64_t ldb, double* dX, int64_t ldx, int64_t* iterations,
                                double* rel_residual) {
    cb_solve_options o;
    cb_solve_options_default(&o);
    if (opts) o = *opts;
    ScratchArena arena;
    bool graphs = false;
    const int st = guarded([&] {
        CB_REQUIRE(M && dB && dX && r > 0, "cb_mme_solve_device_noalloc: bad argument");
        CB_REQUIRE(M->qkind == 0 || M->qkind == 1,
                   "cb_mme_solve_device_noalloc: dense or CSR Qinv only (the split form gathers into growing buffers)");
        CB_REQUIRE(o.precision == CB_PREC_FP64 || o.precision == CB_PREC_MIXED,
                   "cb_mme_solve_device_noalloc: CB_PREC_FP64 or CB_PREC_MIXED only");
        CB_REQUIRE(M->dinv.p && M->Binv.p && (!M->stiff || M->stiff_y.p),
                   "cb_mme_solve_device_noalloc: preconditioner buffers missing (set Qinv first)");
        unsigned char* base = scratch_ptr(M->ctx->device);
        CB_REQUIRE(base, "cb_mme_solve_device_noalloc: no scratch on this device");
        const size_t need = noalloc_need(M, r, o.precision);
        if (need > NOALLOC_SCRATCH_BYTES)
            throw Error(CB_ERR_INVALID, "cb_mme_solve_device_noalloc: " + std::to_string(r) + " columns of " +
                                            std::to_string(M->dim) + " unknowns need about " + std::to_string(need) +
                                            " bytes of scratch, more than its " +
                                            std::to_string(NOALLOC_SCRATCH_BYTES) + "; use cb_mme_solve_device");
        arena.base = reinterpret_cast<char*>(base);
        arena.cap = NOALLOC_SCRATCH_BYTES;
        graphs = M->ctx->graphs;
    });
    if (st != CB_OK) return st;
    // scratch from the arena; no graph capture (instantiation may allocate)
    struct Scope {
        cb_mme* M;
        bool graphs;
        ~Scope() {
            M->arena = nullptr;
            M->ctx->graphs = graphs;
        }
    } scope{M, graphs};
    M->arena = &arena;
    M->ctx->graphs = false;
    return mme_solve_blocks(M, lambda, &o, r, dB, ldb, dX, ldx, true, iterations, rel_residual,
                            "cb_mme_solve_device_noalloc");
}

// GSSolve's fixed-effect standard errors (mixed_model_solver.cpp:346-417):
// column i of C^-1 for each of the p fixed effects, from C x = e_i, as
// right-hand-side blocks of at most BETA_COV_COLS unit columns.
int cb_mme_beta_cov(cb_mme* M, double lambda, double var_e, const cb_solve_options* opts, double* cinv_beta,
                    double* beta_se, int64_t* iterations) {
    cb_solve_options o;
    cb_solve_options_default(&o);
    if (opts) o = *opts;
    return guarded([&] {
        CB_REQUIRE(M && (cinv_beta || beta_se), "cb_mme_beta_cov: NULL argument");
        CB_REQUIRE(M->qkind >= 0, "cb_mme_beta_cov: Qinv not set");
        CB_REQUIRE(lambda >= 0 && std::isfinite(lambda) && var_e > 0, "cb_mme_beta_cov: bad lambda or var_e");
        set_device(M->ctx);
        hipStream_t st = M->ctx->stream;
        const int64_t p = M->p, ld = M->ldv;
        constexpr int64_t BETA_COV_COLS = 256;
        std::vector<double> cinv((size_t)p * p);
        int64_t its = 0;
        for (int64_t c0 = 0; c0 < p; c0 += BETA_COV_COLS) {
            const int r = (int)std::min<int64_t>(BETA_COV_COLS, p - c0);
            const WsPtr<double> dB = ws<double>(M, WS_IO_B, (size_t)ld * r), dX = ws<double>(M, WS_IO_X, (size_t)ld * r);
            CB_HIP(hipMemsetAsync(dB.p, 0, (size_t)ld * r * 8, st));
            k_unit_cols<<<blocks(r), 256, 0, st>>>(dB.p, ld, c0, r);
            CB_HIP(hipGetLastError());
            const SolveStats S = solve_block(M, lambda, o, dB.p, dX.p, r);
            its = std::max(its, S.iterations);
            // rows 0..p of each solution column
            std::vector<double> hx((size_t)p * r);
            copy_host_2d(st, hx.data(), p * 8, dX.p, ld * 8, p * 8, r, hipMemcpyDeviceToHost);
            CB_HIP(hipStreamSynchronize(st));
            for (int c = 0; c < r; ++c)
                for (int64_t j = 0; j < p; ++j) cinv[(size_t)(c0 + c) * p + j] = hx[(size_t)c * p + j];
        }
        if (cinv_beta) std::copy(cinv.begin(), cinv.end(), cinv_beta);  // column i = C^-1 e_i (p x p)
        if (beta_se)
            for (int64_t i = 0; i < p; ++i) beta_se[i] = std::sqrt(std::max(0.0, var_e * cinv[(size_t)i * p + i]));
        if (iterations) *iterations = its;
    });
}

}  // extern "C"

namespace {

// C(lambda) = W'W + lambda blockdiag(0, Q) as a dense column-major dim x dim
// matrix (C must hold dim * dim doubles).
void assemble_dense_C(cb_mme* M, double lambda, double* C) {
    hipStream_t st = M->ctx->stream;
    const int64_t dim = M->dim, p = M->p, q = M->q;
    CB_HIP(hipMemsetAsync(C, 0, (size_t)dim * dim * 8, st));
    k_dense_from_csr<<<blocks(dim), 256, 0, st>>>(M->wp.p, M->wc.p, M->wv64.p, dim, 0, 1.0, C, dim);
    if (M->qkind == 0) {
        k_dense_add_q<<<dim3(blocks(q), (unsigned)q), 256, 0, st>>>(M->Qd64.p, M->ldq, q, lambda, C + p * dim + p, dim);
    } else {
        k_dense_from_csr<<<blocks(q), 256, 0, st>>>(M->qp.p, M->qc.p, M->qv64.p, q, 0, lambda, C + p * dim + p, dim);
        if (M->qkind == 2)
            k_dense_add_block<<<dim3(blocks(M->ng), (unsigned)M->ng), 256, 0, st>>>(M->Qd64.p, M->ldq, M->ng,
                                                                                  M->gmap.p, lambda, C + p * dim + p,
                                                                                  dim);
    }
    CB_HIP(hipGetLastError());
}

// W = [X Z] dense (n x dim, column-major) from the design rows
__global__ void k_build_w(const int64_t* __restrict__ xp, const int32_t* __restrict__ xc,
                          const double* __restrict__ xv, const int32_t* __restrict__ aid, int64_t n, int64_t p,
                          double* __restrict__ W) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int64_t k = xp[r]; k < xp[r + 1]; ++k) W[(int64_t)xc[k] * n + r] += xv[k];
    if (aid[r] >= 0) W[(p + aid[r]) * n + r] = 1.0;
}

// out (cols x rows) = in^T, in rows x cols, both column-major
__global__ void k_transpose(const double* __restrict__ in, int64_t rows, int64_t cols, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= rows) return;
    out[i * cols + j] = in[j * rows + i];
}

// Z A Z' for incidence Z: ZAZ[i, j] = A[aid_i, aid_j] (0 for records without a level)
__global__ void k_zazt(const double* __restrict__ A, int64_t lda, const int32_t* __restrict__ aid, int64_t n,
                       double* __restrict__ ZAZ) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    const int32_t ai = aid[i], aj = aid[j];
    ZAZ[j * n + i] = (ai >= 0 && aj >= 0) ? A[(int64_t)aj * lda + ai] : 0.0;
}

// copy the lower triangle onto the upper (column-major)
__global__ void k_sym_from_lower(double* __restrict__ A, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n || i >= j) return;
    A[j * n + i] = A[i * n + j];
}

// P = (I - H) / ve in place
__global__ void k_p_from_h(double* __restrict__ H, int64_t n, double ve) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    const double h = H[j * n + i];
    H[j * n + i] = ((i == j ? 1.0 : 0.0) - h) / ve;
}

// partial sums over column chunks: sum Mu_ij Mu_ji, sum Mu_ij P_ij, sum P_ij^2, tr Mu, tr P
__global__ __launch_bounds__(256) void k_exact_traces(const double* __restrict__ Mu, const double* __restrict__ P,
                                                      int64_t n, double* __restrict__ part) {
    __shared__ double red[5][256];
    const int64_t per = (n + DOT_CHUNKS - 1) / DOT_CHUNKS;
    const int64_t j0 = (int64_t)blockIdx.x * per, j1 = min(n, j0 + per);
    double uu = 0, ue = 0, ee = 0, tu = 0, te = 0;
    for (int64_t j = j0; j < j1; ++j)
        for (int64_t i = threadIdx.x; i < n; i += 256) {
            const double mij = Mu[j * n + i], mji = Mu[i * n + j], pij = P[j * n + i];
            uu += mij * mji;
            ue += mij * pij;
            ee += pij * pij;
            if (i == j) {
                tu += mij;
                te += pij;
            }
        }
    red[0][threadIdx.x] = uu;
    red[1][threadIdx.x] = ue;
    red[2][threadIdx.x] = ee;
    red[3][threadIdx.x] = tu;
    red[4][threadIdx.x] = te;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5) part[blockIdx.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

#define CB_RB(call, what)                                                                              \
    do {                                                                                               \
        if ((call) != rocblas_status_success) throw Error(CB_ERR_DEVICE, std::string(what " failed")); \
        blas_drain();                                                                                  \
    } while (0)

// run_dense_exact_step (vce.cpp:2859-3195) on dense device matrices.
struct DenseExact {
    cb_mme* M;
    int64_t n, p, q, dim;
    std::unique_ptr<BlasHandle> bh;  // workspace reserved per call (CB_BLAS)
    rocblas_handle h = nullptr;      // bh->h
    DevBuf<double> W, Wt, ZAZ, C, B, H, Mu, b, e, t, part;
    DevBuf<int> info;
    std::vector<double> y;
    double logdetA = 0;
    explicit DenseExact(cb_mme* m) : M(m), n(m->n), p(m->p), q(m->q), dim(m->dim) {}

    hipStream_t st() const { return M->ctx->stream; }

    void init() {
        CB_REQUIRE(n <= 40000 && dim <= 40000, "dense-exact AI-REML: n and p + q are limited to 40000");
        bh = std::make_unique<BlasHandle>(st());
        h = bh->h;
        info.alloc(1);
        // W and W'
        W.alloc((size_t)n * dim);
        CB_HIP(hipMemsetAsync(W.p, 0, W.bytes(), st()));
        k_build_w<<<blocks(n), 256, 0, st()>>>(M->xp.p, M->xc.p, M->xv.p, M->aid.p, n, p, W.p);
        Wt.alloc((size_t)dim * n);
        k_transpose<<<dim3(blocks(n), (unsigned)dim), 256, 0, st()>>>(W.p, n, dim, Wt.p);
        CB_HIP(hipGetLastError());
        // A = Q^-1 (LLT) and log|A| = -log|Q| (initialize_dense, vce.cpp:2888-2895,2963-2970)
        DevBuf<double> A;
        A.alloc((size_t)q * q);
        CB_HIP(hipMemsetAsync(A.p, 0, A.bytes(), st()));
        if (M->qkind == 0) {
            k_dense_add_q<<<dim3(blocks(q), (unsigned)q), 256, 0, st()>>>(M->Qd64.p, M->ldq, q, 1.0, A.p, q);
        } else {
            k_dense_from_csr<<<blocks(q), 256, 0, st()>>>(M->qp.p, M->qc.p, M->qv64.p, q, 0, 1.0, A.p, q);
            if (M->qkind == 2)
                k_dense_add_block<<<dim3(blocks(M->ng), (unsigned)M->ng), 256, 0, st()>>>(M->Qd64.p, M->ldq, M->ng,
                                                                                        M->gmap.p, 1.0, A.p, q);
        }
        CB_HIP(hipGetLastError());
        CB_BLAS(*bh, rocsolver_dpotrf(h, rocblas_fill_lower, (rocblas_int)q, A.p, (rocblas_int)q, info.p), "dpotrf(Q)");
        check_info("Qinv is not positive definite");
        logdetA = -logdiag2(A.p, q);
        CB_BLAS(*bh, rocsolver_dpotri(h, rocblas_fill_lower, (rocblas_int)q, A.p, (rocblas_int)q, info.p), "dpotri(Q)");
        k_sym_from_lower<<<dim3(blocks(q), (unsigned)q), 256, 0, st()>>>(A.p, q);
        ZAZ.alloc((size_t)n * n);
        k_zazt<<<dim3(blocks(n), (unsigned)n), 256, 0, st()>>>(A.p, q, M->aid.p, n, ZAZ.p);
        CB_HIP(hipGetLastError());
        C.alloc((size_t)dim * dim);
        B.alloc((size_t)dim * n);
        H.alloc((size_t)n * n);
        Mu.alloc((size_t)n * n);
        b.alloc(dim);
        e.alloc(n);
        t.alloc(n);
        part.alloc((size_t)DOT_CHUNKS * 5);
        y.resize(n);
        copy_host(st(), y.data(), M->y.p, n * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st()));
    }
    void check_info(const char* msg) {
        int hi = 0;
        copy_host(st(), &hi, info.p, sizeof(int), hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st()));
        CB_REQUIRE(hi == 0, msg);
    }
    // 2 sum log L_ii of a Cholesky factor
    double logdiag2(const double* L, int64_t m) {
        DevBuf<double> pp;
        pp.alloc(DOT_CHUNKS);
        k_logdiag_part<<<DOT_CHUNKS, 256, 0, st()>>>(L, m, pp.p);
        std::vector<double> hp(DOT_CHUNKS);
        copy_host(st(), hp.data(), pp.p, DOT_CHUNKS * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st()));
        double sum = 0;
        for (double v : hp) sum += v;
        return 2.0 * sum;
    }
    // factor C(ve/vu), solve the MME, residuals e; returns false when C is
    // not positive definite. logL of the dense-exact step (vce.cpp:2972-3003).
    bool solve_logl(double vu, double ve, double* logl, std::vector<double>* sol) {
        assemble_dense_C(M, ve / vu, C.p);
        CB_BLAS(*bh, rocsolver_dpotrf(h, rocblas_fill_lower, (rocblas_int)dim, C.p, (rocblas_int)dim, info.p), "dpotrf(C)");
        int hi = 0;
        copy_host(st(), &hi, info.p, sizeof(int), hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st()));
        if (hi != 0) return false;
        const double ldC = logdiag2(C.p, dim);
        copy_host(st(), b.p, M->rhs.data(), dim * 8, hipMemcpyHostToDevice);
        CB_BLAS(*bh, rocsolver_dpotrs(h, rocblas_fill_lower, (rocblas_int)dim, 1, C.p, (rocblas_int)dim, b.p,
                               (rocblas_int)dim), "dpotrs(C, rhs)");
        // e = y - W sol
        CB_HIP(hipMemcpyAsync(e.p, M->y.p, n * 8, hipMemcpyDeviceToDevice, st()));
        const double one = 1.0, mone = -1.0;
        CB_RB(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "pointer mode");
        CB_BLAS(*bh, rocblas_dgemv(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)dim, &mone, W.p, (rocblas_int)n,
                            b.p, 1, &one, e.p, 1), "dgemv(W sol)");
        std::vector<double> he(n);
        copy_host(st(), he.data(), e.p, n * 8, hipMemcpyDeviceToHost);
        if (sol) {
            sol->resize(dim);
            copy_host(st(), sol->data(), b.p, dim * 8, hipMemcpyDeviceToHost);
        }
        CB_HIP(hipStreamSynchronize(st()));
        double ye = 0;
        for (int64_t i = 0; i < n; ++i) ye += y[i] * he[i];
        *logl = -0.5 * (ldC + ((double)n - (double)p - (double)q) * std::log(ve) + (double)q * std::log(vu) + ye / ve +
                        logdetA);
        return true;
    }
    // gradient and AI terms at (vu, ve) (vce.cpp:3048-3120); solve_logl at the
    // same point must have run just before (C holds its factor, e its residuals)
    void terms(double ve, double grad[2], double ai[4]) {
        CB_HIP(hipMemcpyAsync(B.p, Wt.p, B.bytes(), hipMemcpyDeviceToDevice, st()));
        CB_BLAS(*bh, rocsolver_dpotrs(h, rocblas_fill_lower, (rocblas_int)dim, (rocblas_int)n, C.p, (rocblas_int)dim, B.p,
                               (rocblas_int)dim), "dpotrs(C, W')");
        const double one = 1.0, zero = 0.0;
        CB_BLAS(*bh, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                            (rocblas_int)dim, &one, W.p, (rocblas_int)n, B.p, (rocblas_int)dim, &zero, H.p,
                            (rocblas_int)n), "dgemm(W C^-1 W')");
        k_p_from_h<<<dim3(blocks(n), (unsigned)n), 256, 0, st()>>>(H.p, n, ve);  // H now holds P
        CB_HIP(hipGetLastError());
        CB_BLAS(*bh, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                            (rocblas_int)n, &one, H.p, (rocblas_int)n, ZAZ.p, (rocblas_int)n, &zero, Mu.p,
                            (rocblas_int)n), "dgemm(P ZAZ')");
        k_exact_traces<<<DOT_CHUNKS, 256, 0, st()>>>(Mu.p, H.p, n, part.p);
        CB_HIP(hipGetLastError());
        // Py = e / ve ; t = ZAZ' Py
        std::vector<double> he(n), py(n), ht(n), hp((size_t)DOT_CHUNKS * 5);
        copy_host(st(), he.data(), e.p, n * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st()));
        for (int64_t i = 0; i < n; ++i) py[i] = he[i] / ve;
        copy_host(st(), e.p, py.data(), n * 8, hipMemcpyHostToDevice);
        CB_BLAS(*bh, rocblas_dgemv(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n, &one, ZAZ.p, (rocblas_int)n,
                            e.p, 1, &zero, t.p, 1), "dgemv(ZAZ' Py)");
        copy_host(st(), ht.data(), t.p, n * 8, hipMemcpyDeviceToHost);
        copy_host(st(), hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st()));
        double s[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < DOT_CHUNKS; ++k)
            for (int j = 0; j < 5; ++j) s[j] += hp[(size_t)k * 5 + j];
        double qu = 0, qe = 0;
        for (int64_t i = 0; i < n; ++i) {
            qu += py[i] * ht[i];
            qe += py[i] * py[i];
        }
        ai[0] = 0.5 * s[0];
        ai[1] = ai[2] = 0.5 * s[1];
        ai[3] = 0.5 * s[2];
        grad[0] = -0.5 * s[3] + 0.5 * qu;
        grad[1] = -0.5 * s[4] + 0.5 * qe;
    }
};


// calc_logL_internal (vce.cpp:3533-3588) with C(ve/vu) assembled densely on
// the device and factorised by Cholesky; buffers kept in lw between calls.
// *logl = -1e20 (the reference's failed-determinant value) when C is not
// positive definite; returns log|C| (NaN then).
struct LoglWork {
    DevBuf<double> C, b, part;
    DevBuf<int> info;
};

static double reml_logl(cb_mme* M, LoglWork& lw, double var_u, double var_e, double* logl) {
    hipStream_t st = M->ctx->stream;
    const int64_t dim = M->dim, p = M->p, q = M->q;
    const double lambda = var_e / std::max(var_u, 1e-9);
    if (lw.C.n < (size_t)dim * dim) lw.C.alloc((size_t)dim * dim);
    if (lw.b.n < (size_t)dim) lw.b.alloc(dim);
    if (!lw.info.p) lw.info.alloc(1);
    if (!lw.part.p) lw.part.alloc(DOT_CHUNKS);
    assemble_dense_C(M, lambda, lw.C.p);
    copy_host(st, lw.b.p, M->rhs.data(), dim * 8, hipMemcpyHostToDevice);
    BlasHandle bh(st);
    const rocblas_handle h = bh.h;
    CB_BLAS(bh, rocsolver_dpotrf(h, rocblas_fill_lower, (rocblas_int)dim, lw.C.p, (rocblas_int)dim, lw.info.p),
            "REML logL: dpotrf");
    int hinfo = 0;
    copy_host(st, &hinfo, lw.info.p, sizeof(int), hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    if (hinfo != 0) {
        *logl = -1e20;
        return NAN;
    }
    CB_BLAS(bh, rocsolver_dpotrs(h, rocblas_fill_lower, (rocblas_int)dim, 1, lw.C.p, (rocblas_int)dim, lw.b.p,
                                 (rocblas_int)dim),
            "REML logL: dpotrs");
    k_logdiag_part<<<DOT_CHUNKS, 256, 0, st>>>(lw.C.p, dim, lw.part.p);
    std::vector<double> hp(DOT_CHUNKS), sol(dim);
    copy_host(st, hp.data(), lw.part.p, DOT_CHUNKS * 8, hipMemcpyDeviceToHost);
    copy_host(st, sol.data(), lw.b.p, dim * 8, hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    double ld = 0;
    for (double v : hp) ld += v;
    ld *= 2.0;  // log|C| = 2 sum log L_ii
    double sr = 0;
    for (int64_t i = 0; i < dim; ++i) sr += sol[i] * M->rhs[i];
    const double yPy = (M->yty - sr) / var_e;
    const double n = (double)M->n;
    *logl = -0.5 * (ld + (n - (double)p - (double)q) * std::log(var_e) +
                    (double)q * std::log(std::max(var_u, 1e-9)) + yPy);
    return ld;
}

}  // namespace

extern "C" {

int cb_reml_score(cb_mme* M, const cb_reml_options* opts, double var_u, double var_e, double* grad,
                  double* ai, double* trace, double* solution) {
    return guarded([&] {
        CB_REQUIRE(M && opts, "cb_reml_score: NULL argument");
        CB_REQUIRE(M->qkind >= 0, "cb_reml_score: Qinv not set");
        CB_REQUIRE(var_u > 0 && var_e > 0, "cb_reml_score: variances must be positive");
        set_device(M->ctx);
        const RemlPlan pl = reml_plan(M, *opts);
        if (pl.first == CB_VCE_AI_EXACT) {
            DenseExact de(M);
            de.init();
            double lg = 0, g[2], a[4];
            std::vector<double> sol;
            CB_REQUIRE(de.solve_logl(var_u, var_e, &lg, solution ? &sol : nullptr),
                       "cb_reml_score: C is not positive definite");
            de.terms(var_e, g, a);
            if (grad) { grad[0] = g[0]; grad[1] = g[1]; }
            if (ai) std::copy(a, a + 4, ai);
            if (trace) *trace = std::nan("");
            if (solution) std::copy(sol.begin(), sol.end(), solution);
            return;
        }
        cb_reml_work w;
        reml_setup(M, *opts, w, pl);
        RemlState s;
        reml_eval(M, *opts, w, var_u, var_e, ai != nullptr, s, solution);
        if (grad) { grad[0] = s.grad[0]; grad[1] = s.grad[1]; }
        if (ai) for (int i = 0; i < 4; ++i) ai[i] = s.ai[i];
        if (trace) *trace = s.tr;
    });
}

// AI_REML::solve (vce.cpp:741-837) with the data-driven AI step
// (vce.cpp:2414-2857), the EM + Hutchinson step (vce.cpp:1490-1574), the
// dense-exact step (vce.cpp:2859-3195), or the reference's own choice among
// them per iteration (CB_VCE_AUTO, reml_plan).
int cb_reml_run(cb_mme* M, const cb_reml_options* opts, double var_u0, double var_e0, double* var_u,
                double* var_e, int64_t* iterations, int* converged, double* hist_vu, double* hist_ve,
                double* solution, double* ai_last, double* se) {
    return guarded([&] {
        CB_REQUIRE(M && opts && var_u && var_e, "cb_reml_run: NULL argument");
        CB_REQUIRE(M->qkind >= 0, "cb_reml_run: Qinv not set");
        CB_REQUIRE(var_u0 > 0 && var_e0 > 0, "cb_reml_run: initial variances must be positive");
        CB_REQUIRE(opts->algorithm >= CB_VCE_AI && opts->algorithm <= CB_VCE_AUTO, "cb_reml_run: unknown algorithm");
        set_device(M->ctx);
        const cb_reml_options& o = *opts;
        const RemlPlan pl = reml_plan(M, o);
        if (pl.first == CB_VCE_AI_EXACT) {
            // AI_REML::solve with run_dense_exact_step (vce.cpp:741-837,2949-3195)
            DenseExact de(M);
            de.init();
            double vu = var_u0, ve = var_e0, A[4] = {0, 0, 0, 0};
            int64_t it = 0;
            int conv = 0;
            bool have_ai = false;
            for (; it < o.max_iter;) {
                double cur = 0, g[2], a[4];
                if (!de.solve_logl(vu, ve, &cur, nullptr)) break;  // C not PD: step failed
                de.terms(ve, g, a);
                std::copy(a, a + 4, A);
                project_spd_2x2(A);
                have_ai = true;
                const double det = A[0] * A[3] - A[1] * A[2];
                double d0 = (A[3] * g[0] - A[1] * g[1]) / det;
                double d1 = (A[0] * g[1] - A[2] * g[0]) / det;
                bool ok = false;
                double nu = vu, ne = ve;
                for (int k = 0; k < 6; ++k) {  // vce.cpp:3141-3171
                    nu = vu + d0;
                    ne = ve + d1;
                    if (nu > 0 && ne > 0) {
                        double nl = 0;
                        if (de.solve_logl(nu, ne, &nl, nullptr) && (nl > cur || k == 5)) {
                            ok = true;
                            break;
                        }
                    }
                    d0 *= 0.5;
                    d1 *= 0.5;
                }
                if (!ok) break;  // failed step: stop, unconverged
                const double diff = std::fabs(ne - ve) + std::fabs(nu - vu);
                vu = nu;
                ve = ne;
                if (hist_vu) hist_vu[it] = vu;
                if (hist_ve) hist_ve[it] = ve;
                ++it;
                if (diff < o.tol) {
                    conv = 1;
                    break;
                }
            }
            *var_u = vu;
            *var_e = ve;
            if (iterations) *iterations = it;
            if (converged) *converged = conv;
            if (solution || ai_last || se) {
                double lg = 0, g[2], a[4];
                std::vector<double> sol;
                CB_REQUIRE(de.solve_logl(vu, ve, &lg, &sol), "cb_reml_run: C is not positive definite");
                if (solution) std::copy(sol.begin(), sol.end(), solution);
                if (!have_ai) {
                    de.terms(ve, g, a);
                    std::copy(a, a + 4, A);
                    project_spd_2x2(A);
                }
                if (ai_last) std::copy(A, A + 4, ai_last);
                if (se) {
                    const double det = A[0] * A[3] - A[1] * A[2];
                    se[0] = std::sqrt(std::max(0.0, A[3] / det));
                    se[1] = std::sqrt(std::max(0.0, A[0] / det));
                }
            }
            return;
        }
        cb_reml_work w;
        reml_setup(M, o, w, pl);
        // variance stabiliser of vce.cpp:1370-1389
        const int64_t n = M->n;
        std::vector<double> hy(n);
        copy_host(M->ctx->stream, hy.data(), M->y.p, n * 8, hipMemcpyDeviceToHost);
        double ymean = 0;
        for (double v : hy) ymean += v;
        ymean = n > 0 ? ymean / n : 0.0;
        double yvar = 0;
        for (double v : hy) yvar += (v - ymean) * (v - ymean);
        yvar = n > 1 ? std::max(1e-12, yvar / (double)(n - 1)) : std::max(1e-12, M->yty);
        const double vfloor = 1e-12, vceil = std::max(1.0, yvar * 1e6);
        auto stabilize = [&](double proposed, double previous) {
            if (!std::isfinite(proposed) || proposed <= 0.0) return std::max(previous, vfloor);
            double lo = vfloor, hi = vceil;
            if (std::isfinite(previous) && previous > 0.0) {
                lo = std::max(lo, previous * 0.1);
                hi = std::min(hi, previous * 10.0);
            }
            if (lo > hi) lo = vfloor;
            return std::min(std::max(proposed, lo), hi);
        };
        double vu = var_u0, ve = var_e0;
        int64_t it = 0;
        int conv = 0;
        double last_ai[4] = {0, 0, 0, 0};
        bool have_ai = false;
        LoglWork lw;
        for (; it < o.max_iter;) {
            const double old_u = vu, old_e = ve;
            const int kind = it == 0 ? pl.first : pl.later;
            RemlState s;
            reml_eval(M, o, w, vu, ve, kind == CB_VCE_AI, s, nullptr);
            bool em = kind == CB_VCE_EM;
            if (!em) {
                double A[4] = {s.ai[0], s.ai[1], s.ai[2], s.ai[3]};
                project_spd_2x2(A);
                std::copy(A, A + 4, last_ai);  // last_AI_mat (vce.cpp:3245-3252)
                have_ai = true;
                const double det = A[0] * A[3] - A[1] * A[2];
                double d0 = (A[3] * s.grad[0] - A[1] * s.grad[1]) / det;
                double d1 = (A[0] * s.grad[1] - A[2] * s.grad[0]) / det;
                bool ok = std::isfinite(d0) && std::isfinite(d1) &&
                          std::fabs(d0) <= 100.0 * vu && std::fabs(d1) <= 100.0 * ve;
                if (ok && pl.logl) {
                    // Newton step accepted only where logL improves, <= 15
                    // halvings (vce.cpp:2797-2848); unstabilised, as there
                    double cur = 0;
                    (void)reml_logl(M, lw, vu, ve, &cur);
                    if (!std::isfinite(cur)) cur = -1e18;
                    ok = false;
                    double nu = vu, ne = ve;
                    for (int h = 0; h < 15; ++h) {
                        nu = vu + d0;
                        ne = ve + d1;
                        if (nu > 0.0 && ne > 0.0 && std::fabs(d0) <= 5.0 * vu && std::fabs(d1) <= 5.0 * ve) {
                            double nl = 0;
                            (void)reml_logl(M, lw, nu, ne, &nl);
                            if (std::isfinite(nl) && nl > cur) {
                                ok = true;
                                break;
                            }
                        }
                        d0 *= 0.5;
                        d1 *= 0.5;
                    }
                    if (ok) {
                        vu = nu;
                        ve = ne;
                    }
                } else if (ok) {
                    // no logL: halve until positive and within 5x (vce.cpp:2797-2813)
                    int h = 0;
                    for (; h < 15; ++h) {
                        const double nu = vu + d0, ne = ve + d1;
                        if (nu > 0 && ne > 0 && std::fabs(d0) <= 5.0 * vu && std::fabs(d1) <= 5.0 * ve) break;
                        d0 *= 0.5;
                        d1 *= 0.5;
                    }
                    ok = h < 15;
                    if (ok) {
                        vu = stabilize(vu + d0, old_u);
                        ve = stabilize(ve + d1, old_e);
                    }
                }
                em = !ok;  // the EM step for this iteration (vce.cpp:1441-1454)
            }
            if (em) {
                const double lambda = ve / std::max(vu, 1e-9);
                vu = stabilize((s.uQu + s.tr * ve) / (double)M->q, old_u);
                ve = stabilize((s.esse + ve * ((double)(M->p + M->q) - lambda * s.tr)) / (double)n, old_e);
            }
            if (hist_vu) hist_vu[it] = vu;
            if (hist_ve) hist_ve[it] = ve;
            ++it;
            const double diff = std::fabs(ve - old_e) + std::fabs(vu - old_u);
            if (diff < o.tol) {
                conv = 1;
                break;
            }
        }
        *var_u = vu;
        *var_e = ve;
        if (iterations) *iterations = it;
        if (converged) *converged = conv;
        if (ai_last || se) {
            if (!have_ai) {  // EM run: the AI matrix at the estimates
                RemlState s;
                reml_eval(M, o, w, vu, ve, true, s, nullptr);
                std::copy(s.ai, s.ai + 4, last_ai);
                project_spd_2x2(last_ai);
            }
            if (ai_last) std::copy(last_ai, last_ai + 4, ai_last);
            if (se) {
                // calculate_SE (vce.cpp:3342-3352): Cov = AI^-1, SE = sqrt(max(0, diag))
                const double det = last_ai[0] * last_ai[3] - last_ai[1] * last_ai[2];
                se[0] = std::sqrt(std::max(0.0, last_ai[3] / det));
                se[1] = std::sqrt(std::max(0.0, last_ai[0] / det));
            }
        }
        if (solution) {
            // final solve at the estimates (GSSolve uses lambda = ve/vu)
            const double lambda = ve / std::max(vu, 1e-9);
            CB_HIP(hipMemcpyAsync(w.B.p, w.rhs.p, M->ldv * 8, hipMemcpyDeviceToDevice, M->ctx->stream));
            solve_block(M, lambda, o.solve, w.B.p, w.S.p, 1, w.have_s && !w.adaptive);
            copy_host(M->ctx->stream, solution, w.S.p, M->dim * 8, hipMemcpyDeviceToHost);
        }
    });
}

int cb_reml_branch(const cb_mme* M, const cb_reml_options* opts, int* first_step, int* later_step, int* trace_rule,
                   int* logl_checked, int64_t* min_probes, int64_t* max_probes) {
    return guarded([&] {
        CB_REQUIRE(M && opts, "cb_reml_branch: NULL argument");
        CB_REQUIRE(M->qkind >= 0, "cb_reml_branch: Qinv not set");
        const RemlPlan pl = reml_plan(M, *opts);
        if (first_step) *first_step = pl.first;
        if (later_step) *later_step = pl.later;
        if (trace_rule) *trace_rule = pl.first == CB_VCE_AI_EXACT ? 2 : pl.trace;
        if (logl_checked) *logl_checked = pl.logl || pl.first == CB_VCE_AI_EXACT ? 1 : 0;
        if (min_probes) *min_probes = pl.min_k;
        if (max_probes) *max_probes = pl.max_k;
    });
}

int cb_reml_logl(cb_mme* M, double var_u, double var_e, double* logl, double* logdet_c) {
    return guarded([&] {
        CB_REQUIRE(M && logl, "cb_reml_logl: NULL argument");
        CB_REQUIRE(M->qkind >= 0, "cb_reml_logl: Qinv not set");
        CB_REQUIRE(var_u > 0 && var_e > 0, "cb_reml_logl: variances must be positive");
        CB_REQUIRE(M->dim <= 65536, "cb_reml_logl: dense factorisation limited to p + q <= 65536");
        set_device(M->ctx);
        LoglWork lw;
        const double ld = reml_logl(M, lw, var_u, var_e, logl);
        if (logdet_c) *logdet_c = ld;
    });
}

int cb_reml_copy_probes(cb_mme* M, const cb_reml_options* opts, double* out, int64_t ld) {
    return guarded([&] {
        CB_REQUIRE(M && opts && out && ld >= M->q, "cb_reml_copy_probes: bad argument");
        set_device(M->ctx);
        cb_reml_work w;
        reml_setup(M, *opts, w, reml_plan(M, *opts));
        // probes bw at a time through B's columns (the run's own stream)
        for (int64_t f = 0; f < w.k; f += w.bw) {
            const int cnt = (int)std::min<int64_t>(w.bw, w.k - f);
            make_probes(M, *opts, w, f, cnt, 0, false);
            copy_host_2d(M->ctx->stream, out + f * ld, ld * 8, w.B.p + M->p, M->ldv * 8, M->q * 8, cnt, hipMemcpyDeviceToHost);
            CB_HIP(hipStreamSynchronize(M->ctx->stream));
        }
    });
}

}  // extern "C"

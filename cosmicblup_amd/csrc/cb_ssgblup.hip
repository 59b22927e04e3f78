// ssGBLUP H^-1 blocks — the counterpart of the split-mode set-up of
// src/io/relationship_builder.cpp:481-633: A22 of the genotyped animals,
// G aligned to A22 (tuneG), blended, both inverted, and the dense block
// D = tau G^-1 - omega A22^-1 that SplitMatrixAdapter adds on the genotyped
// rows of A^-1.
//
// A22. The reference solves A^-1 x = e_j by a sparse Cholesky for every
// genotyped animal. Here A e_j comes from the pedigree itself:
// A = (I - P/2)^-1 D (I - P/2)^-T with P the parent incidence and D = diag(d_ii)
// (the same d_ii, with inbreeding, that build A^-1), so a block of unit
// columns takes two sweeps over the animals:
//   up   (offspring before parents): w_i = x_i + 1/2 sum_{offspring k} w_k
//   down (parents before offspring): z_i = d_ii w_i + (z_sire + z_dam) / 2
// Animals are grouped by generation level; one launch per level and sweep,
// thread per (animal, column), columns contiguous per animal. Both sweeps
// are in place on one n x b block; the genotyped rows of z are A22's columns.
//
// The dense steps: the tuneG means (two-stage reductions), the blend
// (elementwise, in the reference's operation order), G^-1 by the eigen-bent
// inverse, A22^-1 by rocSOLVER Cholesky (dpotrf + dpotri), then D.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "cb_blas.h"
#include "cb_internal.h"

using namespace cbamd;

namespace {

constexpr int RED_CHUNKS = 256;

// W[gmap[j0 + c], c] = 1 for the batch's genotyped columns (W zeroed before)
__global__ void k_unit_cols(double* __restrict__ W, int64_t b, const int32_t* __restrict__ gmap, int64_t j0,
                            int64_t nb) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nb) return;
    const int32_t g = gmap[j0 + c];
    if (g >= 0) W[(int64_t)g * b + c] = 1.0;
}

// up sweep over the animals order[lo .. lo + cnt)
__global__ void k_sweep_up(const int64_t* __restrict__ order, int64_t lo, int64_t cnt,
                           const int64_t* __restrict__ off_ptr, const int32_t* __restrict__ off_idx,
                           double* __restrict__ W, int64_t b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t a = t / b, c = t - a * b;
    if (a >= cnt) return;
    const int64_t i = order[lo + a];
    double s = 0.0;
    for (int64_t k = off_ptr[i]; k < off_ptr[i + 1]; ++k) s += W[(int64_t)off_idx[k] * b + c];
    W[i * b + c] += 0.5 * s;
}

// down sweep over the animals order[lo .. lo + cnt)
__global__ void k_sweep_down(const int64_t* __restrict__ order, int64_t lo, int64_t cnt,
                             const int32_t* __restrict__ sire, const int32_t* __restrict__ dam,
                             const double* __restrict__ dii, double* __restrict__ W, int64_t b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t a = t / b, c = t - a * b;
    if (a >= cnt) return;
    const int64_t i = order[lo + a];
    const int32_t s = sire[i], d = dam[i];
    const double zs = s >= 0 ? W[(int64_t)s * b + c] : 0.0;
    const double zd = d >= 0 ? W[(int64_t)d * b + c] : 0.0;
    W[i * b + c] = dii[i] * W[i * b + c] + 0.5 * (zs + zd);
}

// A22[r, j0 + c] (column-major, lda) = Z[gmap[r], c]; rows or columns of
// animals outside the pedigree are those of the identity
__global__ void k_extract_a22(const double* __restrict__ W, int64_t b, const int32_t* __restrict__ gmap, int64_t ng,
                              int64_t j0, double* __restrict__ A22, int64_t lda) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (r >= ng) return;
    const int64_t j = j0 + c;
    const int32_t gr = gmap[r], gj = gmap[j];
    double v;
    if (gj < 0 || gr < 0)
        v = r == j ? 1.0 : 0.0;
    else
        v = W[(int64_t)gr * b + c];
    A22[j * lda + r] = v;
}

// partial sums over column chunks: [0] diag, [1] off-diagonal, of G and A22
__global__ __launch_bounds__(256) void k_means_part(const double* __restrict__ G, int64_t ldg,
                                                    const double* __restrict__ A, int64_t lda, int64_t n,
                                                    double* __restrict__ part) {
    __shared__ double red[4][256];
    const int64_t per = (n + RED_CHUNKS - 1) / RED_CHUNKS;
    const int64_t j0 = (int64_t)blockIdx.x * per, j1 = min(n, j0 + per);
    double gd = 0, go = 0, ad = 0, ao = 0;
    for (int64_t j = j0; j < j1; ++j)
        for (int64_t i = threadIdx.x; i < n; i += 256) {
            const double g = G[j * ldg + i], a = A[j * lda + i];
            if (i == j) {
                gd += g;
                ad += a;
            } else {
                go += g;
                ao += a;
            }
        }
    red[0][threadIdx.x] = gd;
    red[1][threadIdx.x] = go;
    red[2][threadIdx.x] = ad;
    red[3][threadIdx.x] = ao;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// G <- (1 - alpha) (b G + a) + alpha A22, elementwise in that order
__global__ void k_blend(double* __restrict__ G, int64_t ldg, const double* __restrict__ A, int64_t lda, int64_t n,
                        double a, double bb, double alpha) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    double g = bb * G[j * ldg + i];
    g = g + a;
    G[j * ldg + i] = (1.0 - alpha) * g + alpha * A[j * lda + i];
}

__global__ void k_add_diag2(double* A, int64_t n, int64_t ld, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[i * ld + i] += v;
}

// D[i, j] = tau Ginv[i, j] - omega A22inv(i, j), A22inv from the lower
// triangle that dpotri leaves (column-major)
__global__ void k_combine(const double* __restrict__ Gi, int64_t ldg, const double* __restrict__ Ai, int64_t lda,
                          int64_t n, double tau, double omega, double* __restrict__ D, int64_t ldd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    const double a = i >= j ? Ai[j * lda + i] : Ai[i * lda + j];
    D[j * ldd + i] = tau * Gi[j * ldg + i] - omega * a;
}

}  // namespace

extern "C" {

void cb_hinv_options_default(cb_hinv_options* o) {
    if (!o) return;
    o->tune_g = 1;      // config.h:120
    o->alpha = 0.05;    // config.h:121
    o->tau = 1.0;       // config.h:88
    o->omega = 1.0;     // config.h:89
    o->ridge = 0.0;     // config.h:118
}

int cb_hinv_blocks(cb_context* ctx, int64_t n, const int32_t* dam, const int32_t* sire, int64_t ng,
                   const int32_t* gmap, const double* dG, int64_t ldg, const cb_hinv_options* opts, double* dD,
                   int64_t ldd, double* dA22, int64_t lda22, double* stats) {
    return guarded([&] {
        CB_REQUIRE(ctx && dam && sire && gmap && dG && opts && dD, "cb_hinv_blocks: NULL argument");
        CB_REQUIRE(n > 0 && n < INT32_MAX && ng > 0 && ldg >= ng && ldd >= ng, "cb_hinv_blocks: bad dimensions");
        CB_REQUIRE(!dA22 || lda22 >= ng, "cb_hinv_blocks: bad A22 leading dimension");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        // pedigree on the host: sanitised parents, d_ii with inbreeding, levels, offspring lists
        std::vector<int32_t> d(dam, dam + n), s(sire, sire + n);
        for (int64_t i = 0; i < n; ++i) {
            if (!(d[i] >= 0 && d[i] < n)) d[i] = -1;
            if (!(s[i] >= 0 && s[i] < n)) s[i] = -1;
            CB_REQUIRE(d[i] < i && s[i] < i, "cb_hinv_blocks: parents must precede their offspring");
        }
        for (int64_t j = 0; j < ng; ++j) CB_REQUIRE(gmap[j] >= -1 && gmap[j] < n, "cb_hinv_blocks: map out of range");
        std::vector<double> F, dii;
        pedigree_inbreeding(d, s, F, dii);
        std::vector<int32_t> level(n, 0);
        int32_t maxl = 0;
        for (int64_t i = 0; i < n; ++i) {
            int32_t l = 0;
            if (s[i] >= 0) l = std::max(l, level[s[i]] + 1);
            if (d[i] >= 0) l = std::max(l, level[d[i]] + 1);
            level[i] = l;
            maxl = std::max(maxl, l);
        }
        std::vector<int64_t> lstart(maxl + 2, 0), order(n);
        for (int64_t i = 0; i < n; ++i) ++lstart[level[i] + 1];
        for (int32_t l = 0; l <= maxl; ++l) lstart[l + 1] += lstart[l];
        {
            std::vector<int64_t> fill(lstart.begin(), lstart.end() - 1);
            for (int64_t i = 0; i < n; ++i) order[fill[level[i]]++] = i;
        }
        std::vector<int64_t> optr(n + 1, 0);
        for (int64_t i = 0; i < n; ++i) {
            if (s[i] >= 0) ++optr[s[i] + 1];
            if (d[i] >= 0) ++optr[d[i] + 1];
        }
        for (int64_t i = 0; i < n; ++i) optr[i + 1] += optr[i];
        std::vector<int32_t> oidx(std::max<int64_t>(1, optr[n]));
        {
            std::vector<int64_t> fill(optr.begin(), optr.end() - 1);
            for (int64_t i = 0; i < n; ++i) {
                if (s[i] >= 0) oidx[fill[s[i]]++] = (int32_t)i;
                if (d[i] >= 0) oidx[fill[d[i]]++] = (int32_t)i;
            }
        }
        DevBuf<int32_t> ds, dd, doff, dmap;
        DevBuf<int64_t> dorder, doptr;
        DevBuf<double> ddii;
        ds.alloc(n); dd.alloc(n); dorder.alloc(n); doptr.alloc(n + 1); doff.alloc(oidx.size()); ddii.alloc(n);
        dmap.alloc(ng);
        copy_host(st, ds.p, s.data(), n * 4, hipMemcpyHostToDevice);
        copy_host(st, dd.p, d.data(), n * 4, hipMemcpyHostToDevice);
        copy_host(st, dorder.p, order.data(), n * 8, hipMemcpyHostToDevice);
        copy_host(st, doptr.p, optr.data(), (n + 1) * 8, hipMemcpyHostToDevice);
        copy_host(st, doff.p, oidx.data(), oidx.size() * 4, hipMemcpyHostToDevice);
        copy_host(st, ddii.p, dii.data(), n * 8, hipMemcpyHostToDevice);
        copy_host(st, dmap.p, gmap, ng * 4, hipMemcpyHostToDevice);
        // A22 in batches of b unit columns
        DevBuf<double> A22;
        A22.alloc((size_t)ng * ng);
        const int64_t b = std::min<int64_t>(512, round_up(ng, 64));
        DevBuf<double> W;
        W.alloc((size_t)n * b);
        for (int64_t j0 = 0; j0 < ng; j0 += b) {
            const int64_t nb = std::min<int64_t>(b, ng - j0);
            CB_HIP(hipMemsetAsync(W.p, 0, W.bytes(), st));
            k_unit_cols<<<(unsigned)ceil_div(nb, 256), 256, 0, st>>>(W.p, b, dmap.p, j0, nb);
            for (int32_t l = maxl; l >= 0; --l) {
                const int64_t cnt = lstart[l + 1] - lstart[l];
                k_sweep_up<<<(unsigned)ceil_div(cnt * b, 256), 256, 0, st>>>(dorder.p, lstart[l], cnt, doptr.p,
                                                                            doff.p, W.p, b);
            }
            for (int32_t l = 0; l <= maxl; ++l) {
                const int64_t cnt = lstart[l + 1] - lstart[l];
                k_sweep_down<<<(unsigned)ceil_div(cnt * b, 256), 256, 0, st>>>(dorder.p, lstart[l], cnt, ds.p, dd.p,
                                                                              ddii.p, W.p, b);
            }
            k_extract_a22<<<dim3((unsigned)ceil_div(ng, 256), (unsigned)nb), 256, 0, st>>>(W.p, b, dmap.p, ng, j0,
                                                                                          A22.p, ng);
            CB_HIP(hipGetLastError());
        }
        W.release();
        if (dA22) CB_HIP(hipMemcpy2DAsync(dA22, lda22 * 8, A22.p, ng * 8, ng * 8, ng, hipMemcpyDeviceToDevice, st));
        // tuneG (relationship_builder.cpp:539-565) and blend (:567-569) on a copy of G
        const cb_hinv_options& o = *opts;
        DevBuf<double> Gb, part;
        Gb.alloc((size_t)ng * ng);
        CB_HIP(hipMemcpy2DAsync(Gb.p, ng * 8, dG, ldg * 8, ng * 8, ng, hipMemcpyDeviceToDevice, st));
        double a = 0.0, bb = 1.0, mgd = 0, mgo = 0, mad = 0, mao = 0;
        if (o.tune_g) {
            part.alloc(RED_CHUNKS * 4);
            k_means_part<<<RED_CHUNKS, 256, 0, st>>>(Gb.p, ng, A22.p, ng, ng, part.p);
            std::vector<double> hp(RED_CHUNKS * 4);
            copy_host(st, hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost);
            CB_HIP(hipStreamSynchronize(st));
            double sgd = 0, sgo = 0, sad = 0, sao = 0;
            for (int k = 0; k < RED_CHUNKS; ++k) {
                sgd += hp[4 * k];
                sgo += hp[4 * k + 1];
                sad += hp[4 * k + 2];
                sao += hp[4 * k + 3];
            }
            const double noff = (double)ng * (double)(ng - 1);
            mgd = sgd / ng;
            mad = sad / ng;
            mgo = noff > 0 ? sgo / noff : 0.0;
            mao = noff > 0 ? sao / noff : 0.0;
            if (std::fabs(mgd - mgo) > 1e-12) bb = (mad - mao) / (mgd - mgo);
            a = mad - bb * mgd;
        }
        k_blend<<<dim3((unsigned)ceil_div(ng, 256), (unsigned)ng), 256, 0, st>>>(Gb.p, ng, A22.p, ng, ng, a, bb,
                                                                                 o.alpha);
        CB_HIP(hipGetLastError());
        // G^-1 (eigen-bent, with the ridge) in place
        invert_psd(ctx, Gb.p, ng, ng, o.ridge, Gb.p, ng, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice, nullptr,
                   nullptr);
        // A22^-1 by Cholesky (relationship_builder.cpp:574-581)
        if (o.ridge > 0.0) k_add_diag2<<<(unsigned)ceil_div(ng, 256), 256, 0, st>>>(A22.p, ng, ng, o.ridge);
        BlasHandle h(st);
        DevBuf<int> info;
        info.alloc(1);
        int hinfo = 0;
        CB_BLAS(h, rocsolver_dpotrf(h.h, rocblas_fill_lower, (rocblas_int)ng, A22.p, (rocblas_int)ng, info.p),
                "cb_hinv_blocks: dpotrf");
        copy_host(st, &hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        CB_REQUIRE(hinfo == 0, "cb_hinv_blocks: A22 is not positive definite");
        CB_BLAS(h, rocsolver_dpotri(h.h, rocblas_fill_lower, (rocblas_int)ng, A22.p, (rocblas_int)ng, info.p),
                "cb_hinv_blocks: dpotri");
        k_combine<<<dim3((unsigned)ceil_div(ng, 256), (unsigned)ng), 256, 0, st>>>(Gb.p, ng, A22.p, ng, ng, o.tau,
                                                                                  o.omega, dD, ldd);
        CB_HIP(hipGetLastError());
        CB_HIP(hipStreamSynchronize(st));
        if (stats) {
            stats[0] = a;
            stats[1] = bb;
            stats[2] = mad;
            stats[3] = mao;
            stats[4] = mgd;
            stats[5] = mgo;
        }
    });
}

}  // extern "C"

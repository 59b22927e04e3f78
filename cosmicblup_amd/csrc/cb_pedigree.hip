// Pedigree relationship inverse: Henderson's rules with inbreeding —
// the counterpart of genmatrix::invertA_henderson (src/core/a_matrix.cpp:392-439)
// and of its inbreeding sweep ml() (a_matrix.cpp:127-186).
//
// ml() computes F_i = sum_j L_ij^2 d_j - 1 for each animal with both parents
// known, visiting the ancestors j of i in descending index order (its scan
// runs j = max(s, d) .. 0 over a dense L). Here the same visit order comes
// from a max-heap of the touched ancestors, so every F_i is the same
// floating-point sum in the same order (bit-identical) at O(#ancestors log)
// instead of O(n) per animal. Animals are processed level by level
// (level = 1 + max level of the parents), each level's animals in parallel
// with thread-private L and heap: an animal reads only d_j and parents of
// lower levels.
//
// Assembly follows the reference's triplet order (i, i), (i, s), (s, i),
// (s, s), (i, d), (d, i), (d, d), (s, d), (d, s) for i = 0..n-1 and sums
// duplicates in that insertion order (stable sort), then applies the
// symmetrisation, the diagonal floor (< 1e-15 -> 1e-12, missing -> 1e-12)
// and the prune (|v| <= 1e-24) of a_matrix.cpp:58-80,434-438.
#include <algorithm>
#include <cmath>
#include <queue>
#include <thread>
#include <vector>

#include "cb_internal.h"

using namespace cbamd;

struct cb_ainv {
    int64_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> col;
    std::vector<double> val;
    std::vector<double> F, dii;
};

namespace {

// F and d_ii for animals [lo, hi) of `order` (one level), thread-private L.
void ml_range(const std::vector<int32_t>& dam, const std::vector<int32_t>& sire, const std::vector<int64_t>& order,
              int64_t lo, int64_t hi, std::vector<double>& F, std::vector<double>& dii) {
    const int64_t n = (int64_t)dam.size();
    std::vector<double> L(n, 0.0);
    std::priority_queue<int32_t> heap;
    std::vector<int32_t> touched;
    for (int64_t t = lo; t < hi; ++t) {
        const int64_t i = order[t];
        const int32_t s = sire[i], d = dam[i];
        const bool s_ok = s >= 0, d_ok = d >= 0;
        const double Fs = s_ok ? F[s] : 0.0, Fd = d_ok ? F[d] : 0.0;
        if (s_ok && d_ok)
            dii[i] = 1.0 - 0.25 * (1.0 + Fs) - 0.25 * (1.0 + Fd);
        else if (s_ok)
            dii[i] = 1.0 - 0.25 * (1.0 + Fs);
        else if (d_ok)
            dii[i] = 1.0 - 0.25 * (1.0 + Fd);
        else
            dii[i] = 1.0;
        if (!(s_ok && d_ok)) {
            F[i] = 0.0;
            continue;
        }
        auto touch = [&](int32_t j, double add) {
            if (L[j] == 0.0) heap.push(j);
            L[j] += add;
        };
        touch(s, 0.5);
        touch(d, 0.5);
        double Fc = 0.0;
        int32_t last = -1;
        while (!heap.empty()) {
            const int32_t j = heap.top();
            heap.pop();
            if (j == last) continue;  // pushed twice (L went back to 0 is impossible: L > 0)
            last = j;
            const double Lj = L[j];
            Fc += Lj * Lj * dii[j];
            if (sire[j] >= 0) touch(sire[j], 0.5 * Lj);
            if (dam[j] >= 0) touch(dam[j], 0.5 * Lj);
            L[j] = 0.0;
        }
        double f = Fc + dii[i] - 1.0;
        if (f < 0.0) f = 0.0;
        F[i] = f;
    }
}

struct Trip {
    int32_t r, c;
    double v;
};

void assemble(cb_ainv* A, const std::vector<int32_t>& dam, const std::vector<int32_t>& sire) {
    const int64_t n = A->n;
    std::vector<Trip> t;
    t.reserve((size_t)n * 9);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t d = dam[i], s = sire[i];
        const double di = 1.0 / A->dii[i];
        const int32_t ii = (int32_t)i;
        t.push_back({ii, ii, di});
        if (s >= 0) {
            t.push_back({ii, s, -0.5 * di});
            t.push_back({s, ii, -0.5 * di});
            t.push_back({s, s, 0.25 * di});
        }
        if (d >= 0) {
            t.push_back({ii, d, -0.5 * di});
            t.push_back({d, ii, -0.5 * di});
            t.push_back({d, d, 0.25 * di});
        }
        if (s >= 0 && d >= 0) {
            t.push_back({s, d, 0.25 * di});
            t.push_back({d, s, 0.25 * di});
        }
    }
    std::stable_sort(t.begin(), t.end(),
                     [](const Trip& a, const Trip& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
    // sum duplicates in insertion order
    std::vector<Trip> u;
    u.reserve(t.size());
    for (const Trip& x : t) {
        if (!u.empty() && u.back().r == x.r && u.back().c == x.c)
            u.back().v += x.v;
        else
            u.push_back(x);
    }
    t.clear();
    t.shrink_to_fit();
    // symmetrise 0.5 (M + M^T): look the transposed entry up in its row
    std::vector<int64_t> rp(n + 1, 0);
    for (const Trip& x : u) ++rp[x.r + 1];
    for (int64_t r = 0; r < n; ++r) rp[r + 1] += rp[r];
    auto find = [&](int32_t r, int32_t c) -> double {
        auto b = u.begin() + rp[r], e = u.begin() + rp[r + 1];
        auto it = std::lower_bound(b, e, c, [](const Trip& x, int32_t cc) { return x.c < cc; });
        return (it != e && it->c == c) ? it->v : 0.0;
    };
    A->ptr.assign(n + 1, 0);
    A->col.clear();
    A->val.clear();
    A->col.reserve(u.size() + n);
    A->val.reserve(u.size() + n);
    for (int64_t r = 0; r < n; ++r) {
        bool has_diag = false;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
            const Trip& x = u[k];
            double v = 0.5 * (x.v + find(x.c, x.r));
            if (x.c == r) {
                has_diag = true;
                if (v < 1e-15) v = 1e-12;
            }
            if (!has_diag && x.c > r) {  // missing diagonal goes in column order
                A->col.push_back((int32_t)r);
                A->val.push_back(1e-12);
                has_diag = true;
            }
            if (std::fabs(v) <= 1e-24) continue;  // Eigen prune(1e-12): |v| <= 1e-12 * 1e-12
            A->col.push_back(x.c);
            A->val.push_back(v);
        }
        if (!has_diag) {
            A->col.push_back((int32_t)r);
            A->val.push_back(1e-12);
        }
        A->ptr[r + 1] = (int64_t)A->col.size();
    }
}

}  // namespace

namespace cbamd {

void pedigree_inbreeding(const std::vector<int32_t>& dam, const std::vector<int32_t>& sire, std::vector<double>& F,
                std::vector<double>& dii) {
    const int64_t n = (int64_t)dam.size();
    std::vector<int32_t> level(n, 0);
    int32_t maxl = 0;
    for (int64_t i = 0; i < n; ++i) {
        int32_t l = 0;
        if (sire[i] >= 0) l = std::max(l, level[sire[i]] + 1);
        if (dam[i] >= 0) l = std::max(l, level[dam[i]] + 1);
        level[i] = l;
        maxl = std::max(maxl, l);
    }
    // animals grouped by level, index order inside a level
    std::vector<int64_t> start(maxl + 2, 0), order(n);
    for (int64_t i = 0; i < n; ++i) ++start[level[i] + 1];
    for (int32_t l = 0; l <= maxl; ++l) start[l + 1] += start[l];
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; ++i) order[fill[level[i]]++] = i;
    }
    F.assign(n, 0.0);
    dii.assign(n, 1.0);
    const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int32_t l = 0; l <= maxl; ++l) {
        const int64_t lo = start[l], hi = start[l + 1], cnt = hi - lo;
        const int nt = (int)std::min<int64_t>(hw, std::max<int64_t>(1, cnt / 2048));
        if (nt <= 1) {
            ml_range(dam, sire, order, lo, hi, F, dii);
            continue;
        }
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) {
            const int64_t a = lo + cnt * t / nt, b = lo + cnt * (t + 1) / nt;
            th.emplace_back([&, a, b] { ml_range(dam, sire, order, a, b, F, dii); });
        }
        for (auto& x : th) x.join();
    }
}

}  // namespace cbamd

extern "C" {

int cb_ainv_henderson(int64_t n, const int32_t* dam, const int32_t* sire, cb_ainv** out) {
    return guarded([&] {
        CB_REQUIRE(out && n > 0 && dam && sire, "cb_ainv_henderson: bad arguments");
        CB_REQUIRE(n < (int64_t)INT32_MAX, "cb_ainv_henderson: too many animals");
        *out = nullptr;
        std::vector<int32_t> d(dam, dam + n), s(sire, sire + n);
        for (int64_t i = 0; i < n; ++i) {  // complete_missing_parents_add_phantoms (a_matrix.cpp:43-52)
            if (!(d[i] >= 0 && d[i] < n)) d[i] = -1;
            if (!(s[i] >= 0 && s[i] < n)) s[i] = -1;
            CB_REQUIRE(d[i] < i && s[i] < i,
                       "cb_ainv_henderson: parents must precede their offspring (sort the pedigree)");
        }
        auto* A = new cb_ainv;
        A->n = n;
        try {
            pedigree_inbreeding(d, s, A->F, A->dii);
            assemble(A, d, s);
        } catch (...) {
            delete A;
            throw;
        }
        *out = A;
    });
}

int cb_ainv_nnz(const cb_ainv* a, int64_t* nnz) {
    return guarded([&] {
        CB_REQUIRE(a && nnz, "cb_ainv_nnz: NULL argument");
        *nnz = a->ptr[a->n];
    });
}

int cb_ainv_copy_csr(const cb_ainv* a, int64_t* ptr, int32_t* col, double* val) {
    return guarded([&] {
        CB_REQUIRE(a && ptr && col && val, "cb_ainv_copy_csr: NULL argument");
        std::copy(a->ptr.begin(), a->ptr.end(), ptr);
        std::copy(a->col.begin(), a->col.end(), col);
        std::copy(a->val.begin(), a->val.end(), val);
    });
}

int cb_ainv_copy_inbreeding(const cb_ainv* a, double* F, double* dii) {
    return guarded([&] {
        CB_REQUIRE(a, "cb_ainv_copy_inbreeding: NULL handle");
        if (F) std::copy(a->F.begin(), a->F.end(), F);
        if (dii) std::copy(a->dii.begin(), a->dii.end(), dii);
    });
}

void cb_ainv_destroy(cb_ainv* a) { delete a; }

}  // extern "C"

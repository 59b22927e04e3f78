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

namespace {

cb_ainv* ainv_sorted(int64_t n, const std::vector<int32_t>& d, const std::vector<int32_t>& s) {
    auto* A = new cb_ainv;
    A->n = n;
    try {
        pedigree_inbreeding(d, s, A->F, A->dii);
        assemble(A, d, s);
    } catch (...) {
        delete A;
        throw;
    }
    return A;
}

void ainv_renumbered(int64_t n, const std::vector<int32_t>& d, const std::vector<int32_t>& s, cb_ainv** out) {
    // Kahn's algorithm over parent -> offspring edges, smallest index first
    std::vector<int32_t> indeg(n, 0);
    std::vector<int64_t> cptr(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        CB_REQUIRE(d[i] != i && s[i] != i, "cb_ainv_henderson: an animal is its own parent");
        if (d[i] >= 0) ++cptr[d[i] + 1];
        if (s[i] >= 0) ++cptr[s[i] + 1];
        indeg[i] = (d[i] >= 0) + (s[i] >= 0);
    }
    for (int64_t i = 0; i < n; ++i) cptr[i + 1] += cptr[i];
    std::vector<int32_t> kids(cptr[n]);
    {
        std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
        for (int64_t i = 0; i < n; ++i) {
            if (d[i] >= 0) kids[fill[d[i]]++] = (int32_t)i;
            if (s[i] >= 0) kids[fill[s[i]]++] = (int32_t)i;
        }
    }
    std::priority_queue<int32_t, std::vector<int32_t>, std::greater<int32_t>> ready;
    for (int64_t i = 0; i < n; ++i)
        if (indeg[i] == 0) ready.push((int32_t)i);
    std::vector<int32_t> order;  // new -> old
    order.reserve(n);
    std::vector<int32_t> pos(n, -1);  // old -> new
    while (!ready.empty()) {
        const int32_t v = ready.top();
        ready.pop();
        pos[v] = (int32_t)order.size();
        order.push_back(v);
        for (int64_t k = cptr[v]; k < cptr[v + 1]; ++k)
            if (--indeg[kids[k]] == 0) ready.push(kids[k]);
    }
    CB_REQUIRE((int64_t)order.size() == n, "cb_ainv_henderson: the pedigree has a cycle");
    std::vector<int32_t> d2(n), s2(n);
    for (int64_t k = 0; k < n; ++k) {
        const int32_t o = order[k];
        d2[k] = d[o] >= 0 ? pos[d[o]] : -1;
        s2[k] = s[o] >= 0 ? pos[s[o]] : -1;
    }
    cb_ainv* B = ainv_sorted(n, d2, s2);
    auto* A = new cb_ainv;
    A->n = n;
    A->F.resize(n);
    A->dii.resize(n);
    A->ptr.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        A->F[i] = B->F[pos[i]];
        A->dii[i] = B->dii[pos[i]];
        A->ptr[i + 1] = A->ptr[i] + (B->ptr[pos[i] + 1] - B->ptr[pos[i]]);
    }
    A->col.resize(B->col.size());
    A->val.resize(B->val.size());
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t i = 0; i < n; ++i) {
        row.clear();
        for (int64_t k = B->ptr[pos[i]]; k < B->ptr[pos[i] + 1]; ++k) row.emplace_back(order[B->col[k]], B->val[k]);
        std::sort(row.begin(), row.end());
        for (size_t k = 0; k < row.size(); ++k) {
            A->col[A->ptr[i] + k] = row[k].first;
            A->val[A->ptr[i] + k] = row[k].second;
        }
    }
    delete B;
    *out = A;
}

}  // namespace

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
        }
        bool sorted = true;
        for (int64_t i = 0; i < n && sorted; ++i) sorted = d[i] < i && s[i] < i;
        if (!sorted) {
            // ml()'s descending scan (a_matrix.cpp:148-175) is exact only when
            // every parent precedes its offspring; the reference runs it on
            // whatever order its parser produced. Here an unsorted pedigree is
            // renumbered parents-first (Kahn's order, lowest index first, the
            // identity on a sorted pedigree), solved, and mapped back.
            *out = nullptr;
            ainv_renumbered(n, d, s, out);
            return;
        }
        *out = ainv_sorted(n, d, s);
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

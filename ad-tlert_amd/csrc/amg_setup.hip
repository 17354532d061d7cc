// Host-side symbolic setup of the smoothed-aggregation hierarchy (host code only).
#include "amg_setup.hpp"

#include <algorithm>
#include <cmath>

#include "adt_common.hpp"

namespace adt {
namespace amg {
namespace {

// Greedy aggregation over strong connections, rows visited in index order (the device
// numbering follows a space-filling curve, so aggregates are spatially compact):
//   1. a row whose strong neighbours are all free starts an aggregate with them;
//   2. a free row joins the aggregate of its strongest already-aggregated strong neighbour;
//   3. a row still free starts an aggregate with its free strong neighbours.
// Rows without any strong off-diagonal connection (e.g. Dirichlet rows) stay out (-1):
// the smoother treats them exactly.
int aggregate(const HostLevel& L, double theta, std::vector<int>& agg) {
    const int n = L.n;
    std::vector<int> sptr(n + 1, 0), sidx;
    std::vector<double> sval;
    sidx.reserve(L.idx.size());
    sval.reserve(L.idx.size());
    for (int i = 0; i < n; ++i) {
        const double aii = std::fabs(L.rep[L.diag[i]]);
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k) {
            const int j = L.idx[k];
            if (j == i) continue;
            const double ajj = std::fabs(L.rep[L.diag[j]]);
            const double a = std::fabs(L.rep[k]);
            if (a > 0.0 && a >= theta * std::sqrt(aii * ajj)) {
                sidx.push_back(j);
                sval.push_back(a / std::sqrt(aii * ajj));
            }
        }
        sptr[i + 1] = static_cast<int>(sidx.size());
    }
    agg.assign(n, -1);
    int count = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0 || sptr[i] == sptr[i + 1]) continue;
        bool free = true;
        for (int k = sptr[i]; k < sptr[i + 1] && free; ++k) free = agg[sidx[k]] < 0;
        if (!free) continue;
        agg[i] = count;
        for (int k = sptr[i]; k < sptr[i + 1]; ++k) agg[sidx[k]] = count;
        ++count;
    }
    std::vector<int> first = agg;  // pass-1 aggregates only
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0 || sptr[i] == sptr[i + 1]) continue;
        double best = -1.0;
        int pick = -1;
        for (int k = sptr[i]; k < sptr[i + 1]; ++k)
            if (first[sidx[k]] >= 0 && sval[k] > best) {
                best = sval[k];
                pick = first[sidx[k]];
            }
        agg[i] = pick;
    }
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0 || sptr[i] == sptr[i + 1]) continue;
        agg[i] = count;
        for (int k = sptr[i]; k < sptr[i + 1]; ++k)
            if (agg[sidx[k]] < 0) agg[sidx[k]] = count;
        ++count;
    }
    return count;
}

// lambda_max(D^-1 A) by power iteration (D^-1 A is similar to the SPD D^-1/2 A D^-1/2).
double lambda_max(const HostLevel& L) {
    const int n = L.n;
    std::vector<double> v(n), w(n);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < n; ++i) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        v[i] = static_cast<double>((s >> 11) & ((1ull << 52) - 1)) / static_cast<double>(1ull << 52) + 0.5;
    }
    double lam = 1.0;
    for (int it = 0; it < 30; ++it) {
        double nw = 0.0, vw = 0.0, vv = 0.0;
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k) acc += L.rep[k] * v[L.idx[k]];
            w[i] = acc / L.rep[L.diag[i]];
        }
        // Rayleigh quotient in the D inner product
        for (int i = 0; i < n; ++i) {
            const double d = L.rep[L.diag[i]];
            vw += v[i] * d * w[i];
            vv += v[i] * d * v[i];
            nw += w[i] * w[i];
        }
        lam = vv > 0.0 ? vw / vv : 1.0;
        const double scale = nw > 0.0 ? 1.0 / std::sqrt(nw) : 1.0;
        for (int i = 0; i < n; ++i) v[i] = w[i] * scale;
    }
    return lam > 0.0 && std::isfinite(lam) ? lam : 2.0;
}

// Prolongator pattern, its gather map and the transpose.
void prolongator(HostLevel& L) {
    const int n = L.n, nc = L.nc;
    L.pptr.assign(n + 1, 0);
    L.pidx.clear();
    L.pmap_ptr.assign(1, 0);
    L.pmap.clear();
    L.pself.clear();
    std::vector<std::pair<int, int>> cols;  // (aggregate, A slot)
    for (int i = 0; i < n; ++i) {
        cols.clear();
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k) {
            const int J = L.agg[L.idx[k]];
            if (J >= 0) cols.emplace_back(J, k);
        }
        std::stable_sort(cols.begin(), cols.end(),
                         [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.first < b.first; });
        for (size_t c = 0; c < cols.size();) {
            const int J = cols[c].first;
            L.pidx.push_back(J);
            L.pself.push_back(J == L.agg[i] ? 1 : 0);
            for (; c < cols.size() && cols[c].first == J; ++c) L.pmap.push_back(cols[c].second);
            L.pmap_ptr.push_back(static_cast<int>(L.pmap.size()));
        }
        L.pptr[i + 1] = static_cast<int>(L.pidx.size());
    }
    L.rptr.assign(nc + 1, 0);
    for (int J : L.pidx) L.rptr[J + 1]++;
    for (int J = 0; J < nc; ++J) L.rptr[J + 1] += L.rptr[J];
    L.rrow.assign(L.pidx.size(), 0);
    L.rslot.assign(L.pidx.size(), 0);
    std::vector<int> fill(L.rptr.begin(), L.rptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int s = L.pptr[i]; s < L.pptr[i + 1]; ++s) {
            const int J = L.pidx[s];
            L.rrow[fill[J]] = i;
            L.rslot[fill[J]] = s;
            fill[J]++;
        }
}

// T = A P and A_next = R T: patterns and product pairs (fixed order: A slots, then P slots).
void galerkin(HostLevel& L, HostLevel& next) {
    const int n = L.n, nc = L.nc;
    std::vector<int> mark(nc, -1), pos(nc, 0);
    L.tptr.assign(n + 1, 0);
    L.tidx.clear();
    L.tpair_ptr.assign(1, 0);
    L.tpair_a.clear();
    L.tpair_p.clear();
    std::vector<int> row_cols;
    std::vector<std::vector<std::pair<int, int>>> row_pairs;
    for (int i = 0; i < n; ++i) {
        row_cols.clear();
        row_pairs.clear();
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k) {
            const int j = L.idx[k];
            for (int s = L.pptr[j]; s < L.pptr[j + 1]; ++s) {
                const int K = L.pidx[s];
                if (mark[K] != i) {
                    mark[K] = i;
                    pos[K] = static_cast<int>(row_cols.size());
                    row_cols.push_back(K);
                    row_pairs.emplace_back();
                }
                row_pairs[pos[K]].emplace_back(k, s);
            }
        }
        std::vector<int> order(row_cols.size());
        for (size_t c = 0; c < order.size(); ++c) order[c] = static_cast<int>(c);
        std::sort(order.begin(), order.end(), [&](int a, int b) { return row_cols[a] < row_cols[b]; });
        for (int c : order) {
            L.tidx.push_back(row_cols[c]);
            for (const auto& pr : row_pairs[c]) {
                L.tpair_a.push_back(pr.first);
                L.tpair_p.push_back(pr.second);
            }
            L.tpair_ptr.push_back(static_cast<int>(L.tpair_a.size()));
        }
        L.tptr[i + 1] = static_cast<int>(L.tidx.size());
    }
    // next level: row J of R T
    next.n = nc;
    next.ptr.assign(nc + 1, 0);
    next.idx.clear();
    L.cpair_ptr.assign(1, 0);
    L.cpair_p.clear();
    L.cpair_t.clear();
    std::fill(mark.begin(), mark.end(), -1);
    for (int J = 0; J < nc; ++J) {
        row_cols.clear();
        row_pairs.clear();
        for (int e = L.rptr[J]; e < L.rptr[J + 1]; ++e) {
            const int i = L.rrow[e], s = L.rslot[e];
            for (int t = L.tptr[i]; t < L.tptr[i + 1]; ++t) {
                const int K = L.tidx[t];
                if (mark[K] != J) {
                    mark[K] = J;
                    pos[K] = static_cast<int>(row_cols.size());
                    row_cols.push_back(K);
                    row_pairs.emplace_back();
                }
                row_pairs[pos[K]].emplace_back(s, t);
            }
        }
        std::vector<int> order(row_cols.size());
        for (size_t c = 0; c < order.size(); ++c) order[c] = static_cast<int>(c);
        std::sort(order.begin(), order.end(), [&](int a, int b) { return row_cols[a] < row_cols[b]; });
        for (int c : order) {
            next.idx.push_back(row_cols[c]);
            for (const auto& pr : row_pairs[c]) {
                L.cpair_p.push_back(pr.first);
                L.cpair_t.push_back(pr.second);
            }
            L.cpair_ptr.push_back(static_cast<int>(L.cpair_p.size()));
        }
        next.ptr[J + 1] = static_cast<int>(next.idx.size());
    }
    next.diag.assign(nc, -1);
    for (int J = 0; J < nc; ++J)
        for (int k = next.ptr[J]; k < next.ptr[J + 1]; ++k)
            if (next.idx[k] == J) next.diag[J] = k;
    // representative numeric values (same formulas and orders as the device kernels)
    std::vector<double> pv(L.pidx.size()), tv(L.tidx.size());
    for (int i = 0; i < n; ++i) {
        const double dinv = 1.0 / L.rep[L.diag[i]];
        for (int s = L.pptr[i]; s < L.pptr[i + 1]; ++s) {
            double acc = 0.0;
            for (int m = L.pmap_ptr[s]; m < L.pmap_ptr[s + 1]; ++m) acc += L.rep[L.pmap[m]];
            pv[s] = (L.pself[s] ? 1.0 : 0.0) - L.omega * dinv * acc;
        }
    }
    for (size_t t = 0; t < tv.size(); ++t) {
        double acc = 0.0;
        for (int m = L.tpair_ptr[t]; m < L.tpair_ptr[t + 1]; ++m) acc += L.rep[L.tpair_a[m]] * pv[L.tpair_p[m]];
        tv[t] = acc;
    }
    next.rep.assign(next.idx.size(), 0.0);
    for (size_t c = 0; c < next.rep.size(); ++c) {
        double acc = 0.0;
        for (int m = L.cpair_ptr[c]; m < L.cpair_ptr[c + 1]; ++m) acc += pv[L.cpair_p[m]] * tv[L.cpair_t[m]];
        next.rep[c] = acc;
    }
}

}  // namespace

std::vector<HostLevel> build_hierarchy(int n, const int32_t* ptr, const int32_t* idx, const double* values,
                                       const Options& opt) {
    std::vector<HostLevel> levels;
    if (n <= 0 || !ptr || !idx || !values) return levels;
    levels.emplace_back();
    HostLevel& L0 = levels.back();
    L0.n = n;
    L0.ptr.assign(ptr, ptr + n + 1);
    L0.idx.assign(idx, idx + ptr[n]);
    L0.rep.assign(values, values + ptr[n]);
    L0.diag.assign(n, -1);
    for (int i = 0; i < n; ++i)
        for (int k = ptr[i]; k < ptr[i + 1]; ++k)
            if (idx[k] == i) L0.diag[i] = k;
    for (int i = 0; i < n; ++i)
        if (L0.diag[i] < 0 || !(L0.rep[L0.diag[i]] > 0.0)) return {};
    while (static_cast<int>(levels.size()) < opt.max_levels && levels.back().n > opt.coarse_max) {
        HostLevel& L = levels.back();
        std::vector<int> agg;
        const int nc = aggregate(L, opt.theta, agg);
        if (nc == 0 || nc > L.n * 9 / 10) break;  // coarsening stalled: this level is the coarsest
        L.agg = std::move(agg);
        L.nc = nc;
        L.omega = 4.0 / (3.0 * lambda_max(L));
        prolongator(L);
        HostLevel next;
        galerkin(L, next);
        bool ok = true;
        for (int J = 0; J < next.n && ok; ++J) ok = next.diag[J] >= 0 && next.rep[next.diag[J]] > 0.0;
        if (!ok) {  // should not happen for SPD input; keep the current level as the coarsest
            L.nc = 0;
            break;
        }
        levels.push_back(std::move(next));
    }
    HostLevel& last = levels.back();
    last.nc = 0;
    last.agg.clear();
    return levels;
}

}  // namespace amg
}  // namespace adt

// Host-side symbolic setup of the smoothed-aggregation hierarchy (host code only).
#include "amg_setup.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "adt_common.hpp"

namespace adt {
namespace amg {
namespace {

// strong[k] for every slot k of A: off-diagonal with |a_ij| >= theta sqrt(|a_ii a_jj|), a_ij != 0.
std::vector<char> strength(const HostLevel& L, double theta) {
    std::vector<char> strong(L.idx.size(), 0);
    for (int i = 0; i < L.n; ++i) {
        const double aii = std::fabs(L.rep[L.diag[i]]);
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k) {
            const int j = L.idx[k];
            if (j == i) continue;
            const double a = std::fabs(L.rep[k]);
            strong[k] = a > 0.0 && a >= theta * std::sqrt(aii * std::fabs(L.rep[L.diag[j]]));
        }
    }
    return strong;
}

// Greedy aggregation over strong connections, rows visited in index order (the device
// numbering follows a space-filling curve, so aggregates are spatially compact):
//   1. a row whose strong neighbours are all free starts an aggregate with them;
//   2. a free row joins the aggregate of its strongest pass-1 strong neighbour;
//   3. a row still free starts an aggregate with its free strong neighbours.
// Rows without any strong off-diagonal connection (e.g. Dirichlet rows) stay out (-1):
// the smoother treats them exactly.
int aggregate(const HostLevel& L, const std::vector<char>& strong, std::vector<int>& agg) {
    const int n = L.n;
    auto has_strong = [&](int i) {
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k)
            if (strong[k]) return true;
        return false;
    };
    auto weight = [&](int i, int k) {
        return std::fabs(L.rep[k]) / std::sqrt(std::fabs(L.rep[L.diag[i]] * L.rep[L.diag[L.idx[k]]]));
    };
    agg.assign(n, -1);
    int count = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0 || !has_strong(i)) continue;
        bool free = true;
        for (int k = L.ptr[i]; k < L.ptr[i + 1] && free; ++k)
            if (strong[k]) free = agg[L.idx[k]] < 0;
        if (!free) continue;
        agg[i] = count;
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k)
            if (strong[k]) agg[L.idx[k]] = count;
        ++count;
    }
    const std::vector<int> first = agg;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0 || !has_strong(i)) continue;
        double best = -1.0;
        int pick = -1;
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k)
            if (strong[k] && first[L.idx[k]] >= 0 && weight(i, k) > best) {
                best = weight(i, k);
                pick = first[L.idx[k]];
            }
        agg[i] = pick;
    }
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0 || !has_strong(i)) continue;
        agg[i] = count;
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k)
            if (strong[k] && agg[L.idx[k]] < 0) agg[L.idx[k]] = count;
        ++count;
    }
    return count;
}

// Largest eigenvalue of the symmetric tridiagonal matrix (a, b) (b[i] couples i and i + 1) by
// bisection on the Sturm sequence count of eigenvalues below x.
double tridiagonal_max(const std::vector<double>& a, const std::vector<double>& b) {
    const int m = static_cast<int>(a.size());
    double lo = a[0], hi = a[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    const auto below = [&](double x) {  // eigenvalues < x: negative pivots of T - x I
        int count = 0;
        double q = 1.0;
        for (int i = 0; i < m; ++i) {
            q = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / q : 0.0);
            if (q == 0.0) q = -1e-300;
            count += q < 0.0;
        }
        return count;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-12 * std::max(1.0, std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        (below(mid) >= m ? hi : lo) = mid;
    }
    return hi;
}

// lambda_max(D^-1 B) for a matrix B on A's pattern with values bv: D^-1 B is similar to the
// symmetric D^-1/2 B D^-1/2, whose largest eigenvalue is the largest Ritz value of 40 Lanczos
// steps (C. Lanczos, J. Res. Nat. Bur. Standards 45 (1950) 255-282) from a fixed
// pseudo-random start.  Ritz values approach the extremal eigenvalues from inside much faster
// than power iteration: 30 power steps were 8-10% low on ten-node tetrahedra, which put a
// scaled Jacobi damping close to the omega lambda < 2 limit.
double lambda_max(const HostLevel& L, const std::vector<double>& bv, const std::vector<double>& d) {
    const int n = L.n;
    std::vector<double> rs(n), v(n), vp(n, 0.0), w(n), y(n);
    for (int i = 0; i < n; ++i) rs[i] = d[i] > 0.0 ? 1.0 / std::sqrt(d[i]) : 0.0;
    uint64_t s = 0x9E3779B97F4A7C15ull;
    double nv = 0.0;
    for (int i = 0; i < n; ++i) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        v[i] = static_cast<double>((s >> 11) & ((1ull << 52) - 1)) / static_cast<double>(1ull << 52) + 0.5;
        nv += v[i] * v[i];
    }
    nv = std::sqrt(nv);
    for (int i = 0; i < n; ++i) v[i] /= nv;
    std::vector<double> alpha, beta;
    double b = 0.0;
    for (int it = 0; it < 40; ++it) {
        for (int i = 0; i < n; ++i) y[i] = rs[i] * v[i];
        double a = 0.0;
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k) acc += bv[k] * y[L.idx[k]];
            w[i] = rs[i] * acc - b * vp[i];
            a += w[i] * v[i];
        }
        double nw = 0.0;
        for (int i = 0; i < n; ++i) {
            w[i] -= a * v[i];
            nw += w[i] * w[i];
        }
        alpha.push_back(a);
        b = std::sqrt(nw);
        if (!(b > 1e-12 * std::fabs(a)) || it == 39) break;
        beta.push_back(b);
        for (int i = 0; i < n; ++i) {
            vp[i] = v[i];
            v[i] = w[i] / b;
        }
    }
    const double lam = tridiagonal_max(alpha, beta);
    return lam > 0.0 && std::isfinite(lam) ? lam : 2.0;
}

// R = P^T: per coarse row, its (fine row, P slot) entries in fine-row order.
void restriction(HostLevel& L) {
    const int n = L.n, nc = L.nc;
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

// Dampings, the filtered diagonal map, the prolongator pattern with its gather map, and R = P^T.
void prolongator(HostLevel& L, const std::vector<char>& strong, double smoother_scale) {
    const int n = L.n;
    std::vector<double> d(n);
    for (int i = 0; i < n; ++i) d[i] = L.rep[L.diag[i]];
    L.omega = smoother_scale * 4.0 / (3.0 * lambda_max(L, L.rep, d));
    // filtered matrix: strong off-diagonals, weak ones lumped into the diagonal
    L.dmap_ptr.assign(n + 1, 0);
    L.dmap.clear();
    std::vector<double> fv(L.rep.size(), 0.0), fd(n, 0.0);
    for (int i = 0; i < n; ++i) {
        L.dmap.push_back(L.diag[i]);
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k)
            if (k != L.diag[i] && !strong[k]) L.dmap.push_back(k);
        L.dmap_ptr[i + 1] = static_cast<int>(L.dmap.size());
        double di = 0.0;
        for (int m = L.dmap_ptr[i]; m < L.dmap_ptr[i + 1]; ++m) di += L.rep[L.dmap[m]];
        fd[i] = di > 0.0 ? di : d[i];
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k) fv[k] = strong[k] ? L.rep[k] : 0.0;
        fv[L.diag[i]] = fd[i];
    }
    L.omega_p = 4.0 / (3.0 * lambda_max(L, fv, fd));
    L.pptr.assign(n + 1, 0);
    L.pidx.clear();
    L.pmap_ptr.assign(1, 0);
    L.pmap.clear();
    L.pself.clear();
    std::vector<std::pair<int, int>> cols;  // (aggregate, strong A slot or -1 for the row itself)
    for (int i = 0; i < n; ++i) {
        cols.clear();
        if (L.agg[i] >= 0) cols.emplace_back(L.agg[i], -1);
        for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k)
            if (strong[k] && L.agg[L.idx[k]] >= 0) cols.emplace_back(L.agg[L.idx[k]], k);
        std::stable_sort(cols.begin(), cols.end(),
                         [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.first < b.first; });
        for (size_t c = 0; c < cols.size();) {
            const int J = cols[c].first;
            L.pidx.push_back(J);
            L.pself.push_back(J == L.agg[i] ? 1 : 0);
            for (; c < cols.size() && cols[c].first == J; ++c)
                if (cols[c].second >= 0) L.pmap.push_back(cols[c].second);
            L.pmap_ptr.push_back(static_cast<int>(L.pmap.size()));
        }
        L.pptr[i + 1] = static_cast<int>(L.pidx.size());
    }
    restriction(L);
}

// A given prolongator (FixedProlongator) on this level: no aggregation, no filtered diagonal.
void fixed_prolongator(HostLevel& L, const FixedProlongator& fp, double smoother_scale) {
    const int n = L.n;
    std::vector<double> d(n);
    for (int i = 0; i < n; ++i) d[i] = L.rep[L.diag[i]];
    L.omega = smoother_scale * 4.0 / (3.0 * lambda_max(L, L.rep, d));
    L.omega_p = 0.0;
    L.nc = fp.nc;
    L.agg.assign(n, -1);
    L.dmap_ptr.resize(n + 1);
    L.dmap.resize(n);
    for (int i = 0; i <= n; ++i) L.dmap_ptr[i] = i;
    for (int i = 0; i < n; ++i) L.dmap[i] = L.diag[i];
    L.pptr.assign(fp.ptr, fp.ptr + n + 1);
    L.pidx.assign(fp.idx, fp.idx + fp.ptr[n]);
    L.pfix.assign(fp.val, fp.val + fp.ptr[n]);
    L.pself.assign(L.pidx.size(), 0);
    L.pmap_ptr.assign(L.pidx.size() + 1, 0);
    L.pmap.clear();
    restriction(L);
}

// Sparse product pattern with pair lists: for every output row, the (x, y) contributions
// produced by `visit(row, emit)` (emit(col, x, y)) are grouped by output column (ascending),
// keeping the visiting order inside each column.
struct PairBuilder {
    std::vector<int> mark, loc, cols, count, start, order, tx, ty, tl;
    explicit PairBuilder(int ncols) : mark(ncols, -1), loc(ncols, 0) {}
    template <typename Visit>
    void build(int nrows, Visit visit, std::vector<int>& optr, std::vector<int>& oidx, std::vector<int>& pptr,
               std::vector<int>& px, std::vector<int>& py) {
        optr.assign(nrows + 1, 0);
        oidx.clear();
        pptr.assign(1, 0);
        px.clear();
        py.clear();
        for (int r = 0; r < nrows; ++r) {
            cols.clear();
            count.clear();
            tx.clear();
            ty.clear();
            tl.clear();
            visit(r, [&](int c, int x, int y) {
                if (mark[c] != r) {
                    mark[c] = r;
                    loc[c] = static_cast<int>(cols.size());
                    cols.push_back(c);
                    count.push_back(0);
                }
                count[loc[c]]++;
                tx.push_back(x);
                ty.push_back(y);
                tl.push_back(loc[c]);
            });
            order.resize(cols.size());
            std::iota(order.begin(), order.end(), 0);
            std::sort(order.begin(), order.end(), [&](int a, int b) { return cols[a] < cols[b]; });
            // start offset of each local column in sorted order
            start.assign(cols.size(), 0);
            int base = static_cast<int>(px.size());
            for (int o : order) {
                oidx.push_back(cols[o]);
                start[o] = base;
                base += count[o];
                pptr.push_back(base);
            }
            px.resize(base);
            py.resize(base);
            for (size_t e = 0; e < tx.size(); ++e) {
                const int at = start[tl[e]]++;
                px[at] = tx[e];
                py[at] = ty[e];
            }
            optr[r + 1] = static_cast<int>(oidx.size());
        }
        std::fill(mark.begin(), mark.end(), -1);
    }
};

// T = A P and A_next = R T (patterns, pair lists) and the representative next-level values.
void galerkin(HostLevel& L, HostLevel& next) {
    const int n = L.n, nc = L.nc;
    PairBuilder pb(nc);
    pb.build(
        n,
        [&](int i, auto emit) {
            for (int k = L.ptr[i]; k < L.ptr[i + 1]; ++k) {
                const int j = L.idx[k];
                for (int s = L.pptr[j]; s < L.pptr[j + 1]; ++s) emit(L.pidx[s], k, s);
            }
        },
        L.tptr, L.tidx, L.tpair_ptr, L.tpair_a, L.tpair_p);
    next.n = nc;
    pb.build(
        nc,
        [&](int J, auto emit) {
            for (int e = L.rptr[J]; e < L.rptr[J + 1]; ++e) {
                const int i = L.rrow[e], s = L.rslot[e];
                for (int t = L.tptr[i]; t < L.tptr[i + 1]; ++t) emit(L.tidx[t], s, t);
            }
        },
        next.ptr, next.idx, L.cpair_ptr, L.cpair_p, L.cpair_t);
    next.diag.assign(nc, -1);
    for (int J = 0; J < nc; ++J)
        for (int k = next.ptr[J]; k < next.ptr[J + 1]; ++k)
            if (next.idx[k] == J) next.diag[J] = k;
    // representative numeric values (same formulas and orders as the device kernels)
    std::vector<double> pv(L.pidx.size()), tv(L.tidx.size());
    for (int i = 0; i < n && L.pfix.empty(); ++i) {
        double df = 0.0;
        for (int m = L.dmap_ptr[i]; m < L.dmap_ptr[i + 1]; ++m) df += L.rep[L.dmap[m]];
        const double dinv = 1.0 / (df > 0.0 ? df : L.rep[L.diag[i]]);
        for (int s = L.pptr[i]; s < L.pptr[i + 1]; ++s) {
            double acc = 0.0;
            for (int m = L.pmap_ptr[s]; m < L.pmap_ptr[s + 1]; ++m) acc += L.omega_p * L.rep[L.pmap[m]];
            pv[s] = (L.pself[s] ? 1.0 - L.omega_p : 0.0) - dinv * acc;
        }
    }
    if (!L.pfix.empty()) pv = L.pfix;
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
        if (levels.size() == 1 && opt.first) {
            fixed_prolongator(L, *opt.first, opt.smoother_scale);
        } else {
            const std::vector<char> strong = strength(L, opt.theta);
            std::vector<int> agg;
            const int nc = aggregate(L, strong, agg);
            if (nc == 0 || nc > L.n * 9 / 10) break;  // coarsening stalled: this level is the coarsest
            L.agg = std::move(agg);
            L.nc = nc;
            prolongator(L, strong, levels.size() == 1 ? opt.smoother_scale : 1.0);
        }
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

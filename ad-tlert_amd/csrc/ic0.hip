// Batched FP64 conjugate gradients preconditioned by an incomplete Cholesky factor IC(0).
//
// Systems share one CSR pattern (values n_sys x nnz, a stride of nnz apart); each system's
// right-hand sides and iterates are an interleaved (n, rp) block, r of the rp columns used.
//
// Factor (fine-grained fixed point, Chow and Patel 2015): on the lower-triangular pattern of A,
//   l_ij = (a_ij - sum_{k<j} l_ik l_jk) / l_jj  (i > j),   l_ii = sqrt(a_ii - sum_{k<i} l_ik^2),
// every entry recomputed from the previous sweep's values (Jacobi style: all entries in parallel,
// the same result whatever the thread order), starting from l_ij = a_ij / sqrt(a_jj).  A few
// sweeps give the incomplete factor to the accuracy a preconditioner needs.
// Apply z = M^-1 r with M = L L^T: L y = r and L^T z = y by a fixed number of Jacobi sweeps each
//   y <- D^-1 (r - (L - D) y),   z <- D^-1 (y - (L^T - D) z)      (D = diag L),
// the second exactly the transpose of the first, so M^-1 is a fixed symmetric positive definite
// operator and plain CG applies.  Both sweeps run over rows (L^T by its own row pattern, built
// on the host), one thread per (row, column): no atomics.
// CG per column: alpha = r.z / p.q, beta = r'.z' / r.z; stop when ||r|| <= tol ||b||.  Column
// sums are two-level and fixed-order (per-block partials, then a per-column reduction), so
// iterates are bitwise repeatable.  Scalars stay on the device; the host polls the number of
// active columns once per `check_every` iterations.
#include "adt_common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int kThreads = 256;

struct Ic0Pattern {
    const int32_t* aptr;   // A (n + 1)
    const int32_t* aidx;
    const int32_t* lptr;   // lower pattern of L incl. diagonal (n + 1), columns ascending
    const int32_t* lidx;
    const int32_t* lsrc;   // position in A of every lower entry
    const int32_t* uptr;   // rows of L^T: the entries of column i of L (n + 1)
    const int32_t* uidx;   // their rows
    const int32_t* upos;   // their positions in L
    int n, nnz, lnnz;
};

// one fixed-point sweep of the IC(0) entries of every system: lnew from lold
__global__ __launch_bounds__(kThreads) void ic0_sweep(Ic0Pattern P, const double* values, const double* lold,
                                                     double* lnew) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int s = static_cast<int>(blockIdx.y);
    if (e >= P.lnnz) return;
    const auto lptr = adt::gview(P.lptr);
    const auto lidx = adt::gview(P.lidx);
    const double* L = lold + static_cast<int64_t>(s) * P.lnnz;
    // row of entry e: binary search in lptr
    int lo = 0, hi = P.n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lptr[mid] <= e) lo = mid;
        else hi = mid;
    }
    const int i = lo, j = lidx[e];
    // sum_{k < j} l_ik l_jk over the common lower columns (both rows ascending), fixed order
    double sum = 0.0;
    int a = lptr[i], b = lptr[j];
    const int aend = lptr[i + 1] - 1, bend = lptr[j + 1] - 1;  // the diagonals are last
    while (a < aend && b < bend) {
        const int ca = lidx[a], cb = lidx[b];
        if (ca >= j) break;
        if (ca == cb) {
            sum = fma(L[a], L[b], sum);
            ++a;
            ++b;
        } else if (ca < cb) {
            ++a;
        } else {
            ++b;
        }
    }
    const double aij = adt::gview(values)[static_cast<int64_t>(s) * P.nnz + adt::gview(P.lsrc)[e]];
    double out;
    if (i == j) {
        const double d = aij - sum;
        out = d > 0.0 ? sqrt(d) : sqrt(fabs(aij));  // a non-positive pivot keeps the scaled diagonal
    } else {
        out = (aij - sum) / L[lptr[j + 1] - 1];
    }
    adt::gview(lnew)[static_cast<int64_t>(s) * P.lnnz + e] = out;
}

__global__ __launch_bounds__(kThreads) void ic0_start(Ic0Pattern P, const double* values, double* L) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    const int s = static_cast<int>(blockIdx.y);
    if (e >= P.lnnz) return;
    const auto lptr = adt::gview(P.lptr);
    int lo = 0, hi = P.n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lptr[mid] <= e) lo = mid;
        else hi = mid;
    }
    const int i = lo, j = adt::gview(P.lidx)[e];
    const double* v = values + static_cast<int64_t>(s) * P.nnz;
    const double aij = adt::gview(v)[adt::gview(P.lsrc)[e]];
    const double ajj = adt::gview(v)[adt::gview(P.lsrc)[lptr[j + 1] - 1]];
    adt::gview(L)[static_cast<int64_t>(s) * P.lnnz + e] = i == j ? sqrt(fabs(aij)) : aij / sqrt(fabs(ajj));
}

struct Ic0Vecs {
    int n, rp, r, n_sys, nb;  // systems use right-hand side block s % nb
    int64_t block;            // n * rp
    const double* b;
    double* x;
    double* rr;               // residual
    double* z;
    double* p;
    double* q;
    double* y;                // sweep scratch
    double* t;                // sweep scratch
    double* part;             // per (system, row block, column) partials
    double* col;              // per (system, column): [rz, pq, rr, bb, alpha, beta, active, iters]
    int rblocks;              // row blocks per system for the column sums
};

constexpr int kCol = 8;
enum { kRz = 0, kPq, kRr, kBb, kAlpha, kBeta, kActive, kIters };

__device__ __forceinline__ double* colv(const Ic0Vecs& V, int s, int c) {
    return V.col + (static_cast<int64_t>(s) * V.rp + c) * kCol;
}

// Per-(row block, column) partial sums of f(row, column) over kThreads elements of the block's
// rows, deterministic: the block's products go to LDS and thread c sums its column in row order.
template <typename F>
__device__ __forceinline__ void column_partials(const Ic0Vecs& V, int s, F f, int slot) {
    __shared__ double prod[kThreads];
    const int64_t e0 = static_cast<int64_t>(blockIdx.x) * kThreads;
    const int64_t e = e0 + threadIdx.x;
    prod[threadIdx.x] = e < V.block ? f(e) : 0.0;
    __syncthreads();
    // the block covers elements [e0, e0 + 256): column c appears at e0 + k with (e0 + k) % rp == c
    for (int c = threadIdx.x; c < V.rp; c += kThreads) {
        double a = 0.0;
        int k = static_cast<int>((c - e0 % V.rp + V.rp) % V.rp);
        for (; k < kThreads; k += V.rp) a += prod[k];
        adt::gview(V.part)[((static_cast<int64_t>(s) * V.rblocks + blockIdx.x) * V.rp + c) * 4 + slot] = a;
    }
}

__device__ __forceinline__ bool col_active(const Ic0Vecs& V, int s, int64_t e) {
    const int c = static_cast<int>(e % V.rp);
    return c < V.r && adt::gview(colv(V, s, c))[kActive] != 0.0;
}

// z = M^-1 r in 2 (sweeps + 1) passes.  Lower: y0 = D^-1 r; y <- D^-1 (r - (L - D) y).
__global__ __launch_bounds__(kThreads) void ic0_lower(Ic0Pattern P, const double* L, Ic0Vecs V, const double* in,
                                                     double* out, int first) {
    const int s = static_cast<int>(blockIdx.y);
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (e >= V.block || !col_active(V, s, e)) return;
    const int row = static_cast<int>(e / V.rp), c = static_cast<int>(e % V.rp);
    const double* Ls = L + static_cast<int64_t>(s) * P.lnnz;
    const int64_t base = static_cast<int64_t>(s) * V.block;
    const int k0 = adt::gview(P.lptr)[row], k1 = adt::gview(P.lptr)[row + 1] - 1;
    double acc = adt::gview(V.rr)[base + e];
    if (!first)
        for (int k = k0; k < k1; ++k)
            acc -= adt::gview(Ls)[k] * adt::gview(in)[base + static_cast<int64_t>(adt::gview(P.lidx)[k]) * V.rp + c];
    adt::gview(out)[base + e] = acc / adt::gview(Ls)[k1];
}

// Upper (L^T by its rows = the columns of L): z <- D^-1 (y - (L^T - D) z).
__global__ __launch_bounds__(kThreads) void ic0_upper(Ic0Pattern P, const double* L, Ic0Vecs V, const double* y,
                                                     const double* in, double* out, int first) {
    const int s = static_cast<int>(blockIdx.y);
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (e >= V.block || !col_active(V, s, e)) return;
    const int row = static_cast<int>(e / V.rp), c = static_cast<int>(e % V.rp);
    const double* Ls = L + static_cast<int64_t>(s) * P.lnnz;
    const int64_t base = static_cast<int64_t>(s) * V.block;
    const int k0 = adt::gview(P.uptr)[row], k1 = adt::gview(P.uptr)[row + 1];
    double acc = adt::gview(y)[base + e], diag = 1.0;
    for (int k = k0; k < k1; ++k) {
        const int j = adt::gview(P.uidx)[k];
        const double l = adt::gview(Ls)[adt::gview(P.upos)[k]];
        if (j == row) diag = l;
        else if (!first) acc -= l * adt::gview(in)[base + static_cast<int64_t>(j) * V.rp + c];
    }
    adt::gview(out)[base + e] = acc / diag;
}

// q = A p (active columns), partials of p.q
__global__ __launch_bounds__(kThreads) void ic0_spmv(Ic0Pattern P, const double* values, Ic0Vecs V) {
    const int s = static_cast<int>(blockIdx.y);
    const int64_t base = static_cast<int64_t>(s) * V.block;
    const double* A = values + static_cast<int64_t>(s) * P.nnz;
    auto f = [&](int64_t e) {
        if (!col_active(V, s, e)) return 0.0;
        const int row = static_cast<int>(e / V.rp), c = static_cast<int>(e % V.rp);
        double acc = 0.0;
        for (int k = adt::gview(P.aptr)[row]; k < adt::gview(P.aptr)[row + 1]; ++k)
            acc = fma(adt::gview(A)[k], adt::gview(V.p)[base + static_cast<int64_t>(adt::gview(P.aidx)[k]) * V.rp + c],
                      acc);
        adt::gview(V.q)[base + e] = acc;
        return acc * adt::gview(V.p)[base + e];
    };
    column_partials(V, s, f, 0);
}

// per-column reductions of partial slot `slot` into column scalar `field`, then `step`
__global__ void ic0_column_reduce(Ic0Vecs V, int slot, int step, double tol2) {
    const int s = static_cast<int>(blockIdx.x);
    for (int c = threadIdx.x; c < V.r; c += blockDim.x) {
        double* cv = colv(V, s, c);
        auto g = adt::gview(cv);
        double a = 0.0;
        for (int blk = 0; blk < V.rblocks; ++blk)
            a += adt::gview(V.part)[((static_cast<int64_t>(s) * V.rblocks + blk) * V.rp + c) * 4 + slot];
        if (step == 0) {  // start of a (re)start: r.z of the first preconditioned residual
            g[kRz] = a;
            continue;
        }
        if (g[kActive] == 0.0) continue;
        if (step == 1) {  // alpha = rz / pq
            g[kPq] = a;
            if (!(a > 0.0) || !std::isfinite(a)) {
                g[kActive] = 0.0;
                g[kIters] = -g[kIters] - 1.0;  // breakdown marker
            } else {
                g[kAlpha] = g[kRz] / a;
            }
        } else if (step == 2) {  // rr after the update: convergence
            g[kRr] = a;
            g[kIters] = g[kIters] + 1.0;
            if (a <= tol2 * g[kBb]) g[kActive] = 0.0;
        } else {  // step 3: beta = rz' / rz
            g[kBeta] = a / g[kRz];
            g[kRz] = a;
        }
    }
}

// x += alpha p, r -= alpha q (active columns), partials of r.r
__global__ __launch_bounds__(kThreads) void ic0_update(Ic0Vecs V) {
    const int s = static_cast<int>(blockIdx.y);
    const int64_t base = static_cast<int64_t>(s) * V.block;
    auto f = [&](int64_t e) {
        if (!col_active(V, s, e)) return 0.0;
        const double alpha = adt::gview(colv(V, s, static_cast<int>(e % V.rp)))[kAlpha];
        adt::gview(V.x)[base + e] = fma(alpha, adt::gview(V.p)[base + e], adt::gview(V.x)[base + e]);
        const double rn = fma(-alpha, adt::gview(V.q)[base + e], adt::gview(V.rr)[base + e]);
        adt::gview(V.rr)[base + e] = rn;
        return rn * rn;
    };
    column_partials(V, s, f, 1);
}

// partials of r.z (after the preconditioner)
__global__ __launch_bounds__(kThreads) void ic0_rz(Ic0Vecs V) {
    const int s = static_cast<int>(blockIdx.y);
    const int64_t base = static_cast<int64_t>(s) * V.block;
    auto f = [&](int64_t e) {
        if (!col_active(V, s, e)) return 0.0;
        return adt::gview(V.rr)[base + e] * adt::gview(V.z)[base + e];
    };
    column_partials(V, s, f, 2);
}

// p = z + beta p (first: p = z)
__global__ __launch_bounds__(kThreads) void ic0_direction(Ic0Vecs V, int first) {
    const int s = static_cast<int>(blockIdx.y);
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (e >= V.block || !col_active(V, s, e)) return;
    const int64_t i = static_cast<int64_t>(s) * V.block + e;
    if (first) {  // p = z: the workspace's earlier contents (possibly not finite) are never read
        adt::gview(V.p)[i] = adt::gview(V.z)[i];
        return;
    }
    const double beta = adt::gview(colv(V, s, static_cast<int>(e % V.rp)))[kBeta];
    adt::gview(V.p)[i] = fma(beta, adt::gview(V.p)[i], adt::gview(V.z)[i]);
}

// r = b - A x (all used columns), partials of b.b (slot 1 is reused for r.r at the start: both
// enter through step 0); columns >= r stay zero
__global__ __launch_bounds__(kThreads) void ic0_residual(Ic0Pattern P, const double* values, Ic0Vecs V) {
    const int s = static_cast<int>(blockIdx.y);
    const int64_t base = static_cast<int64_t>(s) * V.block;
    const int64_t bbase = static_cast<int64_t>(s % V.nb) * V.block;
    const double* A = values + static_cast<int64_t>(s) * P.nnz;
    auto f = [&](int64_t e) {
        const int row = static_cast<int>(e / V.rp), c = static_cast<int>(e % V.rp);
        if (c >= V.r) {
            adt::gview(V.rr)[base + e] = 0.0;
            return 0.0;
        }
        double acc = adt::gview(V.b)[bbase + e];
        for (int k = adt::gview(P.aptr)[row]; k < adt::gview(P.aptr)[row + 1]; ++k)
            acc -= adt::gview(A)[k] * adt::gview(V.x)[base + static_cast<int64_t>(adt::gview(P.aidx)[k]) * V.rp + c];
        adt::gview(V.rr)[base + e] = acc;
        return acc * acc;
    };
    column_partials(V, s, f, 3);
}

__global__ void ic0_bb(Ic0Vecs V) {  // partials of b.b into slot 1
    const int s = static_cast<int>(blockIdx.y);
    const int64_t bbase = static_cast<int64_t>(s % V.nb) * V.block;
    auto f = [&](int64_t e) {
        const int c = static_cast<int>(e % V.rp);
        if (c >= V.r) return 0.0;
        const double v = adt::gview(V.b)[bbase + e];
        return v * v;
    };
    column_partials(V, s, f, 1);
}

// start: bb (slot 1), rr (slot 3) -> kBb, kRr, active
__global__ void ic0_start_cols(Ic0Vecs V, double tol2, int first) {
    const int s = static_cast<int>(blockIdx.x);
    for (int c = threadIdx.x; c < V.rp; c += blockDim.x) {
        auto g = adt::gview(colv(V, s, c));
        double bb = 0.0, rr = 0.0;
        for (int blk = 0; blk < V.rblocks; ++blk) {
            const int64_t o = ((static_cast<int64_t>(s) * V.rblocks + blk) * V.rp + c) * 4;
            bb += adt::gview(V.part)[o + 1];
            rr += adt::gview(V.part)[o + 3];
        }
        const double iters = g[kIters];
        for (int k = 0; k < kCol; ++k) g[k] = 0.0;
        g[kIters] = first ? 0.0 : iters;
        g[kBb] = bb;
        g[kRr] = rr;
        // a column that broke down (negative iteration marker) stays stopped
        g[kActive] = (c < V.r && bb > 0.0 && rr > tol2 * bb && (first || iters >= 0.0)) ? 1.0 : 0.0;
    }
}

__global__ void ic0_count_active(Ic0Vecs V, double* counter) {
    __shared__ double s[256];
    double a = 0.0;
    for (int64_t i = threadIdx.x; i < static_cast<int64_t>(V.n_sys) * V.rp; i += 256)
        a += adt::gview(V.col)[i * kCol + kActive];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *counter = s[0];
}

}  // namespace

extern "C" int64_t adt_ic0_work_doubles(int n_sys, int n, int nnz, int lnnz, int rp) {
    const int64_t block = static_cast<int64_t>(n) * rp;
    const int64_t rblocks = adt::blocks_for(block, kThreads);
    return 2 * static_cast<int64_t>(n_sys) * lnnz + 6 * static_cast<int64_t>(n_sys) * block +
           static_cast<int64_t>(n_sys) * rblocks * rp * 4 + static_cast<int64_t>(n_sys) * rp * kCol + 8 + 0 * nnz;
}

extern "C" int adt_ic0_pcg(int n_sys, int n, int nnz, int r, int rp, const int32_t* d_indptr,
                           const int32_t* d_indices, const int32_t* d_lptr, const int32_t* d_lidx,
                           const int32_t* d_lsrc, int lnnz, const int32_t* d_uptr, const int32_t* d_uidx,
                           const int32_t* d_upos, const double* d_values, const double* d_b, int n_b_blocks,
                           double* d_x, double tol, int max_iter, int factor_sweeps, int solve_sweeps,
                           int check_every, double* d_work, int32_t* h_iters, double* h_relres, void* stream) {
    ADT_REQUIRE(n_sys >= 1 && n_sys <= 65535 && n >= 1 && nnz >= n && lnnz >= n && r >= 1 && r <= rp,
                "adt_ic0_pcg: invalid sizes");
    ADT_REQUIRE(n_b_blocks >= 1 && n_b_blocks <= n_sys && n_sys % n_b_blocks == 0,
                "adt_ic0_pcg: systems must cycle through the right-hand-side blocks");
    ADT_REQUIRE(tol > 0.0 && max_iter >= 0 && factor_sweeps >= 1 && solve_sweeps >= 0 && check_every >= 1,
                "adt_ic0_pcg: invalid iteration settings");
    ADT_REQUIRE(d_indptr && d_indices && d_lptr && d_lidx && d_lsrc && d_uptr && d_uidx && d_upos && d_values &&
                    d_b && d_x && d_work,
                "adt_ic0_pcg: null pointer");
    hipStream_t st = adt::as_stream(stream);
    Ic0Pattern P{d_indptr, d_indices, d_lptr, d_lidx, d_lsrc, d_uptr, d_uidx, d_upos, n, nnz, lnnz};
    Ic0Vecs V{};
    V.n = n;
    V.rp = rp;
    V.r = r;
    V.n_sys = n_sys;
    V.nb = n_b_blocks;
    V.block = static_cast<int64_t>(n) * rp;
    V.rblocks = adt::blocks_for(V.block, kThreads);
    double* L0 = d_work;
    double* L1 = L0 + static_cast<int64_t>(n_sys) * lnnz;
    double* vec = L1 + static_cast<int64_t>(n_sys) * lnnz;
    const int64_t all = static_cast<int64_t>(n_sys) * V.block;
    V.b = d_b;
    V.x = d_x;
    V.rr = vec;
    V.z = vec + all;
    V.p = vec + 2 * all;
    V.q = vec + 3 * all;
    V.y = vec + 4 * all;
    V.t = vec + 5 * all;
    V.part = vec + 6 * all;
    V.col = V.part + static_cast<int64_t>(n_sys) * V.rblocks * rp * 4;
    double* counter = V.col + static_cast<int64_t>(n_sys) * rp * kCol;
    // factor: start, then factor_sweeps Jacobi-style sweeps (ping-pong)
    const dim3 lgrid(adt::blocks_for(lnnz, kThreads), n_sys);
    hipLaunchKernelGGL(ic0_start, lgrid, dim3(kThreads), 0, st, P, d_values, L0);
    double* L = L0;
    double* Lnext = L1;
    for (int k = 0; k < factor_sweeps; ++k) {
        hipLaunchKernelGGL(ic0_sweep, lgrid, dim3(kThreads), 0, st, P, d_values, L, Lnext);
        std::swap(L, Lnext);
    }
    const dim3 vgrid(V.rblocks, n_sys);
    const double tol2 = tol * tol;
    // z = M^-1 r: lower sweeps y <- D^-1 (r - (L - D) y) from y = D^-1 r, then upper sweeps
    // z <- D^-1 (y - (L^T - D) z) from z = D^-1 y; buffers alternate so that z ends in V.z
    auto precondition = [&]() {
        double* cur = V.y;
        double* spare = V.t;
        hipLaunchKernelGGL(ic0_lower, vgrid, dim3(kThreads), 0, st, P, L, V, V.rr, cur, 1);
        for (int k = 0; k < solve_sweeps; ++k) {
            hipLaunchKernelGGL(ic0_lower, vgrid, dim3(kThreads), 0, st, P, L, V, cur, spare, 0);
            std::swap(cur, spare);
        }
        const double* y = cur;  // spare is free now
        double* zcur = (solve_sweeps % 2 == 0) ? V.z : spare;
        double* znext = (zcur == V.z) ? spare : V.z;
        hipLaunchKernelGGL(ic0_upper, vgrid, dim3(kThreads), 0, st, P, L, V, y, zcur, zcur, 1);
        for (int k = 0; k < solve_sweeps; ++k) {
            hipLaunchKernelGGL(ic0_upper, vgrid, dim3(kThreads), 0, st, P, L, V, y, zcur, znext, 0);
            std::swap(zcur, znext);
        }
    };
    auto restart = [&](int first) {  // r = b - A x; bb, rr, active; z, r.z, p = z
        hipLaunchKernelGGL(ic0_bb, vgrid, dim3(kThreads), 0, st, V);
        hipLaunchKernelGGL(ic0_residual, vgrid, dim3(kThreads), 0, st, P, d_values, V);
        hipLaunchKernelGGL(ic0_start_cols, dim3(n_sys), dim3(128), 0, st, V, tol2, first);
        precondition();
        hipLaunchKernelGGL(ic0_rz, vgrid, dim3(kThreads), 0, st, V);
        hipLaunchKernelGGL(ic0_column_reduce, dim3(n_sys), dim3(128), 0, st, V, 2, 0, tol2);
        hipLaunchKernelGGL(ic0_direction, vgrid, dim3(kThreads), 0, st, V, 1);
    };
    double active = 0.0;
    auto poll = [&]() -> int {
        hipLaunchKernelGGL(ic0_count_active, dim3(1), dim3(256), 0, st, V, counter);
        ADT_HIP_CHECK(hipMemcpyAsync(&active, counter, sizeof(double), hipMemcpyDeviceToHost, st));
        ADT_HIP_CHECK(hipStreamSynchronize(st));
        return ADT_OK;
    };
    int rc, it = 0;
    restart(1);
    // CG, then the true residual; columns whose true residual misses tol continue from their
    // iterate (at most two restarts)
    for (int round = 0; round < 3; ++round) {
        if ((rc = poll())) return rc;
        while (active > 0.0 && it < max_iter) {
            const int chunk = std::min(check_every, max_iter - it);
            for (int k = 0; k < chunk; ++k) {
                hipLaunchKernelGGL(ic0_spmv, vgrid, dim3(kThreads), 0, st, P, d_values, V);
                hipLaunchKernelGGL(ic0_column_reduce, dim3(n_sys), dim3(128), 0, st, V, 0, 1, tol2);
                hipLaunchKernelGGL(ic0_update, vgrid, dim3(kThreads), 0, st, V);
                hipLaunchKernelGGL(ic0_column_reduce, dim3(n_sys), dim3(128), 0, st, V, 1, 2, tol2);
                precondition();
                hipLaunchKernelGGL(ic0_rz, vgrid, dim3(kThreads), 0, st, V);
                hipLaunchKernelGGL(ic0_column_reduce, dim3(n_sys), dim3(128), 0, st, V, 2, 3, tol2);
                hipLaunchKernelGGL(ic0_direction, vgrid, dim3(kThreads), 0, st, V, 0);
            }
            ADT_HIP_CHECK(hipGetLastError());
            it += chunk;
            if ((rc = poll())) return rc;
        }
        if (active > 0.0) break;  // iteration limit
        restart(0);               // true residuals; the columns above tol continue
        if ((rc = poll())) return rc;
        if (active == 0.0) break;
    }
    std::vector<double> col(static_cast<size_t>(n_sys) * rp * kCol);
    ADT_HIP_CHECK(hipMemcpyAsync(col.data(), V.col, sizeof(double) * col.size(), hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipStreamSynchronize(st));
    int worst = ADT_OK;
    for (int sidx = 0; sidx < n_sys; ++sidx)
        for (int c = 0; c < r; ++c) {
            const double* g = col.data() + (static_cast<size_t>(sidx) * rp + c) * kCol;
            const double iters = g[kIters];
            const double rel = g[kBb] > 0.0 ? std::sqrt(g[kRr] / g[kBb]) : 0.0;
            if (h_iters) h_iters[static_cast<int64_t>(sidx) * r + c] = static_cast<int32_t>(iters < 0 ? -iters - 1 : iters);
            if (h_relres) h_relres[static_cast<int64_t>(sidx) * r + c] = rel;
            if (iters < 0 || !std::isfinite(rel))
                worst = ADT_ERR_NONFINITE;
            else if (g[kActive] != 0.0 && worst == ADT_OK)
                worst = ADT_ERR_NOT_CONVERGED;
        }
    if (worst == ADT_ERR_NONFINITE) return adt::set_error(worst, "IC(0)-PCG broke down (p.Ap <= 0 or non-finite)");
    if (worst == ADT_ERR_NOT_CONVERGED)
        return adt::set_error(worst, "IC(0)-PCG reached max_iter=%d before tol=%g", max_iter, tol);
    return ADT_OK;
}

// Device kernels of the smoothed-aggregation multigrid preconditioner (V(1,1) cycle with
// damped Jacobi smoothing, dense inverse on the coarsest level), used by the flexible
// conjugate-gradient solver in pcg.hip.  Included by pcg.hip only.
//
// Per solve, for every system (time step x wavenumber), the numeric setup recomputes on the
// device, in fixed orders, the prolongator values P_l, the products T_l = A_l P_l, the
// Galerkin operators A_{l+1} = P_l^T T_l and the dense inverse of the coarsest operator.
// Per iteration, the cycle applied to the Krylov residual r gives z ~ A^-1 r:
//   down, level l:  r1 = b - A (w D^-1 b)            (pre-smoothing from a zero guess)
//                   b_{l+1} = P^T r1
//   coarsest:       e = A_c^-1 b
//   up, level l:    z2 = w D^-1 b + P e_{l+1}
//                   e_l = z2 + w D^-1 (b - A z2)      (post-smoothing)
// The finest level stages w D^-1 r (down) and z2 (up) for its rows and halo in LDS, like the
// conjugate-gradient product, so each staged pass reads its input once per tile.
#pragma once

#include "krylov.hpp"

namespace adt {
namespace krylov {

// ------------------------------------------------------------------------------------------
// numeric setup
// ------------------------------------------------------------------------------------------

// Thread per row i of system blockIdx.y: inverse diagonal (for the smoother), then the
// prolongator row: a given prolongator's values (pfix), or from the filtered matrix (strong
// entries, weak ones lumped into D_F):
//   P[i, J] = [J == agg(i)] (1 - omega_ii) - D_F,ii^-1 sum_{strong k in row i, agg(col k) = J} omega_k A_k
// with omega_k the level's constant damping or, when given, the learned per-slot damping.
__global__ __launch_bounds__(256) void amg_pvals(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    const LvSys S = adt::gview(a.ls)[blockIdx.y];
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= S.rows) return;
    const auto V = adt::gview(a.val + S.val);
    const int dslot = adt::gview(a.diag)[S.diag + i];
    if (a.dinv_out) adt::gview(a.dinv_out)[S.dinv + i] = 1.0 / V[dslot];
    if (S.nc == 0) return;
    if (a.pfix) {  // a given prolongator: the same values for every system of the pattern
        const auto pptr = adt::gview(a.pptr + S.pptr);
        const auto fix = adt::gview(a.pfix + S.pidx);
        const auto P = adt::gview(a.pval + S.pval);
        for (int t = pptr[i]; t < pptr[i + 1]; ++t) P[t] = fix[t];
        return;
    }
    const auto dmap_ptr = adt::gview(a.dmap_ptr + S.dmap_ptr);
    const auto dmap = adt::gview(a.dmap + S.dmap);
    double df = 0.0;
    for (int m = dmap_ptr[i]; m < dmap_ptr[i + 1]; ++m) df += V[dmap[m]];
    const double dfinv = 1.0 / (df > 0.0 ? df : V[dslot]);
    const auto pptr = adt::gview(a.pptr + S.pptr);
    const auto pmap_ptr = adt::gview(a.pmap_ptr + S.pmap_ptr);
    const auto pmap = adt::gview(a.pmap + S.pmap);
    const auto pself = adt::gview(a.pself + S.pidx);
    const auto om = adt::gview(a.omega_slot ? a.omega_slot + S.val : nullptr);
    const double wself = om ? om[dslot] : S.omega_p;
    const auto P = adt::gview(a.pval + S.pval);
    for (int t = pptr[i]; t < pptr[i + 1]; ++t) {
        double acc = 0.0;
        for (int m = pmap_ptr[t]; m < pmap_ptr[t + 1]; ++m) {
            const int k = pmap[m];
            acc = fma(om ? om[k] : S.omega_p, V[k], acc);
        }
        P[t] = (pself[t] ? 1.0 - wself : 0.0) - dfinv * acc;
    }
}

// Thread per slot of T = A P.
__global__ __launch_bounds__(256) void amg_tvals(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    const LvSys S = adt::gview(a.ls)[blockIdx.y];
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= S.n_t) return;
    const auto tp = adt::gview(a.tpair_ptr + S.tpair_ptr);
    const auto ta = adt::gview(a.tpair_a + S.tpair);
    const auto tq = adt::gview(a.tpair_p + S.tpair);
    const auto V = adt::gview(a.val + S.val);
    const auto P = adt::gview(a.pval + S.pval);
    double acc = 0.0;
    for (int m = tp[t]; m < tp[t + 1]; ++m) acc = fma(V[ta[m]], P[tq[m]], acc);
    adt::gview(a.tval)[S.tval + t] = acc;
}

// Thread per slot of the next level's matrix: A_{l+1}[J, K] = sum_i P[i, J] T[i, K].
__global__ __launch_bounds__(256) void amg_cvals(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    const LvSys S = adt::gview(a.ls)[blockIdx.y];
    const int c = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= S.n_next) return;
    const auto cp = adt::gview(a.cpair_ptr + S.cpair_ptr);
    const auto cs = adt::gview(a.cpair_p + S.cpair);
    const auto ct = adt::gview(a.cpair_t + S.cpair);
    const auto P = adt::gview(a.pval + S.pval);
    const auto T = adt::gview(a.tval + S.tval);
    double acc = 0.0;
    for (int m = cp[c]; m < cp[c + 1]; ++m) acc = fma(P[cs[m]], T[ct[m]], acc);
    adt::gview(a.next_valp)[S.next_val + c] = acc;
}

// One workgroup per system: dense copy of the coarsest operator inverted by Gauss-Jordan
// elimination without pivoting (the operator is SPD).  A non-positive pivot (a singular coarse
// operator) zeroes that row and column of the inverse instead.  Each step stages the pivot
// column and the scaled pivot row in LDS, so the rank-one update reads them from LDS and only
// touches other rows of the matrix.  With `lds_matrix` the whole matrix lives in LDS and is
// written out once (operators of up to kDenseLdsRows rows); otherwise it is updated in place
// in global memory.  Dynamic LDS: 2 rows (+ rows^2 with lds_matrix) doubles.
constexpr int kDenseLdsRows = 128;

template <typename MP>
__device__ __forceinline__ void gauss_jordan(MP M, int n, double* colk, double* rowk) {
    const int tid = static_cast<int>(threadIdx.x);
    for (int k = 0; k < n; ++k) {
        const double piv = M[static_cast<int64_t>(k) * n + k];
        const double inv = piv > 0.0 && std::isfinite(piv) ? 1.0 / piv : 0.0;
        for (int i = tid; i < n; i += 256) colk[i] = M[static_cast<int64_t>(i) * n + k];
        for (int j = tid; j < n; j += 256) rowk[j] = (j == k) ? inv : M[static_cast<int64_t>(k) * n + j] * inv;
        __syncthreads();
        for (int j = tid; j < n; j += 256) M[static_cast<int64_t>(k) * n + j] = rowk[j];
        // 16 x 16 thread layout over (row, column), four columns per thread with their loads
        // issued together
        for (int i = tid / 16; i < n; i += 16) {
            if (i == k) continue;
            const double f = colk[i];
            const int64_t base = static_cast<int64_t>(i) * n;
            for (int j0 = tid % 16; j0 < n; j0 += 64) {
                double m[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) m[u] = M[base + (j0 + 16 * u < n ? j0 + 16 * u : j0)];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = j0 + 16 * u;
                    if (j < n) M[base + j] = fma(-f, rowk[j], (j == k) ? 0.0 : m[u]);
                }
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void amg_dense_inverse(const LevelArgs* __restrict__ la, int lds_matrix) {
    extern __shared__ double sh[];
    const LevelArgs& a = *la;
    const LvSys S = adt::gview(a.ls)[blockIdx.x];
    const int n = S.rows, tid = static_cast<int>(threadIdx.x);
    const auto Mg = adt::gview(a.dense + S.dense);
    const auto ptr = adt::gview(a.ptr + S.ptr);
    const auto idx = adt::gview(a.idx + S.idx);
    const auto V = adt::gview(a.val + S.val);
    const int64_t nn = static_cast<int64_t>(n) * n;
    double* colk = sh;
    double* rowk = sh + n;
    if (lds_matrix) {
        double* M = sh + 2 * n;
        for (int64_t e = tid; e < nn; e += 256) M[e] = 0.0;
        __syncthreads();
        for (int i = tid; i < n; i += 256)
            for (int k = ptr[i]; k < ptr[i + 1]; ++k) M[static_cast<int64_t>(i) * n + idx[k]] = V[k];
        __syncthreads();
        gauss_jordan(M, n, colk, rowk);
        for (int64_t e = tid; e < nn; e += 256) Mg[e] = M[e];
    } else {
        for (int64_t e = tid; e < nn; e += 256) Mg[e] = 0.0;
        __syncthreads();
        for (int i = tid; i < n; i += 256)
            for (int k = ptr[i]; k < ptr[i + 1]; ++k) Mg[static_cast<int64_t>(i) * n + idx[k]] = V[k];
        __syncthreads();
        gauss_jordan(Mg, n, colk, rowk);
    }
}

// ------------------------------------------------------------------------------------------
// cycle kernels, gather versions (any level)
// ------------------------------------------------------------------------------------------
// A thread owns the rows row0 + t / L + i * (256 / L) of its tile and works on kGatherRows of
// them at once (gather_rows), so the index, value and gathered-vector loads of several rows
// are in flight together.

constexpr int kGatherRows = 4;
constexpr int kGatherSlots = 4;

struct LevelTile {
    Tile t;
    LvSys S;
    int col;
};

template <int L>
__device__ __forceinline__ bool level_tile(const LevelArgs& a, const Tile* tiles, int n_tiles, LevelTile& c) {
    c.t = adt::gview(tiles)[adt::xcd_tile(blockIdx.x, n_tiles)];
    if (!adt::gview(a.grp_active)[c.t.grp]) return false;
    c.S = adt::gview(a.ls)[c.t.sys];
    c.col = (c.t.grp % a.groups) * L + static_cast<int>(threadIdx.x % L);
    return true;
}

// Row ranges [k0, k1) of a batch of kGatherRows rows of one thread (empty past row1).  The
// row-pointer loads are unconditional (rows past row1 read row r0's), selected afterwards.
template <typename PtrT>
__device__ __forceinline__ void batch_ranges(PtrT ptr, int r0, int step, int row1, int (&k0)[kGatherRows],
                                             int (&k1)[kGatherRows]) {
#pragma unroll
    for (int i = 0; i < kGatherRows; ++i) {
        const int row = r0 + i * step;
        const int rs = row < row1 ? row : r0;
        const int a0 = ptr[rs], a1 = ptr[rs + 1];
        k0[i] = row < row1 ? a0 : 0;
        k1[i] = row < row1 ? a1 : 0;
    }
}

// r1 = b - A (w D^-1 b)
template <int L>
__global__ __launch_bounds__(256) void amg_smooth_residual(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    constexpr int RPP = kRowsPerPass(L);
    const int rp = a.rp;
    const auto ptr = adt::gview(a.ptr + c.S.ptr);
    const auto idx = adt::gview(a.idx + c.S.idx);
    const auto val = adt::gview(a.val + c.S.val);
    const auto D = adt::gview(a.dinv + c.S.dinv);
    const auto B = adt::gview(a.b + c.S.vec + c.col);
    const auto R1 = adt::gview(a.r1 + c.S.vec + c.col);
    const double w = c.S.omega;
    for (int r0 = c.t.row0 + static_cast<int>(threadIdx.x / L); r0 < c.t.row1; r0 += kGatherRows * RPP) {
        int k0[kGatherRows], k1[kGatherRows];
        double acc[kGatherRows], bo[kGatherRows];
        batch_ranges(ptr, r0, RPP, c.t.row1, k0, k1);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int row = r0 + i * RPP;
            bo[i] = B[static_cast<int64_t>(row < c.t.row1 ? row : r0) * rp];  // unused past row1
            acc[i] = 0.0;
        }
        gather_rows<kGatherRows, kGatherSlots>(
            k0, k1, [&](int k) { return idx[k]; }, [&](int k) { return val[k]; },
            [&](int j) { return w * D[j] * B[static_cast<int64_t>(j) * rp]; }, acc);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int row = r0 + i * RPP;
            if (row < c.t.row1) R1[static_cast<int64_t>(row) * rp] = bo[i] - acc[i];
        }
    }
}

// b_{l+1} = P^T r1 (tiles of the next level)
template <int L>
__global__ __launch_bounds__(256) void amg_restrict(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.next_tiles, a.n_next_tiles, c)) return;
    constexpr int RPP = kRowsPerPass(L);
    const int rp = a.rp;
    const auto rptr = adt::gview(a.rptr + c.S.rptr);
    const auto rrow = adt::gview(a.rrow + c.S.rent);
    const auto rslot = adt::gview(a.rslot + c.S.rent);
    const auto P = adt::gview(a.pval + c.S.pval);
    const auto R1 = adt::gview(a.r1 + c.S.vec + c.col);
    const auto NB = adt::gview(a.next_b + c.S.next_vec + c.col);
    for (int J0 = c.t.row0 + static_cast<int>(threadIdx.x / L); J0 < c.t.row1; J0 += kGatherRows * RPP) {
        int k0[kGatherRows], k1[kGatherRows];
        double acc[kGatherRows] = {};
        batch_ranges(rptr, J0, RPP, c.t.row1, k0, k1);
        gather_rows<kGatherRows, kGatherSlots>(
            k0, k1, [&](int e) { return rrow[e]; }, [&](int e) { return P[rslot[e]]; },
            [&](int i) { return R1[static_cast<int64_t>(i) * rp]; }, acc);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int J = J0 + i * RPP;
            if (J < c.t.row1) NB[static_cast<int64_t>(J) * rp] = acc[i];
        }
    }
}

// b_{l+1} = P^T r1 with the tile's restriction slice (fine row and P value of every entry of
// its coarse rows) staged in LDS by coalesced loads: only the r1 rows are gathered per entry.
template <int L>
__global__ __launch_bounds__(256) void amg_restrict_st(const LevelArgs* __restrict__ la) {
    extern __shared__ double smem[];
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.next_tiles, a.n_next_tiles, c)) return;
    constexpr int RPP = kRowsPerPass(L);
    const int rp = a.rp, tid = static_cast<int>(threadIdx.x);
    const auto rptr = adt::gview(a.rptr + c.S.rptr);
    const auto rrow = adt::gview(a.rrow + c.S.rent);
    const auto rslot = adt::gview(a.rslot + c.S.rent);
    const auto P = adt::gview(a.pval + c.S.pval);
    const auto R1 = adt::gview(a.r1 + c.S.vec + c.col);
    const auto NB = adt::gview(a.next_b + c.S.next_vec + c.col);
    const int rows = c.t.row1 - c.t.row0;
    const int e0 = rptr[c.t.row0], ne = rptr[c.t.row1] - e0;
    double* sval = smem;
    int* srow = reinterpret_cast<int*>(sval + a.r_lds_nnz);
    int* sptr = srow + a.r_lds_nnz;
    for (int e = tid; e < ne; e += 256) {
        srow[e] = rrow[e0 + e];
        sval[e] = P[rslot[e0 + e]];
    }
    for (int i = tid; i <= rows; i += 256) sptr[i] = rptr[c.t.row0 + i] - e0;
    __syncthreads();
    for (int q0 = tid / L; q0 < rows; q0 += kGatherRows * RPP) {
        int k0[kGatherRows], k1[kGatherRows];
        double acc[kGatherRows] = {};
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int q = q0 + i * RPP;
            k0[i] = q < rows ? sptr[q] : 0;
            k1[i] = q < rows ? sptr[q + 1] : 0;
        }
        gather_rows<kGatherRows, kGatherSlots>(
            k0, k1, [&](int e) { return srow[e]; }, [&](int e) { return sval[e]; },
            [&](int i) { return R1[static_cast<int64_t>(i) * rp]; }, acc);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int q = q0 + i * RPP;
            if (q < rows) NB[static_cast<int64_t>(c.t.row0 + q) * rp] = acc[i];
        }
    }
}

// e = A_c^-1 b on the coarsest level: workgroup (lane group, block of 256 / L rows), one
// thread per (row, lane)
template <int L>
__global__ __launch_bounds__(256) void amg_coarse_solve(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    const int gid = blockIdx.x;
    if (!adt::gview(a.grp_active)[gid]) return;
    const LvSys S = adt::gview(a.ls)[gid / a.groups];
    const int col = (gid % a.groups) * L + static_cast<int>(threadIdx.x % L);
    const int n = S.rows, rp = a.rp;
    const int i = static_cast<int>(blockIdx.y) * kRowsPerPass(L) + static_cast<int>(threadIdx.x / L);
    if (i >= n) return;
    const auto Mi = adt::gview(a.dense + S.dense + static_cast<int64_t>(i) * n);
    const auto B = adt::gview(a.b + S.vec + col);
    double a0 = 0.0, a1 = 0.0;
    int j = 0;
    for (; j + 1 < n; j += 2) {
        a0 = fma(Mi[j], B[static_cast<int64_t>(j) * rp], a0);
        a1 = fma(Mi[j + 1], B[static_cast<int64_t>(j + 1) * rp], a1);
    }
    if (j < n) a0 = fma(Mi[j], B[static_cast<int64_t>(j) * rp], a0);
    adt::gview(a.e)[S.vec + col + static_cast<int64_t>(i) * rp] = a0 + a1;
}

// z2 = w D^-1 b + P e_{l+1}.  On the finest level (FINE) z2 goes to the direction buffer that
// the coming product overwrites (p[parity ^ 1] holds the direction of two iterations ago).
template <int L, bool FINE>
__global__ __launch_bounds__(256) void amg_prolong(const PcgArgs* __restrict__ pa, const LevelArgs* __restrict__ la,
                                                   int parity) {
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    constexpr int RPP = kRowsPerPass(L);
    const int rp = a.rp;
    const auto pptr = adt::gview(a.pptr + c.S.pptr);
    const auto pidx = adt::gview(a.pidx + c.S.pidx);
    const auto P = adt::gview(a.pval + c.S.pval);
    const auto D = adt::gview(a.dinv + c.S.dinv);
    const auto B = adt::gview((FINE ? residual(*pa, parity) : a.b) + c.S.vec + c.col);
    const auto NE = adt::gview(a.next_e + c.S.next_vec + c.col);
    const auto Z2 = adt::gview((FINE ? pa->p[parity ^ 1] : a.z2) + c.S.vec + c.col);
    const double w = c.S.omega;
    for (int r0 = c.t.row0 + static_cast<int>(threadIdx.x / L); r0 < c.t.row1; r0 += kGatherRows * RPP) {
        int k0[kGatherRows], k1[kGatherRows];
        double acc[kGatherRows], bo[kGatherRows], dv[kGatherRows];
        batch_ranges(pptr, r0, RPP, c.t.row1, k0, k1);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int rs = r0 + i * RPP < c.t.row1 ? r0 + i * RPP : r0;  // rows past row1: unused
            bo[i] = B[static_cast<int64_t>(rs) * rp];
            dv[i] = D[rs];
            acc[i] = 0.0;
        }
        gather_rows<kGatherRows, kGatherSlots>(
            k0, k1, [&](int s) { return pidx[s]; }, [&](int s) { return P[s]; },
            [&](int J) { return NE[static_cast<int64_t>(J) * rp]; }, acc);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int row = r0 + i * RPP;
            if (row < c.t.row1) Z2[static_cast<int64_t>(row) * rp] = fma(w * dv[i], bo[i], acc[i]);
        }
    }
}

// e = z2 + w D^-1 (b - A z2)
template <int L>
__global__ __launch_bounds__(256) void amg_post_smooth(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    constexpr int RPP = kRowsPerPass(L);
    const int rp = a.rp;
    const auto ptr = adt::gview(a.ptr + c.S.ptr);
    const auto idx = adt::gview(a.idx + c.S.idx);
    const auto val = adt::gview(a.val + c.S.val);
    const auto D = adt::gview(a.dinv + c.S.dinv);
    const auto B = adt::gview(a.b + c.S.vec + c.col);
    const auto Z2 = adt::gview(a.z2 + c.S.vec + c.col);
    const auto E = adt::gview(a.e + c.S.vec + c.col);
    const double w = c.S.omega;
    for (int r0 = c.t.row0 + static_cast<int>(threadIdx.x / L); r0 < c.t.row1; r0 += kGatherRows * RPP) {
        int k0[kGatherRows], k1[kGatherRows];
        double acc[kGatherRows], bo[kGatherRows], zo[kGatherRows], dv[kGatherRows];
        batch_ranges(ptr, r0, RPP, c.t.row1, k0, k1);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int rs = r0 + i * RPP < c.t.row1 ? r0 + i * RPP : r0;  // rows past row1: unused
            const int64_t o = static_cast<int64_t>(rs) * rp;
            bo[i] = B[o];
            zo[i] = Z2[o];
            dv[i] = D[rs];
            acc[i] = 0.0;
        }
        gather_rows<kGatherRows, kGatherSlots>(
            k0, k1, [&](int k) { return idx[k]; }, [&](int k) { return val[k]; },
            [&](int j) { return Z2[static_cast<int64_t>(j) * rp]; }, acc);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int row = r0 + i * RPP;
            if (row < c.t.row1) E[static_cast<int64_t>(row) * rp] = fma(w * dv[i], bo[i] - acc[i], zo[i]);
        }
    }
}

// Finest level, gather version: z = z2 + w D^-1 (r - A z2) plus the flexible-CG partials
// z.r (part1) and z.q_old (part0).
template <int L>
__global__ __launch_bounds__(256) void amg_post_smooth_fine(const PcgArgs* __restrict__ pa,
                                                            const LevelArgs* __restrict__ la, int parity) {
    const PcgArgs& g = *pa;
    const LevelArgs& a = *la;
    const int tix = adt::xcd_tile(blockIdx.x, a.n_tiles);
    const Tile t = adt::gview(a.tiles)[tix];
    if (!adt::gview(a.grp_active)[t.grp]) return;
    constexpr int RPP = kRowsPerPass(L);
    const LvSys S = adt::gview(a.ls)[t.sys];
    const int col = (t.grp % a.groups) * L + static_cast<int>(threadIdx.x % L);
    const int rp = a.rp;
    const auto ptr = adt::gview(a.ptr + S.ptr);
    const auto idx = adt::gview(a.idx + S.idx);
    const auto val = adt::gview(a.val + S.val);
    const auto D = adt::gview(a.dinv + S.dinv);
    const auto R = adt::gview(a.b + S.vec + col);
    const auto Z2 = adt::gview(g.p[parity ^ 1] + S.vec + col);
    const auto Q = adt::gview(g.q + S.vec + col);
    const auto Z = adt::gview(a.e + S.vec + col);
    const double w = S.omega;
    double zr = 0.0, zq = 0.0;
    for (int r0 = t.row0 + static_cast<int>(threadIdx.x / L); r0 < t.row1; r0 += kGatherRows * RPP) {
        int k0[kGatherRows], k1[kGatherRows];
        double acc[kGatherRows], rv[kGatherRows], qv[kGatherRows], zo[kGatherRows], dv[kGatherRows];
        batch_ranges(ptr, r0, RPP, t.row1, k0, k1);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int rs = r0 + i * RPP < t.row1 ? r0 + i * RPP : r0;  // rows past row1: unused
            const int64_t o = static_cast<int64_t>(rs) * rp;
            rv[i] = R[o];
            qv[i] = Q[o];
            zo[i] = Z2[o];
            dv[i] = D[rs];
            acc[i] = 0.0;
        }
        gather_rows<kGatherRows, kGatherSlots>(
            k0, k1, [&](int k) { return idx[k]; }, [&](int k) { return val[k]; },
            [&](int j) { return Z2[static_cast<int64_t>(j) * rp]; }, acc);
#pragma unroll
        for (int i = 0; i < kGatherRows; ++i) {
            const int row = r0 + i * RPP;
            if (row < t.row1) {
                const double zi = fma(w * dv[i], rv[i] - acc[i], zo[i]);
                Z[static_cast<int64_t>(row) * rp] = zi;
                zr = fma(zi, rv[i], zr);
                zq = fma(zi, qv[i], zq);
            }
        }
    }
    block_sum2<L>(zq, zr);
    if (threadIdx.x < L) {
        adt::gview(g.part0)[static_cast<int64_t>(tix) * L + threadIdx.x] = zq;
        adt::gview(g.part1)[static_cast<int64_t>(tix) * L + threadIdx.x] = zr;
    }
}

// ------------------------------------------------------------------------------------------
// LDS-staged versions
// ------------------------------------------------------------------------------------------
// A staged product y_i = f(i, sum_k val_k * x[col_k]) over the rows i of a tile reads every
// needed row of x once per tile into LDS instead of gathering it per nonzero.  The staged rows
// are the tile's own rows (smoothing; not for the prolongation, whose x lives on the coarser
// level) followed by the tile's halo list (host-built per pattern tile, sorted), and every
// slot of the tile's CSR slice carries the LDS row of its column (16-bit).  One pass, one
// barrier:
//   * the CSR slice (values, local column slots, row pointers) is copied into LDS while
//   * each thread loads its own rows' operands into registers (kept for the product pass) and
//     the staged values of its share of the halo rows (row ids straight from the global halo
//     list, halo_batch(L) rows with loads in flight together);
//   * barrier; the product reads matrix and staged vector from LDS only.
// Used for the finest level's smoothing (the conjugate-gradient product's tables), the coarse
// levels' smoothing and every level's prolongation.

// Where a staged product's matrix and tables live (pointers already offset to the system).
struct StageSrc {
    const int32_t* ptr;     // row pointers (system rows)
    const double* val;      // slot values
    const uint16_t* lidx;   // slot -> staged row of its column
    const int* halo_ptr;    // per pattern tile: first entry in halo_rows
    const int* halo_rows;   // staged rows outside the tile (or, for the prolongation, all of them)
    int lds_rows, lds_nnz, lds_own;  // LDS capacities: staged rows, slots, tile rows
};

__device__ __forceinline__ StageSrc fine_src(const PcgArgs& g, const Sys& S) {
    return StageSrc{g.indptr + S.ptr_off, g.values + S.val_off, g.lidx + S.lidx_off, g.halo_ptr, g.halo_rows,
                    g.lds_rows, g.lds_nnz, g.lds_own};
}

__device__ __forceinline__ StageSrc level_src(const LevelArgs& a, const LvSys& S) {
    return StageSrc{a.ptr + S.ptr, a.val + S.val, a.lidx + S.idx, a.halo_ptr, a.halo_rows,
                    a.lds_rows, a.lds_nnz, a.lds_own};
}

__device__ __forceinline__ StageSrc prolong_src(const LevelArgs& a, const LvSys& S) {
    return StageSrc{a.pptr + S.pptr, a.pval + S.pval, a.plidx + S.pidx, a.pc_ptr, a.pc_rows,
                    a.p_lds_rows, a.p_lds_nnz, a.p_lds_own};
}

struct LdsTile {
    double* stage;  // staged rows x L values
    double* sval;
    int* sptr;
    uint16_t* sidx;
    int rows;       // rows of the tile
    int own;        // own rows staged first (rows, or 0)
    int nh, h0, e0, ne;
};

__device__ __forceinline__ LdsTile lds_tile(const StageSrc& src, const Tile& t, int L, double* smem, bool stage_own) {
    LdsTile s;
    const auto ptr = adt::gview(src.ptr);
    s.rows = t.row1 - t.row0;
    s.own = stage_own ? s.rows : 0;
    s.h0 = adt::gview(src.halo_ptr)[t.ptile];
    s.nh = adt::gview(src.halo_ptr)[t.ptile + 1] - s.h0;
    s.e0 = ptr[t.row0];
    s.ne = ptr[t.row1] - s.e0;
    s.stage = smem;
    s.sval = smem + static_cast<int64_t>(src.lds_rows) * L;
    s.sptr = reinterpret_cast<int*>(s.sval + src.lds_nnz);
    s.sidx = reinterpret_cast<uint16_t*>(s.sptr + src.lds_own + 1);
    return s;
}

// The tile's CSR slice in registers: the first 512 slots and row pointer t of thread t (loaded
// unconditionally, stored predicated).
struct CsrRegs {
    double va, vb;
    int pi;
    uint16_t ia, ib;
};

__device__ __forceinline__ CsrRegs load_csr(const StageSrc& src, const Tile& t, const LdsTile& s) {
    const auto ptr = adt::gview(src.ptr);
    const auto val = adt::gview(src.val);
    const auto lidx = adt::gview(src.lidx);
    const int tid = static_cast<int>(threadIdx.x);
    // positions past the slice read the system's slot 0 (the slice itself may be empty)
    const int sa = tid < s.ne ? s.e0 + tid : 0, sb = 256 + tid < s.ne ? s.e0 + 256 + tid : 0;
    CsrRegs c;
    c.va = val[sa];
    c.vb = val[sb];
    c.ia = lidx[sa];
    c.ib = lidx[sb];
    c.pi = ptr[t.row0 + (tid <= s.rows ? tid : 0)];
    return c;
}

// Stores the slice into LDS (published by the caller's barrier); slices longer than 512 slots
// or tiles of more than 255 rows continue with plain loops.
__device__ __forceinline__ void store_csr(const StageSrc& src, const Tile& t, const LdsTile& s, const CsrRegs& c) {
    const int tid = static_cast<int>(threadIdx.x);
    if (tid < s.ne) {
        s.sval[tid] = c.va;
        s.sidx[tid] = c.ia;
    }
    if (256 + tid < s.ne) {
        s.sval[256 + tid] = c.vb;
        s.sidx[256 + tid] = c.ib;
    }
    if (tid <= s.rows) s.sptr[tid] = c.pi - s.e0;
    if (s.ne > 512 || s.rows >= 256) {
        const auto ptr = adt::gview(src.ptr);
        const auto val = adt::gview(src.val + s.e0);
        const auto lidx = adt::gview(src.lidx + s.e0);
        for (int e = 512 + tid; e < s.ne; e += 256) {
            s.sval[e] = val[e];
            s.sidx[e] = lidx[e];
        }
        for (int i = 256 + tid; i <= s.rows; i += 256) s.sptr[i] = ptr[t.row0 + i] - s.e0;
    }
}

// Halo row ids of one staging batch: list positions base + t / L + u * (256 / L), u <
// halo_batch(L).  Positions past the list get row 0 (a valid row whose value is not stored).
template <int L>
__device__ __forceinline__ void halo_ids(const StageSrc& src, const LdsTile& s, int base,
                                         int (&jj)[halo_batch(L)]) {
    constexpr int RPP = kRowsPerPass(L);
    const auto halo = adt::gview(src.halo_rows);
    const int h0 = base + static_cast<int>(threadIdx.x / L);
#pragma unroll
    for (int u = 0; u < halo_batch(L); ++u) {
        const bool in = h0 + u * RPP < s.nh;
        const int raw = halo[in ? s.h0 + h0 + u * RPP : 0];
        jj[u] = in ? raw : 0;
    }
}

// Staged values of one halo batch (loads only).
template <int L, typename ValueF>
__device__ __forceinline__ void halo_values(const int (&jj)[halo_batch(L)], double (&v)[halo_batch(L)], ValueF value) {
#pragma unroll
    for (int u = 0; u < halo_batch(L); ++u) v[u] = value(jj[u]);
}

template <int L>
__device__ __forceinline__ void store_halo(const LdsTile& s, int base, const double (&v)[halo_batch(L)]) {
    constexpr int RPP = kRowsPerPass(L);
    const int lane = static_cast<int>(threadIdx.x % L), h0 = base + static_cast<int>(threadIdx.x / L);
#pragma unroll
    for (int u = 0; u < halo_batch(L); ++u)
        if (h0 + u * RPP < s.nh) s.stage[(s.own + h0 + u * RPP) * L + lane] = v[u];
}

// Halo batches after the first (tiles with a large halo only).
template <int L, typename ValueF>
__device__ __forceinline__ void stage_halo_rest(const StageSrc& src, const LdsTile& s, ValueF value) {
    constexpr int step = halo_batch(L) * kRowsPerPass(L);
    for (int base = step; base < s.nh; base += step) {
        int jj[halo_batch(L)];
        double v[halo_batch(L)];
        halo_ids<L>(src, s, base, jj);
        halo_values<L>(jj, v, value);
        store_halo<L>(s, base, v);
        adt::wait_vector_memory();  // the batch's loads are not left pending past this optional loop
    }
}

// sum_k sval[k] * stage[sidx[k]] over tile row q, in slot order, 8 slots in flight; slots past
// the row re-read its first slot and are discarded (an empty row sums to zero).
template <int L>
__device__ __forceinline__ double lds_row_product(const LdsTile& s, int q, int lane) {
    constexpr int U = 8;
    const int k0 = s.sptr[q], k1 = s.sptr[q + 1];
    double acc = 0.0;
    for (int kb = k0; kb < k1; kb += U) {
        int jj[U];
        double vv[U], tt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = kb + u < k1 ? kb + u : k0;
            jj[u] = static_cast<int>(s.sidx[k]);
            vv[u] = s.sval[k];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) tt[u] = s.stage[jj[u] * L + lane];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double next = fma(vv[u], tt[u], acc);
            acc = kb + u < k1 ? next : acc;
        }
    }
    return acc;
}

// Row of own-row slot i of the calling thread (rows past the tile read the tile's first row;
// their values are never stored).
template <int L>
__device__ __forceinline__ int own_row(const Tile& t, int rows, int i) {
    const int q = static_cast<int>(threadIdx.x / L) + i * kRowsPerPass(L);
    return t.row0 + (q < rows ? q : 0);
}

// Finest level: r1 = r - A (w D^-1 r), written into the z buffer.
template <int L>
__global__ __launch_bounds__(256) void amg_smooth_residual_lds(const PcgArgs* __restrict__ pa,
                                                               const LevelArgs* __restrict__ la) {
    extern __shared__ double smem[];
    const PcgArgs& g = *pa;
    const LevelArgs& a = *la;
    const int tix = adt::xcd_tile(blockIdx.x, g.n_tiles);
    const Tile t = adt::gview(g.tiles)[tix];
    if (!adt::gview(g.grp_active)[t.grp]) return;
    constexpr int RPP = kRowsPerPass(L);
    const Sys S = adt::gview(g.sys)[t.sys];
    const double w = adt::gview(a.ls)[t.sys].omega;
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % g.groups) * L + lane;
    const int rp = g.rp;
    const auto R = adt::gview(g.r + S.vec_off + col);
    const auto R1 = adt::gview(g.z + S.vec_off + col);
    const auto D = adt::gview(g.dinv + S.diag_off);
    const StageSrc src = fine_src(g, S);
    const LdsTile s = lds_tile(src, t, L, smem, true);
    const auto value = [&](int j) { return w * D[j] * R[static_cast<int64_t>(j) * rp]; };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double rv[own_max(L)], dv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int row = own_row<L>(t, s.rows, i);
        rv[i] = R[static_cast<int64_t>(row) * rp];
        dv[i] = D[row];
    }
    const CsrRegs cr = load_csr(src, t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, t, s, cr);
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) s.stage[q * L + lane] = w * dv[i] * rv[i];
    }
    store_halo<L>(s, 0, hv);
    stage_halo_rest<L>(src, s, value);
    adt::wait_vector_memory();  // the product loops below then never wait for the stores of earlier rows
    __syncthreads();
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) R1[static_cast<int64_t>(t.row0 + q) * rp] = rv[i] - lds_row_product<L>(s, q, lane);
    }
}

// Finest level: z = z2 + w D^-1 (r - A z2) with z2 (written by the finest prolongation into
// p[parity ^ 1]) staged for the tile and its halo; partials z.q_old (part0) and z.r (part1) for
// the flexible-CG coefficients.
template <int L>
__global__ __launch_bounds__(256) void amg_post_smooth_lds(const PcgArgs* __restrict__ pa,
                                                           const LevelArgs* __restrict__ la, int parity) {
    extern __shared__ double smem[];
    const PcgArgs& g = *pa;
    const LevelArgs& a = *la;
    const int tix = adt::xcd_tile(blockIdx.x, g.n_tiles);
    const Tile t = adt::gview(g.tiles)[tix];
    if (!adt::gview(g.grp_active)[t.grp]) return;
    constexpr int RPP = kRowsPerPass(L);
    const Sys S = adt::gview(g.sys)[t.sys];
    const double w = adt::gview(a.ls)[t.sys].omega;
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % g.groups) * L + lane;
    const int rp = g.rp;
    const auto R = adt::gview(residual(g, parity) + S.vec_off + col);
    const auto Q = adt::gview(g.q + S.vec_off + col);
    const auto Z2 = adt::gview(g.p[parity ^ 1] + S.vec_off + col);
    const auto Z = adt::gview(g.z + S.vec_off + col);
    const auto D = adt::gview(g.dinv + S.diag_off);
    const StageSrc src = fine_src(g, S);
    const LdsTile s = lds_tile(src, t, L, smem, true);
    const auto value = [&](int j) { return Z2[static_cast<int64_t>(j) * rp]; };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double zv[own_max(L)], rv[own_max(L)], qv[own_max(L)], dv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int row = own_row<L>(t, s.rows, i);
        const int64_t o = static_cast<int64_t>(row) * rp;
        zv[i] = Z2[o];
        rv[i] = R[o];
        qv[i] = Q[o];
        dv[i] = D[row];
    }
    const CsrRegs cr = load_csr(src, t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, t, s, cr);
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) s.stage[q * L + lane] = zv[i];
    }
    store_halo<L>(s, 0, hv);
    stage_halo_rest<L>(src, s, value);
    __syncthreads();
    double zr = 0.0, zq = 0.0;
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) {
            const double acc = lds_row_product<L>(s, q, lane);
            const double zi = fma(w * dv[i], rv[i] - acc, zv[i]);
            Z[static_cast<int64_t>(t.row0 + q) * rp] = zi;
            zr = fma(zi, rv[i], zr);
            zq = fma(zi, qv[i], zq);
        }
    }
    block_sum2<L>(zq, zr);
    if (threadIdx.x < L) {
        adt::gview(g.part0)[static_cast<int64_t>(tix) * L + threadIdx.x] = zq;
        adt::gview(g.part1)[static_cast<int64_t>(tix) * L + threadIdx.x] = zr;
    }
}

// Coarse levels: r1 = b - A (w D^-1 b) with w D^-1 b staged.
template <int L>
__global__ __launch_bounds__(256) void amg_smooth_residual_st(const LevelArgs* __restrict__ la) {
    extern __shared__ double smem[];
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    constexpr int RPP = kRowsPerPass(L);
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int rp = a.rp;
    const auto D = adt::gview(a.dinv + c.S.dinv);
    const auto B = adt::gview(a.b + c.S.vec + c.col);
    const auto R1 = adt::gview(a.r1 + c.S.vec + c.col);
    const double w = c.S.omega;
    const StageSrc src = level_src(a, c.S);
    const LdsTile s = lds_tile(src, c.t, L, smem, true);
    const auto value = [&](int j) { return w * D[j] * B[static_cast<int64_t>(j) * rp]; };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double bv[own_max(L)], dv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int row = own_row<L>(c.t, s.rows, i);
        bv[i] = B[static_cast<int64_t>(row) * rp];
        dv[i] = D[row];
    }
    const CsrRegs cr = load_csr(src, c.t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, c.t, s, cr);
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) s.stage[q * L + lane] = w * dv[i] * bv[i];
    }
    store_halo<L>(s, 0, hv);
    stage_halo_rest<L>(src, s, value);
    adt::wait_vector_memory();  // the product loops below then never wait for the stores of earlier rows
    __syncthreads();
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) R1[static_cast<int64_t>(c.t.row0 + q) * rp] = bv[i] - lds_row_product<L>(s, q, lane);
    }
}

// Coarse levels: e = z2 + w D^-1 (b - A z2) with z2 staged.
template <int L>
__global__ __launch_bounds__(256) void amg_post_smooth_st(const LevelArgs* __restrict__ la) {
    extern __shared__ double smem[];
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    constexpr int RPP = kRowsPerPass(L);
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int rp = a.rp;
    const auto D = adt::gview(a.dinv + c.S.dinv);
    const auto B = adt::gview(a.b + c.S.vec + c.col);
    const auto Z2 = adt::gview(a.z2 + c.S.vec + c.col);
    const auto E = adt::gview(a.e + c.S.vec + c.col);
    const double w = c.S.omega;
    const StageSrc src = level_src(a, c.S);
    const LdsTile s = lds_tile(src, c.t, L, smem, true);
    const auto value = [&](int j) { return Z2[static_cast<int64_t>(j) * rp]; };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double zv[own_max(L)], bv[own_max(L)], dv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int row = own_row<L>(c.t, s.rows, i);
        const int64_t o = static_cast<int64_t>(row) * rp;
        zv[i] = Z2[o];
        bv[i] = B[o];
        dv[i] = D[row];
    }
    const CsrRegs cr = load_csr(src, c.t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, c.t, s, cr);
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) s.stage[q * L + lane] = zv[i];
    }
    store_halo<L>(s, 0, hv);
    stage_halo_rest<L>(src, s, value);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) {
            const double acc = lds_row_product<L>(s, q, lane);
            E[static_cast<int64_t>(c.t.row0 + q) * rp] = fma(w * dv[i], bv[i] - acc, zv[i]);
        }
    }
}

// Any level: z2 = w D^-1 b + P e_{l+1} with the coarse rows that the tile's P rows reference
// staged (the finest level writes z2 into p[parity ^ 1], see amg_prolong).
template <int L, bool FINE>
__global__ __launch_bounds__(256) void amg_prolong_st(const PcgArgs* __restrict__ pa, const LevelArgs* __restrict__ la,
                                                      int parity) {
    extern __shared__ double smem[];
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    constexpr int RPP = kRowsPerPass(L);
    const int rp = a.rp;
    const auto D = adt::gview(a.dinv + c.S.dinv);
    const auto B = adt::gview((FINE ? residual(*pa, parity) : a.b) + c.S.vec + c.col);
    const auto NE = adt::gview(a.next_e + c.S.next_vec + c.col);
    const auto Z2 = adt::gview((FINE ? pa->p[parity ^ 1] : a.z2) + c.S.vec + c.col);
    const double w = c.S.omega;
    const StageSrc src = prolong_src(a, c.S);
    const LdsTile s = lds_tile(src, c.t, L, smem, false);
    const auto value = [&](int J) { return NE[static_cast<int64_t>(J) * rp]; };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double bv[own_max(L)], dv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int row = own_row<L>(c.t, s.rows, i);
        bv[i] = B[static_cast<int64_t>(row) * rp];
        dv[i] = D[row];
    }
    const CsrRegs cr = load_csr(src, c.t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, c.t, s, cr);
    store_halo<L>(s, 0, hv);
    stage_halo_rest<L>(src, s, value);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows)
            Z2[static_cast<int64_t>(c.t.row0 + q) * rp] = fma(w * dv[i], bv[i], lds_row_product<L>(s, q, lane));
    }
}

}  // namespace krylov
}  // namespace adt

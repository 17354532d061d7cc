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
// prolongator row from the filtered matrix (strong entries, weak ones lumped into D_F):
//   P[i, J] = [J == agg(i)] (1 - omega_ii) - D_F,ii^-1 sum_{strong k in row i, agg(col k) = J} omega_k A_k
// with omega_k the level's constant damping or, when given, the learned per-slot damping.
__global__ __launch_bounds__(256) void amg_pvals(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    const LvSys S = a.ls[blockIdx.y];
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= S.rows) return;
    const double* V = a.val + S.val;
    const int dslot = a.diag[S.diag + i];
    if (a.dinv_out) a.dinv_out[S.dinv + i] = 1.0 / V[dslot];
    if (S.nc == 0) return;
    const int* dmap_ptr = a.dmap_ptr + S.dmap_ptr;
    const int* dmap = a.dmap + S.dmap;
    double df = 0.0;
    for (int m = dmap_ptr[i]; m < dmap_ptr[i + 1]; ++m) df += V[dmap[m]];
    const double dfinv = 1.0 / (df > 0.0 ? df : V[dslot]);
    const int* pptr = a.pptr + S.pptr;
    const int* pmap_ptr = a.pmap_ptr + S.pmap_ptr;
    const int* pmap = a.pmap + S.pmap;
    const int* pself = a.pself + S.pidx;
    const double* om = a.omega_slot ? a.omega_slot + S.val : nullptr;
    const double wself = om ? om[dslot] : S.omega_p;
    double* P = a.pval + S.pval;
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
    const LvSys S = a.ls[blockIdx.y];
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= S.n_t) return;
    const int* tp = a.tpair_ptr + S.tpair_ptr;
    const int* ta = a.tpair_a + S.tpair;
    const int* tq = a.tpair_p + S.tpair;
    const double* V = a.val + S.val;
    const double* P = a.pval + S.pval;
    double acc = 0.0;
    for (int m = tp[t]; m < tp[t + 1]; ++m) acc = fma(V[ta[m]], P[tq[m]], acc);
    a.tval[S.tval + t] = acc;
}

// Thread per slot of the next level's matrix: A_{l+1}[J, K] = sum_i P[i, J] T[i, K].
__global__ __launch_bounds__(256) void amg_cvals(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    const LvSys S = a.ls[blockIdx.y];
    const int c = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= S.n_next) return;
    const int* cp = a.cpair_ptr + S.cpair_ptr;
    const int* cs = a.cpair_p + S.cpair;
    const int* ct = a.cpair_t + S.cpair;
    const double* P = a.pval + S.pval;
    const double* T = a.tval + S.tval;
    double acc = 0.0;
    for (int m = cp[c]; m < cp[c + 1]; ++m) acc = fma(P[cs[m]], T[ct[m]], acc);
    a.next_valp[S.next_val + c] = acc;
}

// One workgroup per system: dense copy of the coarsest operator inverted in place by
// Gauss-Jordan elimination without pivoting (the operator is SPD).  A non-positive pivot
// (a singular coarse operator) zeroes that row and column of the inverse instead.
// Dynamic LDS: one column (rows doubles).
__global__ __launch_bounds__(256) void amg_dense_inverse(const LevelArgs* __restrict__ la) {
    extern __shared__ double colk[];
    const LevelArgs& a = *la;
    const LvSys S = a.ls[blockIdx.x];
    const int n = S.rows;
    double* M = a.dense + S.dense;
    const int32_t* ptr = a.ptr + S.ptr;
    const int32_t* idx = a.idx + S.idx;
    const double* V = a.val + S.val;
    const int64_t nn = static_cast<int64_t>(n) * n;
    for (int64_t e = threadIdx.x; e < nn; e += blockDim.x) M[e] = 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) M[static_cast<int64_t>(i) * n + idx[k]] = V[k];
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        double* rowk = M + static_cast<int64_t>(k) * n;
        const double piv = rowk[k];
        const double inv = piv > 0.0 && std::isfinite(piv) ? 1.0 / piv : 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) colk[i] = M[static_cast<int64_t>(i) * n + k];
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += blockDim.x) rowk[j] = (j == k) ? inv : rowk[j] * inv;
        __syncthreads();
        // 16 x 16 thread layout over (row, column)
        for (int i = static_cast<int>(threadIdx.x) / 16; i < n; i += 16) {
            if (i == k) continue;
            const double f = colk[i];
            double* Mi = M + static_cast<int64_t>(i) * n;
            for (int j = static_cast<int>(threadIdx.x) % 16; j < n; j += 16) {
                const double m = (j == k) ? 0.0 : Mi[j];
                Mi[j] = fma(-f, rowk[j], m);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// cycle kernels, gather versions (any level)
// ------------------------------------------------------------------------------------------

// sum_k val(k) * term(idx(k)) over k in [k0, k1), U entries at a time: all index and value
// loads first, then all gathers, so one thread keeps U independent requests in flight.  The
// accumulation order is k0, k0+1, ... (padding entries add 0 * term(pad) with a valid pad row).
template <int U, typename IdxF, typename ValF, typename TermF>
__device__ __forceinline__ double gather_sum(int k0, int k1, IdxF idx, ValF val, TermF term, int64_t pad) {
    double acc = 0.0;
    for (int kb = k0; kb < k1; kb += U) {
        int64_t jj[U];
        double vv[U], tt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = kb + u < k1;
            jj[u] = ok ? idx(kb + u) : pad;
            vv[u] = ok ? val(kb + u) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) tt[u] = term(jj[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) acc = fma(vv[u], tt[u], acc);
    }
    return acc;
}

struct LevelTile {
    Tile t;
    LvSys S;
    int col;
};

template <int L>
__device__ __forceinline__ bool level_tile(const LevelArgs& a, const Tile* tiles, int n_tiles, LevelTile& c) {
    c.t = tiles[adt::xcd_tile(blockIdx.x, n_tiles)];
    if (!a.grp_active[c.t.grp]) return false;
    c.S = a.ls[c.t.sys];
    c.col = (c.t.grp % a.groups) * L + static_cast<int>(threadIdx.x % L);
    return true;
}

// r1 = b - A (w D^-1 b)
template <int L>
__global__ __launch_bounds__(256) void amg_smooth_residual(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    const int rp = a.rp;
    const int32_t* ptr = a.ptr + c.S.ptr;
    const int32_t* idx = a.idx + c.S.idx;
    const double* val = a.val + c.S.val;
    const double* D = a.dinv + c.S.dinv;
    const double* B = a.b + c.S.vec + c.col;
    double* R1 = a.r1 + c.S.vec + c.col;
    const double w = c.S.omega;
    for (int row = c.t.row0 + static_cast<int>(threadIdx.x / L); row < c.t.row1; row += kRowsPerPass(L)) {
        const double acc = gather_sum<8>(
            ptr[row], ptr[row + 1], [&](int k) { return static_cast<int64_t>(idx[k]); }, [&](int k) { return val[k]; },
            [&](int64_t j) { return w * D[j] * B[j * rp]; }, row);
        const int64_t o = static_cast<int64_t>(row) * rp;
        R1[o] = B[o] - acc;
    }
}

// b_{l+1} = P^T r1 (tiles of the next level)
template <int L>
__global__ __launch_bounds__(256) void amg_restrict(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.next_tiles, a.n_next_tiles, c)) return;
    const int rp = a.rp;
    const int* rptr = a.rptr + c.S.rptr;
    const int* rrow = a.rrow + c.S.rent;
    const int* rslot = a.rslot + c.S.rent;
    const double* P = a.pval + c.S.pval;
    const double* R1 = a.r1 + c.S.vec + c.col;
    double* NB = a.next_b + c.S.next_vec + c.col;
    for (int J = c.t.row0 + static_cast<int>(threadIdx.x / L); J < c.t.row1; J += kRowsPerPass(L)) {
        const int e0 = rptr[J];
        const double acc = gather_sum<8>(
            e0, rptr[J + 1], [&](int e) { return static_cast<int64_t>(e); }, [&](int e) { return P[rslot[e]]; },
            [&](int64_t e) { return R1[static_cast<int64_t>(rrow[e]) * rp]; }, e0);
        NB[static_cast<int64_t>(J) * rp] = acc;
    }
}

// e = A_c^-1 b on the coarsest level: workgroup (lane group, block of 256 / L rows), one
// thread per (row, lane)
template <int L>
__global__ __launch_bounds__(256) void amg_coarse_solve(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    const int gid = blockIdx.x;
    if (!a.grp_active[gid]) return;
    const LvSys S = a.ls[gid / a.groups];
    const int col = (gid % a.groups) * L + static_cast<int>(threadIdx.x % L);
    const int n = S.rows, rp = a.rp;
    const int i = static_cast<int>(blockIdx.y) * kRowsPerPass(L) + static_cast<int>(threadIdx.x / L);
    if (i >= n) return;
    const double* Mi = a.dense + S.dense + static_cast<int64_t>(i) * n;
    const double* B = a.b + S.vec + col;
    double a0 = 0.0, a1 = 0.0;
    int j = 0;
    for (; j + 1 < n; j += 2) {
        a0 = fma(Mi[j], B[static_cast<int64_t>(j) * rp], a0);
        a1 = fma(Mi[j + 1], B[static_cast<int64_t>(j + 1) * rp], a1);
    }
    if (j < n) a0 = fma(Mi[j], B[static_cast<int64_t>(j) * rp], a0);
    a.e[S.vec + col + static_cast<int64_t>(i) * rp] = a0 + a1;
}

// z2 = w D^-1 b + P e_{l+1}.  On the finest level (FINE) z2 goes to the direction buffer that
// the coming product overwrites (p[parity ^ 1] holds the direction of two iterations ago).
template <int L, bool FINE>
__global__ __launch_bounds__(256) void amg_prolong(const PcgArgs* __restrict__ pa, const LevelArgs* __restrict__ la,
                                                   int parity) {
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    const int rp = a.rp;
    const int* pptr = a.pptr + c.S.pptr;
    const int* pidx = a.pidx + c.S.pidx;
    const double* P = a.pval + c.S.pval;
    const double* D = a.dinv + c.S.dinv;
    const double* B = a.b + c.S.vec + c.col;
    const double* NE = a.next_e + c.S.next_vec + c.col;
    double* Z2 = (FINE ? pa->p[parity ^ 1] : a.z2) + c.S.vec + c.col;
    const double w = c.S.omega;
    for (int row = c.t.row0 + static_cast<int>(threadIdx.x / L); row < c.t.row1; row += kRowsPerPass(L)) {
        const int64_t o = static_cast<int64_t>(row) * rp;
        const double acc = gather_sum<4>(
            pptr[row], pptr[row + 1], [&](int s) { return static_cast<int64_t>(pidx[s]); }, [&](int s) { return P[s]; },
            [&](int64_t J) { return NE[J * rp]; }, 0);
        Z2[o] = fma(w * D[row], B[o], acc);
    }
}

// e = z2 + w D^-1 (b - A z2)
template <int L>
__global__ __launch_bounds__(256) void amg_post_smooth(const LevelArgs* __restrict__ la) {
    const LevelArgs& a = *la;
    LevelTile c;
    if (!level_tile<L>(a, a.tiles, a.n_tiles, c)) return;
    const int rp = a.rp;
    const int32_t* ptr = a.ptr + c.S.ptr;
    const int32_t* idx = a.idx + c.S.idx;
    const double* val = a.val + c.S.val;
    const double* D = a.dinv + c.S.dinv;
    const double* B = a.b + c.S.vec + c.col;
    const double* Z2 = a.z2 + c.S.vec + c.col;
    double* E = a.e + c.S.vec + c.col;
    const double w = c.S.omega;
    for (int row = c.t.row0 + static_cast<int>(threadIdx.x / L); row < c.t.row1; row += kRowsPerPass(L)) {
        const double acc = gather_sum<8>(
            ptr[row], ptr[row + 1], [&](int k) { return static_cast<int64_t>(idx[k]); }, [&](int k) { return val[k]; },
            [&](int64_t j) { return Z2[j * rp]; }, row);
        const int64_t o = static_cast<int64_t>(row) * rp;
        E[o] = fma(w * D[row], B[o] - acc, Z2[o]);
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
    const Tile t = a.tiles[tix];
    if (!a.grp_active[t.grp]) return;
    const LvSys S = a.ls[t.sys];
    const int col = (t.grp % a.groups) * L + static_cast<int>(threadIdx.x % L);
    const int rp = a.rp;
    const int32_t* ptr = a.ptr + S.ptr;
    const int32_t* idx = a.idx + S.idx;
    const double* val = a.val + S.val;
    const double* D = a.dinv + S.dinv;
    const double* R = a.b + S.vec + col;
    const double* Z2 = g.p[parity ^ 1] + S.vec + col;
    const double* Q = g.q + S.vec + col;
    double* Z = a.e + S.vec + col;
    const double w = S.omega;
    double zr = 0.0, zq = 0.0;
    for (int row = t.row0 + static_cast<int>(threadIdx.x / L); row < t.row1; row += kRowsPerPass(L)) {
        const double acc = gather_sum<8>(
            ptr[row], ptr[row + 1], [&](int k) { return static_cast<int64_t>(idx[k]); }, [&](int k) { return val[k]; },
            [&](int64_t j) { return Z2[j * rp]; }, row);
        const int64_t o = static_cast<int64_t>(row) * rp;
        const double ri = R[o];
        const double zi = fma(w * D[row], ri - acc, Z2[o]);
        Z[o] = zi;
        zr = fma(zi, ri, zr);
        zq = fma(zi, Q[o], zq);
    }
    block_sum2<L>(zq, zr);
    if (threadIdx.x < L) {
        g.part0[static_cast<int64_t>(tix) * L + threadIdx.x] = zq;
        g.part1[static_cast<int64_t>(tix) * L + threadIdx.x] = zr;
    }
}

// ------------------------------------------------------------------------------------------
// finest level, LDS-staged versions (tiles and halo tables of the conjugate-gradient product)
// ------------------------------------------------------------------------------------------

struct LdsTile {
    double* stage;
    double* sval;
    int* sptr;
    int* shalo;
    uint16_t* sidx;
    int own, nh, staged;
};

// Phase 1 of every staged product: the tile's CSR slice and halo list into LDS.
__device__ __forceinline__ LdsTile stage_csr(const PcgArgs& g, const Tile& t, const Sys& S, int L, double* smem) {
    LdsTile s;
    const int32_t* ptr = g.indptr + S.ptr_off;
    s.own = t.row1 - t.row0;
    const int h0 = g.halo_ptr[t.ptile];
    s.nh = g.halo_ptr[t.ptile + 1] - h0;
    s.staged = s.own + s.nh;
    const int e0 = ptr[t.row0], ne = ptr[t.row1] - e0;
    s.stage = smem;
    s.sval = smem + static_cast<int64_t>(g.lds_rows) * L;
    s.sptr = reinterpret_cast<int*>(s.sval + g.lds_nnz);
    s.shalo = s.sptr + g.lds_own + 1;
    s.sidx = reinterpret_cast<uint16_t*>(s.shalo + g.lds_halo);
    const double* val = g.values + S.val_off + e0;
    const uint16_t* lidx = g.lidx + S.lidx_off + e0;
    for (int e = static_cast<int>(threadIdx.x); e < ne; e += 256) {
        s.sval[e] = val[e];
        s.sidx[e] = lidx[e];
    }
    for (int i = static_cast<int>(threadIdx.x); i <= s.own; i += 256) s.sptr[i] = ptr[t.row0 + i] - e0;
    for (int i = static_cast<int>(threadIdx.x); i < s.nh; i += 256) s.shalo[i] = g.halo_rows[h0 + i];
    __syncthreads();
    return s;
}

// r1 = r - A (w D^-1 r), written into the z buffer.
template <int L>
__global__ __launch_bounds__(256) void amg_smooth_residual_lds(const PcgArgs* __restrict__ pa,
                                                               const LevelArgs* __restrict__ la) {
    extern __shared__ double smem[];
    const PcgArgs& g = *pa;
    const LevelArgs& a = *la;
    const int tix = adt::xcd_tile(blockIdx.x, g.n_tiles);
    const Tile t = g.tiles[tix];
    if (!g.grp_active[t.grp]) return;
    const Sys S = g.sys[t.sys];
    const double w = a.ls[t.sys].omega;
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % g.groups) * L + lane;
    const int rp = g.rp;
    const double* R = g.r + S.vec_off + col;
    double* R1 = g.z + S.vec_off + col;
    const double* D = g.dinv + S.diag_off;
    const LdsTile s = stage_csr(g, t, S, L, smem);
    constexpr int U = 4;
    for (int s0 = rsub; s0 < s.staged; s0 += U * kRowsPerPass(L)) {
        int64_t jj[U];
        double dj[U], rj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = s0 + u * kRowsPerPass(L);
            jj[u] = q < s.own ? t.row0 + q : (q < s.staged ? s.shalo[q - s.own] : t.row0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            dj[u] = D[jj[u]];
            rj[u] = R[jj[u] * rp];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = s0 + u * kRowsPerPass(L);
            if (q < s.staged) s.stage[q * L + lane] = w * dj[u] * rj[u];
        }
    }
    __syncthreads();
    for (int q = rsub; q < s.own; q += kRowsPerPass(L)) {
        const int k1 = s.sptr[q + 1];
        double acc = 0.0;
        for (int k = s.sptr[q]; k < k1; ++k) acc = fma(s.sval[k], s.stage[static_cast<int>(s.sidx[k]) * L + lane], acc);
        const int64_t o = static_cast<int64_t>(t.row0 + q) * rp;
        R1[o] = R[o] - acc;
    }
}

// z = z2 + w D^-1 (r - A z2) with z2 (written by the finest prolongation into p[parity ^ 1])
// staged for the tile and its halo; partials z.q_old (part0) and z.r (part1) for the
// flexible-CG coefficients.
template <int L>
__global__ __launch_bounds__(256) void amg_post_smooth_lds(const PcgArgs* __restrict__ pa,
                                                           const LevelArgs* __restrict__ la, int parity) {
    extern __shared__ double smem[];
    const PcgArgs& g = *pa;
    const LevelArgs& a = *la;
    const int tix = adt::xcd_tile(blockIdx.x, g.n_tiles);
    const Tile t = g.tiles[tix];
    if (!g.grp_active[t.grp]) return;
    const Sys S = g.sys[t.sys];
    const double w = a.ls[t.sys].omega;
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % g.groups) * L + lane;
    const int rp = g.rp;
    const double* R = g.r + S.vec_off + col;
    const double* Q = g.q + S.vec_off + col;
    const double* Z2 = g.p[parity ^ 1] + S.vec_off + col;
    double* Z = g.z + S.vec_off + col;
    const double* D = g.dinv + S.diag_off;
    const LdsTile s = stage_csr(g, t, S, L, smem);
    constexpr int U = 4;
    for (int s0 = rsub; s0 < s.staged; s0 += U * kRowsPerPass(L)) {
        int64_t jj[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = s0 + u * kRowsPerPass(L);
            jj[u] = q < s.own ? t.row0 + q : (q < s.staged ? s.shalo[q - s.own] : t.row0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = Z2[jj[u] * rp];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = s0 + u * kRowsPerPass(L);
            if (q < s.staged) s.stage[q * L + lane] = v[u];
        }
    }
    __syncthreads();
    double zr = 0.0, zq = 0.0;
    for (int q = rsub; q < s.own; q += kRowsPerPass(L)) {
        const int k1 = s.sptr[q + 1];
        double acc = 0.0;
        for (int k = s.sptr[q]; k < k1; ++k) acc = fma(s.sval[k], s.stage[static_cast<int>(s.sidx[k]) * L + lane], acc);
        const int row = t.row0 + q;
        const int64_t o = static_cast<int64_t>(row) * rp;
        const double ri = R[o];
        const double zi = fma(w * D[row], ri - acc, s.stage[q * L + lane]);
        Z[o] = zi;
        zr = fma(zi, ri, zr);
        zq = fma(zi, Q[o], zq);
    }
    block_sum2<L>(zq, zr);
    if (threadIdx.x < L) {
        g.part0[static_cast<int64_t>(tix) * L + threadIdx.x] = zq;
        g.part1[static_cast<int64_t>(tix) * L + threadIdx.x] = zr;
    }
}

}  // namespace krylov
}  // namespace adt

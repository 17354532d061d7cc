// (b) Batched multi-RHS FP64 sparse products, Jacobi-preconditioned conjugate gradients and
// multigrid-preconditioned flexible conjugate gradients.
//
// Jacobi PCG, one iteration = four launches (captured once in a hipGraph and replayed):
//   spmv   : p_new = z + beta p_old (formed on the fly for every neighbour, z = D^-1 r),
//            q = A p_new, partial sums of p_new.q per (tile, lane)
//   alpha  : one workgroup per lane group reduces the tile partials in a fixed order
//   update : x += alpha p_new, r -= alpha q, partials of r.z and r.r
//   beta   : reduce, convergence test ||r|| <= tol ||b||, beta = rho_new / rho
// Flexible CG (Y. Notay, "Flexible conjugate gradients", SIAM J. Sci. Comput. 22 (2000)
// 1444-1460, with one retained direction) preconditioned by one smoothed-aggregation V-cycle
// (amg_kernels.hpp) per iteration:
//   cycle  : z = B r, partials z.r and z.q_old
//   fbeta  : beta = -(z.q_old) / (p_old.q_old), rho = z.r
//   spmv   : p_new = z + beta p_old, q = A p_new, p_new.q
//   alpha  : alpha = rho / (p_new.q)
//   update : x += alpha p_new, r -= alpha q, ||r||^2
//   check  : convergence; a right-hand side whose residual norm has not dropped by 10% for
//            `stall_limit` iterations is marked stalled and finished by Jacobi PCG from its
//            current iterate (as is any breakdown), to the same tolerance.
// p is double-buffered (ping-pong) because neighbours' p_old is still being read while
// p_new is written.  Converged lane groups skip all work; the host polls one counter of
// active right-hand sides between graph replays (one replay in flight ahead).
#include "krylov.hpp"
#include "amg_kernels.hpp"
#include "amg_setup.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

namespace {

using namespace adt::krylov;


// ------------------------------------------------------------------------------------------
// plain multi-RHS product (adt_spmm)
// ------------------------------------------------------------------------------------------

template <int L>
__global__ __launch_bounds__(256) void spmm_kernel(const Sys* __restrict__ sys, int rp, int rb,
                                                   const int32_t* __restrict__ indptr,
                                                   const int32_t* __restrict__ indices,
                                                   const double* __restrict__ values,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   int mode) {
    const Sys S = sys[blockIdx.y];
    const int lane = threadIdx.x % L, rsub = threadIdx.x / L;
    const int col = blockIdx.z * L + lane;
    const int row0 = blockIdx.x * rb;
    const int row1 = static_cast<int>(std::min<int64_t>(row0 + rb, S.rows));
    const auto ptr = adt::gview(indptr + S.ptr_off);
    const auto idx = adt::gview(indices + S.idx_off);
    const auto val = adt::gview(values + S.val_off);
    const auto X = adt::gview(x + S.vec_off + col);
    const auto Y = adt::gview(y + S.vec_off + col);
    for (int row = row0 + rsub; row < row1; row += kRowsPerPass(L)) {
        double acc = 0.0;
        const int k1 = ptr[row + 1];
        for (int k = ptr[row]; k < k1; ++k) acc = fma(val[k], X[static_cast<int64_t>(idx[k]) * rp], acc);
        const int64_t o = static_cast<int64_t>(row) * rp;
        Y[o] = mode == 0 ? acc : Y[o] - acc;
    }
}

template <int L>
int launch_spmm(int n_sys, int rp, int max_rows, const Sys* d_sys, const int32_t* indptr,
                const int32_t* indices, const double* values, const double* x, double* y, int mode,
                hipStream_t st) {
    const int rb = 1024;
    dim3 grid(adt::blocks_for(max_rows, rb), n_sys, rp / L);
    hipLaunchKernelGGL(spmm_kernel<L>, grid, dim3(256), 0, st, d_sys, rp, rb, indptr, indices, values, x, y,
                       mode);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

// ------------------------------------------------------------------------------------------
// CG kernels
// ------------------------------------------------------------------------------------------

struct TileCtx {
    Tile t;
    Sys S;
    int col;
    int64_t sc;
};

template <int L>
__device__ __forceinline__ bool load_tile(const PcgArgs& a, TileCtx& c) {
    const int tix = adt::xcd_tile(blockIdx.x, a.n_tiles);
    c.t = adt::gview(a.tiles)[tix];
    c.S = adt::gview(a.sys)[c.t.sys];
    const int g = c.t.grp % a.groups;
    c.col = g * L + static_cast<int>(threadIdx.x % L);
    c.sc = static_cast<int64_t>(c.t.sys) * a.rp + c.col;
    return true;
}

__device__ __forceinline__ int tile_index(const PcgArgs& a) { return adt::xcd_tile(blockIdx.x, a.n_tiles); }

template <typename IP, typename VP>
__device__ __forceinline__ double diag_inverse(IP ptr, IP idx, VP val, int row) {
    double d = 0.0;
    for (int k = ptr[row]; k < ptr[row + 1]; ++k)
        if (idx[k] == row) d = val[k];
    return 1.0 / d;
}

// r = b - A x, dinv, p_0 = q = 0; partials ||b||^2, r.z, ||r||^2.
template <int L>
__global__ __launch_bounds__(256) void cg_init(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    TileCtx c;
    load_tile<L>(a, c);
    const int tix = tile_index(a);
    const auto ptr = adt::gview(a.indptr + c.S.ptr_off);
    const auto idx = adt::gview(a.indices + c.S.idx_off);
    const auto val = adt::gview(a.values + c.S.val_off);
    const int rp = a.rp;
    const bool real = c.col < a.r_real;
    const bool first_group = (c.t.grp % a.groups) == 0;
    const auto B = adt::gview(a.b + c.S.b_off + c.col);
    const auto X = adt::gview(a.x + c.S.vec_off + c.col);
    const auto R = adt::gview(a.r + c.S.vec_off + c.col);
    const auto P0 = adt::gview(a.p[0] + c.S.vec_off + c.col);
    const auto P1 = adt::gview(a.p[1] + c.S.vec_off + c.col);
    const auto Q = adt::gview(a.q + c.S.vec_off + c.col);
    const auto D = adt::gview(a.dinv + c.S.diag_off);
    double sb = 0.0, sz = 0.0, sr = 0.0;
    for (int row = c.t.row0 + static_cast<int>(threadIdx.x / L); row < c.t.row1; row += kRowsPerPass(L)) {
        const double di = diag_inverse(ptr, idx, val, row);
        if (first_group && (threadIdx.x % L) == 0) D[row] = di;
        const int64_t o = static_cast<int64_t>(row) * rp;
        P0[o] = 0.0;
        P1[o] = 0.0;
        Q[o] = 0.0;
        if (!real) {
            R[o] = 0.0;
            continue;
        }
        double acc = 0.0;
        for (int k = ptr[row]; k < ptr[row + 1]; ++k) acc = fma(val[k], X[static_cast<int64_t>(idx[k]) * rp], acc);
        const double bi = B[o];
        const double ri = bi - acc;
        R[o] = ri;
        sb = fma(bi, bi, sb);
        sz = fma(ri * ri, di, sz);
        sr = fma(ri, ri, sr);
    }
    double s1 = sb, s2 = sz;
    block_sum2<L>(s1, s2);
    double s3 = block_sum1<L>(sr);
    if (threadIdx.x < L) {
        adt::gview(a.part0)[static_cast<int64_t>(tix) * L + threadIdx.x] = s1;
        adt::gview(a.part1)[static_cast<int64_t>(tix) * L + threadIdx.x] = s2;
        adt::gview(a.part2)[static_cast<int64_t>(tix) * L + threadIdx.x] = s3;
    }
}

// cg_init for staged plans: x of the tile and its halo rows in LDS (as cg_spmv_lds stages p), the
// product from LDS in slot order (the same sums as cg_init's gathers), the diagonal from the
// staged CSR slice (an own row's staged slot is its tile row); same outputs and partials.
// Padding columns read column 0 of b and store r = 0.
template <int L>
__global__ __launch_bounds__(256) void cg_init_lds(const PcgArgs* __restrict__ pa) {
    extern __shared__ double smem[];
    const PcgArgs& a = *pa;
    const int tix = tile_index(a);
    const Tile t = adt::gview(a.tiles)[tix];
    constexpr int RPP = kRowsPerPass(L);
    const Sys S = adt::gview(a.sys)[t.sys];
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % a.groups) * L + lane;
    const bool real = col < a.r_real;
    const bool first_group = (t.grp % a.groups) == 0;
    const int rp = a.rp;
    const auto B = adt::gview(a.b + S.b_off + (real ? col : 0));
    const auto X = adt::gview(a.x + S.vec_off + col);
    const auto R = adt::gview(a.r + S.vec_off + col);
    const auto P0 = adt::gview(a.p[0] + S.vec_off + col);
    const auto P1 = adt::gview(a.p[1] + S.vec_off + col);
    const auto Q = adt::gview(a.q + S.vec_off + col);
    const auto D = adt::gview(a.dinv + S.diag_off);
    const StageSrc src = fine_src(a, S);
    const LdsTile s = lds_tile(src, t, L, smem, true);
    const auto value = [&](int j) { return X[static_cast<int64_t>(j) * rp]; };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double xv[own_max(L)], bv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int64_t o = static_cast<int64_t>(own_row<L>(t, s.rows, i)) * rp;
        xv[i] = X[o];
        bv[i] = B[o];
    }
    const CsrRegs cr = load_csr(src, t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, t, s, cr);
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) s.stage[q * L + lane] = xv[i];
    }
    store_halo<L>(s, 0, hv);
    stage_halo_rest<L>(src, s, value);
    adt::wait_vector_memory();
    __syncthreads();
    double sb = 0.0, sz = 0.0, sr = 0.0;
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q >= s.rows) continue;
        double d = 0.0;
        for (int k = s.sptr[q]; k < s.sptr[q + 1]; ++k)
            if (s.sidx[k] == q) d = s.sval[k];
        const double di = 1.0 / d;
        const int row = t.row0 + q;
        if (first_group && lane == 0) D[row] = di;
        const int64_t o = static_cast<int64_t>(row) * rp;
        P0[o] = 0.0;
        P1[o] = 0.0;
        Q[o] = 0.0;
        const double acc = lds_row_product<L>(s, q, lane);
        const double bi = bv[i];
        const double ri = bi - acc;
        R[o] = real ? ri : 0.0;
        if (real) {
            sb = fma(bi, bi, sb);
            sz = fma(ri * ri, di, sz);
            sr = fma(ri, ri, sr);
        }
    }
    double s1 = sb, s2 = sz;
    block_sum2<L>(s1, s2);
    double s3 = block_sum1<L>(sr);
    if (threadIdx.x < L) {
        adt::gview(a.part0)[static_cast<int64_t>(tix) * L + threadIdx.x] = s1;
        adt::gview(a.part1)[static_cast<int64_t>(tix) * L + threadIdx.x] = s2;
        adt::gview(a.part2)[static_cast<int64_t>(tix) * L + threadIdx.x] = s3;
    }
}

template <int L>
__global__ __launch_bounds__(kReduceThreads) void cg_init_reduce(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    const double bb = group_sum<L>(a, 0, gid);
    const double rz = group_sum<L>(a, 1, gid);
    const double rr = group_sum<L>(a, 2, gid);
    bool act = false;
    if (threadIdx.x < L) {
        const int sys = gid / a.groups, g = gid % a.groups;
        const int col = g * L + threadIdx.x;
        const int64_t sc = static_cast<int64_t>(sys) * a.rp + col;
        int flag = 0;
        if (col < a.r_real) {
            if (!std::isfinite(bb) || !std::isfinite(rr) || !std::isfinite(rz))
                flag = -1;
            else if (bb == 0.0)
                flag = 2;
            else if (rr > a.tol2 * bb)
                flag = 1;
        }
        adt::gview(a.flags)[sc] = flag;
        adt::gview(a.bb)[sc] = bb;
        adt::gview(a.rr)[sc] = rr;
        adt::gview(a.rr_best)[sc] = rr;
        adt::gview(a.rho)[sc] = rz;
        adt::gview(a.rz_old)[sc] = rz;
        adt::gview(a.beta)[sc] = 0.0;
        adt::gview(a.alpha)[sc] = 0.0;
        adt::gview(a.pq_old)[sc] = 1.0;
        adt::gview(a.iters)[sc] = 0;
        adt::gview(a.stall)[sc] = 0;
        act = flag == 1;
        if (act) atomicAdd(a.active_total, 1);
    }
    if (threadIdx.x == 0) adt::gview(a.grp_iters)[gid] = 0;
    refresh_group<L>(a, gid, act);
}

// x = 0 for zero right-hand sides (the solution of A x = 0).
template <int L>
__global__ __launch_bounds__(256) void cg_zero_solution(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    TileCtx c;
    load_tile<L>(a, c);
    if (adt::gview(a.flags)[c.sc] != 2) return;
    const auto X = adt::gview(a.x + c.S.vec_off + c.col);
    for (int row = c.t.row0 + static_cast<int>(threadIdx.x / L); row < c.t.row1; row += kRowsPerPass(L))
        X[static_cast<int64_t>(row) * a.rp] = 0.0;
}

// Search direction of row j: Jacobi p_new = D^-1 r + beta p_old; flexible p_new = z + beta p_old.
template <bool FLEX, typename DP, typename VP>
__device__ __forceinline__ double direction(DP D, VP R, VP Z, VP Pold, double beta, int64_t j, int rp) {
    if (FLEX) return fma(beta, Pold[j * rp], Z[j * rp]);
    return fma(beta, Pold[j * rp], D[j] * R[j * rp]);
}

// Step 1: p_new on the fly, q = A p_new, partial p_new.q.
template <int L, bool FLEX>
__global__ __launch_bounds__(256) void cg_spmv(const PcgArgs* __restrict__ pa, int parity) {
    const PcgArgs& a = *pa;
    const int tix = tile_index(a);
    const Tile t = adt::gview(a.tiles)[tix];
    if (!adt::gview(a.grp_active)[t.grp]) return;
    const Sys S = adt::gview(a.sys)[t.sys];
    const int col = (t.grp % a.groups) * L + static_cast<int>(threadIdx.x % L);
    const int64_t sc = static_cast<int64_t>(t.sys) * a.rp + col;
    const bool act = adt::gview(a.flags)[sc] == 1;
    const double beta = act ? adt::gview(a.beta)[sc] : 0.0;
    const auto ptr = adt::gview(a.indptr + S.ptr_off);
    const auto idx = adt::gview(a.indices + S.idx_off);
    const auto val = adt::gview(a.values + S.val_off);
    const int rp = a.rp;
    const auto R = adt::gview(a.r + S.vec_off + col);
    const auto Z = adt::gview(a.z + S.vec_off + col);
    const auto Pold = adt::gview(a.p[parity] + S.vec_off + col);
    const auto Pnew = adt::gview(a.p[parity ^ 1] + S.vec_off + col);
    const auto Q = adt::gview(a.q + S.vec_off + col);
    const auto D = adt::gview(a.dinv + S.diag_off);
    double pq = 0.0;
    if (act) {
        constexpr int RPP = kRowsPerPass(L);
        for (int r0 = t.row0 + static_cast<int>(threadIdx.x / L); r0 < t.row1; r0 += kGatherRows * RPP) {
            int k0[kGatherRows], k1[kGatherRows];
            double acc[kGatherRows], pv[kGatherRows];
            batch_ranges(ptr, r0, RPP, t.row1, k0, k1);
#pragma unroll
            for (int i = 0; i < kGatherRows; ++i) {
                const int rs = r0 + i * RPP < t.row1 ? r0 + i * RPP : r0;  // rows past row1: unused
                pv[i] = direction<FLEX>(D, R, Z, Pold, beta, rs, rp);
                acc[i] = 0.0;
            }
            gather_rows<kGatherRows, kGatherSlots>(
                k0, k1, [&](int k) { return idx[k]; }, [&](int k) { return val[k]; },
                [&](int j) { return direction<FLEX>(D, R, Z, Pold, beta, j, rp); }, acc);
#pragma unroll
            for (int i = 0; i < kGatherRows; ++i) {
                const int row = r0 + i * RPP;
                if (row < t.row1) {
                    const int64_t o = static_cast<int64_t>(row) * rp;
                    Pnew[o] = pv[i];
                    Q[o] = acc[i];
                    pq = fma(pv[i], acc[i], pq);
                }
            }
        }
    }
    pq = block_sum1<L>(pq);
    if (threadIdx.x < L) adt::gview(a.part0)[static_cast<int64_t>(tix) * L + threadIdx.x] = pq;
}

// LDS-staged variant of cg_spmv, one barrier per tile: the tile's CSR slice is copied into
// LDS while each thread computes p_new for its own rows (kept in registers, written out) and
// its share of the halo rows (halo_batch(L) rows with loads in flight together), all staged in
// LDS; after the barrier q = A p_new is formed from LDS only (slot-major with L lanes
// contiguous: conflict-free).  Every vector row is read from memory once per tile.
template <int L, bool FLEX>
__global__ __launch_bounds__(256) void cg_spmv_lds(const PcgArgs* __restrict__ pa, int parity) {
    extern __shared__ double smem[];
    const PcgArgs& a = *pa;
    const int tix = tile_index(a);
    const Tile t = adt::gview(a.tiles)[tix];
    if (!adt::gview(a.grp_active)[t.grp]) return;
    constexpr int RPP = kRowsPerPass(L);
    const Sys S = adt::gview(a.sys)[t.sys];
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % a.groups) * L + lane;
    const int64_t sc = static_cast<int64_t>(t.sys) * a.rp + col;
    const bool act = adt::gview(a.flags)[sc] == 1;
    const double beta_sc = adt::gview(a.beta)[sc];
    const double beta = act ? beta_sc : 0.0;
    const int rp = a.rp;
    const auto R = adt::gview(a.r + S.vec_off + col);
    const auto Z = adt::gview(a.z + S.vec_off + col);
    const auto Pold = adt::gview(a.p[parity] + S.vec_off + col);
    const auto Pnew = adt::gview(a.p[parity ^ 1] + S.vec_off + col);
    const auto Q = adt::gview(a.q + S.vec_off + col);
    const auto D = adt::gview(a.dinv + S.diag_off);
    const StageSrc src = fine_src(a, S);
    const LdsTile s = lds_tile(src, t, L, smem, true);
    // every lane stages (inactive lanes' values are never read); only active lanes write p_new
    const auto value = [&](int j) { return direction<FLEX>(D, R, Z, Pold, beta, j, rp); };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double v[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) v[i] = value(own_row<L>(t, s.rows, i));
    const CsrRegs cr = load_csr(src, t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, t, s, cr);
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) s.stage[q * L + lane] = v[i];
    }
    store_halo<L>(s, 0, hv);
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (act && q < s.rows) Pnew[static_cast<int64_t>(t.row0 + q) * rp] = v[i];
    }
    stage_halo_rest<L>(src, s, value);
    __syncthreads();
    double pq = 0.0;
    if (act) {
#pragma unroll
        for (int i = 0; i < own_max(L); ++i) {
            const int q = rsub + i * RPP;
            if (q < s.rows) {
                const double acc = lds_row_product<L>(s, q, lane);
                Q[static_cast<int64_t>(t.row0 + q) * rp] = acc;
                pq = fma(v[i], acc, pq);
            }
        }
    }
    pq = block_sum1<L>(pq);
    if (threadIdx.x < L) adt::gview(a.part0)[static_cast<int64_t>(tix) * L + threadIdx.x] = pq;
}

template <int L>
__global__ __launch_bounds__(kReduceThreads) void cg_alpha(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    if (!adt::gview(a.grp_active)[gid]) return;
    const double pq = group_sum<L>(a, 0, gid);
    bool act = false;
    if (threadIdx.x < L) {
        const int64_t sc = static_cast<int64_t>(gid / a.groups) * a.rp + (gid % a.groups) * L + threadIdx.x;
        if (adt::gview(a.flags)[sc] == 1) {
            const double alpha = adt::gview(a.rho)[sc] / pq;
            if (!(pq > 0.0) || !std::isfinite(alpha)) {
                adt::gview(a.flags)[sc] = -1;
                adt::gview(a.alpha)[sc] = 0.0;
                atomicSub(a.active_total, 1);
            } else {
                adt::gview(a.alpha)[sc] = alpha;
                adt::gview(a.pq_old)[sc] = pq;
                act = true;
            }
        }
    }
    refresh_group<L>(a, gid, act);
}

// x += alpha p, r -= alpha q; partials r.D^-1 r (Jacobi) and r.r.
template <int L>
__global__ __launch_bounds__(256) void cg_update(const PcgArgs* __restrict__ pa, int parity) {
    const PcgArgs& a = *pa;
    const int tix = tile_index(a);
    const Tile t = adt::gview(a.tiles)[tix];
    if (!adt::gview(a.grp_active)[t.grp]) return;
    const Sys S = adt::gview(a.sys)[t.sys];
    const int col = (t.grp % a.groups) * L + static_cast<int>(threadIdx.x % L);
    const int64_t sc = static_cast<int64_t>(t.sys) * a.rp + col;
    const bool act = adt::gview(a.flags)[sc] == 1;
    const int rp = a.rp;
    const auto P = adt::gview(a.p[parity ^ 1] + S.vec_off + col);
    const auto Q = adt::gview(a.q + S.vec_off + col);
    const auto X = adt::gview(a.x + S.vec_off + col);
    const auto R = adt::gview(a.r + S.vec_off + col);
    const auto D = adt::gview(a.dinv + S.diag_off);
    double sz = 0.0, sr = 0.0;
    if (act) {
        const double alpha = adt::gview(a.alpha)[sc];
        for (int row = t.row0 + static_cast<int>(threadIdx.x / L); row < t.row1; row += kRowsPerPass(L)) {
            const int64_t o = static_cast<int64_t>(row) * rp;
            X[o] = fma(alpha, P[o], X[o]);
            const double ri = fma(-alpha, Q[o], R[o]);
            R[o] = ri;
            sz = fma(ri * ri, D[row], sz);
            sr = fma(ri, ri, sr);
        }
    }
    block_sum2<L>(sz, sr);
    if (threadIdx.x < L) {
        adt::gview(a.part1)[static_cast<int64_t>(tix) * L + threadIdx.x] = sz;
        adt::gview(a.part2)[static_cast<int64_t>(tix) * L + threadIdx.x] = sr;
    }
}

// Flexible CG with the multigrid cycle: the update x += alpha p, r -= alpha q (partial r.r for
// the convergence check) fused with the next cycle's finest pre-smoothing r1 = r - A (w D^-1 r)
// of the updated residual, staged like amg_smooth_residual_lds.  Halo rows are updated on the
// fly with the arithmetic of their own tiles, so every staged value equals the residual that
// its own tile writes, and the residual is read once per iteration instead of twice.  The
// updated residual goes to the other residual buffer (neighbours still read this one).
// Inactive lanes keep x and write no residual.  Three waves per SIMD: x and p are loaded after
// the barrier, so the staging phase holds only r, q and the diagonal of the own rows and the
// halo loads (168 VGPRs with one 8-byte spill).  On the 3D box a launch averaged 0.473 ms at two
// waves per SIMD and 0.420 ms at three; on the 2.5D time-lapse batch 1.210 and 1.187 ms.
template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void fcg_update_smooth(const PcgArgs* __restrict__ pa,
                                                         const LevelArgs* __restrict__ la, int parity) {
    extern __shared__ double smem[];
    const PcgArgs& g = *pa;
    const LevelArgs& a = *la;
    const int tix = tile_index(g);
    const Tile t = adt::gview(g.tiles)[tix];
    if (!adt::gview(g.grp_active)[t.grp]) return;
    constexpr int RPP = kRowsPerPass(L);
    const Sys S = adt::gview(g.sys)[t.sys];
    const double w = adt::gview(a.ls)[t.sys].omega;
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % g.groups) * L + lane;
    const int64_t sc = static_cast<int64_t>(t.sys) * g.rp + col;
    const bool act = adt::gview(g.flags)[sc] == 1;
    const double alpha = adt::gview(g.alpha)[sc];
    const int rp = g.rp;
    const auto P = adt::gview(g.p[parity ^ 1] + S.vec_off + col);
    const auto Q = adt::gview(g.q + S.vec_off + col);
    const auto X = adt::gview(g.x + S.vec_off + col);
    const auto R = adt::gview(residual(g, parity) + S.vec_off + col);        // this iteration's residual
    const auto Rn = adt::gview(residual(g, parity ^ 1) + S.vec_off + col);   // the next one
    const auto R1 = adt::gview(g.z + S.vec_off + col);
    const auto D = adt::gview(g.dinv + S.diag_off);
    const StageSrc src = fine_src(g, S);
    const LdsTile s = lds_tile(src, t, L, smem, true);
    // inactive columns keep r (alpha taken as 0): a select between the loaded operands would be
    // compiled into a branch around the halo loads, with a full wait for memory in each
    const double alpha_act = act ? alpha : 0.0;
    const auto updated = [&](double r, double q) { return fma(-alpha_act, q, r); };
    const auto value = [&](int j) {
        const int64_t o = static_cast<int64_t>(j) * rp;
        return w * D[j] * updated(R[o], Q[o]);
    };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double rv[own_max(L)], qv[own_max(L)], dv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int row = own_row<L>(t, s.rows, i);
        const int64_t o = static_cast<int64_t>(row) * rp;
        rv[i] = R[o];
        qv[i] = Q[o];
        dv[i] = D[row];
    }
    const CsrRegs cr = load_csr(src, t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, t, s, cr);
    double sr = 0.0;
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        rv[i] = updated(rv[i], qv[i]);
        if (q < s.rows) {
            s.stage[q * L + lane] = w * dv[i] * rv[i];
            if (act) {
                Rn[static_cast<int64_t>(t.row0 + q) * rp] = rv[i];
                sr = fma(rv[i], rv[i], sr);
            }
        }
    }
    store_halo<L>(s, 0, hv);
    stage_halo_rest<L>(src, s, value);
    // Rn stores complete here once, instead of the product loop of every row waiting for them
    // (and for the R1 stores of earlier rows)
    adt::wait_vector_memory();
    __syncthreads();
    // x += alpha p: its operands are loaded after the barrier, in flight during the products, so
    // that they do not hold registers through the staging phase
    double xv[own_max(L)], pv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int64_t o = static_cast<int64_t>(own_row<L>(t, s.rows, i)) * rp;
        xv[i] = X[o];
        pv[i] = P[o];
    }
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) R1[static_cast<int64_t>(t.row0 + q) * rp] = rv[i] - lds_row_product<L>(s, q, lane);
    }
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (act && q < s.rows) X[static_cast<int64_t>(t.row0 + q) * rp] = fma(alpha, pv[i], xv[i]);
    }
    sr = block_sum1<L>(sr);
    if (threadIdx.x < L) adt::gview(g.part2)[static_cast<int64_t>(tix) * L + threadIdx.x] = sr;
}

// Jacobi: convergence test and beta = rho_new / rho.
template <int L>
__global__ __launch_bounds__(kReduceThreads) void cg_beta(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    if (!adt::gview(a.grp_active)[gid]) return;
    if (threadIdx.x == 0) adt::gview(a.grp_iters)[gid] += 1;
    const double rz = group_sum<L>(a, 1, gid);
    const double rr = group_sum<L>(a, 2, gid);
    bool act = false;
    if (threadIdx.x < L) {
        const int64_t sc = static_cast<int64_t>(gid / a.groups) * a.rp + (gid % a.groups) * L + threadIdx.x;
        if (adt::gview(a.flags)[sc] == 1) {
            adt::gview(a.iters)[sc] += 1;
            adt::gview(a.rr)[sc] = rr;
            if (!std::isfinite(rr) || !std::isfinite(rz)) {
                adt::gview(a.flags)[sc] = -1;
                atomicSub(a.active_total, 1);
            } else if (rr <= a.tol2 * adt::gview(a.bb)[sc]) {
                adt::gview(a.flags)[sc] = 0;
                atomicSub(a.active_total, 1);
            } else {
                adt::gview(a.beta)[sc] = rz / adt::gview(a.rho)[sc];
                adt::gview(a.rho)[sc] = rz;
                act = true;
            }
        }
    }
    refresh_group<L>(a, gid, act);
}

// Flexible CG: convergence test and stall detection (coefficients come from fcg_beta).
template <int L>
__global__ __launch_bounds__(kReduceThreads) void cg_check(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    if (!adt::gview(a.grp_active)[gid]) return;
    if (threadIdx.x == 0) adt::gview(a.grp_iters)[gid] += 1;
    const double rr = group_sum<L>(a, 2, gid);
    bool act = false;
    if (threadIdx.x < L) {
        const int64_t sc = static_cast<int64_t>(gid / a.groups) * a.rp + (gid % a.groups) * L + threadIdx.x;
        if (adt::gview(a.flags)[sc] == 1) {
            adt::gview(a.iters)[sc] += 1;
            adt::gview(a.rr)[sc] = rr;
            if (!std::isfinite(rr)) {
                adt::gview(a.flags)[sc] = -1;
                atomicSub(a.active_total, 1);
            } else if (rr <= a.tol2 * adt::gview(a.bb)[sc]) {
                adt::gview(a.flags)[sc] = 0;
                atomicSub(a.active_total, 1);
            } else {
                if (rr < 0.81 * adt::gview(a.rr_best)[sc]) {
                    adt::gview(a.rr_best)[sc] = rr;
                    adt::gview(a.stall)[sc] = 0;
                } else {
                    adt::gview(a.stall)[sc] += 1;
                }
                if (a.stall_limit > 0 && adt::gview(a.stall)[sc] >= a.stall_limit) {
                    adt::gview(a.flags)[sc] = -2;
                    atomicSub(a.active_total, 1);
                } else {
                    act = true;
                }
            }
        }
    }
    refresh_group<L>(a, gid, act);
}

// Flexible CG coefficients from the cycle's partials: beta = -(z.q_old)/(p_old.q_old), rho = z.r.
template <int L>
__global__ __launch_bounds__(kReduceThreads) void fcg_beta(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    if (!adt::gview(a.grp_active)[gid]) return;
    const double zq = group_sum<L>(a, 0, gid);
    const double zr = group_sum<L>(a, 1, gid);
    bool act = false;
    if (threadIdx.x < L) {
        const int64_t sc = static_cast<int64_t>(gid / a.groups) * a.rp + (gid % a.groups) * L + threadIdx.x;
        if (adt::gview(a.flags)[sc] == 1) {
            const bool first = adt::gview(a.iters)[sc] == 0;
            const double beta = first ? 0.0 : -zq / adt::gview(a.pq_old)[sc];
            if (!(zr > 0.0) || !std::isfinite(zr) || !std::isfinite(beta)) {
                adt::gview(a.flags)[sc] = -2;  // preconditioner not positive on this residual: finish with Jacobi
                atomicSub(a.active_total, 1);
            } else {
                adt::gview(a.beta)[sc] = beta;
                adt::gview(a.rho)[sc] = zr;
                act = true;
            }
        }
    }
    refresh_group<L>(a, gid, act);
}

// True residual ||b - A x||^2 partials (reported after the solve).
template <int L>
__global__ __launch_bounds__(256) void cg_true_residual(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    TileCtx c;
    load_tile<L>(a, c);
    const int tix = tile_index(a);
    const auto ptr = adt::gview(a.indptr + c.S.ptr_off);
    const auto idx = adt::gview(a.indices + c.S.idx_off);
    const auto val = adt::gview(a.values + c.S.val_off);
    const auto B = adt::gview(a.b + c.S.b_off + c.col);
    const auto X = adt::gview(a.x + c.S.vec_off + c.col);
    double sr = 0.0;
    if (c.col < a.r_real) {
        // rows and slots batched as in cg_init; products in slot order, rows in order
        constexpr int RB = kGatherRows, RPP = kRowsPerPass(L);
        const int rows = c.t.row1 - c.t.row0;
        for (int q0 = static_cast<int>(threadIdx.x / L); q0 < rows; q0 += RB * RPP) {
            int k0[RB], k1[RB];
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const int q = q0 + b * RPP;
                k0[b] = q < rows ? ptr[c.t.row0 + q] : 0;
                k1[b] = q < rows ? ptr[c.t.row0 + q + 1] : 0;
            }
            double acc[RB] = {};
            gather_rows<RB, kGatherSlots>(
                k0, k1, [&](int e) { return idx[e]; }, [&](int e) { return val[e]; },
                [&](int j) { return X[static_cast<int64_t>(j) * a.rp]; }, acc);
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                const int q = q0 + b * RPP;
                if (q >= rows) continue;
                const double ri = B[static_cast<int64_t>(c.t.row0 + q) * a.rp] - acc[b];
                sr = fma(ri, ri, sr);
            }
        }
    }
    sr = block_sum1<L>(sr);
    if (threadIdx.x < L) adt::gview(a.part2)[static_cast<int64_t>(tix) * L + threadIdx.x] = sr;
}

// cg_true_residual for staged plans: x of the tile and its halo in LDS, products in slot order.
template <int L>
__global__ __launch_bounds__(256) void cg_true_residual_lds(const PcgArgs* __restrict__ pa) {
    extern __shared__ double smem[];
    const PcgArgs& a = *pa;
    const int tix = tile_index(a);
    const Tile t = adt::gview(a.tiles)[tix];
    constexpr int RPP = kRowsPerPass(L);
    const Sys S = adt::gview(a.sys)[t.sys];
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % a.groups) * L + lane;
    const bool real = col < a.r_real;
    const int rp = a.rp;
    const auto B = adt::gview(a.b + S.b_off + (real ? col : 0));
    const auto X = adt::gview(a.x + S.vec_off + col);
    const StageSrc src = fine_src(a, S);
    const LdsTile s = lds_tile(src, t, L, smem, true);
    const auto value = [&](int j) { return X[static_cast<int64_t>(j) * rp]; };
    int hid[halo_batch(L)];
    halo_ids<L>(src, s, 0, hid);
    double xv[own_max(L)], bv[own_max(L)];
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int64_t o = static_cast<int64_t>(own_row<L>(t, s.rows, i)) * rp;
        xv[i] = X[o];
        bv[i] = B[o];
    }
    const CsrRegs cr = load_csr(src, t, s);
    double hv[halo_batch(L)];
    halo_values<L>(hid, hv, value);
    store_csr(src, t, s, cr);
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) s.stage[q * L + lane] = xv[i];
    }
    store_halo<L>(s, 0, hv);
    stage_halo_rest<L>(src, s, value);
    adt::wait_vector_memory();
    __syncthreads();
    double sr = 0.0;
#pragma unroll
    for (int i = 0; i < own_max(L); ++i) {
        const int q = rsub + i * RPP;
        if (q < s.rows) {
            const double ri = bv[i] - lds_row_product<L>(s, q, lane);
            sr = real ? fma(ri, ri, sr) : sr;
        }
    }
    sr = block_sum1<L>(sr);
    if (threadIdx.x < L) adt::gview(a.part2)[static_cast<int64_t>(tix) * L + threadIdx.x] = sr;
}

template <int L>
__global__ __launch_bounds__(kReduceThreads) void cg_true_reduce(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    const double rr = group_sum<L>(a, 2, gid);
    if (threadIdx.x < L) {
        const int64_t sc = static_cast<int64_t>(gid / a.groups) * a.rp + (gid % a.groups) * L + threadIdx.x;
        adt::gview(a.rr)[sc] = rr;
    }
}

// Segment pre-reduction: workgroup per segment (run of <= kSegTiles tiles of one lane group)
// sums the selected partial arrays (mask bits 0..2) into the segment partials, fixed order.
// Bit 3 also reduces segments of inactive groups (before the flags exist, and after the solve).
constexpr int kSegTiles = 64;

template <int L>
__global__ __launch_bounds__(256) void reduce_segments(const PcgArgs* __restrict__ pa, int mask) {
    const PcgArgs& a = *pa;
    const int seg = blockIdx.x;
    if (!(mask & 8) && !adt::gview(a.grp_active)[adt::gview(a.seg_grp)[seg]]) return;
    const int t0 = adt::gview(a.seg_tile0)[seg], nt = adt::gview(a.seg_ntiles)[seg];
    const int lane = static_cast<int>(threadIdx.x % L), chunk = static_cast<int>(threadIdx.x / L);
    for (int k = 0; k < 3; ++k) {
        if (!((mask >> k) & 1)) continue;
        const auto part = adt::gview(k == 0 ? a.part0 : (k == 1 ? a.part1 : a.part2));
        double v = 0.0;
        for (int i = chunk; i < nt; i += kRowsPerPass(L)) v += part[static_cast<int64_t>(t0 + i) * L + lane];
        v = block_sum1<L>(v);
        if (threadIdx.x < L) adt::gview(a.spart)[(static_cast<int64_t>(k) * a.n_segs + seg) * L + threadIdx.x] = v;
    }
}

// One multigrid level (all systems) on the device.
struct AmgLevel {
    std::vector<LvSys> ls_h;
    std::vector<Tile> tiles_h;
    LvSys* ls_d = nullptr;
    Tile* tiles_d = nullptr;
    int n_tiles = 0, max_rows = 0, max_p = 0, max_t = 0, max_next = 0;
    LevelArgs args_h{};
    LevelArgs* args_d = nullptr;
    int64_t dense_len = 0;
    int smooth_lds = 0;   // dynamic LDS of the staged smoothing kernels (0: gather kernels)
    int prolong_lds = 0;  // dynamic LDS of the staged prolongation (0: gather kernel)
    int restrict_lds = 0; // dynamic LDS of the staged restriction to the next level (0: gather kernel)
    int tile_rows = 0;    // rows per tile on this level
};

}  // namespace

// ------------------------------------------------------------------------------------------
// plan
// ------------------------------------------------------------------------------------------

constexpr int kMaxChunk = 64;
constexpr int kEvents = 6;  // per profiled iteration: cycle, spmv and update brackets

struct adt_pcg_plan {
    int n_sys = 0, r = 0, L = 0, rp = 0, rb = 0, groups = 0, n_tiles = 0, n_groups = 0;
    int64_t vec_len = 0, diag_len = 0;
    std::vector<Sys> sys_h;
    std::vector<int64_t> rows_h;
    std::vector<int> pattern_of_sys;
    Sys* sys_d = nullptr;
    Tile* tiles_d = nullptr;
    int* grp_tile0_d = nullptr;
    int* grp_ntiles_d = nullptr;
    int* seg_d = nullptr;  // seg_tile0, seg_ntiles, seg_grp, grp_seg0, grp_nseg
    double* spart_d = nullptr;
    int n_segs = 0;
    double* vec = nullptr;  // r, p0, p1, q, z
    double* dinv = nullptr;
    double* part = nullptr;
    double* scal = nullptr;
    int* ints = nullptr;
    int* h_active = nullptr;
    PcgArgs args_h{};
    PcgArgs* args_d = nullptr;
    hipStream_t cap = nullptr;
    // per mode (0 Jacobi, 1 flexible/multigrid): two identical iteration graphs replayed
    // alternately, so a replay in flight never touches the profiling events of the replay being read
    hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int graph_iters[2] = {0, 0};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int profile = 0;
    int lds_rows = 0;  // staged rows of the LDS spmv (0: gather kernel)
    int lds_nnz = 0;
    uint16_t* lidx_d = nullptr;
    int* halo_ptr_d = nullptr;
    int* halo_rows_d = nullptr;
    hipEvent_t pev[2][kMaxChunk][kEvents] = {};
    double prof_ms[3] = {0.0, 0.0, 0.0};
    int64_t prof_launches = 0;
    // multigrid hierarchy (empty: none attached)
    std::vector<AmgLevel> amg;
    std::vector<void*> amg_buffers;
    int stall_limit = 0;
    int last_fallback = 0;
    int last_polished = 0;
    int last_fcg_iters_max = 0;
    ~adt_pcg_plan();
    void free_amg();
};

void adt_pcg_plan::free_amg() {
    for (void* b : amg_buffers)
        if (b) (void)hipFree(b);
    amg_buffers.clear();
    amg.clear();
    args_h.r_alt = nullptr;
    for (auto& g : graph[1])
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    graph_iters[1] = 0;
}

adt_pcg_plan::~adt_pcg_plan() {
    for (auto& set : graph)
        for (auto& g : set)
            if (g) (void)hipGraphExecDestroy(g);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& set : pev)
        for (auto& it : set)
            for (auto& e : it)
                if (e) (void)hipEventDestroy(e);
    if (cap) (void)hipStreamDestroy(cap);
    void* bufs[] = {sys_d, tiles_d, grp_tile0_d, grp_ntiles_d, vec, dinv, part, scal, ints, args_d, lidx_d, halo_ptr_d,
                    halo_rows_d, seg_d, spart_d};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (h_active) (void)hipHostFree(h_active);
    for (void* b : amg_buffers)
        if (b) (void)hipFree(b);
}

namespace {

template <typename T>
int dalloc(T** p, int64_t count) {
    if (count <= 0) count = 1;
    if (hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * static_cast<size_t>(count)) != hipSuccess)
        return adt::set_error(ADT_ERR_NOMEM, "hipMalloc of %lld bytes failed",
                              static_cast<long long>(sizeof(T) * count));
    return ADT_OK;
}

int64_t fine_lds_bytes(const adt_pcg_plan* p) {
    return lds_bytes_for(p->lds_rows, p->lds_nnz, p->rb, p->L);
}

// Flexible CG with the multigrid cycle fuses each update with the next finest pre-smoothing
// when the finest level is staged in LDS.
bool fused_update(const adt_pcg_plan* p, int mode) {
    return mode == 1 && p->lds_rows > 0 && !p->amg.empty() && p->args_h.r_alt;
}

// Restriction from level l to level l + 1 (tiles of the next level).
template <int L>
void launch_restrict(const adt_pcg_plan* p, int l, hipStream_t st) {
    const auto& A = p->amg[l];
    const int tiles = p->amg[l + 1].n_tiles;
    if (A.restrict_lds > 0)
        hipLaunchKernelGGL(amg_restrict_st<L>, dim3(tiles), dim3(256), A.restrict_lds, st, A.args_d);
    else
        hipLaunchKernelGGL(amg_restrict<L>, dim3(tiles), dim3(256), 0, st, A.args_d);
}

// One V-cycle: z = B r (flexible-CG partials z.r, z.q_old written by the finest post-smoothing).
// Staged (LDS) kernels wherever the plan built their tables, gather kernels elsewhere.
// With `presmoothed`, the finest pre-smoothing was done by the previous update
// (fcg_update_smooth, or the standalone pass after the initial residual).
template <int L>
void launch_cycle(const adt_pcg_plan* p, hipStream_t st, int parity, bool presmoothed) {
    const int m = static_cast<int>(p->amg.size());
    const PcgArgs* pa = p->args_d;
    const auto& f = p->amg[0];
    const size_t lds = p->lds_rows > 0 ? static_cast<size_t>(fine_lds_bytes(p)) : 0;
    if (!presmoothed && p->lds_rows > 0)
        hipLaunchKernelGGL(amg_smooth_residual_lds<L>, dim3(p->n_tiles), dim3(256), lds, st, pa, f.args_d);
    else if (!presmoothed)
        hipLaunchKernelGGL(amg_smooth_residual<L>, dim3(p->n_tiles), dim3(256), 0, st, f.args_d);
    launch_restrict<L>(p, 0, st);
    for (int l = 1; l < m - 1; ++l) {
        const auto& A = p->amg[l];
        if (A.smooth_lds > 0)
            hipLaunchKernelGGL(amg_smooth_residual_st<L>, dim3(A.n_tiles), dim3(256), A.smooth_lds, st, A.args_d);
        else
            hipLaunchKernelGGL(amg_smooth_residual<L>, dim3(A.n_tiles), dim3(256), 0, st, A.args_d);
        launch_restrict<L>(p, l, st);
    }
    hipLaunchKernelGGL(amg_coarse_solve<L>, dim3(p->n_groups, adt::blocks_for(p->amg[m - 1].max_rows, 256 / L)),
                       dim3(256), 0, st, p->amg[m - 1].args_d);
    for (int l = m - 2; l >= 1; --l) {
        const auto& A = p->amg[l];
        if (A.prolong_lds > 0)
            hipLaunchKernelGGL((amg_prolong_st<L, false>), dim3(A.n_tiles), dim3(256), A.prolong_lds, st, pa, A.args_d,
                               parity);
        else
            hipLaunchKernelGGL((amg_prolong<L, false>), dim3(A.n_tiles), dim3(256), 0, st, pa, A.args_d, parity);
        if (A.smooth_lds > 0)
            hipLaunchKernelGGL(amg_post_smooth_st<L>, dim3(A.n_tiles), dim3(256), A.smooth_lds, st, A.args_d);
        else
            hipLaunchKernelGGL(amg_post_smooth<L>, dim3(A.n_tiles), dim3(256), 0, st, A.args_d);
    }
    if (f.prolong_lds > 0)
        hipLaunchKernelGGL((amg_prolong_st<L, true>), dim3(p->n_tiles), dim3(256), f.prolong_lds, st, pa, f.args_d,
                           parity);
    else
        hipLaunchKernelGGL((amg_prolong<L, true>), dim3(p->n_tiles), dim3(256), 0, st, pa, f.args_d, parity);
    if (p->lds_rows > 0)
        hipLaunchKernelGGL(amg_post_smooth_lds<L>, dim3(p->n_tiles), dim3(256), lds, st, pa, f.args_d, parity);
    else
        hipLaunchKernelGGL(amg_post_smooth_fine<L>, dim3(p->n_tiles), dim3(256), 0, st, pa, f.args_d, parity);
}

// Reduction kernel for lane groups, preceded by the segment pre-reduction in segmented plans.
template <int L>
void launch_reduce(const adt_pcg_plan* p, hipStream_t st, void (*kernel)(const PcgArgs*), int mask) {
    if (p->n_segs > 0)
        hipLaunchKernelGGL(reduce_segments<L>, dim3(p->n_segs), dim3(256), 0, st, p->args_d, mask);
    hipLaunchKernelGGL(kernel, dim3(p->n_groups), dim3(kReduceThreads), 0, st, p->args_d);
}

template <int L>
void launch_iteration(const adt_pcg_plan* p, hipStream_t st, int parity, int mode, hipEvent_t* events) {
    const PcgArgs* pa = p->args_d;
    const bool fused = fused_update(p, mode);
    if (mode == 1) {
        if (events) (void)hipEventRecord(events[4], st);
        launch_cycle<L>(p, st, parity, fused);
        if (events) (void)hipEventRecord(events[5], st);
        launch_reduce<L>(p, st, fcg_beta<L>, 3);
    }
    if (events) (void)hipEventRecord(events[0], st);
    if (p->lds_rows > 0) {
        const size_t lds = static_cast<size_t>(fine_lds_bytes(p));
        if (mode == 1)
            hipLaunchKernelGGL((cg_spmv_lds<L, true>), dim3(p->n_tiles), dim3(256), lds, st, pa, parity);
        else
            hipLaunchKernelGGL((cg_spmv_lds<L, false>), dim3(p->n_tiles), dim3(256), lds, st, pa, parity);
    } else {
        if (mode == 1)
            hipLaunchKernelGGL((cg_spmv<L, true>), dim3(p->n_tiles), dim3(256), 0, st, pa, parity);
        else
            hipLaunchKernelGGL((cg_spmv<L, false>), dim3(p->n_tiles), dim3(256), 0, st, pa, parity);
    }
    if (events) (void)hipEventRecord(events[1], st);
    launch_reduce<L>(p, st, cg_alpha<L>, 1);
    if (events) (void)hipEventRecord(events[2], st);
    if (fused)
        hipLaunchKernelGGL(fcg_update_smooth<L>, dim3(p->n_tiles), dim3(256), fine_lds_bytes(p), st, pa,
                           p->amg[0].args_d, parity);
    else
        hipLaunchKernelGGL(cg_update<L>, dim3(p->n_tiles), dim3(256), 0, st, pa, parity);
    if (events) (void)hipEventRecord(events[3], st);
    if (mode == 1)
        launch_reduce<L>(p, st, cg_check<L>, 4);
    else
        launch_reduce<L>(p, st, cg_beta<L>, 6);
}

// Capture `chunk` (even) iterations; replays keep the p ping-pong parity at 0.
template <int L>
int ensure_graphs(adt_pcg_plan* p, int chunk, int mode) {
    if (p->graph[mode][0] && p->graph_iters[mode] == chunk) return ADT_OK;
    for (auto& g : p->graph[mode])
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    for (int set = 0; set < 2; ++set) {
        hipGraph_t g = nullptr;
        ADT_HIP_CHECK(hipStreamBeginCapture(p->cap, hipStreamCaptureModeRelaxed));
        for (int i = 0; i < chunk; ++i) launch_iteration<L>(p, p->cap, i & 1, mode, nullptr);
        ADT_HIP_CHECK(hipStreamEndCapture(p->cap, &g));
        hipError_t e = hipGraphInstantiate(&p->graph[mode][set], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return adt::set_error(ADT_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    p->graph_iters[mode] = chunk;
    return ADT_OK;
}

// Profiling mode: the same kernels launched directly on the caller's stream with timing events
// around every cycle, spmv and update launch (HIP event nodes inside graphs cannot be timed).
template <int L>
int launch_profiled_chunk(adt_pcg_plan* p, hipStream_t st, int chunk, int set, int mode) {
    for (int i = 0; i < chunk; ++i) {
        for (auto& e : p->pev[set][i])
            if (!e) ADT_HIP_CHECK(hipEventCreate(&e));
        if (mode == 0) {  // keep the cycle bracket empty
            ADT_HIP_CHECK(hipEventRecord(p->pev[set][i][4], st));
            ADT_HIP_CHECK(hipEventRecord(p->pev[set][i][5], st));
        }
        launch_iteration<L>(p, st, i & 1, mode, p->pev[set][i]);
    }
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

int accumulate_profile(adt_pcg_plan* p, int set, int chunk) {
    for (int i = 0; i < chunk; ++i) {
        float a = 0.f, b = 0.f, c = 0.f;
        ADT_HIP_CHECK(hipEventElapsedTime(&a, p->pev[set][i][0], p->pev[set][i][1]));
        ADT_HIP_CHECK(hipEventElapsedTime(&b, p->pev[set][i][2], p->pev[set][i][3]));
        ADT_HIP_CHECK(hipEventElapsedTime(&c, p->pev[set][i][4], p->pev[set][i][5]));
        p->prof_ms[0] += a;
        p->prof_ms[1] += b;
        p->prof_ms[2] += c;
        p->prof_launches += 1;
    }
    return ADT_OK;
}

// Numeric multigrid setup for the current matrix values: P, A P, Galerkin products per level
// and the dense inverse of the coarsest operator, for every system.
int amg_numeric(adt_pcg_plan* p, hipStream_t st) {
    const int m = static_cast<int>(p->amg.size());
    for (int l = 0; l < m - 1; ++l) {
        const AmgLevel& A = p->amg[l];
        hipLaunchKernelGGL(amg_pvals, dim3(adt::blocks_for(A.max_rows, 256), p->n_sys), dim3(256), 0, st, A.args_d);
        if (A.max_t > 0)
            hipLaunchKernelGGL(amg_tvals, dim3(adt::blocks_for(A.max_t, 256), p->n_sys), dim3(256), 0, st, A.args_d);
        if (A.max_next > 0)
            hipLaunchKernelGGL(amg_cvals, dim3(adt::blocks_for(A.max_next, 256), p->n_sys), dim3(256), 0, st,
                               A.args_d);
    }
    const AmgLevel& C = p->amg[m - 1];
    hipLaunchKernelGGL(amg_pvals, dim3(adt::blocks_for(C.max_rows, 256), p->n_sys), dim3(256), 0, st, C.args_d);
    const bool lds_matrix = C.max_rows <= kDenseLdsRows;
    const size_t lds = (2 + (lds_matrix ? static_cast<size_t>(C.max_rows) : 0)) * C.max_rows * 8;
    hipLaunchKernelGGL(amg_dense_inverse, dim3(p->n_sys), dim3(256), lds, st, C.args_d, lds_matrix ? 1 : 0);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

template <int L>
int solve_impl(adt_pcg_plan* p, hipStream_t st, int max_iter, int chunk, int mode) {
    const PcgArgs* pa = p->args_d;
    ADT_HIP_CHECK(hipMemsetAsync(p->args_h.active_total, 0, sizeof(int), st));
    if (p->lds_rows > 0)
        hipLaunchKernelGGL(cg_init_lds<L>, dim3(p->n_tiles), dim3(256), fine_lds_bytes(p), st, pa);
    else
        hipLaunchKernelGGL(cg_init<L>, dim3(p->n_tiles), dim3(256), 0, st, pa);
    launch_reduce<L>(p, st, cg_init_reduce<L>, 15);
    hipLaunchKernelGGL(cg_zero_solution<L>, dim3(p->n_tiles), dim3(256), 0, st, pa);
    if (fused_update(p, mode))  // the first cycle's finest pre-smoothing (later ones: fcg_update_smooth)
        hipLaunchKernelGGL(amg_smooth_residual_lds<L>, dim3(p->n_tiles), dim3(256), fine_lds_bytes(p), st, pa,
                           p->amg[0].args_d);
    ADT_HIP_CHECK(hipGetLastError());
    if (max_iter > 0) {
        int rc = p->profile ? ADT_OK : ensure_graphs<L>(p, chunk, mode);
        if (rc) return rc;
        int k = 0;
        bool stop = false;
        for (int it = 0; it < max_iter && !stop; it += chunk, ++k) {
            const int set = k & 1;
            if (p->profile) {
                if ((rc = launch_profiled_chunk<L>(p, st, chunk, set, mode))) return rc;
            } else {
                ADT_HIP_CHECK(hipGraphLaunch(p->graph[mode][set], st));
            }
            ADT_HIP_CHECK(hipMemcpyAsync(&p->h_active[set], p->args_h.active_total, sizeof(int),
                                         hipMemcpyDeviceToHost, st));
            ADT_HIP_CHECK(hipEventRecord(p->ev[set], st));
            if (k >= 1) {
                ADT_HIP_CHECK(hipEventSynchronize(p->ev[set ^ 1]));
                if (p->profile && (rc = accumulate_profile(p, set ^ 1, chunk))) return rc;
                stop = p->h_active[set ^ 1] == 0;
            }
        }
        // the last replay launched is still in flight (or never synchronized)
        const int last = (k - 1) & 1;
        ADT_HIP_CHECK(hipEventSynchronize(p->ev[last]));
        if (p->profile && (rc = accumulate_profile(p, last, chunk))) return rc;
    }
    if (p->lds_rows > 0)
        hipLaunchKernelGGL(cg_true_residual_lds<L>, dim3(p->n_tiles), dim3(256), fine_lds_bytes(p), st, pa);
    else
        hipLaunchKernelGGL(cg_true_residual<L>, dim3(p->n_tiles), dim3(256), 0, st, pa);
    launch_reduce<L>(p, st, cg_true_reduce<L>, 12);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

int solve_dispatch(adt_pcg_plan* p, hipStream_t st, int max_iter, int chunk, int mode) {
    switch (p->L) {
        case 8: return solve_impl<8>(p, st, max_iter, chunk, mode);
        case 16: return solve_impl<16>(p, st, max_iter, chunk, mode);
        case 32: return solve_impl<32>(p, st, max_iter, chunk, mode);
        default: return solve_impl<64>(p, st, max_iter, chunk, mode);
    }
}

}  // namespace

namespace {

// Halo tables of one CSR pattern cut into row tiles of rb rows.
struct PatternTiles {
    std::vector<uint16_t> lidx;
    std::vector<int> halo_ptr;  // one entry per tile (start), closed by the caller
    std::vector<int> halo_rows;
    int max_staged = 0, max_nnz = 0;
};

PatternTiles build_pattern_tiles(const int32_t* ptr, const int32_t* idx, int rows, int rb) {
    PatternTiles pt;
    pt.lidx.resize(static_cast<size_t>(ptr[rows]));
    std::vector<int> halo;
    for (int r0 = 0; r0 < rows; r0 += rb) {
        const int r1 = std::min(r0 + rb, rows);
        halo.clear();
        for (int k = ptr[r0]; k < ptr[r1]; ++k)
            if (idx[k] < r0 || idx[k] >= r1) halo.push_back(idx[k]);
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        const int own = r1 - r0;
        pt.max_staged = std::max(pt.max_staged, own + static_cast<int>(halo.size()));
        pt.max_nnz = std::max(pt.max_nnz, ptr[r1] - ptr[r0]);
        for (int k = ptr[r0]; k < ptr[r1]; ++k) {
            const int c = idx[k];
            const int slot = (c >= r0 && c < r1)
                                 ? c - r0
                                 : own + static_cast<int>(std::lower_bound(halo.begin(), halo.end(), c) - halo.begin());
            pt.lidx[k] = static_cast<uint16_t>(slot);
        }
        pt.halo_ptr.push_back(static_cast<int>(pt.halo_rows.size()));
        pt.halo_rows.insert(pt.halo_rows.end(), halo.begin(), halo.end());
    }
    return pt;
}

// Prolongation staging tables of one pattern's level: per tile of rb rows the sorted coarse rows
// that its P rows reference, and per P slot the position of its column in that list.
struct ProlongTiles {
    std::vector<uint16_t> plidx;
    std::vector<int> pc_ptr;  // one entry per tile (start), closed by the caller
    std::vector<int> pc_rows;
    int max_staged = 0, max_nnz = 0;
};

ProlongTiles build_prolong_tiles(const std::vector<int>& pptr, const std::vector<int>& pidx, int rows, int rb) {
    ProlongTiles pt;
    pt.plidx.resize(pidx.size());
    std::vector<int> cols;
    for (int r0 = 0; r0 < rows; r0 += rb) {
        const int r1 = std::min(r0 + rb, rows);
        cols.assign(pidx.begin() + pptr[r0], pidx.begin() + pptr[r1]);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        pt.max_staged = std::max(pt.max_staged, static_cast<int>(cols.size()));
        pt.max_nnz = std::max(pt.max_nnz, pptr[r1] - pptr[r0]);
        for (int k = pptr[r0]; k < pptr[r1]; ++k)
            pt.plidx[k] = static_cast<uint16_t>(std::lower_bound(cols.begin(), cols.end(), pidx[k]) - cols.begin());
        pt.pc_ptr.push_back(static_cast<int>(pt.pc_rows.size()));
        pt.pc_rows.insert(pt.pc_rows.end(), cols.begin(), cols.end());
    }
    return pt;
}

typedef std::tuple<int64_t, int64_t, int> PatKey;

}  // namespace

extern "C" int adt_pcg_create(adt_pcg_plan** out, int n_sys, int r, int lanes, int rows_per_tile,
                              const int32_t* h_rows, const int64_t* h_ptr_off, const int64_t* h_idx_off,
                              const int64_t* h_val_off, const int64_t* h_b_off, const int32_t* h_indptr,
                              const int32_t* h_indices, int lds_bytes) {
    ADT_REQUIRE(out && h_rows && h_ptr_off && h_idx_off && h_val_off, "adt_pcg_create: null pointer");
    ADT_REQUIRE(n_sys >= 1 && r >= 1, "adt_pcg_create: need at least one system and one right-hand side");
    ADT_REQUIRE(n_sys <= 65535, "adt_pcg_create: at most 65535 systems per plan");
    ADT_REQUIRE(lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64, "adt_pcg_create: lanes must be 8/16/32/64");
    ADT_REQUIRE(rows_per_tile >= 1, "adt_pcg_create: rows_per_tile must be positive");
    auto* p = new adt_pcg_plan();
    p->n_sys = n_sys;
    p->r = r;
    p->L = lanes;
    p->rp = (r + lanes - 1) / lanes * lanes;
    p->groups = p->rp / lanes;
    const int rpp = 256 / lanes;
    p->rb = (rows_per_tile + rpp - 1) / rpp * rpp;
    for (int s = 0; s < n_sys; ++s)
        if (h_rows[s] < 1) {
            delete p;
            return adt::set_error(ADT_ERR_INVALID, "adt_pcg_create: system %d has no rows", s);
        }
    // Halo tables per distinct pattern; shrink the tile height until the staged rows fit in LDS.
    std::map<PatKey, int> pattern_of;
    std::vector<PatKey> patterns;
    for (int s = 0; s < n_sys; ++s) {
        PatKey key{h_ptr_off[s], h_idx_off[s], h_rows[s]};
        if (!pattern_of.count(key)) {
            pattern_of[key] = static_cast<int>(patterns.size());
            patterns.push_back(key);
        }
        p->pattern_of_sys.push_back(pattern_of[key]);
    }
    std::vector<PatternTiles> ptiles;
    if (h_indptr && h_indices && lds_bytes > 0) {
        int device = 0, limit = 0;
        if (hipGetDevice(&device) == hipSuccess &&
            hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
            limit > 0)
            lds_bytes = std::min(lds_bytes, limit - 8192);  // the block reductions keep static LDS
        // staged tiles keep their own rows' operands in registers: at most own_max(L) rows per thread
        for (int rb = std::min(p->rb, max_staged_tile_rows(lanes)); rb >= rpp; rb = (rb / 2 + rpp - 1) / rpp * rpp) {
            ptiles.clear();
            int max_staged = 0, max_nnz = 0;
            for (const auto& key : patterns) {
                ptiles.push_back(build_pattern_tiles(h_indptr + std::get<0>(key), h_indices + std::get<1>(key),
                                                     std::get<2>(key), rb));
                max_staged = std::max(max_staged, ptiles.back().max_staged);
                max_nnz = std::max(max_nnz, ptiles.back().max_nnz);
            }
            if (lds_bytes_for(max_staged, max_nnz, rb, lanes) <= lds_bytes && max_staged < 65536) {
                p->rb = rb;
                p->lds_rows = max_staged;
                p->lds_nnz = max_nnz;
                break;
            }
            if (rb == rpp) break;
        }
        if (!p->lds_rows) ptiles.clear();
    }
    // pattern tiles (row ranges of one pattern) are numbered pattern by pattern; per-pattern-tile
    // tables (halo lists here, prolongation lists of the multigrid levels) use that numbering
    std::vector<int64_t> lidx_base(patterns.size(), 0), ptile_base(patterns.size(), 0);
    for (size_t q = 1; q < patterns.size(); ++q)
        ptile_base[q] = ptile_base[q - 1] + (std::get<2>(patterns[q - 1]) + p->rb - 1) / p->rb;
    std::vector<uint16_t> lidx_all;
    std::vector<int> halo_ptr_all, halo_rows_all;
    for (size_t q = 0; q < ptiles.size(); ++q) {
        lidx_base[q] = static_cast<int64_t>(lidx_all.size());
        lidx_all.insert(lidx_all.end(), ptiles[q].lidx.begin(), ptiles[q].lidx.end());
        const int shift = static_cast<int>(halo_rows_all.size());
        for (int h : ptiles[q].halo_ptr) halo_ptr_all.push_back(h + shift);
        halo_rows_all.insert(halo_rows_all.end(), ptiles[q].halo_rows.begin(), ptiles[q].halo_rows.end());
    }
    halo_ptr_all.push_back(static_cast<int>(halo_rows_all.size()));
    std::vector<Tile> tiles;
    std::vector<int> t0, nt;
    int64_t vo = 0, dvo = 0;
    for (int s = 0; s < n_sys; ++s) {
        const int q = p->pattern_of_sys[s];
        Sys S{h_rows[s], h_ptr_off[s], h_idx_off[s], h_val_off[s], vo, dvo, h_b_off ? h_b_off[s] : vo,
              p->lds_rows ? lidx_base[q] : 0};
        p->sys_h.push_back(S);
        vo += static_cast<int64_t>(h_rows[s]) * p->rp;
        dvo += h_rows[s];
        for (int g = 0; g < p->groups; ++g) {
            t0.push_back(static_cast<int>(tiles.size()));
            int count = 0;
            for (int r0 = 0; r0 < h_rows[s]; r0 += p->rb, ++count)
                tiles.push_back(Tile{s, s * p->groups + g, r0, std::min(r0 + p->rb, h_rows[s]),
                                     static_cast<int>(ptile_base[q]) + count, 0});
            nt.push_back(count);
        }
    }
    p->vec_len = vo;
    p->diag_len = dvo;
    p->n_tiles = static_cast<int>(tiles.size());
    p->n_groups = static_cast<int>(t0.size());
    const int64_t nsc = static_cast<int64_t>(n_sys) * p->rp;
    int rc = ADT_OK;
    if ((rc = dalloc(&p->sys_d, n_sys)) || (rc = dalloc(&p->tiles_d, p->n_tiles)) ||
        (rc = dalloc(&p->grp_tile0_d, p->n_groups)) || (rc = dalloc(&p->grp_ntiles_d, p->n_groups)) ||
        (rc = dalloc(&p->vec, 5 * vo)) || (rc = dalloc(&p->dinv, dvo)) ||
        (rc = dalloc(&p->part, 3 * static_cast<int64_t>(p->n_tiles) * lanes)) ||
        (rc = dalloc(&p->scal, 9 * nsc)) || (rc = dalloc(&p->ints, 3 * nsc + 2 * p->n_groups + 1)) ||
        (rc = dalloc(&p->args_d, 1)) || (rc = dalloc(&p->lidx_d, static_cast<int64_t>(lidx_all.size()))) ||
        (rc = dalloc(&p->halo_ptr_d, static_cast<int64_t>(halo_ptr_all.size()))) ||
        (rc = dalloc(&p->halo_rows_d, static_cast<int64_t>(halo_rows_all.size())))) {
        delete p;
        return rc;
    }
    if ((!lidx_all.empty() &&
         hipMemcpy(p->lidx_d, lidx_all.data(), sizeof(uint16_t) * lidx_all.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(p->halo_ptr_d, halo_ptr_all.data(), sizeof(int) * halo_ptr_all.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (!halo_rows_all.empty() && hipMemcpy(p->halo_rows_d, halo_rows_all.data(), sizeof(int) * halo_rows_all.size(),
                                             hipMemcpyHostToDevice) != hipSuccess)) {
        delete p;
        return adt::set_error(ADT_ERR_HIP, "adt_pcg_create: halo table upload failed");
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&p->h_active), 2 * sizeof(int)) != hipSuccess ||
        hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev[1], hipEventDisableTiming) != hipSuccess) {
        delete p;
        return adt::set_error(ADT_ERR_HIP, "adt_pcg_create: host/stream/event allocation failed");
    }
    if (hipMemcpy(p->sys_d, p->sys_h.data(), sizeof(Sys) * n_sys, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->tiles_d, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->grp_tile0_d, t0.data(), sizeof(int) * t0.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->grp_ntiles_d, nt.data(), sizeof(int) * nt.size(), hipMemcpyHostToDevice) != hipSuccess) {
        delete p;
        return adt::set_error(ADT_ERR_HIP, "adt_pcg_create: metadata upload failed");
    }
    // groups of many tiles: segment partials pre-reduced by many workgroups
    const int max_nt = *std::max_element(nt.begin(), nt.end());
    if (max_nt > 512) {
        std::vector<int> st0, snt, sgrp, gs0(p->n_groups), gns(p->n_groups);
        for (int g = 0; g < p->n_groups; ++g) {
            gs0[g] = static_cast<int>(st0.size());
            for (int off = 0; off < nt[g]; off += kSegTiles) {
                st0.push_back(t0[g] + off);
                snt.push_back(std::min(kSegTiles, nt[g] - off));
                sgrp.push_back(g);
            }
            gns[g] = static_cast<int>(st0.size()) - gs0[g];
        }
        p->n_segs = static_cast<int>(st0.size());
        std::vector<int> meta;
        meta.insert(meta.end(), st0.begin(), st0.end());
        meta.insert(meta.end(), snt.begin(), snt.end());
        meta.insert(meta.end(), sgrp.begin(), sgrp.end());
        meta.insert(meta.end(), gs0.begin(), gs0.end());
        meta.insert(meta.end(), gns.begin(), gns.end());
        if ((rc = dalloc(&p->seg_d, static_cast<int64_t>(meta.size()))) ||
            (rc = dalloc(&p->spart_d, 3 * static_cast<int64_t>(p->n_segs) * lanes))) {
            delete p;
            return rc;
        }
        if (hipMemcpy(p->seg_d, meta.data(), sizeof(int) * meta.size(), hipMemcpyHostToDevice) != hipSuccess) {
            delete p;
            return adt::set_error(ADT_ERR_HIP, "adt_pcg_create: segment table upload failed");
        }
    }
    PcgArgs& a = p->args_h;
    if (p->n_segs > 0) {
        a.seg_tile0 = p->seg_d;
        a.seg_ntiles = p->seg_d + p->n_segs;
        a.seg_grp = p->seg_d + 2 * p->n_segs;
        a.grp_seg0 = p->seg_d + 3 * p->n_segs;
        a.grp_nseg = p->seg_d + 3 * p->n_segs + p->n_groups;
        a.spart = p->spart_d;
        a.segmented = 1;
        a.n_segs = p->n_segs;
    }
    a.lidx = p->lidx_d;
    a.halo_ptr = p->halo_ptr_d;
    a.halo_rows = p->halo_rows_d;
    a.sys = p->sys_d;
    a.tiles = p->tiles_d;
    a.grp_tile0 = p->grp_tile0_d;
    a.grp_ntiles = p->grp_ntiles_d;
    a.r = p->vec;
    a.p[0] = p->vec + vo;
    a.p[1] = p->vec + 2 * vo;
    a.q = p->vec + 3 * vo;
    a.z = p->vec + 4 * vo;
    a.dinv = p->dinv;
    const int64_t np = static_cast<int64_t>(p->n_tiles) * lanes;
    a.part0 = p->part;
    a.part1 = p->part + np;
    a.part2 = p->part + 2 * np;
    a.alpha = p->scal;
    a.beta = p->scal + nsc;
    a.rho = p->scal + 2 * nsc;
    a.bb = p->scal + 3 * nsc;
    a.rr = p->scal + 4 * nsc;
    a.rz_old = p->scal + 5 * nsc;
    a.pq_old = p->scal + 6 * nsc;
    a.rr_best = p->scal + 7 * nsc;
    a.iters = p->ints;
    a.flags = p->ints + nsc;
    a.stall = p->ints + 2 * nsc;
    a.grp_active = p->ints + 3 * nsc;
    a.grp_iters = p->ints + 3 * nsc + p->n_groups;
    a.active_total = p->ints + 3 * nsc + 2 * p->n_groups;
    a.rp = p->rp;
    a.lds_rows = p->lds_rows;
    a.lds_nnz = p->lds_nnz;
    a.lds_own = p->rb;
    a.groups = p->groups;
    a.r_real = r;
    a.n_tiles = p->n_tiles;
    a.stall_limit = 0;
    *out = p;
    return ADT_OK;
}

// ------------------------------------------------------------------------------------------
// multigrid hierarchy
// ------------------------------------------------------------------------------------------

namespace {

// Device copy of a host vector, owned by the plan's multigrid buffer list.
template <typename T>
int upload(adt_pcg_plan* p, const std::vector<T>& v, T** out) {
    T* d = nullptr;
    int rc = dalloc(&d, static_cast<int64_t>(v.size()));
    if (rc) return rc;
    p->amg_buffers.push_back(d);
    if (!v.empty() && hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
        return adt::set_error(ADT_ERR_HIP, "multigrid upload failed");
    *out = d;
    return ADT_OK;
}

template <typename T>
int zeros(adt_pcg_plan* p, int64_t count, T** out) {
    T* d = nullptr;
    int rc = dalloc(&d, count);
    if (rc) return rc;
    p->amg_buffers.push_back(d);
    if (hipMemset(d, 0, sizeof(T) * static_cast<size_t>(std::max<int64_t>(count, 1))) != hipSuccess)
        return adt::set_error(ADT_ERR_HIP, "multigrid buffer clear failed");
    *out = d;
    return ADT_OK;
}

template <typename T>
void append(std::vector<T>& dst, const std::vector<T>& src) {
    dst.insert(dst.end(), src.begin(), src.end());
}

int attach_amg(adt_pcg_plan* p, const int32_t* h_indptr, const int32_t* h_indices, const double* h_values,
               const int64_t* h_val_off, const adt::amg::Options& opt) {
    using adt::amg::HostLevel;
    const int n_pat = 1 + *std::max_element(p->pattern_of_sys.begin(), p->pattern_of_sys.end());
    std::vector<int> rep_sys(n_pat, -1);
    for (int s = 0; s < p->n_sys; ++s)
        if (rep_sys[p->pattern_of_sys[s]] < 0) rep_sys[p->pattern_of_sys[s]] = s;
    // hierarchies per pattern, all of the same depth
    std::vector<std::vector<HostLevel>> H(n_pat);
    int depth = 0;
    for (int q = 0; q < n_pat; ++q) {
        const Sys& S = p->sys_h[rep_sys[q]];
        H[q] = adt::amg::build_hierarchy(static_cast<int>(S.rows), h_indptr + S.ptr_off, h_indices + S.idx_off,
                                         h_values + h_val_off[rep_sys[q]], opt);
        if (H[q].size() < 2)
            return adt::set_error(ADT_ERR_INVALID,
                                  "multigrid setup: pattern %d has no coarse level (matrix too small, not SPD or "
                                  "without strong connections)", q);
        if (H[q].back().n > std::max(opt.coarse_max, 2048))
            return adt::set_error(ADT_ERR_INVALID,
                                  "multigrid setup: coarsening of pattern %d stalled at %d rows", q, H[q].back().n);
        depth = std::max(depth, static_cast<int>(H[q].size()));
    }
    for (int q = 0; q < n_pat; ++q)
        if (static_cast<int>(H[q].size()) < depth) {
            adt::amg::Options o2 = opt;
            o2.coarse_max = 0;
            o2.max_levels = depth;
            const Sys& S = p->sys_h[rep_sys[q]];
            H[q] = adt::amg::build_hierarchy(static_cast<int>(S.rows), h_indptr + S.ptr_off, h_indices + S.idx_off,
                                             h_values + h_val_off[rep_sys[q]], o2);
            if (static_cast<int>(H[q].size()) != depth)
                return adt::set_error(ADT_ERR_INVALID, "multigrid setup: patterns coarsen to different depths");
        }
    p->free_amg();
    p->amg.resize(depth);
    const int rp = p->rp, L = p->L;
    const int rbc = std::min(std::max(256 / L, 32), max_staged_tile_rows(L));  // coarse-level tile height
    int lds_limit = 0, device = 0;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        lds_limit = 0;
    // the block reductions keep static LDS.  Up to 96 KB per workgroup (one per CU): the coarse
    // operators of ten-node tetrahedra have ~70 entries per row, and their 16-row tiles stage ~450
    // rows; at a 64 KB cap they fell back to the gather kernels (3D P2 terrain cycle 4.05 -> 1.86 ms)
    lds_limit = std::max(0, std::min(lds_limit - 8192, 98304));
    // tile height per level: the finest level's tiles; coarse levels start at rbc rows and halve
    // (down to one pass of the workgroup) until their staged smoothing tiles fit the LDS budget
    for (int l = 0; l < depth; ++l) {
        if (l == 0) {
            p->amg[0].tile_rows = p->rb;
            continue;
        }
        int th = rbc;
        for (; th > 256 / L; th = std::max(256 / L, th / 2)) {
            int staged = 0, nnz = 0;
            for (int q = 0; q < n_pat; ++q) {
                const HostLevel& h = H[q][l];
                const PatternTiles pt = build_pattern_tiles(h.ptr.data(), h.idx.data(), h.n, th);
                staged = std::max(staged, pt.max_staged);
                nnz = std::max(nnz, pt.max_nnz);
            }
            if (lds_bytes_for(staged, nnz, th, L) <= lds_limit) break;
        }
        p->amg[l].tile_rows = th;
    }
    // pattern-level offsets, per level
    std::vector<std::vector<int64_t>> off_ptr(depth, std::vector<int64_t>(n_pat)), off_idx = off_ptr,
        off_diag = off_ptr, off_pptr = off_ptr, off_pidx = off_ptr, off_pmp = off_ptr, off_pmap = off_ptr,
        off_rptr = off_ptr, off_rent = off_ptr, off_tpp = off_ptr, off_tp = off_ptr, off_cpp = off_ptr,
        off_cp = off_ptr, off_dmp = off_ptr, off_dmap = off_ptr;
    for (int l = 0; l < depth; ++l) {
        AmgLevel& A = p->amg[l];
        std::vector<int> ptr, idx, diag, pptr, pidx, pmap_ptr, pmap, pself, rptr, rrow, rslot, tpp, ta, tq, cpp, cp,
            ct, dmp, dmap;
        std::vector<double> pfix;  // a given prolongator's values (single-pattern plans only)
        for (int q = 0; q < n_pat; ++q) {
            const HostLevel& h = H[q][l];
            off_ptr[l][q] = static_cast<int64_t>(ptr.size());
            off_idx[l][q] = static_cast<int64_t>(idx.size());
            off_diag[l][q] = static_cast<int64_t>(diag.size());
            off_pptr[l][q] = static_cast<int64_t>(pptr.size());
            off_pidx[l][q] = static_cast<int64_t>(pidx.size());
            off_pmp[l][q] = static_cast<int64_t>(pmap_ptr.size());
            off_pmap[l][q] = static_cast<int64_t>(pmap.size());
            off_rptr[l][q] = static_cast<int64_t>(rptr.size());
            off_rent[l][q] = static_cast<int64_t>(rrow.size());
            off_tpp[l][q] = static_cast<int64_t>(tpp.size());
            off_tp[l][q] = static_cast<int64_t>(ta.size());
            off_cpp[l][q] = static_cast<int64_t>(cpp.size());
            off_cp[l][q] = static_cast<int64_t>(cp.size());
            off_dmp[l][q] = static_cast<int64_t>(dmp.size());
            off_dmap[l][q] = static_cast<int64_t>(dmap.size());
            append(dmp, h.dmap_ptr);
            append(dmap, h.dmap);
            if (l > 0) {
                append(ptr, h.ptr);
                append(idx, h.idx);
            }
            append(diag, h.diag);
            append(pptr, h.pptr);
            append(pidx, h.pidx);
            pfix.insert(pfix.end(), h.pfix.begin(), h.pfix.end());
            append(pself, h.pself);
            append(pmap_ptr, h.pmap_ptr);
            append(pmap, h.pmap);
            append(rptr, h.rptr);
            append(rrow, h.rrow);
            append(rslot, h.rslot);
            append(tpp, h.tpair_ptr);
            append(ta, h.tpair_a);
            append(tq, h.tpair_p);
            append(cpp, h.cpair_ptr);
            append(cp, h.cpair_p);
            append(ct, h.cpair_t);
        }
        LevelArgs& a = A.args_h;
        int32_t* d32 = nullptr;
        int* di = nullptr;
        int rc = ADT_OK;
        if (l > 0) {
            if ((rc = upload(p, ptr, &d32))) return rc;
            a.ptr = d32;
            if ((rc = upload(p, idx, &d32))) return rc;
            a.idx = d32;
        }
#define ADT_UP(vec, field)                         \
    if ((rc = upload(p, vec, &di))) return rc;     \
    a.field = di;
        ADT_UP(diag, diag)
        ADT_UP(pptr, pptr)
        ADT_UP(pidx, pidx)
        ADT_UP(pself, pself)
        ADT_UP(pmap_ptr, pmap_ptr)
        ADT_UP(pmap, pmap)
        ADT_UP(rptr, rptr)
        ADT_UP(rrow, rrow)
        ADT_UP(rslot, rslot)
        ADT_UP(tpp, tpair_ptr)
        ADT_UP(ta, tpair_a)
        ADT_UP(tq, tpair_p)
        ADT_UP(cpp, cpair_ptr)
        ADT_UP(cp, cpair_p)
        ADT_UP(ct, cpair_t)
        ADT_UP(dmp, dmap_ptr)
        ADT_UP(dmap, dmap)
#undef ADT_UP
        if (!pfix.empty()) {
            double* dp = nullptr;
            if ((rc = upload(p, pfix, &dp))) return rc;
            a.pfix = dp;
        }
        // LDS staging tables: smoothing on levels >= 1, prolongation on every level but the
        // coarsest; per pattern tile (tiles of th rows, numbered pattern by pattern), kept only
        // when the staged tile fits the LDS budget
        if (l + 1 < depth) {  // restriction: slices of the next level's tiles
            const int thn = p->amg[l + 1].tile_rows;
            int r_nnz = 0;
            for (int q = 0; q < n_pat; ++q) {
                const HostLevel& h = H[q][l];
                for (int r0 = 0; r0 < h.nc; r0 += thn)
                    r_nnz = std::max(r_nnz, h.rptr[std::min(r0 + thn, h.nc)] - h.rptr[r0]);
            }
            const int64_t r_lds = static_cast<int64_t>(r_nnz) * 12 + (thn + 1) * 4;
            if (r_lds <= lds_limit) {
                a.r_lds_nnz = r_nnz;
                A.restrict_lds = static_cast<int>(r_lds);
            }
        }
        const int th = A.tile_rows;
        if (th > max_staged_tile_rows(L)) continue;
        std::vector<uint16_t> lidx_l, plidx_l;
        std::vector<int> hp_l, hr_l, pcp_l, pcr_l;
        int s_rows = 0, s_nnz = 0, p_rows = 0, p_nnz = 0;
        for (int q = 0; q < n_pat; ++q) {
            const HostLevel& h = H[q][l];
            if (l > 0) {
                PatternTiles pt = build_pattern_tiles(h.ptr.data(), h.idx.data(), h.n, th);
                const int shift = static_cast<int>(hr_l.size());
                for (int v : pt.halo_ptr) hp_l.push_back(v + shift);
                append(hr_l, pt.halo_rows);
                lidx_l.insert(lidx_l.end(), pt.lidx.begin(), pt.lidx.end());
                s_rows = std::max(s_rows, pt.max_staged);
                s_nnz = std::max(s_nnz, pt.max_nnz);
            }
            if (l + 1 < depth) {
                ProlongTiles pr = build_prolong_tiles(h.pptr, h.pidx, h.n, th);
                const int shift = static_cast<int>(pcr_l.size());
                for (int v : pr.pc_ptr) pcp_l.push_back(v + shift);
                append(pcr_l, pr.pc_rows);
                plidx_l.insert(plidx_l.end(), pr.plidx.begin(), pr.plidx.end());
                p_rows = std::max(p_rows, pr.max_staged);
                p_nnz = std::max(p_nnz, pr.max_nnz);
            }
        }
        hp_l.push_back(static_cast<int>(hr_l.size()));
        pcp_l.push_back(static_cast<int>(pcr_l.size()));
        const int64_t s_lds = lds_bytes_for(s_rows, s_nnz, th, L), p_lds = lds_bytes_for(p_rows, p_nnz, th, L);
        if (l > 0 && s_rows < 65536 && s_lds <= lds_limit) {
            uint16_t* d16 = nullptr;
            if ((rc = upload(p, lidx_l, &d16)) || (rc = upload(p, hp_l, &di))) return rc;
            a.lidx = d16;
            a.halo_ptr = di;
            if ((rc = upload(p, hr_l, &di))) return rc;
            a.halo_rows = di;
            a.lds_rows = s_rows;
            a.lds_nnz = s_nnz;
            a.lds_own = th;
            A.smooth_lds = static_cast<int>(s_lds);
        }
        if (l + 1 < depth && p_rows < 65536 && p_lds <= lds_limit) {
            uint16_t* d16 = nullptr;
            if ((rc = upload(p, plidx_l, &d16)) || (rc = upload(p, pcp_l, &di))) return rc;
            a.plidx = d16;
            a.pc_ptr = di;
            if ((rc = upload(p, pcr_l, &di))) return rc;
            a.pc_rows = di;
            a.p_lds_rows = p_rows;
            a.p_lds_nnz = p_nnz;
            a.p_lds_own = th;
            A.prolong_lds = static_cast<int>(p_lds);
        }
    }
    // per-system records, tiles and numeric buffers
    std::vector<int64_t> val_len(depth, 0), vec_len(depth, 0), dinv_len(depth, 0), pv_len(depth, 0),
        tv_len(depth, 0);
    for (int l = 0; l < depth; ++l) {
        AmgLevel& A = p->amg[l];
        for (int s = 0; s < p->n_sys; ++s) {
            const int q = p->pattern_of_sys[s];
            const HostLevel& h = H[q][l];
            const Sys& F = p->sys_h[s];
            LvSys S{};
            S.rows = h.n;
            S.nc = h.nc;
            S.n_p = static_cast<int>(h.pidx.size());
            S.n_t = static_cast<int>(h.tidx.size());
            S.n_next = l + 1 < depth ? static_cast<int>(H[q][l + 1].idx.size()) : 0;
            S.omega = h.omega;
            S.omega_p = h.omega_p;
            S.diag = off_diag[l][q];
            S.dmap_ptr = off_dmp[l][q];
            S.dmap = off_dmap[l][q];
            S.pptr = off_pptr[l][q];
            S.pidx = off_pidx[l][q];
            S.pmap_ptr = off_pmp[l][q];
            S.pmap = off_pmap[l][q];
            S.rptr = off_rptr[l][q];
            S.rent = off_rent[l][q];
            S.tpair_ptr = off_tpp[l][q];
            S.tpair = off_tp[l][q];
            S.cpair_ptr = off_cpp[l][q];
            S.cpair = off_cp[l][q];
            if (l == 0) {
                S.ptr = F.ptr_off;
                S.idx = F.idx_off;
                S.val = F.val_off;
                S.vec = F.vec_off;
                S.dinv = F.diag_off;
            } else {
                S.ptr = off_ptr[l][q];
                S.idx = off_idx[l][q];
                S.val = val_len[l];
                S.vec = vec_len[l];
                S.dinv = dinv_len[l];
                val_len[l] += static_cast<int64_t>(h.idx.size());
                vec_len[l] += static_cast<int64_t>(h.n) * rp;
                dinv_len[l] += h.n;
            }
            S.pval = pv_len[l];
            S.tval = tv_len[l];
            pv_len[l] += S.n_p;
            tv_len[l] += S.n_t;
            if (l + 1 == depth) {
                S.dense = A.dense_len;
                A.dense_len += static_cast<int64_t>(h.n) * h.n;
            }
            A.ls_h.push_back(S);
            A.max_rows = std::max(A.max_rows, h.n);
            A.max_p = std::max(A.max_p, S.n_p);
            A.max_t = std::max(A.max_t, S.n_t);
            A.max_next = std::max(A.max_next, S.n_next);
            if (l > 0) {
                const int th = A.tile_rows;
                int base = 0;  // pattern tiles of this level before pattern q
                for (int q2 = 0; q2 < q; ++q2) base += (H[q2][l].n + th - 1) / th;
                for (int g = 0; g < p->groups; ++g)
                    for (int r0 = 0; r0 < h.n; r0 += th)
                        A.tiles_h.push_back(Tile{s, s * p->groups + g, r0, std::min(r0 + th, h.n), base + r0 / th, 0});
            }
        }
    }
    // next-level offsets (value and vector blocks of the same system one level down)
    for (int l = 0; l + 1 < depth; ++l)
        for (int s = 0; s < p->n_sys; ++s) {
            p->amg[l].ls_h[s].next_val = p->amg[l + 1].ls_h[s].val;
            p->amg[l].ls_h[s].next_vec = p->amg[l + 1].ls_h[s].vec;
        }
    std::vector<double*> val_d(depth, nullptr), dinv_d(depth, nullptr), b_d(depth, nullptr), r1_d(depth, nullptr),
        e_d(depth, nullptr), z2_d(depth, nullptr);
    for (int l = 0; l < depth; ++l) {
        AmgLevel& A = p->amg[l];
        LevelArgs& a = A.args_h;
        int rc = ADT_OK;
        if (l == 0) {
            A.n_tiles = p->n_tiles;
            A.tiles_d = p->tiles_d;
        } else {
            A.n_tiles = static_cast<int>(A.tiles_h.size());
            if ((rc = upload(p, A.tiles_h, &A.tiles_d)) || (rc = zeros(p, val_len[l], &val_d[l])) ||
                (rc = zeros(p, dinv_len[l], &dinv_d[l])) || (rc = zeros(p, vec_len[l], &b_d[l])) ||
                (rc = zeros(p, vec_len[l], &r1_d[l])) || (rc = zeros(p, vec_len[l], &e_d[l])) ||
                (rc = zeros(p, vec_len[l], &z2_d[l])))
                return rc;
        }
        double* pv = nullptr;
        double* tv = nullptr;
        if ((rc = upload(p, A.ls_h, &A.ls_d)) || (rc = zeros(p, pv_len[l], &pv)) || (rc = zeros(p, tv_len[l], &tv)))
            return rc;
        a.ls = A.ls_d;
        a.tiles = A.tiles_d;
        a.grp_active = p->args_h.grp_active;
        a.n_tiles = A.n_tiles;
        a.rp = rp;
        a.groups = p->groups;
        a.n_sys = p->n_sys;
        a.coarsest = l + 1 == depth;
        a.fine = l == 0;
        a.pval = pv;
        a.tval = tv;
        if (a.coarsest) {
            double* dn = nullptr;
            if ((rc = zeros(p, A.dense_len, &dn))) return rc;
            a.dense = dn;
        }
    }
    for (int l = 0; l < depth; ++l) {
        LevelArgs& a = p->amg[l].args_h;
        if (l == 0) {
            a.dinv = p->args_h.dinv;
            a.dinv_out = nullptr;
            a.b = p->args_h.r;
            a.r1 = p->args_h.z;
            a.e = p->args_h.z;
            a.z2 = z2_d[0];
        } else {
            a.val = val_d[l];
            a.dinv = dinv_d[l];
            a.dinv_out = dinv_d[l];
            a.b = b_d[l];
            a.r1 = r1_d[l];
            a.e = e_d[l];
            a.z2 = z2_d[l];
        }
        if (l + 1 < depth) {
            a.next_tiles = p->amg[l + 1].tiles_d;
            a.n_next_tiles = p->amg[l + 1].n_tiles;
            a.next_valp = val_d[l + 1];
            a.next_b = b_d[l + 1];
            a.next_e = e_d[l + 1];
        }
        int rc = dalloc(&p->amg[l].args_d, 1);
        if (rc) return rc;
        p->amg_buffers.push_back(p->amg[l].args_d);
    }
    if (p->lds_rows > 0) {  // second residual buffer of the fused update (fused_update)
        double* r_alt = nullptr;
        const int rc = zeros(p, p->vec_len, &r_alt);
        if (rc) return rc;
        p->args_h.r_alt = r_alt;
    }
    return ADT_OK;
}

}  // namespace

extern "C" int adt_pcg_attach_amg(adt_pcg_plan* p, const int32_t* h_indptr, const int32_t* h_indices,
                                  const double* h_values, const int64_t* h_val_off, double theta, int coarse_max,
                                  int max_levels, int stall_limit, double smoother_scale, const int32_t* h_p_ptr,
                                  const int32_t* h_p_idx, const double* h_p_val, int p_cols, int* h_levels) {
    ADT_REQUIRE(p && h_indptr && h_indices && h_values && h_val_off, "adt_pcg_attach_amg: null pointer");
    ADT_REQUIRE(theta >= 0.0 && theta < 1.0, "adt_pcg_attach_amg: theta must be in [0, 1)");
    ADT_REQUIRE(coarse_max >= 1 && coarse_max <= 4096, "adt_pcg_attach_amg: coarse_max must be in [1, 4096]");
    ADT_REQUIRE(max_levels >= 2 && max_levels <= 32, "adt_pcg_attach_amg: max_levels must be in [2, 32]");
    ADT_REQUIRE(stall_limit >= 0, "adt_pcg_attach_amg: stall_limit must be >= 0");
    ADT_REQUIRE(smoother_scale > 0.0 && smoother_scale < 1.5, "adt_pcg_attach_amg: smoother_scale must be in (0, 1.5)");
    adt::amg::Options opt;
    opt.theta = theta;
    opt.coarse_max = coarse_max;
    opt.max_levels = max_levels;
    opt.smoother_scale = smoother_scale;
    adt::amg::FixedProlongator fp;
    if (h_p_ptr) {
        ADT_REQUIRE(h_p_idx && h_p_val && p_cols >= 1, "adt_pcg_attach_amg: incomplete first-level prolongator");
        for (int s = 0; s < p->n_sys; ++s)
            ADT_REQUIRE(p->pattern_of_sys[s] == 0,
                        "adt_pcg_attach_amg: a first-level prolongator needs systems of one CSR pattern");
        const int n = static_cast<int>(p->sys_h[0].rows);
        ADT_REQUIRE(h_p_ptr[0] == 0, "adt_pcg_attach_amg: prolongator row pointers must start at 0");
        for (int i = 0; i < n; ++i) ADT_REQUIRE(h_p_ptr[i + 1] >= h_p_ptr[i], "adt_pcg_attach_amg: prolongator row pointers decrease");
        for (int t = 0; t < h_p_ptr[n]; ++t)
            ADT_REQUIRE(h_p_idx[t] >= 0 && h_p_idx[t] < p_cols && std::isfinite(h_p_val[t]),
                        "adt_pcg_attach_amg: prolongator entry %d out of range or not finite", t);
        fp.nc = p_cols;
        fp.ptr = h_p_ptr;
        fp.idx = h_p_idx;
        fp.val = h_p_val;
        opt.first = &fp;
    }
    const int rc = attach_amg(p, h_indptr, h_indices, h_values, h_val_off, opt);
    if (rc) {
        p->free_amg();
        return rc;
    }
    p->stall_limit = stall_limit;
    if (h_levels) *h_levels = static_cast<int>(p->amg.size());
    return ADT_OK;
}

extern "C" int adt_pcg_amg_level(const adt_pcg_plan* p, int sys, int level, int* h_rows, int64_t* h_nnz,
                                 double* h_omega) {
    ADT_REQUIRE(p, "adt_pcg_amg_level: null plan");
    ADT_REQUIRE(sys >= 0 && sys < p->n_sys, "adt_pcg_amg_level: system out of range");
    ADT_REQUIRE(level >= 0 && level < static_cast<int>(p->amg.size()), "adt_pcg_amg_level: level out of range");
    const LvSys& S = p->amg[level].ls_h[sys];
    if (h_rows) *h_rows = S.rows;
    if (h_nnz) *h_nnz = level == 0 ? -1 : static_cast<int64_t>(p->amg[level - 1].ls_h[sys].n_next);  // -1: caller's
    if (h_omega) *h_omega = S.omega_p;
    return ADT_OK;
}

extern "C" int adt_amg_hierarchy_info(int n, const int32_t* h_indptr, const int32_t* h_indices,
                                      const double* h_values, double theta, int coarse_max, int max_levels,
                                      int* h_levels, int* h_rows, int64_t* h_nnz) {
    ADT_REQUIRE(n >= 1 && h_indptr && h_indices && h_values && h_levels, "adt_amg_hierarchy_info: invalid input");
    ADT_REQUIRE(max_levels >= 1 && max_levels <= 32, "adt_amg_hierarchy_info: max_levels must be in [1, 32]");
    adt::amg::Options opt;
    opt.theta = theta;
    opt.coarse_max = coarse_max;
    opt.max_levels = max_levels;
    const auto H = adt::amg::build_hierarchy(n, h_indptr, h_indices, h_values, opt);
    ADT_REQUIRE(!H.empty(), "adt_amg_hierarchy_info: matrix is not SPD-like (missing or non-positive diagonal)");
    *h_levels = static_cast<int>(H.size());
    for (size_t l = 0; l < H.size(); ++l) {
        if (h_rows) h_rows[l] = H[l].n;
        if (h_nnz) h_nnz[l] = static_cast<int64_t>(H[l].idx.size());
    }
    return ADT_OK;
}

extern "C" int adt_amg_aggregates(int n, const int32_t* h_indptr, const int32_t* h_indices, const double* h_values,
                                  double theta, int32_t* h_agg) {
    ADT_REQUIRE(n >= 1 && h_indptr && h_indices && h_values && h_agg, "adt_amg_aggregates: invalid input");
    adt::amg::Options opt;
    opt.theta = theta;
    opt.coarse_max = 0;
    opt.max_levels = 2;
    const auto H = adt::amg::build_hierarchy(n, h_indptr, h_indices, h_values, opt);
    ADT_REQUIRE(!H.empty(), "adt_amg_aggregates: matrix is not SPD-like (missing or non-positive diagonal)");
    for (int i = 0; i < n; ++i) h_agg[i] = H[0].agg.empty() ? -1 : H[0].agg[i];
    return ADT_OK;
}

extern "C" int adt_pcg_tiling(const adt_pcg_plan* p, int* h_rows_per_tile, int* h_tiles, int* h_lds_rows) {
    ADT_REQUIRE(p, "adt_pcg_tiling: null plan");
    if (h_rows_per_tile) *h_rows_per_tile = p->rb;
    if (h_tiles) *h_tiles = p->n_tiles;
    if (h_lds_rows) *h_lds_rows = p->lds_rows;
    return ADT_OK;
}

extern "C" int adt_pcg_layout(const adt_pcg_plan* p, int* h_rp, int64_t* h_vec_len, int64_t* h_vec_off) {
    ADT_REQUIRE(p, "adt_pcg_layout: null plan");
    if (h_rp) *h_rp = p->rp;
    if (h_vec_len) *h_vec_len = p->vec_len;
    if (h_vec_off)
        for (int s = 0; s < p->n_sys; ++s) h_vec_off[s] = p->sys_h[s].vec_off;
    return ADT_OK;
}

extern "C" int adt_pcg_destroy(adt_pcg_plan* p) {
    delete p;
    return ADT_OK;
}

namespace {

struct SolveReport {
    std::vector<int> iters, flags;
    std::vector<double> rr, bb;
};

int read_report(adt_pcg_plan* p, hipStream_t st, SolveReport& out) {
    const int64_t nsc = static_cast<int64_t>(p->n_sys) * p->rp;
    out.iters.resize(nsc);
    out.flags.resize(nsc);
    out.rr.resize(nsc);
    out.bb.resize(nsc);
    const PcgArgs& a = p->args_h;
    ADT_HIP_CHECK(hipMemcpyAsync(out.iters.data(), a.iters, sizeof(int) * nsc, hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipMemcpyAsync(out.flags.data(), a.flags, sizeof(int) * nsc, hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipMemcpyAsync(out.rr.data(), a.rr, sizeof(double) * nsc, hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipMemcpyAsync(out.bb.data(), a.bb, sizeof(double) * nsc, hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipStreamSynchronize(st));
    return ADT_OK;
}

int upload_args(adt_pcg_plan* p, hipStream_t st, const double* omega_slot) {
    PcgArgs& a = p->args_h;
    ADT_HIP_CHECK(hipMemcpyAsync(p->args_d, &a, sizeof(PcgArgs), hipMemcpyHostToDevice, st));
    for (size_t l = 0; l < p->amg.size(); ++l) {
        LevelArgs& la = p->amg[l].args_h;
        if (l == 0) {
            la.ptr = a.indptr;
            la.idx = a.indices;
            la.val = a.values;
            la.omega_slot = omega_slot;
        }
        ADT_HIP_CHECK(hipMemcpyAsync(p->amg[l].args_d, &la, sizeof(LevelArgs), hipMemcpyHostToDevice, st));
    }
    return ADT_OK;
}

}  // namespace

extern "C" int adt_pcg_solve(adt_pcg_plan* p, const int32_t* d_indptr, const int32_t* d_indices,
                             const double* d_values, const double* d_b, double* d_x, double tol, int max_iter,
                             int precond, const void* precond_data, int check_every, int32_t* h_iters,
                             double* h_relres, void* stream) {
    ADT_REQUIRE(p && d_indptr && d_indices && d_values && d_b && d_x, "adt_pcg_solve: null pointer");
    ADT_REQUIRE(tol > 0.0 && std::isfinite(tol), "adt_pcg_solve: tol must be positive");
    ADT_REQUIRE(max_iter >= 0, "adt_pcg_solve: max_iter must be >= 0");
    ADT_REQUIRE(precond == ADT_PRECOND_JACOBI || precond == ADT_PRECOND_LEARNED || precond == ADT_PRECOND_AMG,
                "adt_pcg_solve: unsupported preconditioner %d", precond);
    ADT_REQUIRE(precond == ADT_PRECOND_JACOBI || !p->amg.empty(),
                "adt_pcg_solve: the multigrid preconditioner needs adt_pcg_attach_amg first");
    ADT_REQUIRE(precond != ADT_PRECOND_LEARNED || precond_data,
                "adt_pcg_solve: the learned preconditioner needs its per-slot damping array");
    const int chunk = std::min(kMaxChunk, std::max(2, (check_every + 1) / 2 * 2));
    hipStream_t st = adt::as_stream(stream);
    PcgArgs& a = p->args_h;
    a.indptr = d_indptr;
    a.indices = d_indices;
    a.values = d_values;
    a.b = d_b;
    a.x = d_x;
    a.tol2 = tol * tol;
    const int mode = precond == ADT_PRECOND_JACOBI ? 0 : 1;
    a.stall_limit = mode == 1 ? p->stall_limit : 0;
    a.r_pingpong = fused_update(p, mode) ? 1 : 0;
    // Kernels read the arguments from device memory so one captured graph serves every solve.
    int rc = upload_args(p, st, precond == ADT_PRECOND_LEARNED ? static_cast<const double*>(precond_data) : nullptr);
    if (rc) return rc;
    if (mode == 1 && (rc = amg_numeric(p, st))) return rc;
    if ((rc = solve_dispatch(p, st, max_iter, chunk, mode))) return rc;
    SolveReport rep;
    if ((rc = read_report(p, st, rep))) return rc;
    p->last_fallback = 0;
    p->last_polished = 0;
    p->last_fcg_iters_max = 0;
    if (mode == 1) {
        std::vector<int> first_iters = rep.iters;
        int need = 0;
        for (int s = 0; s < p->n_sys; ++s)
            for (int c = 0; c < p->r; ++c) {
                const int64_t i = static_cast<int64_t>(s) * p->rp + c;
                p->last_fcg_iters_max = std::max(p->last_fcg_iters_max, rep.iters[i]);
                need += rep.flags[i] == -1 || rep.flags[i] == -2;
            }
        if (need > 0 && max_iter > 0) {
            // Jacobi PCG finishes every right-hand side from its current iterate; converged ones
            // stop at once (their residual already meets the tolerance).
            p->last_fallback = need;
            a.stall_limit = 0;
            a.r_pingpong = 0;
            if ((rc = upload_args(p, st, nullptr)) || (rc = solve_dispatch(p, st, max_iter, chunk, 0)) ||
                (rc = read_report(p, st, rep)))
                return rc;
            for (size_t i = 0; i < rep.iters.size(); ++i) rep.iters[i] += first_iters[i];
        }
    }
    // The recurrence residual that stops CG drifts from b - A x by rounding (a few 1e-13 of |b|
    // here).  Right-hand sides whose true residual misses tol restart once from their iterate,
    // with a freshly computed residual, in Jacobi PCG; every other right-hand side stops at once.
    int polish = 0;
    for (size_t i = 0; i < rep.flags.size(); ++i)
        polish += rep.flags[i] == 0 && rep.bb[i] > 0.0 && rep.rr[i] > a.tol2 * rep.bb[i];
    if (polish > 0 && max_iter > 0) {
        std::vector<int> prior_iters = rep.iters;
        p->last_polished = polish;
        a.stall_limit = 0;
        a.r_pingpong = 0;
        if ((rc = upload_args(p, st, nullptr)) || (rc = solve_dispatch(p, st, max_iter, chunk, 0)) ||
            (rc = read_report(p, st, rep)))
            return rc;
        for (size_t i = 0; i < rep.iters.size(); ++i) rep.iters[i] += prior_iters[i];
    }
    int worst = ADT_OK;
    for (int s = 0; s < p->n_sys; ++s)
        for (int c = 0; c < p->r; ++c) {
            const int64_t i = static_cast<int64_t>(s) * p->rp + c;
            if (h_iters) h_iters[static_cast<int64_t>(s) * p->r + c] = rep.iters[i];
            const double rel = rep.bb[i] > 0.0 ? std::sqrt(rep.rr[i] / rep.bb[i]) : 0.0;
            if (h_relres) h_relres[static_cast<int64_t>(s) * p->r + c] = rel;
            if (rep.flags[i] == -1 || !std::isfinite(rel))
                worst = ADT_ERR_NONFINITE;
            else if ((rep.flags[i] == 1 || rep.flags[i] == -2) && worst == ADT_OK)
                worst = ADT_ERR_NOT_CONVERGED;
        }
    if (worst == ADT_ERR_NONFINITE) return adt::set_error(worst, "PCG produced non-finite values or broke down");
    if (worst == ADT_ERR_NOT_CONVERGED)
        return adt::set_error(worst, "PCG reached max_iter=%d before tol=%g", max_iter, tol);
    return ADT_OK;
}

extern "C" int adt_pcg_stats(const adt_pcg_plan* p, int* h_levels, int* h_fallback_rhs, int* h_polished_rhs) {
    ADT_REQUIRE(p, "adt_pcg_stats: null plan");
    if (h_levels) *h_levels = static_cast<int>(p->amg.size());
    if (h_fallback_rhs) *h_fallback_rhs = p->last_fallback;
    if (h_polished_rhs) *h_polished_rhs = p->last_polished;
    return ADT_OK;
}

extern "C" int adt_spmm(int n_sys, int lanes, int rp, const int32_t* h_rows, const int64_t* h_ptr_off,
                        const int64_t* h_idx_off, const int64_t* h_val_off, const int64_t* h_vec_off,
                        const int32_t* d_indptr, const int32_t* d_indices, const double* d_values,
                        const double* d_x, double* d_y, int mode, void* stream) {
    ADT_REQUIRE(n_sys >= 0 && (lanes == 1 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64) &&
                    rp % lanes == 0 && rp > 0,
                "adt_spmm: invalid lanes/rp");
    ADT_REQUIRE(mode == 0 || mode == 1, "adt_spmm: mode must be 0 or 1");
    if (n_sys == 0) return ADT_OK;
    ADT_REQUIRE(n_sys <= 65535 && rp / lanes <= 65535, "adt_spmm: too many systems");
    std::vector<Sys> sys(n_sys);
    int max_rows = 0;
    for (int s = 0; s < n_sys; ++s) {
        sys[s] = Sys{h_rows[s], h_ptr_off[s], h_idx_off[s], h_val_off[s], h_vec_off[s], 0, h_vec_off[s], 0};
        max_rows = std::max(max_rows, h_rows[s]);
    }
    if (max_rows == 0) return ADT_OK;
    hipStream_t st = adt::as_stream(stream);
    static thread_local Sys* d_sys = nullptr;
    static thread_local int cap = 0;
    if (cap < n_sys) {
        if (d_sys) ADT_HIP_CHECK(hipFree(d_sys));
        ADT_HIP_CHECK(hipMalloc(&d_sys, sizeof(Sys) * n_sys));
        cap = n_sys;
    }
    ADT_HIP_CHECK(hipMemcpyAsync(d_sys, sys.data(), sizeof(Sys) * n_sys, hipMemcpyHostToDevice, st));
    int rc;
    switch (lanes) {
        case 1: rc = launch_spmm<1>(n_sys, rp, max_rows, d_sys, d_indptr, d_indices, d_values, d_x, d_y, mode, st); break;
        case 8: rc = launch_spmm<8>(n_sys, rp, max_rows, d_sys, d_indptr, d_indices, d_values, d_x, d_y, mode, st); break;
        case 16: rc = launch_spmm<16>(n_sys, rp, max_rows, d_sys, d_indptr, d_indices, d_values, d_x, d_y, mode, st); break;
        case 32: rc = launch_spmm<32>(n_sys, rp, max_rows, d_sys, d_indptr, d_indices, d_values, d_x, d_y, mode, st); break;
        default: rc = launch_spmm<64>(n_sys, rp, max_rows, d_sys, d_indptr, d_indices, d_values, d_x, d_y, mode, st); break;
    }
    if (rc) return rc;
    // the metadata buffer is reused by the next call: keep the host copy alive until uploaded
    ADT_HIP_CHECK(hipStreamSynchronize(st));
    return ADT_OK;
}

extern "C" int adt_pcg_profile(adt_pcg_plan* p, int enable) {
    ADT_REQUIRE(p, "adt_pcg_profile: null plan");
    p->profile = enable ? 1 : 0;
    p->prof_ms[0] = p->prof_ms[1] = p->prof_ms[2] = 0.0;
    p->prof_launches = 0;
    return ADT_OK;
}

extern "C" int adt_pcg_profile_read(const adt_pcg_plan* p, double* h_out) {
    ADT_REQUIRE(p && h_out, "adt_pcg_profile_read: null pointer");
    std::vector<int> gi(p->n_groups);
    ADT_HIP_CHECK(hipMemcpy(gi.data(), p->args_h.grp_iters, sizeof(int) * p->n_groups, hipMemcpyDeviceToHost));
    int64_t group_iters = 0;
    for (int v : gi) group_iters += v;
    h_out[0] = p->prof_ms[0];
    h_out[1] = p->prof_ms[1];
    h_out[2] = static_cast<double>(p->prof_launches);
    h_out[3] = static_cast<double>(group_iters);
    h_out[4] = static_cast<double>(p->L);
    h_out[5] = static_cast<double>(p->rp);
    h_out[6] = p->prof_ms[2];
    return ADT_OK;
}

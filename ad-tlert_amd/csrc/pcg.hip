// (b) Batched multi-RHS FP64 sparse products and preconditioned conjugate gradients.
//
// Layout: each system's vectors are an (N, Rp) row-major block (sources contiguous).  A
// workgroup owns a tile = (system, lane group of L right-hand sides, row range).  Lanes
// t % L walk the L right-hand sides of one row, so every nonzero reads one contiguous
// L*8-byte run of the gathered vector; the CSR row itself is read once per lane group.
//
// One CG iteration is four launches (captured once in a hipGraph and replayed):
//   spmv   : p_new = z + beta p_old (formed on the fly for every neighbour, z = D^-1 r),
//            q = A p_new, partial sums of p_new.q per (tile, lane)
//   alpha  : one workgroup per lane group reduces the tile partials in a fixed order
//   update : x += alpha p_new, r -= alpha q, partials of r.z and r.r
//   beta   : reduce, convergence test ||r|| <= tol ||b||, beta = rho_new / rho
// p is double-buffered (ping-pong) because neighbours' p_old is still being read while
// p_new is written.  Converged lane groups skip all work; the host polls one counter of
// active right-hand sides between graph replays (one replay in flight ahead).
#include "adt_common.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

namespace {

struct Sys {
    int64_t rows, ptr_off, idx_off, val_off, vec_off, diag_off, b_off, lidx_off;
};

// ptile: index of the tile's row range in the halo tables of its CSR pattern (-1: none)
struct Tile {
    int sys, grp, row0, row1, ptile, pad;
};

struct PcgArgs {
    const uint16_t* lidx;   // per nonzero: LDS slot of its column within the row's tile
    const int* halo_ptr;    // per pattern tile: range in halo_rows
    const int* halo_rows;   // rows outside a tile that its rows reference
    const Sys* sys;
    const Tile* tiles;
    const int* grp_tile0;
    const int* grp_ntiles;
    const int32_t* indptr;
    const int32_t* indices;
    const double* values;
    const double* b;
    double* x;
    double* r;
    double* p[2];
    double* q;
    double* z;  // flexible CG: preconditioned residual of the learned network
    double* dinv;
    double* part0;
    double* part1;
    double* part2;
    double* alpha;
    double* beta;
    double* rho;
    double* bb;
    double* rr;
    double* rz_old;
    int* iters;
    int* flags;  // 1 active, 0 converged, 2 zero right-hand side, -1 breakdown / non-finite
    int* stall;
    int* grp_active;
    int* grp_iters;
    int* active_total;
    double tol2;
    int rp, groups, r_real, n_tiles;
    int lds_rows, lds_nnz, lds_own, lds_halo;  // LDS-staged product capacities
};

constexpr int kRowsPerPass(int L) { return 256 / L; }
constexpr int kSpmvUnroll = 8;

// Sum over the threads of a 256-wide block that share lane = t % L; thread t < L gets the
// lane total.  Two independent values are reduced together.  Fixed order.
template <int L>
__device__ __forceinline__ void block_sum2(double& v0, double& v1) {
    __shared__ double sh[2][4][64];
    v0 = adt::wave_sum_strided<L>(v0);
    v1 = adt::wave_sum_strided<L>(v1);
    const int wave = threadIdx.x >> 6, wl = threadIdx.x & 63;
    if (wl < L) {
        sh[0][wave][wl] = v0;
        sh[1][wave][wl] = v1;
    }
    __syncthreads();
    if (threadIdx.x < L) {
        const int l = threadIdx.x;
        v0 = ((sh[0][0][l] + sh[0][1][l]) + sh[0][2][l]) + sh[0][3][l];
        v1 = ((sh[1][0][l] + sh[1][1][l]) + sh[1][2][l]) + sh[1][3][l];
    }
    __syncthreads();
}

template <int L>
__device__ __forceinline__ double block_sum1(double v) {
    double dummy = 0.0;
    block_sum2<L>(v, dummy);
    return v;
}

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
    const int32_t* ptr = indptr + S.ptr_off;
    const int32_t* idx = indices + S.idx_off;
    const double* val = values + S.val_off;
    const double* X = x + S.vec_off + col;
    double* Y = y + S.vec_off + col;
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
    c.t = a.tiles[tix];
    c.S = a.sys[c.t.sys];
    const int g = c.t.grp % a.groups;
    c.col = g * L + static_cast<int>(threadIdx.x % L);
    c.sc = static_cast<int64_t>(c.t.sys) * a.rp + c.col;
    return true;
}

__device__ __forceinline__ int tile_index(const PcgArgs& a) { return adt::xcd_tile(blockIdx.x, a.n_tiles); }

template <int L>
__device__ __forceinline__ double diag_inverse(const int32_t* ptr, const int32_t* idx, const double* val,
                                               int row) {
    double d = 0.0;
    for (int k = ptr[row]; k < ptr[row + 1]; ++k)
        if (idx[k] == row) d = val[k];
    return 1.0 / d;
}

// r = b - A x, dinv, p_0 = 0; partials ||b||^2, r.z, ||r||^2.
template <int L>
__global__ __launch_bounds__(256) void cg_init(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    TileCtx c;
    load_tile<L>(a, c);
    const int tix = tile_index(a);
    const int32_t* ptr = a.indptr + c.S.ptr_off;
    const int32_t* idx = a.indices + c.S.idx_off;
    const double* val = a.values + c.S.val_off;
    const int rp = a.rp;
    const bool real = c.col < a.r_real;
    const bool first_group = (c.t.grp % a.groups) == 0;
    const double* B = a.b + c.S.b_off + c.col;
    const double* X = a.x + c.S.vec_off + c.col;
    double* R = a.r + c.S.vec_off + c.col;
    double* P0 = a.p[0] + c.S.vec_off + c.col;
    double* P1 = a.p[1] + c.S.vec_off + c.col;
    double* D = a.dinv + c.S.diag_off;
    double sb = 0.0, sz = 0.0, sr = 0.0;
    for (int row = c.t.row0 + static_cast<int>(threadIdx.x / L); row < c.t.row1; row += kRowsPerPass(L)) {
        const double di = diag_inverse<L>(ptr, idx, val, row);
        if (first_group && (threadIdx.x % L) == 0) D[row] = di;
        const int64_t o = static_cast<int64_t>(row) * rp;
        P0[o] = 0.0;
        P1[o] = 0.0;
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
        a.part0[static_cast<int64_t>(tix) * L + threadIdx.x] = s1;
        a.part1[static_cast<int64_t>(tix) * L + threadIdx.x] = s2;
        a.part2[static_cast<int64_t>(tix) * L + threadIdx.x] = s3;
    }
}

// Fixed-order sum of one lane group's tile partials; thread l < L receives lane l's total.
template <int L>
__device__ __forceinline__ double group_partials(const double* part, int tile0, int ntiles) {
    const int lane = threadIdx.x % L, chunk = threadIdx.x / L;
    double s = 0.0;
    for (int i = chunk; i < ntiles; i += kRowsPerPass(L)) s += part[static_cast<int64_t>(tile0 + i) * L + lane];
    return block_sum1<L>(s);
}

template <int L>
__device__ __forceinline__ void refresh_group(const PcgArgs& a, int gid, bool lane_active) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (threadIdx.x < L && lane_active) atomicOr(&any, 1);
    __syncthreads();
    if (threadIdx.x == 0) a.grp_active[gid] = any;
}

template <int L>
__global__ __launch_bounds__(256) void cg_init_reduce(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    const int t0 = a.grp_tile0[gid], nt = a.grp_ntiles[gid];
    const double bb = group_partials<L>(a.part0, t0, nt);
    const double rz = group_partials<L>(a.part1, t0, nt);
    const double rr = group_partials<L>(a.part2, t0, nt);
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
        a.flags[sc] = flag;
        a.bb[sc] = bb;
        a.rr[sc] = rr;
        a.rho[sc] = rz;
        a.rz_old[sc] = rz;
        a.beta[sc] = 0.0;
        a.alpha[sc] = 0.0;
        a.iters[sc] = 0;
        a.stall[sc] = 0;
        act = flag == 1;
        if (act) atomicAdd(a.active_total, 1);
    }
    if (threadIdx.x == 0) a.grp_iters[gid] = 0;
    refresh_group<L>(a, gid, act);
}

// x = 0 for zero right-hand sides (the solution of A x = 0).
template <int L>
__global__ __launch_bounds__(256) void cg_zero_solution(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    TileCtx c;
    load_tile<L>(a, c);
    if (a.flags[c.sc] != 2) return;
    double* X = a.x + c.S.vec_off + c.col;
    for (int row = c.t.row0 + static_cast<int>(threadIdx.x / L); row < c.t.row1; row += kRowsPerPass(L))
        X[static_cast<int64_t>(row) * a.rp] = 0.0;
}

// Jacobi PCG step 1: p_new = D^-1 r + beta p_old on the fly, q = A p_new, partial p_new.q.
template <int L>
__global__ __launch_bounds__(256) void cg_spmv(const PcgArgs* __restrict__ pa, int parity) {
    const PcgArgs& a = *pa;
    const int tix = tile_index(a);
    const Tile t = a.tiles[tix];
    if (!a.grp_active[t.grp]) return;
    const Sys S = a.sys[t.sys];
    const int col = (t.grp % a.groups) * L + static_cast<int>(threadIdx.x % L);
    const int64_t sc = static_cast<int64_t>(t.sys) * a.rp + col;
    const bool act = a.flags[sc] == 1;
    const double beta = act ? a.beta[sc] : 0.0;
    const int32_t* ptr = a.indptr + S.ptr_off;
    const int32_t* idx = a.indices + S.idx_off;
    const double* val = a.values + S.val_off;
    const int rp = a.rp;
    const double* R = a.r + S.vec_off + col;
    const double* Pold = a.p[parity] + S.vec_off + col;
    double* Pnew = a.p[parity ^ 1] + S.vec_off + col;
    double* Q = a.q + S.vec_off + col;
    const double* D = a.dinv + S.diag_off;
    double pq = 0.0;
    if (act) {
        for (int row = t.row0 + static_cast<int>(threadIdx.x / L); row < t.row1; row += kRowsPerPass(L)) {
            double acc = 0.0;
            const int k1 = ptr[row + 1];
            // kSpmvUnroll entries at a time: all index loads, then all gathers, so one thread keeps
            // several independent memory requests in flight (padding entries re-read the own row
            // with a zero coefficient).
            for (int kb = ptr[row]; kb < k1; kb += kSpmvUnroll) {
                int64_t jj[kSpmvUnroll];
                double vv[kSpmvUnroll];
#pragma unroll
                for (int u = 0; u < kSpmvUnroll; ++u) {
                    const bool ok = kb + u < k1;
                    jj[u] = ok ? idx[kb + u] : row;
                    vv[u] = ok ? val[kb + u] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kSpmvUnroll; ++u) {
                    const double zj = D[jj[u]] * R[jj[u] * rp];
                    acc = fma(vv[u], fma(beta, Pold[jj[u] * rp], zj), acc);
                }
            }
            const int64_t o = static_cast<int64_t>(row) * rp;
            const double zi = D[row] * R[o];
            const double pi = fma(beta, Pold[o], zi);
            Pnew[o] = pi;
            Q[o] = acc;
            pq = fma(pi, acc, pq);
        }
    }
    pq = block_sum1<L>(pq);
    if (threadIdx.x < L) a.part0[static_cast<int64_t>(tix) * L + threadIdx.x] = pq;
}

// LDS-staged variant of cg_spmv, three phases per tile:
//   1. coalesced bulk copy of the tile's CSR slice (values, tile-local column slots, row
//      pointers) and halo row list into LDS;
//   2. stage p_new = D^-1 r + beta p_old for the tile's rows and halo rows (independent
//      loads, unrolled) into LDS, writing p_new of the own rows;
//   3. q = A p_new entirely from LDS (slot-major with L lanes contiguous: conflict-free).
// Every vector row is read from memory once per tile and the inner product loop has no
// global-memory latency in it.
template <int L>
__global__ __launch_bounds__(256) void cg_spmv_lds(const PcgArgs* __restrict__ pa, int parity) {
    extern __shared__ double smem[];
    const PcgArgs& a = *pa;
    const int tix = tile_index(a);
    const Tile t = a.tiles[tix];
    if (!a.grp_active[t.grp]) return;
    const Sys S = a.sys[t.sys];
    const int lane = static_cast<int>(threadIdx.x % L), rsub = static_cast<int>(threadIdx.x / L);
    const int col = (t.grp % a.groups) * L + lane;
    const int64_t sc = static_cast<int64_t>(t.sys) * a.rp + col;
    const bool act = a.flags[sc] == 1;
    const double beta = act ? a.beta[sc] : 0.0;
    const int32_t* ptr = a.indptr + S.ptr_off;
    const int rp = a.rp;
    const double* R = a.r + S.vec_off + col;
    const double* Pold = a.p[parity] + S.vec_off + col;
    double* Pnew = a.p[parity ^ 1] + S.vec_off + col;
    double* Q = a.q + S.vec_off + col;
    const double* D = a.dinv + S.diag_off;
    const int own = t.row1 - t.row0;
    const int h0 = a.halo_ptr[t.ptile], nh = a.halo_ptr[t.ptile + 1] - h0;
    const int staged = own + nh;
    const int e0 = ptr[t.row0], ne = ptr[t.row1] - e0;
    double* stage = smem;                                      // lds_rows x L
    double* sval = smem + static_cast<int64_t>(a.lds_rows) * L;  // lds_nnz
    int* sptr = reinterpret_cast<int*>(sval + a.lds_nnz);        // lds_own + 1
    int* shalo = sptr + a.lds_own + 1;                           // lds_halo
    uint16_t* sidx = reinterpret_cast<uint16_t*>(shalo + a.lds_halo);  // lds_nnz
    // phase 1
    const double* val = a.values + S.val_off + e0;
    const uint16_t* lidx = a.lidx + S.lidx_off + e0;
    for (int e = static_cast<int>(threadIdx.x); e < ne; e += 256) {
        sval[e] = val[e];
        sidx[e] = lidx[e];
    }
    for (int i = static_cast<int>(threadIdx.x); i <= own; i += 256) sptr[i] = ptr[t.row0 + i] - e0;
    for (int i = static_cast<int>(threadIdx.x); i < nh; i += 256) shalo[i] = a.halo_rows[h0 + i];
    __syncthreads();
    // phase 2
    if (act) {
        constexpr int U = 4;
        for (int s0 = rsub; s0 < staged; s0 += U * kRowsPerPass(L)) {
            int64_t jj[U];
            double dj[U], rj[U], pj[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int s = s0 + u * kRowsPerPass(L);
                jj[u] = s < own ? t.row0 + s : (s < staged ? shalo[s - own] : t.row0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                dj[u] = D[jj[u]];
                rj[u] = R[jj[u] * rp];
                pj[u] = Pold[jj[u] * rp];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int s = s0 + u * kRowsPerPass(L);
                if (s < staged) {
                    const double v = fma(beta, pj[u], dj[u] * rj[u]);
                    stage[s * L + lane] = v;
                    if (s < own) Pnew[jj[u] * rp] = v;
                }
            }
        }
    }
    __syncthreads();
    // phase 3
    double pq = 0.0;
    if (act) {
        for (int s = rsub; s < own; s += kRowsPerPass(L)) {
            const int k1 = sptr[s + 1];
            double acc = 0.0;
            for (int k = sptr[s]; k < k1; ++k) acc = fma(sval[k], stage[static_cast<int>(sidx[k]) * L + lane], acc);
            Q[static_cast<int64_t>(t.row0 + s) * rp] = acc;
            pq = fma(stage[s * L + lane], acc, pq);
        }
    }
    pq = block_sum1<L>(pq);
    if (threadIdx.x < L) a.part0[static_cast<int64_t>(tix) * L + threadIdx.x] = pq;
}

template <int L>
__global__ __launch_bounds__(256) void cg_alpha(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    if (!a.grp_active[gid]) return;
    const double pq = group_partials<L>(a.part0, a.grp_tile0[gid], a.grp_ntiles[gid]);
    bool act = false;
    if (threadIdx.x < L) {
        const int64_t sc = static_cast<int64_t>(gid / a.groups) * a.rp + (gid % a.groups) * L + threadIdx.x;
        if (a.flags[sc] == 1) {
            const double alpha = a.rho[sc] / pq;
            if (!(pq > 0.0) || !std::isfinite(alpha)) {
                a.flags[sc] = -1;
                a.alpha[sc] = 0.0;
                atomicSub(a.active_total, 1);
            } else {
                a.alpha[sc] = alpha;
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
    const Tile t = a.tiles[tix];
    if (!a.grp_active[t.grp]) return;
    const Sys S = a.sys[t.sys];
    const int col = (t.grp % a.groups) * L + static_cast<int>(threadIdx.x % L);
    const int64_t sc = static_cast<int64_t>(t.sys) * a.rp + col;
    const bool act = a.flags[sc] == 1;
    const int rp = a.rp;
    const double* P = a.p[parity ^ 1] + S.vec_off + col;
    const double* Q = a.q + S.vec_off + col;
    double* X = a.x + S.vec_off + col;
    double* R = a.r + S.vec_off + col;
    const double* D = a.dinv + S.diag_off;
    double sz = 0.0, sr = 0.0;
    if (act) {
        const double alpha = a.alpha[sc];
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
        a.part1[static_cast<int64_t>(tix) * L + threadIdx.x] = sz;
        a.part2[static_cast<int64_t>(tix) * L + threadIdx.x] = sr;
    }
}

template <int L>
__global__ __launch_bounds__(256) void cg_beta(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    if (!a.grp_active[gid]) return;
    if (threadIdx.x == 0) a.grp_iters[gid] += 1;
    const int t0 = a.grp_tile0[gid], nt = a.grp_ntiles[gid];
    const double rz = group_partials<L>(a.part1, t0, nt);
    const double rr = group_partials<L>(a.part2, t0, nt);
    bool act = false;
    if (threadIdx.x < L) {
        const int64_t sc = static_cast<int64_t>(gid / a.groups) * a.rp + (gid % a.groups) * L + threadIdx.x;
        if (a.flags[sc] == 1) {
            a.iters[sc] += 1;
            a.rr[sc] = rr;
            if (!std::isfinite(rr) || !std::isfinite(rz)) {
                a.flags[sc] = -1;
                atomicSub(a.active_total, 1);
            } else if (rr <= a.tol2 * a.bb[sc]) {
                a.flags[sc] = 0;
                atomicSub(a.active_total, 1);
            } else {
                a.beta[sc] = rz / a.rho[sc];
                a.rho[sc] = rz;
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
    const int32_t* ptr = a.indptr + c.S.ptr_off;
    const int32_t* idx = a.indices + c.S.idx_off;
    const double* val = a.values + c.S.val_off;
    const double* B = a.b + c.S.b_off + c.col;
    const double* X = a.x + c.S.vec_off + c.col;
    double sr = 0.0;
    if (c.col < a.r_real) {
        for (int row = c.t.row0 + static_cast<int>(threadIdx.x / L); row < c.t.row1; row += kRowsPerPass(L)) {
            double acc = 0.0;
            for (int k = ptr[row]; k < ptr[row + 1]; ++k)
                acc = fma(val[k], X[static_cast<int64_t>(idx[k]) * a.rp], acc);
            const double ri = B[static_cast<int64_t>(row) * a.rp] - acc;
            sr = fma(ri, ri, sr);
        }
    }
    sr = block_sum1<L>(sr);
    if (threadIdx.x < L) a.part2[static_cast<int64_t>(tix) * L + threadIdx.x] = sr;
}

template <int L>
__global__ __launch_bounds__(256) void cg_true_reduce(const PcgArgs* __restrict__ pa) {
    const PcgArgs& a = *pa;
    const int gid = blockIdx.x;
    const double rr = group_partials<L>(a.part2, a.grp_tile0[gid], a.grp_ntiles[gid]);
    if (threadIdx.x < L) {
        const int64_t sc = static_cast<int64_t>(gid / a.groups) * a.rp + (gid % a.groups) * L + threadIdx.x;
        a.rr[sc] = rr;
    }
}

}  // namespace

// ------------------------------------------------------------------------------------------
// plan
// ------------------------------------------------------------------------------------------

constexpr int kMaxChunk = 64;

struct adt_pcg_plan {
    int n_sys = 0, r = 0, L = 0, rp = 0, rb = 0, groups = 0, n_tiles = 0, n_groups = 0;
    int64_t vec_len = 0, diag_len = 0;
    std::vector<Sys> sys_h;
    std::vector<int64_t> rows_h;
    Sys* sys_d = nullptr;
    Tile* tiles_d = nullptr;
    int* grp_tile0_d = nullptr;
    int* grp_ntiles_d = nullptr;
    double* vec = nullptr;  // r, p0, p1, q, z
    double* dinv = nullptr;
    double* part = nullptr;
    double* scal = nullptr;
    int* ints = nullptr;
    int* h_active = nullptr;
    PcgArgs args_h{};
    PcgArgs* args_d = nullptr;
    hipStream_t cap = nullptr;
    // two identical iteration graphs replayed alternately, so a replay in flight never touches
    // the profiling events of the replay being read
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_iters = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int profile = 0;
    int lds_rows = 0;  // staged rows of the LDS spmv (0: gather kernel)
    int lds_nnz = 0, lds_halo = 0;
    uint16_t* lidx_d = nullptr;
    int* halo_ptr_d = nullptr;
    int* halo_rows_d = nullptr;
    hipEvent_t pev[2][kMaxChunk][4] = {};
    double prof_ms[2] = {0.0, 0.0};
    int64_t prof_launches = 0;
    ~adt_pcg_plan();
};

adt_pcg_plan::~adt_pcg_plan() {
    for (auto& g : graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& set : pev)
        for (auto& it : set)
            for (auto& e : it)
                if (e) (void)hipEventDestroy(e);
    if (cap) (void)hipStreamDestroy(cap);
    void* bufs[] = {sys_d, tiles_d, grp_tile0_d, grp_ntiles_d, vec, dinv, part, scal, ints, args_d, lidx_d, halo_ptr_d,
                    halo_rows_d};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (h_active) (void)hipHostFree(h_active);
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

int64_t lds_bytes_for(int staged, int nnz, int own, int halo, int lanes);

template <int L>
void launch_iteration(const adt_pcg_plan* p, hipStream_t st, int parity, hipEvent_t* events) {
    const PcgArgs* pa = p->args_d;
    if (events) (void)hipEventRecord(events[0], st);
    if (p->lds_rows > 0)
        hipLaunchKernelGGL(cg_spmv_lds<L>, dim3(p->n_tiles), dim3(256),
                           static_cast<size_t>(lds_bytes_for(p->lds_rows, p->lds_nnz, p->rb, p->lds_halo, L)), st,
                           pa, parity);
    else
        hipLaunchKernelGGL(cg_spmv<L>, dim3(p->n_tiles), dim3(256), 0, st, pa, parity);
    if (events) (void)hipEventRecord(events[1], st);
    hipLaunchKernelGGL(cg_alpha<L>, dim3(p->n_groups), dim3(256), 0, st, pa);
    if (events) (void)hipEventRecord(events[2], st);
    hipLaunchKernelGGL(cg_update<L>, dim3(p->n_tiles), dim3(256), 0, st, pa, parity);
    if (events) (void)hipEventRecord(events[3], st);
    hipLaunchKernelGGL(cg_beta<L>, dim3(p->n_groups), dim3(256), 0, st, pa);
}

// Capture `chunk` (even) iterations; replays keep the p ping-pong parity at 0.
template <int L>
int ensure_graphs(adt_pcg_plan* p, int chunk) {
    if (p->graph[0] && p->graph_iters == chunk) return ADT_OK;
    for (auto& g : p->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    for (int set = 0; set < 2; ++set) {
        hipGraph_t g = nullptr;
        ADT_HIP_CHECK(hipStreamBeginCapture(p->cap, hipStreamCaptureModeRelaxed));
        for (int i = 0; i < chunk; ++i) launch_iteration<L>(p, p->cap, i & 1, nullptr);
        ADT_HIP_CHECK(hipStreamEndCapture(p->cap, &g));
        hipError_t e = hipGraphInstantiate(&p->graph[set], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return adt::set_error(ADT_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    p->graph_iters = chunk;
    return ADT_OK;
}

// Profiling mode: the same kernels launched directly on the caller's stream with timing events
// around every spmv and update launch (HIP event nodes inside graphs cannot be timed).
template <int L>
int launch_profiled_chunk(adt_pcg_plan* p, hipStream_t st, int chunk, int set) {
    for (int i = 0; i < chunk; ++i) {
        for (auto& e : p->pev[set][i])
            if (!e) ADT_HIP_CHECK(hipEventCreate(&e));
        launch_iteration<L>(p, st, i & 1, p->pev[set][i]);
    }
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

int accumulate_profile(adt_pcg_plan* p, int set, int chunk) {
    for (int i = 0; i < chunk; ++i) {
        float a = 0.f, b = 0.f;
        ADT_HIP_CHECK(hipEventElapsedTime(&a, p->pev[set][i][0], p->pev[set][i][1]));
        ADT_HIP_CHECK(hipEventElapsedTime(&b, p->pev[set][i][2], p->pev[set][i][3]));
        p->prof_ms[0] += a;
        p->prof_ms[1] += b;
        p->prof_launches += 1;
    }
    return ADT_OK;
}

template <int L>
int solve_impl(adt_pcg_plan* p, hipStream_t st, int max_iter, int chunk) {
    const PcgArgs* pa = p->args_d;
    ADT_HIP_CHECK(hipMemsetAsync(p->args_h.active_total, 0, sizeof(int), st));
    hipLaunchKernelGGL(cg_init<L>, dim3(p->n_tiles), dim3(256), 0, st, pa);
    hipLaunchKernelGGL(cg_init_reduce<L>, dim3(p->n_groups), dim3(256), 0, st, pa);
    hipLaunchKernelGGL(cg_zero_solution<L>, dim3(p->n_tiles), dim3(256), 0, st, pa);
    ADT_HIP_CHECK(hipGetLastError());
    if (max_iter > 0) {
        int rc = p->profile ? ADT_OK : ensure_graphs<L>(p, chunk);
        if (rc) return rc;
        int k = 0;
        bool stop = false;
        for (int it = 0; it < max_iter && !stop; it += chunk, ++k) {
            const int set = k & 1;
            if (p->profile) {
                if ((rc = launch_profiled_chunk<L>(p, st, chunk, set))) return rc;
            } else {
                ADT_HIP_CHECK(hipGraphLaunch(p->graph[set], st));
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
    hipLaunchKernelGGL(cg_true_residual<L>, dim3(p->n_tiles), dim3(256), 0, st, pa);
    hipLaunchKernelGGL(cg_true_reduce<L>, dim3(p->n_groups), dim3(256), 0, st, pa);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

}  // namespace

namespace {

// Halo tables of one CSR pattern cut into row tiles of rb rows.
struct PatternTiles {
    std::vector<uint16_t> lidx;
    std::vector<int> halo_ptr;  // one entry per tile (start), closed by the caller
    std::vector<int> halo_rows;
    int max_staged = 0, max_nnz = 0, max_halo = 0;
};

int64_t lds_bytes_for(int staged, int nnz, int own, int halo, int lanes) {
    return static_cast<int64_t>(staged) * lanes * 8 + static_cast<int64_t>(nnz) * 8 + (own + 1 + halo) * 4 +
           static_cast<int64_t>(nnz) * 2;
}

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
        pt.max_halo = std::max(pt.max_halo, static_cast<int>(halo.size()));
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

}  // namespace

extern "C" int adt_pcg_create(adt_pcg_plan** out, int n_sys, int r, int lanes, int rows_per_tile,
                              const int32_t* h_rows, const int64_t* h_ptr_off, const int64_t* h_idx_off,
                              const int64_t* h_val_off, const int64_t* h_b_off, const int32_t* h_indptr,
                              const int32_t* h_indices, int lds_bytes) {
    ADT_REQUIRE(out && h_rows && h_ptr_off && h_idx_off && h_val_off, "adt_pcg_create: null pointer");
    ADT_REQUIRE(n_sys >= 1 && r >= 1, "adt_pcg_create: need at least one system and one right-hand side");
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
    typedef std::tuple<int64_t, int64_t, int> PatKey;
    std::map<PatKey, int> pattern_of;
    std::vector<PatKey> patterns;
    for (int s = 0; s < n_sys; ++s) {
        PatKey key{h_ptr_off[s], h_idx_off[s], h_rows[s]};
        if (!pattern_of.count(key)) {
            pattern_of[key] = static_cast<int>(patterns.size());
            patterns.push_back(key);
        }
    }
    std::vector<PatternTiles> ptiles;
    if (h_indptr && h_indices && lds_bytes > 0) {
        int device = 0, limit = 0;
        if (hipGetDevice(&device) == hipSuccess &&
            hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
            limit > 0)
            lds_bytes = std::min(lds_bytes, limit - 8192);  // the block reductions keep static LDS
        for (int rb = p->rb; rb >= rpp; rb = (rb / 2 + rpp - 1) / rpp * rpp) {
            ptiles.clear();
            int max_staged = 0, max_nnz = 0, max_halo = 0;
            for (const auto& key : patterns) {
                ptiles.push_back(build_pattern_tiles(h_indptr + std::get<0>(key), h_indices + std::get<1>(key),
                                                     std::get<2>(key), rb));
                max_staged = std::max(max_staged, ptiles.back().max_staged);
                max_nnz = std::max(max_nnz, ptiles.back().max_nnz);
                max_halo = std::max(max_halo, ptiles.back().max_halo);
            }
            if (lds_bytes_for(max_staged, max_nnz, rb, max_halo, lanes) <= lds_bytes && max_staged < 65536) {
                p->rb = rb;
                p->lds_rows = max_staged;
                p->lds_nnz = max_nnz;
                p->lds_halo = max_halo;
                break;
            }
            if (rb == rpp) break;
        }
        if (!p->lds_rows) ptiles.clear();
    }
    std::vector<int64_t> lidx_base(patterns.size(), 0), ptile_base(patterns.size(), 0);
    std::vector<uint16_t> lidx_all;
    std::vector<int> halo_ptr_all, halo_rows_all;
    for (size_t q = 0; q < ptiles.size(); ++q) {
        lidx_base[q] = static_cast<int64_t>(lidx_all.size());
        ptile_base[q] = static_cast<int64_t>(halo_ptr_all.size());
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
        const int q = pattern_of[PatKey{h_ptr_off[s], h_idx_off[s], h_rows[s]}];
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
                                     p->lds_rows ? static_cast<int>(ptile_base[q]) + count : -1, 0});
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
        (rc = dalloc(&p->scal, 7 * nsc)) || (rc = dalloc(&p->ints, 3 * nsc + 2 * p->n_groups + 1)) ||
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
    PcgArgs& a = p->args_h;
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
    a.lds_halo = p->lds_halo;
    a.groups = p->groups;
    a.r_real = r;
    a.n_tiles = p->n_tiles;
    *out = p;
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

extern "C" int adt_pcg_solve(adt_pcg_plan* p, const int32_t* d_indptr, const int32_t* d_indices,
                             const double* d_values, const double* d_b, double* d_x, double tol, int max_iter,
                             int precond, const void* gnn, int check_every, int32_t* h_iters, double* h_relres,
                             void* stream) {
    ADT_REQUIRE(p && d_indptr && d_indices && d_values && d_b && d_x, "adt_pcg_solve: null pointer");
    ADT_REQUIRE(tol > 0.0 && std::isfinite(tol), "adt_pcg_solve: tol must be positive");
    ADT_REQUIRE(max_iter >= 0, "adt_pcg_solve: max_iter must be >= 0");
    ADT_REQUIRE(precond == ADT_PRECOND_JACOBI, "adt_pcg_solve: unsupported preconditioner %d", precond);
    (void)gnn;
    const int chunk = std::min(kMaxChunk, std::max(2, (check_every + 1) / 2 * 2));
    hipStream_t st = adt::as_stream(stream);
    PcgArgs& a = p->args_h;
    a.indptr = d_indptr;
    a.indices = d_indices;
    a.values = d_values;
    a.b = d_b;
    a.x = d_x;
    a.tol2 = tol * tol;
    // Kernels read the arguments from device memory so one captured graph serves every solve.
    ADT_HIP_CHECK(hipMemcpyAsync(p->args_d, &a, sizeof(PcgArgs), hipMemcpyHostToDevice, st));
    int rc = ADT_OK;
    switch (p->L) {
        case 8: rc = solve_impl<8>(p, st, max_iter, chunk); break;
        case 16: rc = solve_impl<16>(p, st, max_iter, chunk); break;
        case 32: rc = solve_impl<32>(p, st, max_iter, chunk); break;
        default: rc = solve_impl<64>(p, st, max_iter, chunk); break;
    }
    if (rc) return rc;
    const int64_t nsc = static_cast<int64_t>(p->n_sys) * p->rp;
    std::vector<int> iters(nsc), flags(nsc);
    std::vector<double> rr(nsc), bb(nsc);
    ADT_HIP_CHECK(hipMemcpyAsync(iters.data(), a.iters, sizeof(int) * nsc, hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipMemcpyAsync(flags.data(), a.flags, sizeof(int) * nsc, hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipMemcpyAsync(rr.data(), a.rr, sizeof(double) * nsc, hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipMemcpyAsync(bb.data(), a.bb, sizeof(double) * nsc, hipMemcpyDeviceToHost, st));
    ADT_HIP_CHECK(hipStreamSynchronize(st));
    int worst = ADT_OK;
    for (int s = 0; s < p->n_sys; ++s)
        for (int c = 0; c < p->r; ++c) {
            const int64_t i = static_cast<int64_t>(s) * p->rp + c;
            if (h_iters) h_iters[static_cast<int64_t>(s) * p->r + c] = iters[i];
            const double rel = bb[i] > 0.0 ? std::sqrt(rr[i] / bb[i]) : 0.0;
            if (h_relres) h_relres[static_cast<int64_t>(s) * p->r + c] = rel;
            if (flags[i] == -1 || !std::isfinite(rel))
                worst = ADT_ERR_NONFINITE;
            else if (flags[i] == 1 && worst == ADT_OK)
                worst = ADT_ERR_NOT_CONVERGED;
        }
    if (worst == ADT_ERR_NONFINITE) return adt::set_error(worst, "PCG produced non-finite values or broke down");
    if (worst == ADT_ERR_NOT_CONVERGED)
        return adt::set_error(worst, "PCG reached max_iter=%d before tol=%g", max_iter, tol);
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
    p->prof_ms[0] = p->prof_ms[1] = 0.0;
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
    return ADT_OK;
}

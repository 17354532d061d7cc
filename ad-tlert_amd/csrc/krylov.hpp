// Device-side data of the batched Krylov solver shared by the conjugate-gradient kernels
// (pcg.hip) and the multigrid preconditioner kernels (amg_kernels.hpp).
//
// Layout: each system's vectors are an (N, Rp) row-major block (sources contiguous).  A
// workgroup owns a tile = (system, lane group of L right-hand sides, row range).  Lanes
// t % L walk the L right-hand sides of one row, so every nonzero reads one contiguous
// L*8-byte run of the gathered vector; the CSR row itself is read once per lane group.
#pragma once

#include "adt_common.hpp"

namespace adt {
namespace krylov {

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
    // segmented group sums (groups of many tiles): segment = run of consecutive tiles of one group
    const int* seg_tile0;
    const int* seg_ntiles;
    const int* seg_grp;
    const int* grp_seg0;
    const int* grp_nseg;
    double* spart;  // 3 x (n_segs, L) segment partials
    const int32_t* indptr;
    const int32_t* indices;
    const double* values;
    const double* b;
    double* x;
    double* r;
    double* p[2];
    double* q;
    double* z;  // flexible CG: preconditioned residual (also the fine-level smoothed residual)
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
    double* pq_old;   // flexible CG: p.q of the previous iteration
    double* rr_best;  // stall detection: smallest ||r||^2 seen
    int* iters;
    int* flags;  // 1 active, 0 converged, 2 zero right-hand side, -1 breakdown / non-finite, -2 stalled
    int* stall;
    int* grp_active;
    int* grp_iters;
    int* active_total;
    double tol2;
    int rp, groups, r_real, n_tiles;
    int lds_rows, lds_nnz, lds_own, lds_pad_;  // LDS-staged product capacities
    int stall_limit;                           // iterations without a 10% residual decrease (0: off)
    int segmented, n_segs;
    // Flexible CG with the update fused to the next pre-smoothing: neighbouring tiles read the
    // residual while it is updated, so it alternates between r (even iterations) and r_alt (odd)
    int r_pingpong;
    double* r_alt;
};

// The finest-level residual of the iteration with this parity.
__device__ __forceinline__ double* residual(const PcgArgs& g, int parity) {
    return g.r_pingpong && (parity & 1) ? g.r_alt : g.r;
}

// One multigrid level of one system.  Offsets index the level's concatenated arrays.  On the
// finest level the matrix is the caller's (ptr/idx/val offsets are the system's CSR offsets).
struct LvSys {
    int rows, nc;           // rows on this level, rows on the next (0 on the coarsest)
    int n_p, n_t, n_next;   // slots of P, of T = A P and of the next level's matrix
    int pad;
    double omega;           // smoother damping
    double omega_p;         // prolongation damping (constant; per slot when learned)
    int64_t ptr, idx, diag, val;   // matrix pattern, diagonal slots, values
    int64_t dmap_ptr, dmap;        // filtered diagonal: summed slots per row
    int64_t vec, dinv;             // (rows, Rp) vector block, inverse diagonal
    int64_t pptr, pidx, pmap_ptr, pmap, pval;  // prolongator pattern / gather map / values
    int64_t rptr, rent;            // restriction rows (nc + 1) and (fine row, P slot) entries
    int64_t tptr, tpair_ptr, tpair, tval;      // T = A P
    int64_t cpair_ptr, cpair, next_val;        // next level values from (P slot, T slot) pairs
    int64_t next_vec;              // vector block on the next level
    int64_t dense;                 // coarsest level: dense inverse (rows x rows)
};

struct LevelArgs {
    const LvSys* ls;
    const Tile* tiles;
    const Tile* next_tiles;
    const int* grp_active;
    int n_tiles, n_next_tiles, rp, groups, n_sys, coarsest, fine, pad;
    // matrix of this level (finest level: the caller's arrays)
    const int32_t* ptr;
    const int32_t* idx;
    const double* val;
    const int* diag;
    const double* dinv;
    const int* dmap_ptr;
    const int* dmap;
    // prolongator, restriction, Galerkin product lists
    const int* pptr;
    const int* pidx;
    const int* pmap_ptr;
    const int* pmap;
    const int* pself;
    double* pval;
    const int* rptr;
    const int* rrow;
    const int* rslot;
    const int* tpair_ptr;
    const int* tpair_a;
    const int* tpair_p;
    double* tval;
    const int* cpair_ptr;
    const int* cpair_p;
    const int* cpair_t;
    double* next_valp;       // next level values (all systems)
    double* dinv_out;        // this level's inverse diagonal (written by the numeric setup)
    const double* omega_slot;  // finest level: per-slot learned damping (NULL: constant omega)
    const double* pfix;        // a given prolongator's values, indexed like pidx (NULL: aggregation)
    double* dense;           // coarsest: dense inverses
    // vectors
    const double* b;   // right-hand side of this level (finest: the Krylov residual r)
    double* r1;        // residual after pre-smoothing (finest: aliases the z buffer)
    double* e;         // correction computed on this level (finest: z)
    double* z2;        // pre-smoothed + prolonged correction (levels >= 1, or finest without LDS)
    double* next_b;
    const double* next_e;
    // LDS staging of this level's smoothing (levels >= 1; lidx == NULL: gather kernels)
    const uint16_t* lidx;   // per matrix slot: staged row of its column
    const int* halo_ptr;    // per pattern tile: first entry in halo_rows
    const int* halo_rows;
    int lds_rows, lds_nnz, lds_own, lds_pad;
    // LDS staging of this level's prolongation (plidx == NULL: gather kernel): per pattern tile
    // the coarse rows that its P rows reference, and per P slot the staged row of its column
    const uint16_t* plidx;
    const int* pc_ptr;
    const int* pc_rows;
    int p_lds_rows, p_lds_nnz, p_lds_own, p_lds_pad;
    // LDS staging of the restriction to the next level: largest slice of a next-level tile
    int r_lds_nnz, r_lds_pad;
};

constexpr int kRowsPerPass(int L) { return 256 / L; }

// LDS-staged tiles keep each thread's own-row operands in registers from the staging pass to
// the product pass: a staged tile has at most own_max(L) rows per thread (rows_per_tile <=
// own_max(L) * 256 / L).  Halo rows are staged halo_batch(L) rows per thread at a time, all of
// their loads in flight together (one batch covers the halo of the usual 2D and 3D tiles).
// 64-lane tiles (3D) are short and have long halos: fewer own rows, more halo rows per thread.
constexpr int own_max(int L) { return L >= 64 ? 4 : 8; }
constexpr int halo_batch(int L) { return L >= 64 ? 24 : 12; }
constexpr int max_staged_tile_rows(int L) { return own_max(L) * kRowsPerPass(L); }

// Row-batched CSR gather: for the RB rows of one thread, acc[b] += sum_k val(k) * term(idx(k))
// over k in [k0[b], k1[b]), accumulated per row in slot order.  U slots of every row are
// requested before any is used, so up to RB * U independent loads are in flight per thread
// (a row-at-a-time loop would pay index -> value -> gather latency once per row).  Loads are
// unconditional (slots past a row's end read slot 0 of the arrays and are discarded by a
// select), so no branch makes the compiler wait for the loads in flight.
template <int RB, int U, typename IdxF, typename ValF, typename TermF>
__device__ __forceinline__ void gather_rows(const int (&k0)[RB], const int (&k1)[RB], IdxF idx, ValF val, TermF term,
                                            double (&acc)[RB]) {
    int span = 0;
#pragma unroll
    for (int b = 0; b < RB; ++b) span = max(span, k1[b] - k0[b]);
    for (int kb = 0; kb < span; kb += U) {
        int jj[RB][U];
        double vv[RB][U], tt[RB][U];
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0[b] + kb + u;
                const int ks = k < k1[b] ? k : 0;
                jj[b][u] = idx(ks);
                vv[b][u] = val(ks);
            }
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int u = 0; u < U; ++u) tt[b][u] = term(jj[b][u]);
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double next = fma(vv[b][u], tt[b][u], acc[b]);
                acc[b] = k0[b] + kb + u < k1[b] ? next : acc[b];
            }
    }
}

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

// Per-lane-group reduction kernels run kReduceThreads threads: a group's tile partials (one
// contiguous block of ntiles x L doubles) are summed by kReduceThreads / L parallel chunks with
// four independent accumulators each, so groups of many tiles (large 3D systems) do not
// serialize on load latency.
constexpr int kReduceThreads = 1024;

// Fixed-order sum of one lane group's tile partials; thread l < L receives lane l's total.
// Must be called by all kReduceThreads threads of the block.
template <int L>
__device__ __forceinline__ double group_partials(const double* part, int tile0, int ntiles) {
    constexpr int chunks = kReduceThreads / L;
    constexpr int waves = kReduceThreads / 64;
    __shared__ double sh[waves][64];
    const int lane = static_cast<int>(threadIdx.x) % L, chunk = static_cast<int>(threadIdx.x) / L;
    const auto base = adt::gview(part + static_cast<int64_t>(tile0) * L + lane);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = chunk;
    for (; i + 3 * chunks < ntiles; i += 4 * chunks) {
        s0 += base[static_cast<int64_t>(i) * L];
        s1 += base[static_cast<int64_t>(i + chunks) * L];
        s2 += base[static_cast<int64_t>(i + 2 * chunks) * L];
        s3 += base[static_cast<int64_t>(i + 3 * chunks) * L];
    }
    for (; i < ntiles; i += chunks) s0 += base[static_cast<int64_t>(i) * L];
    double s = adt::wave_sum_strided<L>((s0 + s1) + (s2 + s3));
    const int wave = static_cast<int>(threadIdx.x) >> 6, wl = static_cast<int>(threadIdx.x) & 63;
    if (wl < L) sh[wave][wl] = s;
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < L) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < waves; ++w) t += sh[w][threadIdx.x];
        s = t;
    }
    __syncthreads();
    return s;
}

// Lane totals of partial array k (0, 1, 2) over lane group gid: from the tile partials, or from
// the segment partials when the plan pre-reduces runs of tiles (reduce_segments).
template <int L>
__device__ __forceinline__ double group_sum(const PcgArgs& a, int k, int gid) {
    if (a.segmented)
        return group_partials<L>(a.spart + static_cast<int64_t>(k) * a.n_segs * L, adt::gview(a.grp_seg0)[gid],
                                 adt::gview(a.grp_nseg)[gid]);
    const double* part = k == 0 ? a.part0 : (k == 1 ? a.part1 : a.part2);
    return group_partials<L>(part, adt::gview(a.grp_tile0)[gid], adt::gview(a.grp_ntiles)[gid]);
}

template <int L>
__device__ __forceinline__ void refresh_group(const PcgArgs& a, int gid, bool lane_active) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (threadIdx.x < L && lane_active) atomicOr(&any, 1);
    __syncthreads();
    if (threadIdx.x == 0) adt::gview(a.grp_active)[gid] = any;
}

// Dynamic LDS of a staged tile: staged rows x L doubles, the CSR slice (values, 16-bit local
// columns) and its row pointers.
inline int64_t lds_bytes_for(int staged, int nnz, int own, int lanes) {
    return static_cast<int64_t>(staged) * lanes * 8 + static_cast<int64_t>(nnz) * 8 + (own + 1) * 4 +
           static_cast<int64_t>(nnz) * 2;
}

}  // namespace krylov
}  // namespace adt

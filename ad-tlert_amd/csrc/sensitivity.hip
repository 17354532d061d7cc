// (c) fused normal sensitivity, (d) adjoint cell gradient, and the small measurement kernels.
#include "adt_common.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <cmath>

namespace {

constexpr int kMaxWaves = 64;

struct SensArgs {
    const double* phi;
    const int32_t* a;
    const int32_t* b;
    const int32_t* m;
    const int32_t* n;
    const int32_t* dofs;
    const double* stiff;
    const double* mass;
    double* out;
    int64_t n_dofs;
    int W, S, D, C, dch;
    double ws[kMaxWaves];
    double wm[kMaxWaves];
};

template <typename T>
__device__ __forceinline__ double pot(const double* f, int row, int S, int64_t N, int node) {
    return row < S ? f[static_cast<int64_t>(row) * N + node] : 0.0;
}

// One thread owns one (measurement, cell) entry: lanes walk 64 consecutive cells (coalesced
// output rows and neighbouring potential gathers), the four wavefronts of a workgroup take
// interleaved measurements of the block's measurement chunk.  Cell dofs and unit templates
// stay in registers while the thread sweeps its measurements; the sum runs in the fixed order
// (w, i, j).  P = float gives FP32 products with FP64 accumulation over wavenumbers.
template <int K, typename P>
__global__ __launch_bounds__(256) void sensitivity_kernel(SensArgs A) {
    const int c = blockIdx.x * 64 + static_cast<int>(threadIdx.x & 63);
    if (c >= A.C) return;
    const int sub = static_cast<int>(threadIdx.x >> 6);
    const int d0 = blockIdx.y * A.dch;
    const int d1 = min(d0 + A.dch, A.D);
    int dof[K];
#pragma unroll
    for (int k = 0; k < K; ++k) dof[k] = A.dofs[static_cast<int64_t>(c) * K + k];
    constexpr bool kRegs = K <= 4;
    P kt[kRegs ? K * K : 1], mt[kRegs ? K * K : 1];
    const double* Kc = A.stiff + static_cast<int64_t>(c) * K * K;
    const double* Mc = A.mass ? A.mass + static_cast<int64_t>(c) * K * K : nullptr;
    if constexpr (kRegs) {
#pragma unroll
        for (int e = 0; e < K * K; ++e) {
            kt[e] = static_cast<P>(Kc[e]);
            mt[e] = Mc ? static_cast<P>(Mc[e]) : P(0);
        }
    }
    const int64_t N = A.n_dofs;
    for (int d = d0 + sub; d < d1; d += 4) {
        const int ia = A.a[d], ib = A.b[d], im = A.m[d], in = A.n[d];
        double acc = 0.0;
        for (int w = 0; w < A.W; ++w) {
            const double* f = A.phi + static_cast<int64_t>(w) * A.S * N;
            P cur[K], rec[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                cur[k] = static_cast<P>(pot<P>(f, ia, A.S, N, dof[k]) - pot<P>(f, ib, A.S, N, dof[k]));
                rec[k] = static_cast<P>(pot<P>(f, im, A.S, N, dof[k]) - pot<P>(f, in, A.S, N, dof[k]));
            }
            P sk = 0, sm = 0;
#pragma unroll
            for (int i = 0; i < K; ++i) {
                P yk = 0, ym = 0;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    if constexpr (kRegs) {
                        yk = fma(kt[i * K + j], cur[j], yk);
                        ym = fma(mt[i * K + j], cur[j], ym);
                    } else {
                        yk = fma(static_cast<P>(Kc[i * K + j]), cur[j], yk);
                        if (Mc) ym = fma(static_cast<P>(Mc[i * K + j]), cur[j], ym);
                    }
                }
                sk = fma(rec[i], yk, sk);
                sm = fma(rec[i], ym, sm);
            }
            acc = fma(A.ws[w], static_cast<double>(sk), acc);
            acc = fma(A.wm[w], static_cast<double>(sm), acc);
        }
        A.out[static_cast<int64_t>(d) * A.C + c] = -acc;
    }
}

template <int K>
int launch_sensitivity(const SensArgs& a, int precision, hipStream_t st) {
    dim3 grid(adt::blocks_for(a.C, 64), adt::blocks_for(a.D, a.dch));
    if (precision == 1)
        hipLaunchKernelGGL((sensitivity_kernel<K, float>), grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((sensitivity_kernel<K, double>), grid, dim3(256), 0, st, a);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

// ------------------------------------------------------------------------------------------
// (c') Gram form of the normal sensitivity on interleaved (N, Rp) fields, for many wavenumbers:
//   G_c[e][f] = sum_w sum_i phi_w[e, dof_ci] (T_wc phi_w[f, dof_c.])_i,   T_wc = ws_w K_c + wm_w M_c
//   out[d, c] = -(G_c[m][a] - G_c[m][b] - G_c[n][a] + G_c[n][b])          (row S: remote, zero)
// which equals the per-output formula of sensitivity_kernel.  The element templates are symmetric
// positive semidefinite, so T_wc = R^T R with R = D^1/2 L^T from T = L D L^T (L unit lower, the
// pivots of a semidefinite template that vanish to rounding are dropped), and
//   G_c = Y^T Y,   Y = [R_w Phi_w]_w   (rows (w, i), one column per source)
// is a symmetric rank-(W K) product: only its upper-triangle 16 x 16 tiles are formed.  One
// workgroup per cell: the W factors R_w are computed into LDS (one thread per wavenumber), every
// wave gathers whole field rows of its wavenumbers (the cell's K nodes, lanes along the sources,
// all of them in flight together) and stores Y = R Phi to LDS; the 4 waves then accumulate the
// tiles on the FP64 matrix cores, the upper triangle replaces Y in LDS, and the workgroup writes
// column c of every measurement row.  All wavenumbers are staged at once when Y fits the LDS
// budget (one global round trip per cell), else in chunks.  Cells are mapped XCD-contiguously, so
// neighbouring columns of a row are written by workgroups that share one L2.  FP64 with a fixed
// order of rows (w, i); bitwise repeatable.
// ------------------------------------------------------------------------------------------

struct GramArgs {
    const double* x;
    const int32_t* dofs;
    const double* stiff;
    const double* mass;
    const uint4* quad;         // per quadrupole: LDS byte offsets of G(m,a), G(m,b), G(n,a), G(n,b)
    double* out;
    const int32_t* col_ptr;    // optional column aggregation: column p sums cells col_cells[col_ptr[p] ..]
    const int32_t* col_cells;
    int S, Sp, K, W, rp, wchunk, rows_max, C, D;
    int P;                     // output columns: C, or the aggregated column count
    int group_doubles;         // LDS doubles of one column group (factors and staged rows)
    int factor_doubles;        // LDS doubles of the factors R_w of all wavenumbers
    int64_t vec_off[kMaxWaves];
    double ws[kMaxWaves];
    double wm[kMaxWaves];
};

// Cells of output column `col`: itself, or the aggregation list [c0, c1) into col_cells.
__device__ __forceinline__ void column_cells(const int32_t* col_ptr, int col, int& c0, int& c1) {
    if (col_ptr) {
        c0 = adt::gview(col_ptr)[col];
        c1 = adt::gview(col_ptr)[col + 1];
    } else {
        c0 = col;
        c1 = col + 1;
    }
}

__device__ __forceinline__ int tri_index(int e, int f, int S) {  // e <= f
    return e * S - (e * (e - 1)) / 2 + (f - e);
}

// Column c of J from the cell's Gram block: J[d, c] = -((g(m,a) - g(m,b)) - g(n,a)) + g(n,b)) for
// every quadrupole d.  The electrode indices of U quadrupoles per thread are loaded together
// (clamped, branch-free) before any store, so the loop does not wait for a load round trip per
// quadrupole.
template <typename GF>
__device__ __forceinline__ void write_cell_column(const int32_t* a, const int32_t* b, const int32_t* m,
                                                  const int32_t* n, int D, int64_t C, int c, double* out, GF g) {
    constexpr int U = 8;
    const auto ga = adt::gview(a), gb = adt::gview(b), gm = adt::gview(m), gn = adt::gview(n);
    const auto o = adt::gview(out);
    for (int d0 = static_cast<int>(threadIdx.x); d0 < D; d0 += 256 * U) {
        int ia[U], ib[U], im[U], in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int d = d0 + u * 256;
            const int dc = d < D ? d : 0;
            ia[u] = ga[dc];
            ib[u] = gb[dc];
            im[u] = gm[dc];
            in[u] = gn[dc];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int d = d0 + u * 256;
            const double v = ((g(im[u], ia[u]) - g(im[u], ib[u])) - g(in[u], ia[u])) + g(in[u], ib[u]);
            if (d < D) o[static_cast<int64_t>(d) * C + c] = -v;
        }
    }
}

// Byte offsets in the Gram triangle of gram_kernel (upper triangle over S + 1 slots, slot S the
// remote electrode) of the four G entries of every quadrupole: electrode indices >= S or < 0 are
// the remote slot.
__global__ void quad_offsets_kernel(int D, int S, const int32_t* a, const int32_t* b, const int32_t* m,
                                    const int32_t* n, uint4* quad) {
    const int d = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (d >= D) return;
    auto slot = [S](int e) { return e < 0 || e > S ? S : e; };
    auto off = [S](int e, int f) {
        const int lo = min(e, f), hi = max(e, f);
        return static_cast<uint32_t>(tri_index(lo, hi, S + 1)) * 8u;
    };
    const int ia = slot(adt::gview(a)[d]), ib = slot(adt::gview(b)[d]);
    const int im = slot(adt::gview(m)[d]), in = slot(adt::gview(n)[d]);
    adt::gview(quad)[d] = make_uint4(off(im, ia), off(im, ib), off(in, ia), off(in, ib));
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Upper-triangle 16 x 16 tiles of G_c per wave: Sp = 16 Bt sources give Bt (Bt + 1) / 2 tiles
// dealt over the workgroup's 4 waves.
constexpr int gram_tiles_per_wave(int bt) { return (bt * (bt + 1) / 2 + 3) / 4; }

// In place: the upper triangle of the symmetric positive semidefinite K x K matrix in R (row-major)
// becomes R = D^1/2 L^T of T = L D L^T, i.e. T = R^T R (the upper Cholesky factor), by outer
// products row by row.  A pivot at or below 1e-13 of the largest diagonal entry is rounding left
// by a semidefinite template: its row of R is zero.  Entries below the diagonal are zeroed.
__device__ void template_root(double* R, int K) {
    double tmax = 0.0;
    for (int i = 0; i < K; ++i) tmax = fmax(tmax, R[i * K + i]);
    const double tiny = 1e-13 * tmax;
    for (int i = 0; i < K; ++i) {
        const double p = R[i * K + i];
        if (p > tiny) {
            const double s = sqrt(p), inv = 1.0 / s;
            R[i * K + i] = s;
            for (int j = i + 1; j < K; ++j) R[i * K + j] *= inv;
            for (int j = i + 1; j < K; ++j)
                for (int k = j; k < K; ++k) R[j * K + k] = fma(-R[i * K + j], R[i * K + k], R[j * K + k]);
        } else {
            for (int j = i; j < K; ++j) R[i * K + j] = 0.0;
        }
        for (int j = 0; j < i; ++j) R[i * K + j] = 0.0;
    }
}

// template_root on a K x K matrix with K <= KM held in registers (the same operations in the same
// order; the LDS version's chain of dependent LDS accesses is the longest step of the factor
// phase for small K).
template <int KM>
__device__ void template_root_regs(double* R, int K) {
    double t[KM][KM];
#pragma unroll
    for (int i = 0; i < KM; ++i)
#pragma unroll
        for (int j = 0; j < KM; ++j) t[i][j] = (i < K && j < K) ? R[i * K + j] : 0.0;
    double tmax = 0.0;
#pragma unroll
    for (int i = 0; i < KM; ++i)
        if (i < K) tmax = fmax(tmax, t[i][i]);
    const double tiny = 1e-13 * tmax;
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        if (i >= K) break;
        const double p = t[i][i];
        if (p > tiny) {
            const double sq = sqrt(p), inv = 1.0 / sq;
            t[i][i] = sq;
#pragma unroll
            for (int j = i + 1; j < KM; ++j) t[i][j] *= inv;
#pragma unroll
            for (int j = i + 1; j < KM; ++j)
#pragma unroll
                for (int k = j; k < KM; ++k) t[j][k] = fma(-t[i][j], t[i][k], t[j][k]);
        } else {
#pragma unroll
            for (int j = i; j < KM; ++j) t[i][j] = 0.0;
        }
#pragma unroll
        for (int j = 0; j < i; ++j) t[i][j] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < KM; ++i)
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (i < K && j < K) R[i * K + j] = t[i][j];
}

// Output columns per workgroup of gram_kernel.  One: four independent workgroups per CU (40 KB of
// LDS each) overlap their load, product and store phases.  Blocks of 2 or 4 adjacent columns
// (16- or 32-byte row segments per store) were measured slower on the 20k-node time-lapse mesh
// (1.36 and 1.38 ms against 1.17 ms): their triangles take the CU's whole LDS, leaving one
// workgroup per CU whose phases run one after another.  Two columns per 512-thread workgroup at
// two workgroups per CU (as many waves per CU) measured 1.20 against 1.14 ms.
constexpr int kGramColumns = 1;

// Y in LDS is stored in 4-row x 16-source bricks of 64 doubles, brick (r / 4, e / 16) at
// ((r / 4) * BT + e / 16) * 64 and entry (r % 4, e % 16) inside it, so that the operand of one
// matrix-core step (4 rows x 16 sources) is 512 contiguous bytes: conflict-free LDS reads.
template <int BT>
__device__ __forceinline__ int brick_index(int r, int e) {
    return (((r >> 2) * BT + (e >> 4)) << 6) + ((r & 3) << 4) + (e & 15);
}

// G_c on the matrix cores: v_mfma_f64_16x16x4f64 multiplies a 16 x 4 block of Y^T by a 4 x 16
// block of Y; lane l supplies Y[r0 + l / 16][16 I + l % 16] and Y[r0 + l / 16][16 J + l % 16] and
// holds G[16 I + l / 16 + 4 v][16 J + l % 16] in item v of its accumulator.  Only the tiles with
// I <= J are formed.  The staged rows are padded with zero rows to a multiple of 4.
//
// A workgroup of CB x 256 threads owns CB adjacent output columns: group g (waves 4g .. 4g + 3)
// forms the G of column g with its own LDS region (factors R_w and the staged Y rows); wave w of
// every group runs on SIMD w, and the tile sets of the groups' waves rotate so that each SIMD gets
// the same number of tiles.  Every group runs the same number of steps (the longest column's cell
// count) so that all threads meet at every barrier.  Then the CB triangles are stored slot-major
// (G[slot * CB + g]) and the workgroup writes J[d, c0 .. c0 + CB) for every quadrupole d: adjacent
// lanes write adjacent columns of one row, so a store instruction writes 64 / CB row segments of
// CB x 8 contiguous bytes instead of 64 scattered doubles.  KM bounds the local size K.
template <int BT, int KM, int CB>
__global__ __launch_bounds__(256 * CB) void gram_kernel(GramArgs A) {
    constexpr int Sp = 16 * BT, TPW = gram_tiles_per_wave(BT), n_tiles = BT * (BT + 1) / 2;
    constexpr int kCols = (Sp + 63) / 64;    // staged sources per lane and row
    constexpr int kGroup = KM <= 4 ? 2 : 1;  // wavenumbers per wave whose rows are in flight together
    constexpr int NT = 256 * CB;
    extern __shared__ double lds[];
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63;
    const int wave_id = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int grp = wave_id >> 2, wv = wave_id & 3;
    const int n_blocks = (A.P + CB - 1) / CB;
    const int blk = adt::xcd_tile(static_cast<int>(blockIdx.x), n_blocks);
    const int c0 = blk * CB;
    const int K = A.K, S = A.S;
    double* rl = lds + static_cast<int64_t>(grp) * A.group_doubles;  // W x K x K: R_w of the cell
    double* Y = rl + A.factor_doubles;                               // bricks of the staged rows
    // cells of every column of the block: all groups step together through the longest list
    int c_begin = 0, c_end = 0, steps = 0;
#pragma unroll
    for (int k = 0; k < CB; ++k) {
        int b0 = 0, b1 = 0;
        if (c0 + k < A.P) column_cells(A.col_ptr, c0 + k, b0, b1);
        steps = max(steps, b1 - b0);
        if (k == grp) {
            c_begin = b0;
            c_end = b1;
        }
    }
    const int wave = (wv + grp + static_cast<int>(blockIdx.x)) & 3;
    int ti[TPW], tj[TPW];
    int nt = 0;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        int q = wave + 4 * t, I = 0;
        if (q < n_tiles) ++nt;
        else q = 0;
        while (q >= BT - I) q -= BT - I++;
        ti[t] = I;
        tj[t] = I + q;
    }
    f64x4 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
    const auto x = adt::gview(A.x);
    const int gtid = tid & 255;  // thread index within the group
    // G of a column = the sum of its cells' G_c (J is linear in G): the accumulators run on over
    // the cells of an aggregated column, in list order
    for (int step = 0; step < steps; ++step) {
        const bool active = c_begin + step < c_end;
        const int c = active ? __builtin_amdgcn_readfirstlane(A.col_ptr ? adt::gview(A.col_cells)[c_begin + step]
                                                                         : c_begin + step)
                             : 0;
        const double* Kc = A.stiff + static_cast<int64_t>(c) * K * K;
        const double* Mc = A.mass ? A.mass + static_cast<int64_t>(c) * K * K : nullptr;
        const int32_t* dof = A.dofs + static_cast<int64_t>(c) * K;
        __syncthreads();  // the previous step's factors and rows are consumed
        // T_c,w for every wavenumber, one thread per entry (the loads of all entries in flight
        // together), then one thread per wavenumber factors its T in place
        if (active)
            for (int t = gtid; t < A.W * K * K; t += 256) {
                const int w = t / (K * K), e = t - w * (K * K);
                rl[t] = Mc ? fma(A.wm[w], Mc[e], A.ws[w] * Kc[e]) : A.ws[w] * Kc[e];
            }
        int64_t node[KM];  // wave-uniform row offsets of the cell's nodes
#pragma unroll
        for (int k = 0; k < KM; ++k)
            node[k] = (active && k < K) ? static_cast<int64_t>(__builtin_amdgcn_readfirstlane(dof[k])) * A.rp : 0;
        __syncthreads();
        if (active && gtid < A.W) {
            if constexpr (KM <= 4)
                template_root_regs<KM>(rl + gtid * K * K, K);
            else
                template_root(rl + gtid * K * K, K);
        }
        __syncthreads();
        for (int w0 = 0; w0 < A.W; w0 += A.wchunk) {
            const int nw = min(A.wchunk, A.W - w0);
            const int rows = nw * K, rows4 = (rows + 3) & ~3;
            if (w0 > 0) __syncthreads();  // the previous chunk's products have read Y
            if (active) {
                // Y rows of the wave's wavenumbers: the K field rows of each, lanes along the sources
                for (int wl0 = wv; wl0 < nw; wl0 += 4 * kGroup) {
                    double v[kGroup][KM][kCols];
#pragma unroll
                    for (int g = 0; g < kGroup; ++g) {
                        const int wl = wl0 + 4 * g;
                        const int64_t base = wl < nw ? A.vec_off[w0 + wl] : 0;
#pragma unroll
                        for (int k = 0; k < KM; ++k) {
                            const auto row = x + (base + node[k]);  // wave-uniform: lanes add their source
#pragma unroll
                            for (int q = 0; q < kCols; ++q) {
                                const int e = lane + 64 * q;
                                v[g][k][q] = (wl < nw && k < K && e < S) ? row[e] : 0.0;
                            }
                        }
                    }
#pragma unroll
                    for (int g = 0; g < kGroup; ++g) {
                        const int wl = wl0 + 4 * g;
                        if (wl >= nw) continue;
                        const double* R = rl + (w0 + wl) * K * K;
#pragma unroll
                        for (int i = 0; i < KM; ++i) {
                            if (i >= K) continue;
#pragma unroll
                            for (int q = 0; q < kCols; ++q) {
                                const int e = lane + 64 * q;
                                double y = 0.0;
#pragma unroll
                                for (int j = i; j < KM; ++j)
                                    if (j < K) y = fma(R[i * K + j], v[g][j][q], y);
                                if (e < Sp) Y[brick_index<BT>(wl * K + i, e)] = y;
                            }
                        }
                    }
                }
                for (int t = gtid; t < (rows4 - rows) * Sp; t += 256) Y[brick_index<BT>(rows + t / Sp, t % Sp)] = 0.0;
            }
            __syncthreads();
            if (active) {
                const double* yl = Y + lane;
                for (int r0 = 0; r0 < rows4; r0 += 4) {
                    const double* yr = yl + (r0 >> 2) * BT * 64;
#pragma unroll
                    for (int t = 0; t < TPW; ++t)
                        if (t < nt)
                            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(yr[ti[t] * 64], yr[tj[t] * 64], acc[t], 0,
                                                                          0, 0);
                }
            }
        }
    }  // steps
    __syncthreads();
    // upper triangles over S + 1 slots, slot-major over the CB columns (slot S: the remote
    // electrode, a zero row)
    double* G = lds;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        if (t >= nt) continue;
        const int f = 16 * tj[t] + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int e = 16 * ti[t] + (lane >> 4) + 4 * v;
            if (e < S && f < S && e <= f) G[tri_index(e, f, S + 1) * CB + grp] = acc[t][v];
        }
    }
    for (int e = gtid; e <= S; e += 256) G[tri_index(e, S, S + 1) * CB + grp] = 0.0;
    __syncthreads();
    // J[d, c0 + k] for (d, k) = (i / CB, i % CB): four slot reads of column k, three adds
    constexpr int U = 4;
    const auto q = adt::gview(A.quad);
    const auto o = adt::gview(A.out);
    const char* gb = reinterpret_cast<const char*>(G);
    const int D = A.D, total = D * CB;
    for (int i0 = tid; i0 < total; i0 += NT * U) {
        uint4 s[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * NT;
            s[u] = q[i < total ? i / CB : 0];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * NT, d = i / CB, k = i % CB;
            const uint32_t kb = static_cast<uint32_t>(k) * 8u;
            auto at = [&](uint32_t byte) { return *reinterpret_cast<const double*>(gb + byte * CB + kb); };
            const double v = ((at(s[u].x) - at(s[u].y)) - at(s[u].z)) + at(s[u].w);
            if (i < total && c0 + k < A.P) o[static_cast<int64_t>(d) * A.P + c0 + k] = -v;
        }
    }
}

template <int BT, int CB>
int launch_gram_cb(GramArgs& A, size_t lds, hipStream_t st) {
    const dim3 grid((A.P + CB - 1) / CB), block(256 * CB);
    if (A.K <= 4)
        hipLaunchKernelGGL((gram_kernel<BT, 4, CB>), grid, block, lds, st, A);
    else
        hipLaunchKernelGGL((gram_kernel<BT, 10, CB>), grid, block, lds, st, A);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

template <int BT>
int launch_gram(GramArgs& A, size_t lds, hipStream_t st) {
    return launch_gram_cb<BT, kGramColumns>(A, lds, st);
}

// Cross-Gram form for separate receiver and current fields (3D: the receiver fields are solved
// with their own right-hand sides):  G_c = Phi_R^T (T Phi_C)  (S_r x S_c per cell), then
//   J[d, c] = -((G[m, a] - G[m, b]) - G[n, a]) + G[n, b])
// with current indices a, b >= S_c and receiver indices m, n >= S_r reading zero.  No symmetry is
// needed: the full block is kept in LDS.  One workgroup per cell, chunks of wavenumbers:
//   * the chunk's receiver rows Phi_R[(w, i)][e] and current rows Phi_C[(w, i)][f] are staged
//     row by row (wave-uniform rows: the cell's nodes and the row offsets are scalar, lanes run
//     along the electrodes), kCrossRows rows per wave in flight;
//   * Psi = T Phi_C goes to its own rows, one thread per (w, i, f) entry (in place, one thread
//     per (w, f) column, the products ran on one wave for one wavenumber: 2.61 against 2.22 ms
//     for ten-node tetrahedra, 138 720 cells, 64 electrodes);
//   * the 4 waves accumulate the 16 x 16 tiles of G_c on the FP64 matrix cores
//     (v_mfma_f64_16x16x4f64, lane layout as in gram_kernel), rows padded to a multiple of 4;
// then G_c is stored as S_r x S_c (32 KB at 64 electrodes: five workgroups per CU) and each output
// is four clamped lookups, a remote electrode's reading zero by a select (a zero row and column
// in LDS left four workgroups per CU: 2.2 against 2.1 ms on ten-node tetrahedra, 138 720 cells,
// 64 electrodes).  Sums over (w, i) in order.
struct GramCrossArgs {
    const double* xr;
    const double* xc;
    const int32_t* dofs;
    const double* stiff;
    const double* mass;
    const int32_t* a;
    const int32_t* b;
    const int32_t* m;
    const int32_t* n;
    double* out;
    const int32_t* col_ptr;    // optional column aggregation, as in GramArgs
    const int32_t* col_cells;
    int Sr, Sc, K, W, rp_r, rp_c, wchunk, C, D;
    int btr, btc, rows_max;
    int P;                     // output columns (= workgroups)
    int64_t off_r[kMaxWaves];
    int64_t off_c[kMaxWaves];
    double ws[kMaxWaves];
    double wm[kMaxWaves];
};

constexpr int kCrossRows = 4;  // staged rows per wave and array in flight
constexpr int kMaxLocal = 10;

// Tiles per wave of gram_cross_kernel, rounded up to the instantiated counts.
inline int cross_tiles_per_wave(int tiles) {
    const int t = (tiles + 3) / 4;
    return t <= 1 ? 1 : t <= 2 ? 2 : t <= 4 ? 4 : t <= 8 ? 8 : 16;
}

template <int TPW>
__global__ __launch_bounds__(256) void gram_cross_kernel(GramCrossArgs A) {
    extern __shared__ double lds[];
    const int col = adt::xcd_tile(static_cast<int>(blockIdx.x), A.P);
    int c_begin, c_end;
    column_cells(A.col_ptr, col, c_begin, c_end);
    const int K = A.K, Sr = A.Sr, Sc = A.Sc;
    const int spr = 16 * A.btr, spc = 16 * A.btc, rows_max = A.rows_max;
    double* phr = lds;                                          // rows_max x spr
    double* phc = phr + static_cast<int64_t>(rows_max) * spr;   // rows_max x spc: Phi_C
    double* psi = phc + static_cast<int64_t>(rows_max) * spc;   // rows_max x spc: Psi = T Phi_C
    double* tw = psi + static_cast<int64_t>(rows_max) * spc;    // wchunk x K x K: T_wc of the chunk
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wave = (wv + static_cast<int>(blockIdx.x)) & 3;
    const int n_tiles = A.btr * A.btc;
    int ti[TPW], tj[TPW];
    int nt = 0;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        int q = wave + 4 * t;
        if (q < n_tiles) ++nt;
        else q = 0;
        ti[t] = q / A.btc;
        tj[t] = q % A.btc;
    }
    f64x4 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int ci = c_begin; ci < c_end; ++ci) {  // G of the column: its cells' G_c summed in list order
    const int c = A.col_ptr ? adt::gview(A.col_cells)[ci] : ci;
    const double* Kc = A.stiff + static_cast<int64_t>(c) * K * K;
    const double* Mc = A.mass ? A.mass + static_cast<int64_t>(c) * K * K : nullptr;
    const int32_t* dof = A.dofs + static_cast<int64_t>(c) * K;
    for (int w0 = 0; w0 < A.W; w0 += A.wchunk) {
        const int nw = min(A.wchunk, A.W - w0);
        const int rows = nw * K, rows4 = (rows + 3) & ~3;
        __syncthreads();
        // receiver and current rows of the chunk, wave-uniform; padding rows are zero
        for (int r0 = wv; r0 < rows4; r0 += 4 * kCrossRows) {
            double vr[kCrossRows][2], vc[kCrossRows][2];
#pragma unroll
            for (int u = 0; u < kCrossRows; ++u) {
                const int r = r0 + 4 * u;
                const int rc = min(r, rows - 1);  // rows past the chunk: index math only
                const int wl = rc / K, i = rc - wl * K;
                // wave-uniform row: the node index goes to a scalar register
                const int64_t node = __builtin_amdgcn_readfirstlane(adt::gview(dof)[i]);
                const double* rowr = A.xr + A.off_r[w0 + wl] + node * A.rp_r;
                const double* rowc = A.xc + A.off_c[w0 + wl] + node * A.rp_c;
#pragma unroll
                for (int h = 0; h < 2; ++h) {  // sources <= 128
                    const int e = lane + 64 * h;
                    vr[u][h] = r < rows && e < Sr ? rowr[e] : 0.0;
                    vc[u][h] = r < rows && e < Sc ? rowc[e] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < kCrossRows; ++u) {
                const int r = r0 + 4 * u;
                if (r >= rows4) break;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int e = lane + 64 * h;
                    if (e < spr) phr[r * spr + e] = vr[u][h];
                    if (e < spc) phc[r * spc + e] = vc[u][h];
                }
            }
        }
        for (int t = tid; t < nw * K * K; t += 256) {
            const int wl = t / (K * K), ij = t - wl * K * K, w = w0 + wl;
            tw[t] = Mc ? fma(A.wm[w], Mc[ij], A.ws[w] * Kc[ij]) : A.ws[w] * Kc[ij];
        }
        __syncthreads();
        // Psi[w,i,f] = sum_j T_w[i][j] Phi_C[w,j,f], j in order, one thread per (w, i, f); the
        // padding rows and columns of Psi are zero
        for (int t = tid; t < rows4 * spc; t += 256) {
            const int r = t / spc, f = t - r * spc;
            double v = 0.0;
            if (r < rows && f < Sc) {
                const int wl = r / K, ii = r - wl * K;
                const double* x = phc + wl * K * spc + f;
                const double* tr = tw + wl * K * K + ii * K;
#pragma unroll
                for (int jj = 0; jj < kMaxLocal; ++jj) {
                    if (jj >= K) break;
                    v = fma(tr[jj], x[jj * spc], v);
                }
            }
            psi[r * spc + f] = v;
        }
        __syncthreads();
        const double* ph = phr + (lane >> 4) * spr + (lane & 15);
        const double* ps = psi + (lane >> 4) * spc + (lane & 15);
        for (int r0 = 0; r0 < rows4; r0 += 4) {
#pragma unroll
            for (int t = 0; t < TPW; ++t)
                if (t < nt)
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ph[r0 * spr + 16 * ti[t]], ps[r0 * spc + 16 * tj[t]],
                                                                  acc[t], 0, 0, 0);
        }
    }
    }  // cells of the column
    __syncthreads();
    // G over (S_r + 1) x (S_c + 1) slots, receiver-major; slot S_r / S_c (remote) is zero
    const int gs = Sc;
    double* G = lds;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        if (t >= nt) continue;
        const int f = 16 * tj[t] + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int e = 16 * ti[t] + (lane >> 4) + 4 * v;
            if (e < Sr && f < Sc) G[e * gs + f] = acc[t][v];
        }
    }
    __syncthreads();
    // indices past the block (the remote electrodes) read zero: a select after a clamped read
    const unsigned ur = static_cast<unsigned>(Sr), uc = static_cast<unsigned>(Sc);
    auto g = [&](int e, int f) {
        const unsigned ue = static_cast<unsigned>(e), uf = static_cast<unsigned>(f);
        const double v = G[__umul24(min(ue, ur - 1), static_cast<unsigned>(gs)) + min(uf, uc - 1)];
        return ue < ur && uf < uc ? v : 0.0;
    };
    write_cell_column(A.a, A.b, A.m, A.n, A.D, A.P, col, A.out, g);
}

// Cell-block form for separate receiver and current fields (3D): J[d, c] =
//   -sum_{w,i} (phi_R[w,i,m] - phi_R[w,i,n]) (Psi[w,i,a] - Psi[w,i,b]),  Psi = T_w phi_C,
// over the cell's K DOFs.  A block of NT threads takes CB consecutive cells and stages, per cell,
// the receiver rows phi_R (W K x Sr) and Psi (W K x Sc) in LDS, electrode-major: the R = W K rows
// of CB cells of one electrode are R x CB consecutive doubles (cell fastest); a zero electrode
// slot stands for the remote electrode.  Staging is two passes of independent loads: the cells'
// nodes and element templates, then every (cell, w, i) row of both fields (wave-uniform rows,
// kCellBlockRows rows per wave in flight); Psi then replaces phi_C in place, one thread per
// (electrode, cell, w) column.  Output rows come out as CB consecutive doubles per quadrupole.
// With CB = 32 each thread evaluates two neighbouring cells of a quadrupole (16-byte LDS reads and
// 16-byte stores): one wave instruction reads four quadrupoles' 256-byte segments, each of which
// covers every LDS bank once (no bank conflicts whatever the electrodes).  Smaller CB (one cell
// per thread) pad each electrode slot by 8 doubles so that the 8 or 4 segments of an instruction
// spread over the banks.  With K fixed at compile time (KT > 0, one wavenumber) the R reads of one
// electrode are immediate offsets of one address.  Sums over j and over (w, i) in order.
constexpr int kAbmnChunk = 2048;  // quadrupoles whose indices are staged at a time
constexpr int kCellBlockRows = 8;

constexpr int cell_block_pad(int cb) { return cb >= 32 ? 0 : 8; }

template <int CB, int NT, int KT>
__global__ __launch_bounds__(NT) void cell_block_kernel(GramCrossArgs A) {
    extern __shared__ double lds[];
    constexpr int NW = NT / 64;  // waves per block
    const int K = KT > 0 ? KT : A.K;
    const int W = KT > 0 ? 1 : A.W;
    const int R = W * K;
    const int ES = R * CB + cell_block_pad(CB);  // doubles per electrode slot
    const int Pr = A.Sr + 1, Pc = A.Sc + 1;       // electrode slots, the last one zero
    double* phr = lds;                                      // [Pr][R][CB]
    double* psi = phr + static_cast<int64_t>(Pr) * ES;      // [Pc][R][CB] (phi_C first)
    double* tk = psi + static_cast<int64_t>(Pc) * ES;       // [CB][K][K] stiffness templates
    double* tm = tk + CB * K * K;                           // [CB][K][K] mass templates
    int32_t* nodes = reinterpret_cast<int32_t*>(tm + CB * K * K);  // [CB][K]
    // electrode indices of a chunk of quadrupoles, one packed word each (a | b << 8 | m << 16 | n << 24)
    uint32_t* abmn = reinterpret_cast<uint32_t*>(nodes + ((CB * K + 1) & ~1));
    const int nblk = (A.C + CB - 1) / CB;
    const int c0 = adt::xcd_tile(static_cast<int>(blockIdx.x), nblk) * CB;
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    // staging pass 1: nodes and templates of the block's cells (cells past C repeat cell C - 1 and are
    // never stored)
    {
        const int64_t first = static_cast<int64_t>(c0) * K, last = static_cast<int64_t>(A.C) * K - 1;
        for (int t = tid; t < CB * K; t += NT) nodes[t] = A.dofs[min(first + t, last - (K - 1) + t % K)];
        const int64_t tfirst = static_cast<int64_t>(c0) * K * K, tlast = static_cast<int64_t>(A.C) * K * K;
        for (int t = tid; t < CB * K * K; t += NT) {
            const int64_t q = tfirst + t < tlast ? tfirst + t : tlast - K * K + t % (K * K);
            tk[t] = A.stiff[q];
            tm[t] = A.mass ? A.mass[q] : 0.0;
        }
    }
    __syncthreads();
    // staging pass 2: row (cb, w, i) of phi_R (rows 0 .. CB R - 1) and of phi_C (rows CB R .. 2 CB R - 1),
    // lanes along the electrodes (S <= 128)
    const int n_rows = 2 * CB * R;
    for (int q0 = wv; q0 < n_rows; q0 += NW * kCellBlockRows) {
        double v[kCellBlockRows][2];
#pragma unroll
        for (int u = 0; u < kCellBlockRows; ++u) {
            const int q = min(q0 + NW * u, n_rows - 1);  // rows past the end: index math only
            const bool recv = q < CB * R;
            const int qq = recv ? q : q - CB * R;
            const int cb = qq % CB, r = qq / CB, w = r / K, i = r - w * K;
            const int S = recv ? A.Sr : A.Sc;
            const double* row = recv ? A.xr + A.off_r[w] + static_cast<int64_t>(nodes[cb * K + i]) * A.rp_r
                                     : A.xc + A.off_c[w] + static_cast<int64_t>(nodes[cb * K + i]) * A.rp_c;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = lane + 64 * h;
                v[u][h] = q0 + NW * u < n_rows && e < S ? row[e] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < kCellBlockRows; ++u) {
            const int q = q0 + NW * u;
            if (q >= n_rows) break;
            const bool recv = q < CB * R;
            const int qq = recv ? q : q - CB * R;
            const int cb = qq % CB, r = qq / CB;
            double* dst = (recv ? phr : psi) + r * CB + cb;
            const int S = recv ? A.Sr : A.Sc;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = lane + 64 * h;
                if (e < S) dst[e * ES] = v[u][h];
            }
        }
    }
    for (int t = tid; t < R * CB; t += NT) {  // the remote electrode's zero slots
        phr[A.Sr * ES + t] = 0.0;
        psi[A.Sc * ES + t] = 0.0;
    }
    __syncthreads();
    // Psi[w,i,f] = sum_j (ws_w K_c[i][j] + wm_w M_c[i][j]) phi_C[w,j,f], j in order, in place:
    // each thread reads its column's K values before writing the K results
    for (int t = tid; t < A.Sc * CB * W; t += NT) {
        const int f = t % A.Sc, wc = t / A.Sc, cb = wc % CB, w = wc / CB;
        double* col = psi + f * ES + (w * K) * CB + cb;
        const double* kc = tk + cb * K * K;
        const double* mc = tm + cb * K * K;
        const double ws = A.ws[w], wm = A.wm[w];
        double x[KT > 0 ? KT : kMaxLocal];
#pragma unroll
        for (int jj = 0; jj < (KT > 0 ? KT : kMaxLocal); ++jj)
            if (jj < K) x[jj] = col[jj * CB];
#pragma unroll
        for (int ii = 0; ii < (KT > 0 ? KT : kMaxLocal); ++ii) {
            if (ii >= K) break;
            double v = 0.0;
#pragma unroll
            for (int jj = 0; jj < (KT > 0 ? KT : kMaxLocal); ++jj) {
                if (jj >= K) break;
                const double tij = A.mass ? fma(wm, mc[ii * K + jj], ws * kc[ii * K + jj]) : ws * kc[ii * K + jj];
                v = fma(tij, x[jj], v);
            }
            col[ii * CB] = v;
        }
    }
    constexpr int RPP = NT / CB;  // quadrupoles per pass
    const int cb = tid % CB, dl = tid / CB;
    const int cell = c0 + cb;
    const auto ga = adt::gview(A.a), gb = adt::gview(A.b), gm = adt::gview(A.m), gn = adt::gview(A.n);
    const auto out = adt::gview(A.out);
    for (int dc0 = 0; dc0 < A.D; dc0 += kAbmnChunk) {
        const int nd = min(kAbmnChunk, A.D - dc0);
        __syncthreads();  // the previous chunk's indices are consumed (first chunk: Psi done)
        constexpr int PER = (kAbmnChunk + NT - 1) / NT;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int k = tid + u * NT;
            if (k < nd) {
                const int d = dc0 + k;
                // indices >= S name the remote electrode: the zero slot
                const uint32_t ia = min(max(ga[d], 0), A.Sc), ib = min(max(gb[d], 0), A.Sc);
                const uint32_t im = min(max(gm[d], 0), A.Sr), in = min(max(gn[d], 0), A.Sr);
                abmn[k] = ia | ib << 8 | im << 16 | in << 24;
            }
        }
        __syncthreads();
        if constexpr (CB >= 32) {
            // two neighbouring cells per thread: 16-byte LDS reads and stores
            constexpr int HP = CB / 2, RPP2 = NT / HP;
            const int cp = tid % HP, dl2 = tid / HP;
            const int cell0 = c0 + 2 * cp;
            const bool even = (A.C & 1) == 0 && (reinterpret_cast<uintptr_t>(A.out) & 15) == 0;
            for (int k = dl2; k < nd; k += RPP2) {
                const uint32_t wd = abmn[k];
                const double2* pm = reinterpret_cast<const double2*>(phr + ((wd >> 16) & 255) * ES) + cp;
                const double2* pn = reinterpret_cast<const double2*>(phr + (wd >> 24) * ES) + cp;
                const double2* pa = reinterpret_cast<const double2*>(psi + (wd & 255) * ES) + cp;
                const double2* pb = reinterpret_cast<const double2*>(psi + ((wd >> 8) & 255) * ES) + cp;
                double a0 = 0.0, a1 = 0.0;
                if constexpr (KT > 0) {
#pragma unroll
                    for (int r = 0; r < KT; ++r) {
                        const double2 m2 = pm[r * HP], n2 = pn[r * HP], x2 = pa[r * HP], y2 = pb[r * HP];
                        a0 = fma(m2.x - n2.x, x2.x - y2.x, a0);
                        a1 = fma(m2.y - n2.y, x2.y - y2.y, a1);
                    }
                } else {
                    for (int r = 0; r < R; ++r) {
                        const double2 m2 = pm[r * HP], n2 = pn[r * HP], x2 = pa[r * HP], y2 = pb[r * HP];
                        a0 = fma(m2.x - n2.x, x2.x - y2.x, a0);
                        a1 = fma(m2.y - n2.y, x2.y - y2.y, a1);
                    }
                }
                double* o = A.out + static_cast<int64_t>(dc0 + k) * A.C + cell0;
                if (even && cell0 + 1 < A.C) {
                    *reinterpret_cast<double2*>(o) = make_double2(-a0, -a1);
                } else {
                    if (cell0 < A.C) o[0] = -a0;
                    if (cell0 + 1 < A.C) o[1] = -a1;
                }
            }
        } else {
            for (int k = dl; k < nd; k += RPP) {
                const uint32_t wd = abmn[k];
                const double* pm = phr + ((wd >> 16) & 255) * ES + cb;
                const double* pn = phr + (wd >> 24) * ES + cb;
                const double* pa = psi + (wd & 255) * ES + cb;
                const double* pb = psi + ((wd >> 8) & 255) * ES + cb;
                double acc = 0.0;
                if constexpr (KT > 0) {
#pragma unroll
                    for (int r = 0; r < KT; ++r) acc = fma(pm[r * CB] - pn[r * CB], pa[r * CB] - pb[r * CB], acc);
                } else {
                    for (int r = 0; r < R; ++r) acc = fma(pm[r * CB] - pn[r * CB], pa[r * CB] - pb[r * CB], acc);
                }
                if (cell < A.C) out[static_cast<int64_t>(dc0 + k) * A.C + cell] = -acc;
            }
        }
    }
}

// LDS bytes of cell_block_kernel for CB cells per block.
inline size_t cell_block_lds(const GramCrossArgs& A, int cb) {
    const size_t es = static_cast<size_t>(A.W) * A.K * cb + cell_block_pad(cb);
    const size_t templates = static_cast<size_t>(2) * cb * A.K * A.K * 8;
    const size_t nodes = static_cast<size_t>((cb * A.K + 1) & ~1) * 4;
    return (static_cast<size_t>(A.Sr + 1) + (A.Sc + 1)) * es * 8 + templates + nodes + kAbmnChunk * 4;
}

// Cells per block of cell_block_kernel: 32 (1024 threads, conflict-free reads) where its LDS fits
// one block per CU, else 16/8/4/2 (256 threads) within 64 KB; 0: none.  ADT_CELL_BLOCK set below
// 32 (e.g. 8) leaves out the 32-cell width.
inline int cell_block_width(const GramCrossArgs& A) {
    const char* cap = std::getenv("ADT_CELL_BLOCK");
    if (!(cap && std::atoi(cap) < 32) && cell_block_lds(A, 32) <= 160 * 1024) return 32;
    for (int cb = 16; cb >= 2; cb /= 2)
        if (cell_block_lds(A, cb) <= 65536) return cb;
    return 0;
}

template <int CB, int NT>
int launch_cell_block(GramCrossArgs& A, hipStream_t st) {
    const size_t lds = cell_block_lds(A, CB);
    const dim3 grid((A.C + CB - 1) / CB);
    if (A.W == 1 && A.K == 4)  // 3D P1, one field set
        hipLaunchKernelGGL((cell_block_kernel<CB, NT, 4>), grid, dim3(NT), lds, st, A);
    else
        hipLaunchKernelGGL((cell_block_kernel<CB, NT, 0>), grid, dim3(NT), lds, st, A);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

template <int TPW>
int launch_gram_cross(GramCrossArgs& A, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL(gram_cross_kernel<TPW>, dim3(A.P), dim3(256), lds, st, A);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

// ------------------------------------------------------------------------------------------
// (d) adjoint cell gradient on interleaved (N, Rp) fields: one wavefront per cell, lanes over
// right-hand sides, every row access a contiguous Rp*8-byte run.
// ------------------------------------------------------------------------------------------

struct GradArgs {
    const double* lam;
    const double* phi;
    const int32_t* dofs;
    const double* stiff;
    const double* mass;
    const int32_t* bnd_ptr;   // per cell: its boundary edges (CSR), fixed order
    const int32_t* bnd_list;
    const int32_t* bnd_dofs;  // (n_bnd, kb)
    const double* bmass;      // (n_bnd, kb, kb)
    const double* balpha;     // (W, n_bnd)
    double* out;
    int W, rp_lam, rp_phi, r, C, kb, n_bnd;
    int64_t lam_off[kMaxWaves];
    int64_t phi_off[kMaxWaves];
    double ws[kMaxWaves];
    double wm[kMaxWaves];
    double bws[kMaxWaves];
};

template <int K>
__global__ __launch_bounds__(256) void cell_gradient_kernel(GradArgs A) {
    const int c = blockIdx.x * 4 + static_cast<int>(threadIdx.x >> 6);
    const int lane = static_cast<int>(threadIdx.x & 63);
    if (c >= A.C) return;
    int dof[K];
#pragma unroll
    for (int k = 0; k < K; ++k) dof[k] = A.dofs[static_cast<int64_t>(c) * K + k];
    const double* Kc = A.stiff + static_cast<int64_t>(c) * K * K;
    const double* Mc = A.mass ? A.mass + static_cast<int64_t>(c) * K * K : nullptr;
    double acc = 0.0;
    for (int w = 0; w < A.W; ++w) {
        const double* L = A.lam + A.lam_off[w];
        const double* F = A.phi + A.phi_off[w];
        double s_k = 0.0, s_m = 0.0;
        for (int s = lane; s < A.r; s += 64) {
            double lv[K], pv[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                lv[k] = L[static_cast<int64_t>(dof[k]) * A.rp_lam + s];
                pv[k] = F[static_cast<int64_t>(dof[k]) * A.rp_phi + s];
            }
#pragma unroll
            for (int i = 0; i < K; ++i) {
                double yk = 0.0, ym = 0.0;
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    yk = fma(Kc[i * K + j], pv[j], yk);
                    if (Mc) ym = fma(Mc[i * K + j], pv[j], ym);
                }
                s_k = fma(lv[i], yk, s_k);
                s_m = fma(lv[i], ym, s_m);
            }
        }
        acc = fma(A.ws[w], s_k, acc);
        acc = fma(A.wm[w], s_m, acc);
    }
    if (A.bnd_ptr) {
        for (int e = A.bnd_ptr[c]; e < A.bnd_ptr[c + 1]; ++e) {
            const int b = A.bnd_list[e];
            const int32_t* bd = A.bnd_dofs + static_cast<int64_t>(b) * A.kb;
            const double* Mb = A.bmass + static_cast<int64_t>(b) * A.kb * A.kb;
            for (int w = 0; w < A.W; ++w) {
                const double* L = A.lam + A.lam_off[w];
                const double* F = A.phi + A.phi_off[w];
                const double coef = A.bws[w] * A.balpha[static_cast<int64_t>(w) * A.n_bnd + b];
                double sb = 0.0;
                for (int s = lane; s < A.r; s += 64)
                    for (int i = 0; i < A.kb; ++i) {
                        const double li = L[static_cast<int64_t>(bd[i]) * A.rp_lam + s];
                        for (int j = 0; j < A.kb; ++j)
                            sb = fma(li * Mb[i * A.kb + j], F[static_cast<int64_t>(bd[j]) * A.rp_phi + s], sb);
                    }
                acc = fma(coef, sb, acc);
            }
        }
    }
    acc = adt::wave_sum(acc);
    if (lane == 0) A.out[c] = -acc;
}

template <int K>
int launch_gradient(const GradArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(cell_gradient_kernel<K>, dim3(adt::blocks_for(a.C, 4)), dim3(256), 0, st, a);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

// ------------------------------------------------------------------------------------------
// electrode potentials, deinterleave, column aggregation
// ------------------------------------------------------------------------------------------

__global__ void potentials_kernel(const double* __restrict__ x, const int64_t* __restrict__ vec_off,
                                  const int32_t* __restrict__ enode, const double* __restrict__ eweight,
                                  const double* __restrict__ q, double* __restrict__ p, int n_models, int W,
                                  int rp, int r, int n_recv, int kmax) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    const int t = blockIdx.z;
    if (s >= r) return;
    double acc = 0.0;
    for (int w = 0; w < W; ++w) {
        const double* X = x + vec_off[static_cast<int64_t>(t) * W + w];
        double v = 0.0;
        for (int k = 0; k < kmax; ++k) {
            const double wt = eweight[static_cast<int64_t>(f) * kmax + k];
            if (wt != 0.0) v = fma(wt, X[static_cast<int64_t>(enode[static_cast<int64_t>(f) * kmax + k]) * rp + s], v);
        }
        acc = fma(q[w], v, acc);
    }
    p[(static_cast<int64_t>(t) * r + s) * n_recv + f] = acc;
}

__global__ void deinterleave_kernel(const double* __restrict__ x, const int64_t* __restrict__ vec_off,
                                    double* __restrict__ out, int rp, int r, int64_t n, int64_t stride) {
    __shared__ double tile[64][65];
    const int sys = blockIdx.z;
    const int64_t n0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int s0 = blockIdx.y * 64;
    const double* X = x + vec_off[sys];
    double* O = out + static_cast<int64_t>(sys) * stride;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {
        const int64_t node = n0 + i;
        const int s = s0 + tx;
        tile[i][tx] = (node < n && s < r) ? X[node * rp + s] : 0.0;
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int s = s0 + i;
        const int64_t node = n0 + tx;
        if (s < r && node < n) O[static_cast<int64_t>(s) * n + node] = tile[tx][i];
    }
}

__global__ void aggregate_kernel(const double* __restrict__ in, int64_t in_cols, const int32_t* __restrict__ ptr,
                                 const int32_t* __restrict__ idx, double* __restrict__ out, int n_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int d = blockIdx.y;
    if (p >= n_out) return;
    const double* row = in + static_cast<int64_t>(d) * in_cols;
    double acc = 0.0;
    for (int i = ptr[p]; i < ptr[p + 1]; ++i) acc += row[idx[i]];
    out[static_cast<int64_t>(d) * n_out + p] = acc;
}

// small device scratch for per-call offset tables (grow-only, per host thread)
struct Scratch {
    void* ptr = nullptr;
    size_t cap = 0;
};

int scratch_upload(Scratch& s, const void* host, size_t bytes, hipStream_t st) {
    if (s.cap < bytes) {
        if (s.ptr) ADT_HIP_CHECK(hipFree(s.ptr));
        ADT_HIP_CHECK(hipMalloc(&s.ptr, bytes));
        s.cap = bytes;
    }
    ADT_HIP_CHECK(hipMemcpyAsync(s.ptr, host, bytes, hipMemcpyHostToDevice, st));
    return ADT_OK;
}

}  // namespace

extern "C" int adt_sensitivity(int n_waves, int s_rows, int64_t n_dofs, int n_meas, int n_cells, int k_local,
                               const double* d_phi, const int32_t* d_a, const int32_t* d_b, const int32_t* d_m,
                               const int32_t* d_n, const int32_t* d_cell_dofs, const double* d_stiff,
                               const double* d_mass, const double* h_ws, const double* h_wm, int precision,
                               double* d_out, void* stream) {
    ADT_REQUIRE(n_waves >= 1 && n_waves <= kMaxWaves, "adt_sensitivity: 1 <= n_waves <= %d", kMaxWaves);
    ADT_REQUIRE(precision == 0 || precision == 1, "adt_sensitivity: precision must be 0 or 1");
    ADT_REQUIRE(h_ws && (d_mass == nullptr || h_wm), "adt_sensitivity: missing wavenumber weights");
    if (n_meas == 0 || n_cells == 0) return ADT_OK;
    ADT_REQUIRE(d_phi && d_a && d_b && d_m && d_n && d_cell_dofs && d_stiff && d_out, "adt_sensitivity: null pointer");
    SensArgs a{};
    a.phi = d_phi;
    a.a = d_a;
    a.b = d_b;
    a.m = d_m;
    a.n = d_n;
    a.dofs = d_cell_dofs;
    a.stiff = d_stiff;
    a.mass = d_mass;
    a.out = d_out;
    a.n_dofs = n_dofs;
    a.W = n_waves;
    a.S = s_rows;
    a.D = n_meas;
    a.C = n_cells;
    for (int w = 0; w < n_waves; ++w) {
        a.ws[w] = h_ws[w];
        a.wm[w] = d_mass ? h_wm[w] : 0.0;
    }
    // enough workgroups to fill 256 CUs several times; a multiple of 4 measurements each
    const int cell_blocks = adt::blocks_for(n_cells, 64);
    int chunks = std::max(1, (4096 + cell_blocks - 1) / cell_blocks);
    int dch = (n_meas + chunks - 1) / chunks;
    dch = std::max(4, (dch + 3) / 4 * 4);
    a.dch = dch;
    ADT_REQUIRE(adt::blocks_for(n_meas, dch) <= 65535, "adt_sensitivity: too many measurement chunks");
    hipStream_t st = adt::as_stream(stream);
    switch (k_local) {
        case 3: return launch_sensitivity<3>(a, precision, st);
        case 4: return launch_sensitivity<4>(a, precision, st);
        case 6: return launch_sensitivity<6>(a, precision, st);
        case 10: return launch_sensitivity<10>(a, precision, st);
        default: return adt::set_error(ADT_ERR_INVALID, "adt_sensitivity: unsupported local size %d", k_local);
    }
}

extern "C" int adt_cell_gradient(int n_waves, int r, int n_cells, int k_local, const double* d_lam, int rp_lam,
                                 const int64_t* h_lam_off, const double* d_phi, int rp_phi,
                                 const int64_t* h_phi_off, const int32_t* d_cell_dofs, const double* d_stiff,
                                 const double* d_mass, const double* h_ws, const double* h_wm, int n_bnd, int k_bnd,
                                 const int32_t* d_bnd_dofs, const int32_t* d_cell_bnd, const double* d_bmass,
                                 const double* d_balpha, const double* h_bws, double* d_out, void* stream) {
    ADT_REQUIRE(n_waves >= 1 && n_waves <= kMaxWaves, "adt_cell_gradient: 1 <= n_waves <= %d", kMaxWaves);
    ADT_REQUIRE(r >= 1 && rp_lam >= r && rp_phi >= r, "adt_cell_gradient: invalid r/rp");
    if (n_cells == 0) return ADT_OK;
    ADT_REQUIRE(d_lam && d_phi && d_cell_dofs && d_stiff && d_out && h_ws && h_lam_off && h_phi_off,
                "adt_cell_gradient: null pointer");
    ADT_REQUIRE(d_mass == nullptr || h_wm, "adt_cell_gradient: mass needs wm");
    GradArgs a{};
    a.lam = d_lam;
    a.phi = d_phi;
    a.dofs = d_cell_dofs;
    a.stiff = d_stiff;
    a.mass = d_mass;
    a.out = d_out;
    a.W = n_waves;
    a.rp_lam = rp_lam;
    a.rp_phi = rp_phi;
    a.r = r;
    a.C = n_cells;
    a.kb = k_bnd;
    a.n_bnd = n_bnd;
    if (n_bnd > 0) {
        ADT_REQUIRE(d_bnd_dofs && d_cell_bnd && d_bmass && d_balpha && h_bws && k_bnd >= 1,
                    "adt_cell_gradient: boundary arrays missing");
        a.bnd_ptr = d_cell_bnd;
        a.bnd_list = d_cell_bnd + n_cells + 1;
        a.bnd_dofs = d_bnd_dofs;
        a.bmass = d_bmass;
        a.balpha = d_balpha;
    }
    for (int w = 0; w < n_waves; ++w) {
        a.lam_off[w] = h_lam_off[w];
        a.phi_off[w] = h_phi_off[w];
        a.ws[w] = h_ws[w];
        a.wm[w] = d_mass ? h_wm[w] : 0.0;
        a.bws[w] = n_bnd > 0 ? h_bws[w] : 0.0;
    }
    hipStream_t st = adt::as_stream(stream);
    switch (k_local) {
        case 3: return launch_gradient<3>(a, st);
        case 4: return launch_gradient<4>(a, st);
        case 6: return launch_gradient<6>(a, st);
        case 10: return launch_gradient<10>(a, st);
        default: return adt::set_error(ADT_ERR_INVALID, "adt_cell_gradient: unsupported local size %d", k_local);
    }
}

extern "C" int adt_electrode_potentials(int n_models, int n_waves, int rp, int r, int n_recv, int kmax,
                                        const int64_t* h_vec_off, const double* d_x, const int32_t* d_enode,
                                        const double* d_eweight, const double* h_q, double* d_p, void* stream) {
    ADT_REQUIRE(n_models >= 0 && n_waves >= 1 && r >= 1 && rp >= r && n_recv >= 0 && kmax >= 1,
                "adt_electrode_potentials: invalid sizes");
    if (n_models == 0 || n_recv == 0) return ADT_OK;
    ADT_REQUIRE(n_recv <= 65535 && n_models <= 65535, "adt_electrode_potentials: too many rows");
    hipStream_t st = adt::as_stream(stream);
    static thread_local Scratch offs, qs;
    const size_t nsys = static_cast<size_t>(n_models) * n_waves;
    int rc = scratch_upload(offs, h_vec_off, sizeof(int64_t) * nsys, st);
    if (rc) return rc;
    if ((rc = scratch_upload(qs, h_q, sizeof(double) * n_waves, st))) return rc;
    dim3 grid(adt::blocks_for(r, 64), n_recv, n_models);
    hipLaunchKernelGGL(potentials_kernel, grid, dim3(64), 0, st, d_x, static_cast<const int64_t*>(offs.ptr),
                       d_enode, d_eweight, static_cast<const double*>(qs.ptr), d_p, n_models, n_waves, rp, r,
                       n_recv, kmax);
    ADT_HIP_CHECK(hipGetLastError());
    ADT_HIP_CHECK(hipStreamSynchronize(st));  // scratch reuse
    return ADT_OK;
}

extern "C" int adt_deinterleave(int n_sys, int rp, int r, int64_t n, const int64_t* h_vec_off, const double* d_x,
                                double* d_out, int64_t out_stride, void* stream) {
    ADT_REQUIRE(n_sys >= 0 && r >= 1 && rp >= r && n >= 0, "adt_deinterleave: invalid sizes");
    if (n_sys == 0 || n == 0) return ADT_OK;
    ADT_REQUIRE(n_sys <= 65535, "adt_deinterleave: too many systems");
    hipStream_t st = adt::as_stream(stream);
    static thread_local Scratch offs;
    int rc = scratch_upload(offs, h_vec_off, sizeof(int64_t) * n_sys, st);
    if (rc) return rc;
    dim3 grid(static_cast<unsigned>((n + 63) / 64), adt::blocks_for(r, 64), n_sys);
    hipLaunchKernelGGL(deinterleave_kernel, grid, dim3(256), 0, st, d_x, static_cast<const int64_t*>(offs.ptr),
                       d_out, rp, r, n, out_stride);
    ADT_HIP_CHECK(hipGetLastError());
    ADT_HIP_CHECK(hipStreamSynchronize(st));
    return ADT_OK;
}

extern "C" int adt_aggregate_columns(int n_rows, int64_t in_cols, int n_out, const double* d_in,
                                     const int32_t* d_col_ptr, const int32_t* d_col_idx, double* d_out,
                                     void* stream) {
    ADT_REQUIRE(n_rows >= 0 && n_out >= 0 && in_cols >= 0, "adt_aggregate_columns: invalid sizes");
    if (n_rows == 0 || n_out == 0) return ADT_OK;
    ADT_REQUIRE(n_rows <= 65535, "adt_aggregate_columns: at most 65535 rows per call");
    dim3 grid(adt::blocks_for(n_out, 256), n_rows);
    hipLaunchKernelGGL(aggregate_kernel, grid, dim3(256), 0, adt::as_stream(stream), d_in, in_cols, d_col_ptr,
                       d_col_idx, d_out, n_out);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

extern "C" int adt_sensitivity_gram_cross(int n_waves, const double* d_xr, int rp_r, int s_r, const int64_t* h_off_r,
                                          const double* d_xc, int rp_c, int s_c, const int64_t* h_off_c, int n_meas,
                                          int n_cells, int k_local, const int32_t* d_a, const int32_t* d_b,
                                          const int32_t* d_m, const int32_t* d_n, const int32_t* d_cell_dofs,
                                          const double* d_stiff, const double* d_mass, const double* h_ws,
                                          const double* h_wm, const int32_t* d_col_ptr, const int32_t* d_col_cells,
                                          int n_cols, double* d_out, void* stream) {
    ADT_REQUIRE(n_waves >= 1 && n_waves <= kMaxWaves, "adt_sensitivity_gram_cross: 1 <= n_waves <= %d", kMaxWaves);
    ADT_REQUIRE(!d_col_ptr || (d_col_cells && n_cols >= 0),
                "adt_sensitivity_gram_cross: column map needs its cell list");
    ADT_REQUIRE(s_r >= 1 && s_r <= 128 && s_r <= rp_r && s_c >= 1 && s_c <= 128 && s_c <= rp_c,
                "adt_sensitivity_gram_cross: 1 <= receivers <= min(128, rp_r) and 1 <= currents <= min(128, rp_c)");
    ADT_REQUIRE(k_local >= 1 && k_local <= 10, "adt_sensitivity_gram_cross: unsupported local size %d", k_local);
    ADT_REQUIRE(h_off_r && h_off_c && h_ws && (d_mass == nullptr || h_wm),
                "adt_sensitivity_gram_cross: missing host arrays");
    const int cols = d_col_ptr ? n_cols : n_cells;
    if (n_meas == 0 || n_cells == 0 || cols == 0) return ADT_OK;
    ADT_REQUIRE(d_xr && d_xc && d_a && d_b && d_m && d_n && d_cell_dofs && d_stiff && d_out,
                "adt_sensitivity_gram_cross: null pointer");
    GramCrossArgs A{};
    A.col_ptr = d_col_ptr;
    A.col_cells = d_col_cells;
    A.P = cols;
    A.xr = d_xr;
    A.xc = d_xc;
    A.dofs = d_cell_dofs;
    A.stiff = d_stiff;
    A.mass = d_mass;
    A.a = d_a;
    A.b = d_b;
    A.m = d_m;
    A.n = d_n;
    A.out = d_out;
    A.Sr = s_r;
    A.Sc = s_c;
    A.btr = (s_r + 15) / 16;
    A.btc = (s_c + 15) / 16;
    A.K = k_local;
    A.W = n_waves;
    A.rp_r = rp_r;
    A.rp_c = rp_c;
    A.C = n_cells;
    A.D = n_meas;
    // Gram form: the staged receiver and current rows of a wavenumber chunk fill at most the G
    // block's LDS (or 16 KB), which the G block reuses afterwards
    const size_t block = static_cast<size_t>(s_r) * s_c * 8;
    const size_t row_bytes = static_cast<size_t>(16) * (A.btr + 2 * A.btc) * 8;  // Phi_R, Phi_C and Psi rows
    const size_t budget = std::max<size_t>(block, 16384);
    A.wchunk = std::max(1, std::min(n_waves, static_cast<int>(budget / row_bytes) / k_local));
    A.rows_max = (A.wchunk * k_local + 3) & ~3;
    const size_t staging = A.rows_max * row_bytes + static_cast<size_t>(A.wchunk) * k_local * k_local * 8;
    const size_t lds = std::max(staging, block);
    for (int w = 0; w < n_waves; ++w) {
        A.off_r[w] = h_off_r[w];
        A.off_c[w] = h_off_c[w];
        A.ws[w] = h_ws[w];
        A.wm[w] = d_mass ? h_wm[w] : 0.0;
    }
    hipStream_t st = adt::as_stream(stream);
    // Form (kernel times per call, 2000 quadrupoles, one field set unless noted): the cell block
    // reads 4 W K staged values per output, the Gram block 4 (plus the per-cell G block of
    // (S_r + 1) (S_c + 1) doubles).  The cell block is taken where W K <= 4 (linear tetrahedra:
    // 1.15M cells, 64 / 96 / 128 electrodes 6.3 / 12.8 / 15.3 ms against 21.6 / 29.1 / 58.5 ms for
    // the Gram block) or where 32 cells fit one CU; otherwise the Gram block (ten-node
    // tetrahedra, 138 720 cells, 32 / 64 / 96 electrodes: 2.3 / 2.2 / 4.0 ms against
    // 3.0 / 4.1 / 4.0 ms; two field sets of linear tetrahedra 6.7 against 5.9 ms, three 7.5
    // against 30 ms).
    // ADT_CROSS_FORM=gram forces the Gram block, ADT_CROSS_FORM=block the widest cell block
    // that fits (ADT_CELL_BLOCK=8 caps it below 32 cells).
    const char* form = std::getenv("ADT_CROSS_FORM");
    const bool force_gram = (form && std::strcmp(form, "gram") == 0) || d_col_ptr;  // columns of many cells
    const bool force_block = form && std::strcmp(form, "block") == 0;
    const int width = force_gram ? 0 : cell_block_width(A);
    if (width == 32 || (width > 0 && (force_block || n_waves * k_local <= 4))) {
        switch (width) {
            case 32: return launch_cell_block<32, 1024>(A, st);
            case 16: return launch_cell_block<16, 256>(A, st);
            case 8: return launch_cell_block<8, 256>(A, st);
            case 4: return launch_cell_block<4, 256>(A, st);
            case 2: return launch_cell_block<2, 256>(A, st);
            default: break;
        }
    }
    switch (cross_tiles_per_wave(A.btr * A.btc)) {
        case 1: return launch_gram_cross<1>(A, lds, st);
        case 2: return launch_gram_cross<2>(A, lds, st);
        case 4: return launch_gram_cross<4>(A, lds, st);
        case 8: return launch_gram_cross<8>(A, lds, st);
        default: return launch_gram_cross<16>(A, lds, st);
    }
}

extern "C" int adt_gram_quad_offsets(int n_meas, int s_rows, const int32_t* d_a, const int32_t* d_b,
                                     const int32_t* d_m, const int32_t* d_n, uint32_t* d_quad, void* stream) {
    ADT_REQUIRE(n_meas >= 0 && s_rows >= 1 && s_rows <= 128, "adt_gram_quad_offsets: invalid sizes");
    if (n_meas == 0) return ADT_OK;
    ADT_REQUIRE(d_a && d_b && d_m && d_n && d_quad, "adt_gram_quad_offsets: null pointer");
    ADT_REQUIRE(reinterpret_cast<uintptr_t>(d_quad) % 16 == 0, "adt_gram_quad_offsets: d_quad must be 16-byte aligned");
    hipLaunchKernelGGL(quad_offsets_kernel, dim3(adt::blocks_for(n_meas, 256)), dim3(256), 0, adt::as_stream(stream),
                       n_meas, s_rows, d_a, d_b, d_m, d_n, reinterpret_cast<uint4*>(d_quad));
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

extern "C" int adt_sensitivity_gram(int n_waves, int rp, int s_rows, const int64_t* h_vec_off, const double* d_x,
                                    int n_meas, int n_cells, int k_local, const uint32_t* d_quad, const int32_t* d_cell_dofs,
                                    const double* d_stiff, const double* d_mass, const double* h_ws,
                                    const double* h_wm, const int32_t* d_col_ptr, const int32_t* d_col_cells,
                                    int n_cols, double* d_out, void* stream) {
    ADT_REQUIRE(n_waves >= 1 && n_waves <= kMaxWaves, "adt_sensitivity_gram: 1 <= n_waves <= %d", kMaxWaves);
    ADT_REQUIRE(!d_col_ptr || (d_col_cells && n_cols >= 0), "adt_sensitivity_gram: column map needs its cell list");
    ADT_REQUIRE(s_rows >= 1 && s_rows <= 128 && s_rows <= rp, "adt_sensitivity_gram: 1 <= sources <= min(128, rp)");
    ADT_REQUIRE(k_local >= 1 && k_local <= 10, "adt_sensitivity_gram: unsupported local size %d", k_local);
    ADT_REQUIRE(h_vec_off && h_ws && (d_mass == nullptr || h_wm), "adt_sensitivity_gram: missing host arrays");
    for (int w = 0; w < n_waves; ++w)
        ADT_REQUIRE(h_ws[w] >= 0.0 && (d_mass == nullptr || h_wm[w] >= 0.0),
                    "adt_sensitivity_gram: the weights must be non-negative (the Gram form factors T = R^T R)");
    const int cols = d_col_ptr ? n_cols : n_cells;
    if (n_meas == 0 || n_cells == 0 || cols == 0) return ADT_OK;
    ADT_REQUIRE(d_x && d_quad && d_cell_dofs && d_stiff && d_out, "adt_sensitivity_gram: null pointer");
    ADT_REQUIRE(reinterpret_cast<uintptr_t>(d_quad) % 16 == 0, "adt_sensitivity_gram: d_quad must be 16-byte aligned");
    GramArgs A{};
    A.col_ptr = d_col_ptr;
    A.col_cells = d_col_cells;
    A.P = cols;
    A.x = d_x;
    A.dofs = d_cell_dofs;
    A.stiff = d_stiff;
    A.mass = d_mass;
    A.quad = reinterpret_cast<const uint4*>(d_quad);
    A.out = d_out;
    A.S = s_rows;
    const int B = (s_rows + 15) / 16;
    A.Sp = 16 * B;
    A.K = k_local;
    A.W = n_waves;
    A.rp = rp;
    A.C = n_cells;
    A.D = n_meas;
    for (int w = 0; w < n_waves; ++w) {
        A.vec_off[w] = h_vec_off[w];
        A.ws[w] = h_ws[w];
        A.wm[w] = d_mass ? h_wm[w] : 0.0;
    }
    hipStream_t st = adt::as_stream(stream);
    // Each column group of 4 waves has a 40 KB share of LDS (four groups per CU) for the factors R_w
    // and the Y rows of a wavenumber chunk (all wavenumbers at once where they fit); the
    // triangles reuse the area afterwards (a larger share where one triangle needs more).
    const size_t tri = static_cast<size_t>(s_rows + 1) * (s_rows + 2) / 2;  // slots, with the remote one
    const size_t factors = static_cast<size_t>(n_waves) * k_local * k_local;
    A.factor_doubles = static_cast<int>(factors);
    const size_t share = std::max<size_t>(40960 / 8, tri);  // doubles per group
    ADT_REQUIRE(factors + static_cast<size_t>((k_local + 3) & ~3) * A.Sp <= share,
                "adt_sensitivity_gram: %d wavenumbers x local size %d do not fit the LDS", n_waves, k_local);
    A.wchunk = std::min(n_waves, static_cast<int>((share - factors) / (static_cast<size_t>(k_local) * A.Sp)));
    while (A.wchunk > 1 && factors + static_cast<size_t>((A.wchunk * k_local + 3) & ~3) * A.Sp > share) --A.wchunk;
    constexpr int cb = kGramColumns;
    A.rows_max = (A.wchunk * k_local + 3) & ~3;
    A.group_doubles = static_cast<int>(factors + static_cast<size_t>(A.rows_max) * A.Sp);
    const size_t lds = static_cast<size_t>(cb) * std::max(static_cast<size_t>(A.group_doubles), tri) * 8;
    switch (B) {
        case 1: return launch_gram<1>(A, lds, st);
        case 2: return launch_gram<2>(A, lds, st);
        case 3: return launch_gram<3>(A, lds, st);
        case 4: return launch_gram<4>(A, lds, st);
        case 5: return launch_gram<5>(A, lds, st);
        case 6: return launch_gram<6>(A, lds, st);
        case 7: return launch_gram<7>(A, lds, st);
        default: return launch_gram<8>(A, lds, st);
    }
}

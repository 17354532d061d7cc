// CGLS for min ||A x - b||_2 with A and A^T given as CSR (the Gauss-Newton linear step of the
// inversion loop).  Iteration (Hestenes-Stiefel on the normal equations without forming A^T A):
//   x = 0, r = b, s = A^T r, p = s, gamma = gamma0 = s.s
//   repeat:  q = A p;  alpha = gamma / q.q;  x += alpha p;  r -= alpha q;
//            s = A^T r;  gamma' = s.s;  stop if gamma' <= 0 or gamma' / gamma0 < tol;
//            p = s + (gamma' / gamma) p;  gamma = gamma'
// Everything stays on the device: the scalars live in the workspace and every kernel reads them
// there, so an iteration is six launches with no host round trip; the host polls the stop flag
// once per `check_every` iterations (kernels of iterations after the stop return at once).
// Sparse products put one wavefront per row (short rows) or one 256-thread workgroup per row
// (long rows, e.g. the dense Jacobian block), four independent accumulators per lane; every sum
// runs in a fixed order, so the iterates are bitwise repeatable.
#include "adt_common.hpp"

#include <algorithm>
#include <vector>

namespace {

constexpr int kThreads = 256;
constexpr int kReduceThreads = 256;

// workspace scalars
enum { kGamma = 0, kGamma0, kAlpha, kBeta, kDone, kIters, kDen, kScalars = 8 };
// kDone: 0 running, 1 converged (gamma' / gamma0 < tol or gamma' <= 0), 2 q.q <= 0, 3 nothing to do

struct Csr {
    const int32_t* ptr;
    const int32_t* idx;
    const double* val;
    int rows;
    int per_block;  // rows per workgroup: 4 (a wavefront per row) or 1 (the workgroup per row)
};

__device__ __forceinline__ double row_dot(const Csr& A, int row, const double* x, int lane, int width) {
    const auto ptr = adt::gview(A.ptr);
    const auto idx = adt::gview(A.idx);
    const auto val = adt::gview(A.val);
    const auto xv = adt::gview(x);
    const int k0 = ptr[row], k1 = ptr[row + 1];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = k0 + lane;
    for (; k + 3 * width < k1; k += 4 * width) {
        const int i0 = idx[k], i1 = idx[k + width], i2 = idx[k + 2 * width], i3 = idx[k + 3 * width];
        const double v0 = val[k], v1 = val[k + width], v2 = val[k + 2 * width], v3 = val[k + 3 * width];
        a0 = fma(v0, xv[i0], a0);
        a1 = fma(v1, xv[i1], a1);
        a2 = fma(v2, xv[i2], a2);
        a3 = fma(v3, xv[i3], a3);
    }
    for (; k < k1; k += width) a0 = fma(val[k], xv[idx[k]], a0);
    return (a0 + a1) + (a2 + a3);
}

// y = A x and one partial of y.y per workgroup (fixed order over the workgroup's rows).
__global__ __launch_bounds__(kThreads) void cgls_spmv(Csr A, const double* x, double* y, double* part,
                                                     const double* scal) {
    __shared__ double wave_sq[kThreads / adt::kWave];
    if (adt::gview(scal)[kDone] != 0.0) return;
    const int tid = static_cast<int>(threadIdx.x), lane = tid & (adt::kWave - 1), wv = tid / adt::kWave;
    double sq = 0.0;
    if (A.per_block == 1) {  // the workgroup per row
        const int row = static_cast<int>(blockIdx.x);
        double v = adt::wave_sum(row_dot(A, row, x, tid, kThreads));
        if (lane == 0) wave_sq[wv] = v;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < kThreads / adt::kWave; ++w) t += wave_sq[w];
            adt::gview(y)[row] = t;
            adt::gview(part)[blockIdx.x] = t * t;
        }
        return;
    }
    const int row = static_cast<int>(blockIdx.x) * (kThreads / adt::kWave) + wv;  // a wavefront per row
    if (row < A.rows) {
        const double v = adt::wave_sum(row_dot(A, row, x, lane, adt::kWave));
        if (lane == 0) adt::gview(y)[row] = v;
        sq = v * v;
    }
    if (lane == 0) wave_sq[wv] = sq;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < kThreads / adt::kWave; ++w) t += wave_sq[w];
        adt::gview(part)[blockIdx.x] = t;
    }
}

// Fixed-order sum of n partials by one workgroup: thread t sums t, t + 256, ... in order, then a
// fixed pairwise tree.
__device__ __forceinline__ double reduce_parts(const double* part, int n) {
    __shared__ double s[kReduceThreads];
    const int tid = static_cast<int>(threadIdx.x);
    double a = 0.0;
    for (int i = tid; i < n; i += kReduceThreads) a += adt::gview(part)[i];
    s[tid] = a;
    __syncthreads();
    for (int h = kReduceThreads / 2; h > 0; h >>= 1) {
        if (tid < h) s[tid] += s[tid + h];
        __syncthreads();
    }
    return s[0];
}

// start: gamma = gamma0 = s.s (partials of the first A^T b); nothing to do if gamma0 <= 0 or b = 0
__global__ __launch_bounds__(kReduceThreads) void cgls_start(const double* part_s, int n_s, const double* part_b,
                                                             int n_b, double* scal) {
    const double ss = reduce_parts(part_s, n_s);
    __syncthreads();
    const double bb = reduce_parts(part_b, n_b);
    if (threadIdx.x == 0) {
        auto sc = adt::gview(scal);
        sc[kGamma] = ss;
        sc[kGamma0] = ss;
        sc[kIters] = 0.0;
        sc[kDone] = (ss <= 0.0 || bb <= 0.0) ? 3.0 : 0.0;
    }
}

// alpha = gamma / q.q
__global__ __launch_bounds__(kReduceThreads) void cgls_alpha(const double* part, int n, double* scal) {
    if (adt::gview(scal)[kDone] != 0.0) return;
    const double den = reduce_parts(part, n);
    if (threadIdx.x == 0) {
        auto sc = adt::gview(scal);
        sc[kDen] = den;
        if (den <= 0.0)
            sc[kDone] = 2.0;
        else
            sc[kAlpha] = sc[kGamma] / den;
    }
}

// x += alpha p (n entries), r -= alpha q (m entries)
__global__ __launch_bounds__(kThreads) void cgls_update(double* x, const double* p, int n, double* r, const double* q,
                                                       int m, double* scal) {
    if (adt::gview(scal)[kDone] != 0.0) return;
    const double alpha = adt::gview(scal)[kAlpha];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) adt::gview(x)[i] = fma(alpha, adt::gview(p)[i], adt::gview(x)[i]);
    if (i < m) adt::gview(r)[i] = fma(-alpha, adt::gview(q)[i], adt::gview(r)[i]);
}

// gamma' = s.s: stop test, beta
__global__ __launch_bounds__(kReduceThreads) void cgls_beta(const double* part, int n, double tol, double* scal) {
    if (adt::gview(scal)[kDone] != 0.0) return;
    const double g = reduce_parts(part, n);
    if (threadIdx.x == 0) {
        auto sc = adt::gview(scal);
        sc[kIters] = sc[kIters] + 1.0;  // x and r of this iteration are updated
        if (g <= 0.0 || g / sc[kGamma0] < tol) {
            sc[kDone] = 1.0;
        } else {
            sc[kBeta] = g / sc[kGamma];
            sc[kGamma] = g;
        }
    }
}

// p = s + beta p
__global__ __launch_bounds__(kThreads) void cgls_direction(double* p, const double* s, int n, const double* scal) {
    if (adt::gview(scal)[kDone] != 0.0) return;
    const double beta = adt::gview(scal)[kBeta];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) adt::gview(p)[i] = fma(beta, adt::gview(p)[i], adt::gview(s)[i]);
}

// r = b, x = 0, p = 0 for the first A^T r, and partials of b.b
__global__ __launch_bounds__(kThreads) void cgls_init(const double* b, double* r, int m, double* x, int n,
                                                     double* part) {
    __shared__ double s[kThreads];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    double v = 0.0;
    if (i < m) {
        v = adt::gview(b)[i];
        adt::gview(r)[i] = v;
    }
    if (i < n) adt::gview(x)[i] = 0.0;
    s[threadIdx.x] = v * v;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) adt::gview(part)[blockIdx.x] = s[0];
}

__global__ void cgls_copy(double* dst, const double* src, int n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) adt::gview(dst)[i] = adt::gview(src)[i];
}

__global__ void cgls_clear(double* scal) {
    if (threadIdx.x < kScalars) adt::gview(scal)[threadIdx.x] = 0.0;
}

Csr make_csr(const int32_t* ptr, const int32_t* idx, const double* val, int rows, int64_t nnz) {
    Csr c{ptr, idx, val, rows, 4};
    // long rows (a dense Jacobian row has thousands of entries): the whole workgroup per row
    if (rows > 0 && nnz / rows >= 512) c.per_block = 1;
    return c;
}

int spmv_blocks(const Csr& c) { return c.per_block == 1 ? c.rows : adt::blocks_for(c.rows, kThreads / adt::kWave); }

}  // namespace

extern "C" int64_t adt_cgls_work_doubles(int m, int n) {
    const int64_t parts = std::max<int64_t>({adt::blocks_for(m, 4), adt::blocks_for(n, 4), adt::blocks_for(m, 256),
                                             static_cast<int64_t>(m), static_cast<int64_t>(n)});
    return 2 * static_cast<int64_t>(m) + 2 * static_cast<int64_t>(n) + 2 * parts + kScalars;
}

extern "C" int adt_cgls(int m, int n, const int32_t* d_ptr, const int32_t* d_idx, const double* d_val,
                        const int32_t* d_tptr, const int32_t* d_tidx, const double* d_tval, int64_t nnz,
                        const double* d_b, double* d_x, int max_iter, double tol, int check_every, double* d_work,
                        int* h_iters, int* h_status, void* stream) {
    ADT_REQUIRE(m >= 1 && n >= 1, "adt_cgls: empty system (%d x %d)", m, n);
    ADT_REQUIRE(max_iter >= 0 && tol > 0.0 && check_every >= 1, "adt_cgls: invalid iteration settings");
    ADT_REQUIRE(nnz >= 0 && d_ptr && d_idx && d_tptr && d_tidx && d_b && d_x && d_work &&
                    (nnz == 0 || (d_val && d_tval)),
                "adt_cgls: null pointer");
    hipStream_t st = adt::as_stream(stream);
    const int64_t parts = (adt_cgls_work_doubles(m, n) - 2 * static_cast<int64_t>(m) - 2 * n - kScalars) / 2;
    double* r = d_work;
    double* q = r + m;
    double* s = q + m;
    double* p = s + n;
    double* part = p + n;
    double* part2 = part + parts;
    double* scal = part2 + parts;
    const Csr A = make_csr(d_ptr, d_idx, d_val, m, nnz);
    const Csr At = make_csr(d_tptr, d_tidx, d_tval, n, nnz);
    const int vec_m = adt::blocks_for(m, kThreads), vec_n = adt::blocks_for(n, kThreads);
    const int vec_mn = std::max(vec_m, vec_n);
    hipLaunchKernelGGL(cgls_clear, dim3(1), dim3(64), 0, st, scal);
    hipLaunchKernelGGL(cgls_init, dim3(vec_mn), dim3(kThreads), 0, st, d_b, r, m, d_x, n, part2);
    hipLaunchKernelGGL(cgls_spmv, dim3(spmv_blocks(At)), dim3(kThreads), 0, st, At, r, s, part, scal);
    hipLaunchKernelGGL(cgls_start, dim3(1), dim3(kReduceThreads), 0, st, part, spmv_blocks(At), part2, vec_mn, scal);
    hipLaunchKernelGGL(cgls_copy, dim3(vec_n), dim3(kThreads), 0, st, p, s, n);
    ADT_HIP_CHECK(hipGetLastError());
    double host[kScalars] = {};
    int it = 0;
    while (it < max_iter) {
        const int chunk = std::min(check_every, max_iter - it);
        for (int k = 0; k < chunk; ++k) {
            hipLaunchKernelGGL(cgls_spmv, dim3(spmv_blocks(A)), dim3(kThreads), 0, st, A, p, q, part, scal);
            hipLaunchKernelGGL(cgls_alpha, dim3(1), dim3(kReduceThreads), 0, st, part, spmv_blocks(A), scal);
            hipLaunchKernelGGL(cgls_update, dim3(vec_mn), dim3(kThreads), 0, st, d_x, p, n, r, q, m, scal);
            hipLaunchKernelGGL(cgls_spmv, dim3(spmv_blocks(At)), dim3(kThreads), 0, st, At, r, s, part, scal);
            hipLaunchKernelGGL(cgls_beta, dim3(1), dim3(kReduceThreads), 0, st, part, spmv_blocks(At), tol, scal);
            hipLaunchKernelGGL(cgls_direction, dim3(vec_n), dim3(kThreads), 0, st, p, s, n, scal);
        }
        ADT_HIP_CHECK(hipGetLastError());
        it += chunk;
        ADT_HIP_CHECK(hipMemcpyAsync(host, scal, sizeof(host), hipMemcpyDeviceToHost, st));
        ADT_HIP_CHECK(hipStreamSynchronize(st));
        if (host[kDone] != 0.0) break;
    }
    if (max_iter == 0 || it == 0) {
        ADT_HIP_CHECK(hipMemcpyAsync(host, scal, sizeof(host), hipMemcpyDeviceToHost, st));
        ADT_HIP_CHECK(hipStreamSynchronize(st));
    }
    if (h_iters) *h_iters = static_cast<int>(host[kIters]);
    if (h_status) *h_status = static_cast<int>(host[kDone]);
    return ADT_OK;
}

// Learned multigrid damping: a 3-layer graph message-passing network on the matrix graph that
// predicts, for every matrix slot, the damping used by the finest prolongator of the
// smoothed-aggregation preconditioner (ADT_PRECOND_LEARNED).  Weights are stored as bfloat16
// (staged in LDS, widened to FP32), activations and every sum are FP32 with fixed orders.
//
//   node features  x_i = [log10(a_ii / mean a), sum_{j!=i} |a_ij| / a_ii, sum_j a_ij / a_ii, 1]
//   messages       w_ij = |a_ij| / sum_{j'!=i} |a_ij'|
//   layer 0        h0 = relu(x W0 + b0)                                      (64 channels)
//   layers 1..3    m = Wmsg h_{l-1};  h_l = relu([h_{l-1}, m] W_l + b_l)
//   slot head      z_k = [h3_i, h3_j, a_ij / sqrt(a_ii a_jj), [i == j], same aggregate, strong]
//                  s_k = relu(z_k U + c) . v + d;  omega_k = omega_p (1 + tanh(s_k) / 2)
//
// The backward pass (for training) turns dL/domega into FP32 weight gradients: per-block
// partial sums of every outer product, reduced over blocks in a fixed order.  Transposed
// message and head contributions are gathered through the mirror slot (j, i) of each slot
// (i, j), so no atomics are used anywhere.
#include <mutex>
#include "adt_common.hpp"

#include <cmath>

namespace {

constexpr int H = 64;         // hidden channels
constexpr int HD = 32;        // head hidden units
constexpr int NF = 4;         // node features
constexpr int EF = 4;         // slot features
constexpr int ZIN = 2 * H + EF;  // head input
constexpr int OFF_W0 = 0, OFF_B0 = NF * H;
constexpr int OFF_L1 = OFF_B0 + H;              // per message layer: W (2H x H) then b (H)
constexpr int LAYER_SIZE = 2 * H * H + H;
constexpr int OFF_U = OFF_L1 + 3 * LAYER_SIZE;
constexpr int OFF_C = OFF_U + ZIN * HD, OFF_V = OFF_C + HD, OFF_D = OFF_V + HD;
constexpr int PARAMS = OFF_D + 1;
constexpr int NODES_PER_BLOCK = 32;  // rows per block of the node kernels (4 per round)
constexpr int SLOTS_PER_BLOCK = 64;  // slots per block of the head kernels (8 per round)

__device__ __forceinline__ float bf16_to_float(uint16_t v) { return __uint_as_float(static_cast<uint32_t>(v) << 16); }

// per-system FP32 work layout; every region starts on a 16-byte boundary (pq: per node, the
// head's first layer applied to h3 as the row [h3 U[0:H]] and as the column [h3 U[H:2H]] of a
// slot; the head's hidden units are recomputed from it where the backward needs them)
struct Work {
    float *x, *w, *e, *h[4], *m[4], *s, *mean, *pq;
    double* mean_part;  // per-block partial sums of the diagonal
};

constexpr int kMeanParts = 64;  // blocks per system of the diagonal sum

__host__ __device__ inline int64_t round4(int64_t v) { return (v + 3) & ~int64_t(3); }

__host__ __device__ inline int64_t work_floats(int n, int nnz) {
    return static_cast<int64_t>(n) * (NF + 4 * H + 3 * H + 2 * HD) + 2 * round4(nnz) +
           static_cast<int64_t>(nnz) * EF + 4 + 2 * kMeanParts;
}

__host__ __device__ inline Work work_at(float* base, int n, int nnz) {
    Work W;
    const int64_t N = n, Z = nnz;
    W.x = base;
    W.w = W.x + N * NF;
    W.e = W.w + round4(Z);
    W.h[0] = W.e + Z * EF;
    for (int l = 1; l < 4; ++l) W.h[l] = W.h[l - 1] + N * H;
    W.m[0] = nullptr;
    W.m[1] = W.h[3] + N * H;
    W.m[2] = W.m[1] + N * H;
    W.m[3] = W.m[2] + N * H;
    W.s = W.m[3] + N * H;
    W.mean = W.s + round4(Z);
    W.pq = W.mean + 4;
    W.mean_part = reinterpret_cast<double*>(W.pq + N * 2 * HD);
    return W;
}

struct Graph {
    const int32_t* ptr;
    const int32_t* idx;
    const int32_t* rows;
    const int32_t* mirror;
    const int32_t* flags;  // bit 1: same aggregate
    int n, nnz;
};

// ---------------------------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ int diag_slot(const Graph& g, int i) {
    for (int k = g.ptr[i]; k < g.ptr[i + 1]; ++k)
        if (g.idx[k] == i) return k;
    return -1;
}

// mean of the diagonal per system, fixed order: kMeanParts blocks per system sum rows
// i = b * 256 + t (mod kMeanParts * 256) into block partials, then one block sums the partials
__global__ __launch_bounds__(256) void gnn_diag_partials(Graph g, const double* values, int64_t stride, float* work,
                                                         int64_t wstride) {
    __shared__ double sh[256];
    const double* V = values + blockIdx.y * stride;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < g.n; i += kMeanParts * 256) {
        const int k = diag_slot(g, i);
        acc += k >= 0 ? V[k] : 0.0;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) work_at(work + blockIdx.y * wstride, g.n, g.nnz).mean_part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(64) void gnn_mean_diag(Graph g, float* work, int64_t wstride) {
    Work W = work_at(work + blockIdx.x * wstride, g.n, g.nnz);
    const double v = adt::wave_sum(threadIdx.x < kMeanParts ? W.mean_part[threadIdx.x] : 0.0);
    if (threadIdx.x == 0) W.mean[0] = static_cast<float>(v / g.n);
}

// node features, message weights and slot features: thread per row
__global__ __launch_bounds__(256) void gnn_features(Graph g, const double* values, int64_t stride, float* work,
                                                    int64_t wstride, float theta) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n) return;
    const double* V = values + blockIdx.y * stride;
    Work W = work_at(work + blockIdx.y * wstride, g.n, g.nnz);
    const int kd = diag_slot(g, i);
    const double aii = kd >= 0 ? V[kd] : 1.0;
    double s_off = 0.0, s_row = 0.0;
    for (int k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
        s_row += V[k];
        if (k != kd) s_off += fabs(V[k]);
    }
    W.x[static_cast<int64_t>(i) * NF + 0] = static_cast<float>(log10(aii / static_cast<double>(W.mean[0])));
    W.x[static_cast<int64_t>(i) * NF + 1] = static_cast<float>(s_off / aii);
    W.x[static_cast<int64_t>(i) * NF + 2] = static_cast<float>(s_row / aii);
    W.x[static_cast<int64_t>(i) * NF + 3] = 1.0f;
    for (int k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
        const int j = g.idx[k];
        const int kj = diag_slot(g, j);
        const double ajj = kj >= 0 ? V[kj] : 1.0;
        const double a = V[k];
        const bool diag = k == kd;
        W.w[k] = diag || s_off <= 0.0 ? 0.0f : static_cast<float>(fabs(a) / s_off);
        const double scaled = a / sqrt(aii * ajj);
        const bool strong = !diag && a != 0.0 && fabs(scaled) >= theta;
        float* e = W.e + static_cast<int64_t>(k) * EF;
        e[0] = static_cast<float>(scaled);
        e[1] = diag ? 1.0f : 0.0f;
        e[2] = static_cast<float>((g.flags[k] >> 1) & 1);
        e[3] = strong ? 1.0f : 0.0f;
    }
}

// layer 0: h0 = relu(x W0 + b0); block = 4 rows x 64 channels per round
__global__ __launch_bounds__(256) void gnn_layer0(Graph g, const uint16_t* params, float* work, int64_t wstride) {
    __shared__ float Wl[NF * H + H];
    for (int t = threadIdx.x; t < NF * H + H; t += 256) Wl[t] = bf16_to_float(params[OFF_W0 + t]);
    __syncthreads();
    Work W = work_at(work + blockIdx.y * wstride, g.n, g.nnz);
    const int q = threadIdx.x / H, c = threadIdx.x % H;
    for (int r = 0; r < NODES_PER_BLOCK; r += 4) {
        const int i = blockIdx.x * NODES_PER_BLOCK + r + q;
        if (i >= g.n) break;
        float acc = Wl[NF * H + c];
        for (int t = 0; t < NF; ++t) acc = fmaf(W.x[static_cast<int64_t>(i) * NF + t], Wl[t * H + c], acc);
        W.h[0][static_cast<int64_t>(i) * H + c] = fmaxf(acc, 0.0f);
    }
}

// Node GEMM of a block of 16 U nodes: out[i][o] = act(bias[o] + sum_t in[i][t] W[t][o]) for 64
// outputs, inputs in[i][0..K) staged in LDS (row stride K + kInPad floats), weights W[K][64] in
// LDS as FP32, on the matrix cores: v_mfma_f32_16x16x4f32 multiplies a 16 x 4 block of inputs by a
// 4 x 16 block of weights; lane l supplies in[16 u + l % 16][t0 + l / 16] and W[t0 + l / 16][16 w +
// l % 16] and holds out[16 u + 4 (l / 16) + v][16 w + l % 16] in item v of its accumulator.  Wave w
// owns outputs 16 w .. 16 w + 15 for all U node tiles, so each weight fragment serves U products.
// The input stride K + 4 puts the 64 lanes of an input fragment on distinct LDS banks.  FP32 sums
// in a fixed order (the matrix core's, 4 terms per step): repeatable.
constexpr int kInPad = 4;

template <int U>
constexpr int gemm_nodes() { return 16 * U; }

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int K, int U, bool RELU>
__device__ __forceinline__ void node_gemm(const float* in, const float* Wt, const float* bias, int i0, int n,
                                          float* out) {
    static_assert(K % 4 == 0 && H == 64, "16 x 16 x 4 tiles over 4 waves");
    constexpr int KS = K + kInPad;
    const int lane = static_cast<int>(threadIdx.x) & 63, w = static_cast<int>(threadIdx.x) >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int o = 16 * w + r16;
    const float b0 = bias ? bias[o] : 0.0f;
    f32x4 acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) acc[u] = f32x4{b0, b0, b0, b0};
#pragma unroll 4
    for (int t0 = 0; t0 < K; t0 += 4) {
        const float bv = Wt[(t0 + k4) * H + o];
#pragma unroll
        for (int u = 0; u < U; ++u)
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(in[(16 * u + r16) * KS + t0 + k4], bv, acc[u], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + 16 * u + 4 * k4 + v;
            if (i < n) out[static_cast<int64_t>(i) * H + o] = RELU ? fmaxf(acc[u][v], 0.0f) : acc[u][v];
        }
}

constexpr int kLayerU = 2;  // 32 nodes per block: weights (32 KB) + inputs (16.9 KB) in LDS
constexpr int kProjU = 4;   // 64 nodes per block: weights (16 KB) + inputs (17 KB) in LDS
static_assert(2 * HD == H, "the head projection writes rows of H floats");
static_assert(kMeanParts == adt::kWave, "one wavefront sums the partials");

// messages of layer l: m_i = sum_k w_k h_{l-1}[idx_k] in slot order; 16 threads per node, four
// channels each (16-byte loads), neighbour rows loaded eight at a time
constexpr int kMsgNodes = 16;  // nodes per block

__global__ __launch_bounds__(256) void gnn_message(Graph g, float* work, int64_t wstride, int l) {
    Work W = work_at(work + blockIdx.y * wstride, g.n, g.nnz);
    const auto hp = adt::gview(reinterpret_cast<const float4*>(W.h[l - 1]));
    const auto wk = adt::gview(W.w);
    const auto idx = adt::gview(g.idx);
    const auto ptr = adt::gview(g.ptr);
    const int c4 = static_cast<int>(threadIdx.x) % (H / 4);
    const int i = blockIdx.x * kMsgNodes + static_cast<int>(threadIdx.x) / (H / 4);
    if (i >= g.n) return;
    const int k0 = ptr[i], k1 = ptr[i + 1];
    float4 m = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int kb = k0; kb < k1; kb += 8) {
        int j[8];
        float w[8];
        float4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = kb + u < k1 ? kb + u : k0;
            j[u] = idx[k];
            w[u] = kb + u < k1 ? wk[k] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = hp[static_cast<int64_t>(j[u]) * (H / 4) + c4];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            m.x = fmaf(w[u], v[u].x, m.x);
            m.y = fmaf(w[u], v[u].y, m.y);
            m.z = fmaf(w[u], v[u].z, m.z);
            m.w = fmaf(w[u], v[u].w, m.w);
        }
    }
    adt::gview(reinterpret_cast<float4*>(W.m[l]))[static_cast<int64_t>(i) * (H / 4) + c4] = m;
}

// message layer l: h_l = relu([h_{l-1}, m_l] W_l + b_l); block = kLayerBatches x 32 nodes
constexpr int kLayerBatches = 4;

__global__ __launch_bounds__(256) void gnn_layer(Graph g, const uint16_t* params, float* work, int64_t wstride,
                                                 int l) {
    constexpr int NB = gemm_nodes<kLayerU>();
    __shared__ float Wl[2 * H * H];
    __shared__ float bl[H];
    __shared__ float in[NB * (2 * H + kInPad)];
    const uint16_t* P = params + OFF_L1 + (l - 1) * LAYER_SIZE;
    for (int t = threadIdx.x; t < 2 * H * H; t += 256) Wl[t] = bf16_to_float(P[t]);
    if (threadIdx.x < H) bl[threadIdx.x] = bf16_to_float(P[2 * H * H + threadIdx.x]);
    Work W = work_at(work + blockIdx.y * wstride, g.n, g.nnz);
    const auto hp = adt::gview(W.h[l - 1]);
    const auto mp = adt::gview(W.m[l]);
    // each thread's share of a batch's [h, m] rows, loaded together into registers; the next
    // batch's loads are issued before this batch's products
    constexpr int PER = NB * 2 * H / 256;
    float v[PER];
    const auto load = [&](int i0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int t = static_cast<int>(threadIdx.x) + q * 256;
            const int r = t / (2 * H), c = t % (2 * H), i = i0 + r;
            const int ic = i < g.n ? i : g.n - 1;
            v[q] = c < H ? hp[static_cast<int64_t>(ic) * H + c] : mp[static_cast<int64_t>(ic) * H + c - H];
        }
    };
    const int first = blockIdx.x * kLayerBatches * NB;
    if (first < g.n) load(first);
    for (int b = 0; b < kLayerBatches; ++b) {
        const int i0 = first + b * NB;
        if (i0 >= g.n) break;
        __syncthreads();  // weights stored; the previous batch's inputs read
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int t = static_cast<int>(threadIdx.x) + q * 256;
            const int r = t / (2 * H), c = t % (2 * H);
            in[r * (2 * H + kInPad) + c] = i0 + r < g.n ? v[q] : 0.0f;
        }
        __syncthreads();
        if (b + 1 < kLayerBatches && i0 + NB < g.n) load(i0 + NB);
        node_gemm<2 * H, kLayerU, true>(in, Wl, bl, i0, g.n, W.h[l]);
    }
}

// head, node part: pq[i] = [h3_i U[0:H], h3_i U[H:2H]] (the row and column terms of every slot)
__global__ __launch_bounds__(256) void gnn_head_proj(Graph g, const uint16_t* params, float* work, int64_t wstride) {
    __shared__ float Ws[H * 2 * HD];
    constexpr int NB = gemm_nodes<kProjU>();
    __shared__ float in[NB * (H + kInPad)];
    for (int t = threadIdx.x; t < H * 2 * HD; t += 256) {
        const int r = t / (2 * HD), o = t % (2 * HD);
        Ws[t] = bf16_to_float(params[OFF_U + (o < HD ? r : H + r) * HD + (o % HD)]);
    }
    Work W = work_at(work + blockIdx.y * wstride, g.n, g.nnz);
    const int i0 = blockIdx.x * NB;
    constexpr int PER = NB * H / 256;  // loaded together, then stored
    float v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int t = static_cast<int>(threadIdx.x) + q * 256;
        const int r = t / H, c = t % H, i = min(i0 + r, g.n - 1);
        v[q] = W.h[3][static_cast<int64_t>(i) * H + c];
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int t = static_cast<int>(threadIdx.x) + q * 256;
        const int r = t / H, c = t % H;
        in[r * (H + kInPad) + c] = i0 + r < g.n ? v[q] : 0.0f;
    }
    __syncthreads();
    node_gemm<H, kProjU, false>(in, Ws, nullptr, i0, g.n, W.pq);
}

// hidden units j0..j0+3 of slot k before the ReLU: c + pq_row + pq_col + e_k U[2H:2H+4] (the
// backward recomputes them with this same sequence)
__device__ __forceinline__ float4 head_pre(const float* cs, const float* Ue, float4 pr, float4 pc, float4 e, int j0) {
    float a[4] = {pr.x, pr.y, pr.z, pr.w}, b[4] = {pc.x, pc.y, pc.z, pc.w}, ev[4] = {e.x, e.y, e.z, e.w};
    float o[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        float acc = cs[j0 + v] + a[v] + b[v];
#pragma unroll
        for (int t = 0; t < EF; ++t) acc = fmaf(ev[t], Ue[t * HD + j0 + v], acc);
        o[v] = acc;
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}

__device__ __forceinline__ float4 load4(const float* p) {
    return *adt::gview(reinterpret_cast<const float4*>(p));
}

// slot head: thread = 4 hidden units of one slot, 32 slots per round;
//   u = relu(c + pq_row[i] + pq_col[j] + e_k U[2H:2H+4]),  s = u . v + d
constexpr int kHeadThreadsPerSlot = HD / 4;

__global__ __launch_bounds__(256) void gnn_head(Graph g, const uint16_t* params, float* work, int64_t wstride,
                                                const double* omega_p, double* omega, int64_t ostride) {
    __shared__ float Ue[EF * HD];
    __shared__ float cs[HD], vs[HD];
    for (int t = threadIdx.x; t < EF * HD; t += 256) Ue[t] = bf16_to_float(params[OFF_U + 2 * H * HD + t]);
    if (threadIdx.x < HD) {
        cs[threadIdx.x] = bf16_to_float(params[OFF_C + threadIdx.x]);
        vs[threadIdx.x] = bf16_to_float(params[OFF_V + threadIdx.x]);
    }
    __syncthreads();
    const float d0 = bf16_to_float(params[OFF_D]);
    Work W = work_at(work + blockIdx.y * wstride, g.n, g.nnz);
    const double wp = omega_p[blockIdx.y];
    double* O = omega + blockIdx.y * ostride;
    constexpr int SPR = 256 / kHeadThreadsPerSlot;  // slots per round
    const int q = static_cast<int>(threadIdx.x) / kHeadThreadsPerSlot;
    const int j0 = (static_cast<int>(threadIdx.x) % kHeadThreadsPerSlot) * 4;
    int64_t ri[SLOTS_PER_BLOCK / SPR], ci[SLOTS_PER_BLOCK / SPR];
#pragma unroll
    for (int r = 0; r < SLOTS_PER_BLOCK / SPR; ++r) {
        const int k = blockIdx.x * SLOTS_PER_BLOCK + r * SPR + q;
        const int kc = k < g.nnz ? k : g.nnz - 1;
        ri[r] = adt::gview(g.rows)[kc];
        ci[r] = adt::gview(g.idx)[kc];
    }
#pragma unroll
    for (int r = 0; r < SLOTS_PER_BLOCK / SPR; ++r) {
        const int k = blockIdx.x * SLOTS_PER_BLOCK + r * SPR + q;
        const int kc = k < g.nnz ? k : g.nnz - 1;
        const float4 pre = head_pre(cs, Ue, load4(W.pq + ri[r] * (2 * HD) + j0),
                                    load4(W.pq + ci[r] * (2 * HD) + HD + j0),
                                    load4(W.e + static_cast<int64_t>(kc) * EF), j0);
        float part = fmaxf(pre.x, 0.0f) * vs[j0];
        part = fmaf(fmaxf(pre.y, 0.0f), vs[j0 + 1], part);
        part = fmaf(fmaxf(pre.z, 0.0f), vs[j0 + 2], part);
        part = fmaf(fmaxf(pre.w, 0.0f), vs[j0 + 3], part);
        // sum over the slot's eight threads, fixed butterfly order
#pragma unroll
        for (int off = 1; off < kHeadThreadsPerSlot; off <<= 1) part += __shfl_xor(part, off, 64);
        if (k < g.nnz && j0 == 0) {
            const float sv = part + d0;
            W.s[k] = sv;
            O[k] = wp * (1.0 + 0.5 * tanh(static_cast<double>(sv)));
        }
    }
}

// ---------------------------------------------------------------------------------------------
// backward (one system)
// ---------------------------------------------------------------------------------------------

struct Scratch {
    float *dz, *dh, *dn, *dm, *da, *part;
};

__host__ __device__ inline int head_blocks(int nnz) { return (nnz + SLOTS_PER_BLOCK - 1) / SLOTS_PER_BLOCK; }
__host__ __device__ inline int node_blocks(int n) { return (n + NODES_PER_BLOCK - 1) / NODES_PER_BLOCK; }
constexpr int HEAD_PARAMS = ZIN * HD + HD + HD + 1;  // U, c, v, d
constexpr int LAYER0_PARAMS = NF * H + H;

__host__ __device__ inline int64_t partial_floats(int n, int nnz) {
    const int64_t hb = head_blocks(nnz), nb = node_blocks(n);
    const int64_t a = hb * HEAD_PARAMS, b = nb * LAYER_SIZE;
    return a > b ? a : b;
}

__host__ __device__ inline int64_t scratch_floats(int n, int nnz) {
    return static_cast<int64_t>(nnz) * 2 * H + 4 * static_cast<int64_t>(n) * H + partial_floats(n, nnz);
}

__host__ __device__ inline Scratch scratch_at(float* base, int n, int nnz) {
    Scratch S;
    S.dz = base;
    S.dh = S.dz + static_cast<int64_t>(nnz) * 2 * H;
    S.dn = S.dh + static_cast<int64_t>(n) * H;
    S.dm = S.dn + static_cast<int64_t>(n) * H;
    S.da = S.dm + static_cast<int64_t>(n) * H;
    S.part = S.da + static_cast<int64_t>(n) * H;
    return S;
}

// head backward: du, dz per slot, block partials of (U, c, v, d)
__global__ __launch_bounds__(256) void gnn_head_bwd(Graph g, const uint16_t* params, const float* work,
                                                    double omega_p, const double* domega, float* scratch) {
    __shared__ float z[8][ZIN];
    __shared__ float du[8][HD];
    __shared__ float us[8][HD];
    __shared__ float dsv[8];
    __shared__ uint16_t Us[ZIN * HD];
    __shared__ float vs[HD], cs[HD], Ue[EF * HD];
    for (int t = threadIdx.x; t < ZIN * HD; t += 256) Us[t] = params[OFF_U + t];
    for (int t = threadIdx.x; t < EF * HD; t += 256) Ue[t] = bf16_to_float(params[OFF_U + 2 * H * HD + t]);
    if (threadIdx.x < HD) {
        vs[threadIdx.x] = bf16_to_float(params[OFF_V + threadIdx.x]);
        cs[threadIdx.x] = bf16_to_float(params[OFF_C + threadIdx.x]);
    }
    const Work W = work_at(const_cast<float*>(work), g.n, g.nnz);
    Scratch S = scratch_at(scratch, g.n, g.nnz);
    const int q = threadIdx.x / HD, j = threadIdx.x % HD;
    // entries owned by this thread in the block partial: e = threadIdx.x + 256 r
    constexpr int OWN = (HEAD_PARAMS + 255) / 256;
    float acc[OWN];
#pragma unroll
    for (int r = 0; r < OWN; ++r) acc[r] = 0.0f;
    for (int r0 = 0; r0 < SLOTS_PER_BLOCK; r0 += 8) {
        const int k = blockIdx.x * SLOTS_PER_BLOCK + r0 + q;
        __syncthreads();
        if (k < g.nnz) {
            const int64_t i = g.rows[k], col = g.idx[k];
            for (int t = j; t < H; t += HD) {
                z[q][t] = W.h[3][i * H + t];
                z[q][H + t] = W.h[3][col * H + t];
            }
            if (j < EF) z[q][2 * H + j] = W.e[static_cast<int64_t>(k) * EF + j];
            const float s = W.s[k];
            const float th = tanhf(s);
            const float ds = static_cast<float>(domega[k] * omega_p * 0.5) * (1.0f - th * th);
            const int j4 = j & ~3;
            const float4 pre = head_pre(cs, Ue, load4(W.pq + i * (2 * HD) + j4), load4(W.pq + col * (2 * HD) + HD + j4),
                                        load4(W.e + static_cast<int64_t>(k) * EF), j4);
            const float pj = (j & 3) == 0 ? pre.x : (j & 3) == 1 ? pre.y : (j & 3) == 2 ? pre.z : pre.w;
            const float u = fmaxf(pj, 0.0f);
            us[q][j] = u;
            du[q][j] = u > 0.0f ? ds * vs[j] : 0.0f;
            if (j == 0) dsv[q] = ds;
        } else {
            for (int t = j; t < ZIN; t += HD) z[q][t] = 0.0f;
            du[q][j] = 0.0f;
            us[q][j] = 0.0f;
            if (j == 0) dsv[q] = 0.0f;
        }
        __syncthreads();
        if (k < g.nnz) {  // dz[t] = sum_j U[t][j] du[j] for t < 2H
            for (int t = j; t < 2 * H; t += HD) {
                float a = 0.0f;
                for (int jj = 0; jj < HD; ++jj) a = fmaf(bf16_to_float(Us[t * HD + jj]), du[q][jj], a);
                S.dz[static_cast<int64_t>(k) * 2 * H + t] = a;
            }
        }
        // block partials over the 8 slots of this round, fixed slot order
#pragma unroll
        for (int r = 0; r < OWN; ++r) {
            const int e = threadIdx.x + 256 * r;
            if (e >= HEAD_PARAMS) continue;
            float a = acc[r];
            for (int qq = 0; qq < 8; ++qq) {
                float term;
                if (e < ZIN * HD) {
                    term = z[qq][e / HD] * du[qq][e % HD];
                } else if (e < ZIN * HD + HD) {
                    term = du[qq][e - ZIN * HD];
                } else if (e < ZIN * HD + 2 * HD) {
                    term = dsv[qq] * us[qq][e - ZIN * HD - HD];
                } else {
                    term = dsv[qq];
                }
                a += term;
            }
            acc[r] = a;
        }
    }
#pragma unroll
    for (int r = 0; r < OWN; ++r) {
        const int e = threadIdx.x + 256 * r;
        if (e < HEAD_PARAMS) S.part[static_cast<int64_t>(blockIdx.x) * HEAD_PARAMS + e] = acc[r];
    }
}

// fixed-order sum of block partials: out[e] = sum_b part[b][e]
__global__ __launch_bounds__(256) void gnn_reduce_partials(const float* part, int blocks, int entries, float* out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= entries) return;
    float a = 0.0f;
    for (int b = 0; b < blocks; ++b) a += part[static_cast<int64_t>(b) * entries + e];
    out[e] = a;
}

// dh3_i = sum_{k in row i} dz_k[0:H] + dz_{mirror(k)}[H:2H]  (thread per (row, channel))
__global__ __launch_bounds__(256) void gnn_head_to_nodes(Graph g, float* scratch) {
    Scratch S = scratch_at(scratch, g.n, g.nnz);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int i = static_cast<int>(t / H), c = static_cast<int>(t % H);
    if (i >= g.n) return;
    float a = 0.0f;
    for (int k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
        a += S.dz[static_cast<int64_t>(k) * 2 * H + c];
        a += S.dz[static_cast<int64_t>(g.mirror[k]) * 2 * H + H + c];
    }
    S.dh[static_cast<int64_t>(i) * H + c] = a;
}

// message layer backward, part 1 (thread per (row, channel) in rounds of 4 rows):
// da = dh * [h_l > 0];  [dn, dm] = da W_l^T;  block partials of dW_l, db_l
__global__ __launch_bounds__(256) void gnn_layer_bwd(Graph g, const uint16_t* params, const float* work,
                                                     float* scratch, int l) {
    __shared__ uint16_t Wl[2 * H * H];
    __shared__ float in[4][2 * H];
    __shared__ float das[4][H];
    const uint16_t* P = params + OFF_L1 + (l - 1) * LAYER_SIZE;
    for (int t = threadIdx.x; t < 2 * H * H; t += 256) Wl[t] = P[t];
    const Work W = work_at(const_cast<float*>(work), g.n, g.nnz);
    Scratch S = scratch_at(scratch, g.n, g.nnz);
    const int q = threadIdx.x / H, c = threadIdx.x % H;
    constexpr int OWN = (LAYER_SIZE + 255) / 256;
    float acc[OWN];
#pragma unroll
    for (int r = 0; r < OWN; ++r) acc[r] = 0.0f;
    for (int r0 = 0; r0 < NODES_PER_BLOCK; r0 += 4) {
        const int i = blockIdx.x * NODES_PER_BLOCK + r0 + q;
        __syncthreads();
        if (i < g.n) {
            const int64_t o = static_cast<int64_t>(i) * H + c;
            const float da = W.h[l][o] > 0.0f ? S.dh[o] : 0.0f;
            das[q][c] = da;
            S.da[o] = da;
            in[q][c] = W.h[l - 1][o];
            in[q][H + c] = W.m[l][o];
        } else {
            das[q][c] = 0.0f;
            in[q][c] = 0.0f;
            in[q][H + c] = 0.0f;
        }
        __syncthreads();
        if (i < g.n) {  // dinp[t] = sum_c W[t][c] da[c], t = c and c + H
            for (int t = c; t < 2 * H; t += H) {
                float a = 0.0f;
                for (int cc = 0; cc < H; ++cc) a = fmaf(bf16_to_float(Wl[t * H + cc]), das[q][cc], a);
                if (t < H)
                    S.dn[static_cast<int64_t>(i) * H + t] = a;
                else
                    S.dm[static_cast<int64_t>(i) * H + t - H] = a;
            }
        }
#pragma unroll
        for (int r = 0; r < OWN; ++r) {
            const int e = threadIdx.x + 256 * r;
            if (e >= LAYER_SIZE) continue;
            float a = acc[r];
            if (e < 2 * H * H) {
                const int t = e / H, cc = e % H;
                for (int qq = 0; qq < 4; ++qq) a += in[qq][t] * das[qq][cc];
            } else {
                for (int qq = 0; qq < 4; ++qq) a += das[qq][e - 2 * H * H];
            }
            acc[r] = a;
        }
    }
#pragma unroll
    for (int r = 0; r < OWN; ++r) {
        const int e = threadIdx.x + 256 * r;
        if (e < LAYER_SIZE) S.part[static_cast<int64_t>(blockIdx.x) * LAYER_SIZE + e] = acc[r];
    }
}

// message layer backward, part 2: dh_{l-1}[j] = dn[j] + sum_{k in row j} w_{mirror(k)} dm[col k]
__global__ __launch_bounds__(256) void gnn_msg_bwd(Graph g, const float* work, float* scratch) {
    const Work W = work_at(const_cast<float*>(work), g.n, g.nnz);
    Scratch S = scratch_at(scratch, g.n, g.nnz);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int j = static_cast<int>(t / H), c = static_cast<int>(t % H);
    if (j >= g.n) return;
    float a = S.dn[static_cast<int64_t>(j) * H + c];
    for (int k = g.ptr[j]; k < g.ptr[j + 1]; ++k)
        a = fmaf(W.w[g.mirror[k]], S.dm[static_cast<int64_t>(g.idx[k]) * H + c], a);
    S.dh[static_cast<int64_t>(j) * H + c] = a;
}

// layer 0 backward: da = dh * [h0 > 0]; block partials of dW0 (x da^T) and db0
__global__ __launch_bounds__(256) void gnn_layer0_bwd(Graph g, const float* work, float* scratch) {
    __shared__ float xs[4][NF];
    __shared__ float das[4][H];
    const Work W = work_at(const_cast<float*>(work), g.n, g.nnz);
    Scratch S = scratch_at(scratch, g.n, g.nnz);
    const int q = threadIdx.x / H, c = threadIdx.x % H;
    constexpr int OWN = (LAYER0_PARAMS + 255) / 256;
    float acc[OWN];
#pragma unroll
    for (int r = 0; r < OWN; ++r) acc[r] = 0.0f;
    for (int r0 = 0; r0 < NODES_PER_BLOCK; r0 += 4) {
        const int i = blockIdx.x * NODES_PER_BLOCK + r0 + q;
        __syncthreads();
        if (i < g.n) {
            const int64_t o = static_cast<int64_t>(i) * H + c;
            das[q][c] = W.h[0][o] > 0.0f ? S.dh[o] : 0.0f;
            if (c < NF) xs[q][c] = W.x[static_cast<int64_t>(i) * NF + c];
        } else {
            das[q][c] = 0.0f;
            if (c < NF) xs[q][c] = 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < OWN; ++r) {
            const int e = threadIdx.x + 256 * r;
            if (e >= LAYER0_PARAMS) continue;
            float a = acc[r];
            if (e < NF * H) {
                for (int qq = 0; qq < 4; ++qq) a += xs[qq][e / H] * das[qq][e % H];
            } else {
                for (int qq = 0; qq < 4; ++qq) a += das[qq][e - NF * H];
            }
            acc[r] = a;
        }
    }
#pragma unroll
    for (int r = 0; r < OWN; ++r) {
        const int e = threadIdx.x + 256 * r;
        if (e < LAYER0_PARAMS) S.part[static_cast<int64_t>(blockIdx.x) * LAYER0_PARAMS + e] = acc[r];
    }
}

Graph make_graph(int n, int nnz, const int32_t* ptr, const int32_t* idx, const int32_t* rows, const int32_t* mirror,
                 const int32_t* flags) {
    Graph g;
    g.ptr = ptr;
    g.idx = idx;
    g.rows = rows;
    g.mirror = mirror;
    g.flags = flags;
    g.n = n;
    g.nnz = nnz;
    return g;
}

}  // namespace

extern "C" int adt_gnn_param_count(void) { return PARAMS; }

extern "C" int64_t adt_gnn_work_floats(int n, int nnz) { return work_floats(n, nnz); }

extern "C" int64_t adt_gnn_scratch_floats(int n, int nnz) { return scratch_floats(n, nnz); }

extern "C" int adt_gnn_forward(int n_sys, int n, int nnz, const int32_t* d_indptr, const int32_t* d_indices,
                               const int32_t* d_rows, const int32_t* d_flags, const double* d_values,
                               int64_t val_stride, double theta, const double* d_omega_p, const uint16_t* d_params,
                               float* d_work, double* d_omega, void* stream) {
    ADT_REQUIRE(n_sys >= 1 && n_sys <= 65535 && n >= 1 && nnz >= n, "adt_gnn_forward: invalid sizes");
    ADT_REQUIRE(d_indptr && d_indices && d_rows && d_flags && d_values && d_omega_p && d_params && d_work && d_omega,
                "adt_gnn_forward: null pointer");
    ADT_REQUIRE(val_stride >= nnz, "adt_gnn_forward: val_stride must be >= nnz");
    hipStream_t st = adt::as_stream(stream);
    const Graph g = make_graph(n, nnz, d_indptr, d_indices, d_rows, nullptr, d_flags);
    const int64_t ws = work_floats(n, nnz);
    hipLaunchKernelGGL(gnn_diag_partials, dim3(kMeanParts, n_sys), dim3(256), 0, st, g, d_values, val_stride, d_work,
                       ws);
    hipLaunchKernelGGL(gnn_mean_diag, dim3(n_sys), dim3(64), 0, st, g, d_work, ws);
    hipLaunchKernelGGL(gnn_features, dim3(adt::blocks_for(n, 256), n_sys), dim3(256), 0, st, g, d_values, val_stride,
                       d_work, ws, static_cast<float>(theta));
    const dim3 nodes(node_blocks(n), n_sys);
    hipLaunchKernelGGL(gnn_layer0, nodes, dim3(256), 0, st, g, d_params, d_work, ws);
    const dim3 msg_grid(adt::blocks_for(n, kMsgNodes), n_sys);
    const dim3 layer_grid(adt::blocks_for(n, kLayerBatches * gemm_nodes<kLayerU>()), n_sys);
    for (int l = 1; l <= 3; ++l) {
        hipLaunchKernelGGL(gnn_message, msg_grid, dim3(256), 0, st, g, d_work, ws, l);
        hipLaunchKernelGGL(gnn_layer, layer_grid, dim3(256), 0, st, g, d_params, d_work, ws, l);
    }
    hipLaunchKernelGGL(gnn_head_proj, dim3(adt::blocks_for(n, gemm_nodes<kProjU>()), n_sys), dim3(256), 0, st, g,
                       d_params, d_work, ws);
    hipLaunchKernelGGL(gnn_head, dim3(head_blocks(nnz), n_sys), dim3(256), 0, st, g, d_params, d_work, ws, d_omega_p,
                       d_omega, val_stride);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

namespace {

// every scratch entry the backward kernels read is written earlier in the same call (per-slot
// and per-node gradients, per-block partials), so the scratch never has to start at zero
int backward_launch(int n, int nnz, const int32_t* d_indptr, const int32_t* d_indices, const int32_t* d_rows,
                    const int32_t* d_mirror, double omega_p, const uint16_t* d_params, const float* d_work,
                    const double* d_domega, float* d_scratch, float* d_grad, hipStream_t st) {
    const Graph g = make_graph(n, nnz, d_indptr, d_indices, d_rows, d_mirror, nullptr);
    Scratch S = scratch_at(d_scratch, n, nnz);
    const int hb = head_blocks(nnz), nb = node_blocks(n);
    const int node_channels = adt::blocks_for(static_cast<int64_t>(n) * H, 256);
    hipLaunchKernelGGL(gnn_head_bwd, dim3(hb), dim3(256), 0, st, g, d_params, d_work, omega_p, d_domega, d_scratch);
    hipLaunchKernelGGL(gnn_reduce_partials, dim3(adt::blocks_for(HEAD_PARAMS, 256)), dim3(256), 0, st, S.part, hb,
                       HEAD_PARAMS, d_grad + OFF_U);
    hipLaunchKernelGGL(gnn_head_to_nodes, dim3(node_channels), dim3(256), 0, st, g, d_scratch);
    for (int l = 3; l >= 1; --l) {
        hipLaunchKernelGGL(gnn_layer_bwd, dim3(nb), dim3(256), 0, st, g, d_params, d_work, d_scratch, l);
        hipLaunchKernelGGL(gnn_reduce_partials, dim3(adt::blocks_for(LAYER_SIZE, 256)), dim3(256), 0, st, S.part, nb,
                           LAYER_SIZE, d_grad + OFF_L1 + (l - 1) * LAYER_SIZE);
        hipLaunchKernelGGL(gnn_msg_bwd, dim3(node_channels), dim3(256), 0, st, g, d_work, d_scratch);
    }
    hipLaunchKernelGGL(gnn_layer0_bwd, dim3(nb), dim3(256), 0, st, g, d_work, d_scratch);
    hipLaunchKernelGGL(gnn_reduce_partials, dim3(adt::blocks_for(LAYER0_PARAMS, 256)), dim3(256), 0, st, S.part, nb,
                       LAYER0_PARAMS, d_grad + OFF_W0);
    ADT_HIP_CHECK(hipGetLastError());
    return ADT_OK;
}

}  // namespace

extern "C" int adt_gnn_backward(int n, int nnz, const int32_t* d_indptr, const int32_t* d_indices,
                                const int32_t* d_rows, const int32_t* d_mirror, double omega_p,
                                const uint16_t* d_params, const float* d_work, const double* d_domega,
                                float* d_scratch, float* d_grad, void* stream) {
    ADT_REQUIRE(n >= 1 && nnz >= n, "adt_gnn_backward: invalid sizes");
    ADT_REQUIRE(d_indptr && d_indices && d_rows && d_mirror && d_params && d_work && d_domega && d_scratch && d_grad,
                "adt_gnn_backward: null pointer");
    return backward_launch(n, nnz, d_indptr, d_indices, d_rows, d_mirror, omega_p, d_params, d_work, d_domega,
                           d_scratch, d_grad, adt::as_stream(stream));
}

// Scratch of adt_gnn_backward_static: one library-owned device array per GPU (each device that
// runs the library holds its own copy until the process ends).  In host code the array's name
// is only a host-side stand-in; the device address comes from hipGetSymbolAddress, resolved
// once per device with that device current, and only that address is handed to kernels.
static __device__ __align__(256) float static_backward_scratch[ADT_GNN_STATIC_SCRATCH_FLOATS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_static_scratch_mutex;
float* g_static_scratch[kMaxDevices] = {};

int static_scratch_address(float** out) {
    int device = -1;
    ADT_HIP_CHECK(hipGetDevice(&device));
    ADT_REQUIRE(device >= 0 && device < kMaxDevices, "adt_gnn_backward_static: device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_static_scratch_mutex);
    if (!g_static_scratch[device]) {
        void* address = nullptr;
        const hipError_t err = hipGetSymbolAddress(&address, HIP_SYMBOL(static_backward_scratch));
        if (err != hipSuccess || !address)
            return adt::set_error(ADT_ERR_HIP, "adt_gnn_backward_static: hipGetSymbolAddress failed on device %d: %s",
                                  device, hipGetErrorString(err));
        g_static_scratch[device] = static_cast<float*>(address);
    }
    *out = g_static_scratch[device];
    return ADT_OK;
}

}  // namespace

extern "C" int adt_gnn_backward_static(int n, int nnz, const int32_t* d_indptr, const int32_t* d_indices,
                                       const int32_t* d_rows, const int32_t* d_mirror, double omega_p,
                                       const uint16_t* d_params, const float* d_work, const double* d_domega,
                                       float* d_grad, void* stream) {
    ADT_REQUIRE(n >= 1 && nnz >= n, "adt_gnn_backward_static: invalid sizes");
    ADT_REQUIRE(d_indptr && d_indices && d_rows && d_mirror && d_params && d_work && d_domega && d_grad,
                "adt_gnn_backward_static: null pointer");
    const int64_t need = scratch_floats(n, nnz);
    ADT_REQUIRE(need <= static_cast<int64_t>(ADT_GNN_STATIC_SCRATCH_FLOATS),
                "adt_gnn_backward_static: n=%d, nnz=%d need %lld scratch floats, above the static limit of %lld "
                "(use adt_gnn_backward with a workspace)",
                n, nnz, static_cast<long long>(need), static_cast<long long>(ADT_GNN_STATIC_SCRATCH_FLOATS));
    float* scratch = nullptr;
    const int rc = static_scratch_address(&scratch);
    if (rc) return rc;
    return backward_launch(n, nnz, d_indptr, d_indices, d_rows, d_mirror, omega_p, d_params, d_work, d_domega,
                           scratch, d_grad, adt::as_stream(stream));
}

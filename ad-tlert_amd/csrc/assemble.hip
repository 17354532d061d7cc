// (a) Deterministic finite-element assembly by CSR-slot gather over a slot-ordered scatter map.
//
// The reference materialises the (W, C, K, K) local blocks and folds them into CSR slots with
// padded gather levels (forward/kernels.py:30-85).  Here every CSR slot owns the ordered list of
// the local entries that land on it (the transpose of the scatter map), and the map itself holds
// each entry's template values and conductivity column in slot order, built once per mesh: one
// thread per slot streams its entries contiguously (no per-entry gathers of 8-byte template values
// from 128-byte cells), and only the conductivities are gathered (a per-model array that L2
// holds).  The stiffness and mass parts are summed separately per model,
//   A_m = sum_e sigma_m[col_e] K_e,   B_m = sum_e sigma_m[col_e] M_e,
// and every wavenumber's value follows as A_m + k_w^2 B_m, so the entries are read once per
// model chunk, not once per (model, wavenumber) system.  A slot's entries are summed in list
// order: bitwise repeatable.
#include "adt_common.hpp"

namespace {

constexpr int kModels = 4;  // models whose sums one thread keeps in registers

struct AssembleArgs {
    const double* sigma;
    int64_t sigma_stride;
    const int32_t* slot_ptr;
    const int32_t* slot_vol;
    const int32_t* ecol;
    const double* estiff;
    const double* emass;
    const double* k2;
    const int32_t* ebnd;
    const double* balpha;
    double* values;
    int n_models, n_waves, nnz, n_bnd, with_boundary;
};

__global__ __launch_bounds__(256) void assemble_kernel(AssembleArgs a) {
    const int slot = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (slot >= a.nnz) return;
    const auto ptr = adt::gview(a.slot_ptr);
    const auto col = adt::gview(a.ecol);
    const auto ks = adt::gview(a.estiff);
    const auto sig = adt::gview(a.sigma);
    const int p0 = ptr[slot], pv = p0 + adt::gview(a.slot_vol)[slot];
    const int p1 = a.with_boundary ? ptr[slot + 1] : pv;
    for (int m0 = static_cast<int>(blockIdx.y) * kModels; m0 < a.n_models; m0 += static_cast<int>(gridDim.y) * kModels) {
        double A[kModels], B[kModels];
#pragma unroll
        for (int j = 0; j < kModels; ++j) A[j] = B[j] = 0.0;
        for (int i = p0; i < pv; ++i) {
            const int64_t c = col[i];
            const double t = ks[i];
            const double u = a.emass ? adt::gview(a.emass)[i] : 0.0;
#pragma unroll
            for (int j = 0; j < kModels; ++j) {
                if (m0 + j >= a.n_models) break;
                const double s = sig[static_cast<int64_t>(m0 + j) * a.sigma_stride + c];
                A[j] = fma(s, t, A[j]);
                B[j] = fma(s, u, B[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < kModels; ++j) {
            const int m = m0 + j;
            if (m >= a.n_models) break;
            for (int w = 0; w < a.n_waves; ++w) {
                double v = a.emass ? fma(adt::gview(a.k2)[w], B[j], A[j]) : A[j];
                for (int i = pv; i < p1; ++i) {  // boundary entries: a per-wavenumber coefficient
                    const double s = sig[static_cast<int64_t>(m) * a.sigma_stride + col[i]];
                    const double alpha = adt::gview(a.balpha)[static_cast<int64_t>(w) * a.n_bnd + adt::gview(a.ebnd)[i]];
                    v = fma(s * alpha, ks[i], v);
                }
                adt::gview(a.values)[(static_cast<int64_t>(m) * a.n_waves + w) * a.nnz + slot] = v;
            }
        }
    }
}

__global__ void set_fixed_kernel(double* values, int nnz, int n_sys, const int32_t* fixed, int n_fixed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_fixed) return;
    for (int s = blockIdx.y; s < n_sys; s += gridDim.y) values[static_cast<int64_t>(s) * nnz + fixed[i]] = 1.0;
}

}  // namespace

extern "C" int adt_assemble(int n_models, int n_waves, int nnz, const double* d_sigma, int64_t sigma_stride,
                            const int32_t* d_slot_ptr, const int32_t* d_slot_vol, const int32_t* d_ecol,
                            const double* d_estiff, const double* d_emass, const double* d_k2,
                            const int32_t* d_ebnd, const double* d_balpha, int n_bnd, int with_boundary,
                            const int32_t* d_fixed_slots, int n_fixed, double* d_values, void* stream) {
    ADT_REQUIRE(n_models >= 0 && n_waves >= 1 && nnz >= 0 && n_bnd >= 0, "adt_assemble: invalid sizes");
    ADT_REQUIRE(!with_boundary || n_bnd == 0 || (d_ebnd && d_balpha), "adt_assemble: boundary arrays missing");
    ADT_REQUIRE((d_emass == nullptr) == (d_k2 == nullptr), "adt_assemble: mass and k2 go together");
    const int n_sys = n_models * n_waves;
    if (n_sys == 0 || nnz == 0) return ADT_OK;
    ADT_REQUIRE(d_sigma && d_slot_ptr && d_slot_vol && d_ecol && d_estiff && d_values, "adt_assemble: null pointer");
    ADT_REQUIRE(n_fixed == 0 || d_fixed_slots, "adt_assemble: null fixed-slot list");
    AssembleArgs a{d_sigma, sigma_stride, d_slot_ptr, d_slot_vol, d_ecol, d_estiff, d_emass, d_k2, d_ebnd,
                   d_balpha, d_values, n_models, n_waves, nnz, n_bnd, with_boundary ? 1 : 0};
    const int chunks = (n_models + kModels - 1) / kModels;
    const dim3 grid(adt::blocks_for(nnz, 256), chunks < 65535 ? chunks : 65535);
    hipLaunchKernelGGL(assemble_kernel, grid, dim3(256), 0, adt::as_stream(stream), a);
    ADT_HIP_CHECK(hipGetLastError());
    if (n_fixed > 0) {
        dim3 g2(adt::blocks_for(n_fixed, 256), n_sys < 65535 ? n_sys : 65535);
        hipLaunchKernelGGL(set_fixed_kernel, g2, dim3(256), 0, adt::as_stream(stream), d_values, nnz, n_sys,
                           d_fixed_slots, n_fixed);
        ADT_HIP_CHECK(hipGetLastError());
    }
    return ADT_OK;
}

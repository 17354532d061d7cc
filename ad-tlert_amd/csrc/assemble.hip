// (a) Deterministic finite-element assembly by CSR-slot gather.
//
// The reference materialises the (W, C, K, K) local blocks and folds them into CSR slots
// with padded gather levels (forward/kernels.py:30-85).  Here every CSR slot owns the
// ordered list of local entries that land on it (a precomputed scatter map), and one
// thread per slot sums them directly from the unit templates, so no per-wavenumber
// element blocks are ever written.  Each thread handles one slot for a run of systems:
// the template values are read once per system from L2-resident arrays and the output
// row values[s, :] is written fully coalesced.
#include "adt_common.hpp"

namespace {

struct AssembleArgs {
    const double* sigma;
    int64_t sigma_stride;
    const int32_t* cell_col;
    const int32_t* bnd_col;
    const double* stiff;
    const double* mass;
    const double* k2;
    const double* bmass;
    const double* balpha;
    const int32_t* slot_ptr;
    const int32_t* slot_entry;
    double* values;
    int n_models, n_waves, nnz, n_bnd;
    int kk, kbkb;
    int64_t vol_entries;
};

__global__ __launch_bounds__(256) void assemble_kernel(AssembleArgs a) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.nnz) return;
    const int e0 = a.slot_ptr[slot], e1 = a.slot_ptr[slot + 1];
    const int n_sys = a.n_models * a.n_waves;
    for (int s = blockIdx.y; s < n_sys; s += gridDim.y) {
        const int model = s / a.n_waves, w = s - model * a.n_waves;
        const double* sig = a.sigma + static_cast<int64_t>(model) * a.sigma_stride;
        const double k2 = a.k2 ? a.k2[w] : 0.0;
        double acc = 0.0;
        for (int i = e0; i < e1; ++i) {
            const int64_t e = a.slot_entry[i];
            if (e < a.vol_entries) {
                const int c = static_cast<int>(e / a.kk);
                double t = a.stiff[e];
                if (a.mass) t = fma(k2, a.mass[e], t);
                acc = fma(sig[a.cell_col[c]], t, acc);
            } else {
                const int64_t eb = e - a.vol_entries;
                const int b = static_cast<int>(eb / a.kbkb);
                const double coef = sig[a.bnd_col[b]] * a.balpha[static_cast<int64_t>(w) * a.n_bnd + b];
                acc = fma(coef, a.bmass[eb], acc);
            }
        }
        a.values[static_cast<int64_t>(s) * a.nnz + slot] = acc;
    }
}

__global__ void set_fixed_kernel(double* values, int nnz, int n_sys, const int32_t* fixed, int n_fixed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_fixed) return;
    for (int s = blockIdx.y; s < n_sys; s += gridDim.y) values[static_cast<int64_t>(s) * nnz + fixed[i]] = 1.0;
}

}  // namespace

extern "C" int adt_assemble(int n_models, int n_waves, int nnz, int n_cells, int k_local, int n_bnd,
                            int k_bnd, const double* d_sigma, int64_t sigma_stride,
                            const int32_t* d_cell_col, const int32_t* d_bnd_col, const double* d_stiff,
                            const double* d_mass, const double* d_k2, const double* d_bmass,
                            const double* d_balpha, const int32_t* d_slot_ptr, const int32_t* d_slot_entry,
                            const int32_t* d_fixed_slots, int n_fixed, double* d_values, void* stream) {
    ADT_REQUIRE(n_models >= 0 && n_waves >= 1 && nnz >= 0 && n_cells >= 0 && k_local >= 1,
                "adt_assemble: invalid sizes");
    ADT_REQUIRE(n_bnd == 0 || (d_bnd_col && d_bmass && d_balpha && k_bnd >= 1),
                "adt_assemble: boundary arrays missing");
    ADT_REQUIRE(d_stiff && d_slot_ptr && d_slot_entry && d_values && d_cell_col && d_sigma,
                "adt_assemble: null pointer");
    ADT_REQUIRE((d_mass == nullptr) == (d_k2 == nullptr), "adt_assemble: mass and k2 go together");
    const int n_sys = n_models * n_waves;
    if (n_sys == 0 || nnz == 0) return ADT_OK;
    AssembleArgs a{d_sigma, sigma_stride, d_cell_col, d_bnd_col, d_stiff, d_mass, d_k2, d_bmass, d_balpha,
                   d_slot_ptr, d_slot_entry, d_values, n_models, n_waves, nnz, n_bnd,
                   k_local * k_local, k_bnd * k_bnd,
                   static_cast<int64_t>(n_cells) * k_local * k_local};
    dim3 grid(adt::blocks_for(nnz, 256), n_sys < 65535 ? n_sys : 65535);
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

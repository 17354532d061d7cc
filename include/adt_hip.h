/*
 * adt_hip.h -- C ABI of the MI355X (gfx950) ERT forward-and-sensitivity library.
 *
 * Every pointer argument named d_* is a device pointer owned by the caller (in practice
 * a torch tensor's data_ptr()); h_* arguments are host pointers.  `stream` is a
 * hipStream_t passed as void* (0 = legacy default stream); every call only enqueues
 * work on that stream unless documented otherwise.  Functions return ADT_OK (0) or a
 * status code; adt_last_error() gives a message for the calling thread.
 *
 * Field layout ("interleaved"): the solution/right-hand-side block of one linear system
 * is an (N, Rp) row-major array -- node-major, with the Rp right-hand sides (sources)
 * contiguous -- so one multi-RHS sparse product streams each matrix row once and reads
 * Rp contiguous doubles per nonzero.  Rp is R rounded up to a multiple of the lane
 * group width L (8, 16, 32 or 64).  Several systems are concatenated; system s starts
 * at element offset vec_off[s].
 *
 * The reference entry points each call replaces are cited per function
 * (paths relative to the AD-TLERT 0.3.0 source tree).
 */
#ifndef ADT_HIP_H
#define ADT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ADT_OK = 0,
    ADT_ERR_INVALID = 1,       /* argument validation failed (maps to ValueError)           */
    ADT_ERR_HIP = 2,           /* HIP runtime error (RuntimeError)                            */
    ADT_ERR_NONFINITE = 3,     /* non-finite values in a Krylov recurrence (FloatingPointError) */
    ADT_ERR_NOT_CONVERGED = 4, /* iteration limit reached before the tolerance                */
    ADT_ERR_NOMEM = 5          /* device allocation of a workspace failed                     */
};

/* Message of the last failing call on this host thread (static storage, never NULL). */
const char* adt_last_error(void);

/* ABI revision of this header; the Python loader refuses a library with another value. */
int adt_abi_version(void);

/* Number of device-side compute units and the name of device `device` (for reports). */
int adt_device_info(int device, int* h_cus, char* h_name, int name_len);

/* ------------------------------------------------------------------------------------------
 * (a) Deterministic finite-element assembly.
 *
 * Replaces the gather-levels GroupSum of ERTForward2p5D._assemble
 * (forward/ert2p5d.py:275-314, forward/kernels.py:30-85) and ERTForward3D._assemble_batch
 * (forward/ert3d.py:436-461).  The scatter map comes slot-ordered: slot j owns the entries
 * [slot_ptr[j], slot_ptr[j+1]), its slot_vol[j] volume entries first, then its boundary
 * entries; entry i carries its conductivity column ecol[i] and template value estiff[i] (and
 * emass[i]); a boundary entry also its boundary element ebnd[i].  (Built once per mesh from the
 * per-slot lists of flat template indices.)  For system s = model*n_waves + w:
 *
 *   A = sum_{volume i} sigma[model, ecol[i]] * estiff[i],   B = sum_{volume i} sigma[model, ecol[i]] * emass[i]
 *   values[s, j] = A + k2[w] * B  (+ sum_{boundary i} sigma[model, ecol[i]] * balpha[w, ebnd[i]] * estiff[i]
 *                                   when with_boundary)
 *
 * each sum in the fixed order of the slot's list, so results are bitwise repeatable; the entries
 * are read once per model, not once per wavenumber.  sigma rows are sigma_stride apart.  Slots
 * listed in fixed_slots are then set to 1 (Dirichlet rows).  emass/k2 may be NULL (3D: no
 * wavenumber term); n_bnd may be 0.
 * ---------------------------------------------------------------------------------------- */
int adt_assemble(int n_models, int n_waves, int nnz, const double* d_sigma, int64_t sigma_stride,
                 const int32_t* d_slot_ptr, const int32_t* d_slot_vol, const int32_t* d_ecol,
                 const double* d_estiff, const double* d_emass, const double* d_k2, const int32_t* d_ebnd,
                 const double* d_balpha, int n_bnd, int with_boundary, const int32_t* d_fixed_slots, int n_fixed,
                 double* d_values, void* stream);

/* ------------------------------------------------------------------------------------------
 * Batched multi-RHS sparse product  y_s = A_s x_s  (+ optional  y = beta_in * y_in - A x).
 *
 * Replaces SparsePattern.matvec (forward/discretization.py:60-92).  Systems share or
 * concatenate CSR patterns: system s uses indptr[ptr_off[s] ...], indices[idx_off[s] ...],
 * values[val_off[s] ...] and vectors at vec_off[s], with rows[s] rows.  All offset arrays
 * are host arrays of n_sys entries.  mode 0: y = A x;  mode 1: y = b - A x (b = d_y on
 * entry).  Rows are summed in CSR order: deterministic.
 * ---------------------------------------------------------------------------------------- */
int adt_spmm(int n_sys, int lanes, int rp, const int32_t* h_rows, const int64_t* h_ptr_off,
             const int64_t* h_idx_off, const int64_t* h_val_off, const int64_t* h_vec_off,
             const int32_t* d_indptr, const int32_t* d_indices, const double* d_values, const double* d_x,
             double* d_y, int mode, void* stream);

/* ------------------------------------------------------------------------------------------
 * (b) Batched FP64 preconditioned conjugate gradients on SPD systems.
 *
 * Replaces BatchedSolver.solve (forward/cudss.py:63-122), the cuDSS direct solve that
 * ERTForward2p5D._solve (forward/ert2p5d.py:328-347) and ERTForward3D._solve
 * (forward/ert3d.py:463-481) call.  A plan fixes the batch layout (systems, rows, CSR
 * offsets, lane width) and owns the Krylov workspace; solves then take the matrix values,
 * right-hand sides and the initial guess (overwritten by the solution) by pointer.
 *
 * precond: ADT_PRECOND_JACOBI (plain PCG), ADT_PRECOND_AMG (flexible CG preconditioned by one
 * smoothed-aggregation V-cycle per iteration; needs adt_pcg_attach_amg) or ADT_PRECOND_LEARNED
 * (the same cycle whose finest prolongator damping comes per matrix slot from precond_data, a
 * device array laid out like d_values, e.g. the output of the learned graph network).  With
 * either cycle, a right-hand side whose residual stalls or breaks down is finished by Jacobi
 * PCG from its current iterate, to the same tolerance (adt_pcg_stats reports how many); one
 * that only reaches max_iter is reported as not converged after exactly max_iter cycle
 * iterations.
 * Every solve ends with the true residual b - A x; right-hand sides whose true residual misses
 * tol (the recurrence residual drifts by rounding) restart once from their iterate in Jacobi
 * PCG (adt_pcg_stats reports how many), so the reported residuals are the true ones.
 * h_indptr/h_indices (optional): host copies of the concatenated CSR patterns.  With them and
 * lds_bytes > 0 the plan cuts every pattern into row tiles whose rows and halo (referenced rows
 * outside the tile) fit in lds_bytes of LDS per workgroup, and the sparse product stages each
 * tile once in LDS instead of gathering every neighbour row from memory; rows_per_tile is then
 * an upper bound.  adt_pcg_tiling reports the chosen tile height, tile count and staged rows.
 * h_b_off (optional) places system s's right-hand side block at d_b + h_b_off[s] (systems of
 * different models that share one right-hand side block); by default it mirrors the x layout
 * reported by adt_pcg_layout.
 * Convergence per right-hand side: ||b - A x||_2 <= tol * ||b||_2 (b != 0); a zero b gives x = 0.
 * h_iters (R per system, optional) receives iteration counts, h_relres the final
 * relative residuals.  The call synchronizes the stream once per `check_every` iterations.
 * ---------------------------------------------------------------------------------------- */
enum { ADT_PRECOND_JACOBI = 0, ADT_PRECOND_LEARNED = 1, ADT_PRECOND_AMG = 2 };

typedef struct adt_pcg_plan adt_pcg_plan;

int adt_pcg_create(adt_pcg_plan** plan, int n_sys, int r, int lanes, int rows_per_tile,
                   const int32_t* h_rows, const int64_t* h_ptr_off, const int64_t* h_idx_off,
                   const int64_t* h_val_off, const int64_t* h_b_off, const int32_t* h_indptr,
                   const int32_t* h_indices, int lds_bytes);
int adt_pcg_tiling(const adt_pcg_plan* plan, int* h_rows_per_tile, int* h_tiles, int* h_lds_rows);
int adt_pcg_layout(const adt_pcg_plan* plan, int* h_rp, int64_t* h_vec_len, int64_t* h_vec_off);
int adt_pcg_solve(adt_pcg_plan* plan, const int32_t* d_indptr, const int32_t* d_indices,
                  const double* d_values, const double* d_b, double* d_x, double tol, int max_iter,
                  int precond, const void* precond_data, int check_every, int32_t* h_iters, double* h_relres,
                  void* stream);
int adt_pcg_destroy(adt_pcg_plan* plan);

/* Multigrid hierarchy of a plan (no reference counterpart: it preconditions the iterative
 * solves that replace cuDSS).  Built on the host from host copies of the CSR patterns
 * (h_indptr/h_indices as in adt_pcg_create) and representative values h_values + h_val_off[s]
 * (only the first system of each pattern is read): aggregation by strong connections
 * |a_ij| >= theta sqrt(a_ii a_jj), prolongators smoothed with the filtered matrix (strong
 * entries, weak ones lumped into the diagonal), Galerkin coarse operators, until a
 * level has at most coarse_max rows (dense inverse) or max_levels is reached.  Every solve
 * recomputes the numeric prolongators and Galerkin operators of every system on the device.
 * stall_limit: iterations without a 10% residual decrease before Jacobi PCG takes over (0: never).
 * smoother_scale: damping of the finest level's Jacobi smoother in units of 4 / (3 lambda_max(D^-1 A))
 * of the representative matrix, in (0, 1.5) so that the cycle stays symmetric positive definite
 * (1.3 is the library's default; coarse levels use 4 / (3 lambda_max)).
 * h_p_ptr/h_p_idx/h_p_val, p_cols (optional, null: aggregation on every level): a given
 * prolongator for the finest level, CSR over the fine rows with p_cols coarse columns (e.g. the
 * interpolation of ten-node tetrahedra from their vertices); coarser levels aggregate.  Only
 * for plans whose systems share one CSR pattern.
 * h_levels receives the number of levels. */
int adt_pcg_attach_amg(adt_pcg_plan* plan, const int32_t* h_indptr, const int32_t* h_indices,
                       const double* h_values, const int64_t* h_val_off, double theta, int coarse_max,
                       int max_levels, int stall_limit, double smoother_scale, const int32_t* h_p_ptr,
                       const int32_t* h_p_idx, const double* h_p_val, int p_cols, int* h_levels);
/* Rows, matrix slots (-1 on the finest level: the caller's pattern) and the constant prolongation
 * damping of one level (the value a learned per-slot damping array replaces). */
int adt_pcg_amg_level(const adt_pcg_plan* plan, int sys, int level, int* h_rows, int64_t* h_nnz,
                      double* h_omega);
/* Levels of the attached hierarchy, right-hand sides finished by Jacobi PCG after a stall and
 * right-hand sides restarted for their true residual in the last solve (null outputs skipped). */
int adt_pcg_stats(const adt_pcg_plan* plan, int* h_levels, int* h_fallback_rhs, int* h_polished_rhs);
/* Host-only: the hierarchy that adt_pcg_attach_amg would build for one matrix (level rows and
 * matrix slots into h_rows / h_nnz, max_levels entries each).  Needs no GPU. */
int adt_amg_hierarchy_info(int n, const int32_t* h_indptr, const int32_t* h_indices, const double* h_values,
                           double theta, int coarse_max, int max_levels, int* h_levels, int* h_rows,
                           int64_t* h_nnz);

/* Profiling of the captured iterations (no reference counterpart; used by the benchmark to
 * report the live roofline of the dominant kernels).  adt_pcg_profile(plan, 1) switches the
 * iterations from graph replay to direct launches of the same kernels on the caller's stream,
 * with HIP timing events around every spmv and update launch, and zeroes the accumulators;
 * adt_pcg_profile(plan, 0) returns to graph replay.  adt_pcg_profile_read fills h_out[7] =
 * {total spmv ms, total update ms, number of profiled iterations, summed lane-group iterations
 * of the last solve, L, Rp, total multigrid-cycle ms}. */
int adt_pcg_profile(adt_pcg_plan* plan, int enable);
int adt_pcg_profile_read(const adt_pcg_plan* plan, double* h_out);

/* ------------------------------------------------------------------------------------------
 * (c) Fused normal sensitivity (Jacobian rows).
 *
 * Replaces the NVRTC kernel normal_sensitivity (forward/kernels.py:132-178, launched by
 * kernels.py:222-268):
 *
 *   out[d, c] = - sum_w sum_ij (phi_w[m_d] - phi_w[n_d])[cell_dofs[c,i]]
 *                       * (ws[w] * stiff[c,i,j] + wm[w] * mass[c,i,j])
 *                       * (phi_w[a_d] - phi_w[b_d])[cell_dofs[c,j]]
 *
 * phi is (W, S, N) (field row major); a row index equal to S is a remote electrode with
 * zero potential.  mass/wm may be NULL.  precision 0: FP64 products; 1: FP32 products
 * with FP64 accumulation.  Each output is accumulated by one thread in the fixed order
 * (w, i, j): bitwise repeatable.  out is (D, C) row-major.
 * ---------------------------------------------------------------------------------------- */
int adt_sensitivity(int n_waves, int s_rows, int64_t n_dofs, int n_meas, int n_cells, int k_local,
                    const double* d_phi, const int32_t* d_a, const int32_t* d_b, const int32_t* d_m,
                    const int32_t* d_n, const int32_t* d_cell_dofs, const double* d_stiff,
                    const double* d_mass, const double* h_ws, const double* h_wm, int precision,
                    double* d_out, void* stream);

/* Same Jacobian rows from interleaved fields, in Gram form (for many wavenumbers):
 *   G_c[e][f] = sum_w phi_w[e, dofs_c]^T (ws[w] stiff_c + wm[w] mass_c) phi_w[f, dofs_c]
 *   out[d, c] = -(G_c[m][a] - G_c[m][b] - G_c[n][a] + G_c[n][b])      (index s_rows = remote: 0)
 * phi_w is the (N, rp) block at d_x + h_vec_off[w]; its first s_rows (<= 128) columns are the
 * sources; d_cell_dofs index its rows.  The quadrupoles come as d_quad (n_meas x 4 uint32, 16-byte
 * aligned), built once per survey by adt_gram_quad_offsets from the electrode indices (indices
 * >= s_rows or negative are the remote electrode and read zero).  The
 * element templates must be symmetric positive semidefinite and the weights ws, wm >= 0 (as for
 * every stiffness and mass template and quadrature rule): T = ws stiff + wm mass = R^T R (upper
 * Cholesky factor, pivots that vanish to rounding dropped) and G_c = Y^T Y with Y = [R phi_w]_w.
 * One workgroup per cell forms the upper-triangle 16 x 16 tiles of G_c on the FP64 matrix cores
 * and keeps the upper triangle, with a zero slot for the remote electrode, in LDS (no workspace).
 * d_mass may be NULL (no mass term).  Negative weights are rejected with ADT_ERR_INVALID.
 * FP64, every sum in a fixed order: bitwise repeatable.  out is (D, C) row-major.  Replaces the
 * same reference kernel as adt_sensitivity.
 * Column aggregation (CellAggregation, forward/kernels.py:88-113, fused): with d_col_ptr (n_cols + 1
 * device ints) and d_col_cells, output column p is the sum of the Jacobian columns of cells
 * d_col_cells[d_col_ptr[p] .. d_col_ptr[p+1]); J is linear in G_c, so one workgroup per column sums
 * the G_c of its cells in list order on the matrix cores and writes one column: out is then
 * (D, n_cols) and cells in no list are dropped.  d_col_ptr = NULL: one column per cell. */
/* The quadrupole table of adt_sensitivity_gram: for row d the LDS byte offsets of G(m,a), G(m,b),
 * G(n,a), G(n,b) in the kernel's Gram triangle for s_rows sources (a pure function of the indices
 * and s_rows; one small kernel). */
int adt_gram_quad_offsets(int n_meas, int s_rows, const int32_t* d_a, const int32_t* d_b, const int32_t* d_m,
                          const int32_t* d_n, uint32_t* d_quad, void* stream);
int adt_sensitivity_gram(int n_waves, int rp, int s_rows, const int64_t* h_vec_off, const double* d_x, int n_meas,
                         int n_cells, int k_local, const uint32_t* d_quad, const int32_t* d_cell_dofs,
                         const double* d_stiff, const double* d_mass,
                         const double* h_ws, const double* h_wm, const int32_t* d_col_ptr,
                         const int32_t* d_col_cells, int n_cols, double* d_out, void* stream);

/* Same Jacobian rows in cross-Gram form, for receiver fields that differ from the current
 * fields (3D: ERTForward3D.jacobian, forward/ert3d.py:867-927, with the kernel of
 * forward/kernels.py:132-178):
 *   G_c[e][f] = sum_w phiR_w[e, dofs_c]^T (ws[w] stiff_c + wm[w] mass_c) phiC_w[f, dofs_c]
 *   out[d, c] = -(G_c[m][a] - G_c[m][b] - G_c[n][a] + G_c[n][b])
 * phiR_w / phiC_w are the (N, rp_r) / (N, rp_c) blocks at d_xr + h_off_r[w] / d_xc + h_off_c[w];
 * their first s_r / s_c (<= 128) columns are receivers / current sources; a, b >= s_c and
 * m, n >= s_r read zero (remote electrodes).  Templates need not be symmetric.  When the
 * staged rows of 32 cells fit one CU's LDS (160 KB; 1024-thread workgroups), or of 16/8/4/2
 * cells 64 KB with n_waves * k_local <= 4 (linear tetrahedra), G_c is not formed: a workgroup
 * of consecutive cells evaluates
 * out[d, c] = -sum_{w,i} (phiR[m] - phiR[n])_{w,i} (Psi[a] - Psi[b])_{w,i}
 * with Psi_w = T_w phiC_w on the cell's DOFs, and writes each output row as one contiguous
 * segment (ADT_CROSS_FORM=block takes this form wherever a block fits); otherwise (or with
 * ADT_CROSS_FORM=gram) one workgroup per cell forms G_c on the FP64 matrix cores and keeps it
 * in LDS (remote electrodes read zero).  No workspace.  FP64, fixed summation order: bitwise repeatable.  out
 * is (D, C) row-major.  Column aggregation as in adt_sensitivity_gram (d_col_ptr, d_col_cells,
 * n_cols; out (D, n_cols)); it always takes the whole-Gram form, summing its cells' G_c. */
int adt_sensitivity_gram_cross(int n_waves, const double* d_xr, int rp_r, int s_r, const int64_t* h_off_r,
                               const double* d_xc, int rp_c, int s_c, const int64_t* h_off_c, int n_meas,
                               int n_cells, int k_local, const int32_t* d_a, const int32_t* d_b, const int32_t* d_m,
                               const int32_t* d_n, const int32_t* d_cell_dofs, const double* d_stiff,
                               const double* d_mass, const double* h_ws, const double* h_wm,
                               const int32_t* d_col_ptr, const int32_t* d_col_cells, int n_cols, double* d_out,
                               void* stream);

/* ------------------------------------------------------------------------------------------
 * (d) Adjoint cell gradient:  g[c] = - sum_w sum_s lam_w[:, s]^T (ws[w] K_c + wm[w] M_c) phi_w[:, s]
 *     (+ Robin edges:  - sum_w sum_s lam_w[e]^T (balpha[w,b] * bws[w] * Mb_b) phi_w[e]).
 *
 * Replaces ERTForward2p5D._cell_gradient / ERTForward3D._cell_gradient
 * (forward/ert2p5d.py:765-825, forward/ert3d.py:713-734) including their two SDDMMs
 * (forward/kernels.py:116-127).  lam and phi are interleaved blocks of n_waves systems:
 * lam_w = d_lam + h_lam_off[w] with row stride rp_lam, phi_w likewise; the first r columns
 * are used.  The SDDMM q_w(i,j) = sum_s lam_w[i,s] phi_w[j,s] is formed on the fly per local
 * entry.  Robin edges (n_bnd > 0): d_cell_bnd is a per-cell CSR of boundary-edge ids
 * ((C+1) pointers followed by the edge list), d_bnd_dofs (n_bnd, k_bnd), d_bmass
 * (n_bnd, k_bnd, k_bnd), d_balpha (n_waves, n_bnd).  One wavefront per cell; every reduction
 * is in a fixed order.
 * ---------------------------------------------------------------------------------------- */
int adt_cell_gradient(int n_waves, int r, int n_cells, int k_local, const double* d_lam, int rp_lam,
                      const int64_t* h_lam_off, const double* d_phi, int rp_phi, const int64_t* h_phi_off,
                      const int32_t* d_cell_dofs, const double* d_stiff, const double* d_mass,
                      const double* h_ws, const double* h_wm, int n_bnd, int k_bnd, const int32_t* d_bnd_dofs,
                      const int32_t* d_cell_bnd, const double* d_bmass, const double* d_balpha,
                      const double* h_bws, double* d_out, void* stream);

/* ------------------------------------------------------------------------------------------
 * Electrode potentials of interleaved fields:  P[t, s, f] = sum_w q[w] * sum_k ew[f,k] * x_{t,w}[en[f,k], s]
 *
 * Replaces the integration + E x E transfer map of ERTForward2p5D._integrate /
 * _normal_reciprocal (forward/ert2p5d.py:560-589) and ERTForward3D._measurement_resistances
 * (forward/ert3d.py:559-567).  The electrode interpolation is given as up to kmax
 * (node, weight) pairs per receiver row f (weight 0 pads).  Systems are (t, w) with
 * vec_off[t*n_waves + w]; P is (n_models, r, n_recv) row-major.
 * ---------------------------------------------------------------------------------------- */
int adt_electrode_potentials(int n_models, int n_waves, int rp, int r, int n_recv, int kmax,
                             const int64_t* h_vec_off, const double* d_x, const int32_t* d_enode,
                             const double* d_eweight, const double* h_q, double* d_p, void* stream);

/* Interleaved (N, Rp) block of each system -> (R, N) field rows:  out[s][i, n] = x_s[n, i]. */
int adt_deinterleave(int n_sys, int rp, int r, int64_t n, const int64_t* h_vec_off, const double* d_x,
                     double* d_out, int64_t out_stride, void* stream);

/* Deterministic column aggregation (CellAggregation, forward/kernels.py:88-113):
 * out[d, p] = sum_{i in col_ptr[p] .. col_ptr[p+1]} in[d, col_idx[i]] in CSR order. */
int adt_aggregate_columns(int n_rows, int64_t in_cols, int n_out, const double* d_in,
                          const int32_t* d_col_ptr, const int32_t* d_col_idx, double* d_out,
                          void* stream);

/* Host-only: finest-level aggregate of every row (-1: left out) for strength threshold theta,
 * the aggregation adt_pcg_attach_amg uses for these values.  Needs no GPU. */
int adt_amg_aggregates(int n, const int32_t* h_indptr, const int32_t* h_indices, const double* h_values,
                       double theta, int32_t* h_agg);

/* ------------------------------------------------------------------------------------------
 * Gauss-Newton linear step: CGLS for min ||A x - b||_2 on the device.
 *
 * Replaces the GPU CGLS of the inversion's linearized step (inversion/optimizers.py:270-359,
 * _gpu_cgls at 314-359): x = 0, r = b, s = A^T r, p = s, gamma = gamma0 = s.s; per iteration
 * q = A p, alpha = gamma / q.q, x += alpha p, r -= alpha q, s = A^T r, gamma' = s.s, stop when
 * gamma' <= 0 or gamma' / gamma0 < tol, else p = s + (gamma' / gamma) p.  A (m x n) and A^T
 * (n x m) are CSR with nnz entries each (int32 indices); d_work holds adt_cgls_work_doubles(m, n)
 * doubles (vectors, per-workgroup partial sums, scalars).  The scalars stay on the device; the
 * host checks the stop flag once per check_every iterations.  Every sum runs in a fixed order:
 * bitwise repeatable.  h_iters receives the number of completed iterations, h_status 0
 * (max_iter reached), 1 (converged), 2 (q.q <= 0) or 3 (A^T b = 0 or b = 0: x = 0).
 * ---------------------------------------------------------------------------------------- */
int64_t adt_cgls_work_doubles(int m, int n);
int adt_cgls(int m, int n, const int32_t* d_ptr, const int32_t* d_idx, const double* d_val, const int32_t* d_tptr,
             const int32_t* d_tidx, const double* d_tval, int64_t nnz, const double* d_b, double* d_x, int max_iter,
             double tol, int check_every, double* d_work, int* h_iters, int* h_status, void* stream);

/* ------------------------------------------------------------------------------------------
 * IC(0)-preconditioned batched CG: the second preconditioner that replaces the batched sparse
 * Cholesky of the wavenumber systems (forward/cudss.py:63-122, called from
 * forward/ert2p5d.py:532-553 and forward/ert3d.py:463-481), next to ADT_PRECOND_JACOBI.
 *
 * n_sys SPD systems share one CSR pattern (d_indptr, d_indices: n rows, nnz entries); their
 * values are nnz apart in d_values.  Each system's iterate is an interleaved (n, rp) block of
 * d_x (r used columns, the rest untouched) and uses right-hand-side block s % n_b_blocks of d_b.
 * The incomplete factor L (lower pattern of A) comes from factor_sweeps fine-grained fixed-point
 * sweeps started at l_ij = a_ij / sqrt(a_jj); z = M^-1 r applies L y = r and L^T z = y by
 * 1 + solve_sweeps Jacobi sweeps each, so M^-1 is a fixed SPD operator and CG is plain CG.
 * The lower pattern comes from the host: d_lptr / d_lidx (rows, columns ascending, diagonal
 * last), d_lsrc (position in A of each of the lnnz lower entries), and its transpose by rows
 * d_uptr / d_uidx / d_upos (position in L).  Stop per column at ||b - A x|| <= tol ||b||, checked
 * again on the true residual (at most two restarts).  d_work holds
 * adt_ic0_work_doubles(n_sys, n, nnz, lnnz, rp) doubles.  h_iters / h_relres (n_sys x r) receive
 * iterations and relative residuals.  Column sums run in a fixed order: bitwise repeatable.
 * Returns ADT_ERR_NOT_CONVERGED at max_iter, ADT_ERR_NONFINITE on a breakdown.
 * ---------------------------------------------------------------------------------------- */
int64_t adt_ic0_work_doubles(int n_sys, int n, int nnz, int lnnz, int rp);
int adt_ic0_pcg(int n_sys, int n, int nnz, int r, int rp, const int32_t* d_indptr, const int32_t* d_indices,
                const int32_t* d_lptr, const int32_t* d_lidx, const int32_t* d_lsrc, int lnnz,
                const int32_t* d_uptr, const int32_t* d_uidx, const int32_t* d_upos, const double* d_values,
                const double* d_b, int n_b_blocks, double* d_x, double tol, int max_iter, int factor_sweeps,
                int solve_sweeps, int check_every, double* d_work, int32_t* h_iters, double* h_relres,
                void* stream);

/* ------------------------------------------------------------------------------------------
 * Learned multigrid damping (no reference counterpart; it feeds ADT_PRECOND_LEARNED).
 *
 * A 3-layer graph message-passing network on the matrix graph (64 channels, bfloat16 weights,
 * FP32 activations and sums) maps each matrix to a damping factor per slot for the finest
 * prolongator:  x_i = [log10(a_ii / mean a), sum_{j!=i}|a_ij| / a_ii, sum_j a_ij / a_ii, 1],
 * w_ij = |a_ij| / sum_{j'!=i}|a_ij'|, h0 = relu(x W0 + b0), h_l = relu([h_{l-1}, W_msg h_{l-1}] W_l + b_l)
 * (l = 1..3), z_k = [h3_i, h3_j, a_ij / sqrt(a_ii a_jj), [i==j], same aggregate, strong],
 * omega_k = omega_p (1 + tanh(relu(z_k U + c) . v + d) / 2).
 * Parameters are packed (W0, b0, W1, b1, W2, b2, W3, b3, U, c, v, d), row-major, as bfloat16
 * bit patterns (uint16); adt_gnn_param_count gives their number.  The pattern must be
 * symmetric: d_rows gives each slot's row, d_mirror the slot of (j, i) for slot (i, j),
 * d_flags bit 1 marks slots whose two rows share an aggregate.  Systems share the pattern and
 * are val_stride apart in d_values / d_omega; d_omega_p holds each system's constant damping.
 * d_work holds n_sys * adt_gnn_work_floats(n, nnz) floats of activations, kept for the backward
 * pass (one system): given d_domega = dL/domega it writes the FP32 parameter gradients
 * (packed like the parameters) using adt_gnn_scratch_floats(n, nnz) floats of scratch.
 * Every sum runs in a fixed order (no atomics): results are bitwise repeatable.
 * ---------------------------------------------------------------------------------------- */
int adt_gnn_param_count(void);
int64_t adt_gnn_work_floats(int n, int nnz);
int64_t adt_gnn_scratch_floats(int n, int nnz);
int adt_gnn_forward(int n_sys, int n, int nnz, const int32_t* d_indptr, const int32_t* d_indices,
                    const int32_t* d_rows, const int32_t* d_flags, const double* d_values, int64_t val_stride,
                    double theta, const double* d_omega_p, const uint16_t* d_params, float* d_work,
                    double* d_omega, void* stream);
int adt_gnn_backward(int n, int nnz, const int32_t* d_indptr, const int32_t* d_indices, const int32_t* d_rows,
                     const int32_t* d_mirror, double omega_p, const uint16_t* d_params, const float* d_work,
                     const double* d_domega, float* d_scratch, float* d_grad, void* stream);

/* No-allocation variant of adt_gnn_backward.  It has no counterpart in the reference: it is for
 * callers that cannot allocate device memory (training steps captured in a HIP graph, loops that
 * forbid hipMalloc).  Same arguments and results as adt_gnn_backward without d_scratch: the
 * scratch is a library-owned device array of ADT_GNN_STATIC_SCRATCH_FLOATS floats (48 MiB) per
 * GPU, held until the process ends; neither the call nor the library allocates device memory.
 * Sizes whose adt_gnn_scratch_floats(n, nnz) exceeds the array are rejected with ADT_ERR_INVALID.
 * Every call on one GPU shares that GPU's array (it keeps the last call's contents; nothing in
 * it has to start at zero): calls that overlap on one GPU, e.g. on two streams, race. */
#define ADT_GNN_STATIC_SCRATCH_FLOATS (12LL * 1024 * 1024)
int adt_gnn_backward_static(int n, int nnz, const int32_t* d_indptr, const int32_t* d_indices,
                            const int32_t* d_rows, const int32_t* d_mirror, double omega_p,
                            const uint16_t* d_params, const float* d_work, const double* d_domega, float* d_grad,
                            void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ADT_HIP_H */

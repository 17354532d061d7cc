"""ctypes binding of ``_lib/libadt_hip.so`` (the C ABI declared in ``include/adt_hip.h``).

Arguments are torch tensors on the current HIP device; every call enqueues on
``torch.cuda.current_stream()``.  There is no CPU fallback: when the library is missing or
no GPU is visible the calls raise ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np
import torch

ABI_VERSION = 7
LIB_PATH = Path(os.environ.get("ADT_HIP_LIBRARY", Path(__file__).resolve().parent / "_lib" / "libadt_hip.so"))

ADT_OK, ADT_ERR_INVALID, ADT_ERR_HIP, ADT_ERR_NONFINITE, ADT_ERR_NOT_CONVERGED, ADT_ERR_NOMEM = range(6)
PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_AMG = 0, 1, 2
LDS_BYTES = int(os.environ.get("ADT_SPMV_LDS_BYTES", 49152))  # 0 disables the LDS-staged product
# Wider budget the solver's autotuner also times when the default one cuts the tile height (two
# workgroups per CU instead of three): taller tiles for dense rows (ten-node tetrahedra)
LDS_BYTES_WIDE = 81920
# Jacobi smoother damping of the finest level of the multigrid cycle in units of
# 4 / (3 lambda_max(D^-1 A)), lambda_max from 40 Lanczos steps on the plan's representative matrix.
# Measured with flexible CG to 1e-10 (one MI355X, mean iterations, scale 1.0 / 1.3 / 1.45): 2.5D
# time-lapse batch 24.3 / 22.1 / 26.8; 3D P2 terrain 128 / 114 / 108; 3D P1 box 33.9 / 34.4 / 37.0.
# Scaling the coarse levels too took the box to 32.1, but their Galerkin operators vary from
# system to system (lambda_max 35% above the representative's on the terrain mesh's first coarse
# level for one of twelve models), which made that model's cycle indefinite.
SMOOTHER_SCALE = 1.3

c_int, c_i64, c_double, c_void_p = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
P = c_void_p

_SIGNATURES = {
    "adt_last_error": (ctypes.c_char_p, []),
    "adt_abi_version": (c_int, []),
    "adt_device_info": (c_int, [c_int, P, P, c_int]),
    "adt_assemble": (c_int, [c_int, c_int, c_int, P, c_i64, P, P, P, P, P, P, P, P, c_int, c_int, P, c_int, P, P]),
    "adt_spmm": (c_int, [c_int, c_int, c_int, P, P, P, P, P, P, P, P, P, P, c_int, P]),
    "adt_pcg_create": (c_int, [P, c_int, c_int, c_int, c_int, P, P, P, P, P, P, P, c_int]),
    "adt_pcg_tiling": (c_int, [P, P, P, P]),
    "adt_pcg_layout": (c_int, [P, P, P, P]),
    "adt_pcg_solve": (c_int, [P, P, P, P, P, P, c_double, c_int, c_int, P, c_int, P, P, P]),
    "adt_pcg_destroy": (c_int, [P]),
    "adt_pcg_profile": (c_int, [P, c_int]),
    "adt_pcg_profile_read": (c_int, [P, P]),
    "adt_pcg_attach_amg": (c_int, [P, P, P, P, P, c_double, c_int, c_int, c_int, c_double, P, P, P, c_int, P]),
    "adt_pcg_amg_level": (c_int, [P, c_int, c_int, P, P, P]),
    "adt_pcg_stats": (c_int, [P, P, P, P]),
    "adt_amg_hierarchy_info": (c_int, [c_int, P, P, P, c_double, c_int, c_int, P, P, P]),
    "adt_amg_aggregates": (c_int, [c_int, P, P, P, c_double, P]),
    "adt_gnn_param_count": (c_int, []),
    "adt_gnn_work_floats": (c_i64, [c_int, c_int]),
    "adt_gnn_scratch_floats": (c_i64, [c_int, c_int]),
    "adt_gnn_forward": (c_int, [c_int, c_int, c_int, P, P, P, P, P, c_i64, c_double, P, P, P, P, P]),
    "adt_gnn_backward": (c_int, [c_int, c_int, P, P, P, P, c_double, P, P, P, P, P, P]),
    "adt_gnn_backward_static": (c_int, [c_int, c_int, P, P, P, P, c_double, P, P, P, P, P]),
    "adt_cgls_work_doubles": (c_i64, [c_int, c_int]),
    "adt_ic0_work_doubles": (c_i64, [c_int, c_int, c_int, c_int, c_int]),
    "adt_ic0_pcg": (c_int, [c_int, c_int, c_int, c_int, c_int, P, P, P, P, P, c_int, P, P, P, P, P, c_int, P, c_double,
                            c_int, c_int, c_int, c_int, P, P, P, P]),
    "adt_cgls": (c_int, [c_int, c_int, P, P, P, P, P, P, c_i64, P, P, c_int, c_double, c_int, P, P, P, P]),
    "adt_sensitivity": (c_int, [c_int, c_int, c_i64, c_int, c_int, c_int, P, P, P, P, P, P, P, P, P, P, c_int, P, P]),
    "adt_gram_quad_offsets": (c_int, [c_int, c_int, P, P, P, P, P, P]),
    "adt_sensitivity_gram": (c_int, [c_int, c_int, c_int, P, P, c_int, c_int, c_int, P, P, P, P, P, P, P, P,
                                     c_int, P, P]),
    "adt_sensitivity_gram_cross": (c_int, [c_int, P, c_int, c_int, P, P, c_int, c_int, P, c_int, c_int, c_int, P, P,
                                           P, P, P, P, P, P, P, P, P, c_int, P, P]),
    "adt_cell_gradient": (c_int, [c_int, c_int, c_int, c_int, P, c_int, P, P, c_int, P, P, P, P, P, P, c_int, c_int,
                                  P, P, P, P, P, P, P]),
    "adt_electrode_potentials": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, P, P, P, P, P, P, P]),
    "adt_deinterleave": (c_int, [c_int, c_int, c_int, c_i64, P, P, P, c_i64, P]),
    "adt_aggregate_columns": (c_int, [c_int, c_i64, c_int, P, P, P, P, P]),
}

_lib = None


def library():
    """Load the shared library once (raises ``RuntimeError`` when it is not built)."""

    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python ad-tlert_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (restype, argtypes) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = restype, argtypes
    if lib.adt_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{LIB_PATH} has ABI {lib.adt_abi_version()}, expected {ABI_VERSION}; rebuild it")
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int) -> None:
    if rc == ADT_OK:
        return
    message = library().adt_last_error().decode(errors="replace")
    if rc == ADT_ERR_INVALID:
        raise ValueError(message)
    if rc == ADT_ERR_NONFINITE:
        raise FloatingPointError(message)
    raise RuntimeError(f"adt_hip error {rc}: {message}")


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("ADTLERT-MI355X requires an AMD GPU (HIP); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> c_void_p:
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None) -> c_void_p:
    if t is None:
        return c_void_p(0)
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return c_void_p(t.data_ptr())


def host(a: np.ndarray) -> c_void_p:
    return a.ctypes.data_as(c_void_p)


def device_info(index: int = 0) -> tuple[int, str]:
    cus = c_int(0)
    name = ctypes.create_string_buffer(64)
    check(library().adt_device_info(index, ctypes.byref(cus), name, 64))
    return cus.value, name.value.decode()


# ----------------------------------------------------------------------------------------------
# thin wrappers
# ----------------------------------------------------------------------------------------------


class AssemblyMap:
    """Slot-ordered scatter map of one mesh for ``adt_assemble``.

    ``slot_ptr`` / ``slot_entry`` give each CSR slot's ordered local entries: flat indices into the
    ``(C, K, K)`` cell templates (``< n_vol``) or, past them, into the ``(B, Kb, Kb)`` boundary
    templates.  Within a slot the volume entries come first, then the boundary ones (each in list
    order); ``col`` holds each entry's conductivity column and ``bnd`` the boundary element of a
    boundary entry.  :meth:`entries` lays template values out in the same order (once per mesh)."""

    def __init__(self, slot_ptr, slot_entry, n_vol: int, kk: int, cell_col, *, kb: int = 1, bnd_col=None, device=None):
        device = device or torch.device("cuda")
        ptr = np.asarray(slot_ptr, dtype=np.int64)
        ent = np.asarray(slot_entry, dtype=np.int64)
        nnz = ptr.size - 1
        slot_of = np.repeat(np.arange(nnz), np.diff(ptr))
        is_bnd = ent >= n_vol
        order = np.lexsort((is_bnd, slot_of))  # by slot, volume entries first, list order kept
        ent, is_bnd = ent[order], is_bnd[order]
        self.nnz, self.n_entries = nnz, int(ent.size)
        self.ptr = torch.as_tensor(ptr.astype(np.int32), device=device)
        self.vol = torch.as_tensor(np.bincount(slot_of[~is_bnd], minlength=nnz).astype(np.int32), device=device)
        self._is_bnd = torch.as_tensor(is_bnd, device=device)
        self._vol_index = torch.as_tensor(ent[~is_bnd], device=device)
        self._bnd_index = torch.as_tensor(ent[is_bnd] - n_vol, device=device)
        cell_col = torch.as_tensor(cell_col, device=device).long()
        col = torch.empty(self.n_entries, dtype=torch.int32, device=device)
        col[~self._is_bnd] = cell_col[self._vol_index // kk].int()
        bnd = torch.full((self.n_entries,), -1, dtype=torch.int32, device=device)
        if bool(is_bnd.any()):
            b = self._bnd_index // kb
            col[self._is_bnd] = torch.as_tensor(bnd_col, device=device).long()[b].int()
            bnd[self._is_bnd] = b.int()
        self.col, self.bnd = col, bnd

    def entries(self, templates, boundary=None) -> torch.Tensor:
        """Slot-ordered float64 values of ``templates`` (volume entries) and ``boundary`` (boundary
        entries; zero when None)."""

        out = torch.zeros(self.n_entries, dtype=torch.float64, device=self.ptr.device)
        out[~self._is_bnd] = templates.reshape(-1).double()[self._vol_index]
        if boundary is not None and self._bnd_index.numel():
            out[self._is_bnd] = boundary.reshape(-1).double()[self._bnd_index]
        return out


def assemble(amap: AssemblyMap, sigma, estiff, values, *, n_waves, emass=None, k2=None, balpha=None,
             with_boundary=False, fixed=None):
    """values[(t*W + w), :] for conductivity rows sigma[t] (see ``adt_assemble``)."""

    n_bnd = 0 if balpha is None else int(balpha.shape[-1])
    check(library().adt_assemble(int(sigma.shape[0]), int(n_waves), amap.nnz, ptr(sigma), int(sigma.stride(0)),
                                 ptr(amap.ptr), ptr(amap.vol), ptr(amap.col), ptr(estiff), ptr(emass), ptr(k2),
                                 ptr(amap.bnd), ptr(balpha), n_bnd, int(bool(with_boundary)), ptr(fixed),
                                 0 if fixed is None else int(fixed.numel()), ptr(values), stream()))


def spmm(layout, indptr, indices, values, x, y, mode=0):
    rows, ptr_off, idx_off, val_off, vec_off, lanes, rp = layout
    check(library().adt_spmm(len(rows), lanes, rp, host(rows), host(ptr_off), host(idx_off), host(val_off),
                             host(vec_off), ptr(indptr), ptr(indices), ptr(values), ptr(x), ptr(y), mode, stream()))


class PcgPlan:
    """Owner of one ``adt_pcg_plan`` (batch layout + Krylov workspace)."""

    def __init__(self, rows, r, lanes, rows_per_tile, ptr_off, idx_off, val_off, b_off=None, host_pattern=None,
                 lds_bytes=LDS_BYTES):
        self._lib = library()
        self.rows = np.ascontiguousarray(rows, dtype=np.int32)
        self.n_sys = int(self.rows.size)
        self.r, self.lanes = int(r), int(lanes)
        self.ptr_off = np.ascontiguousarray(ptr_off, dtype=np.int64)
        self.idx_off = np.ascontiguousarray(idx_off, dtype=np.int64)
        self.val_off = np.ascontiguousarray(val_off, dtype=np.int64)
        self.b_off = None if b_off is None else np.ascontiguousarray(b_off, dtype=np.int64)
        hp = None
        if host_pattern is not None:
            hp = tuple(np.ascontiguousarray(a, dtype=np.int32) for a in host_pattern)
        self.host_pattern = hp
        self.amg_levels = 0
        self.lds_bytes = int(lds_bytes) if host_pattern is not None else 0
        handle = c_void_p()
        check(self._lib.adt_pcg_create(ctypes.byref(handle), self.n_sys, self.r, self.lanes, int(rows_per_tile),
                                       host(self.rows), host(self.ptr_off), host(self.idx_off), host(self.val_off),
                                       c_void_p(0) if self.b_off is None else host(self.b_off),
                                       c_void_p(0) if host_pattern is None else host(hp[0]),
                                       c_void_p(0) if host_pattern is None else host(hp[1]),
                                       int(lds_bytes) if host_pattern is not None else 0))
        self._handle = handle
        rp, vec_len = c_int(0), c_i64(0)
        self.vec_off = np.zeros(self.n_sys, dtype=np.int64)
        check(self._lib.adt_pcg_layout(handle, ctypes.byref(rp), ctypes.byref(vec_len), host(self.vec_off)))
        self.rp, self.vec_len = rp.value, vec_len.value
        rb, tiles, lds_rows = c_int(0), c_int(0), c_int(0)
        check(self._lib.adt_pcg_tiling(handle, ctypes.byref(rb), ctypes.byref(tiles), ctypes.byref(lds_rows)))
        self.rows_per_tile, self.tiles, self.lds_rows = rb.value, tiles.value, lds_rows.value

    def attach_amg(self, values, *, theta=0.08, coarse_max=256, max_levels=12, stall_limit=64,
                   smoother_scale=SMOOTHER_SCALE, prolongator=None) -> int:
        """Build the multigrid hierarchy from the first system of each CSR pattern in the device
        ``values`` (read back to the host once); returns the number of levels.  ``smoother_scale``
        multiplies the finest level's Jacobi smoother damping 4 / (3 lambda_max).
        ``prolongator`` (optional): ``(indptr, indices, values, ncols)`` of a given finest-level
        prolongator (single-pattern plans); coarser levels aggregate."""

        if self.host_pattern is None:
            raise ValueError("the multigrid preconditioner needs the host CSR pattern")
        indptr_h = self.host_pattern[0]
        flat = values.reshape(-1)
        segments, seg_of, h_val_off = [], {}, np.zeros(self.n_sys, dtype=np.int64)
        offset = 0
        for s in range(self.n_sys):
            key = (int(self.ptr_off[s]), int(self.idx_off[s]), int(self.rows[s]))
            if key not in seg_of:
                nnz = int(indptr_h[key[0] + key[2]])
                v0 = int(self.val_off[s])
                segments.append(flat[v0:v0 + nnz].double().cpu().numpy())
                seg_of[key] = offset
                offset += nnz
            h_val_off[s] = seg_of[key]
        h_values = np.ascontiguousarray(np.concatenate(segments), dtype=np.float64)
        self.amg_representative = (h_values, h_val_off)
        levels = c_int(0)
        fixed = (None, None, None, 0)
        if prolongator is not None:
            pp, pi, pv, ncols = prolongator
            fixed = (np.ascontiguousarray(pp, dtype=np.int32), np.ascontiguousarray(pi, dtype=np.int32),
                     np.ascontiguousarray(pv, dtype=np.float64), int(ncols))
        self._fixed_prolongator = fixed  # the arrays stay alive during the call
        check(self._lib.adt_pcg_attach_amg(self._handle, host(self.host_pattern[0]), host(self.host_pattern[1]),
                                           host(h_values), host(h_val_off), float(theta), int(coarse_max),
                                           int(max_levels), int(stall_limit), float(smoother_scale),
                                           *(c_void_p(0) if a is None else host(a) for a in fixed[:3]), fixed[3],
                                           ctypes.byref(levels)))
        self.amg_levels = levels.value
        return self.amg_levels

    def amg_level(self, level: int, sys: int = 0) -> dict:
        rows, nnz, omega = c_int(0), c_i64(0), c_double(0.0)
        check(self._lib.adt_pcg_amg_level(self._handle, int(sys), int(level), ctypes.byref(rows), ctypes.byref(nnz),
                                          ctypes.byref(omega)))
        return dict(rows=rows.value, nnz=nnz.value, omega=omega.value)

    def stats(self) -> dict:
        levels, fallback, polished = c_int(0), c_int(0), c_int(0)
        check(self._lib.adt_pcg_stats(self._handle, ctypes.byref(levels), ctypes.byref(fallback),
                                      ctypes.byref(polished)))
        return dict(levels=levels.value, fallback_rhs=fallback.value, polished_rhs=polished.value)

    def solve(self, indptr, indices, values, b, x, tol, max_iter, *, precond=PRECOND_JACOBI, precond_data=None,
              check_every=16):
        iters = np.zeros((self.n_sys, self.r), dtype=np.int32)
        relres = np.zeros((self.n_sys, self.r), dtype=np.float64)
        rc = self._lib.adt_pcg_solve(self._handle, ptr(indptr), ptr(indices), ptr(values), ptr(b), ptr(x), float(tol),
                                     int(max_iter), int(precond), ptr(precond_data), int(check_every), host(iters),
                                     host(relres), stream())
        return rc, iters, relres

    def profile(self, enable: bool = True) -> None:
        check(self._lib.adt_pcg_profile(self._handle, int(bool(enable))))

    def profile_read(self) -> dict:
        out = np.zeros(7, dtype=np.float64)
        check(self._lib.adt_pcg_profile_read(self._handle, host(out)))
        return dict(spmv_ms=out[0], update_ms=out[1], iterations=int(out[2]), group_iterations=int(out[3]),
                    lanes=int(out[4]), rp=int(out[5]), cycle_ms=out[6])

    def close(self):
        if getattr(self, "_handle", None):
            self._lib.adt_pcg_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sensitivity(phi, a, b, m, n, cell_dofs, stiff, mass, ws, wm, out, precision=0):
    W, S, N = phi.shape
    D, (C, K) = int(a.numel()), cell_dofs.shape
    ws = np.ascontiguousarray(ws, dtype=np.float64)
    wm = np.ascontiguousarray(wm if wm is not None else np.zeros(W), dtype=np.float64)
    check(library().adt_sensitivity(W, S, N, D, C, K, ptr(phi), ptr(a), ptr(b), ptr(m), ptr(n), ptr(cell_dofs),
                                    ptr(stiff), ptr(mass), host(ws), host(wm), int(precision), ptr(out), stream()))


def _column_map(columns):
    """``(col_ptr, col_cells, n_cols)`` device arguments of the Gram kernels' column aggregation."""

    if columns is None:
        return None, None, 0
    col_ptr, col_cells = columns
    return col_ptr, col_cells, int(col_ptr.numel()) - 1


def sensitivity_gram(x, vec_off, rp, s_rows, a, b, m, n, cell_dofs, stiff, mass, ws, wm, out, columns=None):
    """Jacobian rows from interleaved fields in Gram form (see ``adt_sensitivity_gram``); with
    ``columns = (col_ptr, col_cells)`` (int32 device tensors) column p sums the cells of its list."""

    W = len(vec_off)
    C, K = cell_dofs.shape
    vec_off = np.ascontiguousarray(vec_off, dtype=np.int64)
    ws = np.ascontiguousarray(ws, dtype=np.float64)
    wm = np.ascontiguousarray(wm if wm is not None else np.zeros(W), dtype=np.float64)
    col_ptr, col_cells, n_cols = _column_map(columns)
    quad = gram_quad_offsets(a, b, m, n, s_rows)
    check(library().adt_sensitivity_gram(W, int(rp), int(s_rows), host(vec_off), ptr(x), int(a.numel()), C, K,
                                         ptr(quad), ptr(cell_dofs), ptr(stiff), ptr(mass), host(ws), host(wm),
                                         ptr(col_ptr), ptr(col_cells), n_cols, ptr(out), stream()))


def gram_quad_offsets(a, b, m, n, s_rows):
    """``(D, 4)`` int32 device table of the Gram kernel's LDS byte offsets of G(m,a), G(m,b), G(n,a),
    G(n,b) per quadrupole (``adt_gram_quad_offsets``)."""

    quad = torch.empty((int(a.numel()), 4), dtype=torch.int32, device=a.device)
    check(library().adt_gram_quad_offsets(int(a.numel()), int(s_rows), ptr(a), ptr(b), ptr(m), ptr(n), ptr(quad),
                                          stream()))
    return quad


def sensitivity_gram_cross(xr, off_r, rp_r, s_r, xc, off_c, rp_c, s_c, a, b, m, n, cell_dofs, stiff, mass, ws, wm,
                           out, columns=None):
    """Jacobian rows from separate receiver / current interleaved fields (``adt_sensitivity_gram_cross``);
    ``columns`` as in :func:`sensitivity_gram`."""

    W = len(off_r)
    C, K = cell_dofs.shape
    off_r = np.ascontiguousarray(off_r, dtype=np.int64)
    off_c = np.ascontiguousarray(off_c, dtype=np.int64)
    ws = np.ascontiguousarray(ws, dtype=np.float64)
    wm = np.ascontiguousarray(wm if wm is not None else np.zeros(W), dtype=np.float64)
    col_ptr, col_cells, n_cols = _column_map(columns)
    check(library().adt_sensitivity_gram_cross(W, ptr(xr), int(rp_r), int(s_r), host(off_r), ptr(xc), int(rp_c),
                                               int(s_c), host(off_c), int(a.numel()), C, K, ptr(a), ptr(b), ptr(m),
                                               ptr(n), ptr(cell_dofs), ptr(stiff), ptr(mass), host(ws), host(wm),
                                               ptr(col_ptr), ptr(col_cells), n_cols, ptr(out), stream()))


def cell_gradient(lam, rp_lam, lam_off, phi, rp_phi, phi_off, r, cell_dofs, stiff, mass, ws, wm, out, *,
                  bnd_dofs=None, cell_bnd=None, bmass=None, balpha=None, bws=None):
    """``out[c] = -sum_w sum_s lam_w[:,s]^T (ws_w K_c + wm_w M_c) phi_w[:,s]`` (+ Robin edges)."""

    W = len(lam_off)
    C, K = cell_dofs.shape
    n_bnd = 0 if bmass is None else int(bmass.shape[0])
    k_bnd = 0 if bmass is None else int(bmass.shape[1])
    lam_off = np.ascontiguousarray(lam_off, dtype=np.int64)
    phi_off = np.ascontiguousarray(phi_off, dtype=np.int64)
    ws = np.ascontiguousarray(ws, dtype=np.float64)
    wm = np.ascontiguousarray(wm if wm is not None else np.zeros(W), dtype=np.float64)
    bws = np.ascontiguousarray(bws if bws is not None else np.zeros(W), dtype=np.float64)
    check(library().adt_cell_gradient(W, int(r), C, K, ptr(lam), int(rp_lam), host(lam_off), ptr(phi), int(rp_phi),
                                      host(phi_off), ptr(cell_dofs), ptr(stiff), ptr(mass), host(ws), host(wm), n_bnd,
                                      k_bnd, ptr(bnd_dofs), ptr(cell_bnd), ptr(bmass), ptr(balpha), host(bws), ptr(out),
                                      stream()))


def electrode_potentials(x, vec_off, n_models, n_waves, rp, r, enode, eweight, q, out):
    vec_off = np.ascontiguousarray(vec_off, dtype=np.int64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    n_recv, kmax = enode.shape
    check(library().adt_electrode_potentials(n_models, n_waves, rp, r, n_recv, kmax, host(vec_off), ptr(x), ptr(enode),
                                             ptr(eweight), host(q), ptr(out), stream()))


def deinterleave(x, vec_off, rp, r, n, out):
    vec_off = np.ascontiguousarray(vec_off, dtype=np.int64)
    check(library().adt_deinterleave(len(vec_off), rp, r, n, host(vec_off), ptr(x), ptr(out), r * n, stream()))


def aggregate_columns(inp, col_ptr, col_idx, out):
    rows = int(inp.shape[0])
    for r0 in range(0, rows, 65535):
        r1 = min(rows, r0 + 65535)
        check(library().adt_aggregate_columns(r1 - r0, int(inp.shape[1]), int(out.shape[1]), ptr(inp[r0:r1]),
                                              ptr(col_ptr), ptr(col_idx), ptr(out[r0:r1]), stream()))


def amg_hierarchy_info(indptr, indices, values, *, theta=0.08, coarse_max=256, max_levels=12):
    """Host-only: level sizes of the multigrid hierarchy built for one CSR matrix (no GPU needed)."""

    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    levels = c_int(0)
    rows = np.zeros(max_levels, dtype=np.int32)
    nnz = np.zeros(max_levels, dtype=np.int64)
    check(library().adt_amg_hierarchy_info(int(indptr.size - 1), host(indptr), host(indices), host(values),
                                           float(theta), int(coarse_max), int(max_levels), ctypes.byref(levels),
                                           host(rows), host(nnz)))
    return rows[:levels.value].tolist(), nnz[:levels.value].tolist()


def amg_aggregates(indptr, indices, values, *, theta=0.08):
    """Host-only: finest-level aggregates (-1: left out) of one CSR matrix."""

    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    out = np.zeros(indptr.size - 1, dtype=np.int32)
    check(library().adt_amg_aggregates(int(indptr.size - 1), host(indptr),
                                       host(np.ascontiguousarray(indices, dtype=np.int32)),
                                       host(np.ascontiguousarray(values, dtype=np.float64)), float(theta), host(out)))
    return out


def gnn_forward(graph, values, val_stride, theta, omega_p, params, work, omega):
    """``omega`` (n_sys rows, ``val_stride`` apart) from matrix values; see ``adt_gnn_forward``."""

    n_sys = int(omega_p.numel())
    check(library().adt_gnn_forward(n_sys, graph.n, graph.nnz, ptr(graph.indptr), ptr(graph.indices), ptr(graph.rows),
                                    ptr(graph.flags), ptr(values), int(val_stride), float(theta), ptr(omega_p),
                                    ptr(params), ptr(work), ptr(omega), stream()))


def gnn_backward(graph, omega_p, params, work, domega, scratch, grad):
    check(library().adt_gnn_backward(graph.n, graph.nnz, ptr(graph.indptr), ptr(graph.indices), ptr(graph.rows),
                                     ptr(graph.mirror), float(omega_p), ptr(params), ptr(work), ptr(domega),
                                     ptr(scratch), ptr(grad), stream()))


GNN_STATIC_SCRATCH_FLOATS = 12 * 1024 * 1024
"""Floats of the library-owned scratch array of ``adt_gnn_backward_static`` (per GPU)."""


def gnn_backward_static(graph, omega_p, params, work, domega, grad):
    """``gnn_backward`` without a workspace: the scratch is the library's static device array
    (nothing is allocated; calls that overlap on one GPU race)."""

    check(library().adt_gnn_backward_static(graph.n, graph.nnz, ptr(graph.indptr), ptr(graph.indices),
                                            ptr(graph.rows), ptr(graph.mirror), float(omega_p), ptr(params),
                                            ptr(work), ptr(domega), ptr(grad), stream()))


CGLS_STATUS = {0: "iteration limit", 1: "converged", 2: "zero search direction (q.q <= 0)", 3: "nothing to solve"}


def cgls(a, at, b, x, max_iter: int, tol: float, check_every: int = 8) -> tuple[int, int]:
    """``min ||A x - b||`` by CGLS on the device (``adt_cgls``): ``a`` and ``at`` are
    ``(indptr, indices, values)`` int32/int32/float64 device tensors of A and A^T; ``x`` receives
    the solution.  Returns ``(iterations, status)`` (``CGLS_STATUS``)."""

    m, n = int(a[0].numel()) - 1, int(at[0].numel()) - 1
    lib = library()
    work = torch.empty(int(lib.adt_cgls_work_doubles(m, n)), dtype=torch.float64, device=b.device)
    iters, status = np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32)
    check(lib.adt_cgls(m, n, ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(at[0]), ptr(at[1]), ptr(at[2]), int(a[1].numel()),
                       ptr(b), ptr(x), int(max_iter), float(tol), int(check_every), ptr(work), host(iters),
                       host(status), stream()))
    return int(iters[0]), int(status[0])


def ic0_patterns(indptr: np.ndarray, indices: np.ndarray):
    """Host-built patterns of an IC(0) factor of a symmetric CSR pattern: the lower triangle
    (``lptr``, ``lidx`` ascending with the diagonal last, ``lsrc`` = position in A) and its
    transpose by rows (``uptr``, ``uidx``, ``upos`` = position in the lower pattern)."""

    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    n = indptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(indptr))
    lower = np.flatnonzero(indices <= rows)  # CSR rows are sorted: per row ascending, diagonal last
    lrows, lcols = rows[lower], indices[lower]
    counts = np.bincount(lrows, minlength=n)
    last = np.cumsum(counts) - 1
    if np.any(counts == 0) or not np.array_equal(lcols[last], np.arange(n)):
        raise ValueError("IC(0) needs a stored diagonal in every row")
    lptr = np.concatenate(([0], np.cumsum(np.bincount(lrows, minlength=n)))).astype(np.int32)
    order = np.lexsort((lrows, lcols))  # by column, then row: the rows of L^T
    uptr = np.concatenate(([0], np.cumsum(np.bincount(lcols, minlength=n)))).astype(np.int32)
    return dict(lptr=lptr, lidx=lcols.astype(np.int32), lsrc=lower.astype(np.int32), uptr=uptr,
                uidx=lrows[order].astype(np.int32), upos=order.astype(np.int32), lnnz=int(lower.size))


def ic0_pcg(n_sys, n, r, rp, indptr, indices, patterns, values, b, n_b_blocks, x, tol, max_iter, *, factor_sweeps=3,
            solve_sweeps=2, check_every=8, work=None):
    """IC(0)-preconditioned CG of ``n_sys`` systems sharing one pattern (``adt_ic0_pcg``);
    ``patterns`` are the device tensors of :func:`ic0_patterns`; ``work`` (optional) is the caller's
    workspace of ``adt_ic0_work_doubles`` doubles.  Returns ``(rc, iters, relres)``."""

    lib = library()
    nnz = int(indices.numel())
    lnnz = int(patterns["lidx"].numel())
    size = int(lib.adt_ic0_work_doubles(n_sys, n, nnz, lnnz, rp))
    if work is None:
        work = torch.empty(size, dtype=torch.float64, device=b.device)
    elif work.numel() < size or work.dtype != torch.float64:
        raise ValueError(f"ic0_pcg needs a float64 workspace of {size} doubles")
    iters = np.zeros((n_sys, r), dtype=np.int32)
    relres = np.zeros((n_sys, r), dtype=np.float64)
    rc = lib.adt_ic0_pcg(int(n_sys), int(n), nnz, int(r), int(rp), ptr(indptr), ptr(indices), ptr(patterns["lptr"]),
                         ptr(patterns["lidx"]), ptr(patterns["lsrc"]), lnnz, ptr(patterns["uptr"]),
                         ptr(patterns["uidx"]), ptr(patterns["upos"]), ptr(values), ptr(b), int(n_b_blocks), ptr(x),
                         float(tol), int(max_iter), int(factor_sweeps), int(solve_sweeps), int(check_every),
                         ptr(work), host(iters), host(relres), stream())
    return rc, iters, relres

"""Learned multigrid damping: a small graph network that adapts the finest prolongator.

The multigrid preconditioner smooths its finest prolongator with one damped Jacobi step of
the filtered matrix, ``P = (I - D_F^-1 (Omega o A_F)) P0``.  The constant method uses one
damping ``omega_p`` for every slot; here a 3-layer message-passing network on the matrix graph
(64 channels, bfloat16 weights, FP32 activations; ``csrc/gnn.hip``) predicts a damping per
slot from the local coefficients, ``omega_k = omega_p (1 + tanh(s_k) / 2)``, so it stays within
[omega_p / 2, 3 omega_p / 2] and a network with a zero output layer reproduces the constant
method exactly.

Training follows the learned-AMG recipe of Luz, Galun, Maron, Basri and Yavneh ("Learning
algebraic multigrid using graph neural networks", ICML 2020) with error samples instead of
the full error-propagation norm: plain Jacobi PCG (the solver this preconditioner replaces)
runs a random number of iterations on random right-hand sides of matrices assembled from
random conductivity fields, and its remaining errors e are the smooth modes the coarse space
must capture.  The loss is the mean energy ratio ``||T e||_A^2 / ||e||_A^2`` of the two-grid
cycle T (damped Jacobi, exact coarse solve, damped Jacobi).  The network's forward and backward
passes are the native kernels; the loss and its gradient with respect to the dampings are
dense PyTorch algebra on the small training meshes, and the optimizer is Adam on FP32 master
weights, rounded to bfloat16 for every forward pass.
"""

from __future__ import annotations

import json
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from . import native

HIDDEN, HEAD, NODE_FEATURES, EDGE_FEATURES, LAYERS = 64, 32, 4, 4, 3
LAYOUT = [("W0", (NODE_FEATURES, HIDDEN)), ("b0", (HIDDEN,))]
for _l in range(1, LAYERS + 1):
    LAYOUT += [(f"W{_l}", (2 * HIDDEN, HIDDEN)), (f"b{_l}", (HIDDEN,))]
LAYOUT += [("U", (2 * HIDDEN + EDGE_FEATURES, HEAD)), ("c", (HEAD,)), ("v", (HEAD,)), ("d", (1,))]
PARAM_COUNT = int(sum(np.prod(s) for _, s in LAYOUT))
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "data" / "damping_gnn.npz"


def mirror_slots(indptr: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Slot of (j, i) for every slot (i, j) of a CSR pattern with sorted rows; the pattern must be symmetric."""

    n = indptr.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    keys = rows * n + indices
    if np.any(np.diff(keys) <= 0):
        raise ValueError("CSR rows must have sorted, unique column indices")
    target = indices.astype(np.int64) * n + rows
    pos = np.searchsorted(keys, target)
    if np.any(pos >= keys.size) or np.any(keys[np.minimum(pos, keys.size - 1)] != target):
        raise ValueError("the learned damping needs a symmetric sparsity pattern")
    return pos.astype(np.int32)


@dataclass
class GnnGraph:
    """Device description of one symmetric CSR pattern and its finest aggregation."""

    n: int
    nnz: int
    indptr: torch.Tensor
    indices: torch.Tensor
    rows: torch.Tensor
    mirror: torch.Tensor
    flags: torch.Tensor
    agg: np.ndarray = field(repr=False, default=None)

    @classmethod
    def from_pattern(cls, indptr, indices, agg, device) -> "GnnGraph":
        indptr = np.ascontiguousarray(indptr, dtype=np.int32)
        indices = np.ascontiguousarray(indices, dtype=np.int32)
        n = indptr.size - 1
        rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(indptr))
        agg = np.asarray(agg, dtype=np.int32)
        same = (agg[rows] >= 0) & (agg[rows] == agg[indices])
        flags = (same.astype(np.int32) << 1)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)  # noqa: E731
        return cls(n, int(indices.size), t(indptr), t(indices), t(rows), t(mirror_slots(indptr, indices)), t(flags), agg)


def _he_init(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in LAYOUT:
        if name.startswith("W") or name == "U":
            parts.append(rng.standard_normal(shape) * np.sqrt(2.0 / shape[0]))
        else:
            parts.append(np.zeros(shape))  # biases; v = d = 0: the untrained network returns omega_p
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


class DampingNetwork:
    """FP32 master weights (trainable) and their bfloat16 copy used by the kernels."""

    def __init__(self, params=None, *, seed: int = 0, theta: float = 0.08, device=None, manifest=None):
        if native.library().adt_gnn_param_count() != PARAM_COUNT:
            raise RuntimeError("the native library packs the damping network differently; rebuild it")
        self.device = native.device() if device is None else device
        flat = _he_init(seed) if params is None else np.asarray(params, dtype=np.float32).reshape(-1)
        if flat.size != PARAM_COUNT:
            raise ValueError(f"expected {PARAM_COUNT} parameters, got {flat.size}")
        self.master = torch.nn.Parameter(torch.as_tensor(flat, device=self.device))
        self.theta = float(theta)
        self.manifest = dict(manifest or {})
        self._work = {}

    def bf16_bits(self) -> torch.Tensor:
        """Parameters rounded to bfloat16 (nearest even), as raw 16-bit patterns."""

        return self.master.detach().to(torch.bfloat16).view(torch.int16).contiguous()

    def _workspace(self, n_sys: int, graph: GnnGraph) -> torch.Tensor:
        size = n_sys * int(native.library().adt_gnn_work_floats(graph.n, graph.nnz))
        buf = self._work.get("fwd")
        if buf is None or buf.numel() < size:
            buf = torch.empty(size, dtype=torch.float32, device=self.device)
            self._work["fwd"] = buf
        return buf

    def omega(self, graph: GnnGraph, values: torch.Tensor, omega_p, *, val_stride=None, params=None,
              work=None) -> torch.Tensor:
        """Per-slot dampings (same layout as ``values``: n_sys rows ``val_stride`` apart)."""

        values = values.contiguous()
        stride = graph.nnz if val_stride is None else int(val_stride)
        n_sys = (values.numel() + stride - graph.nnz) // stride
        omega_p = torch.as_tensor(omega_p, dtype=torch.float64, device=self.device).reshape(-1).expand(n_sys).contiguous()
        out = torch.zeros_like(values)
        work = self._workspace(n_sys, graph) if work is None else work
        native.gnn_forward(graph, values, stride, self.theta, omega_p, self.bf16_bits() if params is None else params,
                           work, out)
        return out

    def save(self, path) -> None:
        np.savez(path, params=self.master.detach().cpu().numpy(), theta=self.theta,
                 manifest=json.dumps(self.manifest))

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device=None) -> "DampingNetwork":
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"], theta=float(z["theta"]), device=device, manifest=json.loads(str(z["manifest"])))


class _Damping(torch.autograd.Function):
    """omega = network(A) for one system, differentiable in the FP32 master weights."""

    @staticmethod
    def forward(ctx, master, net: DampingNetwork, graph: GnnGraph, values: torch.Tensor, omega_p: float):
        params = net.bf16_bits()
        lib = native.library()
        work = torch.empty(int(lib.adt_gnn_work_floats(graph.n, graph.nnz)), dtype=torch.float32, device=values.device)
        omega = net.omega(graph, values, omega_p, params=params, work=work)
        ctx.save_for_backward(params, work)
        ctx.graph, ctx.omega_p = graph, float(omega_p)
        return omega

    @staticmethod
    def backward(ctx, domega):
        params, work = ctx.saved_tensors
        graph = ctx.graph
        lib = native.library()
        scratch = torch.empty(int(lib.adt_gnn_scratch_floats(graph.n, graph.nnz)), dtype=torch.float32,
                              device=work.device)
        grad = torch.zeros(PARAM_COUNT, dtype=torch.float32, device=work.device)
        native.gnn_backward(graph, ctx.omega_p, params, work, domega.contiguous().to(torch.float64), scratch, grad)
        return grad, None, None, None, None


def damping(net: DampingNetwork, graph: GnnGraph, values: torch.Tensor, omega_p: float) -> torch.Tensor:
    return _Damping.apply(net.master, net, graph, values, omega_p)


# ----------------------------------------------------------------------------------------------
# training
# ----------------------------------------------------------------------------------------------


@dataclass
class TrainingSystem:
    """One assembled matrix with its aggregation, filtered diagonal and error samples."""

    graph: GnnGraph
    values: torch.Tensor        # (nnz,) float64
    diag: torch.Tensor          # (n,) float64
    errors: torch.Tensor        # (n, samples) float64
    omega_p: float
    omega_s: float
    p_rows: torch.Tensor = None
    p_cols: torch.Tensor = None
    self_slot: torch.Tensor = None   # per P entry: 1 for the (i, agg(i)) term, else 0
    slot: torch.Tensor = None        # matrix slot of each P term
    coef: torch.Tensor = None        # a_k / D_F,i of each P term (0 for self terms)
    nc: int = 0


def prepare_system(indptr, indices, values: torch.Tensor, errors: torch.Tensor, *, theta=0.08,
                   smoother_scale=native.SMOOTHER_SCALE, device=None) -> TrainingSystem:
    """Aggregation (as the multigrid setup builds it), the prolongator terms and dampings."""

    device = values.device if device is None else device
    indptr = np.asarray(indptr, dtype=np.int32)
    indices = np.asarray(indices, dtype=np.int32)
    vals = values.double().cpu().numpy()
    n = indptr.size - 1
    agg = native.amg_aggregates(indptr, indices, vals, theta=theta)
    nc = int(agg.max()) + 1
    rows = np.repeat(np.arange(n), np.diff(indptr))
    diag = np.zeros(n)
    isdiag = rows == indices
    diag[rows[isdiag]] = vals[isdiag]
    scaled = vals / np.sqrt(diag[rows] * diag[indices])
    strong = (~isdiag) & (vals != 0) & (np.abs(scaled) >= theta)
    weak = (~isdiag) & ~strong
    df = diag + np.bincount(rows[weak], weights=vals[weak], minlength=n)
    df = np.where(df > 0, df, diag)
    self_terms = np.flatnonzero(isdiag & (agg[rows] >= 0))
    strong_terms = np.flatnonzero(strong & (agg[indices] >= 0) & (agg[rows] >= 0))
    p_rows = np.concatenate((rows[self_terms], rows[strong_terms]))
    p_cols = np.concatenate((agg[rows[self_terms]], agg[indices[strong_terms]]))
    slot = np.concatenate((self_terms, strong_terms))
    coef = np.concatenate((np.zeros(self_terms.size), vals[strong_terms] / df[rows[strong_terms]]))
    is_self = np.concatenate((np.ones(self_terms.size), np.zeros(strong_terms.size)))
    # damping constants of the multigrid setup: smoother_scale * 4 / (3 lambda_max) of D^-1 A and
    # 4 / (3 lambda_max) of D_F^-1 A_F
    fv = np.where(strong, vals, 0.0)
    fv[isdiag] = df[rows[isdiag]]
    omega_s = smoother_scale * 4.0 / (3.0 * _power_lambda(indptr, indices, vals, diag))
    omega_p = 4.0 / (3.0 * _power_lambda(indptr, indices, fv, df))
    t = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt, device=device)  # noqa: E731
    graph = GnnGraph.from_pattern(indptr, indices, agg, device)
    return TrainingSystem(graph, values.double().contiguous(), t(diag), errors, omega_p, omega_s,
                          t(p_rows, torch.long), t(p_cols, torch.long), t(is_self), t(slot, torch.long), t(coef), nc)


def _power_lambda(indptr, indices, vals, d, iters=30) -> float:
    """lambda_max(D^-1 B) by power iteration with a Rayleigh quotient in the D inner product."""

    import scipy.sparse as sp

    n = indptr.size - 1
    B = sp.csr_matrix((vals, indices, indptr), shape=(n, n))
    v = np.linspace(0.5, 1.5, n)
    lam = 1.0
    for _ in range(iters):
        w = (B @ v) / d
        lam = float((v * d) @ w / ((v * d) @ v))
        v = w / np.linalg.norm(w)
    return lam


def prolongator(sys: TrainingSystem, omega: torch.Tensor) -> torch.Tensor:
    """Dense finest prolongator for per-slot dampings (the device formula, differentiable)."""

    om = omega[sys.slot]
    vals = sys.self_slot * (1.0 - om) - (1.0 - sys.self_slot) * om * sys.coef
    P = torch.zeros((sys.graph.n, sys.nc), dtype=torch.float64, device=omega.device)
    return P.index_put((sys.p_rows, sys.p_cols), vals, accumulate=True)


class _Product(torch.autograd.Function):
    """Y = A X for the (symmetric) training matrix with the native multi-RHS product; the
    backward product is A^T dY = A dY."""

    @staticmethod
    def forward(ctx, X, sys):
        ctx.sys = sys
        return _apply(sys, X)

    @staticmethod
    def backward(ctx, dY):
        return _apply(ctx.sys, dY), None


def _apply(sys: TrainingSystem, X: torch.Tensor) -> torch.Tensor:
    n, m = X.shape
    rp = -(-m // 8) * 8
    x = torch.zeros((1, n, rp), dtype=torch.float64, device=X.device)
    x[0, :, :m] = X
    y = torch.empty_like(x)
    layout = (np.array([n], np.int32), np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64),
              np.zeros(1, np.int64), 8, rp)
    native.spmm(layout, sys.graph.indptr, sys.graph.indices, sys.values.reshape(1, -1), x, y, 0)
    return y[0, :, :m]


def two_grid_loss(sys: TrainingSystem, omega: torch.Tensor) -> torch.Tensor:
    """Mean energy ratio ||T e||_A^2 / ||e||_A^2 of the two-grid cycle over the error samples."""

    A = lambda X: _Product.apply(X, sys)  # noqa: E731
    E = sys.errors
    dinv = sys.omega_s / sys.diag
    P = prolongator(sys, omega)
    Y = E - dinv[:, None] * A(E)
    Ac = P.T @ A(P)
    Y = Y - P @ torch.linalg.solve(Ac, P.T @ A(Y))
    Y = Y - dinv[:, None] * A(Y)
    return ((Y * A(Y)).sum(0) / (E * A(E)).sum(0)).mean()


def error_samples(indptr_d, indices_d, values: torch.Tensor, n: int, samples: int, iterations: int,
                  rng: np.random.Generator, host_pattern) -> torch.Tensor:
    """Remaining errors x* - x_j of plain Jacobi PCG after `iterations` steps from zero on random
    right-hand sides b = A x*: ``(n, samples)``."""

    lanes = 32 if samples % 32 == 0 else 8
    plan = native.PcgPlan(np.array([n]), samples, lanes, 64, np.zeros(1), np.zeros(1), np.zeros(1),
                          host_pattern=host_pattern)
    xs = torch.zeros((1, n, plan.rp), dtype=torch.float64, device=values.device)
    xs[0, :, :samples] = torch.as_tensor(rng.standard_normal((n, samples)), device=values.device)
    b = torch.zeros_like(xs)
    layout = (np.array([n], np.int32), np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64),
              np.zeros(1, np.int64), lanes, plan.rp)
    native.spmm(layout, indptr_d, indices_d, values.reshape(1, -1).contiguous(), xs, b, 0)
    x = torch.zeros_like(xs)
    plan.solve(indptr_d, indices_d, values.reshape(1, -1).contiguous(), b, x, 1e-300, iterations)
    plan.close()
    return (xs - x)[0, :, :samples].contiguous()


def train(net: DampingNetwork, systems: list[TrainingSystem], *, steps: int = 300, lr: float = 3e-3,
          log=None) -> list[float]:
    """Adam on the FP32 master weights, one system per step (round robin); returns the losses."""

    opt = torch.optim.Adam([net.master], lr=lr)
    losses = []
    t0 = time.perf_counter()
    for step in range(steps):
        sys = systems[step % len(systems)]
        opt.zero_grad()
        omega = damping(net, sys.graph, sys.values, sys.omega_p)
        loss = two_grid_loss(sys, omega)
        loss.backward()
        opt.step()
        losses.append(loss.item())
        if log is not None and (step % 50 == 0 or step == steps - 1):
            log(f"step {step}: loss {losses[-1]:.5f} ({time.perf_counter() - t0:.1f}s)")
    return losses

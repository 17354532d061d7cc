"""FP64 restatement of the learned multigrid damping network (test infrastructure only).

This is the oracle for the GPU graph network of ``ad-tlert_amd/csrc/gnn.hip``; it has no
counterpart in the AD-TLERT reference (the reference solves directly with cuDSS).  The
network maps one matrix A (CSR, symmetric pattern) to a damping factor per matrix slot that
replaces the constant prolongation damping of the finest multigrid level:

  node features  x_i = [log10(a_ii / mean_j a_jj), sum_{j!=i} |a_ij| / a_ii, sum_j a_ij / a_ii, 1]
  messages       w_ij = |a_ij| / sum_{j'!=i} |a_ij'|  (off-diagonal slots; 0 on the diagonal)
  layer 0        h0 = relu(x W0 + b0)                                   (64 channels)
  layers 1..3    m = Wmsg h_{l-1};  h_l = relu([h_{l-1}, m] W_l + b_l)
  slot head      z_k = [h3_i, h3_j, a_ij / sqrt(a_ii a_jj), [i == j], same aggregate, strong]
                 (strong: off-diagonal, a_ij != 0, |a_ij| >= theta sqrt(a_ii a_jj))
                 s_k = relu(z_k U + c) . v + d;   omega_k = omega_p (1 + tanh(s_k) / 2)

Weights are used as rounded to bfloat16 (round to nearest even), as the GPU stores them.
``backward`` is the analytic reverse pass (checked against central differences in the tests).
"""

from __future__ import annotations

import numpy as np

HIDDEN = 64
HEAD = 32
NODE_FEATURES = 4
EDGE_FEATURES = 4
LAYERS = 3
LAYOUT = [("W0", (NODE_FEATURES, HIDDEN)), ("b0", (HIDDEN,))]
for _l in range(1, LAYERS + 1):
    LAYOUT += [(f"W{_l}", (2 * HIDDEN, HIDDEN)), (f"b{_l}", (HIDDEN,))]
LAYOUT += [("U", (2 * HIDDEN + EDGE_FEATURES, HEAD)), ("c", (HEAD,)), ("v", (HEAD,)), ("d", (1,))]
PARAM_COUNT = int(sum(np.prod(s) for _, s in LAYOUT))


def unpack(flat):
    out, o = {}, 0
    for name, shape in LAYOUT:
        size = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + size], dtype=np.float64).reshape(shape)
        o += size
    return out


def pack(params):
    return np.concatenate([np.asarray(params[name], dtype=np.float64).reshape(-1) for name, _ in LAYOUT])


def bf16_round(x):
    """Round float values to bfloat16 (nearest, ties to even), returned as float64."""

    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(indptr, indices, values, flags, theta=0.08):
    """Node features (n, 4), message weights (nnz,), slot features (nnz, 4), slot rows (nnz,).
    ``flags`` bit 1 marks slots inside one aggregate; strong slots follow |a_ij| >= theta sqrt(a_ii a_jj)."""

    n = indptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(indptr))
    diag_mask = rows == indices
    diag = np.zeros(n)
    diag[rows[diag_mask]] = values[diag_mask]
    off = ~diag_mask
    s_off = np.bincount(rows[off], weights=np.abs(values[off]), minlength=n)
    s_row = np.bincount(rows, weights=values, minlength=n)
    x = np.column_stack((np.log10(diag / diag.mean()), s_off / diag, s_row / diag, np.ones(n)))
    w = np.where(off, np.abs(values), 0.0)
    denom = s_off[rows]
    w = np.where(denom > 0, w / np.where(denom > 0, denom, 1.0), 0.0)
    scaled = values / np.sqrt(diag[rows] * diag[indices])
    strong = off & (values != 0) & (np.abs(scaled) >= theta)
    e = np.column_stack((scaled, diag_mask.astype(float), ((flags >> 1) & 1).astype(float), strong.astype(float)))
    return x, w, e, rows


def forward(flat, indptr, indices, values, flags, omega_p, theta=0.08):
    p = {k: (bf16_round(v)) for k, v in unpack(flat).items()}
    x, w, e, rows = features(indptr, indices, values, flags, theta)
    n = x.shape[0]
    msg = lambda h: np.stack([np.bincount(rows, weights=w * h[indices, c], minlength=n)  # noqa: E731
                              for c in range(h.shape[1])], axis=1)
    hs, ms = [np.maximum(x @ p["W0"] + p["b0"], 0.0)], [None]
    for l in range(1, LAYERS + 1):
        m = msg(hs[-1])
        ms.append(m)
        hs.append(np.maximum(np.concatenate((hs[-1], m), axis=1) @ p[f"W{l}"] + p[f"b{l}"], 0.0))
    z = np.concatenate((hs[-1][rows], hs[-1][indices], e), axis=1)
    u = np.maximum(z @ p["U"] + p["c"], 0.0)
    s = u @ p["v"] + p["d"][0]
    omega = omega_p * (1.0 + 0.5 * np.tanh(s))
    cache = dict(p=p, x=x, w=w, e=e, rows=rows, hs=hs, ms=ms, z=z, u=u, s=s, omega_p=omega_p, indices=indices)
    return omega, cache


def backward(cache, domega):
    """Gradients of sum(domega * omega) with respect to the (bf16-rounded) weights, packed."""

    p, rows, indices, w = cache["p"], cache["rows"], cache["indices"], cache["w"]
    hs, ms, z, u, s = cache["hs"], cache["ms"], cache["z"], cache["u"], cache["s"]
    n = hs[0].shape[0]
    g = {}
    ds = domega * cache["omega_p"] * 0.5 * (1.0 - np.tanh(s) ** 2)
    g["d"] = np.array([ds.sum()])
    g["v"] = u.T @ ds
    du = np.outer(ds, p["v"]) * (u > 0)
    g["c"] = du.sum(axis=0)
    g["U"] = z.T @ du
    dz = du @ p["U"].T
    H = HIDDEN
    dh = np.zeros((n, H))
    for c in range(H):
        dh[:, c] = np.bincount(rows, weights=dz[:, c], minlength=n) + np.bincount(indices, weights=dz[:, H + c],
                                                                               minlength=n)
    for l in range(LAYERS, 0, -1):
        da = dh * (hs[l] > 0)
        inp = np.concatenate((hs[l - 1], ms[l]), axis=1)
        g[f"W{l}"] = inp.T @ da
        g[f"b{l}"] = da.sum(axis=0)
        dinp = da @ p[f"W{l}"].T
        dh = dinp[:, :H].copy()
        dm = dinp[:, H:]
        # m_i = sum_k w_k h[col_k]  ->  dh[col_k] += w_k dm[row_k]
        for c in range(H):
            dh[:, c] += np.bincount(indices, weights=w * dm[rows, c], minlength=n)
    da = dh * (hs[0] > 0)
    g["W0"] = cache["x"].T @ da
    g["b0"] = da.sum(axis=0)
    return pack(g)

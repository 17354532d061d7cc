"""Seeded synthetic meshes, surveys and conductivity models.

These build the workloads the forward path is measured and tested on:

* the reference test fixtures (flat terrain-quad strip ``tests/conftest.py:27-51`` via
  ``workflows/terrain.py:748-807``, the 3D ``box_mesh``/``line_survey`` of
  ``tests/test_forward3d_gpu.py:38-75``, ``grid_triangles`` of ``tests/test_mesh_fem.py:18-34``);
* the five throughput configurations: graded P1 triangle meshes for the 2.5D cases
  (4k / 20k / 50k nodes), a graded 6-tets-per-hex box for 3D P1 (~200k nodes, ~1.1M tets)
  and a terrain box for 3D P2.

Electrodes always sit on mesh nodes.  Everything is plain numpy; nothing here touches a GPU.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

REMOTE = -1


# --------------------------------------------------------------------------------------------
# surveys
# --------------------------------------------------------------------------------------------


def wenner_alpha(count: int) -> np.ndarray:
    """Wenner-alpha ABMN rows in pyGIMLi 'wa' order (``workflows/terrain.py:303-315``)."""

    if count < 4:
        raise ValueError("Wenner-alpha surveys need at least four electrodes")
    return np.asarray(
        [[a, a + 3 * s, a + s, a + 2 * s] for s in range(1, count // 3 + 1) for a in range(count - 3 * s)],
        dtype=np.int32,
    )


def dipole_dipole(count: int, spacings=(1, 2, 3), separations=range(1, 9), limit=None) -> np.ndarray:
    """Dipole-dipole rows ``A, B=A+s, M=B+n s, N=M+s`` in generation order, truncated to ``limit``."""

    rows = [
        [a, a + s, a + s + n * s, a + 2 * s + n * s]
        for s in spacings
        for n in separations
        for a in range(count)
        if a + 2 * s + n * s < count
    ]
    rows = np.asarray(rows, dtype=np.int32).reshape(-1, 4)
    return rows if limit is None else rows[:limit]


def wenner_schlumberger(count: int, limit=None) -> np.ndarray:
    """Wenner-Schlumberger rows ``A, M=A+n a, N=M+a, B=N+n a`` for spacings ``a`` and factors ``n``."""

    rows = []
    for a in range(1, count):
        for n in range(1, count):
            span = (2 * n + 1) * a
            if span >= count:
                break
            rows += [[start, start + span, start + n * a, start + n * a + a] for start in range(count - span)]
    rows = np.asarray(rows, dtype=np.int32).reshape(-1, 4)
    return rows if limit is None else rows[:limit]


def random_quadrupoles(count: int, rows: int, rng: np.random.Generator, positions=None,
                       max_factor: float = 2.0e4) -> np.ndarray:
    """``rows`` distinct random ABMN quadruples of four distinct electrodes.

    With electrode ``positions``, quadruples whose half-space geometric factor
    ``2 pi / (1/AM - 1/AN - 1/BM + 1/BN)`` exceeds ``max_factor`` metres in magnitude are drawn
    again: on a regular grid some arrangements are null couplings (the denominator vanishes and
    the factor is infinite), which no survey would record."""

    seen, out = set(), []
    while len(out) < rows:
        quad = tuple(int(x) for x in rng.choice(count, 4, replace=False))
        if quad in seen:
            continue
        seen.add(quad)
        if positions is not None:
            a, b, m, n = (np.asarray(positions[i], dtype=float) for i in quad)
            coupling = (1 / np.linalg.norm(a - m) - 1 / np.linalg.norm(a - n) - 1 / np.linalg.norm(b - m)
                        + 1 / np.linalg.norm(b - n))
            if abs(coupling) * max_factor < 2 * np.pi:
                continue
        out.append(quad)
    return np.asarray(out, dtype=np.int32)


# --------------------------------------------------------------------------------------------
# reference test fixtures
# --------------------------------------------------------------------------------------------


@dataclass(frozen=True)
class Case2D:
    nodes: np.ndarray
    cells: np.ndarray
    electrodes: np.ndarray
    measurements: np.ndarray
    conductivity: np.ndarray
    surface_node_ids: np.ndarray | None = None


@dataclass(frozen=True)
class Case3D:
    nodes: np.ndarray
    cells: np.ndarray
    electrodes: np.ndarray
    measurements: np.ndarray
    conductivity: np.ndarray


def quad_strip_case(nx=30, nz=12, n_electrodes=16, anomaly=True, slope=0.0) -> Case2D:
    """Terrain-following quadrilateral strip of ``tests/conftest.py:27-51`` (dx = 2 m, layers
    2.0 -> 0.5 m; ``workflows/terrain.py:748-797``): the surface rises by ``slope`` per metre plus
    ``0.05 sin(column)`` where the slope is not zero; ``slope = 0`` is flat."""

    thickness = np.linspace(2.0, 0.5, nz)[::-1]
    offsets = np.concatenate(([0.0], -np.cumsum(thickness)))
    xs = np.arange(nx + 1, dtype=float) * 2.0
    slopes = np.full(nx, float(slope)) + (0.0 if slope == 0 else 0.05 * np.sin(np.arange(nx)))
    top = np.concatenate(([0.0], np.cumsum(slopes * 2.0)))
    nodes = np.asarray([[x, z + off] for off in offsets for x, z in zip(xs, top)])
    width = nx + 1
    cells = np.asarray(
        [[l * width + c, l * width + c + 1, (l + 1) * width + c + 1, (l + 1) * width + c]
         for l in range(nz) for c in range(nx)], dtype=np.int64)
    rho = np.full((nz, nx), 100.0)
    if anomaly:
        rho[3:7, 10:18] = 30.0
    rho = rho[::-1, :].reshape(-1)
    count = min(int(n_electrodes), width)
    ex = np.linspace(xs.min(), xs.max(), count)
    return Case2D(nodes, cells, np.column_stack((ex, np.interp(ex, xs, top))), wenner_alpha(count), 1.0 / rho,
                  np.arange(width))


def grid_triangles(nx=6, ny=4, slope=0.0):
    """Unit grid split into two triangles per square (``tests/test_mesh_fem.py:18-34``)."""

    xs, ys = np.meshgrid(np.arange(nx + 1, dtype=float), -np.arange(ny + 1, dtype=float))
    nodes = np.column_stack((xs.ravel(), ys.ravel() + slope * xs.ravel()))
    index = np.arange((nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1)
    cells = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = index[j, i], index[j, i + 1], index[j + 1, i + 1], index[j + 1, i]
            cells += [[a, d, b], [b, d, c]]
    return nodes, np.asarray(cells)


HEX_TETS = ((1, 2, 6), (2, 3, 6), (3, 7, 6), (7, 4, 6), (4, 5, 6), (5, 1, 6))


def hex_tets(xs, ys, zs, relief=None):
    """Tensor grid split into six tetrahedra around the 0-6 diagonal of every hexahedron."""

    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    if relief is not None:
        Z = relief(X, Y, Z)
    nodes = np.column_stack((X.ravel(), Y.ravel(), Z.ravel()))
    nx, ny, nz = len(xs) - 1, len(ys) - 1, len(zs) - 1
    idx = np.arange(nodes.shape[0]).reshape(nx + 1, ny + 1, nz + 1)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    v = np.stack([idx[i, j, k], idx[i + 1, j, k], idx[i + 1, j + 1, k], idx[i, j + 1, k],
                  idx[i, j, k + 1], idx[i + 1, j, k + 1], idx[i + 1, j + 1, k + 1], idx[i, j + 1, k + 1]], axis=1)
    tets = np.stack([np.stack([v[:, 0], v[:, a], v[:, b], v[:, c]], axis=1) for a, b, c in HEX_TETS], axis=1)
    return nodes, tets.reshape(-1, 4)


def box_mesh(nx=6, ny=6, nz=4, relief=0.0):
    """12 x 12 x 8 m box of ``tests/test_forward3d_gpu.py:38-65`` (same node and tet order)."""

    xs, ys, zs = np.linspace(0, 12, nx + 1), np.linspace(0, 12, ny + 1), np.linspace(-8, 0, nz + 1)

    def bend(X, Y, Z):
        return Z + relief * (Z - Z.min()) / (Z.max() - Z.min()) * np.sin(X / 4.0)

    return hex_tets(xs, ys, zs, bend)


def line_survey(nodes):
    """Wenner line on y = 6 at the top surface (``tests/test_forward3d_gpu.py:68-75``)."""

    on = np.flatnonzero(np.isclose(nodes[:, 1], 6.0) & (nodes[:, 2] >= nodes[:, 2].max() - 1e-9))
    on = on[np.argsort(nodes[on, 0])]
    return nodes[on], wenner_alpha(len(on))


def box_case() -> Case3D:
    nodes, tets = box_mesh()
    sigma = np.full(tets.shape[0], 0.01)
    sigma[: tets.shape[0] // 3] = 0.03
    electrodes, quads = line_survey(nodes)
    return Case3D(nodes, tets, electrodes, quads, sigma)


# --------------------------------------------------------------------------------------------
# throughput configurations
# --------------------------------------------------------------------------------------------


def graded_axis(inner_lo, inner_hi, step, outer_lo, outer_hi, growth):
    """Uniform ``step`` on ``[inner_lo, inner_hi]`` growing geometrically to the outer bounds."""

    inner = inner_lo + step * np.arange(int(round((inner_hi - inner_lo) / step)) + 1)
    left, right, h = [], [], step
    x = inner[0]
    while x > outer_lo + 1e-9:
        h *= growth
        x = max(x - h, outer_lo)
        if x - outer_lo < 0.5 * h:
            x = outer_lo
        left.append(x)
    x, h = inner[-1], step
    while x < outer_hi - 1e-9:
        h *= growth
        x = min(x + h, outer_hi)
        if outer_hi - x < 0.5 * h:
            x = outer_hi
        right.append(x)
    return np.concatenate((left[::-1], inner, right))


def depth_axis(first, growth, depth):
    z, h, out = 0.0, first, [0.0]
    while z > -depth + 1e-9:
        z = max(z - h, -depth)
        if z + depth < 0.5 * h * growth:
            z = -depth
        out.append(z)
        h *= growth
    return np.asarray(out[::-1])


def tensor_triangles(xs, ys):
    """Rectangular tensor grid, two triangles per rectangle (``grid_triangles`` pattern)."""

    ys = np.sort(ys)[::-1]
    X, Y = np.meshgrid(xs, ys)
    nodes = np.column_stack((X.ravel(), Y.ravel()))
    nx, ny = len(xs) - 1, len(ys) - 1
    index = np.arange(nodes.shape[0]).reshape(ny + 1, nx + 1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    a, b, c, d = index[j, i], index[j, i + 1], index[j + 1, i + 1], index[j + 1, i]
    cells = np.stack((np.stack((a, d, b), -1), np.stack((b, d, c), -1)), axis=2).reshape(-1, 3)
    return nodes, cells


def log_conductivity_2d(centers, rng, lo=1.0, hi=3.0, anomalies=2):
    """log10 resistivity ~ U(lo, hi) per cell plus Gaussian anomalies -> conductivity."""

    log_rho = rng.uniform(lo, hi, centers.shape[0])
    x0, x1 = centers[:, 0].min(), centers[:, 0].max()
    y0 = centers[:, 1].min()
    for _ in range(anomalies):
        cx, cy = rng.uniform(x0 + 0.2 * (x1 - x0), x1 - 0.2 * (x1 - x0)), rng.uniform(0.6 * y0, 0.1 * y0)
        width, amplitude = rng.uniform(3.0, 8.0), rng.uniform(-1.0, 1.0)
        log_rho += amplitude * np.exp(-((centers[:, 0] - cx) ** 2 + (centers[:, 1] - cy) ** 2) / (2 * width ** 2))
    return 10.0 ** (-log_rho)


def smooth_fields_2d(centers, rng, count, length=8.0, modes=48):
    """``count`` seeded smooth Gaussian-random-like fields (random Fourier features), unit variance."""

    fields = []
    for _ in range(count):
        k = rng.normal(0.0, 1.0 / length, size=(modes, 2))
        phase = rng.uniform(0, 2 * np.pi, modes)
        f = np.sqrt(2.0 / modes) * np.cos(centers @ k.T + phase).sum(axis=1)
        fields.append(f / max(f.std(), 1e-12))
    return np.stack(fields)


def config_2p5d(node_target: int, electrodes: int, spacing: float, seed: int, survey: str, rows: int,
                step: float, dd_spacings=(1, 2, 3, 4)):
    """Graded flat P1 triangle mesh with ``electrodes`` on nodes at ``spacing`` from x = 0."""

    rng = np.random.default_rng(seed)
    span = spacing * (electrodes - 1)
    pad = max(0.3 * span, 10.0)
    xs = graded_axis(-2 * spacing, span + 2 * spacing, step, -pad, span + pad, 1.2)
    depth = 0.35 * span + 10.0
    # choose the surface step so the mesh has about node_target nodes
    rows_needed = max(int(node_target / len(xs)), 10)
    lo, hi = 0.01, 1.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if len(depth_axis(step / 2, 1.0 + mid, depth)) > rows_needed:
            lo = mid
        else:
            hi = mid
    zs = depth_axis(step / 2, 1.0 + hi, depth)
    nodes, cells = tensor_triangles(xs, zs)
    ex = spacing * np.arange(electrodes)
    positions = np.column_stack((ex, np.zeros_like(ex)))
    if survey == "dd":
        quads = dipole_dipole(electrodes, spacings=dd_spacings, limit=rows)
    else:
        quads = wenner_schlumberger(electrodes, limit=rows)
    centers = nodes[cells].mean(axis=1)
    sigma = log_conductivity_2d(centers, rng)
    return Case2D(nodes, cells, positions, quads, sigma), rng


def strip_dd_4k(seed=1):
    """~4k nodes, 48 electrodes at 2 m, ~1000 dipole-dipole rows."""

    return config_2p5d(4000, 48, 2.0, seed, "dd", 1000, 1.0)[0]


def timelapse_ws_20k(seed=2, steps=365, node_target=20000):
    """~20k nodes, 96 electrodes at 1 m, ~3000 Wenner-Schlumberger rows, a seeded time-lapse series.

    ``log10 sigma_t = log10 sigma_0 + 0.2 * sum_k a_k(t) G_k / |a|`` with four smooth spatial fields
    ``G_k`` and low-order sinusoidal temporal coefficients.
    """

    case, rng = config_2p5d(node_target, 96, 1.0, seed, "ws", 3000, 0.5)
    centers = case.nodes[case.cells].mean(axis=1)
    fields = smooth_fields_2d(centers, rng, 4)
    t = np.arange(steps)[:, None]
    freq = rng.uniform(0.5, 3.0, 4)[None, :]
    phase = rng.uniform(0, 2 * np.pi, 4)[None, :]
    coeff = np.sin(2 * np.pi * freq * t / 365.0 + phase) / np.sqrt(2.0)
    perturb = 0.2 * (coeff @ fields) / np.sqrt(4 * 0.5)
    series = case.conductivity[None, :] * 10.0 ** perturb
    return case, series


def mixed_meshes_ws(seed=6, node_targets=(5000, 10000, 15000, 20000, 25000, 30000, 35000, 40000), electrodes=96,
                    rows=3000):
    """The time-lapse survey (``electrodes`` at 1 m, Wenner-Schlumberger rows) on graded meshes of
    ``node_targets`` nodes, one seeded conductivity model each (a mixed batch of uneven meshes)."""

    out = []
    for k, target in enumerate(node_targets):
        step = 0.5 * np.sqrt(20000.0 / target)  # finer surface steps on the larger meshes
        out.append(config_2p5d(target, electrodes, 1.0, seed + 100 * k, "ws", rows, step)[0])
    return out


def strip_dd_50k(seed=3):
    """~50k nodes (49 776 nodes, 98 550 triangles), 64 electrodes at 1 m, 2000 dipole-dipole rows."""

    return config_2p5d(50000, 64, 1.0, seed, "dd", 2000, 0.2, dd_spacings=(1, 2, 3, 4, 5, 6, 7))[0]


def grf_3d(points, rng, mean=-2.0, std=0.5, length=20.0, grid=2.0):
    """log10 conductivity Gaussian random field (exponential spectrum via FFT), sampled at points."""

    lo, hi = points.min(axis=0), points.max(axis=0)
    shape = np.maximum(((hi - lo) / grid).astype(int) + 2, 4)
    noise = rng.standard_normal(shape)
    freqs = np.meshgrid(*[np.fft.fftfreq(n, d=grid) for n in shape], indexing="ij")
    k2 = sum(f * f for f in freqs)
    spectrum = (1.0 + (2 * np.pi) ** 2 * k2 * length ** 2) ** (-1.0)
    field = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.sqrt(spectrum)))
    field = (field - field.mean()) / field.std()
    idx = np.clip(((points - lo) / grid).astype(int), 0, np.asarray(shape) - 1)
    return 10.0 ** (mean + std * field[idx[:, 0], idx[:, 1], idx[:, 2]])


def box_p1_200k(seed=4, fine=1.25, growth=1.13, surface_step=0.5, depth_growth=1.06, rows=2000):
    """200 x 200 x 100 m box, 8 x 8 electrodes at 5 m, 202 005 nodes / 1 149 984 tets, random ABMN rows."""

    rng = np.random.default_rng(seed)
    grid = 5.0 * (np.arange(8) - 3.5)
    xs = graded_axis(grid[0] - fine, grid[-1] + fine, fine, -100.0, 100.0, growth)
    zs = depth_axis(surface_step, depth_growth, 100.0)
    nodes, tets = hex_tets(xs, xs, zs)
    gx, gy = np.meshgrid(grid, grid, indexing="ij")
    positions = np.column_stack((gx.ravel(), gy.ravel(), np.zeros(64)))
    quads = random_quadrupoles(64, rows, rng, positions)
    centers = nodes[tets].mean(axis=1)
    return Case3D(nodes, tets, positions, quads, grf_3d(centers, rng))


def sinusoid_relief(rng, count=6, amplitude=5.0):
    amps = rng.uniform(-amplitude, amplitude, count) / np.sqrt(count)
    kx, ky = rng.uniform(-0.05, 0.05, count), rng.uniform(-0.05, 0.05, count)
    ph = rng.uniform(0, 2 * np.pi, count)

    def surface(x, y):
        return sum(a * np.sin(u * x + v * y + p) for a, u, v, p in zip(amps, kx, ky, ph))

    return surface


def terrain_p2_150k(seed=5, fine=2.5, growth=1.25, surface_step=1.0, depth_growth=1.15, rows=2000):
    """Terrain box for ten-node tets: ~25k hexes -> ~150k tets, surface relief from 6 sinusoids."""

    rng = np.random.default_rng(seed)
    grid = 5.0 * (np.arange(8) - 3.5)
    xs = graded_axis(grid[0] - fine, grid[-1] + fine, fine, -100.0, 100.0, growth)
    zs = depth_axis(surface_step, depth_growth, 100.0)
    surface = sinusoid_relief(rng)

    def bend(X, Y, Z):
        return Z + surface(X, Y) * (Z - Z.min()) / (Z.max() - Z.min())

    nodes, tets = hex_tets(xs, xs, zs, bend)
    gx, gy = np.meshgrid(grid, grid, indexing="ij")
    positions = np.column_stack((gx.ravel(), gy.ravel(), surface(gx.ravel(), gy.ravel())))
    quads = random_quadrupoles(64, rows, rng, positions)
    centers = nodes[tets].mean(axis=1)
    return Case3D(nodes, tets, positions, quads, grf_3d(centers, rng))

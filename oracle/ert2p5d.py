"""2.5D flat-surface ERT forward operator, Jacobian and adjoints (float64, direct solves).

Follows ``adtlert/forward/ert2p5d.py`` (line ranges cited per method) on the
native discretization of ``adtlert/forward/discretization.py:485-575``
(``auxiliary=False``: the input P1-triangle or Q1-quad mesh).  Test infrastructure only.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.special import k0 as bessel_k0

from . import fem, mesh as meshlib, quadrature, survey as surveylib

RADIUS_FRACTION = 1.0 / 6.0  # ert2p5d.py:61


def halfspace_primary(nodes, source, wavenumber, surface):
    """Unit-resistivity half-space primary field with image (``ert2p5d.py:78-94``)."""

    distance = np.linalg.norm(nodes - source, axis=1)
    values = np.zeros(nodes.shape[0])
    if abs(source[1] - surface) <= 1e-8:
        ok = distance > 1e-12
        values[ok] = bessel_k0(distance[ok] * wavenumber) / np.pi
        return values
    mirrored = np.linalg.norm(nodes - (source[0], 2.0 * surface - source[1]), axis=1)
    ok = (distance > 1e-12) & (mirrored > 1e-12)
    values[ok] = (bessel_k0(distance[ok] * wavenumber) + bessel_k0(mirrored[ok] * wavenumber)) / (2.0 * np.pi)
    return values


def csr_pattern(cell_dofs, boundary_dofs, n):
    """Unique (row, col) pairs -> CSR and the scatter maps of every local entry (``discretization.py:100-125``)."""

    def pairs(conn):
        w = conn.shape[1]
        return np.stack((np.repeat(conn, w, axis=1).reshape(-1), np.tile(conn, (1, w)).reshape(-1)), axis=1)

    volume = pairs(cell_dofs)
    unique, inverse = np.unique(np.concatenate((volume, pairs(boundary_dofs))), axis=0, return_inverse=True)
    inverse = inverse.reshape(-1)
    indptr = np.concatenate(([0], np.cumsum(np.bincount(unique[:, 0], minlength=n))))
    return indptr, unique[:, 1], inverse[: volume.shape[0]], inverse[volume.shape[0]:]


class Forward2p5D:
    """2.5D operator on a triangle (P1) or quadrilateral (Q1) mesh.

    Flat surfaces use the mesh as the discretization and the analytic primary field.  On terrain
    (``ert2p5d.py:170-190, 404-436, 662-671``) ``terrain`` supplies the auxiliary discretizations
    as arrays -- ``field`` (the secondary-field space), ``potential`` (the space whose
    unit-conductivity potentials, sampled at the field DOFs, are the primary field) and optionally
    ``geometric`` (numerical geometric factors) -- each a dict with ``cell_dofs``, ``dof_nodes``,
    ``boundary_dofs``, ``parents``, ``stiffness``, ``mass``, ``boundary_mass``, ``alpha`` (W, B),
    ``electrodes`` (E, dofs) and ``boundary_cells``; everything from there on (assembly, the unit
    potential and field solves, measurement, derivatives) is computed here."""

    def __init__(self, nodes, cells, positions, quads, surface_node_ids=None, wavenumbers=None, weights=None,
                 terrain=None):
        self.mesh = meshlib.mesh2d(nodes, cells, surface_node_ids)
        if not self.mesh["flat"] and terrain is None:
            raise NotImplementedError("terrain meshes need their auxiliary discretizations (terrain=...)")
        self.positions, self.quads = surveylib.validate(positions, quads)
        self.E, self.D = self.positions.shape[0], self.quads.shape[0]
        if wavenumbers is None:
            wavenumbers, weights = quadrature.inverse_cosine_weights(*quadrature.wavenumber_bounds(self.positions))
        self.k, self.q = np.asarray(wavenumbers, float), np.asarray(weights, float)
        self.W = self.k.size
        m = self.mesh
        nodes, cells = m["nodes"], m["cells"]
        self.C = cells.shape[0]
        cell_ids, bary = meshlib.locate2d(m, self.positions)
        d = np.linalg.norm(nodes[None] - self.positions[:, None], axis=2)
        nearest = np.argmin(d, axis=1)
        self.source_nodes = np.where(d[np.arange(self.E), nearest] <= 1e-2, nearest, -1)  # ert2p5d.py:192-198
        self.source_cells = cell_ids
        self.slots = surveylib.measurement_slots(self.quads, self.E)
        self.terrain = terrain
        if terrain is None:
            if cells.shape[1] == 3:
                stiffness, mass = fem.p1_triangle(nodes, cells)
            else:
                stiffness, mass = fem.q1_quad(nodes, cells)
            center = self.positions.mean(axis=0)
            electrodes = np.zeros((self.E, nodes.shape[0]))
            np.add.at(electrodes, (np.repeat(np.arange(self.E), cells.shape[1]), cells[cell_ids].reshape(-1)),
                      bary.reshape(-1))
            field = dict(cell_dofs=cells, dof_nodes=nodes, boundary_dofs=m["boundary_edges"], parents=np.arange(self.C),
                         stiffness=stiffness, mass=mass, boundary_mass=fem.edge_mass(m["boundary_edge_lengths"]),
                         alpha=np.stack([fem.robin_2p5d(m, center, kw) for kw in self.k]), electrodes=electrodes,
                         boundary_cells=m["boundary_edge_cells"])
        else:
            field = terrain["field"]
        self._use(field)
        self.primary = self._analytic_primary() if terrain is None else self._numerical_primary()
        self.b = np.stack([self.matrix(np.ones(self.C), w) @ self.primary[w].T for w in range(self.W)]).transpose(0, 2, 1)
        self._cache = {}

    def _use(self, d):
        """Make ``d`` the discretization of the operator's assembly, solves and derivatives."""

        self.dcells = np.asarray(d["cell_dofs"], dtype=np.int64)
        self.parents = np.asarray(d["parents"], dtype=np.int64)
        self.N = np.asarray(d["dof_nodes"]).shape[0]
        self.stiffness, self.mass, self.bmass = (np.asarray(d[k], float) for k in ("stiffness", "mass", "boundary_mass"))
        self.alpha = np.asarray(d["alpha"], float)
        self.bdofs = np.asarray(d["boundary_dofs"], dtype=np.int64)
        self.bcells = np.asarray(d["boundary_cells"], dtype=np.int64)
        self.electrodes = np.asarray(d["electrodes"], float)
        self.indptr, self.indices, self.vinv, self.binv = csr_pattern(self.dcells, self.bdofs, self.N)
        self.nnz = self.indices.size

    def unit_potentials(self, d):
        """``(W, E, dofs)`` solutions of ``A_w(1) U = electrode rows`` on discretization ``d``
        (``ert2p5d.py:391-402``), by sparse LU."""

        saved = {k: getattr(self, k) for k in ("dcells", "parents", "N", "stiffness", "mass", "bmass", "alpha",
                                               "bdofs", "bcells", "electrodes", "indptr", "indices", "vinv", "binv",
                                               "nnz")}
        self._use(d)
        try:
            rhs = self.electrodes.T
            return np.stack([spla.splu(self.matrix(np.ones(self.C), w).tocsc()).solve(rhs).T for w in range(self.W)])
        finally:
            for k, v in saved.items():
                setattr(self, k, v)

    def _numerical_primary(self):
        """Unit potentials of the potential discretization at the field DOF nodes (``ert2p5d.py:404-420``)."""

        potential, field = self.terrain["potential"], self.terrain["field"]
        pn, fn = np.asarray(potential["dof_nodes"], float), np.asarray(field["dof_nodes"], float)
        dist = np.linalg.norm(fn[:, None, :] - pn[None, :, :], axis=2)
        selection = np.argmin(dist, axis=1)
        if np.any(dist[np.arange(fn.shape[0]), selection] > 1e-6):
            raise ValueError("field DOF node not found among the primary-potential nodes")
        return self.unit_potentials(potential)[..., selection]

    # -- operator -------------------------------------------------------------------------

    def values(self, sigma, w, robin=True):
        """CSR values of ``A_w(sigma)`` (``ert2p5d.py:252-314``)."""

        sigma = np.asarray(sigma, float)[self.parents]  # per discretization cell
        vol = sigma[:, None, None] * (self.stiffness + self.k[w] ** 2 * self.mass)
        out = np.bincount(self.vinv, weights=vol.reshape(-1), minlength=self.nnz)
        if robin:
            bnd = (sigma[self.bcells] * self.alpha[w])[:, None, None] * self.bmass
            out = out + np.bincount(self.binv, weights=bnd.reshape(-1), minlength=self.nnz)
        return out

    def matrix(self, sigma, w, robin=True):
        return sp.csr_matrix((self.values(sigma, w, robin), self.indices, self.indptr), shape=(self.N, self.N))

    def _analytic_primary(self):
        """``(W, E, N)`` unit primary with the pyGIMLi electrode-node value (``ert2p5d.py:438-468``)."""

        nodes, cells = self.mesh["nodes"], self.mesh["cells"]
        surface = float(np.max(nodes[:, 1]))
        radii = {}
        for node in set(self.source_nodes.tolist()) - {-1}:
            nb = np.setdiff1d(np.unique(cells[np.any(cells == node, axis=1)]), [node])
            radii[node] = np.min(np.linalg.norm(nodes[nb] - nodes[node], axis=1)) if nb.size else None
        out = np.zeros((self.W, self.E, self.N))
        for w, kw in enumerate(self.k):
            for s in range(self.E):
                out[w, s] = halfspace_primary(nodes, self.positions[s], kw, surface)
                node = self.source_nodes[s]
                if node >= 0:
                    r = radii[node]
                    out[w, s, node] = 0.0 if r is None else bessel_k0(r * RADIUS_FRACTION * kw) / np.pi
        return out

    def source_resistivities(self, sigma):
        """Geometric-mean resistivity around each source node (``ert2p5d.py:480-511``)."""

        cells = self.mesh["cells"]
        out = np.empty(self.E)
        for s in range(self.E):
            node = self.source_nodes[s]
            if node >= 0:
                mask = np.any(cells == node, axis=1)
                out[s] = np.exp(np.mean(-np.log(sigma[mask])))
            else:
                out[s] = 1.0 / sigma[self.source_cells[s]]
        return out

    def fields(self, sigma):
        """Total fields ``(W, E, N)`` by the secondary-field solve (``ert2p5d.py:532-553``)."""

        sigma = np.asarray(sigma, float)
        key = sigma.tobytes()
        if key in self._cache:
            return self._cache[key]
        rho = self.source_resistivities(sigma)
        total = np.empty((self.W, self.E, self.N))
        lus = []
        for w in range(self.W):
            A = self.matrix(sigma, w).tocsc()
            lu = spla.splu(A)
            x0 = self.primary[w] * rho[:, None]
            rhs = self.b[w] - (A @ x0.T).T
            total[w] = lu.solve(rhs.T).T + x0
            lus.append(lu)
        self._cache = {key: (total, lus)}
        return total, lus

    # -- measurement ----------------------------------------------------------------------

    def integrate(self, fields):
        return np.tensordot(self.q, fields, axes=([0], [0]))

    def normal_reciprocal(self, integrated):
        """Normal and reciprocal transfer resistances from ``(E, E)`` potentials (``ert2p5d.py:574-589``)."""

        p = surveylib.with_remote_slot(integrated @ self.electrodes.T)
        a, b, m, n = self.slots.T

        def transfer(src, snk, pos, neg):
            return (p[..., src, pos] - p[..., src, neg]) - (p[..., snk, pos] - p[..., snk, neg])

        return transfer(a, b, m, n), transfer(m, n, a, b)

    def resistance(self, sigma):
        normal, reciprocal = self.normal_reciprocal(self.integrate(self.fields(sigma)[0]))
        return np.sign(normal) * np.sqrt(np.abs(normal * reciprocal))  # ert2p5d.py:641-645

    def geometric_factors(self):
        if self.terrain is None or self.terrain.get("geometric") is None:
            return surveylib.geometric_factors(self.positions, self.quads)
        g = self.terrain["geometric"]  # numerical factors (ert2p5d.py:662-671)
        integrated = np.tensordot(self.q, self.unit_potentials(g), axes=([0], [0]))
        p = surveylib.with_remote_slot(integrated @ np.asarray(g["electrodes"], float).T)
        a, b, m, n = self.slots.T
        sources = p[a] - p[b]
        rows = np.arange(self.D)
        return 1.0 / (sources[rows, m] - sources[rows, n])

    def apparent_resistivity(self, sigma, currents=1.0):
        return self.geometric_factors() * self.resistance(sigma) / currents

    # -- derivatives ----------------------------------------------------------------------

    def templates(self):
        """``(W, C, K, K)`` ``K + k_w^2 M`` (``ert2p5d.py:252-263``)."""

        return self.stiffness[None] + (self.k ** 2)[:, None, None, None] * self.mass[None]

    def columns(self, ids=None, count=None):
        if ids is None:
            return np.arange(self.C), self.C
        ids = np.asarray(ids, dtype=np.int64)
        return ids, int(ids.max()) + 1 if count is None else int(count)

    def aggregate(self, per_cell, ids=None, count=None):
        """Sum cell columns into parameters, dropping id -1 (``kernels.py:88-113``)."""

        targets, count = self.columns(ids, count)
        keep = targets >= 0
        out = np.zeros(per_cell.shape[:-1] + (count,))
        for p in range(count):
            out[..., p] = per_cell[..., keep & (targets == p)].sum(axis=-1)
        return out

    def jacobian(self, sigma, ids=None, count=None, chunk=64, rows=None):
        """Normal sensitivity ``-sum_w q_w (u_M-u_N)^T (K+k^2M) (u_A-u_B)`` (``ert2p5d.py:984-1010``, ``kernels.py:132-178``).

        ``rows`` (optional) selects measurement rows; the result then has ``len(rows)`` rows."""

        phi = np.concatenate((self.fields(sigma)[0], np.zeros((self.W, 1, self.N))), axis=1)
        T = self.q[:, None, None, None] * self.templates()
        cells = self.dcells
        slots = self.slots if rows is None else self.slots[np.asarray(rows)]
        a, b, m, n = slots.T
        rows = []
        for s in range(0, slots.shape[0], chunk):
            sl = slice(s, s + chunk)
            cur = (phi[:, a[sl]] - phi[:, b[sl]])[..., cells]
            rec = (phi[:, m[sl]] - phi[:, n[sl]])[..., cells]
            rows.append(self._to_cells(-np.einsum("wdci,wcij,wdcj->dc", rec, T, cur)))
        return self.aggregate(np.concatenate(rows), ids, count)

    def _to_cells(self, per_dcell):
        """Sum discretization-cell columns into their mesh cells (the parents)."""

        if self.dcells.shape[0] == self.C and np.array_equal(self.parents, np.arange(self.C)):
            return per_dcell
        out = np.zeros(per_dcell.shape[:-1] + (self.C,))
        for c in range(self.C):
            out[..., c] = per_dcell[..., self.parents == c].sum(axis=-1)
        return out

    def pair_cotangent(self, g_normal, g_reciprocal=None):
        """Transpose of :meth:`normal_reciprocal` onto ``(E, E)`` (``ert2p5d.py:591-629``)."""

        slots = self.E + 1
        pairs = np.zeros(slots * slots)
        a, b, m, n = self.slots.T
        quads = [((a, b, m, n), g_normal)] + ([((m, n, a, b), g_reciprocal)] if g_reciprocal is not None else [])
        for (src, snk, pos, neg), g in quads:
            np.add.at(pairs, src * slots + pos, g)
            np.add.at(pairs, src * slots + neg, -g)
            np.add.at(pairs, snk * slots + pos, -g)
            np.add.at(pairs, snk * slots + neg, g)
        return pairs.reshape(slots, slots)[: self.E, : self.E]

    def cell_gradient(self, phi, lam, robin=True):
        """``-sum_ws lam^T dA_w/dsigma_c phi`` (``ert2p5d.py:765-825``)."""

        cells = self.dcells
        g = -np.einsum("wsci,wcij,wscj->c", lam[..., cells], self.templates(), phi[..., cells])
        if robin:
            bt = self.alpha[:, :, None, None] * self.bmass[None]
            per_edge = -np.einsum("wsbi,wbij,wsbj->b", lam[..., self.bdofs], bt, phi[..., self.bdofs])
            g = g + np.bincount(self.bcells, weights=per_edge, minlength=cells.shape[0])
        return np.bincount(self.parents, weights=g, minlength=self.C)

    def _combine_scale(self, normal, reciprocal):
        return np.sign(reciprocal), np.maximum(np.sqrt(np.abs(normal * reciprocal)), 1e-30)  # ert2p5d.py:647-655

    def vjp(self, sigma, cotangent):
        """Exact reverse product with the reciprocal-averaged response (``ert2p5d.py:835-871``)."""

        total, lus = self.fields(sigma)
        normal, reciprocal = self.normal_reciprocal(self.integrate(total))
        sign, scale = self._combine_scale(normal, reciprocal)
        common = 0.5 * np.asarray(cotangent, float) * sign / scale
        pairs = self.pair_cotangent(common * reciprocal, common * normal)
        lam = np.stack([lus[w].solve((self.q[w] * (pairs @ self.electrodes)).T).T for w in range(self.W)])
        return self.cell_gradient(total, lam)

    def jvp(self, sigma, direction):
        """Exact forward product (``ert2p5d.py:873-890``)."""

        total, lus = self.fields(sigma)
        normal, reciprocal = self.normal_reciprocal(self.integrate(total))
        delta = np.stack([lus[w].solve(-(self.matrix(direction, w) @ total[w].T)) .T for w in range(self.W)])
        dn, dr = self.normal_reciprocal(self.integrate(delta))
        sign, scale = self._combine_scale(normal, reciprocal)
        return 0.5 * sign * (dn * reciprocal + normal * dr) / scale

    def normal_vjp(self, sigma, cotangent, ids=None, count=None):
        """Transpose of the normal sensitivity (``ert2p5d.py:892-912``)."""

        phi = self.fields(sigma)[0]
        pairs = self.pair_cotangent(np.broadcast_to(np.asarray(cotangent, float), (self.D,))).T
        current = np.einsum("ef,wfn->wen", pairs, phi)
        receiver = self.q[:, None, None] * phi
        return self.aggregate(self.cell_gradient(current, receiver, robin=False), ids, count)

    def normal_jvp(self, sigma, direction):
        """Normal sensitivity times a direction, volume terms only (``ert2p5d.py:914-943``)."""

        phi = self.fields(sigma)[0]
        gram = np.zeros((self.E, self.E))
        for w in range(self.W):
            tangent = self.matrix(direction, w, robin=False)
            gram += self.q[w] * phi[w] @ (-(tangent @ phi[w].T))
        gram = surveylib.with_remote_slot(gram)
        a, b, m, n = self.slots.T
        return gram[m, a] - gram[m, b] - gram[n, a] + gram[n, b]

    def exact_jacobian(self, sigma):
        """Rows ``vjp(e_d)`` of the exact derivative incl. Robin terms (``ert2p5d.py:1012-1050``)."""

        return np.stack([self.vjp(sigma, np.eye(self.D)[d]) for d in range(self.D)])

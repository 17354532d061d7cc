"""3D tetrahedral P1/P2 ERT forward operator, Jacobian and adjoints (float64, direct solves).

Follows ``adtlert/forward/ert3d.py`` (line ranges cited per method).  Test infrastructure only.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import fem, mesh as meshlib, survey as surveylib

RADIUS_FRACTION = 0.5  # ert3d.py:47


def factorize(matrix):
    """Sparse LU of a symmetric matrix with a minimum-degree ordering of ``A^T + A`` (SuperLU's
    column ordering for symmetric patterns: about half the fill of COLAMD on tetrahedral meshes)."""

    return spla.splu(matrix.tocsc(), permc_spec="MMD_AT_PLUS_A")


def electrode_rows(slots, count):
    """Electrodes used by ``slots`` and the survey-slot -> field-row map (``ert3d.py:93-106``)."""

    ids = np.unique(slots)
    ids = ids[ids < count]
    rows = np.full(count + 1, -1, dtype=np.int64)
    rows[ids] = np.arange(ids.size)
    rows[count] = ids.size
    return ids, rows


class Forward3D:
    """``ERTForward3D`` (``ert3d.py:114-218``): mixed/Dirichlet boundary, P1/P2, SR or point sources."""

    def __init__(self, nodes, cells, positions, quads, boundary_mode="mixed", singularity_removal=True,
                 element_order=1, geometric_factor_mode="auto"):
        self.mesh = m = meshlib.mesh3d(nodes, cells)
        self.positions, self.quads = surveylib.validate(positions, quads)
        self.E, self.D = self.positions.shape[0], self.quads.shape[0]
        self.mode, self.sr, self.order, self.gf_mode = boundary_mode, singularity_removal, element_order, geometric_factor_mode
        self.slots = surveylib.measurement_slots(self.quads, self.E)
        self.curr_ids, self.curr_map = electrode_rows(self.slots[:, :2], self.E)
        self.recv_ids, self.recv_map = electrode_rows(self.slots[:, 2:], self.E)
        self.capacity = max(self.curr_ids.size, self.recv_ids.size)
        if element_order == 2:
            self.dof_nodes, self.conn, self.bconn = meshlib.midpoint_topology(
                m["nodes"], m["cells"], m["boundary_faces"], meshlib.TET_EDGES, ((0, 1), (1, 2), (2, 0)))
            self.volume_t = fem.tet_p2_stiffness(m["nodes"], m["cells"])
        else:
            self.dof_nodes, self.conn, self.bconn = m["nodes"], m["cells"], m["boundary_faces"]
            self.volume_t = fem.tet_p1(m["nodes"], m["cells"])[2]
        self.N, self.C = self.dof_nodes.shape[0], m["cells"].shape[0]
        cell_ids, bary = meshlib.locate3d(m, self.positions)
        if element_order == 2:
            bary = fem.tet_p2_shape(bary)
        self.electrodes = np.zeros((self.E, self.N))
        np.add.at(self.electrodes, (np.repeat(np.arange(self.E), self.conn.shape[1]), self.conn[cell_ids].reshape(-1)),
                  bary.reshape(-1))
        self.cell_t = self.volume_t.copy()
        if boundary_mode == "mixed":
            self._add_robin()
        self._routing()
        self.unit = self._assemble_matrix(np.ones(self.C), robin=True, fixed=True)
        self.source_rhs = self._source_rhs()
        self.receiver_rhs = self._receiver_rhs()
        self.shared = bool((not self.sr or not m["flat"]) and np.array_equal(self.curr_ids, self.recv_ids))
        self._cache = {}

    def _add_robin(self):
        """First-order far-field Robin faces added into cell templates (``ert3d.py:224-256``)."""

        m = self.mesh
        far = np.flatnonzero(~m["surface_face_mask"])
        coefficients = fem.robin_coefficient_3d(m["boundary_face_centers"][far], m["boundary_face_normals"][far],
                                                self.positions.mean(axis=0))
        for f, alpha in zip(far, coefficients):
            if alpha == 0.0:
                continue
            c = int(m["boundary_face_cells"][f])
            lookup = {int(node): i for i, node in enumerate(self.conn[c])}
            local = np.array([lookup[int(node)] for node in self.bconn[f]])
            area = m["boundary_face_areas"][f]
            face = fem.p2_face_mass(area) if self.order == 2 else area * fem.P1_FACE_MASS
            self.cell_t[c][np.ix_(local, local)] += alpha * face

    def _dirichlet_nodes(self):
        return np.unique(self.bconn[~self.mesh["surface_face_mask"]].reshape(-1))  # ert3d.py:285-288

    def _routing(self):
        """Local-entry -> CSR slot routing with Dirichlet masking (``ert3d.py:290-351``)."""

        cells, k = self.conn, self.conn.shape[1]
        rows = np.repeat(cells, k, axis=1).reshape(-1)
        cols = np.tile(cells, (1, k)).reshape(-1)
        keys = rows * self.N + cols
        if self.mode == "dirichlet":
            fixed = self._dirichlet_nodes()
            mask = np.zeros(self.N, bool)
            mask[fixed] = True
            volume_pair = ~mask[rows] & ~mask[cols]
            ukeys = np.unique(keys[volume_pair | (rows == cols)])
            slot = np.searchsorted(ukeys, keys)
            self.groups = np.where(volume_pair, slot, -1)
            self.fixed_slots = np.searchsorted(ukeys, fixed * self.N + fixed)
            pm = volume_pair.reshape(self.C, -1)
            self.flat_t = np.where(pm, self.cell_t.reshape(self.C, -1), 0.0)
            self.flat_vt = np.where(pm, self.volume_t.reshape(self.C, -1), 0.0)
            self.entries = np.where(volume_pair, slot, 0).reshape(self.C, -1)
        else:
            ukeys, inv = np.unique(keys, return_inverse=True)
            self.groups = inv.reshape(-1)
            self.fixed_slots = np.empty(0, np.int64)
            self.flat_t = self.cell_t.reshape(self.C, -1)
            self.flat_vt = self.volume_t.reshape(self.C, -1)
            self.entries = self.groups.reshape(self.C, -1)
        urows = ukeys // self.N
        self.indices = ukeys % self.N
        self.indptr = np.concatenate(([0], np.cumsum(np.bincount(urows, minlength=self.N))))
        self.nnz = self.indices.size

    def values(self, sigma, robin=True, fixed=False):
        """``sum_c sigma_c T_c`` (+ unit fixed diagonals) (``ert3d.py:436-461``)."""

        t = self.flat_t if robin else self.flat_vt
        ok = self.groups >= 0
        contrib = (np.asarray(sigma, float)[:, None] * t).reshape(-1)
        out = np.bincount(self.groups[ok], weights=contrib[ok], minlength=self.nnz)
        if fixed and self.fixed_slots.size:
            out[self.fixed_slots] = 1.0
        return out

    def _assemble_matrix(self, sigma, robin=True, fixed=True):
        return sp.csr_matrix((self.values(sigma, robin, fixed), self.indices, self.indptr), shape=(self.N, self.N))

    def primary(self, source):
        """``(1/r + 1/r')/(4 pi)`` with the electrode-node fallback radius (``ert3d.py:367-385``)."""

        nodes = self.dof_nodes
        mirror = source.copy()
        mirror[2] = 2.0 * self.mesh["surface_level"] - source[2]
        dd = np.linalg.norm(nodes - source, axis=1)
        md = np.linalg.norm(nodes - mirror, axis=1)
        fallback = max(float(np.min(dd[dd > 1e-12])) * RADIUS_FRACTION, 1e-12)
        dd = np.where(dd > 1e-12, dd, fallback)
        md = np.where(md > 1e-12, md, fallback)
        return (1.0 / dd + 1.0 / md) / (4.0 * np.pi)

    def _source_rhs(self):
        """SR ``A(1) u_p`` or point-source rows, Dirichlet-zeroed, padded (``ert3d.py:387-405``)."""

        if self.sr and self.mesh["flat"]:
            prim = np.stack([self.primary(self.positions[i]) for i in self.curr_ids])
            rhs = (self.unit @ prim.T).T
        else:
            rhs = self.electrodes[self.curr_ids].copy()
        if self.mode == "dirichlet":
            rhs[:, self._dirichlet_nodes()] = 0.0
        out = np.zeros((self.capacity, self.N))
        out[: rhs.shape[0]] = rhs
        return out

    def _receiver_rhs(self):
        rhs = self.electrodes[self.recv_ids].copy()  # ert3d.py:407-413
        if self.mode == "dirichlet":
            rhs[:, self._dirichlet_nodes()] = 0.0
        out = np.zeros((self.capacity, self.N))
        out[: rhs.shape[0]] = rhs
        return out

    def fields(self, sigma):
        """Current-electrode fields and the factorization (``ert3d.py:508-518``)."""

        sigma = np.asarray(sigma, float)
        key = sigma.tobytes()
        if key not in self._cache:
            lu = factorize(self._assemble_matrix(sigma))
            current = lu.solve(self.source_rhs.T).T[: self.curr_ids.size]
            self._cache = {key: [current, None, lu]}
        return self._cache[key]

    def receivers(self, sigma):
        entry = self.fields(sigma)
        if entry[1] is None:
            entry[1] = entry[0] if self.shared else entry[2].solve(self.receiver_rhs.T).T[: self.recv_ids.size]
        return entry[1]

    def measurement(self, current):
        """``P[a,m] - P[b,m] - P[a,n] + P[b,n]`` on mapped rows (``ert3d.py:559-567``)."""

        p = surveylib.with_remote_slot(current @ self.receiver_rhs[: self.recv_ids.size].T)
        a, b = self.curr_map[self.slots[:, 0]], self.curr_map[self.slots[:, 1]]
        m, n = self.recv_map[self.slots[:, 2]], self.recv_map[self.slots[:, 3]]
        return p[..., a, m] - p[..., b, m] - p[..., a, n] + p[..., b, n]

    def resistance(self, sigma):
        return self.measurement(self.fields(sigma)[0])

    def geometric_factors(self):
        """Analytic, or ``1/R`` of the unit-conductivity solve (``ert3d.py:569-597``)."""

        numerical = self.gf_mode == "numerical" or (self.gf_mode == "auto" and not self.mesh["flat"])
        if not numerical:
            return surveylib.geometric_factors(self.positions, self.quads)
        unit = factorize(self.unit).solve(self.source_rhs.T).T[: self.curr_ids.size]
        return 1.0 / self.measurement(unit)

    def apparent_resistivity(self, sigma, currents=1.0):
        return self.geometric_factors() * self.resistance(sigma) / currents

    def aggregate(self, per_cell, ids=None, count=None):
        if ids is None:
            return per_cell
        ids = np.asarray(ids, np.int64)
        count = int(ids.max()) + 1 if count is None else int(count)
        out = np.zeros(per_cell.shape[:-1] + (count,))
        for p in range(count):
            out[..., p] = per_cell[..., ids == p].sum(axis=-1)
        return out

    def jacobian(self, sigma, robin=False, ids=None, count=None, chunk=64, rows=None):
        """``-(u_M-u_N)^T T_c (u_A-u_B)`` on current+receiver rows (``ert3d.py:867-927``).

        ``rows`` (optional) selects measurement rows; the result then has ``len(rows)`` rows."""

        current, receiver = self.fields(sigma)[0], self.receivers(sigma)
        phi = np.concatenate((current, receiver, np.zeros((1, self.N))))
        nc = current.shape[0]
        slots = self.slots if rows is None else self.slots[np.asarray(rows)]
        a, b = self.curr_map[slots[:, 0]], self.curr_map[slots[:, 1]]
        a, b = np.where(a == nc, phi.shape[0] - 1, a), np.where(b == nc, phi.shape[0] - 1, b)
        m, n = self.recv_map[slots[:, 2]] + nc, self.recv_map[slots[:, 3]] + nc
        k = self.conn.shape[1]
        T = (self.flat_t if robin else self.flat_vt).reshape(self.C, k, k)
        rows = []
        for s in range(0, slots.shape[0], chunk):
            sl = slice(s, s + chunk)
            cur = (phi[a[sl]] - phi[b[sl]])[:, self.conn]
            rec = (phi[m[sl]] - phi[n[sl]])[:, self.conn]
            rows.append(-np.einsum("dci,cij,dcj->dc", rec, T, cur))
        return self.aggregate(np.concatenate(rows), ids, count)

    def pair_cotangent(self, cotangent):
        """Transpose of the quadrupole map onto ``(n_curr, n_recv)`` (``ert3d.py:688-706``)."""

        nc, nr = self.curr_ids.size, self.recv_ids.size
        width = nr + 1
        a, b = self.curr_map[self.slots[:, 0]], self.curr_map[self.slots[:, 1]]
        m, n = self.recv_map[self.slots[:, 2]], self.recv_map[self.slots[:, 3]]
        g = np.asarray(cotangent, float)
        pairs = np.zeros((nc + 1) * width)
        np.add.at(pairs, a * width + m, g)
        np.add.at(pairs, a * width + n, -g)
        np.add.at(pairs, b * width + m, -g)
        np.add.at(pairs, b * width + n, g)
        return pairs.reshape(nc + 1, width)[:nc, :nr]

    def cell_gradient(self, phi, lam, robin=True):
        """``-sum_s lam_s^T T_c phi_s`` (``ert3d.py:713-734``)."""

        k = self.conn.shape[1]
        T = (self.flat_t if robin else self.flat_vt).reshape(self.C, k, k)
        return -np.einsum("sci,cij,scj->c", lam[:, self.conn], T, phi[:, self.conn])

    def vjp(self, sigma, cotangent):
        current, _, lu = self.fields(sigma)
        rhs = self.pair_cotangent(cotangent) @ self.receiver_rhs[: self.recv_ids.size]
        lam = lu.solve(rhs.T).T  # ert3d.py:744-772
        return self.cell_gradient(current, lam, robin=True)

    def jvp(self, sigma, direction):
        current, _, lu = self.fields(sigma)
        tangent = self._assemble_matrix(direction, robin=True, fixed=False)
        return self.measurement(lu.solve(-(tangent @ current.T)).T)  # ert3d.py:774-793

    def normal_vjp(self, sigma, cotangent, ids=None, count=None):
        current, receiver = self.fields(sigma)[0], self.receivers(sigma)
        grad = self.cell_gradient(current, self.pair_cotangent(cotangent) @ receiver, robin=False)
        return self.aggregate(grad, ids, count)  # ert3d.py:795-841

    def normal_jvp(self, sigma, direction):
        current, receiver = self.fields(sigma)[0], self.receivers(sigma)
        tangent = self._assemble_matrix(direction, robin=False, fixed=False)
        gram = surveylib.with_remote_slot(-(tangent @ current.T).T @ receiver.T)
        a, b = self.curr_map[self.slots[:, 0]], self.curr_map[self.slots[:, 1]]
        m, n = self.recv_map[self.slots[:, 2]], self.recv_map[self.slots[:, 3]]
        return gram[a, m] - gram[b, m] - gram[a, n] + gram[b, n]  # ert3d.py:843-865

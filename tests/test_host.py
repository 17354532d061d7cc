"""Host-side logic of the package against the oracle, and the C ABI's exports (CPU only).

Covers the preprocessing that feeds the kernels (mesh metadata, point location, element
templates, Robin terms, sparsity pattern and slot maps, pair scatters, quadrature, DOF
numbering), the validation contract of ``forward/inputs.py`` and ``survey/core.py``, the field
cache semantics of ``forward/field_cache.py``, and that ``libadt_hip.so`` loads and exports every
function ``include/adt_hip.h`` declares (no device calls).
"""

import re

import numpy as np
import pytest
import torch

from oracle import ert2p5d as o2, ert3d as o3, fem as ofem, mesh as omesh, quadrature as oquad
from oracle import survey as osurvey
from tests.conftest import ROOT, adt, relmax

R = -1


def test_c_abi_library_exports_every_declared_function():
    header = (ROOT / "include" / "adt_hip.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(adt_\w+)\s*\(", header, flags=re.M))
    assert declared, "no declarations parsed"
    lib = adt.native.library()
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing
    assert set(adt.native.exported_symbols()) <= declared
    assert lib.adt_abi_version() == adt.native.ABI_VERSION
    assert lib.adt_last_error() is not None


def test_operators_require_a_gpu_like_the_reference():
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    case = adt.cases.quad_strip_case()
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells, surface_node_ids=case.surface_node_ids)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    with pytest.raises(RuntimeError, match="GPU"):
        adt.ERTForward2p5D.from_mesh_survey(mesh, survey)


@pytest.mark.parametrize("quad, message", [([R, R, 1, 2], "needs an electrode"), ([0, 1, R, R], "needs an electrode"),
                                           ([0, -2, 1, 2], "other than -1"), ([0, 0, 1, 2], "distinct"),
                                           ([0, 1, 2, 4], "outside the survey")])
def test_survey_validation_messages(quad, message):
    line = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [3.0, 0.0]])
    with pytest.raises(ValueError, match=message):
        adt.Survey.from_arrays(line, [quad])


def test_survey_geometric_factors_and_slots_match_the_oracle():
    positions = np.column_stack((np.arange(12.0), np.zeros(12)))
    quads = np.vstack((adt.cases.wenner_alpha(12), [[0, R, 3, 4], [2, 3, 5, R], [1, R, 6, R]]))
    s = adt.Survey.from_arrays(positions, quads)
    assert relmax(s.geometric_factors().numpy(), osurvey.geometric_factors(positions, quads)) < 1e-14
    assert np.array_equal(s.measurement_slots().numpy(), osurvey.measurement_slots(quads, 12))


def test_input_validation_contract():
    inputs = importlib_inputs()
    with pytest.raises(ValueError, match="conductivity must be scalar"):
        inputs.conductivity_model(torch.ones(3), 4)
    with pytest.raises(ValueError, match="positive finite"):
        inputs.conductivity_model(torch.tensor([1.0, -1.0]), 2)
    with pytest.raises(ValueError, match="n_steps"):
        inputs.conductivity_series(torch.ones(3), 3)
    with pytest.raises(ValueError, match="parameter_count requires"):
        inputs.parameter_targets(None, 3, 4)
    with pytest.raises(ValueError, match="smaller than"):
        inputs.parameter_targets(np.array([0, 1, 2, -1]), 2, 4)
    with pytest.raises(ValueError, match="must be one of"):
        inputs.option("x", ("analytic", "numerical"), "topographic_geometric_factor_mode")
    with pytest.raises(FloatingPointError):
        inputs.check_finite(torch.tensor([1.0, float("nan")]), "fields")
    ids, count = inputs.parameter_targets(np.array([0, -1, 1, 1]), None, 4)
    assert count == 2 and ids.tolist() == [0, -1, 1, 1]


def importlib_inputs():
    import importlib

    return importlib.import_module("ad-tlert_amd.inputs")


def test_field_cache_semantics():
    fc = importlib_cache()
    cache = fc.FieldCache(2)
    solved = []

    def solve(batch):
        solved.append(batch.shape[0])
        return [float(m.sum()) for m in batch]

    models = torch.stack([torch.full((3,), float(i)) for i in range(3)])
    values, keys = cache.solve_series(models, 2, solve)
    assert values == [0.0, 3.0, 6.0] and solved == [2, 1] and cache.info()["misses"] == 3
    values, _ = cache.solve_series(models[1:], 2, solve)
    assert cache.info()["hits"] == 2 and cache.info()["entries"] == 2
    assert fc.step_batches(list(range(5)), 2) == [[0, 1], [2, 3], [4]]
    assert fc.newest_batches_first(4, 2) == [[2, 3], [0, 1]]
    assert fc.FieldCache(0).key(models[0]) is None
    with pytest.raises(ValueError, match="non-negative"):
        fc.FieldCache(-1)


def test_field_cache_key_digests(monkeypatch):
    """Keys are a 16-byte digest of the model's bytes plus its dtype: equal models share a key,
    a one-ulp change or another dtype does not; the blake2b fallback (no xxhash) keeps that."""

    fc = importlib_cache()
    model = torch.linspace(0.5, 2.0, 1001, dtype=torch.float64)
    nudged = model.clone()
    nudged[500] = np.nextafter(nudged[500].item(), 3.0)
    for backend in ("default", "blake2b"):
        if backend == "blake2b":
            monkeypatch.setattr(fc, "xxhash", None)
        cache = fc.FieldCache(4)
        key = cache.key(model)
        assert isinstance(key, bytes) and len(key) == 16 + len(b"<f8")
        assert cache.key(model.clone()) == key
        assert cache.key(nudged) != key
        assert cache.key(model.float()) != key
    import hashlib

    assert fc.digest16(b"abc") == hashlib.blake2b(b"abc", digest_size=16).digest()


def importlib_cache():
    import importlib

    return importlib.import_module("ad-tlert_amd.field_cache")


def test_step_batches_split_like_numpy_array_split():
    fc = importlib_cache()
    assert fc.step_batches(list(range(7)), 3) == [a.tolist() for a in np.array_split(np.arange(7), 3)]


def test_mesh2d_metadata_matches_the_oracle():
    for nodes, cells, surface in ((*adt.cases.grid_triangles(9, 5), None),
                                  (adt.cases.quad_strip_case().nodes, adt.cases.quad_strip_case().cells,
                                   np.arange(31))):
        m = adt.Mesh.from_arrays(nodes, cells, surface_node_ids=surface)
        o = omesh.mesh2d(nodes, cells, surface)
        assert np.array_equal(m.boundary_edges.numpy(), o["boundary_edges"])
        assert np.array_equal(m.boundary_edge_cells.numpy(), o["boundary_edge_cells"])
        assert relmax(m.boundary_edge_normals.numpy(), o["boundary_edge_normals"]) < 1e-14
        assert np.array_equal(m.surface_edge_mask.numpy(), o["surface_edge_mask"])
        assert np.array_equal(m.surface_node_ids.numpy(), o["surface_node_ids"])
        assert m.is_flat_surface == o["flat"]
        assert float(m.surface_reference_level) == pytest.approx(o["surface_level"])


def test_point_location_matches_the_oracle():
    nodes, cells = adt.cases.grid_triangles(9, 5)
    m = adt.Mesh.from_arrays(nodes, cells)
    pts = np.array([[0.3, -0.2], [2.5, -1.7], [5.9, -3.9], [3.0, -2.0], [4.0, 0.0]])
    ids, w = m.locate_points(pts)
    oids, ow = omesh.locate2d(omesh.mesh2d(nodes, cells), pts)
    assert np.array_equal(ids.numpy(), oids) and relmax(w.numpy(), ow) < 1e-12
    nodes3, tets = adt.cases.box_mesh()
    m3 = adt.Mesh3D.from_arrays(nodes3, tets)
    pts3 = np.array([[6.0, 6.0, 0.0], [2.5, 3.3, -1.2], [11.9, 0.1, -7.9]])
    ids3, w3 = m3.locate_points(pts3)
    oids3, ow3 = omesh.locate3d(omesh.mesh3d(nodes3, tets), pts3)
    assert np.array_equal(ids3.numpy(), oids3) and relmax(w3.numpy(), ow3) < 1e-12


def test_mesh3d_metadata_matches_the_oracle():
    nodes, tets = adt.cases.box_mesh(relief=0.8)
    m = adt.Mesh3D.from_arrays(nodes, tets)
    o = omesh.mesh3d(nodes, tets)
    assert np.array_equal(m.boundary_faces.numpy(), o["boundary_faces"])
    assert np.array_equal(m.boundary_face_cells.numpy(), o["boundary_face_cells"])
    assert relmax(m.boundary_face_normals.numpy(), o["boundary_face_normals"]) < 1e-13
    assert np.array_equal(m.surface_face_mask.numpy(), o["surface_face_mask"])
    assert m.is_flat_surface == o["flat"] is False


def test_element_templates_and_robin_terms_match_the_oracle():
    fem = importlib_fem()
    nodes, cells = adt.cases.grid_triangles(7, 4)
    k, mm = fem.p1_triangle(nodes.astype(float), cells)
    ok, om = ofem.p1_triangle(nodes.astype(float), cells)
    assert relmax(k, ok) < 1e-14 and relmax(mm, om) < 1e-14
    q = adt.cases.quad_strip_case()
    assert relmax(fem.q1_quad(q.nodes, q.cells)[0], ofem.q1_quad(q.nodes, q.cells)[0]) < 1e-14
    m = omesh.mesh2d(nodes, cells)
    for kw in (0.01, 0.3, 2.0):
        a = fem.robin_2p5d(m["boundary_edge_centers"], m["boundary_edge_normals"], m["surface_edge_mask"],
                           m["surface_level"], np.array([3.0, 0.0]), kw)
        assert relmax(a, ofem.robin_2p5d(m, np.array([3.0, 0.0]), kw)) < 1e-14
    nodes3, tets = adt.cases.box_mesh()
    assert relmax(fem.tet_p1(nodes3, tets), ofem.tet_p1(nodes3, tets)[2]) < 1e-14
    assert relmax(fem.tet_p2(nodes3, tets), ofem.tet_p2_stiffness(nodes3, tets)) < 1e-13


def importlib_fem():
    import importlib

    return importlib.import_module("ad-tlert_amd.fem")


@pytest.mark.parametrize("order", [1, 2])
def test_3d_robin_faces_match_the_oracle_loop(order):
    case = adt.cases.box_case()
    ref = o3.Forward3D(case.nodes, case.cells, case.electrodes, case.measurements, element_order=order)
    fem = importlib_fem()
    m = ref.mesh
    cell_t = ref.volume_t.copy()
    fem.add_robin_faces(cell_t, ref.conn, ref.bconn, m["boundary_face_cells"], m["boundary_face_centers"],
                        m["boundary_face_areas"], m["boundary_face_normals"], m["surface_face_mask"],
                        case.electrodes.mean(axis=0), order)
    assert relmax(cell_t, ref.cell_t) < 1e-14


def test_sparse_pattern_slot_maps_and_pair_scatter_match_the_oracle():
    e2 = importlib_e2()
    c = adt.cases.quad_strip_case()
    ref = o2.Forward2p5D(c.nodes, c.cells, c.electrodes, c.measurements, c.surface_node_ids)
    pat = e2.sparse_pattern(ref.mesh["cells"], ref.bdofs, ref.N)
    assert np.array_equal(pat.indptr, ref.indptr) and np.array_equal(pat.indices, ref.indices)
    assert np.array_equal(pat.volume_inverse, ref.vinv) and np.array_equal(pat.boundary_inverse, ref.binv)
    inverse = np.concatenate((pat.volume_inverse, pat.boundary_inverse))
    ptr, entries = e2.slot_map(inverse, pat.nnz)
    assert np.array_equal(np.sort(entries), np.arange(inverse.size))
    for slot in (0, 7, pat.nnz - 1):
        listed = entries[ptr[slot]:ptr[slot + 1]]
        assert np.all(inverse[listed] == slot) and np.all(np.diff(listed) > 0)
    rng = np.random.default_rng(0)
    gn, gr = rng.standard_normal(ref.D), rng.standard_normal(ref.D)
    assert relmax(e2.pair_scatter(ref.slots, ref.E, gn[None], gr[None])[0], ref.pair_cotangent(gn, gr)) < 1e-15
    e3 = importlib_e3()
    box = adt.cases.box_case()
    r3 = o3.Forward3D(box.nodes, box.cells, box.electrodes, box.measurements)
    a, b = r3.curr_map[r3.slots[:, 0]], r3.curr_map[r3.slots[:, 1]]
    m_, n_ = r3.recv_map[r3.slots[:, 2]], r3.recv_map[r3.slots[:, 3]]
    g = rng.standard_normal(r3.D)
    got = e3.pair_scatter_3d(a, b, m_, n_, r3.curr_ids.size, r3.recv_ids.size, g[None])[0]
    assert relmax(got, r3.pair_cotangent(g)) < 1e-15


def importlib_e2():
    import importlib

    return importlib.import_module("ad-tlert_amd.ert2p5d")


def importlib_e3():
    import importlib

    return importlib.import_module("ad-tlert_amd.ert3d")


def test_quadrature_matches_the_oracle():
    for bounds in ((0.5, 200.0), (1.0, 96.0), (2.5, 190.0)):
        k, q = adt.build_inverse_cosine_weights(*bounds)
        ok, oq = oquad.inverse_cosine_weights(*bounds)
        assert relmax(k.numpy(), ok) < 1e-15 and relmax(q.numpy(), oq) < 1e-15


def test_dof_numbering_is_a_permutation_with_compact_tiles():
    import importlib

    ordering = importlib.import_module("ad-tlert_amd.ordering")
    case = adt.cases.strip_dd_4k()
    perm = ordering.hilbert_order(case.nodes)
    assert np.array_equal(np.sort(perm), np.arange(case.nodes.shape[0]))
    inv = ordering.inverse(perm)
    assert np.array_equal(inv[perm], np.arange(perm.size))
    # 64-node tiles of the Hilbert numbering are compact: small bounding boxes
    pts = case.nodes[perm]
    spans = [np.ptp(pts[i:i + 64], axis=0).max() for i in range(0, pts.shape[0] - 64, 64)]
    natural = [np.ptp(case.nodes[i:i + 64], axis=0).max() for i in range(0, pts.shape[0] - 64, 64)]
    assert np.median(spans) < 0.5 * np.median(natural)


def greedy_aggregates(indptr, indices, values, theta):
    """Restatement of the multigrid setup's three-pass greedy aggregation over strong
    connections |a_ij| >= theta sqrt(a_ii a_jj) (rows in index order), for pinning the library."""

    n = indptr.size - 1
    diag = np.array([values[indptr[i] + np.flatnonzero(indices[indptr[i]:indptr[i + 1]] == i)[0]] for i in range(n)])
    strong = [[] for _ in range(n)]
    weight = [[] for _ in range(n)]
    for i in range(n):
        for k in range(indptr[i], indptr[i + 1]):
            j = indices[k]
            s = abs(values[k]) / np.sqrt(abs(diag[i] * diag[j]))
            if j != i and values[k] != 0 and s >= theta:
                strong[i].append(j)
                weight[i].append(s)
    agg = -np.ones(n, int)
    count = 0
    for i in range(n):
        if agg[i] < 0 and strong[i] and all(agg[j] < 0 for j in strong[i]):
            agg[i] = count
            agg[strong[i]] = count
            count += 1
    first = agg.copy()
    for i in range(n):
        if agg[i] < 0 and strong[i]:
            cand = [(w, first[j]) for j, w in zip(strong[i], weight[i]) if first[j] >= 0]
            if cand:
                agg[i] = max(cand, key=lambda t: t[0])[1]
    for i in range(n):
        if agg[i] < 0 and strong[i]:
            agg[i] = count
            for j in strong[i]:
                if agg[j] < 0:
                    agg[j] = count
            count += 1
    return agg, count


def test_multigrid_hierarchy_matches_a_restated_aggregation_and_galerkin_pattern():
    import scipy.sparse as sp

    rng = np.random.default_rng(7)
    side = 23
    n = side * side
    A = sp.diags([4.0], [0], shape=(n, n)).tolil()
    for i in range(side):
        for j in range(side):
            p = i * side + j
            for di, dj in ((0, 1), (1, 0)):
                if i + di < side and j + dj < side:
                    q = (i + di) * side + j + dj
                    w = -rng.uniform(0.2, 1.0) * (10.0 if (di == 1 and i < 5) else 1.0)  # anisotropic band
                    A[p, q] = A[q, p] = w
    A = A.tocsr()
    A.setdiag(-np.asarray(A.sum(axis=1)).ravel() + A.diagonal() + 0.05)
    A[0, 1] = A[1, 0] = 0.0  # explicit zeros stay weak
    A = A.tocsr()
    A.sort_indices()
    rows, nnz = adt.native.amg_hierarchy_info(A.indptr, A.indices, A.data, theta=0.08, coarse_max=20)
    agg, count = greedy_aggregates(A.indptr, A.indices, A.data, 0.08)
    assert rows[0] == n and rows[1] == count
    assert rows[-1] <= 20 and all(a > b for a, b in zip(rows, rows[1:]))
    # Galerkin pattern of the first coarse level: P row i has the aggregates of i and of its
    # strong neighbours (the prolongator is smoothed with the filtered matrix)
    coo = A.tocoo()
    d = A.diagonal()
    strong = (coo.row == coo.col) | (np.abs(coo.data) >= 0.08 * np.sqrt(d[coo.row] * d[coo.col]))
    keep = (agg[coo.col] >= 0) & strong & (agg[coo.row] >= 0)
    P = sp.csr_matrix((np.ones(keep.sum()), (coo.row[keep], agg[coo.col[keep]])), shape=(n, count))
    Ac = (P.T @ abs(A) @ P).tocsr()
    Ac.data[:] = 1.0
    assert nnz[1] == Ac.nnz


def test_multigrid_hierarchy_leaves_unconnected_rows_out_and_rejects_bad_diagonals():
    import scipy.sparse as sp

    side = 16
    n = side * side
    lap = sp.diags([4.0, -1.0, -1.0, -1.0, -1.0], [0, 1, -1, side, -side], shape=(n, n)).tolil()
    for r in range(0, n, 7):  # Dirichlet-like rows: unit diagonal, no couplings
        lap[r, :] = 0.0
        lap[:, r] = 0.0
        lap[r, r] = 1.0
    lap = lap.tocsr()
    lap.eliminate_zeros()
    rows, _ = adt.native.amg_hierarchy_info(lap.indptr, lap.indices, lap.data, coarse_max=8)
    agg, count = greedy_aggregates(lap.indptr, lap.indices, lap.data, 0.08)
    assert rows[1] == count and (agg[::7] == -1).all()
    bad = lap.copy()
    bad.data[bad.indptr[3]:bad.indptr[4]][bad.indices[bad.indptr[3]:bad.indptr[4]] == 3] = -1.0
    with pytest.raises(ValueError, match="SPD"):
        adt.native.amg_hierarchy_info(bad.indptr, bad.indices, bad.data)


def test_mirror_slots_of_symmetric_patterns_and_rejection_of_others():
    import importlib

    import scipy.sparse as sp

    learned = importlib.import_module("ad-tlert_amd.learned")
    A = sp.random(40, 40, density=0.1, random_state=3, format="csr")
    A = (A + A.T + sp.identity(40)).tocsr()
    A.sort_indices()
    mirror = learned.mirror_slots(A.indptr, A.indices)
    rows = np.repeat(np.arange(40), np.diff(A.indptr))
    assert np.array_equal(rows[mirror], A.indices) and np.array_equal(A.indices[mirror], rows)
    assert np.array_equal(mirror[mirror], np.arange(A.nnz))
    B = sp.csr_matrix(np.triu(np.ones((5, 5))))
    with pytest.raises(ValueError, match="symmetric"):
        learned.mirror_slots(B.indptr, B.indices)


def test_ctypes_signatures_match_the_header_argument_counts():
    """Every function the binding declares takes as many arguments as ``include/adt_hip.h``
    declares for it (a mismatch would shift every later argument)."""

    import re

    header = (ROOT / "include" / "adt_hip.h").read_text()
    native = adt.native
    checked = 0
    for name, (_, argtypes) in native._SIGNATURES.items():
        found = re.search(r"^\s*(?:int|int64_t|const char\*)\s+" + name + r"\s*\(([^;]*?)\)\s*;", header,
                          flags=re.S | re.M)
        assert found, f"{name} is not declared in the header"
        params = [p for p in found.group(1).split(",") if p.strip() and p.strip() != "void"]
        assert len(params) == len(argtypes), f"{name}: header {len(params)} arguments, binding {len(argtypes)}"
        checked += 1
    assert checked >= 25


def test_vertex_prolongator_of_ten_node_tets_interpolates_linear_functions():
    """The finest prolongator of P2 operators (``ERTForward3D._vertex_prolongator``) in a
    shuffled device numbering: rows sum to one, vertex rows are unit rows, and interpolating the
    vertex values of any linear function reproduces it at every ten-node DOF (edge midpoints are
    the averages of their ends); columns are the vertices in device-row order."""

    import types

    case = adt.cases.box_case()
    mesh = adt.Mesh3D.from_arrays(case.nodes, case.cells)
    points, conn, _ = (np.asarray(a) for a in mesh.build_quadratic_topology())
    n = points.shape[0]
    inv = np.random.default_rng(2).permutation(n)  # mesh DOF -> device row
    stub = types.SimpleNamespace(mesh=mesh, _cell_connectivity=conn, N=n)
    indptr, indices, values, nc = adt.ERTForward3D._vertex_prolongator(stub, inv)
    nv = mesh.node_count
    assert nc == nv and indptr.size == n + 1 and indptr[-1] == indices.size == values.size
    import scipy.sparse as sp

    P = sp.csr_matrix((values, indices, indptr), shape=(n, nv)).toarray()
    np.testing.assert_allclose(P.sum(axis=1), 1.0)
    order = np.argsort(inv[:nv], kind="stable")  # column j = the j-th vertex in device-row order
    for f in (points[:, 0], points[:, 1], points[:, 2], 0.3 + points @ np.array([0.7, -0.4, 1.1])):
        coarse = f[order]
        fine = np.empty(n)
        fine[inv] = f
        np.testing.assert_allclose(P @ coarse, fine, atol=1e-12)
    assert np.all((P[inv[:nv]] == 1.0).sum(axis=1) == 1)

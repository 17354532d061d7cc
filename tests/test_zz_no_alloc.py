"""``adt_gnn_backward_static`` (the damping network's backward pass with the library's static
device scratch, no allocation) against ``adt_gnn_backward`` (caller workspace), through the C
ABI: two calls in a row, the second on a smaller graph, bitwise equal gradients; sizes beyond
the static array are rejected before any launch."""

import importlib

import numpy as np
import pytest
import torch

from tests.conftest import adt
from tests.test_gpu_kernels import random_spd_batch

pytestmark = pytest.mark.gpu
native = adt.native
learned = importlib.import_module("ad-tlert_amd.learned")


def case(side: int, seed: int):
    rng = np.random.default_rng(seed)
    indptr, indices, values, _ = random_spd_batch(rng, side=side, S=1)
    agg = native.amg_aggregates(indptr, indices, values[0])
    device = torch.device("cuda", 0)
    graph = learned.GnnGraph.from_pattern(indptr, indices, agg, device)
    net = learned.DampingNetwork(seed=seed)
    flat = net.master.detach().clone()
    flat += torch.as_tensor(rng.standard_normal(flat.numel()) * 0.05, dtype=flat.dtype, device=flat.device)
    net = learned.DampingNetwork(flat.cpu().numpy())
    params = net.bf16_bits()
    lib = native.library()
    work = torch.empty(int(lib.adt_gnn_work_floats(graph.n, graph.nnz)), dtype=torch.float32, device=device)
    values_d = torch.as_tensor(values[0], device=device)
    net.omega(graph, values_d, 0.6, params=params, work=work)
    domega = torch.as_tensor(rng.standard_normal(graph.nnz), device=device)
    return graph, params, work, domega


def backward_both(graph, params, work, domega):
    lib = native.library()
    device = work.device
    scratch = torch.empty(int(lib.adt_gnn_scratch_floats(graph.n, graph.nnz)), dtype=torch.float32, device=device)
    expected = torch.zeros(learned.PARAM_COUNT, dtype=torch.float32, device=device)
    native.gnn_backward(graph, 0.6, params, work, domega, scratch, expected)
    actual = torch.full_like(expected, float("nan"))
    native.gnn_backward_static(graph, 0.6, params, work, domega, actual)
    torch.cuda.synchronize()
    return expected, actual


def test_static_scratch_backward_equals_the_workspace_backward_on_two_calls():
    lib = native.library()
    large, small = case(26, 3), case(13, 4)
    assert lib.adt_gnn_scratch_floats(large[0].n, large[0].nnz) > lib.adt_gnn_scratch_floats(small[0].n, small[0].nnz)
    for graph, params, work, domega in (large, small):
        expected, actual = backward_both(graph, params, work, domega)
        print(f"n={graph.n} nnz={graph.nnz}: max |grad| {float(expected.abs().max()):.3e}")
        assert torch.isfinite(expected).all() and float(expected.abs().max()) > 0
        assert torch.equal(actual, expected)


def test_static_scratch_rejects_sizes_beyond_its_array():
    graph, params, work, domega = case(8, 5)
    grad = torch.zeros(learned.PARAM_COUNT, dtype=torch.float32, device=work.device)
    n = nnz = 200_000  # needs ~38M scratch floats, above the 12M-float array: rejected before any launch
    assert native.library().adt_gnn_scratch_floats(n, nnz) > native.GNN_STATIC_SCRATCH_FLOATS
    lib = native.library()
    rc = lib.adt_gnn_backward_static(n, nnz, native.ptr(graph.indptr), native.ptr(graph.indices),
                                     native.ptr(graph.rows), native.ptr(graph.mirror), 0.6, native.ptr(params),
                                     native.ptr(work), native.ptr(domega), native.ptr(grad), native.stream())
    assert rc == native.ADT_ERR_INVALID
    assert b"static limit" in lib.adt_last_error()
    assert torch.all(grad == 0)

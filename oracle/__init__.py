"""FP64 CPU oracle for the AD-TLERT 2.5D/3D ERT forward-and-sensitivity path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` use it, and only as the checker (or the timed CPU baseline),
never as the thing measured or shipped.

It is a plain numpy/scipy restatement of the reference algorithm, written
from the reference's source (AD-TLERT 0.3.0), with every function citing the
reference ``file:line`` it follows.  It always computes in float64 and solves
with a sparse direct factorization (``scipy.sparse.linalg.splu``), the CPU
counterpart of the reference's cuDSS direct solve (``forward/cudss.py:24-122``).

Modules
-------
``survey``      ABMN slots, remote electrode, analytic geometric factors.
``quadrature``  inverse-cosine-transform wavenumbers/weights.
``mesh``        2D/3D boundary topology, surface metadata, point location.
``fem``         local element templates (P1/P2 tri, Q1 quad, P1/P2 tet, edges/faces).
``ert2p5d``     2.5D forward, Jacobian, vjp/jvp, normal_vjp/normal_jvp.
``ert3d``       3D forward, Jacobian, vjp/jvp, normal_vjp/normal_jvp.

Each function's pinning (closed form, invariant or known-answer test that a
plausible mistake would fail) lives in ``tests/test_oracle_*.py``.
"""

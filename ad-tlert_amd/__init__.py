"""MI355X-native ERT forward-and-sensitivity path (drop-in for AD-TLERT's forward operators).

The package directory is ``ad-tlert_amd``; import it with
``importlib.import_module("ad-tlert_amd")`` (its submodules use relative imports only).

Public names mirror ``adtlert.forward`` / ``adtlert.mesh`` / ``adtlert.survey``:
``ERTForward2p5D``, ``ERTForward3D``, ``ForwardResponse``, ``ForwardResponse3D``,
``apparent_resistivity_autograd`` and its Function aliases, ``Mesh``, ``Mesh3D``, ``Survey``,
``REMOTE_ELECTRODE``, ``with_remote_slot``, ``FLOAT_DTYPE``.  All device work runs in the
hand-written gfx950 HIP library ``_lib/libadt_hip.so`` (see ``include/adt_hip.h``).
"""

from .dtypes import FLOAT_DTYPE, INT_DTYPE, NP_FLOAT_DTYPE
from .survey import REMOTE_ELECTRODE, Survey, with_remote_slot
from .mesh import Mesh, Mesh3D
from .quadrature import build_inverse_cosine_weights, survey_wavenumber_bounds
from .ert2p5d import ERTForward2p5D, ForwardResponse
from .ert3d import ERTForward3D, ForwardResponse3D
from .autograd import (ERT2p5DApparentResistivityFunction, ERT3DApparentResistivityFunction,
                       ERTApparentResistivityFunction, apparent_resistivity_autograd)
from . import cases, native, solver

__all__ = [
    "FLOAT_DTYPE", "INT_DTYPE", "NP_FLOAT_DTYPE", "REMOTE_ELECTRODE", "Survey", "with_remote_slot", "Mesh", "Mesh3D",
    "build_inverse_cosine_weights", "survey_wavenumber_bounds", "ERTForward2p5D", "ForwardResponse", "ERTForward3D",
    "ForwardResponse3D", "ERT2p5DApparentResistivityFunction", "ERT3DApparentResistivityFunction",
    "ERTApparentResistivityFunction", "apparent_resistivity_autograd", "cases", "native", "solver",
]

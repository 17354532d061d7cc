"""Default floating-point precision of returned tensors.

Mirrors ``adtlert/utils/dtypes.py:10-23``: ``ADTLERT_ENABLE_FLOAT64=1`` (read at import)
selects float64 as torch's default dtype and as the dtype of returned results; otherwise
results are float32.  Device arithmetic is float64 regardless (solves, assembly, sensitivity
accumulation).
"""

from __future__ import annotations

import os

import numpy as np
import torch

if os.environ.get("ADTLERT_ENABLE_FLOAT64", "").strip().lower() not in {"", "0", "false", "no", "off"}:
    torch.set_default_dtype(torch.float64)

FLOAT_DTYPE = torch.float64 if torch.get_default_dtype() == torch.float64 else torch.float32
NP_FLOAT_DTYPE = np.float64 if FLOAT_DTYPE is torch.float64 else np.float32
INT_DTYPE = torch.int32

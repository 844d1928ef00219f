"""Precision form: FP32 storage against the FP64 oracle along a 100-step
trajectory of the 32x22x31 GYRE-like basin (tests/precision_study.py).

Tolerance: the design target for the tracers is a max relative error of
1e-5 after 100 steps; momentum and sea level are reported with the bound
they meet (1e-3), the barotropic time averaging being FP64 in both forms.
"""
import json

import numpy as np
import pytest

from precision_study import errors


@pytest.mark.gpu
def test_precision_form_error_growth(hip):
    res = errors((1, 10, 100))
    print(json.dumps(res))
    for kt, e in res.items():
        assert all(np.isfinite(v) for v in e.values()), f"step {kt}: {e}"
    assert res[100]["T"] <= 1e-5 and res[100]["S"] <= 1e-5, res[100]
    assert max(res[100]["u"], res[100]["v"], res[100]["ssh"]) <= 1e-3, res[100]

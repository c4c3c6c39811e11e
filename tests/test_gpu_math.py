"""The device build of the renderer's float elementary functions (dmath.h)
against the host build of the same source, bit for bit, and against the
oracle (the C library's double functions rounded to float), through the C
ABI (az_gpu_math_eval)."""
import numpy as np
import pytest

import oracle
from azimuth_amd import native as N
from test_math import SPECIAL, samples, same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(N.MATH_FUNCTIONS))
def test_device_routines_match_host_and_oracle(gpu, name):
    x, y = samples(name, 1 << 21, 7)
    dev = gpu.math_eval(name, x, y)
    host = N.math_eval_host(name, x, y)
    want = oracle.math_eval(name, x, y)
    d_h = int((~same(dev, host)).sum())
    d_o = int((~same(dev, want)).sum())
    print(f"{name}: device vs host {d_h}, device vs oracle {d_o} of {x.size}", flush=True)
    # Exact operations only (fma, division, sqrt, rint, frexp, ldexp), so the
    # device and host builds agree everywhere.
    assert d_h == 0
    assert d_o <= 2


def test_device_special_values(gpu):
    xs, ys = np.meshgrid(SPECIAL, SPECIAL, indexing="ij")
    xs, ys = xs.reshape(-1).copy(), ys.reshape(-1).copy()
    for name in N.MATH_FUNCTIONS:
        two = name in ("atan2", "pow")
        x = xs if two else SPECIAL
        if name in ("sin", "cos"):
            x = SPECIAL[np.abs(SPECIAL) < 1e20]
        y = ys if two else None
        dev = gpu.math_eval(name, x, y)
        want = oracle.math_eval(name, x, y)
        bad = np.flatnonzero(~same(dev, want))
        assert bad.size == 0, [(name, x[i], None if y is None else y[i], dev[i], want[i]) for i in bad[:8]]

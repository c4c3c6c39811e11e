"""The renderer's float elementary functions (the reference's Sin, Cos, Atan,
Power, Exp, ArcSine and ArcCosine, math.cpp:136-260), each defined as the
double-precision function of the float argument rounded to float.

- The oracle's (the C library's double functions) against numpy's float64
  ufuncs, an independent implementation, rounded the same way.
- The product's own double-precision routines, compiled for the host from
  the device source (az_math_eval), against the oracle bit for bit, on the
  argument ranges the renderer uses and on the special values of C99 Annex F.
The device build of the same routines is checked in test_gpu_math.py."""
import numpy as np
import pytest

import oracle
from azimuth_amd import native as N

SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 0.5, -0.5, 2.0, -2.0, 3.0, -3.0, 4.0, 1e-40,
                    -1e-40, 1e30, -1e30, 1.0000001, 0.99999994, np.pi, -np.pi, np.pi / 2], dtype=np.float32)


def samples(name, n, seed):
    """Arguments over the ranges the renderer passes (angles in [-pi, 2pi],
    cosines in [-1, 1], glossy exponents up to 2000, attenuation and filter
    exponents) plus wider ones."""
    r = np.random.default_rng(seed)
    f = lambda lo, hi, k: r.uniform(lo, hi, k).astype(np.float32)
    q = n // 4
    if name in ("sin", "cos"):
        return np.concatenate([f(0, 2 * np.pi, q), f(-np.pi, np.pi, q), f(0, np.pi / 2, q), f(-1e4, 1e4, n - 3 * q)]), None
    if name == "atan2":
        y, x = f(-1, 1, n), f(-1, 1, n)
        x[:q] *= 1e-3
        y[q:2 * q] *= 1e4
        return y, x
    if name == "pow":
        return f(0, 1, n), np.concatenate([f(0, 20, q), f(0, 2000, q), f(0, 100, q), f(-4, 4, n - 3 * q)])
    if name == "exp":
        return np.concatenate([f(-10, 0, q), f(-104, 89, q), f(-1, 1, n - 2 * q)]), None
    if name in ("asin", "acos"):
        return np.concatenate([f(-1, 1, 2 * q), f(0, 1, q) ** 0.5, (1 - f(0, 1e-4, n - 3 * q))]), None
    return f(-1.2, 1.2, n), None


NUMPY = {
    "sin": lambda x, y: np.sin(x.astype(np.float64)),
    "cos": lambda x, y: np.cos(x.astype(np.float64)),
    "atan2": lambda x, y: np.arctan2(x.astype(np.float64), y.astype(np.float64)),
    "pow": lambda x, y: np.power(x.astype(np.float64), y.astype(np.float64)),
    "exp": lambda x, y: np.exp(x.astype(np.float64)),
    "asin": lambda x, y: np.arcsin(x.astype(np.float64)),
    "acos": lambda x, y: np.arccos(x.astype(np.float64)),
    "pow4": lambda x, y: x.astype(np.float64) ** 4,
    "pow5": lambda x, y: x.astype(np.float64) ** 5,
}


def same(a, b):
    """Elementwise bit equality, NaNs equal."""
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


@pytest.mark.parametrize("name", list(oracle.MATH_FUNCTIONS))
def test_oracle_functions_match_numpy_float64(name):
    x, y = samples(name, 1 << 19, 3)
    with np.errstate(all="ignore"):
        want = NUMPY[name](x, y).astype(np.float32)
    got = oracle.math_eval(name, x, y)
    bad = int((~same(got, want)).sum())
    print(f"{name}: {bad} of {x.size} differ from numpy float64 rounded to float", flush=True)
    # Two double implementations rounded to float part only where the value
    # lies within a few 53-bit units of a float rounding boundary.
    assert bad <= 2


@pytest.mark.parametrize("name", list(oracle.MATH_FUNCTIONS))
def test_product_routines_match_the_oracle_bit_for_bit(name):
    x, y = samples(name, 1 << 20, 5)
    got = N.math_eval_host(name, x, y)
    want = oracle.math_eval(name, x, y)
    bad = ~same(got, want)
    print(f"{name}: {int(bad.sum())} of {x.size} differ", flush=True)
    if bad.any():
        i = int(np.flatnonzero(bad)[0])
        print(f"  first: x={x[i]!r} y={None if y is None else y[i]!r} got={got[i]!r} want={want[i]!r}")
    # The routines stay within a few 53-bit units of the exact value, so a
    # difference after rounding to float is expected about once in 2^26.
    assert int(bad.sum()) <= 1


def test_special_values():
    xs, ys = np.meshgrid(SPECIAL, SPECIAL, indexing="ij")
    xs, ys = xs.reshape(-1).copy(), ys.reshape(-1).copy()
    for name in oracle.MATH_FUNCTIONS:
        two = name in ("atan2", "pow")
        x = xs if two else SPECIAL
        y = ys if two else None
        if name in ("sin", "cos"):
            x = SPECIAL[np.abs(SPECIAL) < 1e20]   # the device reduction covers |x| < 2^30
        got = N.math_eval_host(name, x, y)
        want = oracle.math_eval(name, x, y)
        bad = np.flatnonzero(~same(got, want))
        assert bad.size == 0, [(name, x[i], None if y is None else y[i], got[i], want[i]) for i in bad[:8]]


def test_closed_forms():
    # Values the mathematics fixes, through the product routines.
    f = N.math_eval_host
    assert f("sin", np.float32([0.0]))[0] == 0.0
    assert f("cos", np.float32([0.0]))[0] == 1.0
    assert f("exp", np.float32([0.0]))[0] == 1.0
    assert f("pow", np.float32([0.7]), np.float32([0.0]))[0] == 1.0
    assert f("pow", np.float32([2.0]), np.float32([10.0]))[0] == 1024.0
    assert f("pow", np.float32([0.25]), np.float32([0.5]))[0] == 0.5
    assert f("acos", np.float32([-1.0]))[0] == np.float32(np.pi)
    assert f("asin", np.float32([1.0]))[0] == np.float32(np.pi / 2)
    assert f("atan2", np.float32([1.0]), np.float32([1.0]))[0] == np.float32(np.pi / 4)
    assert f("atan2", np.float32([0.0]), np.float32([-1.0]))[0] == np.float32(np.pi)
    t = np.linspace(-3, 6, 10001, dtype=np.float32)
    s, c = f("sin", t).astype(np.float64), f("cos", t).astype(np.float64)
    assert np.max(np.abs(s * s + c * c - 1)) < 2.5e-7


def test_bad_function_code():
    x = np.zeros(4, dtype=np.float32)
    out = np.zeros(4, dtype=np.float32)
    assert N.load().az_math_eval(99, x.ctypes.data, None, 4, out.ctypes.data) == 1   # AZ_ERR_INVALID_ARGUMENT
    assert N.load().az_math_eval(N.MATH_FUNCTIONS["pow"], x.ctypes.data, None, 4, out.ctypes.data) == \
        1

"""The output stage on the GPU (az_gpu_output_pixels) against the oracle's
restatement of OutputPixelsToFile's colour path and the closed forms that pin
it.  The GPU sums the log luminances in the reference's pixel order, so the
scale and the white point match, and takes powf as the correctly rounded
binary32 value of the double power; on these seeded inputs every byte
matches.  (libm's powf and log are not correctly rounded in every case, so
other inputs could move a byte by one where a value lands within an ulp of a
multiple of 1/255; none of these does.)"""
import numpy as np
import pytest

import analytic_pins as P
import oracle
from azimuth_amd import native as N, workloads

pytestmark = pytest.mark.gpu


def tonemap(key=0.18, burnout=1.0, saturation=1.0):
    return P._tonemap(key, burnout, saturation)


def hdr_pixels(w, h, seed):
    """Log-normal HDR colours with weights in (0.5, 4], a few black and a few
    very bright pixels."""
    rng = np.random.default_rng(seed)
    n = w * h
    px = np.zeros(n, dtype=N.PIXEL_DTYPE)
    wgt = rng.uniform(0.5, 4.0, n).astype(np.float32)
    col = np.exp(rng.normal(-1.5, 1.5, (n, 3))).astype(np.float32) * wgt[:, None]
    col[rng.integers(0, n, n // 100)] = 0.0
    col[rng.integers(0, n, n // 1000)] *= 1000.0
    px["Color"] = col
    px["TotalWeight"] = wgt
    return px


def compare(gpu_out, cpu_out, max_frac):
    d = np.abs(gpu_out.astype(int) - cpu_out.astype(int))
    frac = float(np.count_nonzero(d)) / d.size
    print(f"bytes {d.size}: differ {np.count_nonzero(d)} ({frac:.2e}), max {int(d.max()) if d.size else 0}", flush=True)
    assert int(d.max(initial=0)) <= 1
    assert frac <= max_frac


def test_output_closed_forms(gpu):
    P.check_output_bytes_and_srgb(gpu.output_pixels)
    P.check_photographic_operator(gpu.output_pixels)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (640, 360), (1920, 1080)])
def test_output_matches_oracle_exactly_without_powers(gpu, w, h):
    """Saturation 1 and gamma 1 leave only log and exp (in double, rounded to
    binary32) between the GPU and the oracle: every byte matches."""
    px = hdr_pixels(w, h, 7)
    for burnout in (0.0, 1.0, 50.0, 100.0):
        g = gpu.output_pixels(px, w, h, tonemap(0.18, burnout, 1.0), 1.0, False)
        c = oracle.output_pixels(px, w, h, tonemap(0.18, burnout, 1.0), 1.0, False)
        compare(g, c, 0.0)


@pytest.mark.parametrize("tm,inv_gamma,srgb", [(tonemap(0.18, 1.0, 0.8), 1.0 / 2.4, True),
                                               (tonemap(0.36, 5.0, 1.0), 1.0 / 2.2, False),
                                               (None, 1.0 / 2.2, False), (None, 1.0 / 2.4, True)])
def test_output_matches_oracle(gpu, tm, inv_gamma, srgb):
    w, h = 1920, 1080
    px = hdr_pixels(w, h, 11)
    g = gpu.output_pixels(px, w, h, tm, inv_gamma, srgb)
    c = oracle.output_pixels(px, w, h, tm, inv_gamma, srgb)
    compare(g, c, 0.0)


def test_output_of_a_render(gpu):
    """The interior render's pixel buffer, tonemapped and sRGB-encoded as a
    camera with the photographic operator would save it."""
    sc, cam = workloads.interior_scene(240, 135, spp=16, max_depth=4, shells=12)
    gpu.upload_scene(sc)
    px = gpu.render(cam, filt=N.FilterType_Gaussian, seed=1)
    for tm in (tonemap(0.18, 1.0, 1.0), tonemap(0.18, 2.0, 0.7)):
        g = gpu.output_pixels(px, 240, 135, tm, 1.0 / 2.4, True)
        c = oracle.output_pixels(px, 240, 135, tm, 1.0 / 2.4, True)
        compare(g, c, 0.0)
        assert g.std() > 10.0     # an image, not a flat field


def test_output_rejects_a_burnout_outside_the_image(gpu):
    px = hdr_pixels(4, 4, 1)
    for burnout in (-1.0, 101.0, float("nan")):
        with pytest.raises(N.AzError):
            gpu.output_pixels(px, 4, 4, tonemap(0.18, burnout, 1.0), 1.0, False)
    # An empty image is a no-op.
    assert gpu.output_pixels(px[:0], 0, 0, tonemap(), 1.0, False).size == 0


def sequential_sum(x):
    """numpy's float32 accumulate adds in sequence order, one rounding per step."""
    x = np.asarray(x, np.float32)
    return np.add.accumulate(x, dtype=np.float32)[-1] if x.size else np.float32(0.0)


def sum_inputs(kind, n, rng):
    if kind == "logs":
        return np.log(rng.exponential(0.3, n).astype(np.float32) + np.float32(1e-5))
    if kind == "dyadic":            # few mantissa bits: many exact ties
        return (rng.integers(-40, 8, n) / 8.0).astype(np.float32)
    if kind == "near_zero":         # the running sum keeps crossing zero
        return (rng.integers(-3, 4, n) * 0.5).astype(np.float32)
    if kind == "wide":              # magnitudes over many binades
        return (rng.normal(0, 1, n) * np.exp2(rng.integers(-20, 20, n))).astype(np.float32)
    if kind == "with_inf":
        x = rng.normal(-1.0, 1.0, n).astype(np.float32)
        x[n // 2] = np.inf
        return x
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["logs", "dyadic", "near_zero", "wide", "with_inf"])
def test_sequential_sum_bit_exact(gpu, kind):
    """The parallel fold reproduces the sequential float32 sum bit for bit,
    at sizes from one value to a 4K image's worth, including partial chunks
    and inputs full of rounding ties."""
    rng = np.random.default_rng(5)
    for n in (1, 2, 63, 64, 65, 1023, 1024, 1025, 4097, 100003, 2073600, 8294400):
        x = sum_inputs(kind, n, rng)
        want = sequential_sum(x)
        got = gpu.sum_f32_sequential(x)
        assert (got == want) or (np.isnan(got) and np.isnan(want)), (kind, n, got, want)
        assert np.signbit(got) == np.signbit(want), (kind, n, got, want)
    assert gpu.sum_f32_sequential(np.zeros(0, np.float32)) == 0.0

"""Committed golden fixtures (tests/golden/, written by make_golden.py from
the oracle): the oracle must keep reproducing them bit for bit, and the GPU
library must match them to the parity bar (closest-hit ids, t, beta, gamma
exact; occlusion flags exact; image within the render tolerance)."""
import os

import numpy as np
import pytest

import oracle
from azimuth_amd import native as N

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

import importlib.util  # noqa: E402
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)


def load_trace():
    z = np.load(os.path.join(HERE, "trace_icosphere_s3.npz"))
    return z["rays"].view(N.RAY_DTYPE), z["closest"].view(N.HIT_DTYPE), z["any"].view(N.HIT_DTYPE)


def load_cornell():
    return np.load(os.path.join(HERE, "cornell_32x32_4spp.npz"))["pixels"].view(N.PIXEL_DTYPE)


def test_fixture_rays_are_the_generated_batch():
    _, rays = make_golden.trace_inputs()
    want, _, _ = load_trace()
    assert rays.tobytes() == want.tobytes()


def test_oracle_reproduces_trace_fixture():
    sc, _ = make_golden.trace_inputs()
    rays, closest, anyh = load_trace()
    assert oracle.trace(sc, rays).tobytes() == closest.tobytes()
    assert oracle.trace(sc, rays, any_hit=True).tobytes() == anyh.tobytes()


def test_oracle_reproduces_cornell_fixture():
    sc, cam = make_golden.cornell_inputs()
    px = oracle.render(sc, cam, filt=N.FilterType_Box, seed=5)
    assert px.tobytes() == load_cornell().tobytes()


def test_oracle_reproduces_pcg32_fixture():
    z = np.load(os.path.join(HERE, "pcg32.npz"))
    got = [oracle.rng_draw(int(a), int(b), int(c), int(d), int(e))
           for a, b, c, d, e in zip(z["seed"], z["pixel"], z["sample"], z["node"], z["dim"])]
    assert np.array_equal(np.array(got, dtype=np.uint32), z["value"])


@pytest.mark.gpu
def test_gpu_matches_trace_fixture(gpu):
    sc, _ = make_golden.trace_inputs()
    rays, closest, anyh = load_trace()
    gpu.upload_scene(sc)
    got = gpu.trace(rays)
    assert np.array_equal(got["Prim"], closest["Prim"])
    hit = closest["Prim"] != N.AZ_MISS
    for f in ("t", "Beta", "Gamma"):
        assert np.array_equal(got[f][hit], closest[f][hit]), f
    occ = gpu.trace(rays, any_hit=True)
    assert np.array_equal(occ["Prim"] != N.AZ_MISS, anyh["Prim"] != N.AZ_MISS)


@pytest.mark.gpu
def test_gpu_matches_cornell_fixture(gpu):
    sc, cam = make_golden.cornell_inputs()
    gpu.upload_scene(sc)
    got = gpu.render(cam, filt=N.FilterType_Box, seed=5)
    want = load_cornell()
    assert np.array_equal(got["TotalWeight"], want["TotalWeight"])
    g = got["Color"] / got["TotalWeight"][:, None]
    o = want["Color"] / want["TotalWeight"][:, None]
    scale = float(np.mean(np.abs(o)))
    rmse = float(np.sqrt(np.mean((g - o) ** 2))) / scale
    print(f"cornell fixture: relative RMSE {rmse:.3e}", flush=True)
    assert rmse < 1e-3

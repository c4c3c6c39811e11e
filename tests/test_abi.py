"""The C-ABI library: it loads, exports every function include/*.h declares,
its structure sizes match the bindings, and its host-side error paths return
status codes.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from azimuth_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(az_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = N.load()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_declarations():
    assert declared_functions() <= set(N.EXPORTS) | {"az_abi_sizes"}


def test_abi_struct_sizes():
    lib = N.load()
    buf = (C.c_uint32 * 16)()
    n = lib.az_abi_sizes(buf, 16)
    assert [buf[i] for i in range(n)] == [s for _, s in N.ABI_SIZES]
    assert N.RAY_DTYPE.itemsize == 32 and N.HIT_DTYPE.itemsize == 16 and N.PIXEL_DTYPE.itemsize == 16


def test_builder_rejects_bad_input():
    lib = N.load()
    h = C.c_void_p()
    assert lib.az_sb_create(C.byref(h)) == N.AZ_OK
    out = C.c_void_p()
    assert lib.az_sb_build(h, C.byref(out)) == 1            # nothing added
    v = np.zeros((3, 3), np.float32)
    f = np.array([[0, 1, 5]], np.uint32)                    # index out of range
    d = N.az_mesh_desc()
    d.Vertices, d.VertexCount, d.Faces, d.FaceCount = v.ctypes.data, 3, f.ctypes.data, 1
    d.Transformation = N.mat4.identity()
    assert lib.az_sb_add_mesh(h, C.byref(d)) == 1
    s = N.az_sphere_desc()
    s.Radius = -1.0
    s.Transformation = N.mat4.identity()
    assert lib.az_sb_add_sphere(h, C.byref(s)) == 1
    sing = N.az_sphere_desc()
    sing.Radius = 1.0                                       # all-zero transform: singular
    assert lib.az_sb_add_sphere(h, C.byref(sing)) == N.AZ_OK
    assert lib.az_sb_build(h, C.byref(out)) == 1
    lib.az_sb_destroy(h)


def test_calls_without_context_or_scene():
    lib = N.load()
    assert lib.az_gpu_trace_closest(None, None, 0, None, 0) == 1
    assert lib.az_gpu_upload_scene(None, None, None) == 1
    assert lib.az_gpu_get_stats(None, None) == 1
    assert lib.az_gpu_last_error(None) == b"no context"
    ctx = C.c_void_p()
    st = lib.az_gpu_create(C.byref(ctx), 9999)               # no such device
    assert st in (3,) and not ctx.value


def test_mat4_inverse():
    lib = N.load()
    g = np.random.default_rng(0)
    M = np.eye(4)
    M[:3, :3] = g.normal(size=(3, 3)) + 3 * np.eye(3)
    M[:3, 3] = g.normal(size=3)
    a, b = N.mat4.from_numpy(M), N.mat4()
    assert lib.az_mat4_inverse(C.byref(a), C.byref(b)) == N.AZ_OK
    inv = np.array(b.E[:]).reshape(4, 4)
    np.testing.assert_allclose(inv @ M, np.eye(4), atol=1e-5)

"""ctypes mirror of the C ABI in include/azimuth_types.h, azimuth_scene.h and
azimuth_gpu.h, and the loader of libazimuth_gpu.so.

The library is loaded from this package directory; nothing here falls back to
another implementation when it is missing: `load()` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libazimuth_gpu.so")

f32, u32, s32, u64, f64 = C.c_float, C.c_uint32, C.c_int32, C.c_uint64, C.c_double


class vec2(C.Structure):
    _fields_ = [("x", f32), ("y", f32)]


class vec2u(C.Structure):
    _fields_ = [("x", u32), ("y", u32)]


class vec3(C.Structure):
    _fields_ = [("x", f32), ("y", f32), ("z", f32)]

    def __init__(self, *v):
        if len(v) == 1:
            v = tuple(v[0])
        super().__init__(*[float(a) for a in v]) if v else super().__init__()

    def tuple(self):
        return (self.x, self.y, self.z)


class vec4(C.Structure):
    _fields_ = [("x", f32), ("y", f32), ("z", f32), ("w", f32)]


class mat4(C.Structure):
    _fields_ = [("E", f32 * 16)]

    @staticmethod
    def from_numpy(m):
        r = mat4()
        flat = np.asarray(m, dtype=np.float32).reshape(16)
        for i in range(16):
            r.E[i] = float(flat[i])
        return r

    @staticmethod
    def identity():
        return mat4.from_numpy(np.eye(4, dtype=np.float32))


# brdf_type (brdf.h:1-9)
BRDF_Phong, BRDF_ModifiedPhong, BRDF_NormalizedModifiedPhong, BRDF_BlinnPhong, \
    BRDF_ModifiedBlinnPhong, BRDF_NormalizedModifiedBlinnPhong, BRDF_TorranceSparrow = range(7)
# light_type (light.h:1-7)
LightType_PointLight, LightType_AreaLight, LightType_DirectionalLight, LightType_SpotLight, \
    LightType_EntityLight = range(5)
ShadingMode_Flat, ShadingMode_Smooth = 0, 1
FilterType_Box, FilterType_Gaussian = 0, 1
TonemapOperator_None, TonemapOperator_Photographic = 0, 1
CameraType_Detailed, CameraType_Simple = 0, 1
TextureType_2D, TextureType_PerlinNoise = 0, 1
PixelFormat_RGB_U8, PixelFormat_RGB_F32 = 0, 1
TextureWrapMode_Clamp, TextureWrapMode_Repeat = 0, 1
TextureBlendMode_ReplaceAll, TextureBlendMode_ReplaceDiffuse, TextureBlendMode_BlendDiffuse = 0, 1, 2
TextureFilter_Nearest, TextureFilter_Bilinear = 0, 1
TexturePerlinAppearance_Patch, TexturePerlinAppearance_Vein = 0, 1

AZ_OK = 0
AZ_MISS = 0xFFFFFFFF
STATUS_NAMES = {0: "AZ_OK", 1: "AZ_ERR_INVALID_ARGUMENT", 2: "AZ_ERR_OUT_OF_MEMORY", 3: "AZ_ERR_DEVICE",
                4: "AZ_ERR_NO_SCENE", 5: "AZ_ERR_UNSUPPORTED", 6: "AZ_ERR_NOT_BUILT"}


class brdf(C.Structure):
    _fields_ = [("Exponent", f32), ("RefractiveIndex", f32), ("Type", s32), ("ID", u32)]


class material(C.Structure):
    _fields_ = [("BRDF", brdf), ("AmbientReflectance", vec3), ("DiffuseReflectance", vec3),
                ("SpecularReflectance", vec3), ("MirrorReflectance", vec3), ("Transparency", vec3),
                ("RefractionIndex", f32), ("Glossiness", f32), ("Degamma", u32), ("BRDFID", u32), ("ID", u32)]


class _light_area(C.Structure):
    _fields_ = [("Position", vec3), ("Intensity", vec3), ("EdgeVectors", vec3 * 2), ("Normal", vec3)]


class _light_spot(C.Structure):
    _fields_ = [("Position", vec3), ("Intensity", vec3), ("SpotDirection", vec3), ("CoverageAngle", f32),
                ("FallOffAngle", f32), ("CosHalfCoverage", f32), ("CosHalfFallOff", f32)]


class _light_union(C.Union):
    _fields_ = [("area", _light_area), ("spot", _light_spot), ("raw", C.c_uint64 * 36)]


class light(C.Structure):
    """light.h:9-42; the entity-light member is only reachable as raw bytes."""
    _fields_ = [("u", _light_union), ("ID", u32), ("Type", s32)]


class _tex_image(C.Structure):
    _fields_ = [("Pixels", C.c_void_p), ("PixelFormat", s32), ("Width", u32), ("Height", u32),
                ("InvWidth", f32), ("InvHeight", f32), ("Normalizer", f32), ("Loaded", u32),
                ("ImageFileFormat", s32)]


class _tex_perlin(C.Structure):
    _fields_ = [("Perm", u32 * 256), ("ScalingFactor", f32), ("PerlinAppearance", s32)]


class _tex_union(C.Union):
    _fields_ = [("img", _tex_image), ("perlin", _tex_perlin)]


class texture(C.Structure):
    _fields_ = [("ImagePath", C.c_char * 256), ("u", _tex_union), ("BumpMapMultiplier", f32),
                ("BumpMap", u32), ("Degamma", u32), ("ID", u32), ("WrapMode", s32), ("BlendMode", s32),
                ("Filter", s32), ("Type", s32)]


class tonemap(C.Structure):
    _fields_ = [("KeyValue", f32), ("BurnoutPercentage", f32), ("Saturation", f32), ("Operator", s32)]


class camera(C.Structure):
    _fields_ = [("Tonemap", tonemap), ("ImageName", C.c_char * 256), ("SaveFormat", s32),
                ("NearPlane", vec4), ("Position", vec3), ("Gaze", vec3), ("Up", vec3),
                ("BackgroundColor", vec3), ("Resolution", vec2u), ("NearDistance", f32),
                ("FocusDistance", f32), ("ApertureSize", f32), ("Gamma", f32), ("InvGamma", f32),
                ("GammaCorrection_sRGB", u32), ("GazeIsPoint", u32), ("IsLeftHanded", u32), ("SaveEXR", u32),
                ("SampleCount", u32), ("ID", u32), ("Type", s32)]


class epsilon_values(C.Structure):
    _fields_ = [("Intersection", f32), ("ShadowRay", f32)]


class pixel(C.Structure):
    _fields_ = [("Color", vec3), ("TotalWeight", f32)]


class az_mesh_desc(C.Structure):
    _fields_ = [("Vertices", C.c_void_p), ("VertexCount", u32), ("Faces", C.c_void_p), ("FaceCount", u32),
                ("UVs", C.c_void_p), ("UVCount", u32), ("UVOffset", s32), ("Transformation", mat4),
                ("MotionVector", vec3), ("Radiance", vec3), ("IsEmitter", u32), ("TextureIndex", u32),
                ("MaterialIndex", u32), ("ShadingMode", s32)]


class az_sphere_desc(C.Structure):
    _fields_ = [("Center", vec3), ("Radius", f32), ("Transformation", mat4), ("MotionVector", vec3),
                ("Radiance", vec3), ("IsEmitter", u32), ("TextureIndex", u32), ("MaterialIndex", u32)]


class az_scene_params(C.Structure):
    _fields_ = [("AmbientLight", vec3), ("BackgroundColor", vec3), ("MaxRecursionDepth", u32),
                ("PathTrace", u32), ("ImportanceSampling", u32)]


AZ_ENV_COSINE, AZ_ENV_LUMINANCE = 0, 1
AZ_SCATTER_REFERENCE, AZ_SCATTER_BRDF = 0, 1
AZ_GUIDE_ENV, AZ_GUIDE_SCATTER, AZ_GUIDE_BOTH = 0, 1, 2
AZ_BUILD_AUTO, AZ_BUILD_HOST, AZ_BUILD_DEVICE = 0, 1, 2
AZ_LEVELS_AUTO, AZ_LEVELS_HOST = 0, 1


class az_render_opts(C.Structure):
    _fields_ = [("Seed", u64), ("SampleCountOverride", u32), ("RowBegin", u32), ("RowEnd", u32),
                ("AccumulateF64", u32), ("GuideEnable", u32), ("GuideFraction", f32), ("GuideTrainSpp", u32),
                ("GuideStepsPerChunk", u32), ("EnvSampling", u32), ("GuideMaxDepth", u32), ("GuideChunkSpp", u32),
                ("ScatterSampling", u32), ("GuideTarget", u32)]


class az_stats(C.Structure):
    _fields_ = [("LastKernelMs", f64), ("NodesVisited", u64), ("TrianglesTested", u64),
                ("PrimitivesTested", u64), ("NodeBytes", u32), ("TriangleBytes", u32), ("BvhWidth", u32),
                ("StackDepth", u32), ("BlockSize", u32), ("NodeCount", u32), ("LeafCount", u32),
                ("TriangleCount", u32), ("SphereCount", u32), ("InstanceCount", u32)]


class az_render_stats(C.Structure):
    _fields_ = [("CameraSamples", u64), ("PathRays", u64), ("ShadowRays", u64), ("ExtendNodes", u64),
                ("ExtendTriangles", u64), ("ExtendItems", u64), ("ConnectNodes", u64), ("ConnectTriangles", u64),
                ("ConnectItems", u64), ("Pixels", u64), ("Gaussian", u32), ("Chunks", u32)]


class az_tune_entry(C.Structure):
    _fields_ = [("Block", u32), ("LdsStackEntries", u32), ("MinWaves", s32), ("ChunkSamples", u32),
                ("LoopForm", u32), ("BvhWidth", u32), ("MaxLeafItems", u32), ("Reserved", u32),
                ("Milliseconds", f64)]


class az_tune_result(C.Structure):
    _fields_ = [("Count", u32), ("Best", u32), ("Entries", az_tune_entry * 32)]


# az_ray / az_hit as numpy record layouts (32 and 16 bytes).
RAY_DTYPE = np.dtype([("Origin", "<f4", 3), ("Time", "<f4"), ("Direction", "<f4", 3), ("TMax", "<f4")])
HIT_DTYPE = np.dtype([("Prim", "<u4"), ("t", "<f4"), ("Beta", "<f4"), ("Gamma", "<f4")])
PIXEL_DTYPE = np.dtype([("Color", "<f4", 3), ("TotalWeight", "<f4")])

ABI_SIZES = [("entity", 288), ("mesh", 200), ("material", C.sizeof(material)), ("light", C.sizeof(light)),
             ("texture", C.sizeof(texture)), ("camera", C.sizeof(camera)), ("scene", 2744),
             ("az_mesh_desc", C.sizeof(az_mesh_desc)), ("az_sphere_desc", C.sizeof(az_sphere_desc)),
             ("az_scene_params", C.sizeof(az_scene_params)), ("az_render_opts", C.sizeof(az_render_opts)),
             ("az_stats", C.sizeof(az_stats)), ("az_ray", RAY_DTYPE.itemsize), ("az_hit", HIT_DTYPE.itemsize),
             ("az_tune_result", C.sizeof(az_tune_result))]

_P = C.c_void_p
# Exported functions: name -> (restype, argtypes).  Every name here is also
# declared in include/*.h.
EXPORTS = {
    # azimuth_scene.h
    "az_sb_create": (s32, [C.POINTER(_P)]),
    "az_sb_destroy": (None, [_P]),
    "az_sb_set_params": (s32, [_P, C.POINTER(az_scene_params)]),
    "az_sb_add_material": (s32, [_P, C.POINTER(material), C.POINTER(u32)]),
    "az_sb_add_texture": (s32, [_P, C.POINTER(texture), C.POINTER(u32)]),
    "az_sb_set_environment_map": (s32, [_P, C.POINTER(texture)]),
    "az_sb_set_background_texture": (s32, [_P, C.POINTER(texture)]),
    "az_sb_add_light": (s32, [_P, C.POINTER(light)]),
    "az_sb_add_mesh": (s32, [_P, C.POINTER(az_mesh_desc)]),
    "az_sb_add_sphere": (s32, [_P, C.POINTER(az_sphere_desc)]),
    "az_sb_build": (s32, [_P, C.POINTER(_P)]),
    "az_scene_free": (None, [_P]),
    "az_mat4_inverse": (s32, [C.POINTER(mat4), C.POINTER(mat4)]),
    "az_ply_read": (s32, [C.c_char_p, _P, C.POINTER(u32), _P, C.POINTER(u32), _P, C.POINTER(u32)]),
    "az_sb_add_ply_mesh": (s32, [_P, C.c_char_p, C.POINTER(az_mesh_desc)]),
    "az_read_binary_vertices": (s32, [C.c_char_p, _P, C.POINTER(u32)]),
    "az_read_binary_faces": (s32, [C.c_char_p, u32, u32, _P, C.POINTER(u32)]),
    "az_scene_triangle_count": (u32, [_P]),
    "az_abi_sizes": (u32, [C.POINTER(u32), u32]),
    # azimuth_platform_posix.h
    "az_platform_posix_api": (None, [_P]),
    # azimuth_gpu.h
    "az_gpu_create": (s32, [C.POINTER(_P), C.c_int]),
    "az_gpu_destroy": (None, [_P]),
    "az_gpu_last_error": (C.c_char_p, [_P]),
    "az_gpu_upload_scene": (s32, [_P, _P, C.POINTER(epsilon_values)]),
    "az_gpu_trace_closest": (s32, [_P, _P, u32, _P, u32]),
    "az_gpu_trace_any": (s32, [_P, _P, u32, _P]),
    "az_gpu_render": (s32, [_P, C.POINTER(camera), s32, u32, C.POINTER(az_render_opts), _P]),
    "az_gpu_render_device": (s32, [_P, C.POINTER(camera), s32, u32, C.POINTER(az_render_opts), _P]),
    "az_gpu_get_stats": (s32, [_P, C.POINTER(az_stats)]),
    "az_gpu_output_pixels": (s32, [_P, _P, u32, u32, C.POINTER(tonemap), f32, u32, _P]),
    "az_gpu_sum_f32_sequential": (s32, [_P, _P, u32, _P]),
    "az_gpu_sum_f32_sequential_noalloc": (s32, [_P, _P, u32, _P]),
    "az_gpu_set_counting": (s32, [_P, u32]),
    "az_gpu_get_render_stats": (s32, [_P, C.POINTER(az_render_stats)]),
    "az_gpu_set_launch": (s32, [_P, u32, u32, s32, s32]),
    "az_gpu_autotune_trace": (s32, [_P, _P, u32, u32, _P]),
    "az_gpu_render_batch": (s32, [_P, _P, u32, s32, u32, _P, _P]),
    "az_gpu_set_build": (s32, [_P, u32, f32, f32]),
    "az_gpu_autotune_render": (s32, [_P, _P, s32, u32, _P, _P]),
    "az_gpu_set_render_tuning": (s32, [_P, u32, u32, s32, s32]),
    "az_gpu_trim": (s32, [_P]),
    "az_gpu_set_bvh_width": (s32, [_P, u32]),
    "az_gpu_set_build_mode": (s32, [_P, u32]),
    "az_gpu_set_level_mode": (s32, [_P, u32]),
    "az_gpu_level_graph_counts": (s32, [_P, C.POINTER(u64), C.POINTER(u64)]),
    "az_gpu_autotune_build": (s32, [_P, _P, _P, _P, u32, u32, _P]),
    "az_gpu_set_render_memory_limit": (s32, [_P, u64]),
    "az_gpu_env_luminance_sample": (s32, [_P, _P, u32, _P]),
    "az_gpu_env_luminance_pdf": (s32, [_P, _P, u32, _P]),
    "az_gpu_math_eval": (s32, [_P, u32, _P, _P, u32, _P]),
    "az_math_eval": (s32, [u32, _P, _P, u32, _P]),
    "az_gpu_malloc": (s32, [_P, C.POINTER(_P), u64]),
    "az_gpu_free": (s32, [_P, _P]),
    "az_gpu_upload": (s32, [_P, _P, _P, u64]),
    "az_gpu_download": (s32, [_P, _P, _P, u64]),
    "az_gpu_memset": (s32, [_P, _P, C.c_int, u64]),
    "az_gpu_synchronize": (s32, [_P]),
    "az_gpu_device_copy": (s32, [_P, _P, _P, u64]),
    "az_gpu_record_launch_times": (s32, [_P, u32]),
    "az_gpu_collect_launch_times": (s32, [_P, C.POINTER(f64), u32, C.POINTER(u32)]),
    "az_gpu_timer_start": (s32, [_P]),
    "az_gpu_timer_stop": (s32, [_P, C.POINTER(f64)]),
    "az_render_scene": (u32, [_P, C.POINTER(camera), C.POINTER(epsilon_values), u32, u32, _P]),
    # azimuth_guide.h
    "az_guide_param_count": (u32, []),
    "az_guide_create": (s32, [C.POINTER(_P), C.c_int, u64]),
    "az_guide_destroy": (None, [_P]),
    "az_guide_last_error": (C.c_char_p, [_P]),
    "az_guide_set_params": (s32, [_P, _P, u32]),
    "az_guide_get_params": (s32, [_P, _P, u32]),
    "az_guide_get_grad": (s32, [_P, _P, u32]),
    "az_guide_zero_grad": (s32, [_P]),
    "az_guide_set_learning_rate": (s32, [_P, f32]),
    "az_guide_infer": (s32, [_P, _P, u32, _P]),
    "az_guide_infer_device": (s32, [_P, _P, u32, _P]),
    "az_guide_infer_packed": (s32, [_P, _P, u32, _P]),
    "az_guide_infer_packed_device": (s32, [_P, _P, u32, _P]),
    "az_guide_synchronize": (s32, [_P]),
    "az_guide_train": (s32, [_P, _P, u32, u32, C.POINTER(f32)]),
    "az_guide_train_device": (s32, [_P, _P, u32, u32, C.POINTER(f32)]),
    "az_guide_mfma_selftest": (s32, [_P, _P, _P, _P]),
}

_lib = None


class AzError(RuntimeError):
    pass


def load():
    """Loads libazimuth_gpu.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AzError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    buf = (u32 * 16)()
    n = lib.az_abi_sizes(buf, 16)
    got = [buf[i] for i in range(n)]
    want = [s for _, s in ABI_SIZES]
    if got != want:
        raise AzError(f"ABI size mismatch: library {got} vs bindings {want}")
    _lib = lib
    return lib


def check(status, what="", ctx=None):
    if status != AZ_OK:
        msg = STATUS_NAMES.get(status, str(status))
        if ctx is not None:
            msg += ": " + load().az_gpu_last_error(ctx).decode(errors="replace")
        raise AzError(f"{what} failed with {msg}")


# az_math_fn codes of az_math_eval / az_gpu_math_eval.
MATH_FUNCTIONS = {"sin": 0, "cos": 1, "atan2": 2, "pow": 3, "exp": 4, "asin": 5, "acos": 6, "pow4": 7, "pow5": 8}


def math_eval_host(name, x, y=None):
    """The renderer's float elementary function `name` of float32 arrays x
    (and y), evaluated on the host from the device routines' source."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    yp = None
    if y is not None:
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        assert y.size == x.size
        yp = y.ctypes.data
    out = np.empty_like(x)
    check(load().az_math_eval(MATH_FUNCTIONS[name], x.ctypes.data, yp, x.size, out.ctypes.data), "az_math_eval")
    return out


def ply_read(path):
    """az_ply_read: (vertices (V, 3) f32, faces (F, 3) u32, uvs (V, 2) f32 or
    None) of a PLY file."""
    lib = load()
    p = str(path).encode()
    nv, nf, nuv = u32(0), u32(0), u32(0)
    check(lib.az_ply_read(p, None, C.byref(nv), None, C.byref(nf), None, C.byref(nuv)), "az_ply_read")
    v = np.zeros((nv.value, 3), np.float32)
    f = np.zeros((nf.value, 3), np.uint32)
    uv = np.zeros((nuv.value, 2), np.float32)
    check(lib.az_ply_read(p, v.ctypes.data, C.byref(nv), f.ctypes.data, C.byref(nf),
                          uv.ctypes.data if nuv.value else None, C.byref(nuv)), "az_ply_read")
    return v, f, (uv if nuv.value else None)


def read_binary_vertices(path):
    lib = load()
    p = str(path).encode()
    n = u32(0)
    check(lib.az_read_binary_vertices(p, None, C.byref(n)), "az_read_binary_vertices")
    v = np.zeros((n.value, 3), np.float32)
    check(lib.az_read_binary_vertices(p, v.ctypes.data, C.byref(n)), "az_read_binary_vertices")
    return v


def read_binary_faces(path, vertex_offset=0, zero_based=False):
    lib = load()
    p = str(path).encode()
    n = u32(0)
    check(lib.az_read_binary_faces(p, vertex_offset, 1 if zero_based else 0, None, C.byref(n)),
          "az_read_binary_faces")
    f = np.zeros((n.value, 3), np.uint32)
    check(lib.az_read_binary_faces(p, vertex_offset, 1 if zero_based else 0, f.ctypes.data, C.byref(n)),
          "az_read_binary_faces")
    return f

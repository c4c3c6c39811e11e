"""Host-side objects over the C ABI: a built `scene` and a GPU context.

`SceneBuilder` drives az_sb_* (the importer post-process and the reference
BVH build, include/azimuth_scene.h); `GpuContext` drives az_gpu_* (upload,
batch closest/any hit, render).  Errors raise `native.AzError`.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import native as N


def _vec3(v):
    return N.vec3(*[float(a) for a in v])


def make_material(kd=(0.5, 0.5, 0.5), ks=(0, 0, 0), ka=(0, 0, 0), mirror=(0, 0, 0), transparency=(0, 0, 0),
                  ior=1.0, exponent=1.0, brdf_type=N.BRDF_NormalizedModifiedBlinnPhong, brdf_ior=0.0,
                  glossiness=0.0):
    m = N.material()
    m.BRDF.Type = brdf_type
    m.BRDF.Exponent = float(exponent)
    m.BRDF.RefractiveIndex = float(brdf_ior)
    m.DiffuseReflectance = _vec3(kd)
    m.SpecularReflectance = _vec3(ks)
    m.AmbientReflectance = _vec3(ka)
    m.MirrorReflectance = _vec3(mirror)
    m.Transparency = _vec3(transparency)
    m.RefractionIndex = float(ior)
    m.Glossiness = float(glossiness)
    return m


def point_light(position, intensity):
    L = N.light()
    L.Type = N.LightType_PointLight
    L.u.area.Position = _vec3(position)
    L.u.area.Intensity = _vec3(intensity)
    return L


def area_light(position, intensity, edge0, edge1, normal):
    L = N.light()
    L.Type = N.LightType_AreaLight
    L.u.area.Position = _vec3(position)
    L.u.area.Intensity = _vec3(intensity)
    L.u.area.EdgeVectors[0] = _vec3(edge0)
    L.u.area.EdgeVectors[1] = _vec3(edge1)
    L.u.area.Normal = _vec3(normal)
    return L


def directional_light(direction, intensity):
    L = N.light()
    L.Type = N.LightType_DirectionalLight
    L.u.area.Position = _vec3(direction)       # Direction aliases Position (light.h:22-25)
    L.u.area.Intensity = _vec3(intensity)
    return L


def spot_light(position, intensity, direction, coverage_deg, falloff_deg):
    """Spot light with the cosines the importer derives (half angles)."""
    L = N.light()
    L.Type = N.LightType_SpotLight
    s = L.u.spot
    s.Position = _vec3(position)
    s.Intensity = _vec3(intensity)
    s.SpotDirection = _vec3(direction)
    s.CoverageAngle = float(coverage_deg)
    s.FallOffAngle = float(falloff_deg)
    s.CosHalfCoverage = float(np.float32(np.cos(np.float32(coverage_deg / 2.0 / 180.0) * np.float32(np.pi))))
    s.CosHalfFallOff = float(np.float32(np.cos(np.float32(falloff_deg / 2.0 / 180.0) * np.float32(np.pi))))
    return L


def image_texture(pixels_rgb, normalizer=255.0, wrap=N.TextureWrapMode_Repeat, filt=N.TextureFilter_Nearest,
                  blend=N.TextureBlendMode_ReplaceDiffuse, bump=False, bump_multiplier=1.0):
    """RGB texture from an (H, W, 3) float32 or uint8 array.  The array must
    stay alive until the builder has copied it (az_sb_add_texture)."""
    arr = np.ascontiguousarray(pixels_rgb)
    t = N.texture()
    t.Type = N.TextureType_2D
    t.u.img.Pixels = arr.ctypes.data
    t.u.img.PixelFormat = N.PixelFormat_RGB_F32 if arr.dtype == np.float32 else N.PixelFormat_RGB_U8
    t.u.img.Height, t.u.img.Width = int(arr.shape[0]), int(arr.shape[1])
    t.u.img.InvWidth = float(np.float32(1.0) / np.float32(arr.shape[1]))
    t.u.img.InvHeight = float(np.float32(1.0) / np.float32(arr.shape[0]))
    t.u.img.Normalizer = float(normalizer)
    t.u.img.Loaded = 1
    t.WrapMode, t.Filter, t.BlendMode = wrap, filt, blend
    t.BumpMap = 1 if bump else 0
    t.BumpMapMultiplier = float(bump_multiplier)
    return t, arr


def perlin_texture(seed, scaling=1.0, appearance=N.TexturePerlinAppearance_Patch, bump=False,
                   bump_multiplier=1.0, blend=N.TextureBlendMode_BlendDiffuse):
    """Perlin texture whose permutation is a seeded Fisher-Yates shuffle."""
    t = N.texture()
    t.Type = N.TextureType_PerlinNoise
    perm = np.random.default_rng(seed).permutation(256).astype(np.uint32)
    for i in range(256):
        t.u.perlin.Perm[i] = int(perm[i])
    t.u.perlin.ScalingFactor = float(scaling)
    t.u.perlin.PerlinAppearance = appearance
    t.BumpMap = 1 if bump else 0
    t.BumpMapMultiplier = float(bump_multiplier)
    t.BlendMode = blend
    return t


class Scene:
    """An owned `scene*` produced by az_sb_build."""

    def __init__(self, ptr, meta=None):
        self.ptr = ptr
        self.meta = meta or {}

    @property
    def triangle_count(self):
        return N.load().az_scene_triangle_count(self.ptr)

    def close(self):
        if self.ptr:
            N.load().az_scene_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SceneBuilder:
    def __init__(self, max_depth=1, path_trace=False, importance_sampling=False, ambient=(0, 0, 0),
                 background=(0, 0, 0)):
        self.lib = N.load()
        self.h = C.c_void_p()
        N.check(self.lib.az_sb_create(C.byref(self.h)), "az_sb_create")
        p = N.az_scene_params()
        p.AmbientLight = _vec3(ambient)
        p.BackgroundColor = _vec3(background)
        p.MaxRecursionDepth = int(max_depth)
        p.PathTrace = 1 if path_trace else 0
        p.ImportanceSampling = 1 if importance_sampling else 0
        N.check(self.lib.az_sb_set_params(self.h, C.byref(p)), "az_sb_set_params")
        self._keep = []

    def add_material(self, m):
        idx = C.c_uint32()
        N.check(self.lib.az_sb_add_material(self.h, C.byref(m), C.byref(idx)), "az_sb_add_material")
        return idx.value

    def add_texture(self, t):
        idx = C.c_uint32()
        N.check(self.lib.az_sb_add_texture(self.h, C.byref(t), C.byref(idx)), "az_sb_add_texture")
        return idx.value

    def set_environment_map(self, t):
        N.check(self.lib.az_sb_set_environment_map(self.h, C.byref(t)), "az_sb_set_environment_map")

    def set_background_texture(self, t):
        N.check(self.lib.az_sb_set_background_texture(self.h, C.byref(t)), "az_sb_set_background_texture")

    def add_light(self, L):
        N.check(self.lib.az_sb_add_light(self.h, C.byref(L)), "az_sb_add_light")

    def add_mesh(self, vertices, faces, material=0, transform=None, motion=(0, 0, 0), emitter=False,
                 radiance=(0, 0, 0), smooth=False, uvs=None, texture=0):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        d = N.az_mesh_desc()
        d.Vertices = v.ctypes.data
        d.VertexCount = v.shape[0]
        d.Faces = f.ctypes.data
        d.FaceCount = f.shape[0]
        if uvs is not None:
            uv = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
            d.UVs = uv.ctypes.data
            d.UVCount = uv.shape[0]
            self._keep.append(uv)
        d.Transformation = N.mat4.from_numpy(np.eye(4) if transform is None else transform)
        d.MotionVector = _vec3(motion)
        d.Radiance = _vec3(radiance)
        d.IsEmitter = 1 if emitter else 0
        d.TextureIndex = int(texture)
        d.MaterialIndex = int(material)
        d.ShadingMode = N.ShadingMode_Smooth if smooth else N.ShadingMode_Flat
        N.check(self.lib.az_sb_add_mesh(self.h, C.byref(d)), "az_sb_add_mesh")

    def add_ply_mesh(self, path, material=0, transform=None, motion=(0, 0, 0), emitter=False, radiance=(0, 0, 0),
                     smooth=False, texture=0, uv_offset=0):
        """A mesh read from a PLY file (az_sb_add_ply_mesh)."""
        d = N.az_mesh_desc()
        d.Transformation = N.mat4.from_numpy(np.eye(4) if transform is None else transform)
        d.MotionVector = _vec3(motion)
        d.Radiance = _vec3(radiance)
        d.IsEmitter = 1 if emitter else 0
        d.TextureIndex = int(texture)
        d.MaterialIndex = int(material)
        d.UVOffset = int(uv_offset)
        d.ShadingMode = N.ShadingMode_Smooth if smooth else N.ShadingMode_Flat
        N.check(self.lib.az_sb_add_ply_mesh(self.h, str(path).encode(), C.byref(d)), "az_sb_add_ply_mesh")

    def add_sphere(self, center, radius, material=0, transform=None, motion=(0, 0, 0), emitter=False,
                   radiance=(0, 0, 0), texture=0):
        d = N.az_sphere_desc()
        d.Center = _vec3(center)
        d.Radius = float(radius)
        d.Transformation = N.mat4.from_numpy(np.eye(4) if transform is None else transform)
        d.MotionVector = _vec3(motion)
        d.Radiance = _vec3(radiance)
        d.IsEmitter = 1 if emitter else 0
        d.TextureIndex = int(texture)
        d.MaterialIndex = int(material)
        N.check(self.lib.az_sb_add_sphere(self.h, C.byref(d)), "az_sb_add_sphere")

    def build(self, meta=None):
        out = C.c_void_p()
        N.check(self.lib.az_sb_build(self.h, C.byref(out)), "az_sb_build")
        return Scene(out, meta)

    def __del__(self):
        try:
            if self.h:
                self.lib.az_sb_destroy(self.h)
                self.h = None
        except Exception:
            pass


class DeviceArray:
    """Device allocation owned through a GpuContext."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        N.check(ctx.lib.az_gpu_malloc(ctx.h, C.byref(self.ptr), self.nbytes), "az_gpu_malloc", ctx.h)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        N.check(self.ctx.lib.az_gpu_upload(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "az_gpu_upload", self.ctx.h)

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        N.check(self.ctx.lib.az_gpu_download(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "az_gpu_download",
                self.ctx.h)
        return out

    def free(self):
        if self.ptr:
            self.ctx.lib.az_gpu_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GpuContext:
    def __init__(self, device=0):
        self.lib = N.load()
        self.h = C.c_void_p()
        N.check(self.lib.az_gpu_create(C.byref(self.h), int(device)), "az_gpu_create")

    def upload_scene(self, scene, eps_intersection=0.0, eps_shadow=0.000275):
        e = N.epsilon_values(float(eps_intersection), float(eps_shadow))
        N.check(self.lib.az_gpu_upload_scene(self.h, scene.ptr, C.byref(e)), "az_gpu_upload_scene", self.h)

    def alloc(self, nbytes):
        return DeviceArray(self, nbytes)

    def trace_closest_device(self, d_rays, n, d_hits, cull=False):
        N.check(self.lib.az_gpu_trace_closest(self.h, d_rays.ptr, int(n), d_hits.ptr, 1 if cull else 0),
                "az_gpu_trace_closest", self.h)

    def trace_any_device(self, d_rays, n, d_hits):
        N.check(self.lib.az_gpu_trace_any(self.h, d_rays.ptr, int(n), d_hits.ptr), "az_gpu_trace_any", self.h)

    def trace(self, rays, any_hit=False, cull=False):
        """Host convenience: uploads rays, traces, downloads hits."""
        rays = np.ascontiguousarray(rays, dtype=N.RAY_DTYPE)
        n = rays.shape[0]
        dr = self.alloc(max(1, rays.nbytes))
        dh = self.alloc(max(1, n * 16))
        dr.upload(rays)
        if any_hit:
            self.trace_any_device(dr, n, dh)
        else:
            self.trace_closest_device(dr, n, dh, cull)
        out = dh.download(N.HIT_DTYPE, n)
        dr.free()
        dh.free()
        return out

    def stats(self):
        s = N.az_stats()
        N.check(self.lib.az_gpu_get_stats(self.h, C.byref(s)), "az_gpu_get_stats", self.h)
        return s

    def render_stats(self):
        """Counts of the last render while counting is on (az_render_stats)."""
        r = N.az_render_stats()
        N.check(self.lib.az_gpu_get_render_stats(self.h, C.byref(r)), "az_gpu_get_render_stats", self.h)
        return {name: getattr(r, name) for name, _ in r._fields_}

    def set_counting(self, on):
        N.check(self.lib.az_gpu_set_counting(self.h, 1 if on else 0), "az_gpu_set_counting", self.h)

    def set_build(self, leaf=4, node_cost=1.0, item_cost=1.0):
        """BVH build parameters of the next upload_scene."""
        N.check(self.lib.az_gpu_set_build(self.h, int(leaf), float(node_cost), float(item_cost)),
                "az_gpu_set_build", self.h)

    def set_bvh_width(self, width=4):
        """Children per BVH node (2..4) of the next upload_scene."""
        N.check(self.lib.az_gpu_set_bvh_width(self.h, int(width)), "az_gpu_set_bvh_width", self.h)

    def autotune_build(self, scene, d_rays, n, repeats=3, eps_intersection=0.0, eps_shadow=0.000275):
        """Uploads the scene once per (node width, leaf size) candidate, times
        the current closest-hit launch on the device rays and keeps the
        fastest build; returns (entries, best index)."""
        e = N.epsilon_values(float(eps_intersection), float(eps_shadow))
        res = N.az_tune_result()
        N.check(self.lib.az_gpu_autotune_build(self.h, scene.ptr, C.byref(e), d_rays.ptr, int(n), int(repeats),
                                               C.byref(res)), "az_gpu_autotune_build", self.h)
        return self._tune_list(res)

    def set_launch(self, block=0, ring=0, waves=-1, loop=-1):
        N.check(self.lib.az_gpu_set_launch(self.h, int(block), int(ring), int(waves), int(loop)),
                "az_gpu_set_launch", self.h)

    def synchronize(self):
        N.check(self.lib.az_gpu_synchronize(self.h), "az_gpu_synchronize", self.h)

    def device_copy(self, d_dst, d_src, nbytes):
        N.check(self.lib.az_gpu_device_copy(self.h, d_dst.ptr, d_src.ptr, int(nbytes)), "az_gpu_device_copy", self.h)

    def record_launch_times(self, on=True):
        N.check(self.lib.az_gpu_record_launch_times(self.h, 1 if on else 0), "az_gpu_record_launch_times", self.h)

    def collect_launch_times(self, capacity=65536):
        buf = (C.c_double * capacity)()
        n = C.c_uint32()
        N.check(self.lib.az_gpu_collect_launch_times(self.h, buf, capacity, C.byref(n)),
                "az_gpu_collect_launch_times", self.h)
        return np.array(buf[:min(n.value, capacity)])

    def timer_start(self):
        N.check(self.lib.az_gpu_timer_start(self.h), "az_gpu_timer_start", self.h)

    def timer_stop(self):
        ms = C.c_double()
        N.check(self.lib.az_gpu_timer_stop(self.h, C.byref(ms)), "az_gpu_timer_stop", self.h)
        return ms.value

    def render(self, camera, filt=N.FilterType_Box, cull=False, seed=0, spp=0, row_begin=0, row_end=0,
               guide=False, guide_fraction=0.5, env_sampling=N.AZ_ENV_COSINE, guide_max_depth=0,
               guide_train_spp=0, guide_chunk_spp=0, scatter_sampling=N.AZ_SCATTER_REFERENCE, guide_steps_per_chunk=0,
               guide_target=N.AZ_GUIDE_ENV):
        o = N.az_render_opts()
        o.ScatterSampling = int(scatter_sampling)
        o.GuideTarget = int(guide_target)
        o.EnvSampling = int(env_sampling)
        o.GuideMaxDepth = int(guide_max_depth)
        o.GuideTrainSpp = int(guide_train_spp)
        o.GuideChunkSpp = int(guide_chunk_spp)
        o.GuideStepsPerChunk = int(guide_steps_per_chunk)
        o.Seed = int(seed)
        o.SampleCountOverride = int(spp)
        o.RowBegin, o.RowEnd = int(row_begin), int(row_end)
        o.GuideEnable = 1 if guide else 0
        o.GuideFraction = float(guide_fraction)
        W, H = camera.Resolution.x, camera.Resolution.y
        out = np.zeros(W * H, dtype=N.PIXEL_DTYPE)
        N.check(self.lib.az_gpu_render(self.h, C.byref(camera), int(filt), 1 if cull else 0, C.byref(o),
                                       out.ctypes.data), "az_gpu_render", self.h)
        return out

    @staticmethod
    def _tune_list(res):
        return [{"block": e.Block, "ring": e.LdsStackEntries, "waves": e.MinWaves, "chunk": e.ChunkSamples,
                 "loop": e.LoopForm, "width": e.BvhWidth, "leaf": e.MaxLeafItems, "ms": e.Milliseconds}
                for e in res.Entries[:res.Count]], int(res.Best)

    def autotune_trace(self, d_rays, n, repeats=3):
        """Times every traversal variant on the device rays and keeps the
        fastest; returns (entries, best index)."""
        res = N.az_tune_result()
        N.check(self.lib.az_gpu_autotune_trace(self.h, d_rays.ptr, int(n), int(repeats), C.byref(res)),
                "az_gpu_autotune_trace", self.h)
        return self._tune_list(res)

    def autotune_render(self, camera, filt=N.FilterType_Box, cull=False, seed=0):
        """Times a render (enough samples per pixel to fill the largest
        candidate chunk twice, at most the camera's count) for every (chunk
        size, traversal variant) and keeps the fastest; returns (entries,
        best index)."""
        o = N.az_render_opts()
        o.Seed = int(seed)
        res = N.az_tune_result()
        N.check(self.lib.az_gpu_autotune_render(self.h, C.byref(camera), int(filt), 1 if cull else 0, C.byref(o),
                                                C.byref(res)), "az_gpu_autotune_render", self.h)
        return self._tune_list(res)

    def set_render_tuning(self, chunk=0, block=0, waves=-1, loop=-1):
        N.check(self.lib.az_gpu_set_render_tuning(self.h, int(chunk), int(block), int(waves), int(loop)),
                "az_gpu_set_render_tuning", self.h)

    def output_pixels_device(self, d_pixels, width, height, d_out, tonemap=None, inv_gamma=1.0, srgb=False):
        """OutputPixelsToFile's colour path on the device: d_pixels holds
        width*height pixels, d_out receives 3 bytes per pixel."""
        N.check(self.lib.az_gpu_output_pixels(self.h, d_pixels.ptr, int(width), int(height),
                                              C.byref(tonemap) if tonemap is not None else None, float(inv_gamma),
                                              1 if srgb else 0, d_out.ptr), "az_gpu_output_pixels", self.h)

    def output_pixels(self, pixels, width, height, tonemap=None, inv_gamma=1.0, srgb=False):
        """Host convenience: uploads a PIXEL_DTYPE array, runs the output
        stage, returns (height, width, 3) uint8."""
        px = np.ascontiguousarray(pixels, dtype=N.PIXEL_DTYPE)
        n = int(width) * int(height)
        d_px, d_out = self.alloc(max(1, n) * 16), self.alloc(max(1, n) * 3)
        try:
            d_px.upload(px)
            self.output_pixels_device(d_px, width, height, d_out, tonemap, inv_gamma, srgb)
            return d_out.download(np.uint8, n * 3).reshape(int(height), int(width), 3)
        finally:
            d_px.free()
            d_out.free()

    def sum_f32_sequential(self, x):
        """The float32 sum of x in sequence order, bit for bit."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        d_x, d_s = self.alloc(max(4, x.nbytes)), self.alloc(4)
        try:
            d_x.upload(x)
            N.check(self.lib.az_gpu_sum_f32_sequential(self.h, d_x.ptr, int(x.size), d_s.ptr),
                    "az_gpu_sum_f32_sequential", self.h)
            return d_s.download(np.float32, 1)[0]
        finally:
            d_x.free()
            d_s.free()

    def set_build_mode(self, mode):
        """Where the next uploads build their binary BVHs (N.AZ_BUILD_*)."""
        N.check(self.lib.az_gpu_set_build_mode(self.h, int(mode)), "az_gpu_set_build_mode", self.h)

    def set_level_mode(self, mode):
        """How renders launch their bounce levels (N.AZ_LEVELS_*): replayed
        from a recorded graph per chunk, or one by one from the host."""
        N.check(self.lib.az_gpu_set_level_mode(self.h, int(mode)), "az_gpu_set_level_mode", self.h)

    def level_graph_counts(self):
        """(graphs recorded, graph launches) of this context's renders."""
        rec, lau = C.c_uint64(0), C.c_uint64(0)
        N.check(self.lib.az_gpu_level_graph_counts(self.h, C.byref(rec), C.byref(lau)),
                "az_gpu_level_graph_counts", self.h)
        return int(rec.value), int(lau.value)

    def set_render_memory_limit(self, nbytes):
        """Caps the render queues' device memory (0: no cap); frees them first."""
        N.check(self.lib.az_gpu_set_render_memory_limit(self.h, int(nbytes)), "az_gpu_set_render_memory_limit",
                self.h)

    def env_luminance_sample(self, u):
        """Directions and solid-angle densities of the environment luminance
        sampler at uniforms u (n, 2); returns (n, 4) float32."""
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        n = u.shape[0]
        d_u, d_o = self.alloc(max(8, u.nbytes)), self.alloc(max(16, n * 16))
        try:
            d_u.upload(u)
            N.check(self.lib.az_gpu_env_luminance_sample(self.h, d_u.ptr, n, d_o.ptr), "az_gpu_env_luminance_sample",
                    self.h)
            return d_o.download(np.float32, n * 4).reshape(n, 4)
        finally:
            d_u.free()
            d_o.free()

    def env_luminance_pdf(self, dirs):
        """Solid-angle density of the environment luminance sampler at unit
        directions dirs (n, 3)."""
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = d.shape[0]
        d_d, d_p = self.alloc(max(12, d.nbytes)), self.alloc(max(4, n * 4))
        try:
            d_d.upload(d)
            N.check(self.lib.az_gpu_env_luminance_pdf(self.h, d_d.ptr, n, d_p.ptr), "az_gpu_env_luminance_pdf", self.h)
            return d_p.download(np.float32, n)
        finally:
            d_d.free()
            d_p.free()

    def math_eval(self, name, x, y=None):
        """The device's float elementary function `name` (MATH_FUNCTIONS) of
        float32 arrays x (and y), elementwise."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
        n = x.size
        d_x, d_o = self.alloc(max(4, x.nbytes)), self.alloc(max(4, x.nbytes))
        d_y = None
        try:
            d_x.upload(x)
            if y is not None:
                y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
                assert y.size == n
                d_y = self.alloc(max(4, y.nbytes))
                d_y.upload(y)
            N.check(self.lib.az_gpu_math_eval(self.h, N.MATH_FUNCTIONS[name], d_x.ptr, d_y.ptr if d_y else None, n,
                                              d_o.ptr), "az_gpu_math_eval", self.h)
            return d_o.download(np.float32, n)
        finally:
            d_x.free()
            d_o.free()
            if d_y:
                d_y.free()

    def trim(self):
        """Frees the context's render queues (the scene stays uploaded)."""
        N.check(self.lib.az_gpu_trim(self.h), "az_gpu_trim", self.h)

    def render_f64(self, camera, filt=N.FilterType_Box, seed=0, spp=0, guide=False, guide_fraction=0.5,
                   env_sampling=N.AZ_ENV_COSINE, guide_max_depth=0, guide_train_spp=0, guide_chunk_spp=0,
                   guide_steps_per_chunk=0,
                   scatter_sampling=N.AZ_SCATTER_REFERENCE, guide_target=N.AZ_GUIDE_ENV):
        """Device render accumulated in double precision; returns (W*H, 4)
        float64 (colour sums, weight sums).  The wall time of the render call
        (guide training included) is left in last_render_s."""
        W, H = camera.Resolution.x, camera.Resolution.y
        d = self.alloc(W * H * 32)
        N.check(self.lib.az_gpu_memset(self.h, d.ptr, 0, W * H * 32), "az_gpu_memset", self.h)
        o = N.az_render_opts()
        o.Seed = int(seed)
        o.SampleCountOverride = int(spp)
        o.AccumulateF64 = 1
        o.ScatterSampling = int(scatter_sampling)
        o.GuideTarget = int(guide_target)
        o.EnvSampling = int(env_sampling)
        o.GuideMaxDepth = int(guide_max_depth)
        o.GuideTrainSpp = int(guide_train_spp)
        o.GuideChunkSpp = int(guide_chunk_spp)
        o.GuideStepsPerChunk = int(guide_steps_per_chunk)
        o.GuideEnable = 1 if guide else 0
        o.GuideFraction = float(guide_fraction)
        self.synchronize()
        t0 = time.perf_counter()
        N.check(self.lib.az_gpu_render_device(self.h, C.byref(camera), int(filt), 0, C.byref(o), d.ptr),
                "az_gpu_render_device", self.h)
        self.synchronize()
        self.last_render_s = time.perf_counter() - t0
        out = d.download(np.float64, W * H * 4).reshape(W * H, 4)
        d.free()
        return out

    @staticmethod
    def render_batch(contexts, cameras, filt=N.FilterType_Box, cull=False, seed=0, spp=0, f64=False):
        """Renders cameras[k] against the scene of contexts[k], all in one
        batched render on contexts[0]'s stream; returns one (W*H, 4) array per
        scene (float32 pixel records, or float64 sums with f64) and the wall
        time of the render call."""
        assert len(contexts) == len(cameras) and contexts
        lib = contexts[0].lib
        dt = np.float64 if f64 else np.float32
        bufs = []
        for ctx, cam in zip(contexts, cameras):
            nb = cam.Resolution.x * cam.Resolution.y * 4 * np.dtype(dt).itemsize
            d = contexts[0].alloc(nb)
            N.check(lib.az_gpu_memset(contexts[0].h, d.ptr, 0, nb), "az_gpu_memset", contexts[0].h)
            bufs.append(d)
        o = N.az_render_opts()
        o.Seed = int(seed)
        o.SampleCountOverride = int(spp)
        o.AccumulateF64 = 1 if f64 else 0
        hs = (C.c_void_p * len(contexts))(*[c.h.value for c in contexts])
        cs = (C.c_void_p * len(cameras))(*[C.addressof(cam) for cam in cameras])
        ps = (C.c_void_p * len(bufs))(*[b.ptr for b in bufs])
        contexts[0].synchronize()
        t0 = time.perf_counter()
        N.check(lib.az_gpu_render_batch(hs, cs, len(contexts), int(filt), 1 if cull else 0, C.byref(o), ps),
                "az_gpu_render_batch", contexts[0].h)
        contexts[0].synchronize()
        seconds = time.perf_counter() - t0
        out = []
        for b, cam in zip(bufs, cameras):
            out.append(b.download(dt, cam.Resolution.x * cam.Resolution.y * 4).reshape(-1, 4))
            b.free()
        return out, seconds

    def close(self):
        if self.h:
            self.lib.az_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

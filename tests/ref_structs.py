"""ctypes views of the reference's entity, mesh and scene structs
(include/azimuth_types.h, entity.h:1-102, scene.h:6-30) for tests that walk a
built scene: its entity tree, its meshes' triangle entities and textures."""
import ctypes as C

from azimuth_amd import native as N

f32, u32, s32 = C.c_float, C.c_uint32, C.c_int32


class aabb(C.Structure):
    _fields_ = [("Min", N.vec3), ("Max", N.vec3), ("Center", N.vec3)]


class entity(C.Structure):
    pass


class triangle(C.Structure):
    _fields_ = [("P0", N.vec3), ("P1", N.vec3), ("P2", N.vec3), ("N", N.vec3)]


class mesh(C.Structure):
    _fields_ = [("InvTransformationMatrix", N.mat4), ("TransformationMatrix", N.mat4), ("MotionVector", N.vec3),
                ("Radiance", N.vec3), ("BVH", C.POINTER(entity)), ("Triangles", C.POINTER(triangle)),
                ("TotalArea", f32), ("UVOffset", s32), ("IsEmitter", u32), ("TriangleCount", u32),
                ("TextureIndex", u32), ("MaterialIndex", u32), ("ID", u32), ("ShadingMode", s32)]


class _sphere(C.Structure):
    _fields_ = [("Center", N.vec3), ("Radius", f32)]


class _tri(C.Structure):
    _fields_ = [("P0", N.vec3), ("P1", N.vec3), ("P2", N.vec3), ("N0", N.vec3), ("N1", N.vec3), ("N2", N.vec3),
                ("FaceNormal", N.vec3), ("UV0", N.vec2), ("UV1", N.vec2), ("UV2", N.vec2), ("ShadingMode", s32)]


class _inst(C.Structure):
    _fields_ = [("Mesh", C.POINTER(mesh)), ("BaseMeshID", u32), ("ResetTransform", u32)]


class _node(C.Structure):
    _fields_ = [("Right", C.POINTER(entity)), ("Left", C.POINTER(entity))]


class _ptrs(C.Union):
    _fields_ = [("inst", _inst), ("node", _node)]


class _boxed(C.Structure):
    _fields_ = [("AABB", aabb), ("p", _ptrs)]


class _eunion(C.Union):
    _fields_ = [("sphere", _sphere), ("tri", _tri), ("boxed", _boxed)]


entity._fields_ = [("u", _eunion), ("InvTransformationMatrix", N.mat4), ("TransformationMatrix", N.mat4),
                   ("MotionVector", N.vec3), ("Radiance", N.vec3), ("IsEmitter", u32), ("TextureIndex", u32),
                   ("MaterialIndex", u32), ("ID", u32), ("UniqueEntityID", u32), ("Type", s32)]

EntityType_Sphere, EntityType_Triangle, EntityType_MeshInstance, EntityType_BVHNode = range(4)


class scene(C.Structure):
    _fields_ = [("AmbientLight", N.vec3), ("BackgroundColor", N.vec3), ("BackgroundTexture", N.texture),
                ("BVHRoot", C.POINTER(entity)), ("Meshes", C.POINTER(mesh)), ("Materials", C.POINTER(N.material)),
                ("Lights", C.POINTER(N.light)), ("Textures", C.POINTER(N.texture)), ("EnvironmentMap", N.texture),
                ("MeshCount", u32), ("MaterialCount", u32), ("LightCount", u32), ("TextureCount", u32),
                ("MaxRecursionDepth", u32), ("UseBackgroundTexture", u32), ("UseEnvironmentMap", u32),
                ("PathTrace", u32), ("ImportanceSampling", u32)]


assert C.sizeof(entity) == 288 and C.sizeof(mesh) == 200 and C.sizeof(scene) == 2744


def scene_of(built):
    """The `scene` behind a scene.Scene object."""
    return C.cast(built.ptr, C.POINTER(scene)).contents


def vec(v):
    return (v.x, v.y, v.z)

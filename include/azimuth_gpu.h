/*
 * azimuth_gpu.h -- C ABI of the MI355X (gfx950) render hot path.
 *
 * Reference citations (file:line) name the reference's src/ files without
 * their azimuth_ prefix: main.cpp is src/azimuth_main.cpp.
 *
 * The reference has no plugin or FFI layer: its render path is entered once
 * per camera through
 *     static b32 RenderScene(scene *Scene, camera *Camera,
 *                            epsilon_values *Epsilons, u32 ThreadCount,
 *                            b32 BackfaceCulling)        (main.cpp:1799-1901)
 * which spawns ThreadCount workers running ProduceImageInLinearSpace over
 * scanline bands (main.cpp:1256-1397) and then writes the image.  The entry
 * points below replace the body of that call: the caller keeps its scene,
 * camera and epsilon structures (azimuth_types.h), uploads the scene once and
 * receives the same `pixel {Color, TotalWeight}` buffer (main.cpp:45-48) that
 * the reference hands to OutputPixelsToFile.
 *
 * Conventions: every call returns an az_status and never aborts; the caller
 * owns all structures passed in; one host thread drives a context; all device
 * work of a context is ordered on the context's own HIP stream.  Pointers named
 * d_* are device pointers (from az_gpu_malloc or any HIP allocation); all
 * other pointers are host memory.
 */
#ifndef AZIMUTH_GPU_H
#define AZIMUTH_GPU_H

#include "azimuth_types.h"
#include "azimuth_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A ray of a batch query: the reference ray3 {Origin, Direction, T}
 * (ray.h:1-5) plus an upper distance bound used by the any-hit query.
 * Direction need not be unit length. 32 bytes. */
typedef struct {
    f32 Origin[3];
    f32 Time;           /* ray3.T: shutter time in [0,1) for motion blur */
    f32 Direction[3];
    f32 TMax;           /* any-hit: occluded iff a hit with t < TMax exists */
} az_ray;

/* Closest-hit record, 16 bytes.  Prim is the primitive id (below) or
 * AZ_MISS; t is the world distance along Direction as the reference reports
 * it (ray_hit.t, ray.h:7-13); Beta/Gamma are the barycentrics of P1 and P2
 * computed by the reference's Cramer's-rule test (main.cpp:154-189), 0 for
 * spheres.
 *
 * Primitive ids: the triangles of scene->Meshes[m] are numbered after those of
 * meshes 0..m-1, in the order of that mesh's triangle entity array (the order
 * its BVH leaves have, left to right); spheres follow all triangles as
 * triangle_count + UniqueEntityID. */
typedef struct {
    u32 Prim;
    f32 t;
    f32 Beta;
    f32 Gamma;
} az_hit;

#define AZ_MISS 0xFFFFFFFFu

/* Per-render options the reference fixes elsewhere. */
typedef struct {
    u64 Seed;               /* per-pixel PCG32 streams are seeded from (Seed, pixel index) */
    u32 SampleCountOverride;/* 0: camera->SampleCount; else used instead (still truncated to a square) */
    u32 RowBegin, RowEnd;   /* scanline range to render; RowEnd 0 = all rows */
    u32 AccumulateF64;      /* 1: accumulate in double (long reference renders) */
    u32 GuideEnable;        /* 1: draw directions from a learned guide mixed with the base
                               distribution (one-sample MIS; GuideTarget says which directions);
                               the guide starts from seeded weights and trains online */
    f32 GuideFraction;      /* probability of drawing the direction from the guide (0 = 0.5) */
    u32 GuideTrainSpp;      /* samples per pixel rendered while training (0: a quarter of the
                               render's, at least 4 and at most 32) */
    u32 GuideStepsPerChunk; /* Adam steps after each training chunk (0: spread 512 steps over the
                               training chunks) */
    u32 EnvSampling;        /* environment next-event directions: AZ_ENV_COSINE (the reference's
                               cosine distribution about the normal) or AZ_ENV_LUMINANCE (the
                               environment map's luminance x sin(theta) distribution; unbiased,
                               different random directions) */
    u32 GuideMaxDepth;      /* path depth up to which the guide is used (0 = every depth); deeper
                               vertices use the base distribution alone (still unbiased) */
    u32 GuideChunkSpp;      /* samples per pixel per training chunk (0: about 2M camera samples
                               per chunk, at most GuideTrainSpp / 8); the guide is updated after
                               each chunk */
    u32 ScatterSampling;    /* hemisphere scatter directions: AZ_SCATTER_REFERENCE (the reference's
                               cosine or uniform hemisphere) or AZ_SCATTER_BRDF (a one-sample mixture
                               of the cosine lobe and the material's specular lobe, weighted by the
                               mixture pdf; unbiased, different random directions) */
    u32 GuideTarget;        /* what the guide proposes (with GuideEnable): AZ_GUIDE_ENV (environment
                               next-event directions, mixed with EnvSampling's distribution),
                               AZ_GUIDE_SCATTER (hemisphere scatter directions, mixed with
                               ScatterSampling's distribution; trained on the radiance each scatter
                               ray's path collects) or AZ_GUIDE_BOTH (one network for each) */
} az_render_opts;

#define AZ_ENV_COSINE    0u
#define AZ_ENV_LUMINANCE 1u

#define AZ_SCATTER_REFERENCE 0u
#define AZ_SCATTER_BRDF      1u

#define AZ_GUIDE_ENV     0u
#define AZ_GUIDE_SCATTER 1u
#define AZ_GUIDE_BOTH    2u

typedef struct {
    f64 LastKernelMs;       /* device time of the last trace/render call's kernels */
    u64 NodesVisited;       /* summed over the last trace call (if counting enabled) */
    u64 TrianglesTested;
    u64 PrimitivesTested;
    u32 NodeBytes;          /* bytes fetched per node visit */
    u32 TriangleBytes;      /* bytes fetched per triangle test */
    u32 BvhWidth;
    u32 StackDepth;
    u32 BlockSize;
    u32 NodeCount;
    u32 LeafCount;
    u32 TriangleCount;
    u32 SphereCount;
    u32 InstanceCount;
} az_stats;

typedef struct az_gpu_ctx az_gpu_ctx;

/* Context on HIP device `DeviceOrdinal`; creates the context's stream. */
az_status az_gpu_create(az_gpu_ctx **Out, int DeviceOrdinal);
void      az_gpu_destroy(az_gpu_ctx *Ctx);
/* Text of the last error of this context (static storage, never NULL). */
const char *az_gpu_last_error(const az_gpu_ctx *Ctx);

/* Flattens the scene (entities, meshes, materials, lights, textures, env map)
 * into device arrays and builds the device BVH.  Replaces any earlier upload.
 * Reads Scene and Epsilons only during the call. */
az_status az_gpu_upload_scene(az_gpu_ctx *Ctx, const scene *Scene, const epsilon_values *Epsilons);

/* Batch closest hit for the reference query RayEntityIntersection(BVHRoot,
 * Cull, Epsilons->Intersection) (main.cpp:53-385), N rays. */
az_status az_gpu_trace_closest(az_gpu_ctx *Ctx, const az_ray *d_Rays, u32 N, az_hit *d_Hits,
                               b32 BackfaceCulling);
/* Batch occlusion in the form of the reference shadow query (cull off,
 * epsilon 0, main.cpp:697-712): Prim of d_Hits[i] is AZ_MISS when no
 * primitive is hit with t < TMax, else the id of one such primitive with its
 * t and barycentrics. */
az_status az_gpu_trace_any(az_gpu_ctx *Ctx, const az_ray *d_Rays, u32 N, az_hit *d_Hits);

/* The render call: fills Pixels (host, Width*Height entries, or the rows of
 * Opts->RowBegin..RowEnd) exactly as RenderScene's pixel buffer before
 * OutputPixelsToFile (main.cpp:1814-1880).  Opts may be NULL. */
az_status az_gpu_render(az_gpu_ctx *Ctx, const camera *Camera, filter_type Filter,
                        b32 BackfaceCulling, const az_render_opts *Opts, pixel *Pixels);

/* Device-resident variant: accumulates into d_Pixels (Width*Height pixel,
 * zeroed by the caller) without a copy back; d_PixelsF64 when AccumulateF64. */
az_status az_gpu_render_device(az_gpu_ctx *Ctx, const camera *Camera, filter_type Filter,
                               b32 BackfaceCulling, const az_render_opts *Opts, void *d_Pixels);

/* Batched render: Count scenes, one per context, rendered together.  Each
 * chunk's stages (camera rays; per path level the closest hits, the shading
 * and the shadow rays; the pixel splat) are single launches over the queues
 * of every scene, each workgroup working on one scene.  Ctxs[k] holds scene k
 * (az_gpu_upload_scene) and renders Cameras[k] into d_Pixels[k] (device
 * memory, Width*Height entries zeroed by the caller; doubles when
 * Opts->AccumulateF64).  All contexts must be on one device; the work is
 * ordered on Ctxs[0]'s stream and uses its render tuning, the scenes sharing
 * its chunk (at least one sample per pixel per scene).  Count <= 32.  The
 * learned guide is not available in batches (AZ_ERR_UNSUPPORTED).  No
 * reference counterpart: the reference renders one camera per call. */
az_status az_gpu_render_batch(az_gpu_ctx *const *Ctxs, const camera *const *Cameras, u32 Count, filter_type Filter,
                              b32 BackfaceCulling, const az_render_opts *Opts, void *const *d_Pixels);

/* The colour path of OutputPixelsToFile (main.cpp:1620-1797) for the 8-bit
 * formats (PNG, BMP, TGA, JPG), ending in their byte conversion
 * (main.cpp:1461-1477): each colour divided by its TotalWeight; with
 * Tonemap->Operator == TonemapOperator_Photographic the photographic
 * operator (log-average luminance summed in pixel order as the reference
 * does, white at the BurnoutPercentage percentile of the luminances,
 * Saturation); gamma (sRGB with exponent InvGamma when sRGB, else the power
 * InvGamma); d_Out[3 i + c] = (u8)(Clamp01(C) * 255).  Replaces the pixel
 * loops of OutputPixelsToFile; the image writer (stb) stays with the caller.
 * d_Pixels (Width*Height pixel, TotalWeight > 0 where tonemapped) and d_Out
 * (3 bytes per pixel) are device memory; d_Pixels is read only (the
 * reference normalizes it in place).  Tonemap may be NULL (no tonemapping);
 * a BurnoutPercentage outside [0, 100] returns AZ_ERR_INVALID_ARGUMENT.
 * Images of fewer than 2^31 pixels. */
az_status az_gpu_output_pixels(az_gpu_ctx *Ctx, const pixel *d_Pixels, u32 Width, u32 Height,
                               const tonemap *Tonemap, f32 InvGamma, b32 sRGB, u8 *d_Out);

/* The f32 sum of d_X[0..N) bit for bit as the loop `s = 0; for i: s += x[i]`
 * computes it in binary32 (the order OutputPixelsToFile sums its log
 * luminances in, main.cpp:1643-1661), into *d_Sum (device memory); runs in
 * parallel (chunks whose running sum stays in one binade are folded as
 * integer offsets of the mantissa grid).  N < 2^31.  No reference
 * counterpart as an entry point: the primitive behind az_gpu_output_pixels's
 * log-average. */
az_status az_gpu_sum_f32_sequential(az_gpu_ctx *Ctx, const f32 *d_X, u32 N, f32 *d_Sum);
/* az_gpu_sum_f32_sequential without any device allocation: the same kernels
 * and result, with their chunk workspace in a fixed 2.5 MiB array that the
 * library holds on each device it runs on until the process ends.  N <= 2^26
 * (larger N returns AZ_ERR_INVALID_ARGUMENT).  All calls on one device share
 * that workspace, so calls that overlap on one device (several streams or
 * host threads) race.  No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc). */
az_status az_gpu_sum_f32_sequential_noalloc(az_gpu_ctx *Ctx, const f32 *d_X, u32 N, f32 *d_Sum);

az_status az_gpu_get_stats(az_gpu_ctx *Ctx, az_stats *Stats);
/* Enables per-ray node/triangle counting in trace calls and renders
 * (slower; a counted render launches its levels from the host). */
az_status az_gpu_set_counting(az_gpu_ctx *Ctx, b32 Enable);

/* What the last render of a context did, filled while counting is enabled
 * (az_gpu_set_counting), zero otherwise: the work behind a render's
 * algorithmic bytes.  No reference counterpart. */
typedef struct {
    u64 CameraSamples;      /* camera rays generated */
    u64 PathRays;           /* rays the closest-hit stage traced, camera rays included */
    u64 ShadowRays;         /* shadow rays the connect stage traced */
    u64 ExtendNodes, ExtendTriangles, ExtendItems;      /* node visits, triangle tests, top-level item tests */
    u64 ConnectNodes, ConnectTriangles, ConnectItems;
    u64 Pixels;             /* pixels splatted */
    u32 Gaussian;           /* 1: the Gaussian filter */
    u32 Chunks;             /* chunks of camera samples */
} az_render_stats;
az_status az_gpu_get_render_stats(az_gpu_ctx *Ctx, az_render_stats *Stats);

/* Traversal launch parameters, as az_gpu_autotune_trace chooses them: workgroup
 * size (64, 128, 256), LDS stack entries per lane (8, 16, 32: the top of
 * each lane's stack; older entries spill to scratch), the waves per SIMD the
 * register allocation targets (0 = compiler's choice, 6) and the loop form
 * (0: one node or leaf per step; 1: while-while, all of a lane's node steps
 * before its next leaf).  0 (-1 for MinWaves and LoopForm) keeps the current
 * value; combinations that were not compiled, or whose stack (ring plus
 * spill) is too shallow for the uploaded BVH, return AZ_ERR_INVALID_ARGUMENT.
 * An upload whose BVH needs more than the current ring holds resets the
 * launch to the default. */
az_status az_gpu_set_launch(az_gpu_ctx *Ctx, u32 BlockSize, u32 LdsStackEntries, s32 MinWaves, s32 LoopForm);

/* BVH build parameters of the next az_gpu_upload_scene: the most primitives
 * a mesh BVH leaf holds (1..16, default 4) and the binned-SAH costs of
 * visiting a 4-wide node and of testing a primitive (default 1, 1). */
az_status az_gpu_set_build(az_gpu_ctx *Ctx, u32 MaxLeafItems, f32 NodeCost, f32 ItemCost);
/* Children per BVH node of the next upload, 2..4 (default 4): every node is
 * one 128-byte line with four child slots; narrower trees leave slots
 * empty and are deeper. */
az_status az_gpu_set_bvh_width(az_gpu_ctx *Ctx, u32 Width);
/* Where the binned-SAH binary tree of each BVH of the next uploads is built:
 * AZ_BUILD_AUTO (default: on the device for meshes and top levels of 65,536
 * items or more, on the host below), AZ_BUILD_HOST or AZ_BUILD_DEVICE.  Both
 * builders make the same tree (the device one repeats the host's float
 * operations and stable partitions), which the host then collapses into the
 * wide nodes.  Replaces ConstructBVH* (bvh.cpp:1-328) for the device BVH. */
az_status az_gpu_set_build_mode(az_gpu_ctx *Ctx, u32 Mode);
#define AZ_BUILD_AUTO   0u
#define AZ_BUILD_HOST   1u
#define AZ_BUILD_DEVICE 2u
/* How a render's bounce levels are launched (no counterpart in the
 * reference, whose ProduceImageInLinearSpace follows each path to its end
 * per sample, main.cpp:1256-1397):
 * AZ_LEVELS_AUTO (default) replays the levels of each chunk (extend, shade,
 * segment scan and connect per bounce) as one recorded hipGraph, one graph
 * per bounce count and chunk size, reused while the scene, camera, tuning
 * and buffers stay the same.  Each level is sized from per-level bounds
 * learned on a host-driven chunk of the same scene, camera and options (its
 * rays and shadow rays plus an eighth and 1,024), and the kernels read each
 * level's count on the device; a level past its bound makes the chunk run
 * again host-driven, which learns the bounds afresh.  Guided and batched
 * renders, and AZ_LEVELS_HOST, launch each level from the host after
 * reading back the previous level's count.  Both give the same image, bit
 * for bit.
 * az_gpu_level_graph_counts reports how many graphs were recorded and
 * launched on the context. */
az_status az_gpu_set_level_mode(az_gpu_ctx *Ctx, u32 Mode);
az_status az_gpu_level_graph_counts(az_gpu_ctx *Ctx, u64 *Recorded, u64 *Launched);
#define AZ_LEVELS_AUTO 0u
#define AZ_LEVELS_HOST 1u

/* Autotuning.  Each candidate launch configuration is timed with HIP events
 * on the context stream; the fastest becomes the context's setting (and
 * stays until the next tuning call or az_gpu_set_launch /
 * az_gpu_set_render_tuning).  Out may be NULL. */
typedef struct {
    u32 Block;            /* workgroup size of the traversal kernels */
    u32 LdsStackEntries;  /* LDS stack entries per lane (trace tuning) */
    s32 MinWaves;         /* register target in waves per SIMD (0: compiler's choice) */
    u32 ChunkSamples;     /* camera samples per wavefront chunk (render tuning) */
    u32 LoopForm;         /* 0: one node or leaf per step; 1: while-while */
    u32 BvhWidth;         /* children per BVH node (build tuning; else the current build) */
    u32 MaxLeafItems;     /* primitives per mesh BVH leaf (likewise) */
    u32 Reserved;
    f64 Milliseconds;     /* measured time of the candidate */
} az_tune_entry;

typedef struct {
    u32 Count;            /* entries filled */
    u32 Best;             /* index of the chosen entry */
    az_tune_entry Entries[32];
} az_tune_result;

/* BVH build: uploads Scene once per (node width 2, 3, 4) x (leaf size 2, 4,
 * 8) candidate, times the context's closest-hit launch on d_Rays (N rays,
 * Repeats launches after one warm-up) and leaves the scene uploaded with the
 * fastest build (its width and leaf size become the context's settings). */
az_status az_gpu_autotune_build(az_gpu_ctx *Ctx, const scene *Scene, const epsilon_values *Epsilons,
                                const az_ray *d_Rays, u32 N, u32 Repeats, az_tune_result *Out);

/* Traversal: times closest-hit batches of d_Rays (N rays, Repeats launches
 * per candidate after one warm-up) for every compiled (LDS stack entries,
 * workgroup size, register target, loop form) variant. */
az_status az_gpu_autotune_trace(az_gpu_ctx *Ctx, const az_ray *d_Rays, u32 N, u32 Repeats, az_tune_result *Out);

/* Renderer: times a render of Camera (Opts' rows and seed; guide off) at
 * enough samples per pixel to fill the largest candidate chunk twice (at
 * most the camera's count) for every (chunk size 8M, 16M, 32M samples,
 * traversal variant) candidate; the pixels written are discarded. */
az_status az_gpu_autotune_render(az_gpu_ctx *Ctx, const camera *Camera, filter_type Filter, b32 BackfaceCulling,
                                 const az_render_opts *Opts, az_tune_result *Out);
/* Sets the renderer's chunk size (camera samples; 0 keeps it; default 16M,
 * whose queues take 3 GB of device memory per child a shading point can emit
 * -- scatter, mirror, refraction, reflection -- plus 2 GB per light) and
 * its traversal workgroup size / register target / loop form (0, -1 and -1
 * keep them; default 128, 6, while-while). */
az_status az_gpu_set_render_tuning(az_gpu_ctx *Ctx, u32 ChunkSamples, u32 TraceBlock, s32 TraceMinWaves,
                                   s32 TraceLoopForm);

/* Waits for the context's work, then frees its render queues, pixel staging
 * and learned-guide buffers (the next render allocates them again); the
 * scene stays uploaded.  For callers that keep many contexts on one device.
 * No reference counterpart: the reference frees its pixel buffer at the end
 * of RenderScene (main.cpp:1899). */
az_status az_gpu_trim(az_gpu_ctx *Ctx);

/* Caps the device memory a context's render queues, pixel staging and
 * learned-guide buffers may hold (0: no cap, the default); frees them first
 * (as az_gpu_trim).  An allocation past the cap fails as it would on a full
 * device: a render halves its chunk of samples per pixel until its queues
 * fit (the image is the same up to the float rounding of the per-chunk pixel
 * sums) and returns AZ_ERR_OUT_OF_MEMORY when one sample per pixel does not
 * fit.  For callers that share a device with other work.  No reference
 * counterpart: the reference allocates one pixel buffer per render
 * (main.cpp:1808-1812). */
az_status az_gpu_set_render_memory_limit(az_gpu_ctx *Ctx, u64 Bytes);

/* The environment luminance sampler of AZ_ENV_LUMINANCE on caller inputs:
 * for i < N, the direction drawn at the uniforms d_U[2i], d_U[2i+1] in [0,1)
 * into d_Out[4i..4i+2] and its solid-angle density into d_Out[4i+3]
 * (az_gpu_env_luminance_sample), or the density of the unit direction
 * d_Dirs[3i..3i+2] into d_Pdf[i] (az_gpu_env_luminance_pdf).  Device
 * arrays.  AZ_ERR_UNSUPPORTED unless the uploaded scene has an image
 * environment map.  No reference counterpart: the reference draws
 * environment directions from the cosine distribution only
 * (main.cpp:724-778). */
az_status az_gpu_env_luminance_sample(az_gpu_ctx *Ctx, const f32 *d_U, u32 N, f32 *d_Out);
az_status az_gpu_env_luminance_pdf(az_gpu_ctx *Ctx, const f32 *d_Dirs, u32 N, f32 *d_Pdf);

/* The renderer's float elementary functions, which replace the reference's
 * Sin, Cos, Atan2, Power, Exp, ArcSine and ArcCosine (math.cpp:136-260):
 * each is the double-precision function of the float argument rounded to
 * float.  Out[i] = f(X[i]) or f(X[i], Y[i]) for i < N; Y is read only by
 * AZ_MATH_ATAN2 (X the y-coordinate) and AZ_MATH_POW.  az_gpu_math_eval runs
 * the device routines on device arrays, ordered on the context stream;
 * az_math_eval runs the same source on the host on host arrays. */
typedef enum az_math_fn {
    AZ_MATH_SIN = 0,
    AZ_MATH_COS = 1,
    AZ_MATH_ATAN2 = 2,
    AZ_MATH_POW = 3,
    AZ_MATH_EXP = 4,
    AZ_MATH_ASIN = 5,
    AZ_MATH_ACOS = 6,
    AZ_MATH_POW4 = 7,   /* X^4, the spot light falloff (main.cpp:665) */
    AZ_MATH_POW5 = 8,   /* X^5, the Fresnel terms (main.cpp:560, 896) */
} az_math_fn;
az_status az_gpu_math_eval(az_gpu_ctx *Ctx, u32 Fn, const f32 *d_X, const f32 *d_Y, u32 N, f32 *d_Out);
az_status az_math_eval(u32 Fn, const f32 *X, const f32 *Y, u32 N, f32 *Out);

/* Device memory helpers, ordered on the context stream. */
az_status az_gpu_malloc(az_gpu_ctx *Ctx, void **d_Ptr, u64 Bytes);
az_status az_gpu_free(az_gpu_ctx *Ctx, void *d_Ptr);
az_status az_gpu_upload(az_gpu_ctx *Ctx, void *d_Dst, const void *Src, u64 Bytes);
az_status az_gpu_download(az_gpu_ctx *Ctx, void *Dst, const void *d_Src, u64 Bytes);
az_status az_gpu_memset(az_gpu_ctx *Ctx, void *d_Dst, int Value, u64 Bytes);
az_status az_gpu_synchronize(az_gpu_ctx *Ctx);
/* Copies Bytes (a multiple of 16) with a plain streaming kernel on the
 * context stream: the reference point for HBM bandwidth and counters. */
az_status az_gpu_device_copy(az_gpu_ctx *Ctx, void *d_Dst, const void *d_Src, u64 Bytes);

/* Per-launch kernel timing: while enabled, every trace call records a HIP
 * event pair around its kernel on the context stream.  Collect waits for the
 * stream and returns (up to Capacity) the durations in milliseconds of the
 * launches recorded since the previous collect, in launch order. */
az_status az_gpu_record_launch_times(az_gpu_ctx *Ctx, b32 Enable);
az_status az_gpu_collect_launch_times(az_gpu_ctx *Ctx, f64 *Milliseconds, u32 Capacity, u32 *Count);

/* Event timing of an arbitrary sequence of calls on the context stream. */
az_status az_gpu_timer_start(az_gpu_ctx *Ctx);
az_status az_gpu_timer_stop(az_gpu_ctx *Ctx, f64 *Milliseconds);

/* RenderScene-shaped entry: uploads Scene, renders Camera with the box
 * filter (main.cpp:1847) and writes Pixels.  ThreadCount is accepted for
 * signature compatibility and ignored. */
b32 az_render_scene(scene *Scene, camera *Camera, epsilon_values *Epsilons, u32 ThreadCount,
                    b32 BackfaceCulling, pixel *Pixels);

#ifdef __cplusplus
}
#endif

#endif /* AZIMUTH_GPU_H */

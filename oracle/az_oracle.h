/*
 * az_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the azimuth render hot path, used by the tests, the
 * smoke check and the CPU-baseline leg of the benchmark to check the GPU
 * library.  Nothing in the product links or calls it.
 *
 * It follows the reference functions line by line (citations in
 * az_oracle.cpp) with three defined deviations, shared with the GPU path:
 *   - rand() is replaced by counter-addressed PCG32 streams (az_oracle_rng_*);
 *   - the closest-hit query additionally reports the primitive id and the
 *     barycentrics that the reference computes and discards;
 *   - Perlin hashing uses a proper modulo 256 and the gradient table is a
 *     fixed PCG32-generated table (the reference fills it from rand()).
 */
#ifndef AZ_ORACLE_H
#define AZ_ORACLE_H

#include "azimuth_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Closest hit through the reference BVH (no pruning, ties to the right
 * child) for N rays, on Threads host threads. */
void azo_trace_closest(const scene *Scene, const az_ray *Rays, u32 N, az_hit *Hits,
                       b32 Cull, f32 HitEpsilon, u32 Threads);
/* Shadow-ray form: closest hit with cull off and epsilon 0; reported only if
 * its t < TMax (else Prim = AZ_MISS). */
void azo_trace_any(const scene *Scene, const az_ray *Rays, u32 N, az_hit *Hits, u32 Threads);
/* Closest hit by testing every primitive (no BVH), same per-primitive
 * arithmetic; pins the traversal. */
void azo_brute_closest(const scene *Scene, const az_ray *Rays, u32 N, az_hit *Hits,
                       b32 Cull, f32 HitEpsilon, u32 Threads);

/* ProduceImageInLinearSpace over rows [Opts->RowBegin, RowEnd) on Threads
 * band threads (RenderScene's split), into Pixels (Width*Height, zeroed by
 * the caller; only the rendered rows and, for the Gaussian filter, their
 * 1-pixel neighbourhood are written). */
void azo_render(const scene *Scene, const camera *Camera, const epsilon_values *Eps,
                filter_type Filter, b32 Cull, const az_render_opts *Opts, pixel *Pixels,
                u32 Threads);
/* Same, accumulating Color and TotalWeight in double (4 doubles per pixel). */
void azo_render_f64(const scene *Scene, const camera *Camera, const epsilon_values *Eps,
                    filter_type Filter, b32 Cull, const az_render_opts *Opts, f64 *Pixels,
                    u32 Threads);

/* Threads threads (through the POSIX platform table) each adding Value
 * AddsPerThread times to one float with PlatformAtomicAddF32; returns the
 * total. */
f32 azo_platform_atomic_sum(u32 Threads, u32 AddsPerThread, f32 Value);

/* The float elementary functions of the renderer on arrays: Fn 0 Sin(X),
 * 1 Cos(X), 2 Atan2(X, Y) (X the y-coordinate), 3 Power(X, Y), 4 Exp(X),
 * 5 ArcSine(X), 6 ArcCosine(X), 7 Power(X, 4), 8 Power(X, 5); Y may be null
 * for the one-argument functions.  Returns 0 for an unknown Fn. */
int azo_math_eval(u32 Fn, const f32 *X, const f32 *Y, u32 N, f32 *Out);

/* Counter-addressed random stream: the float drawn at (Seed, Pixel, Sample,
 * Node, Dim), Dim counted within one path vertex. */
f32 azo_rng_draw(u64 Seed, u32 Pixel, u32 Sample, u32 Node, u32 Dim);
/* Raw PCG32 output after seeding with (InitState, InitSeq) and advancing by
 * Delta steps. */
u32 azo_pcg32_at(u64 InitState, u64 InitSeq, u64 Delta);

/* Perlin gradient table entry i (x,y,z into Out). */
void azo_perlin_gradient(u32 Index, f32 *Out);
f32  azo_perlin_sample(const texture *Texture, f32 X, f32 Y, f32 Z);

/* BRDF evaluation exactly as EvaluateBRDF, with Diffuse/Specular starting at
 * 1 as in CalculateLightContribution. */
void azo_evaluate_brdf(const brdf *Brdf, const f32 *LightDir, const f32 *HitN,
                       const f32 *ViewDir, f32 CosThetaIncident, f32 *Diffuse, f32 *Specular);

/* OutputPixelsToFile's colour path for the 8-bit formats: colour over
 * weight, the photographic operator when Tonemap->Operator is not None
 * (Tonemap may be NULL), gamma (sRGB or power InvGamma), then
 * (u8)(Clamp01(C) * 255) into Out (3 bytes per pixel).  Returns -1 (Out
 * untouched) when the burnout percentile indexes outside the image. */
int azo_output_pixels(const pixel *Pixels, u32 Width, u32 Height, const tonemap *Tonemap, f32 InvGamma,
                      b32 SRGB, u8 *Out);

#ifdef __cplusplus
}
#endif

#endif

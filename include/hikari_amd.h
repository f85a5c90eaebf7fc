/*
 * hikari_amd.h -- C ABI of the MI355X path-tracing core (libhikari_amd.so).
 *
 * Drop-in boundary for Hikari's pathMIS hot path.  Each entry point names the
 * reference interface it replaces.  Conventions:
 *   - every call returns HK_OK (0) or a negative HK_ERR_* code; the message of
 *     the last failure on a context is returned by hk_last_error();
 *   - host pointers are caller-owned and only read/written during the call;
 *     *_device variants take device pointers (HBM-resident inputs/outputs);
 *   - one context = one device + one HIP stream; calls on one context are
 *     serialised by the caller (the reference's Scene/Accel are read-only
 *     after Activate(), include/core/Integrator.hpp:35; the same holds here).
 *   - no CPU fallback: if no HIP device is usable, hk_ctx_create fails.
 */
#ifndef HIKARI_AMD_H
#define HIKARI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HK_OK 0
#define HK_ERR_INVALID (-1)
#define HK_ERR_HIP (-2)
#define HK_ERR_OOM (-3)
#define HK_ERR_UNSUPPORTED (-4)
#define HK_ERR_IO (-5)
#define HK_ERR_BUDGET (-6)
#define HK_ERR_CANCELLED (-7)

typedef struct hk_ctx hk_ctx;
typedef struct hk_accel hk_accel;
typedef struct hk_scene hk_scene;

/* ------------------------------------------------------------ context */
/* Opens `device` and creates the context's stream.  No reference counterpart
 * (the reference is single-process CPU code, src/Main.cpp:151-205). */
int hk_ctx_create(int device, hk_ctx **out);
void hk_ctx_destroy(hk_ctx *ctx);
/* Message of the last failed call on ctx (or of the last failed
 * hk_ctx_create when ctx is NULL). */
const char *hk_last_error(const hk_ctx *ctx);
const char *hk_version(void);
/* Synchronises the context stream. */
int hk_ctx_sync(hk_ctx *ctx);

/* ------------------------------------------------------------ acceleration */
/* One ray: origin, tmin, direction (need not be normalised), tmax. */
typedef struct hk_ray {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
} hk_ray;
/* Closest (or any) hit: t, barycentric u, v of Moller-Trumbore, and the
 * global triangle index (mesh-major, face-minor, as the reference's shape list,
 * src/core/Acceleration.cpp:13-28); prim = -1 on a miss. */
typedef struct hk_hit {
  float t, u, v;
  int32_t prim;
} hk_hit;

typedef struct hk_accel_info {
  uint32_t n_triangles, n_nodes, n_leaves, max_depth;
  uint64_t node_bytes, triangle_bytes;
  double build_ms;
} hk_accel_info;

/* Replaces Acceleration::AddMesh + Build (include/core/Acceleration.hpp:27-45,
 * src/acceleration/BVHAcceleration.cpp:46-248,
 * src/acceleration/HLBVHAcceleration.cpp:81-204): builds a 4-wide BVH over nt
 * triangles (idx: nt x 3 into v: nv x 3) in HBM, with the binned-SAH host
 * builder or, under option "bvh_builder" = 1, the device linear-BVH builder
 * (both give the same closest hits). */
int hk_accel_build(hk_ctx *ctx, const float *v, uint32_t nv, const uint32_t *idx, uint32_t nt, hk_accel **out);
void hk_accel_destroy(hk_accel *accel);
int hk_accel_get_info(const hk_accel *accel, hk_accel_info *info);

/* Replaces Acceleration::RayIntersect(ray, isect, bShadowRay)
 * (include/core/Acceleration.hpp:57, src/core/Acceleration.cpp:45-80) for a
 * batch of n rays.  Closest hit (any_hit = 0) returns the Moller-Trumbore hit
 * of the reference bit for bit; among equal t the largest prim wins (the
 * reference's brute-force order, where the last tested triangle wins).  Any
 * hit (any_hit = 1) reports some hit in [tmin, tmax] (prim >= 0) or a miss.
 * Host buffers. */
int hk_accel_intersect(hk_accel *accel, const hk_ray *rays, uint32_t n, hk_hit *hits, int any_hit);
/* Same on device buffers (d_rays, d_hits in HBM) on the context stream;
 * asynchronous.  Both take n <= 2^31 rays per call (HK_ERR_INVALID beyond) and
 * trace them on persistent waves tuned by the context's "refill" (closest
 * hit), "refill_shadow" (any hit) and "fetch_batch" options.  Calls on one
 * accelerator must not overlap (they share its fetch counter). */
int hk_accel_intersect_device(hk_accel *accel, const hk_ray *d_rays, uint32_t n, hk_hit *d_hits, int any_hit);

/* No counterpart in the reference.  hk_accel_intersect_device for callers
 * that cannot allocate device memory: code under HIP stream capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same rays, hits, options
 * and results (closest hits bit for bit); it allocates nothing: the
 * persistent kernels' fetch counter lives in a fixed 256-B scratch block of
 * the library, one per device, held until the process ends (its device
 * address is taken when the device's first context is created), and is
 * cleared by a kernel at every call (no memset node in a captured graph).
 * Every call on one device shares that block: calls that overlap on one
 * device, from any accelerator or context, race and must be ordered by the
 * caller.  At most HK_NOALLOC_MAX_RAYS rays per call (HK_ERR_INVALID beyond).
 * Asynchronous, on the accelerator's context stream. */
#define HK_NOALLOC_MAX_RAYS 2147483648u
int hk_accel_intersect_device_noalloc(hk_accel *accel, const hk_ray *d_rays, uint32_t n, hk_hit *d_hits,
                                      int any_hit);

/* ------------------------------------------------------------ BSDFs */
/* Material records: 32 floats each.  [0] type (0 diffuse, 1 conductor,
 * 2 dielectric, 3 plastic, 4 roughConductor, 5 roughDielectric, 6 roughPlastic,
 * 7 coating, 8 roughCoating, 9 mirror, 10 microfacet, 11 roughDiffuse,
 * 12 twoSided), [1] distribution (0 beckmann, 1 ggx), [2] nested material
 * index (coatings; a twoSided's upper side), [3] flags (1 nonlinear,
 * 2 anisotropic, 4 fastApprox), [4..6] colour A, [8..10] colour B, [12] a
 * twoSided's lower-side material index (-1: same as the upper), [20] intIOR,
 * [21] extIOR, [22] alpha or alphaU, [23] alphaV, [24] thickness.  Colour A/B:
 * diffuse albedo/-; (rough)conductor ks/k; (rough)dielectric
 * ksReflect/ksRefract; (rough)plastic ks/kd; (rough)coating sigmaA/ks;
 * microfacet kd/-; roughDiffuse albedo/-.  The XML properties of the src/bsdf
 * classes.  Coatings nest one non-coating, non-twoSided BSDF; a twoSided nests
 * BSDFs without transmission (TwoSidedBSDF.cpp:106-122). */
#define HK_MATERIAL_STRIDE 32

/* Batched BSDF::Eval / Pdf / Sample (include/core/BSDF.hpp:67-94) of
 * material `index` on the GPU, for parity tests.  wi, wo: n x 3 (local frame);
 * measure: n (1 solid angle, 2 discrete); u: n x 4 uniforms (u[0], u[1] the
 * Sample() point, u[2], u[3] the extra draws Sample() makes from the sampler).
 * Outputs (n x 3 / n): eval, pdf, and for Sample(): f, wo, measure, eta, and
 * Pdf() of the sampled record.  rft_dir: directory holding
 * BeckmannRFTData.bin / GGXRFTData.bin (NULL: the library's data dir). */
int hk_bsdf_batch(hk_ctx *ctx, const float *materials, uint32_t n_materials, uint32_t index, const char *rft_dir,
                  uint32_t n, const float *wi, const float *wo, const int32_t *measure, const float *u,
                  float *out_eval, float *out_pdf, float *out_f, float *out_wo, int32_t *out_measure,
                  float *out_eta, float *out_spdf);

/* ------------------------------------------------------------ scenes */
typedef struct hk_mesh_desc {
  const float *v; uint32_t nv;       /* nv x 3 positions (world space) */
  const float *n; const float *uv;   /* nullable: nv x 3 normals, nv x 2 uv */
  const uint32_t *f; uint32_t nt;    /* nt x 3 vertex indices */
  int32_t material;                  /* index into the material table */
  int32_t is_emitter; float radiance[3]; /* area light (src/emitter/AreaLight.cpp) */
} hk_mesh_desc;

typedef struct hk_scene_desc {
  const hk_mesh_desc *meshes; int32_t n_meshes;
  const float *materials; int32_t n_materials;
  const float *env; int32_t env_w, env_h;      /* nullable lat-long RGB32F map, rows top to bottom */
  float env_scale; float env_to_world[16];      /* row-major 4x4 */
  float cam_to_world[16];                       /* row-major 4x4 */
  float fov, near_clip, far_clip; int32_t width, height;
  int32_t filter;                               /* 0 box, 1 gaussian, 2 tent, 3 mitchell */
  float filter_radius, filter_stddev, filter_b, filter_c;
  int32_t max_depth; float background[3]; int32_t force_background;
  const char *rft_dir;                          /* nullable */
  int32_t integrator;                           /* 0 pathMIS, 1 pathEMS, 2 pathMATS, 3 whitted */
  int32_t rr_start_depth;                       /* Russian roulette from this path depth on (0: from the
                                                   first vertex, as the reference) */
} hk_scene_desc;

/* Integrators (estimators of the bounce kernel): pathMIS
 * (src/integrator/PathMISIntegrator.cpp:16-157), pathEMS (emitter sampling
 * only, PathEMSIntegrator.cpp:16-119), pathMATS (BSDF sampling only,
 * PathMATSIntegrator.cpp:16-87) and whitted (WhittedIntegrator.cpp:26-116).
 * max_depth: negative means unlimited (the reference reads it as uint32);
 * clamped to 262143.  rr_start_depth: no counterpart in the reference, whose
 * estimators play Russian roulette (continue with probability 0.95, weight
 * 1 / 0.95) at every vertex; vertices of depth below it continue with weight
 * 1 instead (the roulette's random number is drawn all the same).  Unbiased
 * either way; 0 (the default) is the reference's schedule, bit for bit.
 *
 * Replaces Scene::Activate (src/core/Scene.cpp:128-165) plus the Activate()
 * of every mesh / BSDF / emitter / camera: uploads geometry, BVH, material
 * table (with reduced rough-transmittance tables), emitter CDFs and the
 * environment map to HBM.  Emitters are ordered area lights in mesh order,
 * then the environment. */
int hk_scene_create(hk_ctx *ctx, const hk_scene_desc *desc, hk_scene **out);
void hk_scene_destroy(hk_scene *scene);
/* Bytes of HBM the scene holds. */
uint64_t hk_scene_device_bytes(const hk_scene *scene);

/* ------------------------------------------------------------ rendering */
/* One render request: camera pose and film of the request; spp samples per
 * pixel numbered [spp_offset, spp_offset + spp); `seed` keys the per-path
 * pcg32 streams together with pixel and sample index. */
typedef struct hk_request {
  float cam_to_world[16]; float fov; int32_t width, height;
  uint32_t spp, spp_offset, seed;
  int32_t use_scene_camera;                     /* 1: take pose/fov/size from the scene */
} hk_request;

typedef struct hk_render_stats {
  uint64_t paths, extension_rays, shadow_rays, bounces_launched;
  double gpu_ms;          /* device time of the render (HIP events) */
  double trace_ms;        /* device time inside the closest-hit kernel */
  uint64_t trace_launches;
  uint64_t node_visits, triangle_tests;  /* BVH work; non-zero only in builds with HK_TRAVERSAL_STATS */
  uint64_t graph_bounces;                /* bounces replayed from a captured HIP graph */
  uint64_t invalid_samples;              /* samples the film dropped (NaN / inf / negative radiance,
                                            src/core/Block.cpp:97-101) */
} hk_render_stats;

/* Whole-image render with the gpuPathMIS integrator (the wavefront form of
 * PathMISIntegrator::Li, src/integrator/PathMISIntegrator.cpp:16-157, driven
 * as src/Main.cpp:11-147).  out_rgb: host, height x width x 3, the filtered
 * image (ImageBlock::ToBitmap).  Synchronous.  The bounce loop is driven from
 * the host: every bounce enqueues k_regen / k_extend / k_shade / k_shadow
 * (replayed from a captured HIP graph after the first bounces) and the host
 * reads the queue sizes between bounces; it is not one persistent launch. */
int hk_render(hk_scene *scene, const hk_request *req, float *out_rgb, hk_render_stats *stats);

/* ------------------------------------------------------------ autotuning */
/* No counterpart in the reference.  Picks the context's wavefront size
 * ("max_paths"), refill thresholds and fetch batch of the persistent
 * traversal kernels ("refill", "refill_shadow", "fetch_batch"), camera-ray
 * tile edge ("tile") and queue-kernel grid cap ("grid") for requests like `req` on `scene`: each candidate renders a
 * calibration slice of req (about calib_samples path samples, 0: 64 M; at
 * least one sample per pixel), timed with HIP events; the fastest setting is
 * stored in the context options and reported.  No film depends on these
 * options (the same bits for any setting), so tuning changes only speed. */
#define HK_TUNE_MAX 32
typedef struct hk_tune_result {
  uint32_t max_paths, grid, tile, refill, refill_shadow, fetch_batch;   /* the chosen setting */
  uint32_t n_tried;
  uint32_t tried_max_paths[HK_TUNE_MAX], tried_grid[HK_TUNE_MAX], tried_tile[HK_TUNE_MAX];
  uint32_t tried_refill[HK_TUNE_MAX], tried_refill_shadow[HK_TUNE_MAX], tried_fetch_batch[HK_TUNE_MAX];
  double samples_per_s[HK_TUNE_MAX];
  double total_ms;                  /* wall time of the search */
} hk_tune_result;
int hk_autotune(hk_scene *scene, const hk_request *req, uint64_t calib_samples, hk_tune_result *out);

/* Per-sample radiance of one pixel for samples [s0, s0 + n) (the t-test
 * harness of src/core/StudentTTest.cpp:137-175): out n x 5 (r, g, b, px, py). */
int hk_render_pixel_samples(hk_scene *scene, uint32_t seed, uint32_t x, uint32_t y, uint32_t s0, uint32_t n,
                            float *out);

/* Batched renders: n requests over one resident scene share one wavefront
 * (paths tagged by request).  out_rgb[i]: host buffer of request i
 * (height x width x 3).  Results equal the requests rendered alone, bit for
 * bit.  max_paths: paths in flight (0: the context's).  Host-driven bounce
 * loop, as hk_render. */
int hk_render_batch(hk_scene *scene, const hk_request *reqs, uint32_t n, float **out_rgb, uint32_t max_paths,
                    hk_render_stats *stats);

/* ---------------------------------------------------------------- serving
 * A request loop with continuous batching (no counterpart in the reference,
 * whose Render() handles one scene file per process).  Each resident scene
 * has two lanes, each a wavefront on its own stream: a bulk lane (64M paths,
 * or max_paths when that option is set; a sample buffer of 4 x max_samples;
 * first come first
 * served; the scenes' bulk lanes, which never run at once, share one
 * workspace) and an express lane (4M paths, or max_paths / 16 when that
 * option is set; max_samples / 8
 * samples) for requests of at most max_samples / 16 path samples, fewest
 * samples left first.  A worker thread advances every lane as soon as its
 * previous bounce has finished on the device.  Bulk requests run one at a
 * time in arrival order over all scenes (a request's Russian-roulette tail
 * would otherwise be stretched by the next one's full bounces), while small
 * requests pass them on the express lanes.  A request that arrives while
 * others run is admitted, first come first served, once its film fits the
 * HBM budget; an express request's paths join its lane's next bounce.
 * hbm_budget bounds the server's path state and sample buffers
 * (the shared bulk workspace and each scene's express lane) plus the films
 * of admitted requests; 0 =
 * unbounded.  Each result equals hk_render of the same request on a context
 * with the same options, bit for bit.  Scenes must stay alive until
 * hk_server_destroy. */
typedef struct hk_server hk_server;
typedef struct hk_server_result {
  int32_t status;          /* HK_OK or the request's error */
  double queue_ms;         /* submit -> first samples admitted */
  double latency_ms;       /* submit -> film in host memory */
  uint64_t paths;          /* path samples of the request */
  double submit_ms;        /* submit time on the server clock (ms since hk_server_create) */
  int32_t lane;            /* 0 bulk, 1 express */
  uint64_t invalid_samples; /* samples the film dropped (NaN / inf / negative radiance) */
} hk_server_result;
typedef struct hk_server_stats {
  uint64_t submitted, completed, bounces, paths;
  uint64_t fixed_bytes, film_bytes;   /* HBM held: workspaces, admitted films */
  uint64_t shared_bounces;            /* bounces launched while two or more requests were in flight */
  uint64_t max_in_flight;             /* most requests in flight at once */
  uint64_t cancelled;                 /* requests ended by hk_server_cancel */
  uint64_t requests_in_flight;        /* requests holding a lane's request slot now */
} hk_server_stats;

int hk_server_create(hk_ctx *ctx, uint64_t hbm_budget, hk_server **out);
/* before the first submit; scene_id receives the scene's index */
int hk_server_add_scene(hk_server *server, hk_scene *scene, int32_t *scene_id);
/* out_rgb (height x width x 3) must stay valid until hk_server_wait returns */
int hk_server_submit(hk_server *server, int32_t scene_id, const hk_request *req, float *out_rgb, uint64_t *ticket);
/* blocks until the request is done; its status is also the return value */
int hk_server_wait(hk_server *server, uint64_t ticket, hk_server_result *result);
/* No counterpart in the reference (one render per process there).  Cancels a
 * submitted request that has not been waited for: one still queued is dropped
 * at once; one in flight starts no further samples, and its chunk slots,
 * sample-buffer ranges, film memory and request slot are reclaimed as soon as
 * the paths it already started have finished.  Its out_rgb is left as it was
 * and hk_server_wait returns HK_ERR_CANCELLED.  A request that finished before
 * the call keeps its result.  HK_ERR_INVALID for an unknown ticket. */
int hk_server_cancel(hk_server *server, uint64_t ticket);
int hk_server_get_stats(hk_server *server, hk_server_stats *stats);
/* finishes every submitted request, then frees the server */
int hk_server_destroy(hk_server *server);

/* Tuning knobs (value 0 restores the default):
 *   "max_paths"   paths in flight (default 256M for renders and batches; a
 *                 server's bulk wavefront 64M and express lanes 4M each,
 *                 or this value and 1/16 of it when set); at every bounce
 *                 finished paths are replaced by the next samples
 *   "max_samples" samples per segment (default 256M, 16 B of HBM each): a
 *                 request's samples are cut into chunks of
 *                 floor(max_samples / pixels) samples per pixel, resolved
 *                 into the film when their segment ends
 *   "grid"        cap on blocks per queue-kernel launch (0: one lane per item)
 *   "refill"      idle lanes (1..64, default 40) that make a wave of the
 *                 persistent closest-hit kernel fetch new rays
 *   "refill_shadow" the same for the shadow-ray kernel (default 16)
 *   "fetch_batch" queue entries (in 64s, 1..64, default 4) a wave of the
 *                 persistent kernels reserves per atomic on the shared fetch
 *                 counter
 *   "schedule"    order in which chunks are admitted and their samples
 *                 started: 1 (default) requests with the fewest samples
 *                 left first (a chunk that does not fit is passed over for
 *                 smaller ones); 0 first come first served.  Films do not
 *                 depend on it
 *   "express"     read by a server when a scene is added: 1 (default) requests of at
 *                 most max_samples / 16 samples take an express lane beside
 *                 the bulk lane (low median latency); 0: every request waits
 *                 its turn on its scene's bulk lane, first come first served
 *                 (no interference with large requests: lowest tail latency)
 *   "graph"       1 (default): after a few eager bounces a render replays
 *                 each bounce as a captured HIP graph; 0: eager launches
 *                 (profiled renders always launch eagerly)
 *   "tile"        edge of the pixel tiles camera rays start in (1, 2, 4, 8,
 *                 16 or 32; 1: scanlines; default 8)
 *   "profile"     1: time the closest-hit kernel with HIP events
 *   "resolve_tiles" 0: resolve films with the plain per-pixel gather instead
 *                 of LDS tiles (same bits; default 1)
 *   "bvh_builder" 0: binned-SAH build on the host (default); 1: linear BVH
 *                 built on the device (Morton order, SAH leaf collapse, a
 *                 bucketed SAH tree over the Morton cells), for callers that
 *                 add scenes while rendering.  Both give the same closest hits
 *                 bit for bit.
 *   "bvh_cell_bits" device builder: Morton bits of the cells under its SAH
 *                 upper tree, 0..30 (0: the radix tree all the way up;
 *                 default 24; the reference's hlbvh uses 12).  More bits
 *                 trace faster and build slower; hits do not change. */
int hk_set_option(hk_ctx *ctx, const char *name, int64_t value);

#ifdef __cplusplus
}
#endif
#endif /* HIKARI_AMD_H */

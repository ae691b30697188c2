/*
 * keepsake_amd: MI355X (gfx950) implementation of keepsake's small_pt path tracer.
 *
 * C ABI of the GPU library libkeepsake_amd.so. Plain pointers and sizes only; every function returns
 * an int status (KS_OK = 0, negative = error class) and never throws or aborts across the ABI.
 * ks_last_error() returns a thread-local message for the last failing call on the calling thread.
 *
 * The reference has no FFI for this path; its "operator API" is the C++ task/type surface listed next
 * to each entry point below. The C++ wrapper (include/ks/small_pt.h) restores the reference's C++
 * names and its failure behaviour (a non-zero status aborts, as the reference's ASSERT does).
 *
 * Threading: one host thread drives one HIP stream per device handle; calls on distinct handles are
 * thread-safe, calls on one handle are not re-entrant.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KS_OK 0
#define KS_ERR_INVALID (-1)     /* bad argument (the reference would ASSERT or std::abort) */
#define KS_ERR_HIP (-2)         /* HIP runtime failure */
#define KS_ERR_OOM (-3)         /* device or host allocation failed */
#define KS_ERR_UNSUPPORTED (-4) /* feature outside this build's scope */
#define KS_ERR_NO_DEVICE (-5)   /* no HIP device / library loaded without a GPU */

typedef struct ks_device ks_device;
typedef struct ks_scene ks_scene;
typedef struct ks_film ks_film;
typedef struct ks_host_scene ks_host_scene;

/* Closure inputs of PrincipledBSDF (principled_bsdf.h:78-93, fields evaluated by
 * PrincipledBSDF::eval_closure principled_bsdf.cpp:246-275) plus Material::emission (material.h:45).
 * All fields are constant per material (ConstantField, shader_field.h:37-60). */
typedef struct ks_material_desc {
    float basecolor[3];
    float roughness;
    float metallic;
    float ior;
    float specular_r0_mul; /* field value in [0,1]; the closure doubles it */
    float specular_trans;
    float diffuse_trans;
    float diffuse_trans_fwd;
    float emission[3];
    int32_t has_emission;
} ks_material_desc;

#define KS_LIGHT_SKY_AMBIENT 0 /* SkyLight(const color3 &ambient) light.cpp:52-57 */
#define KS_LIGHT_DIRECTIONAL 1 /* DirectionalLight light.h:91-105 */
#define KS_LIGHT_POINT 2       /* PointLight light.h:107-120 */
#define KS_LIGHT_SKY_MAP 3     /* SkyLight(map, l2w, transform_y_up, strength) light.cpp:19-49: the scene's sky_map;
                                  color scales nothing (the map's strength does) */

/* Non-mesh lights as created by create_light/create_sky_light (light.cpp:237-265). */
typedef struct ks_light_desc {
    int32_t type;
    float color[3]; /* ambient radiance, L, or I */
    float vec[3];   /* direction (directional) or position (point) */
} ks_light_desc;

/* An equirectangular HDR sky (SkyLight from a .hdr / .exr map, light.cpp:19-49; create_sky_light
 * light.cpp:258-272). The library builds the thresholded luminance distribution itself. */
typedef struct ks_sky_map_desc {
    const float *rgb;        /* width * height * 3 linear RGB, row 0 at theta = 0 (the map's up pole) */
    int32_t width, height;
    float to_world[9];       /* row-major 3x3 direction part of the map's l2w transform */
    int32_t transform_y_up;  /* 1: the map's up axis is +y (to_spherical_yup), 0: +z (to_spherical) */
    float strength;          /* radiance scale */
} ks_sky_map_desc;

/* A texture (Texture, texture.h:46-66; create_texture texture.cpp:1068-1084): texels row by row, each
 * `channels` values of `data_type`, levels back to back (each level max(1, (n + 1) / 2) of the one
 * before in both directions, as the reference's NPOT chain). Values are stored as given and converted
 * on fetch: u8 / 255, u16 / 65535, f32 as is, then sRGB -> linear when color_space is sRGB
 * (Texture::fetch_as_float, texture.cpp:403-460). */
#define KS_TEX_U8 0
#define KS_TEX_U16 1
#define KS_TEX_F32 2
#define KS_COLORSPACE_LINEAR 0
#define KS_COLORSPACE_SRGB 1
typedef struct ks_texture_desc {
    const void *data;
    int32_t width, height;
    int32_t channels;      /* 1..4 */
    int32_t data_type;     /* KS_TEX_* */
    int32_t color_space;   /* KS_COLORSPACE_* */
    int32_t levels;        /* levels present in data (1 .. 1 + ceil(log2(max(width, height)))) */
    int32_t build_mipmaps; /* with levels == 1: build the whole chain as Texture(bytes, ..., true) does */
} ks_texture_desc;

/* A shader field (ShaderField<color<N>>, shader_field.h:11-86): a constant, or a texture lookup through
 * a sampler (create_texture_sampler texture.cpp:1086-1121) with the TextureField's flip_v, uv_scale,
 * uv_offset, swizzle and scale. */
#define KS_SAMPLER_NEAREST 0
#define KS_SAMPLER_LINEAR 1 /* bilinear, trilinear across levels by the primary ray's uv footprint */
#define KS_SAMPLER_CUBIC 2
#define KS_SAMPLER_EWA 3
#define KS_WRAP_REPEAT 0
#define KS_WRAP_CLAMP 1
#define KS_CUBIC_MITCHELL 0
#define KS_CUBIC_BSPLINE 1
#define KS_CUBIC_CATMULL_ROM 2
typedef struct ks_field_desc {
    int32_t texture;       /* index into ks_scene_desc.textures, or -1: the constant `value` */
    int32_t sampler;       /* KS_SAMPLER_* */
    int32_t cubic_kernel;  /* KS_CUBIC_* */
    float anisotropy;      /* EWA, clamped to [1, 16] (EWASampler ctor, texture.h:127) */
    int32_t wrap_u, wrap_v; /* KS_WRAP_* */
    int32_t flip_v;
    int32_t swizzle[4];    /* texture channel of each component, -1: 0 */
    float uv_scale[2], uv_offset[2];
    float scale[4];
    float value[4];        /* the constant (texture == -1) */
} ks_field_desc;

/* A material whose closure inputs come from shader fields (PrincipledBSDF::eval_closure,
 * principled_bsdf.cpp:246-275, with Material::normal_map / opacity_map / emission, material.h:39-43).
 * The clamps of eval_closure apply to the field values at each hit. */
#define KS_NORMAL_MAP_TANGENT 0
#define KS_NORMAL_MAP_WORLD 1
typedef struct ks_material_fields {
    ks_field_desc basecolor, roughness, metallic, ior, specular_r0_mul, specular_trans, diffuse_trans,
        diffuse_trans_fwd;
    int32_t has_emission; /* the emission field makes the geometry's triangles mesh lights */
    ks_field_desc emission;
    int32_t has_normal_map; /* NormalMap (normal_map.h:11-30, normal_map.cpp:8-43) */
    ks_field_desc normal_map;
    int32_t normal_space;  /* KS_NORMAL_MAP_* */
    int32_t normal_range01;
    float normal_strength;
    float normal_to_world[9]; /* world-space maps: row-major linear part of the map's to_world (normals go
                                 through its inverse transpose, Transform::normal maths.h:987) */
    int32_t has_opacity_map; /* OpacityMap (opacity_map.h:11-21): stochastic alpha test in traversal */
    ks_field_desc opacity;
} ks_material_fields;

/* Flat scene: the geometry Scene::intersect1/occlude1 see (scene.h:49-87) with an identity instance
 * transform and one material per geometry (SubScene::materials, scene.h:37). Triangles of one geometry
 * are contiguous and geometries appear in increasing id order (mesh-light order, scene.cpp:315-336).
 * All meshes are two-sided (MeshData::twosided default, geometry.h:84). */
typedef struct ks_scene_desc {
    const float *vertices; /* 3 * n_vertices */
    uint32_t n_vertices;
    const float *normals; /* optional smooth-shading normals, 3 * n_vertices, or NULL */
    const uint32_t *indices; /* 3 * n_tris */
    uint32_t n_tris;
    const uint32_t *tri_geom; /* geometry id of each triangle */
    const uint32_t *geom_material; /* material of each geometry */
    uint32_t n_geoms;
    const ks_material_desc *materials;
    uint32_t n_materials;
    const ks_light_desc *lights;
    uint32_t n_lights;
    const uint8_t *geom_flags; /* per geometry (KS_GEOM_*), or NULL: two-sided, smooth normals when given */
    const ks_sky_map_desc *sky_map; /* the map of the (one) KS_LIGHT_SKY_MAP light, or NULL */
    const float *texcoords; /* 2 * n_vertices texture coordinates, or NULL: no geometry has them */
    const ks_texture_desc *textures;
    uint32_t n_textures;
    /* n_materials entries: a textured material replaces materials[m], NULL keeps it; or NULL */
    const ks_material_fields *const *material_fields;
} ks_scene_desc;

#define KS_GEOM_ONE_SIDED 1u /* MeshData::twosided = false: the frame is not flipped towards wo (geometry.cpp:126) */
#define KS_GEOM_FLAT 2u      /* MeshData::use_smooth_normal = false: vertex normals ignored (geometry.cpp:135) */
#define KS_GEOM_NO_TEXCOORD 4u /* the geometry has no texture coordinates: uv is the barycentric (u, v) and dpdu
                                  the edge v1 - v0 (geometry.cpp:75-116) even when ks_scene_desc.texcoords is set */

/* SBVHBuildOption (cuda/sbvh.cuh:9-16). Zeroed fields take the reference defaults
 * (maxPrimsInNode 4, bucketCount 12, travWeight 1, alpha 1e-5). */
typedef struct ks_bvh_options {
    uint32_t max_prims_in_node;
    uint32_t bucket_count;
    float trav_weight;
    float alpha;
    int32_t threads; /* host build threads, 0 = all */
    int32_t branching; /* node width on the device: 4 (default; the SBVH collapsed to 128-byte nodes) or 2 */
    int32_t builder;   /* KS_BVH_BUILD_* (default HOST_SBVH) */
} ks_bvh_options;

/* BVH builders. HOST_SBVH: binned SAH with spatial splits on host threads (the reference's SBVH,
 * cuda/sbvh.cu:301-440). DEVICE: built by kernels on the scene's GPU, clustering triangles in Morton
 * order (PLOC) and collapsing subtrees into leaves by SAH (no reference counterpart; the reference has
 * no device BVH build). */
#define KS_BVH_BUILD_HOST_SBVH 0
#define KS_BVH_BUILD_DEVICE 1

/* create_camera (camera.cpp:172-205) perspective branch: look_at(pos, target, up) with vfov in
 * degrees and aspect; when use_matrix != 0, to_world (row-major 4x4 camera-to-world) replaces look_at. */
typedef struct ks_camera_desc {
    float pos[3];
    float target[3];
    float up[3];
    float vfov_deg;
    float aspect;
    int32_t use_matrix;
    float to_world[16];
} ks_camera_desc;

#define KS_MIS_REFERENCE 0  /* NEE weight without light-selection probability (nee.cpp:49) */
#define KS_MIS_CONSISTENT 1 /* both strategies weighted with the selection probability */

#define KS_GUIDE_OFF 0
#define KS_GUIDE_TRAIN_AND_SAMPLE 1 /* online-trained vMF-mixture guide, one-sample MIS with the BSDF */

/* SmallPTInput (small_pt.h:17-41) plus this build's additions (mis_mode, guide_mode, queue size). */
typedef struct ks_small_pt_params {
    int32_t bounces;
    int32_t render_width, render_height;
    int32_t crop_start_x, crop_start_y, crop_width, crop_height;
    int32_t spp;
    int32_t rng_seed;
    int32_t include_background;
    float backdrop[3];
    int32_t scale_ray_diff; /* camera ray differentials scaled by 1/sqrt(spp) (Camera::spawn_ray, camera.cpp:162-167) */
    float clamp_direct;
    float clamp_indirect;
    int32_t spp_prog_interval;
    int32_t mis_mode;
    int32_t guide_mode;
    int32_t queue_size; /* wavefront path slots; 0 = automatic: the pass's paths, at most 2^26 */
    /* SmallPTInput::pixel_filter (create_pixel_filter, render_target.cpp:28-44; render_target.h:79-184):
     * KS_FILTER_BOX draws the jitter uniformly over [-w/2, w/2]^2; the gaussian and Blackman-Harris
     * filters draw it from their 32-per-unit sample table (PixelFilterSampleTable). */
    int32_t pixel_filter;
    float filter_width; /* 0: the type's default (box 1, gaussian 3, blackman_harris 3) */
    float filter_sigma; /* gaussian only, 0: 0.5 */
} ks_small_pt_params;

#define KS_FILTER_BOX 0
#define KS_FILTER_GAUSSIAN 1
#define KS_FILTER_BLACKMAN_HARRIS 2

/* Kernel classes timed with HIP events on the render stream. */
#define KS_K_RAYGEN 0
#define KS_K_CLOSEST 1
#define KS_K_SHADE 2
#define KS_K_SHADOW 3
#define KS_K_FILM 4
#define KS_K_GUIDE 5
#define KS_K_TAIL 6 /* the last paths of a render, each carried to its end by one thread */
#define KS_K_COUNT 7

typedef struct ks_render_stats {
    double wall_ms;                     /* host wall time of the render call */
    double kernel_ms[KS_K_COUNT];       /* summed event time per kernel class (0 unless timing on) */
    uint64_t launches[KS_K_COUNT];
    uint64_t paths;                     /* camera samples traced (W*H*spp over the crop) */
    uint64_t rays_closest;              /* primary + extension rays of the wavefront's traversal kernel */
    uint64_t rays_shadow;               /* NEE shadow rays of the wavefront's traversal kernel */
    uint64_t rays_tail;                 /* closest-hit and shadow rays of the tail kernel (KS_K_TAIL) */
    uint64_t node_visits_closest, tri_tests_closest;
    uint64_t node_visits_shadow, tri_tests_shadow;
    uint64_t nonfinite;                 /* non-finite throughput events (thread_monitor_check, small_pt.cpp:97) */
    uint64_t stack_overflows;           /* traversal stack entries spilled past LDS */
    int32_t iterations;
    int32_t queue_size;
    uint64_t guide_samples;             /* training samples consumed by Adam steps (guided renders) */
    int32_t trace_blocks_per_cu;        /* traversal blocks per CU of a batched render */
    int32_t batch_chunks;               /* wavefront chunks of a batched render */
} ks_render_stats;

typedef void (*ks_progress_fn)(const ks_film *film, int32_t spp_finished, void *user);

/* ---- device ---------------------------------------------------------------------------------- */
const char *ks_last_error(void);
int ks_device_count(int32_t *count);
/* Opens a device handle: its stream, pinned control blocks and the sampler's constant permutation
   table (128 MiB of device memory, filled once here and freed by ks_device_close). */
int ks_device_open(int32_t ordinal, ks_device **out);
int ks_device_close(ks_device *dev);
/* Record HIP-event kernel timings into ks_render_stats (costs one event pair per launch). */
int ks_device_set_timing(ks_device *dev, int32_t enable);
/* Shading order of the wavefront (no counterpart in the reference, whose CPU loop shades one path at a time,
 * Material::sample_with_nee, material.cpp:101-167): 0 (default) shades every path with one kernel; 1 sorts
 * each bounce's paths by material class (escaped rays, opaque dielectrics, metals, other materials) and
 * shades each class with a kernel compiled for its BSDF lobes. Results are the same either way; textured
 * scenes always use one kernel. */
int ks_device_set_shading(ks_device *dev, int32_t sorted);
/* The end of a render's wavefront (no counterpart in the reference, whose CPU loop traces each path to its
 * end): once every work item has started and at most max_paths paths remain, one kernel carries each of
 * them to its end in one thread instead of one wavefront iteration per bounce. Results are the same.
 * Applies to plain renders (not guided ones, not HIP-graph batches) of scenes without alpha-tested
 * triangles; 0 turns it off. Default 0: with the default 2^26-path queue the wavefront's last iterations
 * cost less than the tail kernel (40.5 against 42.0 ms for a 1080p 64-spp render). */
int ks_device_set_tail(ks_device *dev, int32_t max_paths);

/* ---- scene: replaces create_scene + Scene::create_rtc_scene (rtcCommitScene, scene.cpp:144-160),
 *      Scene::build_mesh_lights (scene.cpp:315-336) and PowerLightSampler::build (light.cpp:636-677).
 *      Builds the SBVH on the host (cuda/sbvh.cu:301-440) and uploads scene + BVH + light tables. ---- */
int ks_scene_create(ks_device *dev, const ks_scene_desc *desc, const ks_bvh_options *opt, ks_scene **out);
int ks_scene_free(ks_scene *scene);

typedef struct ks_scene_info {
    uint32_t n_tris, n_vertices, n_nodes, n_refs, n_lights, max_depth;
    uint64_t device_bytes;
    double build_ms;
    float bound_lo[3], bound_hi[3]; /* Scene::bound (scene.cpp:162-169) */
    uint32_t branching;             /* node width of the device BVH (n_nodes and max_depth refer to it) */
    int32_t builder;                /* the KS_BVH_BUILD_* that built it */
    double bvh_build_ms;            /* wall time of the BVH build alone (device builds: kernels included) */
} ks_scene_info;
int ks_scene_info_get(const ks_scene *scene, ks_scene_info *info);
/* Copies the device BVH out: info.n_nodes nodes (128-byte BVH4 or 64-byte BVH2 records, layouts in
 * keepsake_amd/csrc/ks_internal.h) and info.n_refs 48-byte triangle records in leaf order. For
 * inspection and tests; the reference's Embree BVH is opaque (no counterpart). */
int ks_scene_bvh_get(const ks_scene *scene, void *nodes, void *refs);
/* PowerLightSampler tables in flat light order: prim (-1 for non-mesh), power, cdf (n+1). */
int ks_scene_light_table(const ks_scene *scene, int32_t *n_lights, int32_t *prim, float *power, float *cdf);

/* ---- film: RenderTarget2 (render_target.h:224-293): FP64 RGB sums + FP64 (hit, miss) weights ---- */
int ks_film_create(ks_device *dev, int32_t width, int32_t height, const float backdrop[3], ks_film **out);
int ks_film_clear(ks_film *film);                                  /* RenderTarget2::clear */
int ks_film_download(const ks_film *film, double *rgb, double *weights); /* W*H*3, W*H*2 */
int ks_film_composite(const ks_film *film, float *rgb);            /* composite (render_target.cpp:187-192) */
int ks_film_free(ks_film *film);
/* relMSE of two films' composites: mean over pixels and channels of (x - r)^2 / (r^2 + 0.01), in FP64 on
 * the device (the error metric of the guided-vs-plain comparison; no reference counterpart) */
int ks_film_relmse(const ks_film *film, const ks_film *reference, double *relmse);
/* ks_film_relmse without any allocation (no counterpart in the reference): for callers that cannot
 * allocate device memory, such as tight progressive loops that check convergence every pass or code
 * around HIP graph capture. The partial sums live in a static device array of 2049 doubles (16 KiB) that
 * each GPU holds for the life of the process; every call on one GPU shares it, so calls that overlap on
 * one GPU race (as they would on the device's stream and pinned host block). Films above 2^32 pixels
 * are rejected (KS_ERR_INVALID). The result equals ks_film_relmse's up to FP64 summation order. */
int ks_film_relmse_noalloc(const ks_film *film, const ks_film *reference, double *relmse);

/* ---- the render call: SmallPT::run (small_pt.cpp:15-70). Accumulates into `film` (which must be
 *      render_width x render_height); calls progress after each spp batch of spp_prog_interval on the
 *      calling thread (small_pt.cpp:66-68). ---- */
int ks_small_pt_render(ks_scene *scene, const ks_camera_desc *cam, const ks_small_pt_params *params,
                       ks_film *film, ks_progress_fn progress, void *user, ks_render_stats *stats);

/* ---- batched renders: many views of one scene (each its own camera, resolution, crop and spp) in one
 *      wavefront, iterations replayed from a HIP graph, queue size and traversal occupancy chosen by an
 *      autotuner. Each view's film equals what ks_small_pt_render gives for that view alone (the same
 *      Sobol samples, folded in sample order). No reference counterpart: the reference renders one
 *      frame per task. spp_prog_interval and progress callbacks do not apply. ---- */
typedef struct ks_batch_view {
    ks_camera_desc camera;
    ks_small_pt_params params;
    ks_film *film;
} ks_batch_view;
typedef struct ks_batch_options {
    int32_t queue_size;          /* 0: autotune over 2^18..2^25 (cached per scene and workload size) */
    int32_t trace_blocks_per_cu; /* with a fixed queue_size: traversal blocks per CU, 1..8 (8) */
    int32_t disable_graph;       /* 1: launch every iteration eagerly instead of replaying a HIP graph */
    int32_t max_chunk_items;     /* work items (path samples) per chunk; 0: 2^28, i.e. 4 GiB of per-sample results.
                                    A view's samples may be split across chunks at sample boundaries */
} ks_batch_options;
int ks_small_pt_render_batch(ks_scene *scene, int32_t n_views, const ks_batch_view *views,
                             const ks_batch_options *opt, ks_render_stats *stats);

/* ---- measurement ---- */
/* The rays a render issues: traces the listed (pixel, sample) paths exactly as ks_small_pt_render would
 * (the same Sobol samples, kernels and params) and returns every closest-hit ray and every NEE shadow ray
 * they issue, up to max_rays of each, as n x 8 floats (o, tmin, d, tmax). Replaces the reference's
 * per-ray loop in SmallPT::trace (small_pt.cpp:72-178), which issues them one at a time through
 * Scene::intersect1 / occlude1 (scene.cpp:206-244). */
int ks_trace_capture(ks_scene *scene, const ks_camera_desc *cam, const ks_small_pt_params *params, int64_t n,
                     const int32_t *px, const int32_t *py, const int32_t *sample, int64_t max_rays, float *closest_rays,
                     int64_t *n_closest, float *shadow_rays, int64_t *n_shadow);
/* The work of the reference traversal order on rays (n x 8): node fetches and triangle tests of
 * SBVH::intersect (any_hit 0) or SBVH::intersectBool (1) (cuda/sbvh.cu:512-603), with the reference's
 * slab and watertight triangle tests, over the scene's host-built SBVH in BVH2 form; counted on the host.
 * t_out (optional, n): the closest hit distance (-1: miss), or 1 / -1 occluded / not. Needs a scene built
 * with KS_BVH_BUILD_HOST_SBVH (else KS_ERR_UNSUPPORTED). */
int ks_scene_traversal_counts(const ks_scene *scene, const float *rays, int64_t n, int32_t any_hit,
                              uint64_t *node_fetches, uint64_t *tri_tests, float *t_out);

/* ---- component entry points (parity tests; same kernels as the render path) ---- */
/* PTRenderSampler (sobol.h:235-248) streams for n (pixel, sample) pairs; kinds[c] = 1 (next) or
 * 2 (next2d); out holds n * n_values values (uint32 before the float map, and the float). */
int ks_sobol_stream(ks_device *dev, int32_t width, int32_t height, int32_t spp, int32_t seed, int64_t n,
                    const int32_t *px, const int32_t *py, const int32_t *sample, int32_t n_calls,
                    const int32_t *kinds, uint32_t *out_u32, float *out_f);
/* Camera::spawn_ray after the pixel-filter jitter (small_pt.cpp:98-105); rays n x 6 (o, d). */
int ks_camera_rays(ks_device *dev, const ks_camera_desc *cam, int32_t width, int32_t height, int32_t spp,
                   int32_t seed, int64_t n, const int32_t *px, const int32_t *py, const int32_t *sample,
                   float *rays);
/* Scene::intersect1 (scene.cpp:206-236): rays n x 8 (o.xyz, tmin, d.xyz, tmax) ->
 * hits n x 4 (t, u, v, prim id as int32 bits, -1 = miss); u, v are Embree's barycentrics. */
int ks_intersect(ks_scene *scene, const float *rays, int64_t n, float *hits);
/* Scene::occlude1 (scene.cpp:238-244) */
int ks_occluded(ks_scene *scene, const float *rays, int64_t n, uint8_t *occluded);
/* PrincipledBSDF::internal::eval/pdf (principled_bsdf.cpp:308-421) on local directions. */
int ks_bsdf_eval(ks_device *dev, const ks_material_desc *m, const float *wo, const float *wi, int64_t n,
                 float *f, float *pdf);
/* PrincipledBSDF::internal::sample (principled_bsdf.cpp:326-393); u: n x 3 (u_lobe, u_wi). */
int ks_bsdf_sample(ks_device *dev, const ks_material_desc *m, const float *wo, const float *u, int64_t n,
                   float *wi, float *beta, float *pdf);
/* SmallPT::trace (small_pt.cpp:72-178) for a list of (pixel, sample): out n x 4 (L.rgb, hit). */
int ks_trace_samples(ks_scene *scene, const ks_camera_desc *cam, const ks_small_pt_params *params, int64_t n,
                     const int32_t *px, const int32_t *py, const int32_t *sample, float *out);

/* ---- path guide: hash-grid encoding + 64-64-64-48 BF16 MLP (MFMA, FP32 accumulation) -> 8-lobe vMF
 *      mixture, trained online with Adam. The reference has no network: the lobe math is VMF/WVMF/
 *      sample_vmm (vmf.h:10-127) and the update rule is adam_optimizer_step (gpu/shaders/adam.slang:25-51).
 *      Zeroed config fields take the defaults given in the comments. ---- */
typedef struct ks_guide ks_guide;
typedef struct ks_guide_config {
    int32_t log2_table_size;    /* hash-table entries per level, log2 (19) */
    float base_resolution;      /* coarsest grid resolution (16) */
    float max_resolution;       /* finest grid resolution, per unit of the scene bound (256) */
    float learning_rate;        /* Adam lr (1e-2) */
    float beta1, beta2, eps;    /* Adam moments (0.9, 0.999, 1e-8 as adam.slang) */
    float guide_fraction;       /* one-sample MIS: probability of sampling the guide (0.5) */
    int32_t train_batch;        /* samples per Adam step (2^18) */
    int32_t max_steps_per_pass; /* Adam steps after each 1-spp pass of a guided render (4) */
    uint64_t seed;
    int32_t train_spp;          /* guided renders train during their first train_spp samples per pixel, in
                                   1-spp passes, then render the rest frozen in batched passes (0: train
                                   throughout) */
    float weight_clamp;         /* render training: sample weights divided by the pass mean, clamped at
                                   this (20) */
    int32_t inference_fp8;      /* 1: sample the guide with FP8 (OCP e4m3) MLP inference, per-tensor power-of-two
                                   scales; training stays BF16 with FP32 master weights (0: BF16) */
    int32_t max_guided_bounce;  /* vertices at this bounce or deeper sample the BSDF only and skip inference
                                   (0: every vertex is guided) */
} ks_guide_config;

int ks_guide_create(ks_device *dev, const ks_guide_config *cfg, const float bound_lo[3], const float bound_hi[3],
                    ks_guide **out);
int ks_guide_free(ks_guide *guide);
/* grid parameters (16 levels x 2^log2 entries x 2 features) and MLP weights (11264) */
int ks_guide_param_count(const ks_guide *guide, int64_t *n_grid, int64_t *n_mlp);
/* sampling fraction and Adam steps per pass of later guided renders (max_steps_per_pass <= 0: frozen) */
int ks_guide_set_mode(ks_guide *guide, float guide_fraction, int32_t max_steps_per_pass);
/* training progress; the last_* sums cover the most recent Adam step: sum c (-log q(wi)), sum c, and
 * sum c (-log(cos(n, wi) / pi)) (the same cross-entropy for a cosine lobe, for comparison) */
typedef struct ks_guide_stats {
    int64_t steps, samples;
    double last_loss, last_weight, last_loss_cosine;
} ks_guide_stats;
int ks_guide_get_stats(const ks_guide *guide, ks_guide_stats *stats);
/* inference precision of later forwards and guided renders: 1 FP8 (e4m3), 0 BF16. Switching to FP8 builds
 * the FP8 copy of the network from the current weights and activation maxima. */
int ks_guide_set_inference(ks_guide *guide, int32_t fp8);
/* FP8 calibration: records the largest |input| of each MLP layer over n queries (position, normal, wo;
 * BF16 forward) and rebuilds the FP8 copy with input scales from them. Guided renders calibrate by
 * themselves: every training step records the maxima over its batch. */
int ks_guide_calibrate_fp8(ks_guide *guide, int64_t n, const float *queries);
/* the current FP8 scales: 1/s_x, 1/s_h1, 1/s_h2, s_x s_w1, s_h1 s_w2, s_h2 s_w3, s_w1, s_w2, s_w3, s_x, s_h1,
 * s_h2 (all powers of two; s_w = the smallest power of two with max|W| / s_w <= 448, s_in = twice that for
 * the recorded input maximum) */
int ks_guide_fp8_scales(const ks_guide *guide, float scales[12]);
/* grid resolutions of the 16 levels and the position normalisation u = (x - lo) * inv_ext */
int ks_guide_grid_info(const ks_guide *guide, int32_t res[16], float lo[3], float inv_ext[3]);
/* FP32 master parameters: [grid][W1 64x64][W2 64x64][W3 64x48], weights [in][out] */
int ks_guide_get_params(const ks_guide *guide, float *params);
int ks_guide_set_params(ks_guide *guide, const float *params);
/* queries n x 9 (position, normal, wo) -> mixtures n x 40: w[8], then (mu xyz, kappa) x 8 */
int ks_guide_forward(ks_guide *guide, int64_t n, const float *queries, float *mixtures);
/* loss = sum_i -weight_i log q(wi_i | query_i) and its gradient w.r.t. every parameter */
int ks_guide_gradient(ks_guide *guide, int64_t n, const float *queries, const float *wi, const float *weight,
                      double *loss, float *grad);
/* one Adam step on that loss */
int ks_guide_train_step(ks_guide *guide, int64_t n, const float *queries, const float *wi, const float *weight,
                        double *loss);
/* small_pt with the guide: passes of one sample per pixel; after each pass the guide takes up to
 * max_steps_per_pass Adam steps on training samples recorded at the first three vertices of each path
 * (weight luminance of the radiance the continuation brought back, divided by the path throughput before
 * the vertex). From the first trained pass on, vertices with roughness >= 0.05 pick the mixture with
 * probability guide_fraction and the BSDF otherwise (one-sample MIS, continuation pdf
 * alpha q + (1 - alpha) p_bsdf, also in the NEE and emitter MIS weights). Needs mis_mode consistent.
 * No reference counterpart: the reference has no learned guide. */
int ks_small_pt_render_guided(ks_scene *scene, const ks_camera_desc *cam, const ks_small_pt_params *params,
                              ks_guide *guide, ks_film *film, ks_progress_fn progress, void *user,
                              ks_render_stats *stats);
/* mixture pdf at directions / samples from mixtures (sample_vmm, vmf.h:121-127); u n x 2 */
int ks_vmm_pdf(ks_device *dev, int64_t n, const float *mixtures, const float *dirs, float *pdf);
int ks_vmm_sample(ks_device *dev, int64_t n, const float *mixtures, const float *u, float *dirs, float *pdf);

/* ---- procedural scenes of the benchmark configurations (no reference counterpart: the reference
 *      loads meshes from files, mesh_asset.cpp). name: "cornell", "spheres", "spheres_small", "room",
 *      "room_spheres" (the room plus 30 icospheres, 10M triangles). ---- */
int ks_host_scene_create(const char *name, uint64_t seed, ks_host_scene **out);
int ks_host_scene_get(const ks_host_scene *hs, ks_scene_desc *desc, ks_camera_desc *cam);
int ks_host_scene_free(ks_host_scene *hs);

#ifdef __cplusplus
}
#endif

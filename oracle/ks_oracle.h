/*
 * CPU oracle for the keepsake small_pt path (test infrastructure only).
 *
 * This is a plain C++ restatement of the reference CPU integrator and the pieces it calls
 * (small_pt.cpp, sobol.h, hash.h, rng.h, camera.cpp, ray.h, geometry.cpp, light.cpp, nee.cpp,
 * material.cpp, principled_bsdf.cpp, microfacet.h, fresnel.h, distrib.cpp, render_target.cpp,
 * parallel.h, cuda/ray_tri.cuh, cuda/aabb.cuh). Each function in ks_oracle.cpp cites the reference
 * file:line it follows.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
 * only as the checker or as the timed CPU baseline -- never as the product path.
 *
 * Embree (the reference's closest-hit/occlusion engine) is not available; the oracle uses its own
 * binned-SAH BVH (object splits only, independent of the product's SBVH) with the reference's
 * watertight triangle test and traversal rules (strict t < tmax update, first-found wins on ties).
 *
 * Pinning status per function (what checks each one; details in the tests):
 *   ora_sobol          - MurmurHash64A KAT, generator-matrix closed forms checked against the reference
 *                        table text, (0,m,2)-net stratification per pixel, digit-permutation bijection.
 *                        The split of `dimension` between digit permutation and scramble hash is pinned
 *                        only by the restatement itself (parity partially unpinned).
 *   ora_camera_rays    - closed-form pinhole direction (independent of the matrix inverse path).
 *   ora_intersect      - brute force over all triangles; closed meshes never leak rays aimed at their
 *                        vertices and edges; strict interval ends; the FP64 edge fallback decides a
 *                        2^-24 case exactly (tests/test_oracle_geometry.py).
 *   ora_bsdf_*, ora_ggx_*, ora_fresnel
 *                      - GGX normalisation and projected-area identity by quadrature, angle-form D and
 *                        Lambda, Beta-function G2, VNDF chi^2; two-index Fresnel, Brewster, TIR; lobes
 *                        against FP64 closed forms, reciprocity; every lobe sampler by chi^2 against the
 *                        density derived from it, weight == eval/pdf (tests/test_oracle_bsdf.py).
 *   ora_light_*        - powers as closed forms of the scene, CDF as normalised prefix sums, upper_bound
 *                        selection, uniform-area triangle samples with pdf d^2/(A|cos|), ambient sky
 *                        pdf 1/(2 pi^2 sin theta) (tests/test_oracle_lights.py).
 *   ora_render         - equal-spp comparison is the parity check for the GPU path.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors PrincipledBSDF closure inputs (principled_bsdf.h:78-93) + Material::emission (material.h:45). */
typedef struct ora_material {
    float basecolor[3];
    float roughness;
    float metallic;
    float ior;
    float specular_r0_mul; /* raw [0,1] field value; the closure doubles it (principled_bsdf.cpp:258) */
    float specular_trans;
    float diffuse_trans;
    float diffuse_trans_fwd;
    float emission[3];
    int32_t has_emission;
} ora_material;

/* Non-mesh lights (light.h:34-120): 0 = ambient SkyLight, 1 = DirectionalLight, 2 = PointLight,
 * 3 = SkyLight from the scene's equirectangular map (ora_scene_desc.sky_map). */
typedef struct ora_light {
    int32_t type;
    float color[3]; /* ambient radiance / L / I */
    float vec[3];   /* direction (directional) or position (point) */
} ora_light;

/* SkyLight(path, to_world, transform_y_up, strength) (light.cpp:19-49): h rows of w RGB texels, row 0
 * at theta = 0; to_world is the row-major linear part of the light's transform. */
typedef struct ora_sky_map {
    const float *rgb;
    int32_t width, height;
    float to_world[9];
    int32_t transform_y_up;
    float strength;
} ora_sky_map;

/* Texture (texture.h:46-66): levels back to back (each max(1, (n + 1) / 2) of the one before), texels of
 * `channels` values of type 0 u8 / 1 u16 / 2 f32; color_space 0 linear, 1 sRGB; with levels == 1 and
 * build_mipmaps the chain is built as Texture(bytes, ..., true) does. */
typedef struct ora_texture {
    const void *data;
    int32_t width, height, channels, data_type, color_space, levels, build_mipmaps;
} ora_texture;

/* ShaderField (shader_field.h:11-86): texture -1 is the constant `value`; sampler 0 nearest, 1 linear,
 * 2 cubic (kernel 0 Mitchell-Netravali, 1 B-spline, 2 Catmull-Rom), 3 EWA; wrap 0 repeat, 1 clamp. */
typedef struct ora_field {
    int32_t texture, sampler, cubic_kernel;
    float anisotropy;
    int32_t wrap_u, wrap_v, flip_v;
    int32_t swizzle[4];
    float uv_scale[2], uv_offset[2], scale[4], value[4];
} ora_field;

/* PrincipledBSDF fields + Material::emission / normal_map / opacity_map (material.h:39-43). */
typedef struct ora_material_fields {
    ora_field basecolor, roughness, metallic, ior, specular_r0_mul, specular_trans, diffuse_trans,
        diffuse_trans_fwd;
    int32_t has_emission;
    ora_field emission;
    int32_t has_normal_map;
    ora_field normal_map;
    int32_t normal_space, normal_range01; /* space 0 tangent, 1 world */
    float normal_strength;
    float normal_to_world[9];
    int32_t has_opacity_map;
    ora_field opacity;
} ora_material_fields;

typedef struct ora_scene_desc {
    const float *vertices; /* 3 * n_vertices */
    uint32_t n_vertices;
    const float *normals; /* optional per-vertex normals, 3 * n_vertices, or NULL */
    const uint32_t *indices; /* 3 * n_tris */
    uint32_t n_tris;
    const uint32_t *tri_geom; /* geometry id per triangle, triangles of one geometry contiguous */
    const uint32_t *geom_material;
    uint32_t n_geoms;
    const ora_material *materials;
    uint32_t n_materials;
    const ora_light *lights;
    uint32_t n_lights;
    const uint8_t *geom_flags; /* per geometry: 1 one-sided (MeshData::twosided = false), 2 flat normals; or NULL */
    const ora_sky_map *sky_map; /* texels of the type-3 light, or NULL */
    const float *texcoords;     /* 2 per vertex, or NULL */
    const ora_texture *textures;
    uint32_t n_textures;
    const ora_material_fields *const *material_fields; /* per material, NULL entries: constant material */
} ora_scene_desc;

/* create_camera (camera.cpp:172-205): look_at(pos, target, up) + perspective vfov (degrees) / aspect. */
typedef struct ora_camera {
    float pos[3];
    float target[3];
    float up[3];
    float vfov_deg;
    float aspect;
} ora_camera;

/* SmallPTInput (small_pt.h:17-41) plus mis_mode (0 = reference weights, 1 = consistent). */
typedef struct ora_params {
    int32_t bounces;
    int32_t render_width, render_height;
    int32_t crop_start_x, crop_start_y, crop_width, crop_height;
    int32_t spp;
    int32_t rng_seed;
    int32_t include_background;
    float backdrop[3];
    float clamp_direct, clamp_indirect;
    int32_t mis_mode;
    int32_t scale_ray_diff; /* camera ray differentials scaled by 1/sqrt(spp) (camera.cpp:162-167) */
    int32_t pixel_filter;   /* 0 box, 1 gaussian, 2 Blackman-Harris (create_pixel_filter, render_target.cpp:28-44) */
    float filter_width;     /* 0: the type's default (1, 3, 3) */
    float filter_sigma;     /* gaussian, 0: 0.5 */
} ora_params;

void *ora_scene_create(const ora_scene_desc *desc);
void ora_scene_free(void *scene);
/* rays: n x 8 floats (o.xyz, tmin, d.xyz, tmax); out: n x 4 (t, u, v, prim as int bits, -1 = miss) */
int ora_intersect(void *scene, const float *rays, int64_t n, float *out, int threads);
int ora_intersect_brute(void *scene, const float *rays, int64_t n, float *out, int threads);
int ora_occluded(void *scene, const float *rays, int64_t n, uint8_t *out, int threads);
/* Light table as the sampler sees it: flat order, prim id (-1 non-mesh), cdf (n+1). */
int ora_light_table(void *scene, int32_t *n_lights, int32_t *prim, float *power, float *cdf);
/* The scene's sky-map light: radiance (n x 3) and solid-angle pdf (n) at world directions (n x 3). */
int ora_sky_eval(void *scene, const float *dirs, int64_t n, float *le, float *pdf);
/* Light samples from u (n x 2): direction (n x 3), Le / pdf (n x 3), pdf (n). */
int ora_sky_sample(void *scene, const float *u, int64_t n, float *wi, float *le_over_pdf, float *pdf);
/* ShaderField evaluation of texture-backed field `f` of the scene at n (u, v, dudx, dvdx, dudy, dvdy)
 * rows: out n x 4. */
int ora_field_eval(void *scene, const ora_field *f, const float *uvd, int64_t n, float *out);
/* Level `level` of texture `tex` as float texels (Texture::fetch_as_float): out w * h * channels. */
int ora_texture_level(void *scene, int tex, int level, int *w, int *h, float *out);
/* The oracle's sRGB u8 -> linear table (256) and EWA weights (128). */
void ora_tex_tables(float *srgb256, float *ewa128);
/* Surface at the camera ray's first hit for (pixel, sample): out n x 16 = (u, v, dudx, dvdx, dudy, dvdy,
 * shading n xyz, shading t xyz, geometric n xyz, prim as int bits); prim -1 on a miss. */
int ora_primary_surface(void *scene, const ora_camera *cam, const ora_params *p, const int32_t *px,
                        const int32_t *py, const int32_t *sample, int64_t n, float *out);
/* rgb: W*H*3 FP64 sums, weights: W*H*2 FP64 (hit, miss) -- RenderTarget2 (render_target.h:224-293). */
int ora_render(void *scene, const ora_camera *cam, const ora_params *p, double *rgb, double *weights,
               int threads);
/* Scalar per-sample radiance trace for given (pixel, sample) list: out n x 4 (rgb, hit). */
int ora_trace_samples(void *scene, const ora_camera *cam, const ora_params *p, const int32_t *px,
                      const int32_t *py, const int32_t *sample, int64_t n, float *out, int threads);

/* PTRenderSampler stream for one (pixel, sample): kinds[i] = 1 (next) or 2 (next2d).
 * out_u32/out_f receive 1 or 2 values per call, in order. */
int ora_sobol(int width, int height, int spp, int seed, int px, int py, int sample, int n_calls,
              const int32_t *kinds, uint32_t *out_u32, float *out_f);
uint64_t ora_murmur64a(const void *key, int len, uint64_t seed);
uint64_t ora_mix_bits(uint64_t v);
uint32_t ora_owen_scramble(uint32_t v, uint32_t seed);
void ora_sobol_matrices(uint32_t *out104);

/* Camera rays for (pixel, sample) list with pixel-filter jitter from the sampler (small_pt.cpp:98-105);
 * rays: n x 6 (o.xyz, d.xyz). */
int ora_camera_rays(const ora_camera *cam, int width, int height, int spp, int seed, const int32_t *px,
                    const int32_t *py, const int32_t *sample, int64_t n, float *rays);
/* Camera matrices: proj_to_world (row-major 4x4) and position. */
int ora_camera_matrices(const ora_camera *cam, float *proj_to_world16, float *position3);

/* PrincipledBSDF on local directions: f (rgb, cosine-weighted), pdf. */
int ora_bsdf_eval(const ora_material *m, const float *wo, const float *wi, int64_t n, float *f, float *pdf);
/* u: n x 3 (u_lobe, u_wi.x, u_wi.y) -> wi (n x 3), beta (n x 3), pdf (n). */
int ora_bsdf_sample(const ora_material *m, const float *wo, const float *u, int64_t n, float *wi, float *beta,
                    float *pdf);

/* Probes of building blocks, used by the tests that pin them (closed forms, quadrature, chi^2).
 * GGX (microfacet.h:20-153) with clamped alphas: what 0 D(a), 1 Lambda(a), 2 G1(a), 3 G2(a, b). */
int ora_ggx_eval(float ax, float ay, int what, const float *a, const float *b, int64_t n, float *out);
/* VNDF sample (microfacet.h:98-118): wo (n x 3), u (n x 2) -> microfacet normal (n x 3). */
int ora_ggx_sample(float ax, float ay, const float *wo, const float *u, int64_t n, float *wm);
/* what 0 fresnel_dielectric(c, eta), 1 fresnel_schlick(c) (fresnel.h:9-30), 2 HG phase(c, g = eta). */
int ora_fresnel(int what, const float *c, float eta, int64_t n, float *out);
/* lobe_sample_weights (principled_bsdf.cpp:518-540): n x 4. */
int ora_bsdf_lobes(const ora_material *m, const float *wo, int64_t n, float *w4);
/* One lobe's f*cos (0 diffuse, 1 diffuse transmission, 2 metallic, 3 dielectric), or lobe -1: the
 * mixture pdf (principled_bsdf.cpp:395-421) in every channel. f: n x 3. */
int ora_bsdf_lobe(const ora_material *m, int lobe, const float *wo, const float *wi, int64_t n, float *f);
/* Light `light` of the flat table sampled at points p (n x 3) with u (n x 2). */
int ora_light_sample(void *scene, int light, const float *p, const float *u, int64_t n, float *wi, float *dist,
                     float *pdf, float *le_over_pdf);
int ora_light_pdf(void *scene, int light, const float *wi, const float *dist, int64_t n, float *pdf);
/* The params' pixel filter: offsets sampled from u (n x 2) and values at offsets xy (n x 2) (either pair may
 * be NULL). Returns the sample-table size n (0 for the box filter). */
int ora_pixel_filter(const ora_params *p, const float *u, const float *xy, int64_t n, float *offset, float *value);
/* Light selection from u (n): index and probability. */
int ora_light_select(void *scene, const float *u, int64_t n, int32_t *idx, float *prob);

#ifdef __cplusplus
}
#endif

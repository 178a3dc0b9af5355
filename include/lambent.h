/*
 * lambent.h - C ABI of the MI355X (gfx950) path-tracing library.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an lmb_status (0 = success).  On failure the message of the last
 * error on the calling thread is available from lmb_last_error().
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the Vecvec/Lambent source tree).
 *
 * Threading: handles are externally synchronised.  Calls that take a
 * `stream` (a hipStream_t; NULL = the default stream) return once the GPU
 * work is enqueued; lmb_sync() waits.  Host arrays are copied during the
 * call that receives them.  Device buffers belong to the library; their
 * device addresses can be exported for zero-copy consumers.
 */
#ifndef LAMBENT_H
#define LAMBENT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
typedef int32_t lmb_status;
#define LMB_OK 0
#define LMB_ERR_INVALID_ARGUMENT (-1)
/* camera.rs:63-66 CameraError::ProjectionUninvertable */
#define LMB_ERR_SINGULAR_PROJECTION (-2)
/* camera.rs:63-66 CameraError::ViewUninvertable */
#define LMB_ERR_SINGULAR_VIEW (-3)
/* lib.rs:165-170: `assert_ne!(u16::MAX, idx)` panics in the reference */
#define LMB_ERR_RESERVED_INDEX (-4)
#define LMB_ERR_OUT_OF_MEMORY (-5)
/* a HIP runtime call failed, or no GPU is present */
#define LMB_ERR_DEVICE (-6)
/* the built BVH is deeper than the traversal stack supports */
#define LMB_ERR_BVH_TOO_DEEP (-7)
/* textures.rs:93-97 TextureError (bad dimensions / data) */
#define LMB_ERR_TEXTURE (-8)

/* Message describing the last failure on this thread ("" if none). */
const char *lmb_last_error(void);
/* Library version string. */
const char *lmb_version(void);

/* ------------------------------------------------------------------ */
/* Materials  (lib.rs:83-203, bindings.wesl:48-57)                     */
/* ------------------------------------------------------------------ */
/* 32-byte record, bit-identical to the reference's `Material`. */
typedef struct lmb_material {
    uint32_t tex_pos_1;                /* 2 x f16, low half = u       */
    uint32_t tex_pos_2;
    uint32_t tex_pos_3;
    uint32_t tex_pos_recolour;         /* 2 x u16 texel coordinate     */
    uint32_t tex_idx_diffuse_emission; /* u16 diffuse | u16 emission<<16 */
    uint32_t tex_idx_attributes_ty;    /* u16 attributes | u16 type<<16  */
    uint32_t emission_scale;           /* bf16 emissive | bf16 diffuse<<16 */
    uint32_t refractive_index;         /* f16 n(760nm) | f16 n(340nm)<<16 */
} lmb_material;

/* lib.rs:194-203 MaterialType */
#define LMB_MATERIAL_DIFFUSE 0u
#define LMB_MATERIAL_METALLIC 1u
#define LMB_MATERIAL_TRANSPARENT 2u
/* bindings for `Option<u16>` arguments: negative = None */
#define LMB_NONE (-1)

/* lib.rs:103-121 EmissionScale::from_f32 + pack: two bf16 (RNE) packed. */
lmb_status lmb_emission_scale_pack(float emissive_texture_scale,
                                   float diffuse_texture_scale,
                                   uint32_t *packed);

/* lib.rs:132-162 Material::new.
 * tex_idx_emission / tex_idx_attributes: LMB_NONE or 0..0xFFFE
 * (0xFFFF is reserved and returns LMB_ERR_RESERVED_INDEX).
 * refractive_index: NULL (= VACUUM 1..1) or {n_760nm, n_340nm}. */
lmb_status lmb_material_new(const float tex_pos_1[2], const float tex_pos_2[2],
                            const float tex_pos_3[2],
                            const uint16_t tex_pos_recolour[2],
                            uint16_t tex_idx_diffuse, int32_t tex_idx_emission,
                            int32_t tex_idx_attributes,
                            uint32_t emission_scale_packed,
                            const float *refractive_index, uint32_t ty,
                            lmb_material *out);

/* ------------------------------------------------------------------ */
/* Camera  (camera.rs:6-70)                                            */
/* ------------------------------------------------------------------ */
/* Column-major 4x4 inverses, 128 bytes, bit-identical layout. */
typedef struct lmb_camera {
    float projection_inverse[16];
    float view_inverse[16];
} lmb_camera;

/* camera.rs:13-25 Camera::from_proj_view (column-major inputs). */
lmb_status lmb_camera_from_proj_view(const float proj[16], const float view[16],
                                     lmb_camera *out);
/* camera.rs:27-42 Camera::update_from_proj_view; NULL leaves a matrix as is. */
lmb_status lmb_camera_update_from_proj_view(lmb_camera *cam, const float *proj,
                                            const float *view);

/* lib.rs:207-219 dispatch_size: (ceil(w/64), h). */
void lmb_dispatch_size(uint32_t width, uint32_t height, uint32_t *groups_x,
                       uint32_t *groups_y);

/* ------------------------------------------------------------------ */
/* Textures  (textures.rs:12-243)                                      */
/* ------------------------------------------------------------------ */
/* A host-side texture collection, the counterpart of TextureLoader +
 * bind_group_from_textures.  Images are RGBA8 UNORM (textures.rs:60),
 * sampled with the default sampler: nearest, clamp-to-edge
 * (textures.rs:196).  Decoding image files is the caller's job (the
 * reference uses the `image` crate); this takes decoded RGBA8 texels. */
typedef struct lmb_textures_s *lmb_textures;
#define LMB_TEXTURE_DIFFUSE 0u
#define LMB_TEXTURE_EMISSION 1u
#define LMB_TEXTURE_ATTRIBUTES 2u

lmb_status lmb_textures_create(lmb_textures *out);
/* Appends an image to the diffuse / emission / attributes array; its
 * index in that array is written to *index. */
lmb_status lmb_textures_add(lmb_textures tex, uint32_t kind, uint32_t width,
                            uint32_t height, const uint8_t *rgba8,
                            uint32_t *index);
/* textures.rs:122-149: without a recolour texture a 1x1 white one is used. */
lmb_status lmb_textures_set_recolour(lmb_textures tex, uint32_t width,
                                     uint32_t height, const uint8_t *rgba8);
/* bindings.wesl:123-124 `bg: texture_cube<f32>`: 6 faces (+x,-x,+y,-y,+z,-z),
 * each size x size RGBA8.  Default: black. */
lmb_status lmb_textures_set_background(lmb_textures tex, uint32_t size,
                                       const uint8_t *rgba8_faces);
lmb_status lmb_textures_destroy(lmb_textures tex);

/* ------------------------------------------------------------------ */
/* Scene: geometry + materials + textures on the GPU                   */
/* Replaces the wgpu BLAS/TLAS build (examples/cornell_box/src/main.rs: */
/* 464-563) and bind group 0 (bindings.wesl:67-100) with a host SAH    */
/* build of a compressed 8-wide BVH.                                   */
/* ------------------------------------------------------------------ */
typedef struct lmb_scene_s *lmb_scene;

typedef struct lmb_scene_desc {
    const float *vertices;       /* vec4 per vertex (xyz used), lib.rs:439-454 */
    uint64_t num_vertices;
    const uint32_t *indices;     /* 3 per triangle; NULL = non-indexed */
    uint64_t num_indices;
    const uint32_t *material_idx; /* 1 per triangle (bindings.wesl:59-61) */
    const lmb_material *materials;
    uint64_t num_materials;
    lmb_textures textures;       /* may be NULL: no textures, black sky */
    uint32_t flags;              /* LMB_SCENE_* (0: opaque geometry)         */
} lmb_scene_desc;

/* Geometry without the OPAQUE flag (main.rs:469 builds opaque geometry):
 * a candidate triangle becomes the hit only when its diffuse texel's alpha,
 * sampled at the hit's texture coordinates, is not near zero (|a| >= 1e-4)
 * or its material is TRANSPARENT (general.wesl:1057-1066); the debugging
 * shaders test the alpha alone (front_face.wesl:82-91).  Costs a material
 * and a texel fetch per candidate hit; batches take opaque scenes only. */
#define LMB_SCENE_ALPHA_TEST 1u

lmb_status lmb_scene_create(const lmb_scene_desc *desc, lmb_scene *out);

/* TLAS instances (wgpu TlasInstance::new(blas, transform, custom_index,
 * mask), examples/cornell_box/src/main.rs:560-563): instance k draws
 * geometry `geometry` through `transform` (3x4 row-major object to world)
 * with materials materials[material_idx[prim]] (general.wesl:1059:
 * material_idx[instance_index].materials[primitive_index]).  The scene is
 * one BVH traversed in a single level: a top tree over the instances above
 * one subtree per instance holding its triangles in world space.  Triangles
 * are numbered instance by instance in hits, over the instances seen at
 * creation; an instance with (mask & 0xFF) == 0 is not seen (the cull mask
 * of the reference's ray queries is 0xFF).  Transforms must be finite and
 * keep the geometry within 1e30 of the origin.  As the
 * reference's vertex fetch (get_vertices, bindings.wesl:102-110, or
 * getCommittedVertexPositions) returns object-space positions, the hit's
 * tangent and normal come from the object-space triangle, and its facing is
 * decided in object space (a mirroring transform reverses the winding seen
 * in world space). */
typedef struct lmb_geometry_desc {
    const float *vertices;      /* vec4 per vertex (xyz used) */
    uint64_t num_vertices;
    const uint32_t *indices;    /* 3 per triangle; NULL = non-indexed */
    uint64_t num_indices;
    /* Procedural geometry instead of triangles (wgpu BlasGeometries::
     * AabbGeometries): 6 floats per box, min xyz then max xyz.  The
     * reference's ray query hands each candidate box to the pipeline's
     * IntersectionHandler (low_level.rs:115-147, general.wesl:1068-1076);
     * the handler this library implements is the default one
     * (default_intersection_handler.wesl), which reports no hit, so boxes
     * never take part in a closest hit.  Boxes must be finite with
     * min <= max. */
    const float *aabbs;
    uint64_t num_aabbs;
} lmb_geometry_desc;
typedef struct lmb_instance {
    uint32_t geometry;
    float transform[12];
    const uint32_t *material_idx; /* 1 per triangle of the geometry */
    uint32_t mask;
} lmb_instance;
typedef struct lmb_instanced_desc {
    const lmb_geometry_desc *geometries;
    uint32_t num_geometries;
    const lmb_instance *instances;
    uint32_t num_instances;
    const lmb_material *materials;
    uint64_t num_materials;
    lmb_textures textures;
    uint32_t flags;             /* LMB_SCENE_* */
} lmb_instanced_desc;
lmb_status lmb_scene_create_instanced(const lmb_instanced_desc *desc, lmb_scene *out);
/* Moves instances: the counterpart of writing new TlasInstance entries and
 * rebuilding the TLAS (wgpu build_acceleration_structures with the TLAS
 * only, examples/cornell_box/src/main.rs:560-563).  `instances` lists every
 * instance of the creation desc, in its order, with the same geometry; each
 * one's transform and mask are taken (material_idx is ignored: an instance
 * keeps the materials it was made with).  An instance hidden at creation
 * (mask 0) cannot be shown later.  The top tree is rebuilt on the host; on
 * the device every triangle is re-placed from its object-space copy and the
 * subtrees are refit from the leaves up, keeping the topology built at
 * creation (hits stay exact; boxes may grow looser than a fresh build's
 * under rotation).  Synchronous: waits for the device's queued work first
 * (which may be reading the scene) and returns when the scene is ready.  On
 * an error reported before any device work (invalid argument, depth beyond
 * the traversal stack) the scene is unchanged. */
lmb_status lmb_scene_update_instances(lmb_scene scene, const lmb_instance *instances,
                                      uint32_t num_instances);
lmb_status lmb_scene_destroy(lmb_scene scene);
/* BVH statistics: node count, depth, leaf triangle count, SAH cost, build ms */
typedef struct lmb_bvh_stats {
    uint64_t num_triangles;
    uint64_t num_nodes;
    uint32_t max_depth;
    uint32_t reserved;
    double sah_cost;
    double build_ms;
    uint64_t device_bytes;
} lmb_bvh_stats;
lmb_status lmb_scene_bvh_stats(lmb_scene scene, lmb_bvh_stats *stats);

/* Closest-hit ray query (replaces rayQueryInitialize/Proceed/GetCommitted,
 * general.wesl:1052-1078).  A ray is 8 floats: origin.xyz, t_min,
 * direction.xyz, t_max.  A hit is 4 x 32 bit: prim (u32, 0xFFFFFFFF = miss),
 * t, u, v (barycentric weights of vertex 1 and 2).  Ties in t resolve to the
 * lower primitive index.  Host arrays; blocks until done. */
typedef struct lmb_hit {
    uint32_t prim;
    float t;
    float u;
    float v;
} lmb_hit;
lmb_status lmb_scene_intersect(lmb_scene scene, const float *rays, uint64_t n,
                               lmb_hit *hits);
/* Same on device buffers (rays_dev: n*32 bytes, hits_dev: n*16 bytes). */
lmb_status lmb_scene_intersect_device(lmb_scene scene, const void *rays_dev,
                                      uint64_t n, void *hits_dev, void *stream);

/* Allocation-free variant of lmb_scene_intersect_device, with the same
 * results.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The traversal's scratch (fetch counters
 * and the per-thread stack spill) lives in a static 16 MiB device array that
 * each GPU the library runs on holds until the process ends; the call
 * allocates nothing.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (several streams or threads) race: serialise them.  The
 * launch uses as many query blocks as fit the array; a scene whose BVH is so
 * deep that fewer than one block per CU fits (more than 39 node levels) is
 * rejected with LMB_ERR_INVALID_ARGUMENT. */
lmb_status lmb_scene_intersect_device_noalloc(lmb_scene scene, const void *rays_dev,
                                              uint64_t n, void *hits_dev, void *stream);

/* ------------------------------------------------------------------ */
/* Tracer  (lib.rs:260-376, path_tracing/mod.rs:7-147,                 */
/* low_level.rs:77-113)                                                */
/* ------------------------------------------------------------------ */
#define LMB_QUALITY_LOW 0u    /* path_tracing/mod.rs:126 Low    */
#define LMB_QUALITY_MEDIUM 1u /* path_tracing/mod.rs:98  Medium */
#define LMB_QUALITY_HIGH 2u   /* path_tracing/mod.rs:70  High   */
/* No reference counterpart: High's spectral path with four wavelengths per
 * path over 380-780 nm (hero-wavelength sampling) and Cauchy dispersion
 * n = A + B / lambda^2 through the material's IOR-range endpoints
 * n(760 nm), n(340 nm); see lmb_material_new's ior argument. */
#define LMB_QUALITY_SPECTRAL 3u
/* src/debug/mod.rs:1-58 debugging shaders: the camera ray's closest hit
 * only (t_min / t_max from the options; samples, bounces, restir, Russian
 * roulette and guiding are ignored); the colour output holds the value, the
 * normal and albedo outputs zero, w = 1. */
#define LMB_QUALITY_DEBUG_FRONT_FACE 16u  /* debug::FrontFace (front_face.wesl):
                                             front (0, 1/(t+1), 0), back (1/(t+1), 0, 0) */
#define LMB_QUALITY_DEBUG_REFLECTANCE 17u /* debug::Reflectance (reflectance.wesl):
                                             the material's IOR pair / 5 in x, y */
#define LMB_QUALITY_DEBUG_TANGENT 18u     /* debug::Tangent (tangent.wesl):
                                             (correct_tangent(normalize(v0 - v1)) + 1) / 2 */

/* path_tracing/mod.rs:25-34 RestirTemporalMode */
#define LMB_RESTIR_DISABLED 0u
#define LMB_RESTIR_PER_SAMPLE 1u
#define LMB_RESTIR_PER_PIXEL 2u

/* Path guiding: the reference's Markov-chain (vMF) guiding with its world
 * hash grid (general.wesl:148-499), or plain BSDF sampling. */
#define LMB_GUIDING_OFF 0u
#define LMB_GUIDING_MARKOV_CHAIN 1u

typedef struct lmb_tracer_opts {
    uint32_t quality;      /* LMB_QUALITY_*                              */
    uint32_t samples;      /* SAMPLES override, default 4 (low_level.rs:101) */
    float t_min;           /* default 1e-4  (low_level.rs:102)           */
    float t_max;           /* default 1000  (low_level.rs:103)           */
    uint32_t restir;       /* LMB_RESTIR_*, default PER_PIXEL            */
    uint32_t bounces;      /* path segments per sample, reference const 5 (general.wesl:50) */
    uint32_t rr_start;     /* Russian roulette from this bounce on; >= bounces = off (reference) */
    uint32_t guiding;      /* LMB_GUIDING_*                              */
} lmb_tracer_opts;

/* Fills the reference defaults: Medium, 4 spp, 1e-4, 1000, PerPixel,
 * 5 bounces, no Russian roulette, Markov-chain guiding. */
void lmb_tracer_opts_default(lmb_tracer_opts *opts);

typedef struct lmb_tracer_s *lmb_tracer;
/* lib.rs:268-318 RayTracer::<Q>::new + create_pipeline */
lmb_status lmb_tracer_create(const lmb_tracer_opts *opts, lmb_tracer *out);
lmb_status lmb_tracer_destroy(lmb_tracer tracer);

/* ------------------------------------------------------------------ */
/* Per-frame buffers  (data_buffer/mod.rs:18-472)                      */
/* ------------------------------------------------------------------ */
/* data_buffer/mod.rs:18-27 BufferType bit flags */
#define LMB_BUFFER_MARKOV_CHAIN_SCREEN_SPACE (1u << 0)
#define LMB_BUFFER_SPATIAL_RESAMPLING (1u << 1)
#define LMB_BUFFER_MARKOV_CHAIN_WORLD_SPACE (1u << 2)
#define LMB_BUFFER_ALL 7u

/* Outputs: bindings.wesl:117-122 (rgba32float each, row-major) */
#define LMB_OUTPUT_COLOR 0u
#define LMB_OUTPUT_NORMAL 1u
#define LMB_OUTPUT_ALBEDO 2u

typedef struct lmb_data_buffers_s *lmb_data_buffers;
/* data_buffer/mod.rs:151-252 DataBuffers::new; also owns the three output
 * images and the wavefront ray queues sized for width x height.  width *
 * height must be below 2^31 (LMB_ERR_INVALID_ARGUMENT otherwise).  A resize
 * that fails - refused, or out of device memory part way - leaves every
 * buffer and the size as they were; it holds the old and the new window's
 * memory at once. */
lmb_status lmb_data_buffers_create(uint32_t width, uint32_t height,
                                   uint32_t world_size, uint32_t buffer_flags,
                                   lmb_data_buffers *out);
/* data_buffer/mod.rs:259-267 DataBuffers::resize */
lmb_status lmb_data_buffers_resize(lmb_data_buffers bufs, uint32_t width,
                                   uint32_t height);
lmb_status lmb_data_buffers_destroy(lmb_data_buffers bufs);
/* Device address and byte size of an output image (LMB_OUTPUT_*). */
lmb_status lmb_data_buffers_output(lmb_data_buffers bufs, uint32_t which,
                                   void **dev_ptr, uint64_t *bytes);
/* Copies an output image (width*height*4 floats) to host memory (blocks). */
lmb_status lmb_data_buffers_read_output(lmb_data_buffers bufs, uint32_t which,
                                        float *host);

/* The example application's frame ring and average (no counterpart in the
 * library API: examples/cornell_box/src/main.rs:377-413, 699-748 and
 * average_to_buf.wgsl:13-22).  A ring of `frames` device images; push
 * copies the buffers' colour output into the next slot (cycling), average
 * writes, per pixel, the sum over the filled slots in slot order of
 * colour / n (n = filled slots) as packed RGB f32 (3 floats per pixel, the
 * layout of the example's denoiser staging buffers). */
typedef struct lmb_frame_ring_s *lmb_frame_ring;
lmb_status lmb_frame_ring_create(uint32_t width, uint32_t height, uint32_t frames,
                                 lmb_frame_ring *out);
lmb_status lmb_frame_ring_destroy(lmb_frame_ring ring);
lmb_status lmb_frame_ring_push(lmb_frame_ring ring, lmb_data_buffers bufs, void *stream);
lmb_status lmb_frame_ring_count(lmb_frame_ring ring, uint32_t *count);
lmb_status lmb_frame_ring_average(lmb_frame_ring ring, void *stream);
/* The average's device buffer (width * height * 3 floats) and a host copy. */
lmb_status lmb_frame_ring_output(lmb_frame_ring ring, void **rgb_dev, uint64_t *bytes);
lmb_status lmb_frame_ring_read_average(lmb_frame_ring ring, float *host);
/* One output (LMB_OUTPUT_*) as packed RGB f32 into a device buffer of
 * width * height * 3 floats: the denoiser's colour / normal / albedo inputs
 * of main.rs:797-963. */
lmb_status lmb_data_buffers_pack_rgb(lmb_data_buffers bufs, uint32_t which, void *rgb_dev,
                                     void *stream);
/* Byte-level access to the persistent state for tests and checkpoints.
 * which: 0 gi_reservoirs, 1 old_gi_reservoirs, 2 info, 3 markov_chains,
 *        4 old_markov_chains, 5 world_markov_chains, 6 old_world_markov_chains */
lmb_status lmb_data_buffers_state_size(lmb_data_buffers bufs, uint32_t which,
                                       uint64_t *bytes);
lmb_status lmb_data_buffers_read_state(lmb_data_buffers bufs, uint32_t which,
                                       void *host, uint64_t bytes);
lmb_status lmb_data_buffers_write_state(lmb_data_buffers bufs, uint32_t which,
                                        const void *host, uint64_t bytes);
/* Path vertices (closest-hit queries) issued by the last lmb_trace. Blocks. */
lmb_status lmb_data_buffers_path_vertices(lmb_data_buffers bufs,
                                          uint64_t *vertices);

/* Kernel timing (no counterpart in the reference, which has no GPU timing;
 * this is its profiling hook).  When enabled, every traversal ("extend") and
 * shading launch of lmb_trace is bracketed by HIP events on its stream.
 * lmb_data_buffers_kernel_stats waits for them, reports the totals since the
 * previous call (or since enabling) and resets them. */
typedef struct lmb_kernel_stats {
    double extend_ms;          /* summed duration of the traversal launches */
    double shade_ms;           /* summed duration of the shading launches   */
    uint64_t extend_launches;
    uint64_t shade_launches;
    uint64_t path_vertices;    /* closest-hit queries issued meanwhile      */
    double cache_infer_ms;     /* neural-cache inference launches (lmb_trace_nrc) */
    double cache_train_ms;     /* neural-cache training steps (lmb_trace_nrc)     */
    uint64_t cache_infer_launches;
    uint64_t cache_train_steps;
} lmb_kernel_stats;
lmb_status lmb_data_buffers_set_profiling(lmb_data_buffers bufs, int32_t enable);
lmb_status lmb_data_buffers_kernel_stats(lmb_data_buffers bufs,
                                         lmb_kernel_stats *stats);

/* ------------------------------------------------------------------ */
/* Frame passes                                                        */
/* ------------------------------------------------------------------ */
/* general.wesl:506-610 rt_main dispatched over dispatch_size(w, h)
 * (examples/cornell_box/src/main.rs:723-731); seed_offset is the 4-byte
 * immediate (general.wesl:322). */
lmb_status lmb_trace(lmb_tracer tracer, lmb_scene scene, const lmb_camera *cam,
                     lmb_data_buffers bufs, uint32_t seed_offset, void *stream);

/* Primary-ray closest hits for every pixel (create_ray, general.wesl:
 * 1158-1173, then the ray query): prim (u32) and t per pixel, host arrays. */
lmb_status lmb_primary_hits(lmb_scene scene, const lmb_camera *cam,
                            uint32_t width, uint32_t height, float t_min,
                            float t_max, uint32_t *prim, float *t);

/* importance_sampling/temporal_resampling.wesl:37-64 */
lmb_status lmb_temporal_resample(lmb_data_buffers bufs, uint32_t seed_offset,
                                 void *stream);
/* importance_sampling/spatial_resampling.wesl:52-194 (IS_SAMPLES, IS_SPACE
 * overrides, defaults 8 / 61) writes the final colour image. */
lmb_status lmb_spatial_resample(lmb_data_buffers bufs, uint32_t is_samples,
                                uint32_t is_space, uint32_t seed_offset,
                                void *stream);
/* data_buffer/mod.rs:272-349 DataBuffers::advance_frame with
 * AdvanceOptions{confidence_cap} (data_buffer/mod.rs:442-472). */
lmb_status lmb_advance_frame(lmb_data_buffers bufs, uint32_t buffers_to_advance,
                             uint32_t confidence_cap, void *stream);

/* ------------------------------------------------------------------ */
/* Neural radiance cache                                               */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference, whose only cache is the world-grid
 * radiance estimate that trains its guiding chains (general.wesl:398-444)
 * and never ends a path.  This cache ends every path at its second vertex
 * with a network estimate of the radiance scattered there, and trains the
 * network online from a sparse set of paths traced to full depth.
 *
 * Network: 13 inputs in [0, 1] (position in the scene box, direction to the
 * previous vertex and normal mapped from [-1, 1], roughness, albedo), a
 * frequency encoding sin / cos (2^k pi x), k = 0..11, plus a constant 1,
 * then 5 hidden layers of 64 with ReLU and 3 linear outputs; bf16 weights on
 * the matrix cores with fp32 accumulation; relative L2 loss
 * (y - t)^2 / (y^2 + 0.01) (y held constant in the denominator); Adam.
 *
 * Weights (LMB_NRC_NUM_WEIGHTS fp32), each layer [in][out] row-major:
 *   [0, 20480)      layer 0: 320 x 64; rows 0..311 encoding feature
 *                   s * 24 + 2 k + c (input s, frequency k, c 0 sin / 1 cos),
 *                   row 312 the constant feature, rows 313..319 unused (0)
 *   [20480, 36864)  layers 1..4: 64 x 64 each
 *   [36864, 37056)  layer 5: 64 x 3 */
#define LMB_NRC_NUM_WEIGHTS 37056u
#define LMB_NRC_INPUTS 13u

typedef struct lmb_nrc_s *lmb_nrc;

typedef struct lmb_nrc_opts {
    float learning_rate;   /* Adam step size                        (1e-3)  */
    float beta1, beta2;    /* Adam moment decays                    (0.9, 0.99) */
    float epsilon;         /* Adam denominator offset               (1e-8)  */
    uint32_t train_paths;  /* training paths per lmb_trace_nrc      (65536) */
    uint32_t train_steps;  /* optimiser steps per lmb_trace_nrc, 1..256 (4) */
    uint32_t seed;         /* weight initialisation (He-uniform)            */
    float ema_decay;       /* inference uses an exponential moving average
                              of the weights with this decay per step; 0 = the
                              trained weights themselves               (0)  */
    uint32_t defer_specular; /* n > 0: a path whose scatters so far were all
                              specular (smooth glass, mirrors: roughness
                              below 0.1, or any at LOW quality) does not end
                              at its second vertex but at the vertex after
                              its first rough scatter, at most n vertices
                              later; 0: always the second vertex (<= 64) (2) */
    uint32_t overlap_training; /* 1: lmb_trace_nrc runs its training (a
                              training pass as lmb_trace_nrc_train's, then the
                              Adam steps) on a stream of the cache, beside
                              the frame's samples, which all end at their
                              cache vertices; 0: the training paths come from
                              the frame's first sample and the steps follow
                              the samples on the caller's stream.  Either way
                              the samples of frame k use the weights trained
                              up to frame k - 1, and the call's work is
                              ordered on the caller's stream.  With 1 and
                              train_paths = 0 (a frozen cache) the samples
                              are traced as with training, up to eight at a
                              time, and nothing else runs               (1) */
    uint32_t stratify;     /* 1: the first scatter of a sample (at the camera
                              ray's hit) draws its direction and Fresnel
                              choice from the pixel's Owen-scrambled Sobol
                              sequence, point seed_offset x samples + sample,
                              so the samples of a frame, and of frames with
                              consecutive seed offsets, are stratified over
                              the lobe; every point is uniform, so each sample
                              stays unbiased.  0: the reference's hash draws,
                              as the plain tracer's                      (1) */
} lmb_nrc_opts;

void lmb_nrc_opts_default(lmb_nrc_opts *opts);
lmb_status lmb_nrc_create(const lmb_nrc_opts *opts, lmb_nrc *out);
lmb_status lmb_nrc_destroy(lmb_nrc nrc);
/* fp32 master weights (LMB_NRC_NUM_WEIGHTS), host memory.  Setting them
 * resets the optimiser state. */
lmb_status lmb_nrc_get_weights(lmb_nrc nrc, float *weights);
lmb_status lmb_nrc_set_weights(lmb_nrc nrc, const float *weights);
/* Network outputs for n input rows (n x 13 floats) -> n x 3 floats. */
lmb_status lmb_nrc_infer(lmb_nrc nrc, const float *inputs, uint64_t n, float *outputs);
/* Gradient of the mean relative L2 loss over n samples (targets n x 3) with
 * respect to the master weights (bf16 forward / backward, fp32 sums); the
 * weights are not changed.  loss (optional) receives the mean loss. */
lmb_status lmb_nrc_gradient(lmb_nrc nrc, const float *inputs, const float *targets,
                            uint64_t n, float *grad, float *loss);
/* One Adam step on n samples. */
lmb_status lmb_nrc_train(lmb_nrc nrc, const float *inputs, const float *targets,
                         uint64_t n, float *loss);
/* One Adam step with a caller-supplied gradient (LMB_NRC_NUM_WEIGHTS). */
lmb_status lmb_nrc_apply_gradient(lmb_nrc nrc, const float *grad);
/* Replaces the options of a live cache (learning-rate schedules, stopping
 * the training with train_paths = 0); the weights and the optimiser state
 * are kept.  The seed is ignored. */
lmb_status lmb_nrc_set_opts(lmb_nrc nrc, const lmb_nrc_opts *opts);

typedef struct lmb_nrc_stats {
    uint64_t steps;          /* optimiser steps taken                     */
    uint64_t records;        /* training records of the last lmb_trace_nrc */
    uint64_t queries;        /* cache queries of its last sample (of its
                                last group of samples traced together)    */
    float loss;              /* mean loss of its training steps           */
    uint32_t reserved;
} lmb_nrc_stats;
/* Blocks until the cache's work is done. */
lmb_status lmb_nrc_stats_get(lmb_nrc nrc, lmb_nrc_stats *stats);

/* One frame of the cached tracer: for each of the tracer's samples, every
 * path is traced to its second vertex (a later one past specular scatters,
 * see defer_specular), where it ends with
 * throughput * cache(vertex).  Training (overlap_training = 0): in the first
 * sample about train_paths of them (chosen at random there) continue to full
 * depth instead and become training records (their full-path estimate also
 * goes to the image), then train_steps Adam steps on the records.  With
 * overlap_training = 1 the records come from a training pass of their own
 * traced beside the samples.  Uses the tracer's quality, samples, bounces, rr_start,
 * t_min / t_max; guiding must be off. */
lmb_status lmb_trace_nrc(lmb_tracer tracer, lmb_scene scene, const lmb_camera *cam,
                         lmb_data_buffers bufs, uint32_t seed_offset, lmb_nrc nrc,
                         void *stream);
/* No reference counterpart.  A training frame of the cached tracer: about
 * train_paths camera paths at random pixels (seeded as in lmb_trace_nrc's
 * first sample) become training records at their cache vertices and are
 * traced to full depth, then train_steps Adam steps.  Writes no image (the
 * buffers' outputs keep their contents); one sample whatever the tracer's
 * sample count.  For warming up a cache faster than with full frames. */
lmb_status lmb_trace_nrc_train(lmb_tracer tracer, lmb_scene scene, const lmb_camera *cam,
                               lmb_data_buffers bufs, uint32_t seed_offset, lmb_nrc nrc,
                               void *stream);

/* ------------------------------------------------------------------ */
/* Batched tracer                                                      */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (one rt_main dispatch per image).  Many
 * scenes and resolutions rendered by one set of persistent wavefront
 * launches: every ray carries its image index, so one traversal / shading
 * launch serves all images and balances their uneven work.  A frame (all
 * passes, samples and bounces) is captured once into a hipGraph and
 * replayed.  Plain mode (guiding off) only; each image equals what lmb_trace
 * renders for its scene and camera with the same options and seed. */
typedef struct lmb_batch_s *lmb_batch;

typedef struct lmb_batch_item {
    lmb_scene scene;      /* not owned; must outlive the batch */
    lmb_camera camera;
    uint32_t width, height;
} lmb_batch_item;

typedef struct lmb_batch_config {
    uint32_t blocks_per_cu;  /* persistent grid of every stage, per CU  (8)  */
    uint32_t refill;         /* idle lanes that make a wave fetch rays  (16) */
    uint32_t tile_w, tile_h; /* pixel tile of ray generation            (8x8) */
    uint32_t pass_pixels;    /* pixels per wavefront pass (queue size); 0 = all */
    uint32_t use_graph;      /* replay the frame as a captured hipGraph (1)  */
} lmb_batch_config;

void lmb_batch_config_default(lmb_batch_config *cfg);
/* Writes up to `cap` configurations of the default autotuner sweep (tile
 * shape x grid size x refill x pass size) and returns their number. */
uint32_t lmb_batch_default_sweep(lmb_batch_config *out, uint32_t cap);
lmb_status lmb_batch_create(const lmb_batch_item *items, uint32_t n,
                            const lmb_tracer_opts *opts, const lmb_batch_config *cfg,
                            lmb_batch *out);
lmb_status lmb_batch_destroy(lmb_batch batch);
lmb_status lmb_batch_set_config(lmb_batch batch, const lmb_batch_config *cfg);
lmb_status lmb_batch_get_config(lmb_batch batch, lmb_batch_config *cfg);
lmb_status lmb_batch_set_camera(lmb_batch batch, uint32_t item, const lmb_camera *cam);
/* One frame of every image; seed_offset as lmb_trace's. */
lmb_status lmb_batch_trace(lmb_batch batch, uint32_t seed_offset, void *stream);
/* Colour image of item `item` (width * height * 4 floats), host memory. */
lmb_status lmb_batch_read_output(lmb_batch batch, uint32_t item, float *host);
/* Path vertices of the last frame (blocks). */
lmb_status lmb_batch_path_vertices(lmb_batch batch, uint64_t *vertices);
/* Times `frames` frames of each candidate (HIP events, after one warm-up
 * frame that also builds the graph), writes the mean ms per frame of
 * candidate i to ms[i], applies and returns the fastest. */
lmb_status lmb_batch_autotune(lmb_batch batch, const lmb_batch_config *candidates,
                              uint32_t n, uint32_t frames, lmb_batch_config *best,
                              double *ms);

/* Waits for all work on `stream` (NULL = device-wide). */
lmb_status lmb_sync(void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LAMBENT_H */

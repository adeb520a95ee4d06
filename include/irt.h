/*
 * irt.h — C ABI of the MI355X (gfx950) per-pixel lighting + denoising path.
 *
 * Replaces the wgpu compute passes that src/renderer.rs schedules for the
 * ReSTIR direct/VPL reservoirs, the SDF soft-shadow resolve and SVGF
 * (reference: src/renderer.rs:1873-1941, src/svgf.rs:140-261).  The caller
 * (renderer.rs, or any host) keeps packing the same POD records it uploads to
 * wgpu today and hands them over through plain pointers; the library owns the
 * device memory and all cross-frame state (reservoirs, SVGF moments/feedback/
 * history, previous G-buffer, VPL archive ring).
 *
 * Conventions
 *   - Every entry point returns IRT_OK (0) or a negative irt_status.  The text
 *     of the last failure of a context is returned by irt_last_error(); errors
 *     raised before a context exists go to irt_global_error().
 *   - Matrices are 4x4 column-major float (glam Mat4 memory order).
 *   - Pixel buffers exchanged through the ABI are row-major, id = y*W + x, in
 *     the reference's texture formats: Depth32Float = float, Rgba16F = four
 *     IEEE binary16 bit patterns (uint16_t), Rg16F = two, Rgba32F = four float.
 *   - Calls on one context are serialised on the context's HIP stream and are
 *     asynchronous unless stated; irt_read* and irt_sync block.  Different
 *     contexts are independent (one per view / session).
 *   - No pointer passed in is retained after a call returns.  Device
 *     pointers handed to irt_set_gbuffer_device / irt_set_vpls_device are
 *     read by work the call enqueues: keep them valid until that work is
 *     done (irt_sync, or the frame's completion).
 */
#ifndef IRT_H
#define IRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IRT_ABI_VERSION 1u

/* ---------------------------------------------------------------- status */
typedef enum irt_status {
  IRT_OK = 0,
  IRT_ERR_INVALID_ARGUMENT = -1, /* null pointer, wrong size, bad enum       */
  IRT_ERR_INVALID_SETTINGS = -2, /* Settings::validate failure (same text)   */
  IRT_ERR_LIMIT = -3,            /* adapter-limit style checks               */
  IRT_ERR_STATE = -4,            /* call order (e.g. no scene/frame yet)     */
  IRT_ERR_DEVICE = -5,           /* HIP runtime failure or no gfx950 device  */
  IRT_ERR_OUT_OF_MEMORY = -6
} irt_status;

/* ------------------------------------------------------------- POD records
 * Byte-for-byte the layouts of src/renderer.rs:8-66 / src/shaders/common.wgsl:3-16
 * (sizes asserted by the reference at src/renderer.rs:2254-2261). */

/* Frame uniform, 384 B (renderer.rs:51-66, packing at renderer.rs:1687-1735). */
typedef struct irt_frame {
  float vp[16];
  float inv_vp[16];
  float previous_vp[16];
  float previous_inv_vp[16];
  float camera[4];     /* xyz, 0 */
  uint32_t size[4];    /* width, height, frame_index, scene generation */
  uint32_t counts[4];  /* lights, vpl_count, vpl_max_bounces, objects */
  uint32_t restir[4];  /* direct candidates, VPL candidates, spatial neighbours, spatial radius */
  float temporal[4];   /* history weight, depth threshold, normal threshold, SVGF max history */
  float sdf[4];        /* max steps, hit epsilon, soft-shadow k, 0 */
  float rr[4];         /* roulette min, roulette max, ReSTIR max history, 0 */
  uint32_t flags[4];   /* history enabled, previous generation, debug mode, 0 */
} irt_frame;

/* GpuObject, 208 B (renderer.rs:8-15, built by objects() renderer.rs:359-384). */
typedef struct irt_object {
  float model[16];
  float previous[16];
  float inverse[16];
  uint32_t info[4]; /* mesh index, material index, bits(scale), bits(voxel diagonal * scale) */
} irt_object;

/* GpuMesh, 128 B (renderer.rs:16-28, filled at renderer.rs:846-885). */
typedef struct irt_mesh {
  float lo[4];               /* SDF grid bounds (node-centred samples) */
  float hi[4];
  uint32_t geometry[4];      /* vertex offset, vertex count, index offset, index count */
  uint32_t field[4];         /* atlas origin x, y, z, resolution */
  float grid[4];             /* 1/spacing xyz, |spacing| */
  float geometry_lo[4];      /* vertex AABB xyz, w = 1 if certified box */
  float geometry_hi[4];      /* vertex AABB xyz, w = 1/Lipschitz bound */
  float enclosing_sphere[4]; /* centre xyz, radius */
} irt_mesh;

/* GpuLight, 48 B (renderer.rs:37-43, renderer.rs:1743-1755). */
typedef struct irt_light {
  float position_strength[4]; /* xyz, W/sr */
  float direction_angle[4];   /* spot axis xyz, full cone angle (rad) */
  float color_type[4];        /* rgb, w = 1 for spot */
} irt_light;

/* BVH node, 48 B, preorder with escape links (renderer.rs:44-50, 385-455). */
typedef struct irt_node {
  float lo[4];
  float hi[4];
  uint32_t info[4]; /* leaf: object id / inner: left, inner: right, is_leaf, escape index */
} irt_node;

/* Virtual point light, 48 B (common.wgsl:15). */
typedef struct irt_vpl {
  float position[4];
  float normal[4];
  float flux[4];
} irt_vpl;

/* Reservoir, 32 B (common.wgsl:16). */
typedef struct irt_reservoir {
  uint32_t id;
  uint32_t source; /* frame index the sample was drawn in */
  uint32_t count;  /* M */
  uint32_t age;
  float sum;             /* w_sum */
  float selected_target; /* p-hat of the selected sample */
  float pdf;
  uint32_t generation;
} irt_reservoir;

/* Mirror of the reference Settings record (src/settings.rs:3-28). */
typedef struct irt_settings {
  uint32_t render_width;
  uint32_t render_height;
  uint32_t vpl_count;
  uint32_t vpl_max_bounces;
  float vpl_rr_min_survival;
  float vpl_rr_max_survival;
  uint32_t restir_direct_candidates;
  uint32_t restir_vpl_candidates;
  uint32_t restir_spatial_neighbors;
  uint32_t restir_spatial_radius;
  float depth_threshold;
  float normal_threshold;
  float temporal_history_weight;
  uint32_t temporal_max_history;
  uint32_t restir_max_history;
  uint32_t svgf_iterations;
  uint32_t sdf_max_steps;
  float sdf_hit_epsilon;
  float sdf_soft_shadow_k;
  float target_frame_ms;
} irt_settings;

/* Storage profile of a context. */
typedef enum irt_precision {
  /* FP32 arithmetic; every buffer the reference keeps in FP32 stays FP32 and
   * the G-buffer is widened to FP32.  Values that the reference writes to an
   * Rgba16F target are rounded to binary16 at the same point, so results
   * follow the reference's storage rounding. */
  IRT_PRECISION_FP32 = 0,
  /* binary16 G-buffer / radiance / SVGF radiance / direct lighting; SVGF
   * moments as binary16 first moment and age with an FP32 second moment
   * (8 B); binary16 SDF samples rounded toward -inf; FP32 arithmetic,
   * reservoir W/M sums and march.  The march contracts multiply-adds and
   * uses the hardware division and square root; the fused schedule's SVGF
   * taps are fused multiply-adds with one exponential per tap.  The
   * reservoir passes keep the FP32 profile's exact arithmetic. */
  IRT_PRECISION_MIXED = 1
} irt_precision;

/* Schedules of irt_encode / irt_lighting / irt_svgf. */
typedef enum irt_schedule {
  IRT_SCHEDULE_PASSES = 0, /* one kernel per reference pass (11 + copies) */
  IRT_SCHEDULE_FUSED = 1   /* fused kernels: reservoirs A/B, SVGF chain */
} irt_schedule;

/* Single passes, in the reference's entry-point order (renderer.rs:1477-1486, svgf.rs:100). */
typedef enum irt_pass_id {
  IRT_PASS_CANDIDATES_DIRECT = 0,
  IRT_PASS_CANDIDATES_INDIRECT = 1,
  IRT_PASS_TEMPORAL_DIRECT = 2,
  IRT_PASS_TEMPORAL_INDIRECT = 3,
  IRT_PASS_SPATIAL_DIRECT = 4,
  IRT_PASS_SPATIAL_INDIRECT = 5,
  IRT_PASS_RESOLVE_DIRECT = 6,
  IRT_PASS_RESOLVE_INDIRECT = 7,
  IRT_PASS_SVGF_TEMPORAL = 8,
  IRT_PASS_SVGF_VARIANCE = 9,
  IRT_PASS_SVGF_ATROUS = 10, /* stride argument of irt_pass_ex selects 1<<i */
  IRT_PASS_COUNT = 11
} irt_pass_id;

/* Buffers visible to irt_read / irt_write (the reference's test hooks,
 * src/renderer.rs:457-712).  Formats are the reference's. */
typedef enum irt_buffer_id {
  IRT_BUF_DIRECT0 = 0,      /* irt_reservoir[W*H] */
  IRT_BUF_DIRECT1 = 1,
  IRT_BUF_DIRECT2 = 2,      /* cross-frame direct reservoirs */
  IRT_BUF_INDIRECT0 = 3,
  IRT_BUF_INDIRECT1 = 4,
  IRT_BUF_INDIRECT2 = 5,    /* cross-frame indirect reservoirs */
  IRT_BUF_DIRECT_LIGHTING = 6, /* Rgba32F direct HDR */
  IRT_BUF_LIGHTING = 7,     /* Rgba16F noisy HDR (resolve output, SVGF input) */
  IRT_BUF_HISTORY = 8,      /* Rgba16F denoised histories[current] */
  IRT_BUF_MOMENTS = 9,      /* Rgba32F moments[current] (m1, m2, age, 0) */
  IRT_BUF_MOMENTS_OLD = 10, /* Rgba32F moments[old] */
  IRT_BUF_FEEDBACK = 11,    /* Rgba16F feedback[current] */
  IRT_BUF_FEEDBACK_OLD = 12,/* Rgba16F feedback[old] */
  IRT_BUF_TEMPORAL = 13,    /* Rgba16F SVGF temporal radiance */
  IRT_BUF_DEPTH = 14,       /* Depth32Float */
  IRT_BUF_NORMAL = 15,      /* Rgba16F */
  IRT_BUF_COLOR = 16,       /* Rgba16F */
  IRT_BUF_PARAMS = 17,      /* Rg16F */
  IRT_BUF_MOTION = 18,      /* Rg16F */
  IRT_BUF_PREVIOUS_DEPTH = 19,  /* Depth32Float */
  IRT_BUF_PREVIOUS_NORMAL = 20, /* Rgba16F */
  IRT_BUF_VPL_ARCHIVE = 21, /* irt_vpl[(restir_max_history+1)*vpl_count] */
  IRT_BUF_SCRATCH0 = 22,    /* Rgba16F SVGF scratch[0] (variance output) */
  IRT_BUF_SCRATCH1 = 23,    /* Rgba16F SVGF scratch[1] */
  IRT_BUF_COUNT = 24
} irt_buffer_id;

typedef struct irt_ctx irt_ctx;

/* Version of this ABI (IRT_ABI_VERSION). */
uint32_t irt_abi_version(void);

/* Text of the last failure that happened without a context (thread-local). */
const char* irt_global_error(void);

/* Settings::validate (src/settings.rs:42-106): same checks, same messages
 * ("invalid render size", "invalid VPL budget", ...). */
int irt_settings_validate(const irt_settings* settings);

/* Renderer::new's hot-path part (src/renderer.rs:713-766, 1205-1333,
 * src/svgf.rs:16-138): validates settings and the reservoir/VPL-archive
 * budget checks of renderer.rs:750-766, allocates all per-view state on the
 * current HIP device.  `precision` is an irt_precision. */
int irt_create(const irt_settings* settings, uint32_t precision, irt_ctx** out);
void irt_destroy(irt_ctx* ctx);
const char* irt_last_error(const irt_ctx* ctx);

/* Selects IRT_SCHEDULE_PASSES or IRT_SCHEDULE_FUSED for irt_encode/irt_lighting/irt_svgf. */
int irt_set_schedule(irt_ctx* ctx, uint32_t schedule);

/* Scene (group 0 bindings, renderer.rs:1115-1204).  The SDF atlas is the
 * reference's R32Float 3-D texture, x fastest: texel(x,y,z) = texels[x + w*(y + h*z)]
 * (renderer.rs:776-845). */
int irt_set_sdf_atlas(irt_ctx* ctx, const float* texels, uint32_t w, uint32_t h, uint32_t d);
int irt_set_meshes(irt_ctx* ctx, const irt_mesh* meshes, uint32_t count);
/* objects() and hierarchy() uploads of renderer.rs:1736-1782 (every frame in the reference). */
int irt_set_objects(irt_ctx* ctx, const irt_object* objects, uint32_t count,
                    const irt_node* nodes, uint32_t node_count);
int irt_set_lights(irt_ctx* ctx, const irt_light* lights, uint32_t count);

/* Frame uniform (renderer.rs:1687-1737).  size[0..1] must equal the context's size. */
int irt_set_frame(irt_ctx* ctx, const irt_frame* frame);

/* G-buffer of the frame being rendered, host memory, reference formats
 * (renderer.rs:1207-1262): depth float[W*H]; normal/color uint16[4*W*H]
 * binary16; params/motion uint16[2*W*H]. */
int irt_set_gbuffer(irt_ctx* ctx, const float* depth, const uint16_t* normal,
                    const uint16_t* color, const uint16_t* params, const uint16_t* motion);
/* Same, from device memory already resident in HBM (read on the context stream). */
int irt_set_gbuffer_device(irt_ctx* ctx, const float* depth, const uint16_t* normal,
                           const uint16_t* color, const uint16_t* params, const uint16_t* motion);

/* This frame's VPL pool (vpl.wgsl output).  Written to archive ring slot
 * frame_index % (restir_max_history + 1) — the copy of renderer.rs:1873-1884.
 * `count` must equal settings.vpl_count.  Call after irt_set_frame. */
int irt_set_vpls(irt_ctx* ctx, const irt_vpl* vpls, uint32_t count);
int irt_set_vpls_device(irt_ctx* ctx, const irt_vpl* vpls, uint32_t count);

/* ------------------------------------------------ VPL pool generation
 * The producer of the VPL pool (vpl.wgsl:2-47).  The reference traces its
 * light paths with hardware ray queries against a TLAS of triangle BLASes
 * (renderer.rs:1061-1106, 1786-1797); gfx950 has no ray-tracing units, so
 * the library builds a bounding-volume hierarchy per mesh on the host and
 * walks it in software.  The instance level is the hierarchy of
 * irt_set_objects: its leaf boxes must contain each object's triangles (the
 * exported SDF boxes do, they pad the vertex bounds). */

/* Vertex, 64 B (common.wgsl:13; vertex stride 64 at renderer.rs:1096). */
typedef struct irt_vertex {
  float position[4];
  float normal[4];
  float tangent[4]; /* xyz, w = handedness */
  float uv[4];
} irt_vertex;

/* GpuMaterial, 64 B (renderer.rs:31-36, filled at renderer.rs:951-981). */
typedef struct irt_material {
  float color[4];      /* base colour rgba */
  float params[4];     /* roughness, metallic, normal strength, 1 if the base colour texture is sRGB */
  int32_t textures[4]; /* texture layer of base colour, roughness, metallic, normal map; -1 = none */
  uint32_t channels[4];/* texel channel read for roughness (y) and metallic (z) */
} irt_material;

/* Triangle geometry and materials (group 0 bindings 3, 5, 6 of
 * renderer.rs:1115-1204).  Each mesh of irt_set_meshes finds its vertices
 * and its mesh-local indices through irt_mesh.geometry (vertex offset and
 * count, index offset and count); objects name their material in info[1].
 * Builds the per-mesh hierarchies (the BLAS build of renderer.rs:1061-1106)
 * and uploads everything; blocking.  Textured materials name layers of the
 * texture array of irt_set_textures.  Call after irt_set_meshes; a later
 * irt_set_meshes drops the geometry. */
int irt_set_geometry(irt_ctx* ctx, const irt_vertex* vertices, uint32_t vertex_count, const uint32_t* indices,
                     uint32_t index_count, const irt_material* materials, uint32_t material_count);

/* Material textures (group 0 bindings 9, 10, 13 of renderer.rs:1115-1204,
 * uploaded at renderer.rs:951-1046): `layers` square RGBA8 layers of
 * size x size texels, layer-major, rows top to bottom, already resized as
 * Renderer::new resizes them (renderer.rs:230-262).  Sampled as the
 * reference's views and sampler: the linear view for roughness, metalness
 * and normal maps and for base colours without the sRGB flag (params.w),
 * the sRGB view for the others; repeat addressing, bilinear filtering,
 * level 0 (material_at, common.wgsl:50-66).  Blocking. */
int irt_set_textures(irt_ctx* ctx, const uint8_t* rgba8, uint32_t size, uint32_t layers);

/* The VPL pass (vpl.wgsl:2-47, dispatched at renderer.rs:1863-1872):
 * frame.counts.y light paths of up to frame.counts.z surface hits each,
 * seeded from the frame index; path i writes pool entry i.  The pool goes
 * to archive slot frame_index % (restir_max_history + 1), as irt_set_vpls
 * does (the copy of renderer.rs:1873-1884); it replaces irt_set_vpls for the
 * frame.  Uses the frame's objects, hierarchy and lights.  Enqueued on
 * irt_stream; call after irt_set_frame and irt_set_geometry. */
int irt_generate_vpls(irt_ctx* ctx);

/* The G-buffer pass (gbuffer.wgsl:1-12; the depth pre-pass and G-buffer
 * pass of renderer.rs:1801-1862 with the pipelines of renderer.rs:1370-1415:
 * back faces culled, counter-clockwise front faces, depth Less then Equal,
 * targets cleared to 0 and depth to 1): rasterises the frame's objects
 * (model and previous matrices of irt_set_objects, triangles of
 * irt_set_geometry, frame.vp and frame.previous_vp) into the current
 * G-buffer, in place of irt_set_gbuffer.  Values are rounded as the
 * reference's Depth32Float / Rgba16F / Rg16F targets hold them.  Enqueued
 * on irt_stream; call after irt_set_frame. */
int irt_render_gbuffer(irt_ctx* ctx);

/* Presentation (present.wgsl:1-8 into the Rgba8UnormSrgb capture target of
 * renderer.rs:1300-1306, drawn at renderer.rs:1961; read back as
 * Renderer::capture does, renderer.rs:2109-2163): the last frame's denoised
 * history as RGBA8, row-major W*H*4 bytes: each rgb channel clamped to
 * [0, 1], sRGB-encoded (IEC 61966-2-1) and rounded to the nearest code
 * (irt_host_srgb_thresholds), alpha 255.  irt_present writes host memory
 * and blocks until the frame is done; irt_present_device writes device
 * memory, enqueued after the frame's SVGF on the stream that runs it
 * (irt_stream_svgf under IRT_SCHEDULE_FUSED, irt_stream otherwise). */
int irt_present(irt_ctx* ctx, uint8_t* rgba8);
int irt_present_device(irt_ctx* ctx, void* rgba8);

/* The six reservoir passes and the two resolves (renderer.rs:1889-1928). */
int irt_lighting(irt_ctx* ctx);
/* Svgf::encode (svgf.rs:140-261): temporal, variance, `iterations` a-trous,
 * feedback and final copy; debug >= 3 copies the noisy image (svgf.rs:154-161). */
int irt_svgf(irt_ctx* ctx, uint32_t iterations, uint32_t debug);
/* End-of-frame history (renderer.rs:1946-1960): current depth/normal become previous. */
int irt_end_frame(irt_ctx* ctx);
/* irt_lighting + irt_svgf(settings.svgf_iterations, flags.z) + irt_end_frame. */
int irt_encode(irt_ctx* ctx);

/* One reference pass with the reference's bindings for the frame schedule
 * (pass_group at renderer.rs:1617-1674; svgf.rs:163-253).  For
 * IRT_PASS_SVGF_ATROUS `arg` is the iteration index i (stride 1<<i, reads
 * scratch[i%2], writes scratch[1-i%2]); ignored otherwise. */
int irt_pass(irt_ctx* ctx, uint32_t pass, uint32_t arg);

/* Read-back / injection of internal state in reference formats (blocking).
 * `bytes` must equal the buffer's size in that format. */
size_t irt_buffer_bytes(const irt_ctx* ctx, uint32_t buffer);
int irt_read(irt_ctx* ctx, uint32_t buffer, void* dst, size_t bytes);
int irt_write(irt_ctx* ctx, uint32_t buffer, const void* src, size_t bytes);

/* Block until every call enqueued on the context has finished. */
int irt_sync(irt_ctx* ctx);

/* HIP streams of the context (hipStream_t as void*): irt_stream carries the
 * ingest, lighting and read-backs; under IRT_SCHEDULE_FUSED, SVGF runs on
 * irt_stream_svgf after everything irt_stream has queued, so it overlaps the
 * next frame's lighting (a frame's last work is on irt_stream_svgf). */
void* irt_stream(irt_ctx* ctx);
void* irt_stream_svgf(irt_ctx* ctx);

/* GPU time of every pass recorded since timing was enabled or last read
 * (read_gpu_time, renderer.rs:2066-2108), one entry per launch in launch
 * order; reading clears the record.  Names: candidates_direct,
 * candidates_indirect, temporal_direct, temporal_indirect, spatial_direct,
 * spatial_indirect, visibility_direct, visibility_indirect, svgf; in the fused
 * schedule the kernels report under reservoirs_a, spatial_setup, march,
 * resolve_combine (reservoirs_b_resolve for scenes beyond the LDS staging
 * limits), svgf_temporal, svgf_variance, svgf_atrous0..4.  Returns the
 * number of entries written. */
int irt_enable_timing(irt_ctx* ctx, int enable);
int irt_pass_times(irt_ctx* ctx, char names[][32], float* ms, int capacity);
/* Timing of one pass only: like irt_enable_timing(ctx, 1), but records just
 * the launches reported under `pass` (for example "march"), two events per
 * launch, so a timed run carries almost no timing work; NULL or "" records
 * every pass.  No counterpart in the reference (its timestamp queries cover
 * every pass); for callers that time one kernel inside a throughput run. */
int irt_enable_timing_for(irt_ctx* ctx, const char* pass);

/* Event counters of the next irt_encode / irt_lighting (instrumented
 * kernels; slower).  counts[] order: brdf, node tests, box object
 * evaluations, grid object evaluations, march steps, SVGF geometry taps,
 * march calls.  Reset on enable. */
#define IRT_COUNTER_COUNT 7
int irt_enable_counters(irt_ctx* ctx, int enable);
int irt_read_counters(irt_ctx* ctx, uint64_t counts[IRT_COUNTER_COUNT]);

/* Device-to-device copy of src's scene (SDF atlas, meshes, objects,
 * hierarchy, lights, candidate grid, and the triangle geometry and
 * textures of irt_set_geometry / irt_set_textures when src has them) into dst; both contexts must use the
 * same precision profile.  Synchronises both contexts' streams.  The
 * reference uploads one scene per Renderer (renderer.rs:713-817); this lets
 * a server give each session its own copy of a prepared scene without a
 * host round trip. */
int irt_copy_scene(irt_ctx* dst, irt_ctx* src);

/* Device bytes held by the context (per-view state + scene). */
size_t irt_device_bytes(const irt_ctx* ctx);
/* The scene part of irt_device_bytes (atlas, records, candidate grid). */
size_t irt_scene_bytes(const irt_ctx* ctx);
/* Per-view state bytes irt_create would allocate for these settings. */
size_t irt_estimate_bytes(const irt_settings* settings, uint32_t precision);

/* ------------------------------------------------ batched views
 * One launch per pass over the tiles of many views (contexts).  The
 * reference renders one view per Renderer::encode (renderer.rs:1675-1984);
 * a batch runs that frame for every view at once with the fused schedule's
 * arithmetic (results equal each view's own irt_encode under
 * IRT_SCHEDULE_FUSED).  Per frame: set each view's frame, G-buffer and VPLs
 * as usual, then call irt_batch_encode instead of irt_encode per view.
 * Views must share the precision profile and the device, hold at most 64
 * objects and 64 meshes, and render debug modes 0..2; sizes and SVGF
 * iteration counts may differ.  The batch's stream waits for the views'
 * streams and the views' streams wait for the batch. */
typedef struct irt_batch irt_batch;
int irt_batch_create(irt_ctx* const* views, uint32_t count, irt_batch** out);
void irt_batch_destroy(irt_batch* batch);
const char* irt_batch_last_error(const irt_batch* batch);
/* Replace the batch's views (1..255 distinct contexts); a no-op when the
 * list is unchanged.  Drops a captured graph. */
int irt_batch_set_views(irt_batch* batch, irt_ctx* const* views, uint32_t count);
int irt_batch_encode(irt_batch* batch);
/* Pass kinds: 0 guide, 1 reservoirs A, 2 spatial set-up, 3 resolve combine,
 * 4 SVGF temporal, 5 SVGF variance, 6 a-trous, 7 history copy.
 * Tile shapes: 0 = 8x8, 1 = 16x8, 2 = 16x16 pixels. */
int irt_batch_set_tile(irt_batch* batch, uint32_t kind, uint32_t shape);
/* Times one frame per tile shape (`repeats` times each) and keeps the
 * fastest shape per pass kind; renders frames from the views' current
 * inputs. */
int irt_batch_autotune(irt_batch* batch, uint32_t repeats, uint32_t shapes_out[8]);
/* Capture the launch sequence in a HIP graph on first encode and replay it
 * (per-frame arguments live in device memory uploaded before the replay). */
int irt_batch_use_graph(irt_batch* batch, int enable);
int irt_batch_sync(irt_batch* batch);
/* The batch's two streams: a step's lighting passes run on the first, its
 * SVGF passes on the second (where they overlap the next step's lighting);
 * a step's last work is on the second. */
void* irt_batch_stream(irt_batch* batch);
void* irt_batch_stream_svgf(irt_batch* batch);
/* One encode with per-pass GPU times (ms[0..7] by pass kind, ms[8] march). */
int irt_batch_encode_timed(irt_batch* batch, float ms[9]);
int irt_batch_pass_times(irt_batch* batch, float ms[9]);

/* ------------------------------------------------ multi-session frame server
 * Sessions keep their reservoirs, SVGF history, previous G-buffer and VPL
 * ring resident on the device, each with its own copy of a prepared scene.
 * Admission: a free slot and room in the HBM budget (per-view state +
 * scene copy).  Clients submit frame requests whose G-buffer and VPL pool
 * are device pointers (valid until the frame completes); irt_server_step
 * renders the oldest pending request of every session that has one in one
 * batched frame, so requests arriving while earlier steps run join the next
 * step.  A session's frames run in submission order and its results are the
 * same whichever sessions share its steps.  The server keeps the frame
 * bookkeeping of Renderer::encode (renderer.rs:1687-1735, 1973-1982) per
 * session.  Not thread-safe: one host thread drives a server. */
typedef struct irt_server irt_server;
typedef struct irt_server_config {
  uint32_t max_sessions;     /* 1..255 */
  uint32_t precision;        /* irt_precision of every session */
  uint64_t hbm_budget_bytes; /* device bytes for sessions: state + scene copies */
  uint32_t max_in_flight;    /* batched steps queued on the device, 1..8 */
  uint32_t use_graph;        /* replay steps of unchanged membership from a HIP graph */
  uint32_t tile_shapes[8];   /* per batch pass kind (irt_batch_set_tile) */
} irt_server_config;
typedef struct irt_frame_request {
  float vp[16];              /* view-projection, column-major */
  float camera[3];
  uint32_t generation;       /* scene generation (renderer.rs:1706) */
  uint32_t light_count;
  uint32_t debug_mode;       /* 0..2 */
  uint32_t vpl_count;        /* must equal the session's settings.vpl_count */
  uint32_t reserved;
  const float* depth;        /* device pointers, reference formats (irt_set_gbuffer); all five */
  const uint16_t* normal;    /* null: the G-buffer pass renders them (irt_render_gbuffer) from */
  const uint16_t* color;     /* the session's scene, which then needs irt_set_geometry */
  const uint16_t* params;
  const uint16_t* motion;
  const irt_vpl* vpls;       /* null: the VPL pass generates the pool (irt_generate_vpls) */
  uint64_t tag;              /* returned in the completion */
} irt_frame_request;
typedef struct irt_completion {
  uint64_t tag;
  uint32_t session;
  uint32_t frame_index;
  uint64_t pixels;
  double submit_s;           /* irt_server_now() clock */
  double launch_s;
  double done_s;             /* when the server saw the step finish */
  float gpu_ms;              /* GPU time of the step that rendered the frame */
  uint32_t batch_views;      /* sessions in that step */
} irt_completion;
typedef struct irt_server_status {
  uint32_t open_sessions;
  uint32_t idle_contexts;
  uint32_t steps_in_flight;
  uint32_t pending_requests;
  uint64_t device_bytes;
  uint64_t budget_bytes;
} irt_server_status;

double irt_server_now(void);
int irt_server_create(const irt_server_config* config, irt_server** out);
void irt_server_destroy(irt_server* server);
const char* irt_server_last_error(const irt_server* server);
/* Admit a session rendering `scene`'s scene with `settings`.  IRT_ERR_LIMIT
 * when no slot is free or the HBM budget would be exceeded. */
int irt_server_open(irt_server* server, const irt_settings* settings, irt_ctx* scene, uint32_t* session);
/* The session's context (read-backs); NULL for a closed session. */
irt_ctx* irt_server_context(irt_server* server, uint32_t session);
/* Ends a session; its pending requests are dropped, frames in flight finish. */
int irt_server_close(irt_server* server, uint32_t session);
int irt_server_submit(irt_server* server, uint32_t session, const irt_frame_request* request);
/* Launches one batched step; returns the number of frames in it (0: nothing
 * pending, or max_in_flight steps already queued). */
int irt_server_step(irt_server* server);
/* Moves up to `capacity` finished frames into `out`; with `wait`, blocks on
 * the oldest step when nothing has finished yet.  Returns the count. */
int irt_server_poll(irt_server* server, irt_completion* out, uint32_t capacity, int wait);
int irt_server_sync(irt_server* server);
int irt_server_stats(const irt_server* server, irt_server_status* out);

/* Clears every per-view buffer and the frame state, as after irt_create
 * (the scene is kept). */
int irt_reset(irt_ctx* ctx);

/* Stream-ordering audit (debug runs).  No counterpart in the reference, whose
 * single wgpu queue orders every pass; it is for checking this library's two
 * streams per context (lighting, and SVGF of the fused schedule).  With
 * IRT_STREAM_AUDIT=1 in the environment when a context is created, every
 * operation the context enqueues declares the buffers it reads and writes,
 * and a read after a write, write after a read or write after a write on the
 * other stream with no join, event wait or host synchronisation between them
 * is recorded.  irt_audit_report returns the number of such violations and
 * writes their text (one per line, then a line with the number of audited
 * operations on each stream) into out; IRT_ERR_STATE when the audit is
 * off or the context has joined a batch (batched steps run on the batch's
 * streams, which the audit does not model).  irt_host_audit_selftest runs
 * scripted schedules with known answers through the audit, without a GPU:
 * 0 when every case is judged right, else the number of the first wrong one. */
int irt_audit_report(irt_ctx* ctx, char* out, size_t capacity);
int irt_host_audit_selftest(void);

/* Self-check of the kernels' shortcuts for correctly rounded division and
 * square root (no counterpart in the reference: it guards the exact FP32
 * contract of this library on the device it runs on).  On the current
 * device, compares for every binary32 input of each domain, and writes the
 * number of inputs whose result bits differ (all 0 on a correct device):
 *   mismatches[0]: the reciprocal shortcut against 1.0f / b, for b 0,
 *                  infinite, NaN or normal with |b| in [2^-100, 2^100];
 *   mismatches[1]: the division by pi against a / pi, for a 0, infinite,
 *                  NaN or normal with |a| in [2^-79, 2^126];
 *   mismatches[2]: the square root shortcut against sqrtf(x), for x NaN or
 *                  in [2^-96, +inf];
 *   mismatches[3]: the ranged quotient against a / b, for every a, each
 *                  over one divisor b drawn from a's hash with |b| in
 *                  [2^-100, 2^100), compared where a is 0, infinite or NaN
 *                  or where |a| >= 2^-79 and a / b's binary exponent is
 *                  within [-125, 126].
 * About a second of GPU time.  IRT_ERR_DEVICE (irt_global_error) on a HIP
 * failure. */
int irt_check_fast_division(uint64_t mismatches[4]);

/* ------------------------------------------------ host precompute (a23)
 * The scene-side helpers renderer.rs runs on the CPU before upload. */

/* is_box_geometry (renderer.rs:287-329). positions: xyz stride `stride_floats`. */
int irt_host_is_box_geometry(const float* positions, uint32_t vertex_count, uint32_t stride_floats,
                             const uint32_t* indices, uint32_t index_count,
                             const float lo[3], const float hi[3]);
/* field_lipschitz_bound (renderer.rs:330-358). */
float irt_host_field_lipschitz_bound(const float* samples, uint32_t resolution, const float spacing[3]);
/* hierarchy (renderer.rs:385-455): world AABBs of the objects' SDF boxes
 * (lo/hi as xyz triples) → preorder nodes with escape links.  `nodes` must
 * hold 2*count-1 records.  Returns the node count. */
int irt_host_hierarchy(const float* box_lo, const float* box_hi, uint32_t count, irt_node* nodes);
/* Object-space SDF box → world AABB of its 8 transformed corners (renderer.rs:387-424). */
void irt_host_transform_box(const float model[16], const float lo[3], const float hi[3],
                            float out_lo[3], float out_hi[3]);
/* SDF atlas placement of renderer.rs:778-817: grid (gx, gy, gz); mesh m's tile
 * origin is ((m%gx)*tile, ((m/gx)%gy)*tile, (m/(gx*gy))*tile). */
void irt_host_atlas_grid(uint32_t mesh_count, uint32_t tile, uint32_t grid[3]);
/* Frame packing of Renderer::encode (renderer.rs:1687-1735).  inv_vp and
 * previous_inv_vp are the inverses of vp / previous_vp; history_enabled is
 * flags.x (frame_index > 0 && previous debug mode == debug mode). */
int irt_host_pack_frame(const irt_settings* settings, const float vp[16], const float previous_vp[16],
                        const float camera[3], uint32_t frame_index, uint32_t generation,
                        uint32_t previous_generation, uint32_t history_enabled, uint32_t debug_mode,
                        uint32_t light_count, uint32_t object_count, irt_frame* out);
/* 4x4 column-major inverse (double internally, rounded to float). Returns 0 if singular. */
int irt_host_inverse(const float m[16], float out[16]);

/* SDF bake of one mesh on the GPU (tools/export_scene.py:121-158) and the
 * Lipschitz bound of the result (field_lipschitz_bound, renderer.rs:330-358),
 * for building per-session scenes at admission.  positions: xyz of vertex v
 * at positions[v * stride_floats]; indices: triangles (index_count a
 * multiple of 3).  The grid is the vertex bounds padded by
 * |hi - lo| / (resolution - 1) * 2 on every axis (returned in bounds_lo /
 * bounds_hi), sampled on its nodes, x fastest: samples[x + r*(y + r*z)] =
 * sign * (distance to the nearest triangle) - |extent| * 2e-6, negative
 * inside by a majority of three ray-parity votes along the exporter's
 * directions.  resolution in [4, 256]; samples holds resolution^3 floats.
 * Blocking; runs on the current HIP device, allocating its own scratch.
 * IRT_ERR_INVALID_ARGUMENT with "SDF classification exceeded 256
 * intersections" when a parity ray crosses 256 surfaces, as the exporter
 * refuses such meshes. */
int irt_bake_sdf(const float* positions, uint32_t stride_floats, uint32_t vertex_count, const uint32_t* indices,
                 uint32_t index_count, uint32_t resolution, float* samples, float bounds_lo[3], float bounds_hi[3],
                 float* lipschitz);

/* irt_bake_sdf without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (a serving loop that
 * prepares a session's scene at admission, code under HIP graph capture on
 * another stream, harnesses that forbid hipMalloc).  Same arguments, results
 * and errors as irt_bake_sdf, bit for bit; it calls no hipMalloc/hipFree.
 * Its scratch (hierarchy, sample grid, flag and reduction words) is one
 * fixed array of IRT_BAKE_NOALLOC_SCRATCH_BYTES in the library's code
 * object, which every device the library runs on holds until the process
 * ends.  Limits: resolution <= IRT_BAKE_NOALLOC_MAX_RESOLUTION and at most
 * IRT_BAKE_NOALLOC_MAX_TRIANGLES triangles; larger meshes return
 * IRT_ERR_INVALID_ARGUMENT ("... exceeds its fixed scratch").  Calls on one
 * device share the array and are serialised by a lock inside the library;
 * the array keeps the last call's contents, and the words that must start
 * from fixed values are set by every call. */
#define IRT_BAKE_NOALLOC_MAX_RESOLUTION 128u
#define IRT_BAKE_NOALLOC_MAX_TRIANGLES 16384u
#define IRT_BAKE_NOALLOC_SCRATCH_BYTES 10223872u /* 256 + 1 MiB nodes + 768 KiB triangles + 8 MiB samples */
int irt_bake_sdf_noalloc(const float* positions, uint32_t stride_floats, uint32_t vertex_count,
                         const uint32_t* indices, uint32_t index_count, uint32_t resolution, float* samples,
                         float bounds_lo[3], float bounds_hi[3], float* lipschitz);

/* Decision thresholds of the display's sRGB encode: a channel value v
 * (clamped to [0, 1]) gets the code #{k : out[k] <= v}.  out[k] is the
 * linear value where the exact encode times 255 reaches k + 0.5, rounded up
 * to a float. */
void irt_host_srgb_thresholds(float out[255]);

/* Node of a mesh's triangle hierarchy, 32 B, in preorder.  count == 0: an
 * inner node whose left child is the next node, link = the node after its
 * subtree (escape); count > 0: a leaf holding triangles link ..
 * link + count - 1, whose escape is the next node.  Boxes are the bounds of
 * their triangles, widened by 4e-5 of their largest coordinate magnitude so
 * that rounding in the slab test never culls a triangle that is hit. */
typedef struct irt_tri_node {
  float lo[3];
  uint32_t link;
  float hi[3];
  uint32_t count;
} irt_tri_node;
/* Triangle of a hierarchy leaf, 48 B: its corner positions; prim = index of
 * the triangle in its mesh (index-buffer order). */
typedef struct irt_tri {
  float a[3];
  uint32_t prim;
  float b[3];
  uint32_t pad0;
  float c[3];
  uint32_t pad1;
} irt_tri;
/* The per-mesh triangle hierarchy irt_set_geometry builds (software form of
 * the BLAS build, renderer.rs:1061-1106): binned SAH over triangle
 * centroids, leaves of at most 4 triangles.  indices are mesh-local
 * (index_count a multiple of 3, each < vertex_count).  `nodes` must hold
 * max(2*index_count/3 - 1, 1) records and `tris` index_count/3.  Returns the
 * node count (0 for no triangles) or a negative irt_status. */
int irt_host_mesh_hierarchy(const irt_vertex* vertices, uint32_t vertex_count, const uint32_t* indices,
                            uint32_t index_count, irt_tri_node* nodes, irt_tri* tris);

#ifdef __cplusplus
}
#endif
#endif /* IRT_H */

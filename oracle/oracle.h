/*
 * oracle.h — CPU restatement of the reference's per-pixel lighting and
 * denoising path, used ONLY as the checker in tests/, bench.py's
 * cpu_baseline leg and __graft_entry__.smoke().  Never linked into the
 * product library.
 *
 * Two instantiations share one template:
 *   bits = 32: FP32 arithmetic in WGSL evaluation order, values written to
 *              the reference's Rgba16F targets rounded to binary16 (the
 *              "CPU restatement" and the CPU baseline);
 *   bits = 64: FP64 arithmetic, no storage rounding (states the error of the
 *              FP32 results).
 * Both consume the same PCG stream (the FP32 random value), so decisions
 * match except where FP rounding separates near-ties.
 *
 * Interfaces carry every value as double (exact for FP32 inputs).
 */
#ifndef IRT_ORACLE_H
#define IRT_ORACLE_H

#include <stdint.h>

#include "../include/irt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_reservoir {
  uint32_t id, source, count, age;
  double sum, selected_target, pdf;
  uint32_t generation, pad;
} oracle_reservoir; /* 48 B */

typedef struct oracle_scene {
  const irt_object* objects;
  const irt_mesh* meshes;
  const irt_node* nodes;
  const irt_light* lights;
  const float* atlas; /* reference atlas, x fastest */
  uint32_t atlas_w, atlas_h, atlas_d;
  uint32_t n_objects, n_meshes, n_nodes, n_lights;
} oracle_scene;

/* Per-pixel planes, row-major W*H, values of the reference formats. */
typedef struct oracle_gbuffer {
  const double* depth;           /* 1 */
  const double* normal;          /* 4 */
  const double* color;           /* 4 */
  const double* params;          /* 2 */
  const double* motion;          /* 2 */
  const double* previous_depth;  /* 1 */
  const double* previous_normal; /* 4 */
} oracle_gbuffer;

typedef struct oracle_io {
  /* lighting passes (group 2 bindings, lighting.wgsl:6-21) */
  const oracle_reservoir* direct_in;
  const oracle_reservoir* indirect_in;
  oracle_reservoir* direct_out;
  oracle_reservoir* indirect_out;
  const oracle_reservoir* direct_history;
  const oracle_reservoir* indirect_history;
  const irt_vpl* current_vpls;
  const irt_vpl* previous_vpls;
  const irt_vpl* archived_vpls;
  double* direct_output;         /* rgba, resolve_direct */
  const double* direct_lighting; /* rgba, resolve_indirect input */
  double* output;                /* rgba, resolve_indirect (Rgba16F) */
  /* SVGF passes (svgf.wgsl:3-17) */
  const double* noisy;
  const double* old_radiance;
  const double* old_moments;
  const double* input_radiance;
  const double* moments;
  double* radiance_out;
  double* moments_out;
  uint32_t stride; /* wavelet.x */
  /* event counters (brdf, node, box object, grid object, step, tap, march); may be null */
  uint64_t* counters;
  /* restrict the pass to rows [row_begin, row_end) (0,0 = all rows) */
  uint32_t row_begin, row_end;
  int threads; /* OpenMP threads, <= 0: runtime default */
} oracle_io;

enum {
  ORACLE_CANDIDATES_DIRECT = 0,
  ORACLE_CANDIDATES_INDIRECT = 1,
  ORACLE_TEMPORAL_DIRECT = 2,
  ORACLE_TEMPORAL_INDIRECT = 3,
  ORACLE_SPATIAL_DIRECT = 4,
  ORACLE_SPATIAL_INDIRECT = 5,
  ORACLE_RESOLVE_DIRECT = 6,
  ORACLE_RESOLVE_INDIRECT = 7,
  ORACLE_SVGF_TEMPORAL = 8,
  ORACLE_SVGF_VARIANCE = 9,
  ORACLE_SVGF_ATROUS = 10
};

int oracle_run_pass(int bits, int pass, const irt_frame* frame, const oracle_scene* scene,
                    const oracle_gbuffer* gbuffer, oracle_io* io);

/* Building blocks, exposed for known-answer tests. */
float oracle_random(uint32_t* state);
uint32_t oracle_seed(uint32_t i, uint32_t salt, uint32_t frame_index);
double oracle_signed_box_distance(int bits, const double p[3], const double lo[3], const double hi[3]);
double oracle_mesh_distance(int bits, const double p[3], const irt_mesh* mesh, const oracle_scene* scene);
double oracle_scene_distance(int bits, const irt_frame* frame, const oracle_scene* scene, const double p[3],
                             int need_spacing, uint32_t hint, double* voxel, uint32_t* closest);
double oracle_visibility(int bits, const irt_frame* frame, const oracle_scene* scene, const double p[3],
                         const double n[3], const double target[3], const double end_normal[3], uint64_t* counters);
void oracle_brdf(int bits, const double surface[11], const double wo[3], const double wi[3], double out[3]);
/* add() with the random draw supplied (random < 0: draw from *rng). */
int oracle_reservoir_add(int bits, oracle_reservoir* r, uint32_t id, uint32_t source, double target, double pdf,
                         double weight, uint32_t count, uint32_t age, double random, uint32_t* rng);
double oracle_normalization(int bits, const oracle_reservoir* r);
/* One soft-shadow update of the march (sdf.wgsl:180-181): returns the new result. */
double oracle_shadow_step(int bits, double result, double h, double previous, double t, double k);
/* `count` spatial-reuse neighbour offsets (lighting.wgsl:86-89) drawn in
 * sequence from one PCG state: out[2i], out[2i+1] = (dx, dy). */
void oracle_neighbor_offsets(int bits, uint32_t state, uint32_t spatial_radius, uint32_t count, int32_t* out);
uint16_t oracle_f32_to_f16(float f);
float oracle_f16_to_f32(uint16_t h);

/* Triangle geometry of the VPL pass (vpl.wgsl bindings, common.wgsl:11-13). */
typedef struct oracle_geometry {
  const irt_vertex* vertices;
  const uint32_t* indices; /* mesh-local, located by irt_mesh.geometry */
  const irt_material* materials;
  uint32_t n_vertices, n_indices, n_materials;
  const uint8_t* textures; /* RGBA8 layers, size x size, rows top to bottom (may be null) */
  uint32_t texture_size, texture_layers;
} oracle_geometry;

/* vpl.wgsl:2-47 (FP32): frame->counts.y paths into out[0..counts.y).  The
 * closest hit is found by testing every triangle of every object in
 * scene->objects.  *rays (may be null) receives the number of rays traced.
 * Returns 0, or -1 for bad arguments. */
int oracle_generate_vpls(const irt_frame* frame, const oracle_scene* scene, const oracle_geometry* geometry,
                         irt_vpl* out, int threads, uint64_t* rays);
/* Closest hit of one ray (world space); returns 1 on a hit. */
int oracle_trace(const oracle_scene* scene, const oracle_geometry* geometry, const float origin[3],
                 const float direction[3], float* t, float* u, float* v, uint32_t* object, uint32_t* prim);
/* The G-buffer pass (gbuffer.wgsl:1-12, raster state renderer.rs:1370-1415)
 * for frame->size[0] x frame->size[1] pixels, in the reference formats:
 * depth float[W*H], normal/color binary16[4*W*H], params/motion
 * binary16[2*W*H].  Untextured materials.  Returns 0, or -1. */
int oracle_render_gbuffer(const irt_frame* frame, const oracle_scene* scene, const oracle_geometry* geometry,
                          float* depth, uint16_t* normal, uint16_t* color, uint16_t* params, uint16_t* motion);
/* tools/export_scene.py:121-158: the SDF grid of one mesh (every triangle
 * tested), res^3 floats x fastest, and its padded bounds.  Returns 0, -1
 * for bad arguments, -2 when a parity ray crosses 256 surfaces. */
int oracle_bake_sdf(const float* positions, uint32_t stride, uint32_t nv, const uint32_t* indices, uint32_t ni,
                    uint32_t res, float* out, float lo[3], float hi[3], int threads);
/* present.wgsl:1-8 into the Rgba8UnormSrgb display: n Rgba16F history
 * texels to RGBA8 (rgb clamped to [0, 1], sRGB-encoded, nearest code;
 * alpha 255). */
void oracle_present(const uint16_t* history, uint32_t n, uint8_t* out);
/* sample_bsdf (common.wgsl:86-97) for an untextured surface. */
void oracle_sample_bsdf(const float n[3], const float color[3], float rough, float metal, const float wo[3],
                        uint32_t* rng, float direction[3], float weight[3]);

#ifdef __cplusplus
}
#endif
#endif

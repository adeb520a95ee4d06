// Host-side helpers shared by the C ABI implementation.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/irt.h"

namespace irt_host {

void set_global_error(const std::string& s);
const char* global_error();
int validate_settings(const irt_settings& s, std::string* msg);
bool is_box_geometry(const float* pos, uint32_t nv, uint32_t stride, const uint32_t* idx, uint32_t ni,
                     const float lo[3], const float hi[3]);
float field_lipschitz_bound(const float* samples, uint32_t res, const float spacing[3]);
int hierarchy(const float* box_lo, const float* box_hi, uint32_t count, irt_node* out);
void transform_box(const float m[16], const float lo[3], const float hi[3], float out_lo[3], float out_hi[3]);
void atlas_grid(uint32_t count, uint32_t grid[3]);
bool inverse4(const float* a, float* out);
int pack_frame(const irt_settings& s, const float* vp, const float* pvp, const float* cam, uint32_t frame_index,
               uint32_t generation, uint32_t previous_generation, uint32_t history_enabled, uint32_t debug_mode,
               uint32_t lights, uint32_t objects, irt_frame* f);
uint16_t f32_to_f16(float f);
float f16_to_f32(uint16_t h);
uint16_t f32_to_f16_down(float f);
// The 255 linear values at which the correctly rounded 8-bit sRGB code of
// the exact encode steps up (threshold k: code k -> k + 1).
void srgb_thresholds(float out[255]);
// Triangle hierarchy of one mesh (mesh-local indices); node and triangle
// links are local to the returned arrays.
int mesh_hierarchy(const irt_vertex* v, uint32_t nv, const uint32_t* idx, uint32_t ni,
                   std::vector<irt_tri_node>& nodes, std::vector<irt_tri>& tris, std::string* msg);

}  // namespace irt_host

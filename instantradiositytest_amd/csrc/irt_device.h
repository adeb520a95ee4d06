// Device-side per-pixel math of the lighting + denoising path (gfx950).
//
// Each function restates the behaviour of one WGSL function of the reference;
// the file:line of that function is cited beside it.  Semantics that matter
// for parity are kept exactly: the PCG stream and its consumption points,
// half-to-even round(), saturating float->int conversions, the reservoir
// add() rule that lets NaN weights through, the order of the BVH walk and the
// march's early-outs.  Everything else is written for the hardware: scalar
// (kernarg) frame constants, LDS-resident scene records where a kernel stages
// them, and explicit FP32 arithmetic (no packed-FP32 forms, see the build).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/irt.h"

// The arithmetic contract of the translation unit: IRT_MARCH_EXACT (one
// IEEE rounding per source operation, correctly rounded division and square
// root) unless the unit defines IRT_MARCH_MATH as IRT_MARCH_CONTRACTED before
// its first include (the mixed profile's march, irt_march_contracted.hip).
#define IRT_MARCH_EXACT 0
#define IRT_MARCH_CONTRACTED 1
#ifndef IRT_MARCH_MATH
#define IRT_MARCH_MATH IRT_MARCH_EXACT
#endif

namespace irt {

#define IRT_PI 3.14159265359f

// sin and cos for angles |x| <= 2 pi (the spatial neighbour's 2 pi u, the VPL
// pass's azimuths and half cone angles), evaluated with the same float
// operations as the oracle so that rounding at pixel and path decisions is
// reproducible bit for bit (platform sin/cos differ by an ulp between the
// device and the host): reduction by quarter turns in two
// constants (after Cody & Waite 1980: pi/2 as its float plus the float of the
// remainder), then Taylor polynomials in r^2 to degree 11 / 12, whose
// truncation error on |r| <= pi/4 is below 2e-9.
__device__ __forceinline__ void shared_sincos(float x, float* sn, float* cs) {
  const float k = rintf(x * 0.63661977236758134f);  // x / (pi/2), to the nearest quarter turn
  const float r = (x - k * 1.57079637050628662f) - k * -4.37113900018624283e-8f;
  const float r2 = r * r;
  const float s = r + r * r2 * (-1.0f / 6.0f + r2 * (1.0f / 120.0f + r2 * (-1.0f / 5040.0f + r2 * (1.0f / 362880.0f + r2 * (-1.0f / 39916800.0f)))));
  const float c = 1.0f + r2 * (-0.5f + r2 * (1.0f / 24.0f + r2 * (-1.0f / 720.0f + r2 * (1.0f / 40320.0f + r2 * (-1.0f / 3628800.0f + r2 * (1.0f / 479001600.0f))))));
  const int q = ((int)k) & 3;
  *sn = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
  *cs = q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}
__device__ __forceinline__ float shared_cos(float x) {
  float s, c;
  shared_sincos(x, &s, &c);
  return c;
}

// Square root.  The exact contract keeps the correctly rounded sqrtf.  The
// contracted unit (built without correctly rounded square roots) lowers sqrtf
// to the hardware square root plus a rescaling of inputs below 2^-96; calling
// the hardware instruction directly gives the same result for every input at
// or above 2^-96 and drops the rescaling (five instructions per root).
__device__ __forceinline__ float fsqrt(float x) {
#if IRT_MARCH_MATH == IRT_MARCH_CONTRACTED
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

// min(max(x, lo), hi) for lo <= hi.  IEEE-mode minimum and maximum quiet a
// signalling NaN, so the compiler puts a canonicalising copy in front of
// every operand it cannot prove is an arithmetic result (values read from
// LDS or memory); the contracted unit takes the median of three, which is
// the same value for every non-NaN x in one instruction.  The exact contract
// keeps the minimum and maximum (a NaN x gives lo there, as in the FP32
// restatement).
__device__ __forceinline__ float fclamp(float x, float lo, float hi) {
#if IRT_MARCH_MATH == IRT_MARCH_CONTRACTED
  return __builtin_amdgcn_fmed3f(x, lo, hi);
#else
  return fminf(fmaxf(x, lo), hi);
#endif
}
// max(x, 0) the same way.
__device__ __forceinline__ float fmax0(float x) {
#if IRT_MARCH_MATH == IRT_MARCH_CONTRACTED
  return __builtin_amdgcn_fmed3f(x, 0.0f, __builtin_inff());
#else
  return fmaxf(x, 0.0f);
#endif
}

// ------------------------------------- correctly rounded division shortcuts
// The exact contract's a / b is an 11-instruction sequence (operand scaling,
// reciprocal, Newton step, two remainder corrections, special-case fixup).
// Where the operands are known to stay in range, fewer steps give the same
// bits:
// - the hardware reciprocal refined by one Newton step is RN(1/b) for every
//   normal |b| in [2^-100, 2^100];
// - with y = RN(1/b), q = RN(a y) and one remainder step RN(q + (a - b q) y)
//   is RN(a/b) (Markstein's correction) while a, a/b and the remainder stay
//   normal;
// - v_div_fixup gives zero, infinite and NaN operands the IEEE result, and
//   -0 numerators their sign.
// irt_check_fast_division compares these helpers with the correctly rounded
// sequences on the GPU for every binary32 input of their domains.

// sqrtf(x) for x >= 2^-96, +inf and NaN: the hardware root and the two
// neighbour residual tests of the correctly rounded sequence, without its
// rescaling of inputs below 2^-96 and its zero/infinity select (9
// instructions for 16).
__device__ __forceinline__ float sqrt_rn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float down = __uint_as_float(__float_as_uint(s) - 1u);
  const float up = __uint_as_float(__float_as_uint(s) + 1u);
  float r = fmaf(-down, s, x) <= 0.0f ? down : s;
  r = fmaf(-up, s, x) > 0.0f ? up : r;
  return r;
}
// 1.0f / b for normal |b| in [2^-100, 2^100] and for +-0, +-inf and NaN.
__device__ __forceinline__ float rcp_rn(float b) {
  const float y = __builtin_amdgcn_rcpf(b);
  return __builtin_amdgcn_div_fixupf(fmaf(fmaf(-b, y, 1.0f), y, y), b, 1.0f);
}
// a / IRT_PI for |a| in [2^-79, 2^126] and for +-0, +-inf and NaN.
__device__ __forceinline__ float div_pi_rn(float a) {
  constexpr float y = 1.0f / IRT_PI;  // RN(1/pi), folded at compile time
  const float q = a * y;
  const float r = fmaf(-IRT_PI, q, a);
  return __builtin_amdgcn_div_fixupf(fmaf(r, y, q), IRT_PI, a);
}
// RN(1/b) for normal |b| in [2^-100, 2^100] (no special-case fixup).
__device__ __forceinline__ float rcp_refined(float b) {
  const float y = __builtin_amdgcn_rcpf(b);
  return fmaf(fmaf(-b, y, 1.0f), y, y);
}
// a / b given y = rcp_refined(b), for normal |b| in [2^-100, 2^100]: the
// correctly rounded quotient when a is 0, infinite or NaN, or when
// |a| >= 2^-79 and the binary exponents satisfy -125 <= e(a) - e(b) <= 126
// (a, the quotient and the remainder then stay normal, so the
// significand-level equality scales to them exactly).  For |b| in
// [2^-17, 2^40) every a in [2^-79, 2^110) qualifies (outside_div_ranged).
__device__ __forceinline__ float div_ranged(float a, float b, float y) {
  const float q = a * y;
  return __builtin_amdgcn_div_fixupf(fmaf(fmaf(-b, q, a), y, q), b, a);
}
// True when one of the three numerators lies outside [2^-79, 2^110), div_ranged's
// domain for divisors in [2^-17, 2^40) (frexp's exponent is 0 for 0,
// infinities and NaN, which are inside).
__device__ __forceinline__ bool outside_div_ranged(float a0, float a1, float a2) {
  const int e0 = __builtin_amdgcn_frexp_expf(a0), e1 = __builtin_amdgcn_frexp_expf(a1),
            e2 = __builtin_amdgcn_frexp_expf(a2);
  return min(e0, min(e1, e2)) < -78 || max(e0, max(e1, e2)) > 110;
}
// a / pdf for a candidate weight, pdf = 1 / count (|pdf| in [2^-32, 1]) and
// y = rcp_refined(pdf): div_ranged's argument holds for 2^-79 <= |a| < 2^95
// (the quotient stays below 2^127), and 0, infinite or NaN a; other a take
// the division sequence.  The branch is skipped by waves whose lanes all fit.
__device__ __forceinline__ float div_weight(float a, float pdf, float y) {
  float w = div_ranged(a, pdf, y);
  const int e = __builtin_amdgcn_frexp_expf(a);
  if (e < -78 || e > 95) {
    asm volatile("" : "+v"(a));  // keeps the division in the branch (not speculated)
    w = a / pdf;
  }
  return w;
}

// ------------------------------------------------------------------ vectors
struct v3 {
  float x, y, z;
};
__device__ __forceinline__ v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ v3 operator/(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(v3 a) { return fsqrt(dot(a, a)); }
// WGSL normalize(e) = e / length(e), divided per component.
__device__ __forceinline__ v3 normalize(v3 a) {
  float l = fsqrt(dot(a, a));
  return v3{a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ v3 vmax(v3 a, v3 b) { return v3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ v3 vmin(v3 a, v3 b) { return v3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ v3 vmaxs(v3 a, float s) { return v3{fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)}; }
__device__ __forceinline__ v3 ld3(const float* p) { return v3{p[0], p[1], p[2]}; }
__device__ __forceinline__ v3 xyz(float4 v) { return v3{v.x, v.y, v.z}; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// Column-major 4x4 times (x, y, z, w).
__device__ __forceinline__ float4 mat_mul(const float* m, float x, float y, float z, float w) {
  float4 r;
  r.x = m[0] * x + m[4] * y + m[8] * z + m[12] * w;
  r.y = m[1] * x + m[5] * y + m[9] * z + m[13] * w;
  r.z = m[2] * x + m[6] * y + m[10] * z + m[14] * w;
  r.w = m[3] * x + m[7] * y + m[11] * z + m[15] * w;
  return r;
}

// ------------------------------------------------------------ conversions
// WGSL f32 -> i32 / u32 conversions saturate (NaN -> 0); C casts do not.
__device__ __forceinline__ int32_t f2i_sat(float x) {
  if (!(x == x)) return 0;
  if (x >= 2147483647.0f) return 2147483647;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)x;
}
__device__ __forceinline__ uint32_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967295.0f) return 4294967295u;
  return (uint32_t)x;
}
// Rgba16F storage rounding (round to nearest even), as a texture write does.
// The sign is taken back from x: the compiler may fold a product that feeds
// the conversion into a mixed-precision FMA with a +0 addend
// (v_fma_mixlo_f16), which turns a -0 product into +0.
__device__ __forceinline__ float round_f16(float x) { return __builtin_copysignf((float)(_Float16)x, x); }
__device__ __forceinline__ float4 round_f16(float4 v) {
  return make_float4(round_f16(v.x), round_f16(v.y), round_f16(v.z), round_f16(v.w));
}

// -------------------------------------------------------- storage formats
// 16-bit channel storage (binary16 bit patterns) for the mixed profile.
// Aligned to their size so that loads and stores are single 64-/32-bit accesses.
struct __align__(8) h4 {
  _Float16 x, y, z, w;
};
struct __align__(4) h2 {
  _Float16 x, y;
};
__device__ __forceinline__ float4 ld4(const float4* p, uint32_t i) { return p[i]; }
__device__ __forceinline__ float4 ld4(const h4* p, uint32_t i) {
  h4 v = p[i];
  return make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
}
__device__ __forceinline__ float2 ld2(const float2* p, uint32_t i) { return p[i]; }
__device__ __forceinline__ float2 ld2(const h2* p, uint32_t i) {
  h2 v = p[i];
  return make_float2((float)v.x, (float)v.y);
}
// Store a value the reference writes to an Rgba16F target.
__device__ __forceinline__ void st_r16(float4* p, uint32_t i, float4 v) { p[i] = round_f16(v); }
__device__ __forceinline__ void st_r16(h4* p, uint32_t i, float4 v) {
  p[i] = h4{(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
}
__device__ __forceinline__ void st4(float4* p, uint32_t i, float4 v) { p[i] = v; }
__device__ __forceinline__ void st4(h4* p, uint32_t i, float4 v) { st_r16(p, i, v); }

// Per-pixel guide record, built once per frame by k_guide: the world
// position of the depth sample (FP32, computed exactly as world() does) and
// the binary16 G-buffer channels that neighbourhood taps read.  The G-buffer
// holds binary16 values in both profiles, so the record is lossless.  32 B.
struct __align__(16) GuidePx {
  float px, py, pz;
  uint32_t valid;  // depth < 1
  _Float16 nx, ny, nz, cr, cg, cb, m0, m1;
};
static_assert(sizeof(GuidePx) == 32, "guide record is 32 bytes");
// Loads one guide record as two 16-byte vector loads and unpacks the halves
// in registers (field-wise access would issue one 16-bit load per channel).
__device__ __forceinline__ _Float16 lo16(uint32_t w) { return __builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)); }
__device__ __forceinline__ _Float16 hi16(uint32_t w) { return __builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }
__device__ __forceinline__ GuidePx unpack_guide(uint4 a, uint4 b) {
  GuidePx r;
  r.px = __uint_as_float(a.x);
  r.py = __uint_as_float(a.y);
  r.pz = __uint_as_float(a.z);
  r.valid = a.w;
  r.nx = lo16(b.x);
  r.ny = hi16(b.x);
  r.nz = lo16(b.y);
  r.cr = hi16(b.y);
  r.cg = lo16(b.z);
  r.cb = hi16(b.z);
  r.m0 = lo16(b.w);
  r.m1 = hi16(b.w);
  return r;
}
__device__ __forceinline__ GuidePx ld_guide(const GuidePx* g, uint32_t i) {
  const uint4* p = reinterpret_cast<const uint4*>(g + i);
  return unpack_guide(p[0], p[1]);
}

// SVGF moments record of the mixed profile, 8 B: the first moment and the age
// in binary16 (ages are integers <= 32, exact), the second moment in FP32.
// The second moment keeps FP32 because the variance m2 - m1^2 cancels: a
// 16-bit m2 would move the variance of a smooth region by more than the
// variance itself.  The first moment is a luminance mean, at most 65000 as
// the radiance is.  The FP32 profile keeps the reference's Rgba32F record.
struct __align__(8) Mom8 {
  _Float16 m0, age;
  float m1;
};
static_assert(sizeof(Mom8) == 8, "mixed moments record is 8 bytes");
__device__ __forceinline__ float4 ld_mom(const float4* p, uint32_t i) { return p[i]; }
__device__ __forceinline__ float4 ld_mom(const Mom8* p, uint32_t i) {
  const Mom8 r = p[i];
  return make_float4((float)r.m0, r.m1, (float)r.age, 0.0f);
}
__device__ __forceinline__ void st_mom(float4* p, uint32_t i, float4 v) { p[i] = v; }
__device__ __forceinline__ void st_mom(Mom8* p, uint32_t i, float4 v) {
  Mom8 r;
  r.m0 = (_Float16)v.x;
  r.age = (_Float16)v.z;
  r.m1 = v.y;
  p[i] = r;
}

// Storage profiles.
struct ProfileF32 {
  typedef float4 N4;  // normal, colour (xyz/rgb, w)
  typedef float2 P2;  // params (rough, metal), motion
  typedef float4 R4;  // radiance targets (values binary16-rounded as the reference stores them)
  typedef float4 L4;  // direct lighting (Rgba32F in the reference)
  typedef float4 M4;  // SVGF moments (Rgba32F in the reference)
};
struct ProfileMixed {
  typedef h4 N4;
  typedef h2 P2;
  typedef h4 R4;
  typedef h4 L4;
  typedef Mom8 M4;
};

// ------------------------------------------------------------- counters
// Event counters for the algorithmic FLOP model (instrumented launches only).
enum { C_BRDF = 0, C_NODE, C_OBJ_BOX, C_OBJ_GRID, C_STEP, C_TAP, C_MARCH, C_N };
template <bool ON>
struct Count {
  __device__ __forceinline__ void add(int, uint32_t) {}
  __device__ __forceinline__ void flush(unsigned long long*) {}
};
template <>
struct Count<true> {
  uint32_t c[C_N] = {0, 0, 0, 0, 0, 0, 0};
  __device__ __forceinline__ void add(int k, uint32_t n) { c[k] += n; }
  __device__ void flush(unsigned long long* out) {
    for (int k = 0; k < C_N; ++k) {
      unsigned long long v = c[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0 && v) atomicAdd(out + k, v);
    }
  }
};

// ------------------------------------------------------------------ RNG
// random(), seed() — src/shaders/common.wgsl:39-44.
__device__ __forceinline__ float rnd(uint32_t& s) {
  s = s * 747796405u + 2891336453u;
  uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (float)((w >> 22u) ^ w) * (1.0f / 4294967296.0f);
}
__device__ __forceinline__ uint32_t seed(uint32_t i, uint32_t salt, uint32_t frame_index) {
  return i * 1973u + frame_index * 9277u + salt * 26699u + 89173u;
}

// luminance, world, pixel_uv — common.wgsl:45-47.
__device__ __forceinline__ float luminance(v3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
__device__ __forceinline__ v3 world(float u, float v, float depth, const float* m) {
  float4 h = mat_mul(m, u * 2.0f - 1.0f, 1.0f - v * 2.0f, depth, 1.0f);
  return v3{h.x / h.w, h.y / h.w, h.z / h.w};
}
__device__ __forceinline__ float pixel_u(int32_t x, uint32_t w) { return ((float)x + 0.5f) / (float)w; }

// ------------------------------------------------------------------ BRDF
struct Surface {
  v3 p, n, color;
  float rough, metal;
  // The parts of brdf (common.wgsl:57-74) that depend only on the surface
  // and the camera, evaluated once per pixel by surface_view() with the
  // same operations the per-sample evaluation would repeat.
  v3 wo, f0;
  float nv, a2, g1v, im;
};

// pow(x, 5) and pow(x, 64) for x >= 0 as fixed multiply chains: (x^2)^2 * x
// and six squarings.  The FP32 restatement evaluates the same chains.
__device__ __forceinline__ float pow5(float x) {
  float x2 = x * x;
  return x2 * x2 * x;
}
__device__ __forceinline__ float pow64(float x) {
  float y = x * x;  // 2
  y = y * y;        // 4
  y = y * y;        // 8
  y = y * y;        // 16
  y = y * y;        // 32
  return y * y;     // 64
}

// brdf — common.wgsl:67-74.
// The wo-dependent terms of brdf for s (wo, n.v, alpha^2, the view Smith
// term, 1 - metal and F0); surface_view() takes wo = normalize(camera - p).
__device__ __forceinline__ void surface_toward(Surface& s, v3 wo) {
  s.wo = wo;
  s.nv = fmaxf(dot(s.n, s.wo), 0.00001f);
  const float a = s.rough * s.rough;
  s.a2 = a * a;
  s.g1v = 2.0f * s.nv / (s.nv + sqrtf(s.a2 + (1.0f - s.a2) * s.nv * s.nv));
  s.im = 1.0f - s.metal;
  s.f0 = v3{0.04f * s.im + s.color.x * s.metal, 0.04f * s.im + s.color.y * s.metal,
            0.04f * s.im + s.color.z * s.metal};
}
__device__ __forceinline__ void surface_view(Surface& s, v3 camera) { surface_toward(s, normalize(camera - s.p)); }

// brdf toward wi for a surface whose view terms surface_view() filled.
// FAST (G-buffer surfaces, whose colour and metalness are binary16 values)
// divides by pi with div_pi_rn, takes the half vector's normalisation and
// the specular division with div_ranged and sets `unsafe` when an operand
// leaves div_ranged's domain (the caller then evaluates the exact form); the
// results are the exact form's bits.
template <bool FAST>
__device__ __forceinline__ v3 brdf_of(const Surface& s, v3 wi, bool& unsafe) {
  const v3 wo = s.wo;
  const float nv = s.nv, a2 = s.a2, g1v = s.g1v, im = s.im;
  const v3 f0 = s.f0;
  float nl = fmaxf(dot(s.n, wi), 0.0f);
  if (nl <= 0.0f) return v3{0.f, 0.f, 0.f};
  v3 h;
  if constexpr (FAST) {
    // |wo + wi| <= 2 for unit wo, wi; below 2^-17, 0 or NaN leaves the domain
    const v3 hs = wo + wi;
    const float l2 = dot(hs, hs);
    // l2 >= 2^-34: sqrt_rn's domain and a length of at least 2^-17
    unsafe = unsafe || !(l2 >= 0x1p-34f) || outside_div_ranged(hs.x, hs.y, hs.z);
    const float l = sqrt_rn(l2);
    const float y = rcp_refined(l);
    h = v3{div_ranged(hs.x, l, y), div_ranged(hs.y, l, y), div_ranged(hs.z, l, y)};
  } else {
    h = normalize(wo + wi);
  }
  float nh = fmaxf(dot(s.n, h), 0.0f);
  float vh = fmaxf(dot(wo, h), 0.0f);
  float q = nh * nh * (a2 - 1.0f) + 1.0f;
  float d = a2 / (IRT_PI * q * q);
  float g1l;
  if constexpr (FAST) {
    // With nl in [2^-40, 1.01] and the root's argument at least 2^-96,
    // 2 nl is in [2^-39, 2.02] and its divisor nl + sqrt(...) in [2^-40, 2^33]
    // (a2 <= 2^64 for binary16 roughness), so the quotient's exponent is in
    // [-72, 1]: inside div_ranged's contract
    const float g = a2 + (1.0f - a2) * nl * nl;
    unsafe = unsafe || !(nl >= 0x1p-40f) || !(g >= 0x1p-96f);
    const float gb = nl + sqrt_rn(g);
    g1l = div_ranged(2.0f * nl, gb, rcp_refined(gb));
  } else {
    g1l = 2.0f * nl / (nl + sqrtf(a2 + (1.0f - a2) * nl * nl));
  }
  float p5 = pow5(1.0f - vh);
  v3 f = v3{f0.x + (1.0f - f0.x) * p5, f0.y + (1.0f - f0.y) * p5, f0.z + (1.0f - f0.z) * p5};
  // (1-f)*(1-metal)*color/PI + f*d*g1v*g1l/max(4*nv*nl, 1e-5), evaluated left to right.
  float den = fmaxf(4.0f * nv * nl, 0.00001f);
  v3 diff;
  if constexpr (FAST) {
    // G-buffer surfaces: the numerators are 0, infinite, NaN or in
    // [2^-59, 2^65]: factors 1 - f (0 or at least 2^-24: exact below 2 by
    // Sterbenz's lemma), 1 - metal and the colour (binary16 values: 0 or at
    // least 2^-24, at most 2^16).
    diff = v3{div_pi_rn((1.0f - f.x) * im * s.color.x), div_pi_rn((1.0f - f.y) * im * s.color.y),
              div_pi_rn((1.0f - f.z) * im * s.color.z)};
  } else {
    diff = v3{(1.0f - f.x) * im * s.color.x / IRT_PI, (1.0f - f.y) * im * s.color.y / IRT_PI,
              (1.0f - f.z) * im * s.color.z / IRT_PI};
  }
  v3 spec;
  if constexpr (FAST) {
    // den in [1e-5, 4.01]: n and wo, wi are unit vectors
    const float s0 = f.x * d * g1v * g1l, s1 = f.y * d * g1v * g1l, s2 = f.z * d * g1v * g1l;
    unsafe = unsafe || outside_div_ranged(s0, s1, s2);
    const float y = rcp_refined(den);
    spec = v3{div_ranged(s0, den, y), div_ranged(s1, den, y), div_ranged(s2, den, y)};
  } else {
    spec = v3{f.x * d * g1v * g1l / den, f.y * d * g1v * g1l / den, f.z * d * g1v * g1l / den};
  }
  return diff + spec;
}

// The exact form, for surfaces of any material values (the VPL pass).
__device__ __forceinline__ v3 brdf(const Surface& s, v3 wi) {
  bool unused = false;
  return brdf_of<false>(s, wi, unused);
}

// direct — common.wgsl:75-79.  FAST as in brdf_of.
template <bool FAST, class C>
__device__ __forceinline__ v3 direct_of(const Surface& s, const irt_light& l, C& cnt, bool& unsafe) {
  v3 lp = v3{l.position_strength[0], l.position_strength[1], l.position_strength[2]};
  v3 delta = lp - s.p;
  float r2 = fmaxf(dot(delta, delta), 0.0001f);
  v3 wi = delta * rcp_rn(sqrt_rn(r2));  // inverseSqrt(r2); r2 >= 1e-4, sqrt(r2) in [0.01, 2^64] or +inf
  if (l.color_type[3] > 0.5f) {
    v3 axis = normalize(v3{l.direction_angle[0], l.direction_angle[1], l.direction_angle[2]});
    if (dot(-wi, axis) < cosf(l.direction_angle[3] * 0.5f)) return v3{0.f, 0.f, 0.f};
  }
  cnt.add(C_BRDF, 1);
  v3 b = brdf_of<FAST>(s, wi, unsafe);  // view terms of the frame's camera (surface_view)
  float k = fmaxf(dot(s.n, wi), 0.0f);
  float strength = l.position_strength[3];
  const float n0 = b.x * k * l.color_type[0] * strength, n1 = b.y * k * l.color_type[1] * strength,
              n2 = b.z * k * l.color_type[2] * strength;
  if constexpr (FAST) {
    // r2 >= 1e-4; above 2^40 (or +inf) it leaves div_ranged's domain
    unsafe = unsafe || !(r2 <= 0x1p40f) || outside_div_ranged(n0, n1, n2);
    const float y = rcp_refined(r2);
    return v3{div_ranged(n0, r2, y), div_ranged(n1, r2, y), div_ranged(n2, r2, y)};
  } else {
    return v3{n0 / r2, n1 / r2, n2 / r2};
  }
}
// The shortcut form, and the exact form for a wave with a lane whose
// operands leave the shortcuts' domains (counted once).
template <class C>
__device__ __forceinline__ v3 direct(const Surface& s, const irt_light& l, C& cnt) {
  bool unsafe = false;
  v3 v = direct_of<true>(s, l, cnt, unsafe);
  if (unsafe) {
    Count<false> none;
    v = direct_of<false>(s, l, none, unsafe);
  }
  return v;
}

// indirect — common.wgsl:80-83.  FAST as in brdf_of.
template <bool FAST, class C>
__device__ __forceinline__ v3 indirect_of(const Surface& s, const irt_vpl& v, C& cnt, bool& unsafe) {
  v3 delta = ld3(v.position) - s.p;
  float r2 = fmaxf(dot(delta, delta), 0.0025f);
  v3 wi = delta * rcp_rn(sqrt_rn(r2));  // inverseSqrt(r2); r2 >= 0.0025, sqrt(r2) in [0.05, 2^64] or +inf
  cnt.add(C_BRDF, 1);
  v3 b = brdf_of<FAST>(s, wi, unsafe);  // view terms of the frame's camera (surface_view)
  float k1 = fmaxf(dot(s.n, wi), 0.0f);
  float k2 = fmaxf(dot(ld3(v.normal), -wi), 0.0f);
  float den = 4.0f * IRT_PI * r2;
  const float n0 = b.x * k1 * k2 * v.flux[0], n1 = b.y * k1 * k2 * v.flux[1], n2 = b.z * k1 * k2 * v.flux[2];
  if constexpr (FAST) {
    // den >= 0.0314; above 2^40 (or +inf) it leaves div_ranged's domain
    unsafe = unsafe || !(den <= 0x1p40f) || outside_div_ranged(n0, n1, n2);
    const float y = rcp_refined(den);
    return v3{div_ranged(n0, den, y), div_ranged(n1, den, y), div_ranged(n2, den, y)};
  } else {
    return v3{n0 / den, n1 / den, n2 / den};
  }
}
template <class C>
__device__ __forceinline__ v3 indirect(const Surface& s, const irt_vpl& v, C& cnt) {
  bool unsafe = false;
  v3 r = indirect_of<true>(s, v, cnt, unsafe);
  if (unsafe) {
    Count<false> none;
    r = indirect_of<false>(s, v, none, unsafe);
  }
  return r;
}

// ------------------------------------------------------------- reservoirs
// empty_reservoir, normalization, add — common.wgsl:98-106; cap — lighting.wgsl:51.
__device__ __forceinline__ irt_reservoir empty_reservoir(uint32_t frame_index, uint32_t generation) {
  irt_reservoir r;
  r.id = 0u;
  r.source = frame_index;
  r.count = 0u;
  r.age = 0u;
  r.sum = 0.0f;
  r.selected_target = 0.0f;
  r.pdf = 0.0f;
  r.generation = generation;
  return r;
}
__device__ __forceinline__ float normalization(const irt_reservoir& r) {
  if (r.count == 0u || r.selected_target <= 0.0f) return 0.0f;
  return r.sum / ((float)r.count * r.selected_target);
}
// A NaN weight passes the range test (as in the reference) and poisons `sum`.
__device__ __forceinline__ bool add(irt_reservoir& r, uint32_t id, uint32_t source, float target, float pdf,
                                    float weight, uint32_t count, uint32_t age, uint32_t& rng) {
  r.count += count;
  if (weight <= 0.0f || weight > 1e30f) return false;
  r.sum += weight;
  if (rnd(rng) * r.sum < weight) {
    r.id = id;
    r.source = source;
    r.selected_target = target;
    r.pdf = pdf;
    r.age = age;
    return true;
  }
  return false;
}
__device__ __forceinline__ irt_reservoir cap(irt_reservoir a, uint32_t limit) {
  if (a.count > limit) {
    a.sum *= (float)limit / (float)a.count;
    a.count = limit;
  }
  return a;
}

// --------------------------------------------------------------- SDF
// Candidate lists of a uniform grid over the scene.  Each cell lists every
// object that can hold the minimum distance anywhere in the cell (objects
// whose lower bound over the cell does not exceed the smallest upper bound
// over the cell), nearest lower bound first.  count == 255: list overflow,
// queries in that cell walk the hierarchy instead.
struct CellList {
  uint8_t count;
  uint8_t ids[7];
};
static_assert(sizeof(CellList) == 8, "cell list is 8 bytes");
constexpr uint32_t CELL_MAX = 7;  // longer lists overflow
constexpr uint8_t CELL_OVERFLOW = 255;
struct GridView {
  const CellList* cells;  // null: no grid, hierarchy walk only
  const float4* boxes;    // per object: world AABB lo (xyz), hi (xyz) as two float4
  float x0, y0, z0, inv_cell;
  uint32_t nx, ny, nz;
};

// Scene records as the kernels see them.
template <class O>
struct SceneViewOf {
  const O* objects;
  const irt_mesh* meshes;
  const irt_node* nodes;
  const irt_light* lights;
  const float* atlas;  // reference atlas layout: x + w*(y + h*z)
  uint32_t atlas_w, atlas_h;
  GridView grid;
  const _Float16* atlas16;  // mixed profile: the same layout, samples rounded toward -inf (else null)
};
using SceneView = SceneViewOf<irt_object>;
// The parts of an object record that the scene queries read (inverse
// transform and info: 80 of its 208 bytes), as the march stages them in LDS.
struct QueryObject {
  float inverse[16];
  uint32_t info[4];
};
using QuerySceneView = SceneViewOf<QueryObject>;

// signed_box_distance — sdf.wgsl:1-4.
__device__ __forceinline__ float signed_box_distance(v3 p, v3 lo, v3 hi) {
  v3 delta = vmax(lo - p, p - hi);
  return length(vmaxs(delta, 0.0f)) + fminf(fmaxf(delta.x, fmaxf(delta.y, delta.z)), 0.0f);
}

// sample_field (manual 8-tap trilinear) — common.wgsl:31-38.
template <class SV>
__device__ __forceinline__ float sample_field(const SV& sc, const uint32_t* origin, v3 q, uint32_t res) {
  // cell = min(floor(q), res - 2) with the clamp taken on the float (a
  // median of three: also 0 below, the saturation of the u32 conversion);
  // res - 2 and the floors are exact integers, so cell and f = q - cell are
  // the values of the reference's integer clamp.
  const float top = (float)(res - 2u);
  const float ux = __builtin_amdgcn_fmed3f(floorf(q.x), 0.0f, top);
  const float uy = __builtin_amdgcn_fmed3f(floorf(q.y), 0.0f, top);
  const float uz = __builtin_amdgcn_fmed3f(floorf(q.z), 0.0f, top);
  const uint32_t cx = (uint32_t)ux, cy = (uint32_t)uy, cz = (uint32_t)uz;
  float fx = q.x - ux, fy = q.y - uy, fz = q.z - uz;
  const uint32_t w = sc.atlas_w, h = sc.atlas_h;
  const size_t off = ((size_t)(origin[2] + cz) * h + (origin[1] + cy)) * w + (origin[0] + cx);
  const size_t sy = w, sz = (size_t)w * h;
  float t000, t100, t010, t110, t001, t101, t011, t111;
  if (sc.atlas16) {
    const _Float16* b = sc.atlas16 + off;
    t000 = (float)b[0], t100 = (float)b[1], t010 = (float)b[sy], t110 = (float)b[sy + 1];
    t001 = (float)b[sz], t101 = (float)b[sz + 1], t011 = (float)b[sz + sy], t111 = (float)b[sz + sy + 1];
  } else {
    const float* b = sc.atlas + off;
    t000 = b[0], t100 = b[1], t010 = b[sy], t110 = b[sy + 1];
    t001 = b[sz], t101 = b[sz + 1], t011 = b[sz + sy], t111 = b[sz + sy + 1];
  }
#if IRT_MARCH_MATH == IRT_MARCH_CONTRACTED
  // The contracted unit blends with seven lerps, x then y then z: the same
  // trilinear value up to rounding, in 14 instructions for the weighted
  // sum's 27.
  const float x00 = fmaf(fx, t100 - t000, t000), x10 = fmaf(fx, t110 - t010, t010);
  const float x01 = fmaf(fx, t101 - t001, t001), x11 = fmaf(fx, t111 - t011, t011);
  const float y0 = fmaf(fy, x10 - x00, x00), y1 = fmaf(fy, x11 - x01, x01);
  return fmaf(fz, y1 - y0, y0);
#else
  float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
  // value += texel * w.x * w.y * w.z, corners in z, y, x order (common.wgsl:33-36).
  float value = 0.0f;
  value += t000 * gx * gy * gz;
  value += t100 * fx * gy * gz;
  value += t010 * gx * fy * gz;
  value += t110 * fx * fy * gz;
  value += t001 * gx * gy * fz;
  value += t101 * fx * gy * fz;
  value += t011 * gx * fy * fz;
  value += t111 * fx * fy * fz;
  return value;
#endif
}

// mesh_distance — sdf.wgsl:5-33.
template <class SV, class C>
__device__ __forceinline__ float mesh_distance(const SV& sc, v3 p, const irt_mesh& mesh, C& cnt) {
  v3 glo = ld3(mesh.geometry_lo), ghi = ld3(mesh.geometry_hi);
  float geometry_bound = signed_box_distance(p, glo, ghi);
  if (mesh.geometry_lo[3] > 0.5f) {
    cnt.add(C_OBJ_BOX, 1);
    return geometry_bound;
  }
  cnt.add(C_OBJ_GRID, 1);
  uint32_t res = mesh.field[3];
  v3 lo = ld3(mesh.lo), hi = ld3(mesh.hi);
  v3 clamped = v3{fclamp(p.x, lo.x, hi.x), fclamp(p.y, lo.y, hi.y), fclamp(p.z, lo.z, hi.z)};
  v3 q = (clamped - lo) * ld3(mesh.grid);
  float value = sample_field(sc, mesh.field, q, res);
  float bound = value - mesh.grid[3] * (0.5f + 1.0f / 128.0f);
  v3 outside = vmaxs(vmax(lo - p, p - hi), 0.0f);
  float outside_squared = dot(outside, outside);
  if (outside_squared > 0.0f) {
    float positive = fmaxf(bound, 0.0f);
    bound = fsqrt(positive * positive + outside_squared);
  }
  if (bound > 0.0f) bound *= mesh.geometry_hi[3];
  float sphere_bound = length(p - ld3(mesh.enclosing_sphere)) - mesh.enclosing_sphere[3];
  return fmaxf(bound, fmaxf(geometry_bound, sphere_bound));
}

// object_distance — sdf.wgsl:35-41.  Returns the distance; *voxel gets the spacing term.
template <class SV, class C>
__device__ __forceinline__ float object_distance(const SV& sc, v3 p, uint32_t id, bool need_spacing,
                                                 float hit_eps, float* voxel, C& cnt) {
  const auto& obj = sc.objects[id];
  const irt_mesh& mesh = sc.meshes[obj.info[0]];
  float4 l = mat_mul(obj.inverse, p.x, p.y, p.z, 1.0f);
  float d = mesh_distance(sc, v3{l.x, l.y, l.z}, mesh, cnt) * __uint_as_float(obj.info[2]);
  if (d > 0.0f) d *= 0.9999f;
  *voxel = need_spacing ? __uint_as_float(obj.info[3]) : hit_eps;
  return d;
}

struct Distance {
  float distance, voxel;
  uint32_t object_id;
};

// scene_distance — sdf.wgsl:42-60.  The reference walks a stackless preorder
// hierarchy; its result is the minimum object distance (pruning only drops
// objects whose conservative lower bound exceeds the running minimum).  When
// the candidate grid is present, the same minimum is taken over the cell's
// candidate list with the same per-object bound test.
template <class SV, class C>
__device__ __forceinline__ bool scene_distance_grid(const SV& sc, float hit_eps, v3 p, bool need_spacing,
                                                    uint32_t hint, C& cnt, Distance* out) {
  const GridView& g = sc.grid;
  const float fx = (p.x - g.x0) * g.inv_cell, fy = (p.y - g.y0) * g.inv_cell, fz = (p.z - g.z0) * g.inv_cell;
  if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)g.nx && fy < (float)g.ny && fz < (float)g.nz))
    return false;
  const uint32_t cx = min((uint32_t)fx, g.nx - 1u), cy = min((uint32_t)fy, g.ny - 1u), cz = min((uint32_t)fz, g.nz - 1u);
  // The 8-byte list as two dwords; byte 0 is the count, byte k+1 the k-th id
  // (register selects rather than a dynamically indexed array).
  const uint2 L = reinterpret_cast<const uint2*>(g.cells)[((size_t)cz * g.ny + cy) * g.nx + cx];
  const uint32_t count = L.x & 0xffu;
  if (count == CELL_OVERFLOW) return false;
  float best = 1e20f, voxel = hit_eps;
  uint32_t closest = hint;
  for (uint32_t k = 0; k < count; ++k) {
    const uint32_t b = k + 1u;
    const uint32_t j = ((b < 4u ? L.x : L.y) >> ((b & 3u) * 8u)) & 0xffu;
    cnt.add(C_NODE, 1);
    const float4 lo = g.boxes[2 * j], hi = g.boxes[2 * j + 1];
    v3 delta = vmaxs(vmax(v3{lo.x - p.x, lo.y - p.y, lo.z - p.z}, v3{p.x - hi.x, p.y - hi.y, p.z - hi.z}), 0.0f);
    v3 lb = delta * 0.9998f;
    float bp = fmax0(best);
    if (dot(lb, lb) > bp * bp) continue;
    float vx;
    float d = object_distance(sc, p, j, need_spacing, hit_eps, &vx, cnt);
    if (d < best) {
      best = d;
      voxel = vx;
      closest = j;
    }
  }
  *out = Distance{best, voxel, closest};
  return true;
}

template <class SV, class C>
__device__ __forceinline__ Distance scene_distance(const SV& sc, uint32_t n_objects, float hit_eps, v3 p,
                                                   bool need_spacing, uint32_t hint, C& cnt) {
  if (sc.grid.cells) {
    Distance d;
    if (scene_distance_grid(sc, hit_eps, p, need_spacing, hint, cnt, &d)) return d;
  }
  float best = 1e20f, voxel = hit_eps;
  uint32_t closest = hint;
  if (hint < n_objects) best = object_distance(sc, p, hint, need_spacing, hit_eps, &voxel, cnt);
  const uint32_t n_nodes = n_objects * 2u - 1u;
  uint32_t cursor = 0u;
  while (cursor < n_nodes) {
    const irt_node& node = sc.nodes[cursor];
    cnt.add(C_NODE, 1);
    v3 delta = vmaxs(vmax(ld3(node.lo) - p, p - ld3(node.hi)), 0.0f);
    v3 lb = delta * 0.9998f;
    float bp = fmax0(best);
    if (dot(lb, lb) > bp * bp) {
      cursor = node.info[3];
      continue;
    }
    if (node.info[2] == 1u && node.info[0] != hint) {
      float vx;
      float d = object_distance(sc, p, node.info[0], need_spacing, hit_eps, &vx, cnt);
      if (d < best) {
        best = d;
        voxel = vx;
        closest = node.info[0];
      }
    }
    cursor++;
  }
  return Distance{best, voxel, closest};
}

// visibility — sdf.wgsl:61-81.  Conservative sphere-traced soft shadow.
template <class C>
__device__ __forceinline__ float visibility(const SceneView& sc, const irt_frame& F, v3 p, v3 n, v3 target,
                                            v3 end_normal, C& cnt) {
  const uint32_t n_objects = F.counts[3];
  const float eps0 = F.sdf[1];
  cnt.add(C_MARCH, 1);
  Distance clearance = scene_distance(sc, n_objects, eps0, p, true, 0xffffffffu, cnt);
  float epsilon = fmaxf(eps0, clearance.voxel * 0.65f);
  uint32_t closest = clearance.object_id;
  v3 origin = p + n * epsilon;
  v3 endpoint = target;
  if (dot(end_normal, end_normal) > 0.5f) {
    Distance e = scene_distance(sc, n_objects, eps0, target, true, 0xffffffffu, cnt);
    endpoint = endpoint + end_normal * fmaxf(eps0, e.voxel * 0.65f);
  }
  v3 delta = endpoint - origin;
  float limit = length(delta) - eps0;
  if (limit <= 0.0f) return 1.0f;
  v3 direction = normalize(delta);
  float t = eps0, result = 1.0f, previous = 1e20f;
  const uint32_t steps = f2u_sat(F.sdf[0]);
  const float k = F.sdf[2];
  for (uint32_t i = 0; i < steps; ++i) {
    if (t >= limit) return clampf(result, 0.0f, 1.0f);
    cnt.add(C_STEP, 1);
    Distance qd = scene_distance(sc, n_objects, eps0, origin + direction * t, false, closest, cnt);
    float h = qd.distance;
    closest = qd.object_id;
    if (h < eps0) return 0.0f;
    float y = h * h / (2.0f * previous);
    float d = fsqrt(fmaxf(h * h - y * y, 0.0f));
    if (t > y) result = fminf(result, k * d / (t - y));
    if (result < 0.001f) return 0.0f;
    previous = h;
    t += h;
  }
  return 0.0f;
}

// ------------------------------------------------------------ ray march
// visibility() split into a set-up and a single step so that marches can be
// scheduled independently of pixels.  Same operations in the same order.
struct MarchRay {
  float ox, oy, oz, limit;
  float dx, dy, dz;
  uint32_t closest;
};
static_assert(sizeof(MarchRay) == 32, "ray record is 32 bytes");

// The clearance queries and ray set-up of sdf.wgsl:62-74.  Returns false when
// the visibility is already decided (limit <= 0 -> 1).
template <class C>
__device__ __forceinline__ bool march_setup(const SceneView& sc, const irt_frame& F, v3 p, v3 n, v3 target,
                                            v3 end_normal, MarchRay* ray, C& cnt) {
  const uint32_t n_objects = F.counts[3];
  const float eps0 = F.sdf[1];
  cnt.add(C_MARCH, 1);
  Distance clearance = scene_distance(sc, n_objects, eps0, p, true, 0xffffffffu, cnt);
  float epsilon = fmaxf(eps0, clearance.voxel * 0.65f);
  v3 origin = p + n * epsilon;
  v3 endpoint = target;
  if (dot(end_normal, end_normal) > 0.5f) {
    Distance e = scene_distance(sc, n_objects, eps0, target, true, 0xffffffffu, cnt);
    endpoint = endpoint + end_normal * fmaxf(eps0, e.voxel * 0.65f);
  }
  v3 delta = endpoint - origin;
  float limit = length(delta) - eps0;
  if (limit <= 0.0f) return false;
  v3 direction = normalize(delta);
  *ray = MarchRay{origin.x, origin.y, origin.z, limit, direction.x, direction.y, direction.z, clearance.object_id};
  return true;
}

struct MarchState {
  float t, result, previous;
  uint32_t closest, i;
};
__device__ __forceinline__ MarchState march_begin(const irt_frame& F, const MarchRay& r) {
  return MarchState{F.sdf[1], 1.0f, 1e20f, r.closest, 0u};
}
// One iteration of the loop of sdf.wgsl:76-86.  Returns true when the march
// has finished, with the visibility in *vis.
template <class C>
__device__ __forceinline__ bool march_step(const SceneView& sc, const irt_frame& F, const MarchRay& r, MarchState& s,
                                           float* vis, C& cnt) {
  const float eps0 = F.sdf[1];
  if (s.i >= f2u_sat(F.sdf[0])) {
    *vis = 0.0f;
    return true;
  }
  if (s.t >= r.limit) {
    *vis = clampf(s.result, 0.0f, 1.0f);
    return true;
  }
  cnt.add(C_STEP, 1);
  const v3 q = v3{r.ox, r.oy, r.oz} + v3{r.dx, r.dy, r.dz} * s.t;
  Distance qd = scene_distance(sc, F.counts[3], eps0, q, false, s.closest, cnt);
  const float h = qd.distance;
  s.closest = qd.object_id;
  if (h < eps0) {
    *vis = 0.0f;
    return true;
  }
  const float y = h * h / (2.0f * s.previous);
  const float d = fsqrt(fmaxf(h * h - y * y, 0.0f));
  if (s.t > y) s.result = fminf(s.result, F.sdf[2] * d / (s.t - y));
  if (s.result < 0.001f) {
    *vis = 0.0f;
    return true;
  }
  s.previous = h;
  s.t += h;
  s.i += 1u;
  return false;
}

}  // namespace irt

// Shared device helpers for the gfx950 kernels: small float3 algebra, binary16
// decode, wave64 reductions and the C-ABI error channel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/powerfoam_amd.h"

namespace pf {

// ---------------------------------------------------------------- errors --
// Sets the thread's last-error text (returned by pf_last_error) and the code.
int set_error(int code, const char* fmt, ...);
int hip_check(hipError_t e, const char* what);

#define PF_HIP(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return pf::hip_check(_e, #call); \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int tiles_h(const pf_camera& c) { return 1 + (c.height - 1) / PF_TILE_WIDTH; }
inline int tiles_w(const pf_camera& c) { return 1 + (c.width - 1) / PF_TILE_WIDTH; }

// --------------------------------------------------------------- algebra --
struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

__device__ __forceinline__ float half_bits_to_float(uint16_t h) {
  return static_cast<float>(__builtin_bit_cast(_Float16, h));
}

constexpr int WAVE = 64;

// Orders this wave's LDS traffic: LDS operations of one wave execute in
// program order, so only the compiler has to be kept from moving them.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------- wave64 reductions --
// x of the lane that DPP control CTRL names (within the lane's row)
template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xf, 0xf, false));
}

// v_permlane32_swap_b32 a, b: the upper 32 lanes of a trade places with the
// lower 32 lanes of b; v_permlane16_swap_b32 a, b: the odd rows (16 lanes) of
// a with the even rows of b (gfx950). Written as instructions that read and
// write both registers: through the compiler builtins, hipcc (ROCm 7.2) adds
// the first result to itself in place of the pair's sum. A swap must follow a
// VALU write of its operands by two wait states (the s_nop 1 the compiler
// puts there for the builtins; it does not look inside asm).
__device__ __forceinline__ void permlane32_swap(float& a, float& b) {
  asm("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(a), "+v"(b));
}
__device__ __forceinline__ void permlane16_swap(float& a, float& b) {
  asm("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(a), "+v"(b));
}

// Full-wave sum, every lane must be active; the result is uniform (each
// step adds a lane's value to its partner's and the partner does the same
// sum). All in registers: DPP pairs i, i^1 and i, i^2 (quad_perm), the two
// quads of a half-row (row_half_mirror), the halves of a row (row_ror:8),
// then the rows and half-waves swap.
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);  // row_half_mirror
  v += dpp<0x128>(v);  // row_ror:8
  float a = v, b = v;
  permlane16_swap(a, b);
  v = a + b;
  a = v;
  b = v;
  permlane32_swap(a, b);
  return a + b;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// run of tile groups so neighbouring tiles (which share sites) share an L2.
// Returns the index of workgroup b's group among n_groups (valid: b maps to
// one; launch a whole number of rounds of 8 workgroups).
__device__ __forceinline__ int xcd_group(int b, int n_groups, bool& valid) {
  const int per = (n_groups + 7) >> 3;
  const int xcd = b & 7, k = b >> 3;
  const int full = n_groups - (per - 1) * 8;  // XCDs that own `per` groups
  valid = xcd < full ? k < per : k < per - 1;
  return xcd < full ? xcd * per + k : full * per + (xcd - full) * (per - 1) + k;
}

// Workgroups for n work items at nw waves each: a whole number of XCD rounds.
static inline int xcd_grid(int n, int nw) { return (((n + nw - 1) / nw + 7) / 8) * 8; }

// Atomic adds with the address space spelled out: p must address LDS
// (lds_add, a ds_add_f32) or global memory (global_add). Plain atomicAdd on
// pointers that two branches select between can be merged into one generic
// (flat) atomic, which goes through the vector memory path even for LDS.
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(1))) float global_float;
__device__ __forceinline__ void lds_add(float* p, float v) {
  __hip_atomic_fetch_add((lds_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void global_add(float* p, float v) {
  __hip_atomic_fetch_add((global_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// a load that p must serve from LDS (ds_read), for the same reason
__device__ __forceinline__ int lds_load(const int* p) { return *(const __attribute__((address_space(3))) int*)p; }
__device__ __forceinline__ float lds_load(const float* p) {
  return *(const __attribute__((address_space(3))) float*)p;
}

// Uniform-broadcast of lane `l`'s 32-bit value.
__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float readlane(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// ---------------------------------------------------------- fast math --
constexpr float LOG2E = 1.4426950408889634f;

__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * LOG2E); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// num / den from v_rcp_f32 plus one Newton step on the quotient: within about
// half an ulp, so interval ends agree with a correctly rounded division except
// at exact ties (the interval bounds decide which sites a ray shades).
__device__ __forceinline__ float acc_div(float num, float den) {
  const float r = __builtin_amdgcn_rcpf(den);
  const float q = num * r;
  return __builtin_fmaf(__builtin_fmaf(-den, q, num), r, q);
}
__device__ __forceinline__ float sel_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float sel_max(float a, float b) { return a > b ? a : b; }

// ray-sphere (rendering_math.py:4-23) with the perpendicular-distance
// discriminant; oc = o - c, |d| = 1
__device__ __forceinline__ bool sphere_hit(f3 oc, f3 d, float r, float& tn, float& tf) {
  const float b = dot(oc, d);
  const f3 cr = cross(oc, d);
  const float h2 = r * r - dot(cr, cr);
  tn = 0.0f;
  tf = 0.0f;
  if (h2 < 0.0f) return false;
  const float h = __builtin_sqrtf(h2);
  const float far = -b + h, near = -b - h;
  if (near < 0.0f && far < 0.0f) return false;
  tn = near < 0.0f ? 0.0f : near;
  tf = far;
  return true;
}



}  // namespace pf

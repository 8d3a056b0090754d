// Scalar float3 arithmetic for the kernels, with a fixed operation order.
//
// Every kernel is compiled with -ffp-contract=off and correctly rounded
// divide/sqrt, so each expression below is one IEEE-754 single-precision
// operation per step.  dot/cross follow the OpenCL definitions evaluated left
// to right; normalize(v) = v * (1/sqrt(dot(v,v))).  fminf/fmaxf are IEEE
// minNum/maxNum (the non-NaN operand wins).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

struct v3 {
    float x, y, z;
};

__host__ __device__ __forceinline__ v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
__host__ __device__ __forceinline__ v3 xyz(float4 a) { return v3{a.x, a.y, a.z}; }
__host__ __device__ __forceinline__ v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
__host__ __device__ __forceinline__ v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ v3 operator*(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ __forceinline__ v3 operator*(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ v3 normalize(v3 v) {
    float inv = 1.0f / __builtin_sqrtf(dot(v, v));
    return v * inv;
}
__device__ __forceinline__ v3 vmin(v3 a, v3 b) { return v3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ v3 vmax(v3 a, v3 b) { return v3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.f), 1.f); }
__device__ __forceinline__ v3 recip(v3 v) { return v3{1.0f / v.x, 1.0f / v.y, 1.0f / v.z}; }

// Order-preserving int key of a float (for integer atomics min/max).
__device__ __forceinline__ int32_t fkey(float f) {
    int32_t i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float funkey(int32_t k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }

// Lane helpers for wave64.
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// p, known to point into global memory, as a pointer the compiler can see is
// global: loads and stores through it become global_* instead of flat_*
// instructions (a flat access also counts against the LDS wait counter, so
// an LDS stack read would wait for outstanding node loads).  For pointers
// read from device tables, whose address space the compiler cannot infer.
// (Through an integer: the compiler folds a plain generic->global->generic
// cast pair back to the generic pointer.)
template <class T>
__device__ __forceinline__ T* gptr(T* p) {
    return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}


// Device-coherent atomics for the look-back protocol.  A status word carries
// its value, so relaxed agent-scope accesses suffice: no release/acquire
// fence (on gfx950 those write back / invalidate the XCD's whole L2).
// Results written next to the protocol are consumed only by later launches.
__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Complex element helpers shared by the LU and triangular-solve kernels:
// T = double (FP64 path) or float (FP32 LU of the mixed-precision path).
#pragma once
#include <hip/hip_runtime.h>

namespace mh {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// element traits: T = double (FP64 LU) or float (FP32 LU of the mixed path)
template <class T>
struct Cx;
template <>
struct Cx<double> {
    typedef double2 t;
    typedef d4 acc;
    // v_mfma_f64_16x16x4f64: result register r of lane l is row 4r + l/16
    __device__ __forceinline__ static int row(int r, int l) { return 4 * r + (l >> 4); }
    __device__ __forceinline__ static acc mfma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Cx<float> {
    typedef float2 t;
    typedef f4 acc;
    // v_mfma_f32_16x16x4f32: result register r of lane l is row 4 (l/16) + r
    __device__ __forceinline__ static int row(int r, int l) { return 4 * (l >> 4) + r; }
    __device__ __forceinline__ static acc mfma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <class C2>
__device__ __forceinline__ C2 mkc(decltype(C2().x) a, decltype(C2().x) b) {
    C2 o;
    o.x = a;
    o.y = b;
    return o;
}

// Raw buffer loads and stores of one complex entry: a per-lane byte offset
// plus a wave-uniform one (scalar register), so strided accesses need no
// 64-bit address registers, and a run of stores writes no address register
// between them (a register that still holds a store's data is not rewritten,
// so the stores are not made to wait for each other's completion).
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float2 buf_load(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float2) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ double2 buf_load(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, double2) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ void buf_store(float2 v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ void buf_store(double2 v, __amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, (int)vo, (int)so, 0);
}


template <class C2>
__device__ __forceinline__ C2 cmul(C2 a, C2 b) {
    return mkc<C2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a - b*c
template <class C2>
__device__ __forceinline__ C2 cfms(C2 a, C2 b, C2 c) {
    return mkc<C2>(a.x - (b.x * c.x - b.y * c.y), a.y - (b.x * c.y + b.y * c.x));
}
template <class C2>
__device__ __forceinline__ C2 crecip(C2 a) {
    // Smith's algorithm for 1/a
    typedef decltype(C2().x) T;
    if (fabs(a.x) >= fabs(a.y)) {
        const T r = a.y / a.x, d = a.x + a.y * r;
        return mkc<C2>(T(1) / d, -r / d);
    }
    const T r = a.x / a.y, d = a.y + a.x * r;
    return mkc<C2>(r / d, T(-1) / d);
}
template <class C2>
__device__ __forceinline__ C2 cdiv(C2 a, C2 b) { return cmul(a, crecip(b)); }

}  // namespace mh

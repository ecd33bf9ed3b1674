// rm_device.h -- device-side number formats shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/rotormap_hip.h"

namespace rm {

// fp32 -> OCP E4M3FN, round to nearest even, |x| > 448 saturates to 448,
// NaN -> 0x7F.  Written from the OCP 8-bit floating point specification:
// below 2^-6 the codes are multiples of 2^-9; above, 3 mantissa bits.
__device__ __forceinline__ uint8_t f32_to_e4m3(float f) {
    uint32_t u = __float_as_uint(f);
    uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    float a = __builtin_fabsf(f);
    if (a != a) return (uint8_t)(sign | 0x7fu);
    if (a > 448.0f) return (uint8_t)(sign | 0x7eu);
    if (a < 0.015625f) {
        float q = __builtin_rintf(a * 512.0f);
        return (uint8_t)(sign | (uint8_t)q);
    }
    int e = (int)((__float_as_uint(a) >> 23) & 0xffu) - 127;
    float mant = __builtin_amdgcn_ldexpf(a, -e);
    int qi = (int)__builtin_rintf((mant - 1.0f) * 8.0f);
    if (qi == 8) { e += 1; qi = 0; }
    int code = ((e + 7) << 3) | qi;
    if (code > 0x7e) code = 0x7e;
    return (uint8_t)(sign | (uint8_t)code);
}

__device__ __forceinline__ float e4m3_to_f32(uint32_t b) {
    uint32_t e = (b >> 3) & 0xfu, m = b & 7u;
    float v;
    if (e == 0) v = (float)m * 0.001953125f;  // m * 2^-9
    else v = __uint_as_float(((e + 120u) << 23) | (m << 20));
    if (e == 15 && m == 7) v = __uint_as_float(0x7fc00000u);
    return (b & 0x80u) ? -v : v;
}

// byte b (0..3) of w as e4m3 -> fp32 with the hardware converter (exact)
__device__ __forceinline__ float e4m3_byte_to_f32(uint32_t w, int b) {
    switch (b) {
    case 0: return __builtin_amdgcn_cvt_f32_fp8((int)w, 0);
    case 1: return __builtin_amdgcn_cvt_f32_fp8((int)w, 1);
    case 2: return __builtin_amdgcn_cvt_f32_fp8((int)w, 2);
    default: return __builtin_amdgcn_cvt_f32_fp8((int)w, 3);
    }
}

__device__ __forceinline__ uint16_t f32_to_f16_bits(float f) {
    _Float16 h = (_Float16)f;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_bits_to_f32(uint16_t h) {
    return (float)__builtin_bit_cast(_Float16, h);
}

// store one value in the requested storage type at element offset i
__device__ __forceinline__ void store_elem(void *p, int dtype, int64_t i, float v) {
    if (dtype == RM_F32) ((float *)p)[i] = v;
    else if (dtype == RM_F16) ((uint16_t *)p)[i] = f32_to_f16_bits(v);
    else ((uint8_t *)p)[i] = f32_to_e4m3(f16_bits_to_f32(f32_to_f16_bits(v)));
}

// (value desc, column asc) total order
__device__ __forceinline__ bool beats(float va, int32_t ca, float vb, int32_t cb) {
    return va > vb || (va == vb && ca < cb);
}

}  // namespace rm

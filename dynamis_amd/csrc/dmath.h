// Device FP32 vector and quaternion math of the step kernels. Expressions keep the operation
// order of the reference WGSL (crates/dynamis-shader/shaders/core.wgsl and WGSL built-ins); the
// library is compiled with FP contraction off and IEEE division / square root, so each product
// and sum rounds once, where it is written.
#pragma once

#include "common.h"

namespace dyn {

using v3 = float3;
using q4 = float4;

__device__ __forceinline__ v3 mk(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ v3 ld3(const float* p) { return make_float3(p[0], p[1], p[2]); }
__device__ __forceinline__ void st3(float* p, v3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
__device__ __forceinline__ q4 ld4(const float* p) { return make_float4(p[0], p[1], p[2], p[3]); }
__device__ __forceinline__ void st4(float* p, q4 q) { p[0] = q.x; p[1] = q.y; p[2] = q.z; p[3] = q.w; }
__device__ __forceinline__ v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ v3 mul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 smul(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ v3 dv(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ v3 mulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 absv(v3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__device__ __forceinline__ float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross3(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len3(v3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ v3 norm3(v3 a) { return dv(a, len3(a)); }
__device__ __forceinline__ float comp(v3 a, uint32_t i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ bool zero3(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
// WGSL max/min on ordinary (non-NaN) values.
__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fmn(fmx(v, lo), hi); }
__device__ __forceinline__ v3 zero3v() { return mk(0.0f, 0.0f, 0.0f); }

// core.wgsl:132-137
__device__ __forceinline__ q4 qmul(q4 a, q4 b) {
    const v3 av = mk(a.x, a.y, a.z), bv = mk(b.x, b.y, b.z);
    const v3 xyz = add(add(smul(a.w, bv), smul(b.w, av)), cross3(av, bv));
    return make_float4(xyz.x, xyz.y, xyz.z, a.w * b.w - dot3(av, bv));
}
__device__ __forceinline__ q4 qconj(q4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }
// core.wgsl:143-147
__device__ __forceinline__ v3 qrot(q4 q, v3 v) {
    const v3 u = mk(q.x, q.y, q.z);
    const float s = q.w;
    return add(add(smul(2.0f * dot3(u, v), u), smul(s * s - dot3(u, u), v)),
               smul(2.0f * s, cross3(u, v)));
}
__device__ __forceinline__ q4 qnorm(q4 q) {
    const float n = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return make_float4(q.x / n, q.y / n, q.z / n, q.w / n);
}
// Sine and cosine of an FP32 argument, correctly rounded: evaluated in FP64 and rounded once.
// (WGSL leaves their precision to the implementation; this choice makes every build of the
// step agree bit for bit.)
__device__ __forceinline__ float sin_rn(float x) { return static_cast<float>(sin(static_cast<double>(x))); }
__device__ __forceinline__ float cos_rn(float x) { return static_cast<float>(cos(static_cast<double>(x))); }

// core.wgsl:170-178
__device__ __forceinline__ v3 rotate_about(v3 rotation, v3 v) {
    const float angle = len3(rotation);
    if (angle < 1e-8f) return v;
    const v3 axis = dv(rotation, angle);
    const v3 offset = cross3(axis, v);
    return add(add(v, smul(sin_rn(angle), offset)), smul(1.0f - cos_rn(angle), cross3(axis, offset)));
}
// core.wgsl:180-186
__device__ __forceinline__ v3 sign_normalize(v3 v) {
    const float n = len3(v);
    if (n > 1e-8f) return dv(v, n);
    return mk(0.0f, 1.0f, 0.0f);
}
// core.wgsl:355-361, tensors stored [xx, xy, xz, yy, yz, zz]
__device__ __forceinline__ v3 sym_apply(const float* t, v3 v) {
    return mk(t[0] * v.x + t[1] * v.y + t[2] * v.z, t[1] * v.x + t[3] * v.y + t[4] * v.z,
              t[2] * v.x + t[4] * v.y + t[5] * v.z);
}
// core.wgsl:6-12: i32 fixed-point words, wrapping adds
__device__ __forceinline__ uint32_t solver_word(float value, float scale) {
    const float c = clampf(value * scale, -2.0e9f, 2.0e9f);
    return static_cast<uint32_t>(static_cast<int32_t>(c));
}
__device__ __forceinline__ float solver_value(uint32_t word, float scale) {
    return static_cast<float>(static_cast<int32_t>(word)) / scale;
}

}  // namespace dyn

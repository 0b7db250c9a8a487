/*
 * omath.h - FP32 vector/quaternion helpers of the CPU oracle. TEST INFRASTRUCTURE ONLY.
 *
 * Each helper spells out the expression of the reference WGSL it restates
 * (crates/dynamis-shader/shaders/core.wgsl and WGSL built-ins), left to right, so that with
 * -ffp-contract=off every product and sum rounds once in the written order.
 */
#ifndef DYN_OMATH_H
#define DYN_OMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } q4;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vload(const float* p) { return V(p[0], p[1], p[2]); }
static inline void vstore(float* p, v3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
static inline q4 qload(const float* p) { q4 q = {p[0], p[1], p[2], p[3]}; return q; }
static inline void qstore(float* p, q4 q) { p[0] = q.x; p[1] = q.y; p[2] = q.z; p[3] = q.w; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline v3 vmul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 smul(float s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline v3 vdiv(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 vmulv(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vabs(v3 a) { return V(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vnorm(v3 a) { return vdiv(a, vlen(a)); }
static inline float vcomp(v3 a, uint32_t i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline int vzero(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
static inline float fmaxo(float a, float b) { return a > b ? a : b; }
static inline float fmino(float a, float b) { return a < b ? a : b; }
static inline float fclamp(float v, float lo, float hi) { return fmino(fmaxo(v, lo), hi); }

/* core.wgsl:132-137 */
static inline q4 qmul(q4 a, q4 b) {
    v3 av = V(a.x, a.y, a.z), bv = V(b.x, b.y, b.z);
    v3 xyz = vadd(vadd(smul(a.w, bv), smul(b.w, av)), vcross(av, bv));
    q4 r = {xyz.x, xyz.y, xyz.z, a.w * b.w - vdot(av, bv)};
    return r;
}
/* core.wgsl:139-141 */
static inline q4 qconj(q4 q) { q4 r = {-q.x, -q.y, -q.z, q.w}; return r; }
/* core.wgsl:143-147: 2 dot(u,v) u + (s^2 - dot(u,u)) v + 2 s cross(u,v) */
static inline v3 qrot(q4 q, v3 v) {
    v3 u = V(q.x, q.y, q.z);
    float s = q.w;
    return vadd(vadd(smul(2.0f * vdot(u, v), u), smul(s * s - vdot(u, u), v)),
                smul(2.0f * s, vcross(u, v)));
}
static inline float qdot(q4 a, q4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
static inline q4 qnorm(q4 q) {
    float n = sqrtf(qdot(q, q));
    q4 r = {q.x / n, q.y / n, q.z / n, q.w / n};
    return r;
}
/* sine and cosine of an FP32 argument, correctly rounded (FP64 evaluation, one rounding) */
static inline float sin_rn(float x) { return (float)sin((double)x); }
static inline float cos_rn(float x) { return (float)cos((double)x); }

/* core.wgsl:170-178 */
static inline v3 rotate_about(v3 rotation, v3 v) {
    float angle = vlen(rotation);
    if (angle < 1e-8f) return v;
    v3 axis = vdiv(rotation, angle);
    v3 offset = vcross(axis, v);
    return vadd(vadd(v, smul(sin_rn(angle), offset)), smul(1.0f - cos_rn(angle), vcross(axis, offset)));
}
/* core.wgsl:180-186 */
static inline v3 sign_normalize(v3 v) {
    float n = vlen(v);
    if (n > 1e-8f) return vdiv(v, n);
    return V(0.0f, 1.0f, 0.0f);
}
/* core.wgsl:355-361 */
static inline v3 symmetric_apply(const float* t, v3 v) {
    return V(t[0] * v.x + t[1] * v.y + t[2] * v.z, t[1] * v.x + t[3] * v.y + t[4] * v.z,
             t[2] * v.x + t[4] * v.y + t[5] * v.z);
}
/* core.wgsl:6-12 */
static inline uint32_t solver_word(float value, float scale) {
    float c = fclamp(value * scale, -2.0e9f, 2.0e9f);
    return (uint32_t)(int32_t)c;
}
static inline float solver_value(uint32_t word, float scale) { return (float)(int32_t)word / scale; }

static inline uint32_t f2bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

#endif

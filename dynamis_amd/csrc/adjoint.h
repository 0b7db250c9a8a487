// The differentiable step: a replay of one step's velocity and position phases over a recorded
// ("frozen") contact set, and its hand-written reverse.
//
// The step it differentiates is the production step (integrate, warm start, Jacobi iterations
// with mass splitting, advance, position corrections; solver_*.wgsl, substep_*.wgsl,
// position_*.wgsl) with the contact set frozen at its forward values: which pairs touch, their
// normals, features, body-local anchors, the relay of warm-start impulses between steps and the
// split counts. What stays differentiable:
//   - every body's pose and velocities, through integration (with the gyroscopic solve), the
//     solver and the advance;
//   - the contact geometry: a point's depth moves with the two anchors along the frozen normal,
//     depth(x) = depth_fwd + n . ((arm_a(x) - arm_a(x_fwd)) - (arm_b(x) - arm_b(x_fwd))), arm =
//     position + rotation * local anchor; the solver's lever arms rotate with the bodies;
//   - the accumulated impulses, within a step and across steps through the relay;
//   - the mass properties (a per-body scale s of mass and inertia at fixed shape) and the
//     friction coefficients of the colliders (through the contact's combined friction).
// Branches (the normal clamp, the friction-cone clamps, the position phase's active points) take
// the derivative of the branch the forward value selects. The fixed-point velocity and position
// words of the production step are read as the real numbers they quantise.
//
// Everything here is templated on the scalar type and on an executor that runs a phase over a
// range and separates phases: one workgroup per world on the device (float), or serial loops on
// the host (double, for the central-difference checks). Soft contacts (relaxation > 0) and rolling
// or spin friction are outside the model: the recorder refuses them.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define ADJ_FN __host__ __device__ inline
#else
#define ADJ_FN inline
#endif

namespace dyn_adj {

// ---- record layouts (float words) -------------------------------------------------------------

constexpr uint32_t kStateFloats = 16;    // position xyz _, orientation xyzw, velocity xyz _, spin xyz _
constexpr uint32_t kBodyFloats = 20;     // see body_* below
constexpr uint32_t kContactFloats = 76;  // 12 of header, 16 per manifold point
constexpr uint32_t kAccFloats = 12;      // normal[4], tangent 1[4], tangent 2[4]
constexpr uint32_t kRowFloats = 48;      // per point: ra xyz, nmass, rb xyz, truth, m1, m2, _, _
constexpr uint32_t kSideFloats = 12;     // per block side: orientation adjoint 4, inverse inertia 6, inverse mass 1, _
constexpr uint32_t kPointAdj = 16;       // per point: ra 3, nmass, rb 3, truth, m1, m2, target, depth, _ x4
constexpr uint32_t kSlotFloats = 16;     // per block side: position 3, orientation 4, velocity 3, spin 3, _ x3

// body record
constexpr uint32_t kBodyIm = 0, kBodyGravityScale = 1, kBodyLinearDamping = 2, kBodyAngularDamping = 3;
constexpr uint32_t kBodyInvInertia = 4, kBodyInertia = 10, kBodyCom = 16, kBodyFlags = 19;
constexpr uint32_t kBodySimulates = 1u, kBodyMoves = 2u, kBodyInert = 4u;

// contact record header: fb, sb, flags, collider a | normal xyz, friction | dmu/da, dmu/db,
// restitution, collider b
constexpr uint32_t kFlagPoints = 7u, kFlagInertF = 8u, kFlagInertS = 16u, kFlagResolves = 32u;
// per point (at 12 + 16 * p): local a xyz, depth | local b xyz, _ | position xyz, source | an, at1, at2, _
constexpr uint32_t kNoSource = 0xFFFFFFFFu;

struct AdjParams {
    float dt, h;
    float gravity[3];
    float slop, relaxation, restitution_threshold, max_velocity, max_angular_velocity;
    uint32_t substeps, iterations, position_iterations, friction_combine;
};

struct AdjRange {  // one world's simulated rows and collider slots
    uint32_t row_base, row_count, collider_base, collider_count;
};

ADJ_FN uint32_t fbits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}
ADJ_FN float bitsf(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// ---- scalars ----------------------------------------------------------------------------------

ADJ_FN float val(float x) { return x; }
ADJ_FN double val(double x) { return x; }
ADJ_FN float asqrt(float x) { return sqrtf(x); }
ADJ_FN double asqrt(double x) { return sqrt(x); }
// The step's sine and cosine are evaluated in FP64 and rounded once.
ADJ_FN float asin_(float x) { return static_cast<float>(sin(static_cast<double>(x))); }
ADJ_FN double asin_(double x) { return sin(x); }
ADJ_FN float acos_(float x) { return static_cast<float>(cos(static_cast<double>(x))); }
ADJ_FN double acos_(double x) { return cos(x); }

// Forward-mode dual numbers: a value and its derivatives along N seeded directions. They give
// the integrator's local Jacobian, whose gyroscopic solve (four Rodrigues rotations) is exact
// this way without a hand-derived reverse.
template <int N, class S = float>
struct Dual {
    S v;
    S d[N];
    ADJ_FN Dual() : v(S(0)) {
        for (int i = 0; i < N; ++i) d[i] = S(0);
    }
    ADJ_FN Dual(S x) : v(x) {  // NOLINT: constants convert implicitly
        for (int i = 0; i < N; ++i) d[i] = S(0);
    }
    template <class U, class = decltype(S(U()))>
    ADJ_FN Dual(U x) : v(S(x)) {  // NOLINT
        for (int i = 0; i < N; ++i) d[i] = S(0);
    }
};
template <int N, class S> ADJ_FN S val(const Dual<N, S>& x) { return x.v; }
template <int N, class S> ADJ_FN Dual<N, S> operator+(const Dual<N, S>& a, const Dual<N, S>& b) {
    Dual<N, S> r;
    r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N, class S> ADJ_FN Dual<N, S> operator-(const Dual<N, S>& a, const Dual<N, S>& b) {
    Dual<N, S> r;
    r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N, class S> ADJ_FN Dual<N, S> operator-(const Dual<N, S>& a) {
    Dual<N, S> r;
    r.v = -a.v;
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N, class S> ADJ_FN Dual<N, S> operator*(const Dual<N, S>& a, const Dual<N, S>& b) {
    Dual<N, S> r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N, class S> ADJ_FN Dual<N, S> operator/(const Dual<N, S>& a, const Dual<N, S>& b) {
    Dual<N, S> r;
    r.v = a.v / b.v;
    const S inv = S(1) / b.v;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
template <int N, class S> ADJ_FN Dual<N, S> asqrt(const Dual<N, S>& a) {
    Dual<N, S> r;
    r.v = asqrt(a.v);
    const S k = r.v > S(0) ? S(0.5) / r.v : S(0);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * k;
    return r;
}
template <int N, class S> ADJ_FN Dual<N, S> asin_(const Dual<N, S>& a) {
    Dual<N, S> r;
    r.v = asin_(a.v);
    const S c = acos_(a.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * c;
    return r;
}
template <int N, class S> ADJ_FN Dual<N, S> acos_(const Dual<N, S>& a) {
    Dual<N, S> r;
    r.v = acos_(a.v);
    const S sn = -asin_(a.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * sn;
    return r;
}

// ---- vectors and quaternions ------------------------------------------------------------------

template <class T> struct V3 { T x, y, z; };
template <class T> struct Q4 { T x, y, z, w; };

template <class T> ADJ_FN V3<T> mk3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> ADJ_FN V3<T> zero3() { return V3<T>{T(0.0f), T(0.0f), T(0.0f)}; }
template <class T> ADJ_FN V3<T> operator+(const V3<T>& a, const V3<T>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> ADJ_FN V3<T> operator-(const V3<T>& a, const V3<T>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> ADJ_FN V3<T> operator-(const V3<T>& a) { return {-a.x, -a.y, -a.z}; }
template <class T> ADJ_FN V3<T> operator*(const V3<T>& a, const T& s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> ADJ_FN V3<T> operator/(const V3<T>& a, const T& s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> ADJ_FN V3<T>& operator+=(V3<T>& a, const V3<T>& b) {
    a = a + b;
    return a;
}
template <class T> ADJ_FN V3<T>& operator-=(V3<T>& a, const V3<T>& b) {
    a = a - b;
    return a;
}
template <class T> ADJ_FN T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> ADJ_FN V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> ADJ_FN T alen(const V3<T>& a) { return asqrt(dot(a, a)); }
template <class T> ADJ_FN V3<T> vnorm(const V3<T>& a) { return a / alen(a); }
template <class T> ADJ_FN bool is_zero(const V3<T>& a) { return val(a.x) == 0 && val(a.y) == 0 && val(a.z) == 0; }

template <class T> ADJ_FN V3<T> qv(const Q4<T>& q) { return {q.x, q.y, q.z}; }
template <class T> ADJ_FN Q4<T> qconj(const Q4<T>& q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class T> ADJ_FN Q4<T> qmul(const Q4<T>& a, const Q4<T>& b) {  // core.wgsl quaternion product
    const V3<T> av = qv(a), bv = qv(b);
    const V3<T> xyz = bv * a.w + av * b.w + cross(av, bv);
    return {xyz.x, xyz.y, xyz.z, a.w * b.w - dot(av, bv)};
}
template <class T> ADJ_FN V3<T> qrot(const Q4<T>& q, const V3<T>& v) {  // core.wgsl:143-147
    const V3<T> u = qv(q);
    const T s = q.w;
    return u * (T(2.0f) * dot(u, v)) + v * (s * s - dot(u, u)) + cross(u, v) * (T(2.0f) * s);
}
template <class T> ADJ_FN Q4<T> qnorm(const Q4<T>& q) {
    const T n = asqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return {q.x / n, q.y / n, q.z / n, q.w / n};
}
template <class T> ADJ_FN V3<T> sym(const T* t, const V3<T>& v) {  // symmetric 3x3 in 6 words
    return {t[0] * v.x + t[1] * v.y + t[2] * v.z, t[1] * v.x + t[3] * v.y + t[4] * v.z,
            t[2] * v.x + t[4] * v.y + t[5] * v.z};
}
// apply_inverse_inertia: q * (I_body^-1 (q^-1 v))
template <class T> ADJ_FN V3<T> inertia_apply(const Q4<T>& q, const T* ib, const V3<T>& v) {
    return qrot(q, sym(ib, qrot(qconj(q), v)));
}

// ---- reverse of the elementary maps -------------------------------------------------------------

// y = qrot(q, v): accumulates q̄ and v̄ for ȳ.
template <class T> ADJ_FN void qrot_vjp(const Q4<T>& q, const V3<T>& v, const V3<T>& yb, Q4<T>* qb, V3<T>* vb) {
    const V3<T> u = qv(q);
    const T s = q.w;
    const T uy = dot(u, yb), uv = dot(u, v), vy = dot(v, yb);
    const V3<T> ub = v * (T(2.0) * uy) + yb * (T(2.0) * uv) - u * (T(2.0) * vy) + cross(v, yb) * (T(2.0) * s);
    const T sb = T(2.0) * s * vy + T(2.0) * dot(yb, cross(u, v));
    if (qb) {
        qb->x += ub.x;
        qb->y += ub.y;
        qb->z += ub.z;
        qb->w += sb;
    }
    if (vb) *vb += u * (T(2.0) * uy) + yb * (s * s - dot(u, u)) + cross(yb, u) * (T(2.0) * s);
}
// y = sym(t, v)
template <class T> ADJ_FN void sym_vjp(const T* t, const V3<T>& v, const V3<T>& yb, T* tb, V3<T>* vb) {
    if (tb) {
        tb[0] += yb.x * v.x;
        tb[1] += yb.x * v.y + yb.y * v.x;
        tb[2] += yb.x * v.z + yb.z * v.x;
        tb[3] += yb.y * v.y;
        tb[4] += yb.y * v.z + yb.z * v.y;
        tb[5] += yb.z * v.z;
    }
    if (vb) *vb += sym(t, yb);
}
// y = inertia_apply(q, ib, v)
template <class T>
ADJ_FN void inertia_vjp(const Q4<T>& q, const T* ib, const V3<T>& v, const V3<T>& yb, Q4<T>* qb, T* ibb, V3<T>* vb) {
    const Q4<T> qc = qconj(q);
    const V3<T> a = qrot(qc, v);
    const V3<T> b = sym(ib, a);
    V3<T> bb = zero3<T>(), ab = zero3<T>();
    Q4<T> q1 = {T(0.0), T(0.0), T(0.0), T(0.0)}, q2 = q1;
    qrot_vjp(q, b, yb, &q1, &bb);
    sym_vjp(ib, a, bb, ibb, &ab);
    qrot_vjp(qc, v, ab, &q2, vb);
    if (qb) {
        qb->x += q1.x - q2.x;
        qb->y += q1.y - q2.y;
        qb->z += q1.z - q2.z;
        qb->w += q1.w + q2.w;
    }
}
// y = u / |u| (four components)
template <class T> ADJ_FN Q4<T> qnorm_vjp(const Q4<T>& u, const Q4<T>& yb) {
    const T n = asqrt(u.x * u.x + u.y * u.y + u.z * u.z + u.w * u.w);
    const Q4<T> y = {u.x / n, u.y / n, u.z / n, u.w / n};
    const T d = y.x * yb.x + y.y * yb.y + y.z * yb.z + y.w * yb.w;
    return {(yb.x - y.x * d) / n, (yb.y - y.y * d) / n, (yb.z - y.z * d) / n, (yb.w - y.w * d) / n};
}

// solver_contact_block.wgsl make_tangents
template <class T> ADJ_FN void make_tangents(const V3<T>& n, V3<T>* t1, V3<T>* t2) {
    const V3<T> t = fabs(static_cast<double>(val(n.x))) > 0.9 ? mk3<T>(T(0.0f), T(1.0f), T(0.0f))
                                                                : mk3<T>(T(1.0f), T(0.0f), T(0.0f));
    *t1 = vnorm(cross(n, t));
    *t2 = cross(n, *t1);
}

// ---- record access ----------------------------------------------------------------------------

template <class T> ADJ_FN V3<T> ld3(const float* p) { return {T(p[0]), T(p[1]), T(p[2])}; }
template <class T> ADJ_FN V3<T> ld3(const T* p, int) { return {p[0], p[1], p[2]}; }
template <class T> ADJ_FN Q4<T> ld4(const T* p) { return {p[0], p[1], p[2], p[3]}; }
template <class T> ADJ_FN void st3(T* p, const V3<T>& v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
template <class T> ADJ_FN void st4(T* p, const Q4<T>& q) {
    p[0] = q.x;
    p[1] = q.y;
    p[2] = q.z;
    p[3] = q.w;
}

ADJ_FN uint32_t body_flags(const float* b) { return fbits(b[kBodyFlags]); }
ADJ_FN uint32_t contact_u(const float* c, uint32_t i) { return fbits(c[i]); }
ADJ_FN uint32_t contact_points(const float* c) { return contact_u(c, 2) & kFlagPoints; }
ADJ_FN bool contact_active(const float* c) {  // resolves, and not both bodies inert
    const uint32_t f = contact_u(c, 2);
    return (f & kFlagResolves) && (f & kFlagPoints) && !((f & kFlagInertF) && (f & kFlagInertS));
}
ADJ_FN const float* point_rec(const float* c, uint32_t p) { return c + 12u + 16u * p; }

// friction_combine (core.wgsl material_combine): combined value and its two partials
template <class T> ADJ_FN T combine(T a, T b, uint32_t mode) {
    if (mode == 1u) return val(a) < val(b) ? a : b;
    if (mode == 2u) return val(a) > val(b) ? a : b;
    if (mode == 3u) return (a + b) * T(0.5f);
    return asqrt(a * b);
}

// ---- the workspace of one step ----------------------------------------------------------------

// All arrays are indexed by global row, contact or collider, so the worlds of a batch share them
// and each world's workgroup touches only its own entries.
template <class T>
struct Work {
    uint32_t n, cap;            // rows, contact capacity
    T* x;                       // n x 16: live state of the replay
    T* vel;                     // substeps x (3 + iterations) x n x 8: velocities at phase boundaries
    T* pose;                    // substeps x 2 x n x 8: pose at substep start, after advance
    T* res;                     // substeps x position_iterations x n x 4: resolution before the iteration, count
    T* acc;                     // substeps x (iterations + 1) x cap x 12: impulses entering each sweep
    T* rows;                    // substeps x cap x 48: per-substep row quantities
    T* depth;                   // cap x 4: the model depth of each point this step
    T* target;                  // cap x 4
    T* slot;                    // cap x 2 x 16: per block side contributions of one phase
    uint32_t* list_start;       // n: first entry of a row's block list (world-relative scan)
    uint32_t* list_count;       // n: resolving blocks of a non-inert row, the split count
    uint32_t* list;             // 2 cap: entries ci << 1 | side in contact order
    // adjoints
    T* xb;                      // n x 16: the state's adjoint (position, orientation, velocity, spin)
    T* resb;                    // n x 4
    T* alb;                     // n x 4: a position iteration's per-row correction adjoint
    T* sbar;                    // n: mass scale adjoint (whole rollout)
    T* mubar;                   // colliders: friction adjoint (whole rollout)
    T* pointb;                  // cap x 4 x 16: per point adjoints of the row quantities
    T* sideb;                   // cap x 2 x 12
    T* accb;                    // cap x 12: impulses' adjoint through the sweeps
    T* mucb;                    // cap: combined friction adjoint this step
};

// ---- phase helpers --------------------------------------------------------------------------

template <class T>
struct SideBody {  // a block side as the solver sees it: frozen (zero masses) when inert
    Q4<T> q;
    V3<T> p, com;
    T im;
    T ib[6];
    bool inert;
};

// x: the row's pose (position xyz _, orientation xyzw)
template <class T>
ADJ_FN SideBody<T> side_body(const float* bodies, const T* s, const T* x, uint32_t row, bool inert) {
    SideBody<T> b;
    b.p = mk3<T>(x[0], x[1], x[2]);
    b.q = ld4(x + 4);
    const float* rec = bodies + static_cast<size_t>(kBodyFloats) * row;
    b.com = b.p + qrot(b.q, ld3<T>(rec + kBodyCom));
    b.inert = inert;
    const T scale = s ? s[row] : T(1.0f);
    b.im = inert ? T(0.0f) : T(rec[kBodyIm]) / scale;
    for (int k = 0; k < 6; ++k) b.ib[k] = inert ? T(0.0f) : T(rec[kBodyInvInertia + k]) / scale;
    return b;
}

// substep_integrate.wgsl:15-76 for one simulated row (no fields: force and torque are zero).
template <class T>
ADJ_FN void integrate_body(const AdjParams& P, const float* rec, T s, const Q4<T>& q, V3<T>& v, V3<T>& w) {
    const T h = T(P.h);
    const float gs = rec[kBodyGravityScale];
    const V3<T> acc = mk3<T>(T(P.gravity[0] * gs), T(P.gravity[1] * gs), T(P.gravity[2] * gs));
    v = v + acc * h;
    V3<T> gw = w;
    const float* i0 = rec + kBodyInertia;
    const bool iso = i0[1] == 0.0f && i0[2] == 0.0f && i0[4] == 0.0f && i0[0] == i0[3] && i0[3] == i0[5];
    if (!iso) {
        T ib[6], in[6];
        for (int k = 0; k < 6; ++k) {
            ib[k] = T(rec[kBodyInvInertia + k]) / s;
            in[k] = T(i0[k]) * s;
        }
        const V3<T> bs = qrot(qconj(q), w);
        const V3<T> mom = sym(in, bs);
        if (!is_zero(cross(bs, mom))) {
            V3<T> solved = mom;
            for (int it = 0; it < 4; ++it) {
                const V3<T> mid = (mom + solved) * T(0.5f);
                const V3<T> rot = -sym(ib, mid) * h;
                // rotate_about (core.wgsl:170-178)
                const T angle = alen(rot);
                if (val(angle) < 1e-8f) {
                    solved = mom;
                } else {
                    const V3<T> axis = rot / angle;
                    const V3<T> off = cross(axis, mom);
                    solved = mom + off * asin_(angle) + cross(axis, off) * (T(1.0f) - acos_(angle));
                }
            }
            gw = qrot(q, sym(ib, solved));
        }
    }
    w = gw;
    const T speed = alen(v);
    if (val(speed) > P.max_velocity) v = v * (T(P.max_velocity) / speed);
    const T spin = alen(w);
    if (val(spin) > P.max_angular_velocity) w = w * (T(P.max_angular_velocity) / spin);
    v = v / (T(1.0f) + T(rec[kBodyLinearDamping]) * h);
    w = w / (T(1.0f) + T(rec[kBodyAngularDamping]) * h);
}

// The per-world contact range of a step: contacts are sorted by (collider a, collider b) and a
// world's colliders are contiguous.
ADJ_FN uint32_t lower_collider(const float* contacts, uint32_t count, uint32_t collider) {
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2u;
        if (contact_u(contacts + static_cast<size_t>(kContactFloats) * mid, 3) < collider) lo = mid + 1u;
        else hi = mid;
    }
    return lo;
}

struct StepIn {
    const float* bodies;    // n x 20 (constant through the rollout)
    const float* contacts;  // count x 76
    const float* fwd;       // n x 16: the step-start states of the recorded forward (the depth model's x_fwd)
    uint32_t c0, c1;        // this world's contacts
};

// The impulses contact c's points enter the step with: the recorded warm start, or, where a point
// was relayed, the final impulses of the caller's previous step (`relayed`, per contact of that
// step) at the point it was relayed from.
template <class T>
ADJ_FN void load_acc0(const float* c, const T* relayed, T* acc) {
    for (uint32_t p = 0; p < 4; ++p) {
        const float* pr = point_rec(c, p);
        const bool has = p < contact_points(c);
        acc[p] = has ? T(pr[12]) : T(0.0f);
        acc[4 + p] = has ? T(pr[13]) : T(0.0f);
        acc[8 + p] = has ? T(pr[14]) : T(0.0f);
        const uint32_t src = fbits(pr[11]);
        if (relayed && has && src != kNoSource) {
            const T* r = relayed + static_cast<size_t>(kAccFloats) * (src >> 2);
            acc[p] = r[src & 3u];
            acc[4 + p] = r[4 + (src & 3u)];
            acc[8 + p] = r[8 + (src & 3u)];
        }
    }
}

// ---- the replay ---------------------------------------------------------------------------------

// One step of world `wr` from the states in W.x (already loaded), with mass scales `s` (null: 1),
// collider frictions `mu` (null: the recorded combined friction) and relayed impulses `relayed`
// (null: the recorded warm start; else per contact, used where the point has a relay source).
// Leaves the end state in W.x and the final impulses in the last acc slot; stores every phase
// boundary the reverse reads.
template <class Ex, class T>
ADJ_FN void replay_step(const Ex& ex, const AdjParams& P, const AdjRange& wr, const StepIn& in, Work<T>& W,
                        const T* s, const T* mu, const T* relayed) {
    const uint32_t r0 = wr.row_base, r1 = wr.row_base + wr.row_count;
    const uint32_t c0 = in.c0, nc = in.c1 - in.c0;
    const uint32_t S = P.substeps, I = P.iterations, Q = P.position_iterations;
    const size_t n = W.n, cap = W.cap;
    const uint32_t nv = 3u + I;
    auto crec = [&](uint32_t ci) { return in.contacts + static_cast<size_t>(kContactFloats) * ci; };
    auto brec = [&](uint32_t row) { return in.bodies + static_cast<size_t>(kBodyFloats) * row; };
    auto inert_row = [&](uint32_t row) { return (body_flags(brec(row)) & kBodyInert) != 0u; };
    // block lists: every non-inert row of the world, its resolving blocks in contact order
    ex.each(wr.row_count, [&](uint32_t i) {
        const uint32_t row = r0 + i;
        uint32_t count = 0;
        for (uint32_t ci = c0; ci < in.c1; ++ci) {
            const float* c = crec(ci);
            if (!(contact_u(c, 2) & kFlagResolves)) continue;
            if (contact_u(c, 0) == row && !(contact_u(c, 2) & kFlagInertF)) ++count;
            if (contact_u(c, 1) == row && !(contact_u(c, 2) & kFlagInertS)) ++count;
        }
        W.list_count[row] = inert_row(row) ? 0u : count;
    });
    ex.sync();
    if (ex.leader()) {
        uint32_t total = 2u * c0;
        for (uint32_t row = r0; row < r1; ++row) {
            W.list_start[row] = total;
            total += W.list_count[row];
        }
    }
    ex.sync();
    ex.each(wr.row_count, [&](uint32_t i) {
        const uint32_t row = r0 + i;
        if (W.list_count[row] == 0u) return;
        uint32_t at = W.list_start[row];
        for (uint32_t ci = c0; ci < in.c1; ++ci) {
            const float* c = crec(ci);
            if (!(contact_u(c, 2) & kFlagResolves)) continue;
            if (contact_u(c, 0) == row && !(contact_u(c, 2) & kFlagInertF)) W.list[at++] = ci << 1;
            if (contact_u(c, 1) == row && !(contact_u(c, 2) & kFlagInertS)) W.list[at++] = (ci << 1) | 1u;
        }
    });
    auto split = [&](uint32_t row) {
        const uint32_t c = W.list_count[row];
        return T(static_cast<float>(c > 1u ? c : 1u));
    };
    auto friction = [&](const float* c) {
        if (!mu) return T(c[7]);
        return combine(mu[contact_u(c, 3)], mu[contact_u(c, 11)], P.friction_combine);
    };
    // rows outside the world's simulated range (its static bodies) never move: their recorded state
    ex.each(nc, [&](uint32_t k) {
        const float* c = crec(c0 + k);
        for (uint32_t side = 0; side < 2; ++side) {
            const uint32_t row = contact_u(c, side);
            if (row >= r0 && row < r1) continue;
            for (uint32_t j = 0; j < kStateFloats; ++j) W.x[kStateFloats * row + j] = T(in.fwd[kStateFloats * row + j]);
        }
    });
    ex.sync();
    // depth model and targets (solver_total.wgsl / solver_blocks.wgsl:13-48) from the start state
    ex.each(nc, [&](uint32_t k) {
        const uint32_t ci = c0 + k;
        const float* c = crec(ci);
        const uint32_t fb = contact_u(c, 0), sb = contact_u(c, 1);
        const V3<T> nrm = ld3<T>(c + 4);
        const T* xf = W.x + kStateFloats * fb;
        const T* xs = W.x + kStateFloats * sb;
        const float* ff = in.fwd + kStateFloats * fb;
        const float* fs = in.fwd + kStateFloats * sb;
        const Q4<T> qf = ld4(xf + 4), qs = ld4(xs + 4);
        const Q4<T> qf0 = {T(ff[4]), T(ff[5]), T(ff[6]), T(ff[7])}, qs0 = {T(fs[4]), T(fs[5]), T(fs[6]), T(fs[7])};
        const V3<T> pf = ld3(xf, 0), ps = ld3(xs, 0), pf0 = ld3<T>(ff), ps0 = ld3<T>(fs);
        const V3<T> comf = pf + qrot(qf, ld3<T>(brec(fb) + kBodyCom)), coms = ps + qrot(qs, ld3<T>(brec(sb) + kBodyCom));
        const V3<T> vf = ld3(xf + 8, 0), wf = ld3(xf + 12, 0), vs = ld3(xs + 8, 0), ws = ld3(xs + 12, 0);
        for (uint32_t p = 0; p < 4; ++p) {
            T d = T(0.0f), target = T(0.0f);
            if (p < contact_points(c)) {
                const float* pr = point_rec(c, p);
                const V3<T> la = ld3<T>(pr), lb = ld3<T>(pr + 4), pos = ld3<T>(pr + 8);
                const V3<T> moved = ((qrot(qf, la) + pf) - (qrot(qf0, la) + pf0)) - ((qrot(qs, lb) + ps) - (qrot(qs0, lb) + ps0));
                d = T(pr[3]) + dot(nrm, moved);
                const V3<T> rel = (vs + cross(ws, pos - coms)) - (vf + cross(wf, pos - comf));
                const T speed = dot(rel, nrm);
                if (val(d) < 0.0f) target = d / T(P.dt);
                else if (val(speed) < -P.restitution_threshold) target = -speed * T(c[10]);
            }
            W.depth[4 * ci + p] = d;
            W.target[4 * ci + p] = target;
        }
        load_acc0(c, relayed, W.acc + kAccFloats * static_cast<size_t>(ci));  // sweep 0 of substep 0
    });
    ex.each(wr.row_count, [&](uint32_t i) {
        T* r = W.res + 4u * (r0 + i);  // iteration 0 of substep 0 reads zero resolution
        r[0] = r[1] = r[2] = r[3] = T(0.0f);
    });
    ex.sync();
    for (uint32_t sub = 0; sub < S; ++sub) {
        T* vel = W.vel + static_cast<size_t>(sub) * nv * n * 8u;
        T* pose = W.pose + static_cast<size_t>(sub) * 2u * n * 8u;
        T* acc = W.acc + static_cast<size_t>(sub) * (I + 1u) * cap * kAccFloats;
        T* rows = W.rows + static_cast<size_t>(sub) * cap * kRowFloats;
        // substep start, integrate
        ex.each(wr.row_count, [&](uint32_t i) {
            const uint32_t row = r0 + i;
            T* x = W.x + kStateFloats * row;
            T* ps = pose + 8u * row;
            for (int k = 0; k < 8; ++k) ps[k] = x[k];
            T* v0 = vel + 8u * row;
            for (int k = 0; k < 3; ++k) {
                v0[k] = x[8 + k];
                v0[4 + k] = x[12 + k];
            }
            const float* rec = brec(row);
            const uint32_t f = body_flags(rec);
            if ((f & kBodySimulates) && !(f & kBodyInert)) {
                V3<T> v = ld3(x + 8, 0), w = ld3(x + 12, 0);
                integrate_body<T>(P, rec, s ? s[row] : T(1.0f), ld4(x + 4), v, w);
                st3(x + 8, v);
                st3(x + 12, w);
            }
            T* v1 = vel + static_cast<size_t>(n) * 8u + 8u * row;
            for (int k = 0; k < 3; ++k) {
                v1[k] = x[8 + k];
                v1[4 + k] = x[12 + k];
            }
        });
        ex.sync();
        // rows of this substep (solver_contact_block.wgsl:31-110), and the carried impulses
        ex.each(nc, [&](uint32_t k) {
            const uint32_t ci = c0 + k;
            const float* c = crec(ci);
            if (sub > 0) {
                const T* from = W.acc + (static_cast<size_t>(sub - 1u) * (I + 1u) + I) * cap * kAccFloats + kAccFloats * ci;
                T* to = acc + kAccFloats * ci;
                for (uint32_t j = 0; j < kAccFloats; ++j) to[j] = from[j];
            }
            if (!contact_active(c)) return;
            const uint32_t fb = contact_u(c, 0), sb = contact_u(c, 1), fl = contact_u(c, 2);
            const SideBody<T> bf = side_body<T>(in.bodies, s, W.x + kStateFloats * fb, fb, (fl & kFlagInertF) != 0u);
            const SideBody<T> bs = side_body<T>(in.bodies, s, W.x + kStateFloats * sb, sb, (fl & kFlagInertS) != 0u);
            const T c1 = split(fb), c2 = split(sb);
            T ibf[6], ibs[6];
            for (int j = 0; j < 6; ++j) {
                ibf[j] = bf.ib[j] * c1;
                ibs[j] = bs.ib[j] * c2;
            }
            const V3<T> nrm = ld3<T>(c + 4);
            V3<T> t1, t2;
            make_tangents(nrm, &t1, &t2);
            const uint32_t np = contact_points(c);
            V3<T> ra[4], rb[4];
            V3<T> reach_a = zero3<T>(), reach_b = zero3<T>();
            for (uint32_t p = 0; p < np; ++p) {
                const float* pr = point_rec(c, p);
                ra[p] = (qrot(bf.q, ld3<T>(pr)) + bf.p) - bf.com;
                rb[p] = (qrot(bs.q, ld3<T>(pr + 4)) + bs.p) - bs.com;
                reach_a += ra[p];
                reach_b += rb[p];
            }
            const T count = T(static_cast<float>(np > 1u ? np : 1u));
            const T imf_split = bf.im * c1, ims_split = bs.im * c2;
            T* row = rows + kRowFloats * ci;
            for (uint32_t p = 0; p < np; ++p) {
                const V3<T> arm_a = cross(ra[p], nrm), arm_b = cross(rb[p], nrm);
                const T nmass = count * (imf_split + ims_split) + dot(inertia_apply(bf.q, ibf, cross(reach_a, nrm)), arm_a) +
                                dot(inertia_apply(bs.q, ibs, cross(reach_b, nrm)), arm_b);
                const T truth = bf.im + bs.im + dot(arm_a, inertia_apply(bf.q, bf.ib, arm_a)) +
                                dot(arm_b, inertia_apply(bs.q, bs.ib, arm_b));
                const V3<T> a1 = cross(ra[p], t1), b1 = cross(rb[p], t1), a2 = cross(ra[p], t2), b2 = cross(rb[p], t2);
                const T m1 = (imf_split + ims_split + dot(a1, inertia_apply(bf.q, ibf, a1)) + dot(b1, inertia_apply(bs.q, ibs, b1))) * count;
                const T m2 = (imf_split + ims_split + dot(a2, inertia_apply(bf.q, ibf, a2)) + dot(b2, inertia_apply(bs.q, ibs, b2))) * count;
                T* q = row + 12u * p;
                st3(q, ra[p]);
                q[3] = nmass;
                st3(q + 4, rb[p]);
                q[7] = truth;
                q[8] = m1;
                q[9] = m2;
            }
        });
        ex.sync();
        // sweeps: the warm start (first substep), then the iterations; block words, then row sums
        const uint32_t sweeps = I + (sub == 0 ? 1u : 0u);
        for (uint32_t sw = 0; sw < sweeps; ++sw) {
            const bool warm = sub == 0 && sw == 0;
            const uint32_t it = sub == 0 ? sw - 1u : sw;  // iteration index when not warm
            const T* a_in = acc + (warm ? 0u : it) * cap * kAccFloats;
            T* a_out = acc + (warm ? 0u : it + 1u) * cap * kAccFloats;
            ex.each(nc, [&](uint32_t k) {
                const uint32_t ci = c0 + k;
                const float* c = crec(ci);
                T* slot = W.slot + 2u * kSlotFloats * ci;
                for (uint32_t j = 0; j < 2u * kSlotFloats; ++j) slot[j] = T(0.0f);
                if (!contact_active(c)) {
                    if (!warm) {
                        for (uint32_t j = 0; j < kAccFloats; ++j) a_out[kAccFloats * ci + j] = a_in[kAccFloats * ci + j];
                    }
                    return;
                }
                const uint32_t fb = contact_u(c, 0), sb = contact_u(c, 1), fl = contact_u(c, 2);
                const SideBody<T> bf = side_body<T>(in.bodies, s, W.x + kStateFloats * fb, fb, (fl & kFlagInertF) != 0u);
                const SideBody<T> bs = side_body<T>(in.bodies, s, W.x + kStateFloats * sb, sb, (fl & kFlagInertS) != 0u);
                const V3<T> nrm = ld3<T>(c + 4);
                V3<T> t1, t2;
                make_tangents(nrm, &t1, &t2);
                const T* row = rows + kRowFloats * ci;
                const T* ain = a_in + kAccFloats * ci;
                T* aout = a_out + kAccFloats * ci;
                const uint32_t np = contact_points(c);
                const T* xf = W.x + kStateFloats * fb;
                const T* xs = W.x + kStateFloats * sb;
                const V3<T> vf = ld3(xf + 8, 0), wf = ld3(xf + 12, 0), vs = ld3(xs + 8, 0), ws = ld3(xs + 12, 0);
                const T muc = friction(c);
                V3<T> dvf = zero3<T>(), xaf = zero3<T>(), dvs = zero3<T>(), xas = zero3<T>();
                for (uint32_t p = 0; p < np; ++p) {
                    const T* q = row + 12u * p;
                    const V3<T> ra = ld3(q, 0), rb = ld3(q + 4, 0);
                    T an = ain[p], at1 = ain[4 + p], at2 = ain[8 + p];
                    V3<T> imp;
                    if (warm) {
                        imp = nrm * an + t1 * at1 + t2 * at2;
                    } else {
                        const T nmass = q[3], m1 = q[8], m2 = q[9];
                        const V3<T> rel = (vs + cross(ws, rb)) - (vf + cross(wf, ra));
                        const T nspeed = dot(rel, nrm);
                        const T delta = (W.target[4 * ci + p] - nspeed) / nmass;
                        const T sum = an + delta;
                        const T next_n = val(T(0.0f)) > val(sum) ? T(0.0f) : sum;
                        imp = nrm * (next_n - an);
                        an = next_n;
                        const T limit = muc * an;
                        const T s1 = dot(rel, t1);
                        T x1 = at1 - s1 / m1;
                        if (!(val(x1) > val(-limit))) x1 = -limit;
                        else if (!(val(x1) < val(limit))) x1 = limit;
                        imp += t1 * (x1 - at1);
                        at1 = x1;
                        const T s2 = dot(rel, t2);
                        const T inner = limit * limit - at1 * at1;
                        const T rem = asqrt(val(inner) > 0.0f ? inner : T(0.0f));
                        T x2 = at2 - s2 / m2;
                        if (!(val(x2) > val(-rem))) x2 = -rem;
                        else if (!(val(x2) < val(rem))) x2 = rem;
                        imp += t2 * (x2 - at2);
                        at2 = x2;
                        aout[p] = an;
                        aout[4 + p] = at1;
                        aout[8 + p] = at2;
                    }
                    dvf -= imp * bf.im;
                    xaf += cross(ra, imp);
                    dvs += imp * bs.im;
                    xas += cross(rb, imp);
                }
                if (!warm) {
                    for (uint32_t p = np; p < 4; ++p) {
                        aout[p] = ain[p];
                        aout[4 + p] = ain[4 + p];
                        aout[8 + p] = ain[8 + p];
                    }
                }
                const V3<T> dwf = -inertia_apply(bf.q, bf.ib, xaf), dws = inertia_apply(bs.q, bs.ib, xas);
                st3(slot + 7, dvf);
                st3(slot + 10, dwf);
                st3(slot + kSlotFloats + 7, dvs);
                st3(slot + kSlotFloats + 10, dws);
            });
            ex.sync();
            ex.each(wr.row_count, [&](uint32_t i) {
                const uint32_t row = r0 + i;
                const uint32_t count = W.list_count[row];
                T* x = W.x + kStateFloats * row;
                V3<T> lin = zero3<T>(), ang = zero3<T>();
                for (uint32_t e = 0; e < count; ++e) {
                    const uint32_t entry = W.list[W.list_start[row] + e];
                    const T* sl = W.slot + 2u * kSlotFloats * (entry >> 1) + kSlotFloats * (entry & 1u);
                    lin += ld3(sl + 7, 0);
                    ang += ld3(sl + 10, 0);
                }
                st3(x + 8, ld3(x + 8, 0) + lin);
                st3(x + 12, ld3(x + 12, 0) + ang);
                T* vo = vel + (static_cast<size_t>(warm ? 2u : 3u + it)) * n * 8u + 8u * row;
                for (int k = 0; k < 3; ++k) {
                    vo[k] = x[8 + k];
                    vo[4 + k] = x[12 + k];
                }
            });
            ex.sync();
        }
        if (sub > 0) {  // no warm start: the velocities after it are those after integration
            ex.each(wr.row_count, [&](uint32_t i) {
                const uint32_t row = r0 + i;
                const T* a = vel + static_cast<size_t>(n) * 8u + 8u * row;
                T* b = vel + static_cast<size_t>(2u) * n * 8u + 8u * row;
                for (int k = 0; k < 8; ++k) b[k] = a[k];
            });
        }
        // advance (substep_advance.wgsl)
        ex.each(wr.row_count, [&](uint32_t i) {
            const uint32_t row = r0 + i;
            const uint32_t f = body_flags(brec(row));
            T* x = W.x + kStateFloats * row;
            if ((f & kBodyMoves) && !(f & kBodyInert)) {
                const T h = T(P.h);
                st3(x, ld3(x, 0) + ld3(x + 8, 0) * h);
                const Q4<T> q = ld4(x + 4);
                const Q4<T> spin = {x[12], x[13], x[14], T(0.0f)};
                const Q4<T> dq = qmul(spin, q);
                const T k = T(0.5f) * h;
                st4(x + 4, qnorm(Q4<T>{q.x + k * dq.x, q.y + k * dq.y, q.z + k * dq.z, q.w + k * dq.w}));
            }
            T* pa = pose + static_cast<size_t>(n) * 8u + 8u * row;
            for (int k = 0; k < 8; ++k) pa[k] = x[k];
        });
        ex.sync();
        // position iterations (position_block.wgsl, position_correction.wgsl:5-46, position_apply.wgsl)
        for (uint32_t j = 0; j < Q; ++j) {
            T* rs = W.res + (static_cast<size_t>(sub) * Q + j) * n * 4u;
            ex.each(nc, [&](uint32_t k) {
                const uint32_t ci = c0 + k;
                const float* c = crec(ci);
                T* slot = W.slot + 2u * kSlotFloats * ci;
                for (uint32_t q = 0; q < 2u * kSlotFloats; ++q) slot[q] = T(0.0f);
                position_block(P, c, ci, in.bodies, s, W, rs, slot);
            });
            ex.sync();
            T* rn = (j + 1u < Q) ? rs + n * 4u : (sub + 1u < S ? W.res + (static_cast<size_t>(sub + 1u) * Q) * n * 4u : nullptr);
            ex.each(wr.row_count, [&](uint32_t i) {
                const uint32_t row = r0 + i;
                const uint32_t count = W.list_count[row];
                T* x = W.x + kStateFloats * row;
                V3<T> sum = zero3<T>();
                uint32_t contributing = 0;
                for (uint32_t e = 0; e < count; ++e) {
                    const uint32_t entry = W.list[W.list_start[row] + e];
                    const T* sl = W.slot + 2u * kSlotFloats * (entry >> 1) + kSlotFloats * (entry & 1u);
                    if (val(sl[15]) != 0.0f) {
                        ++contributing;
                        sum += ld3(sl, 0);
                    }
                }
                rs[4u * row + 3] = T(static_cast<float>(contributing));
                V3<T> r = ld3(rs + 4u * row, 0);
                if (contributing > 0u) {
                    const V3<T> al = sum / T(static_cast<float>(contributing));
                    st3(x, ld3(x, 0) + al);
                    st4(x + 4, qnorm(ld4(x + 4)));
                    r = r + al;
                }
                if (rn) {
                    st3(rn + 4u * row, r);
                    rn[4u * row + 3] = T(0.0f);
                }
            });
            ex.sync();
        }
    }
}

// One position block's forward at resolution `rs`: writes the two sides' linear corrections into
// the slot (position words) and a contribution mark (word 15 of each side).
template <class T>
ADJ_FN void position_block(const AdjParams& P, const float* c, uint32_t ci, const float* bodies, const T* s,
                           const Work<T>& W, const T* rs, T* slot) {
    if (!(contact_u(c, 2) & kFlagResolves) || !contact_points(c)) return;
    const uint32_t fb = contact_u(c, 0), sb = contact_u(c, 1), fl = contact_u(c, 2);
    const float* recf = bodies + static_cast<size_t>(kBodyFloats) * fb;
    const float* recs = bodies + static_cast<size_t>(kBodyFloats) * sb;
    const T imf = (fl & kFlagInertF) ? T(0.0f) : T(recf[kBodyIm]) / (s ? s[fb] : T(1.0f));
    const T ims = (fl & kFlagInertS) ? T(0.0f) : T(recs[kBodyIm]) / (s ? s[sb] : T(1.0f));
    const V3<T> nrm = ld3<T>(c + 4);
    const V3<T> rf = (fl & kFlagInertF) ? zero3<T>() : ld3(rs + 4u * fb, 0);
    const V3<T> rsv = (fl & kFlagInertS) ? zero3<T>() : ld3(rs + 4u * sb, 0);
    const T resolved = dot(nrm, rsv - rf);
    V3<T> total = zero3<T>();
    uint32_t contributing = 0;
    const T k = imf + ims;
    for (uint32_t p = 0; p < contact_points(c); ++p) {
        const T e = W.depth[4 * ci + p] - T(P.slop) - resolved;
        if (!(val(e) > 0.0f)) continue;
        if (val(k) == 0.0f) continue;
        total += nrm * (T(P.relaxation) * e / k);
        ++contributing;
    }
    if (contributing == 0u) return;
    const V3<T> corr = total / T(static_cast<float>(contributing));
    const V3<T> la = -corr * imf, lb = corr * ims;
    if (!(val(dot(la, la)) > 0.0f || val(dot(lb, lb)) > 0.0f)) return;
    st3(slot, la);
    slot[15] = T(1.0f);
    st3(slot + kSlotFloats, lb);
    slot[kSlotFloats + 15] = T(1.0f);
}

// ---- the reverse ------------------------------------------------------------------------------

// The reverse of replay_step, run right after it (its stores are read back). On entry W.xb holds
// the adjoint of the step's end state and accb_end the adjoint of its final impulses (per contact,
// kAccFloats). On exit W.xb holds the adjoint of the step's start state, W.sbar and W.mubar have
// gained this step's parameter adjoints and accb_start (per contact) the adjoint of the impulses
// the step started from.
template <class Ex, class T>
ADJ_FN void reverse_step(const Ex& ex, const AdjParams& P, const AdjRange& wr, const StepIn& in, Work<T>& W,
                         const T* accb_end, T* accb_start) {
    const uint32_t r0 = wr.row_base;
    const uint32_t c0 = in.c0, nc = in.c1 - in.c0;
    const uint32_t S = P.substeps, I = P.iterations, Q = P.position_iterations;
    const size_t n = W.n, cap = W.cap;
    const uint32_t nv = 3u + I;
    auto crec = [&](uint32_t ci) { return in.contacts + static_cast<size_t>(kContactFloats) * ci; };
    auto brec = [&](uint32_t row) { return in.bodies + static_cast<size_t>(kBodyFloats) * row; };
    auto split = [&](uint32_t row) {
        const uint32_t c = W.list_count[row];
        return T(static_cast<float>(c > 1u ? c : 1u));
    };
    const uint32_t r1 = wr.row_base + wr.row_count;
    // A row's pose or velocities at a phase boundary: the world's simulated rows from the phase's
    // store, the others (static) from the live state, which they never leave.
    auto pose_of = [&](const T* arr, uint32_t row) -> const T* {
        return (row >= r0 && row < r1) ? arr + 8u * row : W.x + kStateFloats * row;
    };
    auto vel_of = [&](const T* arr, uint32_t row) -> const T* {
        return (row >= r0 && row < r1) ? arr + 8u * row : W.x + kStateFloats * row + 8u;
    };
    // gathers a phase's slot words into the rows: resolution (slot 0..2) or velocities (7..12)
    auto gather = [&](bool velocities) {
        ex.each(wr.row_count, [&](uint32_t i) {
            const uint32_t row = r0 + i;
            const uint32_t lc = W.list_count[row];
            T* xb = W.xb + kStateFloats * row;
            for (uint32_t e = 0; e < lc; ++e) {
                const uint32_t entry = W.list[W.list_start[row] + e];
                const T* sl = W.slot + 2u * kSlotFloats * (entry >> 1) + kSlotFloats * (entry & 1u);
                for (uint32_t j = 0; j < 3; ++j) {
                    if (velocities) {
                        xb[8 + j] += sl[7 + j];
                        xb[12 + j] += sl[10 + j];
                    } else {
                        W.resb[4u * row + j] += sl[j];
                    }
                }
            }
        });
    };
    auto clear_slots = [&](uint32_t ci) {
        T* slot = W.slot + 2u * kSlotFloats * ci;
        for (uint32_t j = 0; j < 2u * kSlotFloats; ++j) slot[j] = T(0.0f);
    };
    ex.each(nc, [&](uint32_t k) {
        const uint32_t ci = c0 + k;
        T* a = W.accb + kAccFloats * ci;
        for (uint32_t j = 0; j < kAccFloats; ++j) a[j] = accb_end ? accb_end[kAccFloats * ci + j] : T(0.0f);
        T* pb = W.pointb + 4u * kPointAdj * ci;
        for (uint32_t j = 0; j < 4u * kPointAdj; ++j) pb[j] = T(0.0f);
        T* sd = W.sideb + 2u * kSideFloats * ci;
        for (uint32_t j = 0; j < 2u * kSideFloats; ++j) sd[j] = T(0.0f);
        W.mucb[ci] = T(0.0f);
    });
    ex.each(wr.row_count, [&](uint32_t i) {
        T* r = W.resb + 4u * (r0 + i);
        r[0] = r[1] = r[2] = r[3] = T(0.0f);
    });
    ex.sync();
    for (uint32_t sub = S; sub-- > 0;) {
        const T* vel = W.vel + static_cast<size_t>(sub) * nv * n * 8u;
        const T* pose = W.pose + static_cast<size_t>(sub) * 2u * n * 8u;
        const T* acc = W.acc + static_cast<size_t>(sub) * (I + 1u) * cap * kAccFloats;
        const T* rows = W.rows + static_cast<size_t>(sub) * cap * kRowFloats;
        // -- position iterations, last first
        for (uint32_t j = Q; j-- > 0;) {
            const T* rs = W.res + (static_cast<size_t>(sub) * Q + j) * n * 4u;
            ex.each(wr.row_count, [&](uint32_t i) {
                const uint32_t row = r0 + i;
                T* xb = W.xb + kStateFloats * row;
                T* alb = W.alb + 4u * row;
                for (int k = 0; k < 3; ++k) alb[k] = xb[k] + W.resb[4u * row + k];
                if (val(rs[4u * row + 3]) > 0.0f) {
                    const Q4<T> qa = ld4(pose + static_cast<size_t>(n) * 8u + 8u * row + 4u);
                    st4(xb + 4, qnorm_vjp(qa, ld4(xb + 4)));
                }
            });
            ex.sync();
            ex.each(nc, [&](uint32_t k) {
                const uint32_t ci = c0 + k;
                clear_slots(ci);
                const float* c = crec(ci);
                T fwd[2 * kSlotFloats];
                for (uint32_t q = 0; q < 2u * kSlotFloats; ++q) fwd[q] = T(0.0f);
                position_block(P, c, ci, in.bodies, static_cast<const T*>(nullptr), W, rs, fwd);
                if (val(fwd[15]) == 0.0f) return;
                const uint32_t fb = contact_u(c, 0), sb = contact_u(c, 1), fl = contact_u(c, 2);
                const bool fi = (fl & kFlagInertF) != 0u, si = (fl & kFlagInertS) != 0u;
                const T imf = fi ? T(0.0f) : T(brec(fb)[kBodyIm]);
                const T ims = si ? T(0.0f) : T(brec(sb)[kBodyIm]);
                const V3<T> nrm = ld3<T>(c + 4);
                const V3<T> rf = fi ? zero3<T>() : ld3(rs + 4u * fb, 0);
                const V3<T> rsv = si ? zero3<T>() : ld3(rs + 4u * sb, 0);
                const T resolved = dot(nrm, rsv - rf);
                const T kk = imf + ims;
                uint32_t contributing = 0;
                V3<T> total = zero3<T>();
                for (uint32_t p = 0; p < contact_points(c); ++p) {
                    const T e = W.depth[4 * ci + p] - T(P.slop) - resolved;
                    if (val(e) > 0.0f) {
                        ++contributing;
                        total += nrm * (T(P.relaxation) * e / kk);
                    }
                }
                const T cnt = T(static_cast<float>(contributing));
                const V3<T> corr = total / cnt;
                V3<T> corrb = zero3<T>();
                T imfb = T(0.0f), imsb = T(0.0f);
                if (!fi) {
                    const V3<T> lb = ld3(W.alb + 4u * fb, 0) / rs[4u * fb + 3];
                    corrb -= lb * imf;
                    imfb -= dot(corr, lb);
                }
                if (!si) {
                    const V3<T> lb = ld3(W.alb + 4u * sb, 0) / rs[4u * sb + 3];
                    corrb += lb * ims;
                    imsb += dot(corr, lb);
                }
                const T g = dot(nrm, corrb / cnt);  // total's adjoint along the normal
                T resolvedb = T(0.0f), kb = T(0.0f);
                T* pb = W.pointb + 4u * kPointAdj * ci;
                for (uint32_t p = 0; p < contact_points(c); ++p) {
                    const T e = W.depth[4 * ci + p] - T(P.slop) - resolved;
                    if (!(val(e) > 0.0f)) continue;
                    const T eb = T(P.relaxation) * g / kk;
                    kb -= T(P.relaxation) * e * g / (kk * kk);
                    pb[kPointAdj * p + 11] += eb;  // depth
                    resolvedb -= eb;
                }
                if (!fi) imfb += kb;
                if (!si) imsb += kb;
                T* sd = W.sideb + 2u * kSideFloats * ci;
                sd[10] += imfb;
                sd[kSideFloats + 10] += imsb;
                T* slot = W.slot + 2u * kSlotFloats * ci;
                st3(slot, -nrm * resolvedb);
                st3(slot + kSlotFloats, nrm * resolvedb);
            });
            ex.sync();
            gather(false);
            ex.sync();
        }
        // -- advance
        ex.each(wr.row_count, [&](uint32_t i) {
            const uint32_t row = r0 + i;
            const uint32_t f = body_flags(brec(row));
            if (!(f & kBodyMoves) || (f & kBodyInert)) return;
            T* xb = W.xb + kStateFloats * row;
            const T h = T(P.h), k = T(0.5f) * h;
            const Q4<T> q = ld4(pose + 8u * row + 4u);
            const T* vlast = vel + static_cast<size_t>(2u + I) * n * 8u + 8u * row;
            const V3<T> w = ld3(vlast + 4, 0);
            const Q4<T> dq = qmul(Q4<T>{w.x, w.y, w.z, T(0.0f)}, q);
            const Q4<T> u = {q.x + k * dq.x, q.y + k * dq.y, q.z + k * dq.z, q.w + k * dq.w};
            const Q4<T> ub = qnorm_vjp(u, ld4(xb + 4));
            const V3<T> uv = mk3<T>(ub.x * k, ub.y * k, ub.z * k);
            const T uw = ub.w * k;
            const V3<T> qvv = qv(q);
            const V3<T> qvb = cross(uv, w) - w * uw;
            const T qwb = dot(uv, w);
            const V3<T> wb = uv * q.w + cross(qvv, uv) - qvv * uw;
            xb[4] = ub.x + qvb.x;
            xb[5] = ub.y + qvb.y;
            xb[6] = ub.z + qvb.z;
            xb[7] = ub.w + qwb;
            st3(xb + 8, ld3(xb + 8, 0) + ld3(xb, 0) * h);
            st3(xb + 12, ld3(xb + 12, 0) + wb);
        });
        ex.sync();
        // -- the sweeps, last first
        const uint32_t sweeps = I + (sub == 0 ? 1u : 0u);
        for (uint32_t sw = sweeps; sw-- > 0;) {
            const bool warm = sub == 0 && sw == 0;
            const uint32_t it = sub == 0 ? sw - 1u : sw;
            const T* a_in = acc + (warm ? 0u : it) * cap * kAccFloats;
            const T* v_in = vel + static_cast<size_t>(warm ? 1u : 2u + it) * n * 8u;
            ex.each(nc, [&](uint32_t k) {
                const uint32_t ci = c0 + k;
                clear_slots(ci);
                const float* c = crec(ci);
                if (!contact_active(c)) return;
                const uint32_t fb = contact_u(c, 0), sb = contact_u(c, 1), fl = contact_u(c, 2);
                const bool fi = (fl & kFlagInertF) != 0u, si = (fl & kFlagInertS) != 0u;
                const SideBody<T> bf = side_body<T>(in.bodies, nullptr, pose_of(pose, fb), fb, fi);
                const SideBody<T> bs = side_body<T>(in.bodies, nullptr, pose_of(pose, sb), sb, si);
                const V3<T> nrm = ld3<T>(c + 4);
                V3<T> t1, t2;
                make_tangents(nrm, &t1, &t2);
                const T* row = rows + kRowFloats * ci;
                const T* ain = a_in + kAccFloats * ci;
                T* ab = W.accb + kAccFloats * ci;  // adjoint of the sweep's output impulses, becomes its input's
                T* pb = W.pointb + 4u * kPointAdj * ci;
                T* sd = W.sideb + 2u * kSideFloats * ci;
                const uint32_t np = contact_points(c);
                const V3<T> vf = ld3(vel_of(v_in, fb), 0), wf = ld3(vel_of(v_in, fb) + 4u, 0);
                const V3<T> vs = ld3(vel_of(v_in, sb), 0), ws = ld3(vel_of(v_in, sb) + 4u, 0);
                const T muc = T(c[7]);
                const T* xbf = W.xb + kStateFloats * fb;
                const T* xbs = W.xb + kStateFloats * sb;
                const V3<T> dvfb = fi ? zero3<T>() : ld3(xbf + 8, 0), dwfb = fi ? zero3<T>() : ld3(xbf + 12, 0);
                const V3<T> dvsb = si ? zero3<T>() : ld3(xbs + 8, 0), dwsb = si ? zero3<T>() : ld3(xbs + 12, 0);
                // forward once more: each point's impulse and branch values
                V3<T> imp[4], rel[4];
                T an0[4], next_n[4], sum_n[4], x1r[4], x1[4], lim[4], x2r[4], x2[4], rem[4], inner[4], s1[4], s2[4];
                V3<T> xaf = zero3<T>(), xas = zero3<T>(), impsum = zero3<T>();
                for (uint32_t p = 0; p < np; ++p) {
                    const T* q = row + 12u * p;
                    const V3<T> ra = ld3(q, 0), rb = ld3(q + 4, 0);
                    const T an = ain[p], at1 = ain[4 + p], at2 = ain[8 + p];
                    an0[p] = an;
                    if (warm) {
                        imp[p] = nrm * an + t1 * at1 + t2 * at2;
                    } else {
                        const T nmass = q[3], m1 = q[8], m2 = q[9];
                        rel[p] = (vs + cross(ws, rb)) - (vf + cross(wf, ra));
                        const T nspeed = dot(rel[p], nrm);
                        const T delta = (W.target[4 * ci + p] - nspeed) / nmass;
                        sum_n[p] = an + delta;
                        next_n[p] = val(T(0.0f)) > val(sum_n[p]) ? T(0.0f) : sum_n[p];
                        lim[p] = muc * next_n[p];
                        s1[p] = dot(rel[p], t1);
                        x1r[p] = at1 - s1[p] / m1;
                        x1[p] = x1r[p];
                        if (!(val(x1[p]) > val(-lim[p]))) x1[p] = -lim[p];
                        else if (!(val(x1[p]) < val(lim[p]))) x1[p] = lim[p];
                        s2[p] = dot(rel[p], t2);
                        inner[p] = lim[p] * lim[p] - x1[p] * x1[p];
                        rem[p] = asqrt(val(inner[p]) > 0.0f ? inner[p] : T(0.0f));
                        x2r[p] = at2 - s2[p] / m2;
                        x2[p] = x2r[p];
                        if (!(val(x2[p]) > val(-rem[p]))) x2[p] = -rem[p];
                        else if (!(val(x2[p]) < val(rem[p]))) x2[p] = rem[p];
                        imp[p] = nrm * (next_n[p] - an) + t1 * (x1[p] - at1) + t2 * (x2[p] - at2);
                    }
                    xaf += cross(ra, imp[p]);
                    xas += cross(rb, imp[p]);
                    impsum += imp[p];
                }
                // outputs: dv_f = -sum imp im_f, dw_f = -I_f(xaf), dv_s = sum imp im_s, dw_s = I_s(xas)
                V3<T> xafb = zero3<T>(), xasb = zero3<T>();
                Q4<T> qfb = {T(0.0f), T(0.0f), T(0.0f), T(0.0f)}, qsb = qfb;
                T ibfb[6] = {T(0.0f), T(0.0f), T(0.0f), T(0.0f), T(0.0f), T(0.0f)};
                T ibsb[6] = {T(0.0f), T(0.0f), T(0.0f), T(0.0f), T(0.0f), T(0.0f)};
                if (!fi) inertia_vjp(bf.q, bf.ib, xaf, -dwfb, &qfb, ibfb, &xafb);
                if (!si) inertia_vjp(bs.q, bs.ib, xas, dwsb, &qsb, ibsb, &xasb);
                const T imfb = -dot(impsum, dvfb), imsb = dot(impsum, dvsb);
                V3<T> vfb = zero3<T>(), wfb = zero3<T>(), vsb = zero3<T>(), wsb = zero3<T>();
                T mucb = T(0.0f);
                for (uint32_t p = 0; p < np; ++p) {
                    const T* q = row + 12u * p;
                    const V3<T> ra = ld3(q, 0), rb = ld3(q + 4, 0);
                    T* pp = pb + kPointAdj * p;
                    V3<T> impb = dvsb * bs.im - dvfb * bf.im + cross(xafb, ra) + cross(xasb, rb);
                    V3<T> rab = cross(imp[p], xafb), rbb = cross(imp[p], xasb);
                    const T en = dot(nrm, impb), e1 = dot(t1, impb), e2 = dot(t2, impb);
                    if (warm) {
                        ab[p] += en;
                        ab[4 + p] += e1;
                        ab[8 + p] += e2;
                    } else {
                        const T nmass = q[3], m1 = q[8], m2 = q[9];
                        T nnb = en + ab[p], n1b = e1 + ab[4 + p], n2b = e2 + ab[8 + p];
                        T anb = -en, at1b = -e1, at2b = -e2;
                        // tangent 2: clamp to the remaining cone
                        T remb = T(0.0f), x2b = T(0.0f);
                        if (!(val(x2r[p]) > val(-rem[p]))) remb -= n2b;
                        else if (!(val(x2r[p]) < val(rem[p]))) remb += n2b;
                        else x2b = n2b;
                        at2b += x2b;
                        const T s2b = -x2b / m2;
                        const T m2b = x2b * s2[p] / (m2 * m2);
                        T limb = T(0.0f);
                        if (val(inner[p]) > 0.0f) {
                            const T ib = remb / (T(2.0f) * rem[p]);
                            limb += T(2.0f) * lim[p] * ib;
                            n1b -= T(2.0f) * x1[p] * ib;
                        }
                        // tangent 1: clamp to the cone
                        T x1b = T(0.0f);
                        if (!(val(x1r[p]) > val(-lim[p]))) limb -= n1b;
                        else if (!(val(x1r[p]) < val(lim[p]))) limb += n1b;
                        else x1b = n1b;
                        at1b += x1b;
                        const T s1b = -x1b / m1;
                        const T m1b = x1b * s1[p] / (m1 * m1);
                        // limit = friction * next normal
                        mucb += limb * next_n[p];
                        nnb += limb * muc;
                        // normal clamp
                        T deltab = T(0.0f);
                        if (!(val(T(0.0f)) > val(sum_n[p]))) {
                            anb += nnb;
                            deltab = nnb;
                        }
                        const T nspeed = dot(rel[p], nrm);
                        const T delta = (W.target[4 * ci + p] - nspeed) / nmass;
                        pp[10] += deltab / nmass;  // target
                        const T nspeedb = -deltab / nmass;
                        pp[3] -= deltab * delta / nmass;  // nmass
                        pp[8] += m1b;
                        pp[9] += m2b;
                        const V3<T> relb = nrm * nspeedb + t1 * s1b + t2 * s2b;
                        vsb += relb;
                        wsb += cross(rb, relb);
                        rbb += cross(relb, ws);
                        vfb -= relb;
                        wfb -= cross(ra, relb);
                        rab -= cross(relb, wf);
                        ab[p] = anb;
                        ab[4 + p] = at1b;
                        ab[8 + p] = at2b;
                    }
                    pp[0] += rab.x;
                    pp[1] += rab.y;
                    pp[2] += rab.z;
                    pp[4] += rbb.x;
                    pp[5] += rbb.y;
                    pp[6] += rbb.z;
                }
                W.mucb[ci] += mucb;
                for (int j = 0; j < 4; ++j) {
                    sd[j] += j == 0 ? qfb.x : j == 1 ? qfb.y : j == 2 ? qfb.z : qfb.w;
                    sd[kSideFloats + j] += j == 0 ? qsb.x : j == 1 ? qsb.y : j == 2 ? qsb.z : qsb.w;
                }
                for (int j = 0; j < 6; ++j) {
                    sd[4 + j] += ibfb[j];
                    sd[kSideFloats + 4 + j] += ibsb[j];
                }
                sd[10] += imfb;
                sd[kSideFloats + 10] += imsb;
                if (!warm) {
                    T* slot = W.slot + 2u * kSlotFloats * ci;
                    st3(slot + 7, vfb);
                    st3(slot + 10, wfb);
                    st3(slot + kSlotFloats + 7, vsb);
                    st3(slot + kSlotFloats + 10, wsb);
                }
            });
            ex.sync();
            if (!warm) {
                gather(true);
                ex.sync();
            }
        }
        // -- row quantities of this substep: to the orientations and mass properties
        ex.each(nc, [&](uint32_t k) {
            const uint32_t ci = c0 + k;
            const float* c = crec(ci);
            T* pb = W.pointb + 4u * kPointAdj * ci;
            T* sd = W.sideb + 2u * kSideFloats * ci;
            T* slot = W.slot + 2u * kSlotFloats * ci;
            clear_slots(ci);
            if (contact_active(c)) {
                const uint32_t fb = contact_u(c, 0), sb = contact_u(c, 1), fl = contact_u(c, 2);
                const bool fi = (fl & kFlagInertF) != 0u, si = (fl & kFlagInertS) != 0u;
                const SideBody<T> bf = side_body<T>(in.bodies, nullptr, pose_of(pose, fb), fb, fi);
                const SideBody<T> bs = side_body<T>(in.bodies, nullptr, pose_of(pose, sb), sb, si);
                const T c1 = split(fb), c2 = split(sb);
                T ibf[6], ibs[6];
                for (int j = 0; j < 6; ++j) {
                    ibf[j] = bf.ib[j] * c1;
                    ibs[j] = bs.ib[j] * c2;
                }
                const V3<T> nrm = ld3<T>(c + 4);
                V3<T> t1, t2;
                make_tangents(nrm, &t1, &t2);
                const uint32_t np = contact_points(c);
                const T count = T(static_cast<float>(np > 1u ? np : 1u));
                const T* row = rows + kRowFloats * ci;
                V3<T> reach_a = zero3<T>(), reach_b = zero3<T>();
                for (uint32_t p = 0; p < np; ++p) {
                    reach_a += ld3(row + 12u * p, 0);
                    reach_b += ld3(row + 12u * p + 4u, 0);
                }
                V3<T> reachab = zero3<T>(), reachbb = zero3<T>();
                V3<T> rab[4], rbb[4];
                Q4<T> qfb = {T(0.0f), T(0.0f), T(0.0f), T(0.0f)}, qsb = qfb;
                T ibfs[6] = {T(0.0f), T(0.0f), T(0.0f), T(0.0f), T(0.0f), T(0.0f)};
                T ibss[6] = {T(0.0f), T(0.0f), T(0.0f), T(0.0f), T(0.0f), T(0.0f)};
                T imfs = T(0.0f), imss = T(0.0f);  // adjoints of the split inverse masses
                for (uint32_t p = 0; p < np; ++p) {
                    const T* q = row + 12u * p;
                    const V3<T> ra = ld3(q, 0), rb = ld3(q + 4, 0);
                    T* pp = pb + kPointAdj * p;
                    rab[p] = ld3(pp, 0);
                    rbb[p] = ld3(pp + 4, 0);
                    // nmass = count (imf' + ims') + <I'_f (reach_a x n), ra x n> + <I'_s (reach_b x n), rb x n>
                    const T nb = pp[3];
                    imfs += count * nb;
                    imss += count * nb;
                    {
                        const V3<T> u = cross(reach_a, nrm), wv = cross(ra, nrm);
                        const V3<T> y = inertia_apply(bf.q, ibf, u);
                        V3<T> ub = zero3<T>();
                        if (!fi) inertia_vjp(bf.q, ibf, u, wv * nb, &qfb, ibfs, &ub);
                        reachab += cross(nrm, ub);
                        rab[p] += cross(nrm, y * nb);
                    }
                    {
                        const V3<T> u = cross(reach_b, nrm), wv = cross(rb, nrm);
                        const V3<T> y = inertia_apply(bs.q, ibs, u);
                        V3<T> ub = zero3<T>();
                        if (!si) inertia_vjp(bs.q, ibs, u, wv * nb, &qsb, ibss, &ub);
                        reachbb += cross(nrm, ub);
                        rbb[p] += cross(nrm, y * nb);
                    }
                    // m1, m2 = count (imf' + ims' + <x, I'_f x> + <y, I'_s y>), x = ra x t, y = rb x t
                    for (int t = 0; t < 2; ++t) {
                        const V3<T> tt = t == 0 ? t1 : t2;
                        const T mb = (t == 0 ? pp[8] : pp[9]) * count;
                        imfs += mb;
                        imss += mb;
                        const V3<T> xa = cross(ra, tt), xs = cross(rb, tt);
                        V3<T> xab = inertia_apply(bf.q, ibf, xa) * mb, xsb = inertia_apply(bs.q, ibs, xs) * mb;
                        if (!fi) inertia_vjp(bf.q, ibf, xa, xa * mb, &qfb, ibfs, &xab);
                        if (!si) inertia_vjp(bs.q, ibs, xs, xs * mb, &qsb, ibss, &xsb);
                        rab[p] += cross(tt, xab);
                        rbb[p] += cross(tt, xsb);
                    }
                }
                // ra = q la - q com (the position cancels), the reach sums every point's arm
                const V3<T> cfl = ld3<T>(brec(fb) + kBodyCom), csl = ld3<T>(brec(sb) + kBodyCom);
                for (uint32_t p = 0; p < np; ++p) {
                    const float* pr = point_rec(c, p);
                    const V3<T> a = rab[p] + reachab, b = rbb[p] + reachbb;
                    if (!fi) {
                        qrot_vjp(bf.q, ld3<T>(pr), a, &qfb, static_cast<V3<T>*>(nullptr));
                        qrot_vjp(bf.q, cfl, -a, &qfb, static_cast<V3<T>*>(nullptr));
                    }
                    if (!si) {
                        qrot_vjp(bs.q, ld3<T>(pr + 4), b, &qsb, static_cast<V3<T>*>(nullptr));
                        qrot_vjp(bs.q, csl, -b, &qsb, static_cast<V3<T>*>(nullptr));
                    }
                }
                // split masses: I' = c I, m' = c m
                for (int j = 0; j < 6; ++j) {
                    sd[4 + j] += ibfs[j] * c1;
                    sd[kSideFloats + 4 + j] += ibss[j] * c2;
                }
                sd[10] += imfs * c1;
                sd[kSideFloats + 10] += imss * c2;
                sd[0] += qfb.x;
                sd[1] += qfb.y;
                sd[2] += qfb.z;
                sd[3] += qfb.w;
                sd[kSideFloats + 0] += qsb.x;
                sd[kSideFloats + 1] += qsb.y;
                sd[kSideFloats + 2] += qsb.z;
                sd[kSideFloats + 3] += qsb.w;
            }
            // the block's side totals to the slots; the point adjoints of this substep are spent
            for (int side = 0; side < 2; ++side) {
                T* sdd = sd + kSideFloats * side;
                T* sl = slot + kSlotFloats * side;
                for (int j = 0; j < 11; ++j) sl[j < 4 ? 3 + j : j + 4] = sdd[j];  // q at 3..6, ib at 8..13, im at 14
                for (uint32_t j = 0; j < kSideFloats; ++j) sdd[j] = T(0.0f);
            }
            for (uint32_t p = 0; p < 4; ++p) {
                T* pp = pb + kPointAdj * p;
                for (int j = 0; j < 10; ++j) pp[j] = T(0.0f);
            }
        });
        ex.sync();
        // rows: orientation adjoint, and the mass scale from the inverse mass and inertia (both / s)
        ex.each(wr.row_count, [&](uint32_t i) {
            const uint32_t row = r0 + i;
            const uint32_t lc = W.list_count[row];
            T* xb = W.xb + kStateFloats * row;
            const float* rec = brec(row);
            T sb = T(0.0f);
            for (uint32_t e = 0; e < lc; ++e) {
                const uint32_t entry = W.list[W.list_start[row] + e];
                const T* sl = W.slot + 2u * kSlotFloats * (entry >> 1) + kSlotFloats * (entry & 1u);
                for (int j = 0; j < 4; ++j) xb[4 + j] += sl[3 + j];
                for (int j = 0; j < 6; ++j) sb -= T(rec[kBodyInvInertia + j]) * sl[8 + j];
                sb -= T(rec[kBodyIm]) * sl[14];
            }
            W.sbar[row] += sb;
        });
        ex.sync();
        // -- integrate: the local Jacobian by forward-mode duals over (v, w, q, s)
        ex.each(wr.row_count, [&](uint32_t i) {
            const uint32_t row = r0 + i;
            const float* rec = brec(row);
            const uint32_t f = body_flags(rec);
            if (!(f & kBodySimulates) || (f & kBodyInert)) return;
            using D = Dual<11, T>;
            const T* v0 = vel + 8u * row;
            const T* q0 = pose + 8u * row + 4u;
            D in_[11];
            for (int j = 0; j < 3; ++j) {
                in_[j] = D(v0[j]);
                in_[3 + j] = D(v0[4 + j]);
            }
            for (int j = 0; j < 4; ++j) in_[6 + j] = D(q0[j]);
            in_[10] = D(T(1.0f));
            for (int j = 0; j < 11; ++j) in_[j].d[j] = T(1.0f);
            V3<D> v = {in_[0], in_[1], in_[2]}, w = {in_[3], in_[4], in_[5]};
            integrate_body<D>(P, rec, in_[10], Q4<D>{in_[6], in_[7], in_[8], in_[9]}, v, w);
            T* xb = W.xb + kStateFloats * row;
            const T ob[6] = {xb[8], xb[9], xb[10], xb[12], xb[13], xb[14]};
            const D outs[6] = {v.x, v.y, v.z, w.x, w.y, w.z};
            T gin[11];
            for (int j = 0; j < 11; ++j) {
                gin[j] = T(0.0f);
                for (int o = 0; o < 6; ++o) gin[j] += T(outs[o].d[j]) * ob[o];
            }
            xb[8] = gin[0];
            xb[9] = gin[1];
            xb[10] = gin[2];
            xb[12] = gin[3];
            xb[13] = gin[4];
            xb[14] = gin[5];
            for (int j = 0; j < 4; ++j) xb[4 + j] += gin[6 + j];
            W.sbar[row] += gin[10];
        });
        ex.sync();
    }
    // -- step start: targets (depth and approach speed) and the depth model
    ex.each(nc, [&](uint32_t k) {
        const uint32_t ci = c0 + k;
        const float* c = crec(ci);
        clear_slots(ci);
        if (!(contact_u(c, 2) & kFlagResolves)) return;
        const uint32_t fb = contact_u(c, 0), sb = contact_u(c, 1), fl = contact_u(c, 2);
        const bool fi = (fl & kFlagInertF) != 0u, si = (fl & kFlagInertS) != 0u;
        const V3<T> nrm = ld3<T>(c + 4);
        const T* x0f = vel_of(W.vel, fb);  // velocities at the step start (substep 0, before integration)
        const T* x0s = vel_of(W.vel, sb);
        const T* pf0 = pose_of(W.pose, fb);  // pose at the step start
        const T* ps0 = pose_of(W.pose, sb);
        const Q4<T> qf = ld4(pf0 + 4), qs = ld4(ps0 + 4);
        const V3<T> comf = ld3(pf0, 0) + qrot(qf, ld3<T>(brec(fb) + kBodyCom));
        const V3<T> coms = ld3(ps0, 0) + qrot(qs, ld3<T>(brec(sb) + kBodyCom));
        const V3<T> vf = ld3(x0f, 0), wf = ld3(x0f + 4, 0), vs = ld3(x0s, 0), ws = ld3(x0s + 4, 0);
        T* pb = W.pointb + 4u * kPointAdj * ci;
        V3<T> pfb = zero3<T>(), psb = zero3<T>(), vfb = zero3<T>(), wfb = zero3<T>(), vsb = zero3<T>(), wsb = zero3<T>();
        Q4<T> qfb = {T(0.0f), T(0.0f), T(0.0f), T(0.0f)}, qsb = qfb;
        for (uint32_t p = 0; p < contact_points(c); ++p) {
            const float* pr = point_rec(c, p);
            T* pp = pb + kPointAdj * p;
            const T d = W.depth[4 * ci + p];
            T db = pp[11];
            const T tb = pp[10];
            const V3<T> pos = ld3<T>(pr + 8);
            const V3<T> rel = (vs + cross(ws, pos - coms)) - (vf + cross(wf, pos - comf));
            const T speed = dot(rel, nrm);
            if (val(d) < 0.0f) {
                db += tb / T(P.dt);
            } else if (val(speed) < -P.restitution_threshold) {
                const V3<T> relb = nrm * (-tb * T(c[10]));
                vsb += relb;
                wsb += cross(pos - coms, relb);
                const V3<T> comsb = -cross(relb, ws);
                vfb -= relb;
                wfb -= cross(pos - comf, relb);
                const V3<T> comfb = cross(relb, wf);
                psb += comsb;
                pfb += comfb;
                qrot_vjp(qs, ld3<T>(brec(sb) + kBodyCom), comsb, &qsb, static_cast<V3<T>*>(nullptr));
                qrot_vjp(qf, ld3<T>(brec(fb) + kBodyCom), comfb, &qfb, static_cast<V3<T>*>(nullptr));
            }
            // depth = depth_fwd + n . ((q_f la + p_f) - (q_s lb + p_s) - the same at the forward)
            pfb += nrm * db;
            psb -= nrm * db;
            qrot_vjp(qf, ld3<T>(pr), nrm * db, &qfb, static_cast<V3<T>*>(nullptr));
            qrot_vjp(qs, ld3<T>(pr + 4), -nrm * db, &qsb, static_cast<V3<T>*>(nullptr));
            pp[10] = pp[11] = T(0.0f);
        }
        T* slot = W.slot + 2u * kSlotFloats * ci;
        if (!fi) {
            st3(slot, pfb);
            st4(slot + 3, qfb);
            st3(slot + 7, vfb);
            st3(slot + 10, wfb);
        }
        if (!si) {
            st3(slot + kSlotFloats, psb);
            st4(slot + kSlotFloats + 3, qsb);
            st3(slot + kSlotFloats + 7, vsb);
            st3(slot + kSlotFloats + 10, wsb);
        }
    });
    ex.sync();
    ex.each(wr.row_count, [&](uint32_t i) {
        const uint32_t row = r0 + i;
        const uint32_t lc = W.list_count[row];
        T* xb = W.xb + kStateFloats * row;
        for (uint32_t e = 0; e < lc; ++e) {
            const uint32_t entry = W.list[W.list_start[row] + e];
            const T* sl = W.slot + 2u * kSlotFloats * (entry >> 1) + kSlotFloats * (entry & 1u);
            for (int j = 0; j < 3; ++j) xb[j] += sl[j];
            for (int j = 0; j < 4; ++j) xb[4 + j] += sl[3 + j];
            for (int j = 0; j < 3; ++j) xb[8 + j] += sl[7 + j];
            for (int j = 0; j < 3; ++j) xb[12 + j] += sl[10 + j];
        }
    });
    // -- friction: the combined coefficient's adjoint to the colliders
    ex.each(wr.collider_count, [&](uint32_t i) {
        const uint32_t col = wr.collider_base + i;
        T sum = T(0.0f);
        for (uint32_t ci = c0; ci < in.c1; ++ci) {
            const float* c = crec(ci);
            if (contact_u(c, 3) == col) sum += W.mucb[ci] * T(c[8]);
            if (contact_u(c, 11) == col) sum += W.mucb[ci] * T(c[9]);
        }
        W.mubar[col] += sum;
    });
    // -- impulses entering the step
    if (accb_start) {
        ex.each(nc, [&](uint32_t k) {
            const uint32_t ci = c0 + k;
            for (uint32_t j = 0; j < kAccFloats; ++j) accb_start[kAccFloats * ci + j] = W.accb[kAccFloats * ci + j];
        });
    }
    ex.sync();
}

// Sends the adjoint of the impulses a step started from to the points of the previous step they
// were relayed from (contact_relay.wgsl, identity.wgsl:85-117): `prev` must be zero beforehand.
template <class Ex, class T>
ADJ_FN void relay_back(const Ex& ex, const StepIn& in, const T* accb_start, T* prev) {
    const uint32_t nc = in.c1 - in.c0;
    ex.each(nc, [&](uint32_t k) {
        const uint32_t ci = in.c0 + k;
        const float* c = in.contacts + static_cast<size_t>(kContactFloats) * ci;
        for (uint32_t p = 0; p < contact_points(c); ++p) {
            const uint32_t src = fbits(point_rec(c, p)[11]);
            if (src == kNoSource) continue;
            const uint32_t pc = src >> 2, pp = src & 3u;
            prev[kAccFloats * pc + pp] = accb_start[kAccFloats * ci + p];
            prev[kAccFloats * pc + 4 + pp] = accb_start[kAccFloats * ci + 4 + p];
            prev[kAccFloats * pc + 8 + pp] = accb_start[kAccFloats * ci + 8 + p];
        }
    });
    ex.sync();
}

}  // namespace dyn_adj

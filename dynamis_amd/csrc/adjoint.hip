// The differentiable rollout on the device: checkpoints of the production step every k steps,
// recompute of each segment with its tape recorded, and the replay + reverse of adjoint.h per step
// with one workgroup per world.
//
// A world of the differentiation workload (tens of bodies, a hundred contacts) fits one
// workgroup: its phases are separated by workgroup barriers instead of kernel launches, and every
// per-row sum walks the row's block list in contact order, so the adjoint is deterministic. A
// batch's worlds run side by side, one per workgroup (256 worlds fill the 256 CUs).
#include <vector>

#include "adjoint.h"
#include "common.h"

namespace dyn {
hipStream_t world_stream(dyn_world* world);
void world_sizes(dyn_world* world, uint32_t* n, uint32_t* d, uint32_t* k, uint32_t* cap);
void world_lock_growth(dyn_world* world, int delta);
const dyn_step_params& world_params(dyn_world* world);
const dyn_body_state* world_states(dyn_world* world);
const dyn_body_desc* world_descs(dyn_world* world);
void world_record(dyn_world* world, float* states, float* contacts, uint32_t* count, uint32_t* error,
                  uint32_t* relay_src);
}  // namespace dyn

namespace {

using namespace dyn_adj;

constexpr int kAdjThreads = 256;

struct DevExec {
    __device__ void sync() const { __syncthreads(); }
    template <class F>
    __device__ void each(uint32_t n, F&& f) const {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) f(i);
    }
    __device__ bool leader() const { return threadIdx.x == 0; }
};

// Per row: the body record of adjoint.h (constant through a rollout).
__global__ void __launch_bounds__(kAdjThreads) k_adj_bodies(const dyn_body_state* st, const dyn_body_desc* de,
                                                            uint32_t n, float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const dyn_body_desc& d = de[i];
    float* o = out + static_cast<size_t>(kBodyFloats) * i;
    o[kBodyIm] = d.inverse_mass;
    o[kBodyGravityScale] = d.gravity_scale;
    o[kBodyLinearDamping] = d.linear_damping;
    o[kBodyAngularDamping] = d.angular_damping;
    for (int k = 0; k < 6; ++k) {
        o[kBodyInvInertia + k] = d.inverse_inertia[k];
        o[kBodyInertia + k] = d.inertia[k];
    }
    for (int k = 0; k < 3; ++k) o[kBodyCom + k] = d.com[k];
    const bool kinematic = (d.flags & DYN_BODY_KINEMATIC) != 0u;
    const bool sim = !kinematic && d.inverse_mass > 0.0f;
    const bool mov = kinematic || d.inverse_mass > 0.0f;
    const bool inert = !sim || st[i].sleeping != 0u;
    o[kBodyFlags] = bitsf((sim ? kBodySimulates : 0u) | (mov ? kBodyMoves : 0u) | (inert ? kBodyInert : 0u));
}

__global__ void __launch_bounds__(kAdjThreads) k_adj_zero(float* p, size_t count) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0f;
}

// One world's step backwards over the batch-wide arrays (global row and contact indices): its
// replay from the recorded start, then the reverse, then the impulses' adjoint sent to the
// previous step's contacts (accb_prev, zero beforehand).
__device__ void step_world_global(const AdjParams& P, const AdjRange& wr, const float* bodies, const float* contacts,
                                  uint32_t c0, uint32_t c1, const float* states, Work<float> W,
                                  const float* accb_end, float* accb_prev) {
    const DevExec ex;
    StepIn in;
    in.bodies = bodies;
    in.contacts = contacts;
    in.fwd = states;
    in.c0 = c0;
    in.c1 = c1;
    ex.each(wr.row_count, [&](uint32_t i) {
        const size_t row = wr.row_base + i;
        for (uint32_t j = 0; j < kStateFloats; ++j) W.x[kStateFloats * row + j] = states[kStateFloats * row + j];
    });
    ex.sync();
    replay_step(ex, P, wr, in, W, static_cast<const float*>(nullptr), static_cast<const float*>(nullptr),
                static_cast<const float*>(nullptr));
    reverse_step(ex, P, wr, in, W, accb_end, static_cast<float*>(nullptr));
    if (accb_prev) relay_back(ex, in, W.accb, accb_prev);
}

// ---- the world-local step ------------------------------------------------------------------------
//
// A world's rows, contacts and colliders renumbered from zero (its simulated rows, then the static
// rows its contacts touch), so that the step's per-row and per-contact arrays are dense and small:
// the ones every phase reads (states, velocity and pose stores, resolutions, block words, depths,
// lists, the state adjoint, and the contact records when they fit) sit in LDS, the per-sweep
// impulse and row stores in a per-world stretch of the batch-wide arrays. Phases then wait on LDS
// instead of on L2 round trips through arrays strided by the whole batch.

constexpr uint32_t kMaxStaticRows = 8;

__host__ __device__ inline size_t lds_align(size_t b) { return (b + 15u) & ~static_cast<size_t>(15u); }

// LDS bytes of a world with n local rows and c contacts (contacts in LDS or not)
__host__ __device__ inline size_t local_bytes(const AdjParams& P, uint32_t n, uint32_t c, bool contacts_in_lds) {
    const size_t S = P.substeps, I = P.iterations, Q = P.position_iterations;
    size_t b = 0;
    if (contacts_in_lds) b += lds_align(4u * kContactFloats * static_cast<size_t>(c));
    b += lds_align(4u * kBodyFloats * n) + lds_align(4u * kStateFloats * n);  // bodies, recorded start
    b += lds_align(4u * kStateFloats * n);                                     // x
    b += lds_align(4u * S * (3u + I) * n * 8u) + lds_align(4u * S * 2u * n * 8u) + lds_align(4u * S * Q * n * 4u);
    b += 2u * lds_align(4u * 4u * static_cast<size_t>(c));                    // depth, target
    b += lds_align(4u * 2u * kSlotFloats * static_cast<size_t>(c));           // slot
    b += 2u * lds_align(4u * n) + lds_align(4u * 2u * static_cast<size_t>(c)); // list start, count, list
    b += lds_align(4u * kStateFloats * n) + 2u * lds_align(4u * 4u * n);       // xb, resb, alb
    b += lds_align(4u * n) + lds_align(4u * static_cast<size_t>(c));           // sbar, mucb
    return b;
}

__global__ void __launch_bounds__(kAdjThreads) k_adj_step(AdjParams P, const AdjRange* ranges, const float* bodies,
                                                          const float* contacts, const uint32_t* count,
                                                          const float* states, Work<float> work, float* local_contacts,
                                                          const float* accb_end, float* accb_prev, uint32_t lds_bytes) {
    extern __shared__ __align__(16) char lds[];
    __shared__ uint32_t s_static[kMaxStaticRows];
    __shared__ uint32_t s_nstatic, s_mode;
    const AdjRange wr = ranges[blockIdx.x];
    const uint32_t total = *count;
    const uint32_t c0 = lower_collider(contacts, total, wr.collider_base);
    const uint32_t c1 = lower_collider(contacts, total, wr.collider_base + wr.collider_count);
    const uint32_t C = c1 - c0, D = wr.row_count, r0 = wr.row_base, r1 = wr.row_base + wr.row_count;
    if (threadIdx.x == 0) {
        uint32_t ns = 0;
        bool ok = true;
        for (uint32_t ci = c0; ci < c1 && ok; ++ci) {
            const float* c = contacts + static_cast<size_t>(kContactFloats) * ci;
            for (uint32_t side = 0; side < 2; ++side) {
                const uint32_t row = contact_u(c, side);
                if (row >= r0 && row < r1) continue;
                bool seen = false;
                for (uint32_t j = 0; j < ns; ++j) seen |= s_static[j] == row;
                if (seen) continue;
                if (ns == kMaxStaticRows) {
                    ok = false;
                    break;
                }
                s_static[ns++] = row;
            }
        }
        s_nstatic = ns;
        // 0: batch-wide arrays; 1: local, contacts in LDS; 2: local, contacts in the global stretch
        uint32_t mode = 0;
        if (ok && local_bytes(P, D + ns, C, true) <= lds_bytes) mode = 1;
        else if (ok && local_bytes(P, D + ns, C, false) <= lds_bytes) mode = 2;
        s_mode = mode;
    }
    __syncthreads();
    if (s_mode == 0) {
        step_world_global(P, wr, bodies, contacts, c0, c1, states, work, accb_end, accb_prev);
        return;
    }
    const DevExec ex;
    const uint32_t NS = s_nstatic, n = D + NS;
    const size_t S = P.substeps, I = P.iterations, Q = P.position_iterations;
    char* at = lds;
    auto take = [&](size_t bytes) -> float* {
        float* p = reinterpret_cast<float*>(at);
        at += lds_align(bytes);
        return p;
    };
    float* lc = s_mode == 1 ? take(4u * kContactFloats * static_cast<size_t>(C))
                            : local_contacts + static_cast<size_t>(kContactFloats) * c0;
    float* lb = take(4u * kBodyFloats * n);
    float* lf = take(4u * kStateFloats * n);
    Work<float> L;
    L.n = n;
    L.cap = C;
    L.x = take(4u * kStateFloats * n);
    L.vel = take(4u * S * (3u + I) * n * 8u);
    L.pose = take(4u * S * 2u * n * 8u);
    L.res = take(4u * S * Q * n * 4u);
    L.depth = take(4u * 4u * C);
    L.target = take(4u * 4u * C);
    L.slot = take(4u * 2u * kSlotFloats * C);
    L.list_start = reinterpret_cast<uint32_t*>(take(4u * n));
    L.list_count = reinterpret_cast<uint32_t*>(take(4u * n));
    L.list = reinterpret_cast<uint32_t*>(take(4u * 2u * C));
    L.xb = take(4u * kStateFloats * n);
    L.resb = take(4u * 4u * n);
    L.alb = take(4u * 4u * n);
    L.sbar = take(4u * n);
    L.mucb = take(4u * C);
    L.acc = work.acc + S * (I + 1u) * kAccFloats * static_cast<size_t>(c0);
    L.rows = work.rows + S * kRowFloats * static_cast<size_t>(c0);
    L.pointb = work.pointb + 4u * kPointAdj * static_cast<size_t>(c0);
    L.sideb = work.sideb + 2u * kSideFloats * static_cast<size_t>(c0);
    L.accb = work.accb + kAccFloats * static_cast<size_t>(c0);
    L.mubar = work.mubar + wr.collider_base;
    auto global_row = [&](uint32_t local) { return local < D ? r0 + local : s_static[local - D]; };
    auto local_row = [&](uint32_t row) {
        if (row >= r0 && row < r1) return row - r0;
        uint32_t j = 0;
        while (j + 1u < NS && s_static[j] != row) ++j;
        return D + j;
    };
    ex.each(C, [&](uint32_t k) {
        const float* src = contacts + static_cast<size_t>(kContactFloats) * (c0 + k);
        float* dst = lc + static_cast<size_t>(kContactFloats) * k;
        for (uint32_t j = 0; j < kContactFloats; ++j) dst[j] = src[j];
        dst[0] = bitsf(local_row(contact_u(src, 0)));
        dst[1] = bitsf(local_row(contact_u(src, 1)));
        dst[3] = bitsf(contact_u(src, 3) - wr.collider_base);
        dst[11] = bitsf(contact_u(src, 11) - wr.collider_base);  // relay sources stay batch-wide
    });
    ex.each(n, [&](uint32_t i) {
        const size_t g = global_row(i);
        for (uint32_t j = 0; j < kBodyFloats; ++j) lb[kBodyFloats * i + j] = bodies[kBodyFloats * g + j];
        for (uint32_t j = 0; j < kStateFloats; ++j) {
            lf[kStateFloats * i + j] = states[kStateFloats * g + j];
            L.x[kStateFloats * i + j] = states[kStateFloats * g + j];
            L.xb[kStateFloats * i + j] = i < D ? work.xb[kStateFloats * g + j] : 0.0f;
        }
        L.sbar[i] = 0.0f;
    });
    ex.sync();
    const AdjRange lr = {0u, D, 0u, wr.collider_count};
    StepIn in;
    in.bodies = lb;
    in.contacts = lc;
    in.fwd = lf;
    in.c0 = 0;
    in.c1 = C;
    replay_step(ex, P, lr, in, L, static_cast<const float*>(nullptr), static_cast<const float*>(nullptr),
                static_cast<const float*>(nullptr));
    reverse_step(ex, P, lr, in, L, accb_end ? accb_end + kAccFloats * static_cast<size_t>(c0) : nullptr,
                 static_cast<float*>(nullptr));
    if (accb_prev) relay_back(ex, in, L.accb, accb_prev);
    ex.each(D, [&](uint32_t i) {
        const size_t g = r0 + i;
        for (uint32_t j = 0; j < kStateFloats; ++j) work.xb[kStateFloats * g + j] = L.xb[kStateFloats * i + j];
        work.sbar[g] += L.sbar[i];
    });
}

// d loss / d mass = (d loss / d scale) / mass at fixed shape
__global__ void __launch_bounds__(kAdjThreads) k_adj_mass(const float* sbar, const float* bodies, uint32_t n,
                                                          float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* b = bodies + static_cast<size_t>(kBodyFloats) * i;
    out[i] = (fbits(b[kBodyFlags]) & kBodySimulates) ? sbar[i] * b[kBodyIm] : 0.0f;
}

uint32_t grid_for(size_t n) { return static_cast<uint32_t>((n + kAdjThreads - 1) / kAdjThreads); }

}  // namespace

struct dyn_adjoint {
    dyn_world* world = nullptr;
    hipStream_t stream = nullptr;
    uint32_t steps = 0, every = 0, segments = 0, worlds = 0;
    uint32_t n = 0, d = 0, k = 0, cap = 0;
    float dt = 0.0f;
    bool forwarded = false;
    bool last_recorded = false;  // the tape slots hold the last segment, recorded by the forward rollout
    AdjParams P{};
    AdjRange* d_ranges = nullptr;
    size_t snap_bytes = 0;
    char* d_snaps = nullptr;       // segments + 1 checkpoints (the last: the rollout's end)
    float* d_bodies = nullptr;
    float* d_states = nullptr;     // every x n x 16: the segment's recorded start states
    float* d_contacts = nullptr;   // every x cap x 76
    uint32_t* d_counts = nullptr;  // every
    uint32_t* d_error = nullptr;
    uint32_t* d_relay_src = nullptr;
    float* d_accb[2] = {nullptr, nullptr};
    float* d_local_contacts = nullptr;  // cap x 76: renumbered contact records of worlds whose records stay out of LDS
    uint32_t lds_bytes = 0;             // dynamic LDS of k_adj_step
    bool locked = false;                // holds the world's capacity fixed
    Work<float> W{};
    std::vector<void*> owned;
};

namespace {

template <typename T>
int adj_alloc(dyn_adjoint* a, T** ptr, size_t count) {
    void* p = nullptr;
    DYN_HIP(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
    a->owned.push_back(p);
    DYN_HIP(hipMemset(p, 0, (count ? count : 1) * sizeof(T)));
    *ptr = static_cast<T*>(p);
    return DYN_OK;
}

int zero(dyn_adjoint* a, float* p, size_t count) {
    if (!count) return DYN_OK;
    k_adj_zero<<<grid_for(count), kAdjThreads, 0, a->stream>>>(p, count);
    DYN_LAUNCHED("k_adj_zero");
    return DYN_OK;
}

void destroy(dyn_adjoint* a) {
    if (a->locked) dyn::world_lock_growth(a->world, -1);
    for (void* p : a->owned) (void)hipFree(p);
    delete a;
}

}  // namespace

extern "C" {

int dyn_adjoint_create(dyn_adjoint** out, dyn_world* world, uint32_t steps, uint32_t checkpoint_every,
                       const uint32_t* world_ranges, uint32_t world_count) {
    using namespace dyn;
    if (!out || !world || (world_count && !world_ranges)) return fail(DYN_E_INVALID, "dyn_adjoint_create: null argument");
    if (steps == 0 || checkpoint_every == 0) return fail(DYN_E_INVALID, "dyn_adjoint_create: steps and checkpoint_every must be positive");
    if (world_count == 0) return fail(DYN_E_INVALID, "dyn_adjoint_create: no worlds");
    *out = nullptr;
    dyn_adjoint* a = new dyn_adjoint();
    a->world = world;
    a->stream = world_stream(world);
    // the tapes, snapshots and relay sources are sized by the world's contact capacity: the
    // world's streams keep their size while this rollout exists
    world_lock_growth(world, 1);
    a->locked = true;
    a->steps = steps;
    a->every = checkpoint_every < steps ? checkpoint_every : steps;
    a->segments = (steps + a->every - 1) / a->every;
    a->worlds = world_count;
    world_sizes(world, &a->n, &a->d, &a->k, &a->cap);
    std::vector<AdjRange> ranges(world_count);
    for (uint32_t i = 0; i < world_count; ++i) {
        const uint32_t* r = world_ranges + 4u * i;
        ranges[i] = AdjRange{r[0], r[1], r[2], r[3]};
        if (static_cast<uint64_t>(r[0]) + r[1] > a->d || static_cast<uint64_t>(r[2]) + r[3] > a->k) {
            destroy(a);
            return fail(DYN_E_INVALID, "dyn_adjoint_create: a world's rows or colliders reach past the world's simulated rows or colliders");
        }
    }
    const dyn_step_params& p = world_params(world);
    const size_t n = a->n, cap = a->cap, S = p.substeps, I = p.solve_iterations, Q = p.position_iterations;
    int rc = DYN_OK;
    size_t snap = 0;
    if (dyn_world_snapshot_bytes(world, &snap) != DYN_OK) {
        destroy(a);
        return DYN_E_INVALID;
    }
    a->snap_bytes = snap;
#define ADJ_A(ptr, count) if (rc == DYN_OK) rc = adj_alloc(a, &(ptr), (count))
    ADJ_A(a->d_ranges, world_count);
    ADJ_A(a->d_snaps, snap * (a->segments + 1u));
    ADJ_A(a->d_bodies, n * kBodyFloats);
    ADJ_A(a->d_states, static_cast<size_t>(a->every) * n * kStateFloats);
    ADJ_A(a->d_contacts, static_cast<size_t>(a->every) * cap * kContactFloats);
    ADJ_A(a->d_counts, a->every);
    ADJ_A(a->d_error, 1);
    ADJ_A(a->d_relay_src, 4u * cap);
    ADJ_A(a->d_accb[0], cap * kAccFloats);
    ADJ_A(a->d_accb[1], cap * kAccFloats);
    ADJ_A(a->d_local_contacts, cap * kContactFloats);
    Work<float>& W = a->W;
    W.n = a->n;
    W.cap = a->cap;
    ADJ_A(W.x, n * kStateFloats);
    ADJ_A(W.vel, S * (3u + I) * n * 8u);
    ADJ_A(W.pose, S * 2u * n * 8u);
    ADJ_A(W.res, S * Q * n * 4u);
    ADJ_A(W.acc, S * (I + 1u) * cap * kAccFloats);
    ADJ_A(W.rows, S * cap * kRowFloats);
    ADJ_A(W.depth, cap * 4u);
    ADJ_A(W.target, cap * 4u);
    ADJ_A(W.slot, cap * 2u * kSlotFloats);
    ADJ_A(W.list_start, n);
    ADJ_A(W.list_count, n);
    ADJ_A(W.list, 2u * cap);
    ADJ_A(W.xb, n * kStateFloats);
    ADJ_A(W.resb, n * 4u);
    ADJ_A(W.alb, n * 4u);
    ADJ_A(W.sbar, n);
    ADJ_A(W.mubar, a->k);
    ADJ_A(W.pointb, cap * 4u * kPointAdj);
    ADJ_A(W.sideb, cap * 2u * kSideFloats);
    ADJ_A(W.accb, cap * kAccFloats);
    ADJ_A(W.mucb, cap);
#undef ADJ_A
    if (rc == DYN_OK && hipMemcpy(a->d_ranges, ranges.data(), sizeof(AdjRange) * world_count, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(DYN_E_HIP, "dyn_adjoint_create: world ranges");
    }
    if (rc == DYN_OK) {
        int device = 0, lds = 0;
        if (hipGetDevice(&device) != hipSuccess ||
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
            rc = fail(DYN_E_HIP, "dyn_adjoint_create: LDS size");
        } else {
            // the kernel's static LDS (static rows, mode) is a few words
            a->lds_bytes = lds > 1024 ? static_cast<uint32_t>(lds - 1024) : 0u;
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_adj_step),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(a->lds_bytes));
            (void)hipGetLastError();
        }
    }
    if (rc != DYN_OK) {
        destroy(a);
        return rc;
    }
    *out = a;
    return DYN_OK;
}

int dyn_adjoint_forward(dyn_adjoint* a, float dt) {
    using namespace dyn;
    if (!a) return fail(DYN_E_INVALID, "dyn_adjoint_forward: null adjoint");
    if (!(dt > 0.0f)) return fail(DYN_E_INVALID, "dyn_adjoint_forward: dt must be positive");
    const dyn_step_params& p = world_params(a->world);
    AdjParams& P = a->P;
    P.dt = dt;
    P.h = dt / static_cast<float>(p.substeps);
    for (int i = 0; i < 3; ++i) P.gravity[i] = p.gravity[i];
    P.slop = p.slop;
    P.relaxation = p.relaxation;
    P.restitution_threshold = p.restitution_threshold;
    P.max_velocity = p.max_velocity;
    P.max_angular_velocity = p.max_angular_velocity;
    P.substeps = p.substeps;
    P.iterations = p.solve_iterations;
    P.position_iterations = p.position_iterations;
    P.friction_combine = p.friction_combine;
    a->dt = dt;
    k_adj_bodies<<<grid_for(a->n), kAdjThreads, 0, a->stream>>>(world_states(a->world), world_descs(a->world), a->n,
                                                                 a->d_bodies);
    DYN_LAUNCHED("k_adj_bodies");
    // the last segment records its tape as it steps, so the backward pass starts without a recompute
    const uint32_t last = (a->segments - 1u) * a->every;
    const size_t n = a->n, cap = a->cap;
    DYN_TRY(zero(a, reinterpret_cast<float*>(a->d_error), 1));
    a->last_recorded = false;
    for (uint32_t t = 0; t < a->steps; ++t) {
        if (t % a->every == 0) DYN_TRY(dyn_world_snapshot(a->world, a->d_snaps + a->snap_bytes * (t / a->every)));
        if (t >= last) {
            const size_t j = t - last;
            world_record(a->world, a->d_states + j * n * kStateFloats, a->d_contacts + j * cap * kContactFloats,
                         a->d_counts + j, a->d_error, a->d_relay_src);
        }
        const int rc = dyn_world_step(a->world, dt);
        world_record(a->world, nullptr, nullptr, nullptr, nullptr, nullptr);
        DYN_TRY(rc);
    }
    DYN_TRY(dyn_world_snapshot(a->world, a->d_snaps + a->snap_bytes * a->segments));
    a->forwarded = true;
    a->last_recorded = true;
    return DYN_OK;
}

int dyn_adjoint_backward(dyn_adjoint* a, const float* d_state_grad, float* d_state0_grad, float* d_mass_grad,
                         float* d_friction_grad) {
    using namespace dyn;
    if (!a || !d_state_grad) return fail(DYN_E_INVALID, "dyn_adjoint_backward: null argument");
    if (!a->forwarded) return fail(DYN_E_INVALID, "dyn_adjoint_backward: no forward rollout to differentiate");
    hipStream_t st = a->stream;
    Work<float>& W = a->W;
    const size_t n = a->n, cap = a->cap;
    DYN_HIP(hipMemcpyAsync(W.xb, d_state_grad, sizeof(float) * n * kStateFloats, hipMemcpyDeviceToDevice, st));
    DYN_TRY(zero(a, W.sbar, n));
    DYN_TRY(zero(a, W.mubar, a->k));
    DYN_TRY(zero(a, a->d_accb[0], cap * kAccFloats));
    DYN_TRY(zero(a, a->d_accb[1], cap * kAccFloats));
    for (uint32_t seg = a->segments; seg-- > 0;) {
        const uint32_t first = seg * a->every;
        const uint32_t len = (first + a->every <= a->steps) ? a->every : a->steps - first;
        if (!(seg + 1u == a->segments && a->last_recorded)) {
            if (seg + 1u == a->segments) DYN_TRY(zero(a, reinterpret_cast<float*>(a->d_error), 1));
            DYN_TRY(dyn_world_restore(a->world, a->d_snaps + a->snap_bytes * seg));
            for (uint32_t j = 0; j < len; ++j) {
                world_record(a->world, a->d_states + static_cast<size_t>(j) * n * kStateFloats,
                             a->d_contacts + static_cast<size_t>(j) * cap * kContactFloats, a->d_counts + j,
                             a->d_error, a->d_relay_src);
                const int rc = dyn_world_step(a->world, a->dt);
                world_record(a->world, nullptr, nullptr, nullptr, nullptr, nullptr);
                DYN_TRY(rc);
            }
            if (seg + 1u != a->segments) a->last_recorded = false;  // the slots now hold this segment
        }
        if (seg + 1u == a->segments) {  // the model's limits hold for the whole rollout or it is refused
            uint32_t error = 0;
            DYN_HIP(hipMemcpyAsync(&error, a->d_error, 4u, hipMemcpyDeviceToHost, st));
            DYN_HIP(hipStreamSynchronize(st));
            if (error & 1u) return fail(DYN_E_INVALID, "dyn_adjoint_backward: soft contacts or rolling/spin friction are outside the differentiable step");
            if (error & 2u) return fail(DYN_E_INVALID, "dyn_adjoint_backward: a body sleeps (turn sleeping off for differentiation)");
        }
        for (uint32_t j = len; j-- > 0;) {
            const uint32_t t = first + j;
            float* prev = t > 0 ? a->d_accb[(t + 1u) & 1u] : nullptr;
            if (prev) DYN_TRY(zero(a, prev, cap * kAccFloats));
            k_adj_step<<<a->worlds, kAdjThreads, a->lds_bytes, st>>>(
                a->P, a->d_ranges, a->d_bodies, a->d_contacts + static_cast<size_t>(j) * cap * kContactFloats,
                a->d_counts + j, a->d_states + static_cast<size_t>(j) * n * kStateFloats, W, a->d_local_contacts,
                a->d_accb[t & 1u], prev, a->lds_bytes);
            DYN_LAUNCHED("k_adj_step");
        }
    }
    // leave the world where the forward rollout ended (already so when nothing was recomputed)
    if (!a->last_recorded) DYN_TRY(dyn_world_restore(a->world, a->d_snaps + a->snap_bytes * a->segments));
    if (d_state0_grad) {
        DYN_HIP(hipMemcpyAsync(d_state0_grad, W.xb, sizeof(float) * n * kStateFloats, hipMemcpyDeviceToDevice, st));
    }
    if (d_mass_grad) {
        k_adj_mass<<<grid_for(n), kAdjThreads, 0, st>>>(W.sbar, a->d_bodies, a->n, d_mass_grad);
        DYN_LAUNCHED("k_adj_mass");
    }
    if (d_friction_grad && a->k) {
        DYN_HIP(hipMemcpyAsync(d_friction_grad, W.mubar, sizeof(float) * a->k, hipMemcpyDeviceToDevice, st));
    }
    return DYN_OK;
}

int dyn_adjoint_read_tape(dyn_adjoint* a, uint32_t slot, float* states, float* contacts, uint32_t capacity,
                          uint32_t* count) {
    using namespace dyn;
    if (!a || !count) return fail(DYN_E_INVALID, "dyn_adjoint_read_tape: null argument");
    if (slot >= a->every) return fail(DYN_E_INVALID, "dyn_adjoint_read_tape: slot past the segment length");
    DYN_HIP(hipStreamSynchronize(a->stream));
    uint32_t c = 0;
    DYN_HIP(hipMemcpy(&c, a->d_counts + slot, 4u, hipMemcpyDeviceToHost));
    *count = c;
    if (states) {
        DYN_HIP(hipMemcpy(states, a->d_states + static_cast<size_t>(slot) * a->n * kStateFloats,
                          sizeof(float) * a->n * kStateFloats, hipMemcpyDeviceToHost));
    }
    const uint32_t copy = c < capacity ? c : capacity;
    if (contacts && copy) {
        DYN_HIP(hipMemcpy(contacts, a->d_contacts + static_cast<size_t>(slot) * a->cap * kContactFloats,
                          sizeof(float) * copy * kContactFloats, hipMemcpyDeviceToHost));
    }
    return DYN_OK;
}

int dyn_adjoint_read_bodies(dyn_adjoint* a, float* out) {
    using namespace dyn;
    if (!a || !out) return fail(DYN_E_INVALID, "dyn_adjoint_read_bodies: null argument");
    DYN_HIP(hipStreamSynchronize(a->stream));
    DYN_HIP(hipMemcpy(out, a->d_bodies, sizeof(float) * a->n * kBodyFloats, hipMemcpyDeviceToHost));
    return DYN_OK;
}

int dyn_adjoint_destroy(dyn_adjoint* a) {
    if (!a) return DYN_OK;
    (void)hipStreamSynchronize(a->stream);
    destroy(a);
    return DYN_OK;
}

}  // extern "C"

/*
 * adjoint_host.cpp - TEST INFRASTRUCTURE ONLY (see oracle.h): the frozen-contact rollout on the
 * host, for the differentiable step's checks. The product never links or loads it.
 *
 * It runs the replay and reverse templates of dynamis_amd/csrc/adjoint.h with a serial executor:
 *   adj_host_rollout   - the frozen-contact rollout in FP64 from given initial states, mass
 *                        scales and collider frictions, free or anchored to the recording
 *                        (central differences of its loss are the gradient oracle);
 *   adj_host_gradient  - the reverse of that rollout (FP64 or FP32), linearised about its own
 *                        FP64 trajectory or, anchored, about the recorded states (as the device).
 * The rollout relays each step's final impulses into the next step through the recorded relay
 * sources (contact_relay.wgsl); the tapes come from the CPU oracle or from the device recorder.
 */
#include <cstring>
#include <vector>

#include "../dynamis_amd/csrc/adjoint.h"

using namespace dyn_adj;

namespace {

struct HostExec {
    void sync() const {}
    template <class F>
    void each(uint32_t n, F&& f) const {
        for (uint32_t i = 0; i < n; ++i) f(i);
    }
    bool leader() const { return true; }
};

template <class T>
struct HostWork {
    std::vector<T> x, vel, pose, res, acc, rows, depth, target, slot, xb, resb, alb, sbar, mubar, pointb, sideb, accb, mucb;
    std::vector<uint32_t> list_start, list_count, list;
    Work<T> w{};
    HostWork(const AdjParams& P, uint32_t n, uint32_t k, uint32_t cap) {
        const size_t S = P.substeps, I = P.iterations, Q = P.position_iterations;
        x.assign(size_t(n) * kStateFloats, T(0));
        vel.assign(S * (3 + I) * n * 8, T(0));
        pose.assign(S * 2 * n * 8, T(0));
        res.assign(S * Q * n * 4, T(0));
        acc.assign(S * (I + 1) * cap * kAccFloats, T(0));
        rows.assign(S * cap * kRowFloats, T(0));
        depth.assign(size_t(cap) * 4, T(0));
        target.assign(size_t(cap) * 4, T(0));
        slot.assign(size_t(cap) * 2 * kSlotFloats, T(0));
        xb.assign(size_t(n) * kStateFloats, T(0));
        resb.assign(size_t(n) * 4, T(0));
        alb.assign(size_t(n) * 4, T(0));
        sbar.assign(n, T(0));
        mubar.assign(k ? k : 1, T(0));
        pointb.assign(size_t(cap) * 4 * kPointAdj, T(0));
        sideb.assign(size_t(cap) * 2 * kSideFloats, T(0));
        accb.assign(size_t(cap) * kAccFloats, T(0));
        mucb.assign(cap, T(0));
        list_start.assign(n, 0);
        list_count.assign(n, 0);
        list.assign(size_t(cap) * 2, 0);
        w.n = n;
        w.cap = cap;
        w.x = x.data();
        w.vel = vel.data();
        w.pose = pose.data();
        w.res = res.data();
        w.acc = acc.data();
        w.rows = rows.data();
        w.depth = depth.data();
        w.target = target.data();
        w.slot = slot.data();
        w.list_start = list_start.data();
        w.list_count = list_count.data();
        w.list = list.data();
        w.xb = xb.data();
        w.resb = resb.data();
        w.alb = alb.data();
        w.sbar = sbar.data();
        w.mubar = mubar.data();
        w.pointb = pointb.data();
        w.sideb = sideb.data();
        w.accb = accb.data();
        w.mucb = mucb.data();
    }
    const T* final_acc(const AdjParams& P) const {
        return acc.data() + (size_t(P.substeps - 1) * (P.iterations + 1) + P.iterations) * w.cap * kAccFloats;
    }
};

StepIn step_in(const float* bodies, const float* contacts, uint32_t count, const float* fwd, const AdjRange& r) {
    StepIn in;
    in.bodies = bodies;
    in.contacts = contacts;
    in.fwd = fwd;
    in.c0 = lower_collider(contacts, count, r.collider_base);
    in.c1 = lower_collider(contacts, count, r.collider_base + r.collider_count);
    return in;
}

// One step of every world in FP64 from states x (updated in place); fills `fin` with the step's
// final impulses.
void step_all(HostWork<double>& hw, const AdjParams& P, const AdjRange* ranges, uint32_t worlds, const float* bodies,
              const float* contacts, uint32_t count, const float* fwd, std::vector<double>& x, const double* s,
              const double* mu, const double* relayed, std::vector<double>& fin) {
    HostExec ex;
    for (uint32_t wi = 0; wi < worlds; ++wi) {
        const AdjRange& r = ranges[wi];
        const StepIn in = step_in(bodies, contacts, count, fwd, r);
        for (uint32_t row = r.row_base; row < r.row_base + r.row_count; ++row) {
            for (uint32_t j = 0; j < kStateFloats; ++j) hw.x[size_t(row) * kStateFloats + j] = x[size_t(row) * kStateFloats + j];
        }
        replay_step(ex, P, r, in, hw.w, s, mu, relayed);
        for (uint32_t row = r.row_base; row < r.row_base + r.row_count; ++row) {
            for (uint32_t j = 0; j < kStateFloats; ++j) x[size_t(row) * kStateFloats + j] = hw.x[size_t(row) * kStateFloats + j];
        }
        const double* f = hw.final_acc(P);
        for (uint32_t ci = in.c0; ci < in.c1; ++ci) {
            for (uint32_t j = 0; j < kAccFloats; ++j) fin[size_t(ci) * kAccFloats + j] = f[size_t(ci) * kAccFloats + j];
        }
    }
}

// The FP64 rollout; states[t] (t < steps) receive the start of each step, finals[t] each step's
// final impulses, when given.
//
// Anchored: every step is taken about the recorded forward instead, x_{t+1} = fwd_{t+1} +
// F_t(x_t) - F_t(fwd_t), and likewise the relayed impulses (recorded warm start plus the change of
// the previous step's final impulses). The unperturbed anchored rollout is the recording itself,
// and its derivative is the chain of each step's derivative at the recorded states: the gradient
// the device computes from its own recording.
void rollout(const AdjParams& P, const AdjRange* ranges, uint32_t worlds, uint32_t n, uint32_t k, uint32_t cap,
             const float* bodies, uint32_t steps, const float* const* contacts, const uint32_t* counts,
             const float* const* fwd, const double* x0, const double* s, const double* mu, double* out,
             std::vector<std::vector<double>>* states, std::vector<std::vector<double>>* finals, bool anchored) {
    HostWork<double> hw(P, n, k, cap);
    std::vector<double> x(x0, x0 + size_t(n) * kStateFloats);
    const size_t accs = size_t(cap) * kAccFloats;
    std::vector<double> prev(accs, 0.0), next(accs, 0.0), base_prev(accs, 0.0), base_next(accs, 0.0), relayed(accs, 0.0);
    for (uint32_t t = 0; t < steps; ++t) {
        if (states) (*states)[t] = x;
        if (!anchored) {
            step_all(hw, P, ranges, worlds, bodies, contacts[t], counts[t], fwd[t], x, s, mu,
                     t > 0 ? prev.data() : nullptr, next);
        } else {
            // relayed impulses: the recorded warm start moved by the previous step's change
            if (t > 0) {
                for (uint32_t ci = 0; ci < counts[t]; ++ci) {
                    const float* c = contacts[t] + size_t(kContactFloats) * ci;
                    for (uint32_t p = 0; p < contact_points(c); ++p) {
                        const uint32_t src = fbits(point_rec(c, p)[11]);
                        if (src == kNoSource) continue;
                        for (uint32_t j = 0; j < 3; ++j) {
                            const size_t at = size_t(kAccFloats) * (src >> 2) + 4 * j + (src & 3u);
                            relayed[at] = double(point_rec(c, p)[12 + j]) + prev[at] - base_prev[at];
                        }
                    }
                }
            }
            std::vector<double> base(fwd[t], fwd[t] + size_t(n) * kStateFloats);
            step_all(hw, P, ranges, worlds, bodies, contacts[t], counts[t], fwd[t], base, nullptr, nullptr, nullptr,
                     base_next);
            std::vector<double> moved = x;
            step_all(hw, P, ranges, worlds, bodies, contacts[t], counts[t], fwd[t], moved, s, mu,
                     t > 0 ? relayed.data() : nullptr, next);
            const float* target = t + 1 < steps ? fwd[t + 1] : nullptr;
            for (size_t i = 0; i < x.size(); ++i) {
                const double anchor = target ? double(target[i]) : base[i];
                x[i] = anchor + (moved[i] - base[i]);
            }
            base_prev.swap(base_next);
        }
        if (finals) (*finals)[t] = next;
        prev.swap(next);
    }
    std::memcpy(out, x.data(), sizeof(double) * x.size());
}

template <class T>
void gradient(const AdjParams& P, const AdjRange* ranges, uint32_t worlds, uint32_t n, uint32_t k, uint32_t cap,
              const float* bodies, uint32_t steps, const float* const* contacts, const uint32_t* counts,
              const float* const* fwd, const double* x0, const double* xbar, double* x0bar, double* sbar,
              double* mubar, bool anchored) {
    std::vector<std::vector<double>> states(steps), finals(steps);
    std::vector<double> end(size_t(n) * kStateFloats);
    if (!anchored) {
        rollout(P, ranges, worlds, n, k, cap, bodies, steps, contacts, counts, fwd, x0, nullptr, nullptr, end.data(),
                &states, &finals, false);
    }
    HostWork<T> hw(P, n, k, cap);
    for (size_t i = 0; i < hw.xb.size(); ++i) hw.xb[i] = T(xbar[i]);
    std::vector<T> accb_end(size_t(cap) * kAccFloats, T(0)), accb_start(size_t(cap) * kAccFloats, T(0));
    std::vector<T> prev_final(size_t(cap) * kAccFloats, T(0));
    HostExec ex;
    for (uint32_t t = steps; t-- > 0;) {
        std::vector<T> relayed;
        if (t > 0 && !anchored) relayed.assign(finals[t - 1].begin(), finals[t - 1].end());
        std::fill(prev_final.begin(), prev_final.end(), T(0));
        for (uint32_t wi = 0; wi < worlds; ++wi) {
            const AdjRange& r = ranges[wi];
            const StepIn in = step_in(bodies, contacts[t], counts[t], fwd[t], r);
            for (uint32_t row = r.row_base; row < r.row_base + r.row_count; ++row) {
                for (uint32_t j = 0; j < kStateFloats; ++j) {
                    const size_t at = size_t(row) * kStateFloats + j;
                    hw.x[at] = anchored ? T(fwd[t][at]) : T(states[t][at]);
                }
            }
            replay_step(ex, P, r, in, hw.w, static_cast<const T*>(nullptr), static_cast<const T*>(nullptr),
                        relayed.empty() ? static_cast<const T*>(nullptr) : relayed.data());
            reverse_step(ex, P, r, in, hw.w, accb_end.data(), accb_start.data());
            relay_back(ex, in, accb_start.data(), prev_final.data());
        }
        accb_end.swap(prev_final);
    }
    for (size_t i = 0; i < hw.xb.size(); ++i) x0bar[i] = double(hw.xb[i]);
    for (uint32_t i = 0; i < n; ++i) sbar[i] = double(hw.sbar[i]);
    for (uint32_t i = 0; i < k; ++i) mubar[i] = double(hw.mubar[i]);
}

}  // namespace

extern "C" {

int adj_host_rollout(const AdjParams* P, const AdjRange* ranges, uint32_t worlds, uint32_t n, uint32_t k, uint32_t cap,
                     const float* bodies, uint32_t steps, const float* const* contacts, const uint32_t* counts,
                     const float* const* fwd, const double* x0, const double* s, const double* mu, double* out,
                     int anchored) {
    rollout(*P, ranges, worlds, n, k, cap, bodies, steps, contacts, counts, fwd, x0, s, mu, out, nullptr, nullptr,
            anchored != 0);
    return 0;
}

int adj_host_gradient(const AdjParams* P, const AdjRange* ranges, uint32_t worlds, uint32_t n, uint32_t k,
                      uint32_t cap, const float* bodies, uint32_t steps, const float* const* contacts,
                      const uint32_t* counts, const float* const* fwd, const double* x0, const double* xbar,
                      double* x0bar, double* sbar, double* mubar, int single, int anchored) {
    if (single) {
        gradient<float>(*P, ranges, worlds, n, k, cap, bodies, steps, contacts, counts, fwd, x0, xbar, x0bar, sbar, mubar,
                        anchored != 0);
    } else {
        gradient<double>(*P, ranges, worlds, n, k, cap, bodies, steps, contacts, counts, fwd, x0, xbar, x0bar, sbar,
                         mubar, anchored != 0);
    }
    return 0;
}

uint32_t adj_host_sizes(uint32_t* out) {
    out[0] = sizeof(AdjParams);
    out[1] = sizeof(AdjRange);
    out[2] = kContactFloats;
    out[3] = kBodyFloats;
    return 4;
}

}  // extern "C"

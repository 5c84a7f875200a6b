// ref_mitigation_shim.cpp -- C entry points around the reference's header-only zero-noise
// extrapolation (src/Mitigation.h: ll::NoisyBackend, ll::ZNE) over its ll::ReferenceBackend
// (src/Backend.h:49-148), so tests can run the compiled reference from Python (ctypes).
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`) against the reference
// headers where they lie; the output goes to oracle/_ref/ and is never shipped or linked by the
// product.  This file holds no reference code: it replays a gate list on the reference's
// NoisyBackend and forwards ZNE::extrapolate, turning std::invalid_argument into -1 and any
// other exception into -2.
//
// Gate list: kinds 0 h, 1 x, 2 rx, 3 ry, 4 rz, 5 cnot (q[0] control, q[1] target),
// 6 mcz (q[0..nq)); 16 qubit slots per gate in qs; observable sum_i w[i] p_i.
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "Backend.h"
#include "Mitigation.h"

namespace {

struct Circuit {
    int G;
    const int* kind;
    const int* nq;
    const int* qs;
    const double* theta;
    void operator()(ll::Backend& b) const {
        for (int k = 0; k < G; ++k) {
            const int* q = qs + 16 * k;
            switch (kind[k]) {
                case 0: b.h(q[0]); break;
                case 1: b.x(q[0]); break;
                case 2: b.rx(q[0], theta[k]); break;
                case 3: b.ry(q[0], theta[k]); break;
                case 4: b.rz(q[0], theta[k]); break;
                case 5: b.cnot(q[0], q[1]); break;
                default: b.mcz(std::vector<int>(q, q + nq[k])); break;
            }
        }
    }
};

double diag_obs(const ll::Backend& b, const double* w) {
    const std::vector<double> p = b.probabilities();
    double e = 0.0;
    for (std::size_t i = 0; i < p.size(); ++i) e += w[i] * p[i];
    return e;
}

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument&) {
        return -1;
    } catch (...) {
        return -2;
    }
}

}  // namespace

extern "C" {

// Per-trajectory observable values at one scale: NoisyBackend(p, seed), set_scale(scale), then
// n_traj times reset(n) / circuit / observable, one continuing stream (ZNE::extrapolate's loop).
int refm_values(int n, int G, const int* kind, const int* nq, const int* qs, const double* theta, double p,
                uint64_t seed, int scale, int n_traj, const double* w, double* out) {
    return guard([&] {
        ll::ReferenceBackend inner(n);
        ll::NoisyBackend nb(inner, p, seed);
        nb.set_scale(scale);
        const Circuit c{G, kind, nq, qs, theta};
        for (int t = 0; t < n_traj; ++t) {
            nb.reset(n);
            c(nb);
            out[t] = diag_obs(nb, w);
        }
    });
}

// ZNE(trajectories).extrapolate(NoisyBackend(p, seed), circuit, observable, scales).
// out_scalar = {value, std_error, trajectories}; per-scale values, std_errors, weights.
int refm_extrapolate(int n, int G, const int* kind, const int* nq, const int* qs, const double* theta, double p,
                     uint64_t seed, int trajectories, int n_scales, const double* scales, const double* w,
                     double* out_scalar, double* values, double* std_errors, double* weights) {
    return guard([&] {
        ll::ReferenceBackend inner(n);
        ll::NoisyBackend nb(inner, p, seed);
        const Circuit c{G, kind, nq, qs, theta};
        const ll::ZNE zne(trajectories);
        const ll::ZNEResult r = zne.extrapolate(
            nb, [&](ll::Backend& b) { c(b); }, [&](const ll::Backend& b) { return diag_obs(b, w); },
            std::vector<double>(scales, scales + n_scales));
        out_scalar[0] = r.value;
        out_scalar[1] = r.std_error;
        out_scalar[2] = r.trajectories;
        for (int k = 0; k < n_scales; ++k) {
            values[k] = r.values[k];
            std_errors[k] = r.std_errors[k];
            weights[k] = r.weights[k];
        }
    });
}

// ZNE::richardson_weights (src/Mitigation.h:300-314).
int refm_richardson_weights(int n_scales, const double* scales, double* out) {
    return guard([&] {
        const std::vector<double> c = ll::ZNE::richardson_weights(std::vector<double>(scales, scales + n_scales));
        std::memcpy(out, c.data(), sizeof(double) * c.size());
    });
}

}  // extern "C"

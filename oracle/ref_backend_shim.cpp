// ref_backend_shim.cpp -- C entry points around the reference's own Eigen-free
// backends, so tests can drive the compiled reference from Python (ctypes).
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`) against
// the reference headers where they lie (src/Backend.h, src/StateVectorBackend.h,
// src/Algorithms.h); the output goes to oracle/_ref/ and is never shipped or
// linked by the product.  This file holds no reference code: it only forwards
// calls to ll::ReferenceBackend (src/Backend.h:49-148), ll::StateVectorBackend
// (src/StateVectorBackend.h:43-219) and ll::VQE::prepare (src/Algorithms.h:240-252),
// turning std::invalid_argument into a -1 return.
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "Algorithms.h"
#include "Backend.h"
#include "StateVectorBackend.h"

namespace {
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

// kind 0 = ReferenceBackend, 1 = StateVectorBackend.  Returns null on a bad size.
void* refb_create(int kind, int n) {
    try {
        if (kind == 0) return static_cast<ll::Backend*>(new ll::ReferenceBackend(n));
        return static_cast<ll::Backend*>(new ll::StateVectorBackend(n));
    } catch (...) {
        return nullptr;
    }
}
void refb_destroy(void* b) { delete static_cast<ll::Backend*>(b); }
int refb_reset(void* b, int n) { return guard([&] { static_cast<ll::Backend*>(b)->reset(n); }); }
int refb_n_qubits(void* b) { return static_cast<ll::Backend*>(b)->n_qubits(); }
int refb_h(void* b, int q) { return guard([&] { static_cast<ll::Backend*>(b)->h(q); }); }
int refb_x(void* b, int q) { return guard([&] { static_cast<ll::Backend*>(b)->x(q); }); }
int refb_rx(void* b, int q, double t) { return guard([&] { static_cast<ll::Backend*>(b)->rx(q, t); }); }
int refb_ry(void* b, int q, double t) { return guard([&] { static_cast<ll::Backend*>(b)->ry(q, t); }); }
int refb_rz(void* b, int q, double t) { return guard([&] { static_cast<ll::Backend*>(b)->rz(q, t); }); }
int refb_cnot(void* b, int c, int t) { return guard([&] { static_cast<ll::Backend*>(b)->cnot(c, t); }); }
int refb_mcz(void* b, const int* qs, int nq) {
    return guard([&] { static_cast<ll::Backend*>(b)->mcz(std::vector<int>(qs, qs + nq)); });
}
// Copies 2^n complex128 (interleaved re, im) into out.
int refb_state(void* b, double* out) {
    return guard([&] {
        const auto s = static_cast<ll::Backend*>(b)->state();
        std::memcpy(out, s.data(), s.size() * sizeof(std::complex<double>));
    });
}
int refb_probabilities(void* b, double* out) {
    return guard([&] {
        const auto p = static_cast<ll::Backend*>(b)->probabilities();
        std::memcpy(out, p.data(), p.size() * sizeof(double));
    });
}
// VQE ansatz: RY layers + CNOT ring, parameters theta[l * n + q].
int refb_vqe_prepare(void* b, int n, int layers, const double* theta) {
    return guard([&] {
        ll::VQE vqe(n, layers);
        vqe.prepare(*static_cast<ll::Backend*>(b),
                    std::vector<double>(theta, theta + static_cast<std::size_t>(n) * layers));
    });
}
}

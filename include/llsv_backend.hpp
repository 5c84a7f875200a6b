/* llsv_backend.hpp -- the reference's gate-level backend interface (ll::Backend,
 * src/Backend.h:25-45) implemented over the C ABI of llsv.h, header-only.
 *
 * `Base` is the abstract interface the adapter derives from.  In the reference tree that is
 * ll::Backend, so a maintainer adds (src/bindings_algorithms.cpp:46-54 shape):
 *
 *     #include "Backend.h"
 *     #include "llsv_backend.hpp"
 *     using GpuBackend = llsv::GpuBackend<ll::Backend>;
 *     m.def("make_gpu_backend", [](int n) { return std::unique_ptr<ll::Backend>(new GpuBackend(n)); },
 *           py::arg("n_qubits"), "MI355X state vector, up to 34 qubits.");
 *
 * and Grover, QAOA and VQE (src/Algorithms.h) run on the GPU unchanged.  Conventions as the
 * interface states: qubit q is bit q of the index, RX/RY/RZ(t) = exp(-i t P / 2), mcz negates
 * the amplitudes whose listed bits are all 1; invalid qubits, duplicate mcz qubits and
 * non-finite angles throw std::invalid_argument, device failures std::runtime_error.
 * No counterpart in the reference beyond the interface itself.
 */
#pragma once

#include <cmath>
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "llsv.h"

namespace llsv {

template <class Base>
class GpuBackend final : public Base {
public:
    static constexpr int MAX_QUBITS = LLSV_MAX_QUBITS;

    explicit GpuBackend(int n_qubits = 1) { reset(n_qubits); }
    ~GpuBackend() override {
        if (s_) llsv_destroy(s_);
    }
    GpuBackend(const GpuBackend&) = delete;
    GpuBackend& operator=(const GpuBackend&) = delete;

    void reset(int n_qubits) override {
        if (n_qubits < 1 || n_qubits > MAX_QUBITS)
            throw std::invalid_argument("GpuBackend::reset: n_qubits must be in [1, " + std::to_string(MAX_QUBITS) + "]");
        if (s_ && n_qubits == n_) {
            ok(llsv_reset(s_));
            return;
        }
        if (s_) llsv_destroy(s_);
        s_ = nullptr;
        n_ = 0;
        ok(llsv_create(n_qubits, &s_));
        n_ = n_qubits;
    }

    int n_qubits() const override { return n_; }

    void h(int q) override {
        const double r = 1.0 / std::sqrt(2.0);
        one(q, {r, 0.0, r, 0.0, r, 0.0, -r, 0.0});
    }
    void x(int q) override { one(q, {0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0}); }
    void rx(int q, double t) override {
        angle(t, "rx");
        const double c = std::cos(0.5 * t), s = std::sin(0.5 * t);
        one(q, {c, 0.0, 0.0, -s, 0.0, -s, c, 0.0});
    }
    void ry(int q, double t) override {
        angle(t, "ry");
        const double c = std::cos(0.5 * t), s = std::sin(0.5 * t);
        one(q, {c, 0.0, -s, 0.0, s, 0.0, c, 0.0});
    }
    void rz(int q, double t) override {
        angle(t, "rz");
        const double c = std::cos(0.5 * t), s = std::sin(0.5 * t);
        one(q, {c, -s, 0.0, 0.0, 0.0, 0.0, c, s});
    }
    void cnot(int control, int target) override { ok(llsv_apply_cnot(s_, control, target)); }
    void mcz(const std::vector<int>& qubits) override {
        if (qubits.empty()) throw std::invalid_argument("GpuBackend::mcz: empty qubit list");
        ok(llsv_apply_mcz(s_, qubits.data(), int(qubits.size())));
    }

    std::vector<double> probabilities() const override {
        std::vector<double> p(std::size_t(1) << n_);
        ok(llsv_probabilities(s_, p.data(), p.size()));
        return p;
    }
    std::vector<std::complex<double>> state() const override {
        std::vector<std::complex<double>> a(std::size_t(1) << n_);
        ok(llsv_copy_to_host(s_, reinterpret_cast<double*>(a.data()), a.size()));
        return a;
    }

    llsv_state* handle() const { return s_; }

private:
    llsv_state* s_ = nullptr;
    int n_ = 0;

    static void ok(int rc) {
        if (rc == LLSV_OK) return;
        if (rc == LLSV_ERR_ARG) throw std::invalid_argument(llsv_last_error());
        throw std::runtime_error(llsv_last_error());
    }
    static void angle(double t, const char* who) {
        if (!std::isfinite(t)) throw std::invalid_argument(std::string("GpuBackend::") + who + ": non-finite angle");
    }
    void one(int q, std::initializer_list<double> m) {
        double g[8];
        int i = 0;
        for (double v : m) g[i++] = v;
        ok(llsv_apply_1q(s_, q, g));
    }
};

}  // namespace llsv

// py_statevector.cpp -- the quantum_sim_v3 / quantum_sim_v2 pybind11 surface over libllsv.
//
// Compiled twice: -DLLSV_PYMOD=quantum_sim_v3 (method apply_gate_cache_optimized and
// apply_gate_standard, src/StateVectorOptimized.cpp:112-135) and -DLLSV_PYMOD=quantum_sim_v2
// -DLLSV_V2 (method apply_gate, src/bindings_v2.cpp:87-102).  Same class either way.
//
// Host/device coherence.  The state lives in HBM.  Two reference behaviours alias it with
// host memory: the zero-copy constructor from a NumPy array (src/StateVectorOptimized.cpp:26-47)
// and get_numpy_view() (src/StateVectorOptimized.cpp:99-106).  While such a host buffer is
// linked -- the borrowed array, or a view that is still alive -- every state-changing call
// uploads the host buffer first and writes the result back afterwards, so NumPy code that
// reads or writes the array between calls (benchmarks/benchmark_fusion.py:62-65,
// benchmarks/test_zerocopy_proof.py:30-33) sees exactly the reference's semantics.
// defer_host_sync(True) drops the per-call transfers; sync() then writes the host buffer.
// Unlike the reference, the view's base is the live instance, not a copy of it, and a
// borrowing instance keeps its array alive.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/llsv.h"

#ifndef LLSV_PYMOD
#define LLSV_PYMOD quantum_sim_v3
#endif

namespace py = pybind11;
using cplx = std::complex<double>;

namespace {

[[noreturn]] void raise_rc(int rc) {
    const std::string msg = llsv_last_error();
    if (rc == LLSV_ERR_ARG) throw py::value_error(msg);
    if (rc == LLSV_ERR_NOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

inline void check(int rc) {
    if (rc != LLSV_OK) raise_rc(rc);
}

struct ViewAnchor {
    py::object owner;  // keeps the StateVector alive while any view exists
};

class StateVector {
public:
    explicit StateVector(int n) {
        if (n < 1 || n > LLSV_MAX_QUBITS)
            throw py::value_error("StateVector: n must be in [1, " + std::to_string(LLSV_MAX_QUBITS) + "]");
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = llsv_create(n, &st_);
        }
        check(rc);
    }

    explicit StateVector(py::array arr) {
        if (arr.ndim() != 1) throw std::runtime_error("NumPy array must be 1-dimensional");
        // Same acceptance as py::array_t<complex<double>> with forcecast: anything else is
        // converted into a private complex128 copy (which the instance then keeps alive).
        auto a = py::array_t<cplx, py::array::c_style | py::array::forcecast>::ensure(arr);
        if (!a) throw std::runtime_error("NumPy array must be convertible to complex128");
        const uint64_t size = static_cast<uint64_t>(a.shape(0));
        if (size < 2 || (size & (size - 1)) != 0) throw std::runtime_error("Array size must be power of 2");
        borrowed_ = a;
        is_borrowed_ = true;
        host_ = reinterpret_cast<double*>(a.mutable_data());
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = llsv_create_from_host(host_, size, &st_);
            if (rc == LLSV_OK) rc = llsv_sync(st_);
        }
        check(rc);
    }

    ~StateVector() {
        if (st_) {
            py::gil_scoped_release nogil;
            llsv_destroy(st_);
        }
    }

    StateVector(const StateVector&) = delete;
    StateVector& operator=(const StateVector&) = delete;

    int num_qubits() const { return llsv_num_qubits(st_); }
    uint64_t size() const { return llsv_size(st_); }

    // ---- gates
    void apply_1q(int target, const py::list& gate) {
        if (py::len(gate) < 4) throw py::index_error("gate list needs 4 entries");
        double g[8];
        for (int i = 0; i < 4; ++i) {
            const cplx v = gate[i].cast<cplx>();
            g[2 * i] = v.real();
            g[2 * i + 1] = v.imag();
        }
        mutate([&] { return llsv_apply_1q(st_, target, g); });
    }

    void cnot(int c, int t) {
        mutate([&] { return llsv_apply_cnot(st_, c, t); });
    }
    void cz(int a, int b) {
        mutate([&] { return llsv_apply_cz(st_, a, b); });
    }
    void mcz(const std::vector<int>& qs) {
        mutate([&] { return llsv_apply_mcz(st_, qs.data(), static_cast<int>(qs.size())); });
    }

    void apply_fused(const std::vector<int>& targets, py::array matrix) {
        const int k = static_cast<int>(targets.size());
        auto m = py::array_t<cplx, py::array::c_style | py::array::forcecast>::ensure(matrix);
        if (!m) throw py::value_error("apply_fused: matrix must be convertible to complex128");
        const py::ssize_t D = py::ssize_t(1) << (k > 0 && k < 16 ? k : 0);
        if (k < 1 || m.size() != D * D)
            throw py::value_error("apply_fused: matrix must be 2^k x 2^k for k = len(targets)");
        const double* mp = reinterpret_cast<const double*>(m.data());
        mutate([&] { return llsv_apply_dense(st_, k, targets.data(), mp); });
    }

    // Parses a Python op list (see apply_program); keep holds the dense matrices it points into.
    static std::vector<llsv_op> parse_ops(const py::list& ops,
                                          std::vector<py::array_t<cplx, py::array::c_style | py::array::forcecast>>& keep) {
        std::vector<llsv_op> prog;
        prog.reserve(py::len(ops));
        for (const py::handle& h : ops) prog.push_back(parse_op(py::reinterpret_borrow<py::tuple>(h), keep));
        return prog;
    }

    void apply_program(const py::list& ops) {
        std::vector<llsv_op> prog;
        std::vector<py::array_t<cplx, py::array::c_style | py::array::forcecast>> keep;
        prog.reserve(py::len(ops));
        for (const py::handle& h : ops) prog.push_back(parse_op(py::reinterpret_borrow<py::tuple>(h), keep));
        mutate([&] { return llsv_apply_program(st_, prog.data(), prog.size()); });
    }

    void reset() {
        mutate([&] { return llsv_reset(st_); }, /*upload=*/false);
    }

    // ---- readout
    double norm() {
        double v = 0.0;
        read([&] { return llsv_norm2(st_, &v); });
        return v;
    }
    double expect_z(int q) {
        double v = 0.0;
        read([&] { return llsv_expect_z(st_, q, &v); });
        return v;
    }
    // <psi|H|psi> for H = sum of (coeff, xmask, zmask) Pauli terms
    double expect_pauli(const std::vector<std::tuple<double, uint64_t, uint64_t>>& terms) {
        std::vector<llsv_pauli> t(terms.size());
        for (size_t i = 0; i < terms.size(); ++i) t[i] = llsv_pauli{std::get<0>(terms[i]), std::get<1>(terms[i]), std::get<2>(terms[i])};
        double total = 0.0;
        read([&] { return llsv_expect_pauli(st_, t.data(), int(t.size()), nullptr, &total); });
        return total;
    }

    py::array_t<uint64_t> sample(uint64_t shots, uint64_t seed) {
        py::array_t<uint64_t> out(static_cast<py::ssize_t>(shots));
        uint64_t* p = out.mutable_data();
        read([&] { return llsv_sample(st_, shots, seed, p); });
        return out;
    }

    py::array_t<double> probabilities() {
        py::array_t<double> out(static_cast<py::ssize_t>(size()));
        double* p = out.mutable_data();
        const uint64_t n = size();
        read([&] { return llsv_probabilities(st_, p, n); });
        return out;
    }
    py::array_t<cplx> state() {
        py::array_t<cplx> out(static_cast<py::ssize_t>(size()));
        double* p = reinterpret_cast<double*>(out.mutable_data());
        const uint64_t n = size();
        read([&] { return llsv_copy_to_host(st_, p, n); });
        return out;
    }

    py::array get_numpy_view(py::object self) {
        if (!host_) {
            mirror_ = py::array_t<cplx>(static_cast<py::ssize_t>(size()));
            host_ = reinterpret_cast<double*>(mirror_.mutable_data());
        }
        py::object anchor = anchor_alive();
        if (anchor.is_none()) {
            anchor = py::cast(ViewAnchor{self});
            anchor_ = py::weakref(anchor);
            if (!is_borrowed_) pull();  // a fresh view of an owned state starts from the device copy
        }
        return py::array(py::dtype::of<cplx>(), {static_cast<py::ssize_t>(size())},
                         {static_cast<py::ssize_t>(sizeof(cplx))}, host_, anchor);
    }

    void sync() {
        if (host_linked()) pull();
        else device_call([&] { return llsv_sync(st_); });
    }
    void defer_host_sync(bool on) {
        if (!on && deferred_ && host_linked()) pull();  // bring the host buffer up to date
        deferred_ = on;
    }
    bool host_coherent() { return host_linked() && !deferred_; }

    // ---- measurement helpers
    void timer_start() { device_call([&] { return llsv_timer_start(st_); }); }
    float timer_stop() {
        float ms = 0.f;
        device_call([&] { return llsv_timer_stop(st_, &ms); });
        return ms;
    }
    int tile_bits() {
        const int b = llsv_program_tile_bits(st_);
        if (b < 0) raise_rc(b);
        return b;
    }
    void mark(int slot) { device_call([&] { return llsv_mark(st_, slot); }); }
    float mark_elapsed(int a, int b) {
        float ms = 0.f;
        device_call([&] { return llsv_mark_elapsed(st_, a, b, &ms); });
        return ms;
    }
    void copy_from(StateVector& other) {
        other.refresh_device();
        mutate([&] { return llsv_copy_state(st_, other.st_); }, /*upload=*/false);
    }
    uintptr_t device_ptr() { return reinterpret_cast<uintptr_t>(llsv_device_ptr(st_)); }
    uintptr_t stream() { return reinterpret_cast<uintptr_t>(llsv_stream(st_)); }

private:
    llsv_state* st_ = nullptr;
    py::array borrowed_;  // keeps the borrowed array alive (a default py::array is an empty array, not null)
    bool is_borrowed_ = false;
    py::array_t<cplx> mirror_;
    double* host_ = nullptr;
    py::object anchor_ = py::none();
    bool deferred_ = false;

    py::object anchor_alive() {
        if (anchor_.is_none()) return py::none();
        py::object o = py::reinterpret_borrow<py::weakref>(anchor_)();
        return o;
    }

    bool host_linked() {
        if (!host_) return false;
        if (is_borrowed_) return true;
        return !anchor_alive().is_none();
    }

    template <class F>
    void device_call(F&& f) {
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = f();
        }
        check(rc);
    }

    // A call that only reads the state: while a host buffer is linked (and transfers are not
    // deferred) the host copy is the current one -- NumPy may have written it since the last
    // call -- so it goes to the device first.
    template <class F>
    void read(F&& f) {
        const bool linked = host_linked() && !deferred_;
        double* h = host_;
        const uint64_t n = size();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = linked ? llsv_copy_from_host(st_, h, n) : LLSV_OK;
            if (rc == LLSV_OK) rc = f();
        }
        check(rc);
    }
    void refresh_device() {
        if (host_linked() && !deferred_) push();
    }

    void push() {
        double* h = host_;
        const uint64_t n = size();
        device_call([&] { return llsv_copy_from_host(st_, h, n); });
    }
    void pull() {
        double* h = host_;
        const uint64_t n = size();
        device_call([&] { return llsv_copy_to_host(st_, h, n); });
    }

    template <class F>
    void mutate(F&& f, bool upload = true) {
        const bool linked = host_linked() && !deferred_;
        double* h = host_;
        const uint64_t n = size();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = (linked && upload) ? llsv_copy_from_host(st_, h, n) : LLSV_OK;
            if (rc == LLSV_OK) rc = f();
            if (rc == LLSV_OK && linked) rc = llsv_copy_to_host(st_, h, n);
        }
        check(rc);
    }

    static llsv_op parse_op(const py::tuple& t,
                            std::vector<py::array_t<cplx, py::array::c_style | py::array::forcecast>>& keep) {
        if (t.size() < 2) throw py::value_error("op tuple needs a kind and qubits");
        const std::string kind = t[0].cast<std::string>();
        llsv_op op{};
        auto one_q = [&](const py::list& g) {
            op.kind = LLSV_OP_1Q;
            op.nq = 1;
            op.q[0] = t[1].cast<int>();
            for (int i = 0; i < 4; ++i) {
                const cplx v = g[i].cast<cplx>();
                op.m[2 * i] = v.real();
                op.m[2 * i + 1] = v.imag();
            }
        };
        const double r = 1.0 / std::sqrt(2.0);
        if (kind == "u") {
            one_q(py::list(t[2]));
        } else if (kind == "h") {
            one_q(py::cast(std::vector<cplx>{r, r, r, -r}));
        } else if (kind == "x") {
            one_q(py::cast(std::vector<cplx>{0.0, 1.0, 1.0, 0.0}));
        } else if (kind == "rx" || kind == "ry" || kind == "rz") {
            const double th = t[2].cast<double>();
            if (!std::isfinite(th)) throw py::value_error(kind + ": non-finite angle");
            const double c = std::cos(0.5 * th), s = std::sin(0.5 * th);
            std::vector<cplx> g;
            if (kind == "rx") g = {c, cplx(0, -s), cplx(0, -s), c};
            else if (kind == "ry") g = {c, -s, s, c};
            else g = {std::polar(1.0, -0.5 * th), 0.0, 0.0, std::polar(1.0, 0.5 * th)};
            one_q(py::cast(g));
        } else if (kind == "cnot" || kind == "cz") {
            op.kind = kind == "cnot" ? LLSV_OP_CNOT : LLSV_OP_CZ;
            op.nq = 2;
            op.q[0] = t[1].cast<int>();
            op.q[1] = t[2].cast<int>();
        } else if (kind == "mcz" || kind == "dense") {
            const auto qs = t[1].cast<std::vector<int>>();
            if (qs.empty() || qs.size() > 8) throw py::value_error(kind + ": needs 1..8 qubits");
            op.kind = kind == "mcz" ? LLSV_OP_MCZ : LLSV_OP_DENSE;
            op.nq = static_cast<int32_t>(qs.size());
            for (size_t j = 0; j < qs.size(); ++j) op.q[j] = qs[j];
            if (kind == "dense") {
                auto m = py::array_t<cplx, py::array::c_style | py::array::forcecast>::ensure(t[2]);
                const py::ssize_t D = py::ssize_t(1) << qs.size();
                if (!m || m.size() != D * D) throw py::value_error("dense: matrix must be 2^k x 2^k");
                op.dense = reinterpret_cast<const double*>(m.data());
                keep.push_back(m);
            }
        } else {
            throw py::value_error("unknown op kind '" + kind + "'");
        }
        return op;
    }
};

}  // namespace

#define LLSV_STR2(x) #x
#define LLSV_STR(x) LLSV_STR2(x)

PYBIND11_MODULE(LLSV_PYMOD, m) {
    m.doc() = "MI355X (gfx950) complex128 state-vector simulator: the " LLSV_STR(LLSV_PYMOD) " surface over libllsv";

    py::class_<ViewAnchor>(m, "_ViewAnchor", py::module_local()).def_property_readonly("owner", [](const ViewAnchor& a) { return a.owner; });

    py::class_<StateVector> cls(m, "StateVector", py::module_local());
    cls.def(py::init<int>(), py::arg("n"))
        .def(py::init<py::array>(), py::arg("array"))
#ifdef LLSV_V2
        .def("apply_gate", &StateVector::apply_1q, py::arg("target"), py::arg("gate"))
#else
        .def("apply_gate_cache_optimized", &StateVector::apply_1q, py::arg("target"), py::arg("gate"))
        .def("apply_gate_standard", &StateVector::apply_1q, py::arg("target"), py::arg("gate"))
#endif
        .def("get_numpy_view", [](py::object self) { return self.cast<StateVector&>().get_numpy_view(self); })
        .def("get_num_qubits", &StateVector::num_qubits)
        .def("get_size", &StateVector::size)
        // extensions beyond the reference surface
        .def("cnot", &StateVector::cnot, py::arg("control"), py::arg("target"))
        .def("cz", &StateVector::cz, py::arg("a"), py::arg("b"))
        .def("mcz", &StateVector::mcz, py::arg("qubits"))
        .def("apply_fused", &StateVector::apply_fused, py::arg("targets"), py::arg("matrix"))
        .def("apply_program", &StateVector::apply_program, py::arg("ops"))
        .def("reset", &StateVector::reset)
        .def("norm", &StateVector::norm)
        .def("expect_z", &StateVector::expect_z, py::arg("q"))
        .def("probabilities", &StateVector::probabilities)
        .def("expect_pauli", &StateVector::expect_pauli, py::arg("terms"))
        .def("sample", &StateVector::sample, py::arg("shots"), py::arg("seed"))
        .def("state", &StateVector::state)
        .def("sync", &StateVector::sync)
        .def("defer_host_sync", &StateVector::defer_host_sync, py::arg("on") = true)
        .def("host_coherent", &StateVector::host_coherent)
        .def("timer_start", &StateVector::timer_start)
        .def("timer_stop", &StateVector::timer_stop)
        .def("program_tile_bits", &StateVector::tile_bits)
        .def("mark", &StateVector::mark, py::arg("slot"))
        .def("mark_elapsed", &StateVector::mark_elapsed, py::arg("a"), py::arg("b"))
        .def("copy_from", &StateVector::copy_from, py::arg("other"))
        .def("device_ptr", &StateVector::device_ptr)
        .def("stream", &StateVector::stream);
    // the pass compiler of apply_program (process-wide)
    m.def("set_program_jit", [](int mode) { const int r = llsv_set_program_jit(mode); check(r < 0 ? r : 0); return r; },
          py::arg("mode"));
    m.def("program_jit_stats", [] {
        uint64_t c = 0, l = 0, f = 0;
        llsv_program_jit_stats(&c, &l, &f);
        return py::make_tuple(c, l, f);
    });
    m.def("program_jit_wait", [] {
        py::gil_scoped_release nogil;
        llsv_program_jit_wait();
    });
    m.def("program_jit_precompile", [](int n, const py::list& ops) {
        std::vector<py::array_t<cplx, py::array::c_style | py::array::forcecast>> keep;
        const std::vector<llsv_op> prog = StateVector::parse_ops(ops, keep);
        const int r = llsv_program_jit_precompile(n, prog.data(), prog.size());
        check(r < 0 ? r : 0);
        return r;
    }, py::arg("n"), py::arg("ops"));
    m.attr("__backend__") = llsv_version();
}

// Runs Grover search through llsv::GpuBackend (include/llsv_backend.hpp) and prints one JSON
// line per case.  The interface below restates the reference's ll::Backend declarations
// (src/Backend.h:25-45) so the adapter is exercised exactly as the reference's algorithms
// would call it; the Grover loop follows src/Algorithms.h:86-125 (phase flip by X conjugation
// around an all-qubit MCZ, inversion about the mean).
#include <algorithm>
#include <complex>
#include <cstdio>
#include <stdexcept>
#include <vector>

namespace iface {
class Backend {
public:
    virtual ~Backend() = default;
    virtual void reset(int n_qubits) = 0;
    virtual int n_qubits() const = 0;
    virtual void h(int q) = 0;
    virtual void x(int q) = 0;
    virtual void rx(int q, double theta) = 0;
    virtual void ry(int q, double theta) = 0;
    virtual void rz(int q, double theta) = 0;
    virtual void cnot(int control, int target) = 0;
    virtual void mcz(const std::vector<int>& qubits) = 0;
    virtual std::vector<double> probabilities() const = 0;
    virtual std::vector<std::complex<double>> state() const = 0;
};
}  // namespace iface

#include "llsv_backend.hpp"

static void grover(iface::Backend& b, int n, int marked, int iterations) {
    std::vector<int> all(n);
    for (int q = 0; q < n; ++q) all[q] = q;
    b.reset(n);
    for (int q = 0; q < n; ++q) b.h(q);
    for (int it = 0; it < iterations; ++it) {
        for (int q = 0; q < n; ++q)
            if (!((marked >> q) & 1)) b.x(q);
        b.mcz(all);
        for (int q = 0; q < n; ++q)
            if (!((marked >> q) & 1)) b.x(q);
        for (int q = 0; q < n; ++q) {
            b.h(q);
            b.x(q);
        }
        b.mcz(all);
        for (int q = 0; q < n; ++q) {
            b.x(q);
            b.h(q);
        }
    }
}

int main() {
    llsv::GpuBackend<iface::Backend> gpu(2);
    iface::Backend& b = gpu;
    for (int n : {10, 12, 14, 16, 18}) {
        int marked = 0;
        for (int q = 0; q < n; q += 2) marked |= 1 << q;  // 0b...0101
        grover(b, n, marked, 10);
        const std::vector<double> p = b.probabilities();
        const int arg = int(std::max_element(p.begin(), p.end()) - p.begin());
        double tot = 0.0;
        for (double v : p) tot += v;
        std::printf("{\"n\": %d, \"marked\": %d, \"p_success\": %.17g, \"argmax\": %d, \"p_total\": %.17g}\n", n, marked,
                    p[marked], arg, tot);
    }
    // the adapter's error semantics: invalid_argument for a bad qubit, a duplicate MCZ qubit, a
    // non-finite angle
    int caught = 0;
    try { b.h(99); } catch (const std::invalid_argument&) { ++caught; }
    try { b.mcz({1, 1}); } catch (const std::invalid_argument&) { ++caught; }
    try { b.rx(0, 1.0 / 0.0); } catch (const std::invalid_argument&) { ++caught; }
    std::printf("{\"errors_caught\": %d}\n", caught);
    return caught == 3 ? 0 : 1;
}

// sv_api.hip -- the C ABI declared in include/llsv.h.
//
// A state owns one device allocation of 2^n complex128 amplitudes (256-byte aligned by
// hipMalloc), one non-blocking HIP stream at the default priority, a small device
// parameter area (dense matrices, offset tables), a partial-sum buffer for the
// reductions and a pinned host word for scalar results.  Gate calls only enqueue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/llsv.h"
#include "sv_common.h"
#include "sv_dense.h"
#include "sv_kernels.h"
#include "sv_program.h"

#include "sv_state.h"

namespace llsv {

static thread_local std::string g_last_error;

// Allocations left before the injected failure (-1: none armed).
static std::atomic<long> g_fail_after{-1};

static bool injected_failure() {
    long v = g_fail_after.load();
    while (v >= 0)
        if (g_fail_after.compare_exchange_weak(v, v - 1)) return v == 0;
    return false;
}

hipError_t dev_alloc(void** p, size_t bytes) {
    if (injected_failure()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

hipError_t dev_host_alloc(void** p, size_t bytes) {
    if (injected_failure()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipHostMalloc(p, bytes, hipHostMallocDefault);
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define LLSV_HIP(expr)                                                                                  \
    do {                                                                                                \
        const hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(hip_code(e_), "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define LLSV_HANDLE(s) \
    if (!(s)) return fail(LLSV_ERR_HANDLE, "%s: null state handle", __func__)

static int check_qubit(const llsv_state* s, int q, const char* who) {
    if (q < 0 || q >= s->n) return fail(LLSV_ERR_ARG, "%s: qubit %d outside [0, %d)", who, q, s->n);
    return LLSV_OK;
}

static constexpr size_t kParamBytes = 1 << 20;
static constexpr uint64_t kProbChunk = uint64_t(1) << 20;  // doubles per llsv_probabilities step (8 MiB)
static constexpr int kStageSlots = 8;
// one llsv_apply_dense upload: the 32 x 32 complex matrix and 32 offsets at most
static constexpr size_t kStageBytes = 32 * 32 * 16 + 32 * 8;

static void free_state(llsv_state* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    delete s->program;
    if (s->amps) (void)hipFree(s->amps);
    if (s->partials) (void)hipFree(s->partials);
    if (s->result_dev) (void)hipFree(s->result_dev);
    if (s->params) (void)hipFree(s->params);
    if (s->prob_chunk) (void)hipFree(s->prob_chunk);
    if (s->stage) (void)hipHostFree(s->stage);
    for (hipEvent_t& e : s->stage_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->result_host) (void)hipHostFree(s->result_host);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    for (hipEvent_t& e : s->marks)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

static int alloc_state(int n, llsv_state** out) {
    if (!out) return fail(LLSV_ERR_ARG, "llsv_create: out is null");
    *out = nullptr;
    if (n < 1 || n > LLSV_MAX_QUBITS)
        return fail(LLSV_ERR_ARG, "llsv_create: n_qubits must be in [1, %d], got %d", LLSV_MAX_QUBITS, n);
    llsv_state* s = new (std::nothrow) llsv_state;
    if (!s) return fail(LLSV_ERR_NOMEM, "llsv_create: host allocation failed");
    s->n = n;
    s->size = uint64_t(1) << n;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free_state(s);
        return fail(LLSV_ERR_DEVICE, "llsv_create: %s", hipGetErrorString(e));
    }
    e = dev_alloc(reinterpret_cast<void**>(&s->amps), s->size * sizeof(v2d));
    if (e != hipSuccess) {
        free_state(s);
        (void)hipGetLastError();
        return fail(LLSV_ERR_NOMEM, "llsv_create: cannot allocate %llu bytes of device memory for %d qubits",
                    static_cast<unsigned long long>(uint64_t(1) << n) * 16ull, n);
    }
    if ((e = dev_alloc(reinterpret_cast<void**>(&s->partials), 2 * kRedBlocks * sizeof(double))) != hipSuccess ||
        (e = dev_alloc(reinterpret_cast<void**>(&s->result_dev), 4 * sizeof(double))) != hipSuccess ||
        (e = dev_alloc(reinterpret_cast<void**>(&s->params), kParamBytes)) != hipSuccess ||
        (e = dev_host_alloc(reinterpret_cast<void**>(&s->result_host), 4 * sizeof(double))) != hipSuccess ||
        (e = hipEventCreate(&s->ev0)) != hipSuccess || (e = hipEventCreate(&s->ev1)) != hipSuccess) {
        free_state(s);
        (void)hipGetLastError();
        return fail(LLSV_ERR_NOMEM, "llsv_create: %s", hipGetErrorString(e));
    }
    s->params_bytes = kParamBytes;
    s->program = new (std::nothrow) ProgramCache;
    if (!s->program) {
        free_state(s);
        return fail(LLSV_ERR_NOMEM, "llsv_create: host allocation failed");
    }
    *out = s;
    return LLSV_OK;
}

}  // namespace llsv

using namespace llsv;

extern "C" {

const char* llsv_last_error(void) { return g_last_error.c_str(); }

const char* llsv_version(void) { return "llsv 0.1 gfx950"; }

long llsv_test_fail_allocation(long after) { return g_fail_after.exchange(after < 0 ? -1 : after); }

int llsv_create(int n_qubits, llsv_state** out) {
    int rc = alloc_state(n_qubits, out);
    if (rc != LLSV_OK) return rc;
    const hipError_t e = launch_fill_basis((*out)->stream, (*out)->amps, n_qubits, 0);
    if (e != hipSuccess) {
        free_state(*out);
        *out = nullptr;
        return fail(LLSV_ERR_DEVICE, "llsv_create: %s", hipGetErrorString(e));
    }
    return LLSV_OK;
}

int llsv_create_from_host(const double* amps, uint64_t n_amps, llsv_state** out) {
    if (!amps) return fail(LLSV_ERR_ARG, "llsv_create_from_host: amps is null");
    if (n_amps < 2 || (n_amps & (n_amps - 1)) != 0)
        return fail(LLSV_ERR_ARG, "Array size must be power of 2");
    int n = 0;
    while ((uint64_t(1) << n) < n_amps) ++n;
    int rc = alloc_state(n, out);
    if (rc != LLSV_OK) return rc;
    const hipError_t e = hipMemcpyAsync((*out)->amps, amps, n_amps * sizeof(v2d), hipMemcpyHostToDevice, (*out)->stream);
    if (e != hipSuccess) {
        free_state(*out);
        *out = nullptr;
        return fail(LLSV_ERR_DEVICE, "llsv_create_from_host: %s", hipGetErrorString(e));
    }
    return LLSV_OK;
}

int llsv_destroy(llsv_state* s) {
    free_state(s);
    return LLSV_OK;
}

int llsv_reset(llsv_state* s) {
    LLSV_HANDLE(s);
    LLSV_HIP(launch_fill_basis(s->stream, s->amps, s->n, 0));
    return LLSV_OK;
}

int llsv_num_qubits(const llsv_state* s) { return s ? s->n : -1; }
uint64_t llsv_size(const llsv_state* s) { return s ? s->size : 0; }

int llsv_apply_1q(llsv_state* s, int target, const double g[8]) {
    LLSV_HANDLE(s);
    if (!g) return fail(LLSV_ERR_ARG, "llsv_apply_1q: gate is null");
    if (int rc = check_qubit(s, target, "llsv_apply_1q")) return rc;
    Gate2 gate;
    std::memcpy(gate.r, g, sizeof gate.r);
    LLSV_HIP(launch_apply_1q(s->stream, s->amps, s->n, target, gate));
    return LLSV_OK;
}

int llsv_apply_cnot(llsv_state* s, int control, int target) {
    LLSV_HANDLE(s);
    if (int rc = check_qubit(s, control, "llsv_apply_cnot")) return rc;
    if (int rc = check_qubit(s, target, "llsv_apply_cnot")) return rc;
    if (control == target) return fail(LLSV_ERR_ARG, "llsv_apply_cnot: control == target");
    LLSV_HIP(launch_cnot(s->stream, s->amps, s->n, control, target));
    return LLSV_OK;
}

static int make_bitlist(const llsv_state* s, const int* qubits, int nq, const char* who, BitList* out) {
    if (!qubits || nq <= 0) return fail(LLSV_ERR_ARG, "%s: empty qubit list", who);
    if (nq > s->n) return fail(LLSV_ERR_ARG, "%s: more qubits than the state has", who);
    BitList b{};
    b.m = nq;
    for (int j = 0; j < nq; ++j) {
        if (int rc = check_qubit(s, qubits[j], who)) return rc;
        const uint64_t bit = uint64_t(1) << qubits[j];
        if (b.mask & bit) return fail(LLSV_ERR_ARG, "%s: duplicate qubit %d", who, qubits[j]);
        b.mask |= bit;
        b.pos[j] = qubits[j];
    }
    std::sort(b.pos, b.pos + nq);
    *out = b;
    return LLSV_OK;
}

int llsv_apply_mcz(llsv_state* s, const int* qubits, int nq) {
    LLSV_HANDLE(s);
    BitList b;
    if (int rc = make_bitlist(s, qubits, nq, "llsv_apply_mcz", &b)) return rc;
    LLSV_HIP(launch_mcz(s->stream, s->amps, s->n, b));
    return LLSV_OK;
}

int llsv_apply_cz(llsv_state* s, int a, int b) {
    const int qs[2] = {a, b};
    return llsv_apply_mcz(s, qs, 2);
}

int llsv_apply_dense(llsv_state* s, int k, const int* targets, const double* m) {
    LLSV_HANDLE(s);
    if (k < 1 || k > LLSV_MAX_DENSE_K)
        return fail(LLSV_ERR_ARG, "llsv_apply_dense: k must be in [1, %d], got %d", LLSV_MAX_DENSE_K, k);
    if (!m) return fail(LLSV_ERR_ARG, "llsv_apply_dense: matrix is null");
    BitList b;
    if (int rc = make_bitlist(s, targets, k, "llsv_apply_dense", &b)) return rc;
    const int D = 1 << k;
    // params layout: [U: D*D v2d][off: D uint64]
    uint64_t off[32];
    for (int c = 0; c < D; ++c) {
        uint64_t o = 0;
        for (int j = 0; j < k; ++j)
            if ((c >> j) & 1) o |= uint64_t(1) << targets[j];
        off[c] = o;
    }
    const size_t ubytes = size_t(D) * D * sizeof(v2d);
    // Upload through a pinned slot: the copy reads it after this call returns, so a slot is
    // rewritten only once the copy that last read it has completed.
    if (!s->stage) {
        if (dev_host_alloc(reinterpret_cast<void**>(&s->stage), kStageSlots * kStageBytes) != hipSuccess) {
            s->stage = nullptr;
            (void)hipGetLastError();
            return fail(LLSV_ERR_NOMEM, "llsv_apply_dense: cannot allocate %zu bytes of pinned staging",
                        kStageSlots * kStageBytes);
        }
    }
    const int slot = s->stage_next;
    s->stage_next = (slot + 1) % kStageSlots;
    hipEvent_t& ev = s->stage_ev[slot];
    if (ev)
        LLSV_HIP(hipEventSynchronize(ev));
    else
        LLSV_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    unsigned char* stage = s->stage + size_t(slot) * kStageBytes;
    std::memcpy(stage, m, ubytes);
    std::memcpy(stage + ubytes, off, D * sizeof(uint64_t));
    LLSV_HIP(hipMemcpyAsync(s->params, stage, ubytes + D * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream));
    LLSV_HIP(hipEventRecord(ev, s->stream));
    PassDesc pd;
    DenseTile dt;
    if (make_dense_tile(s->n, k, targets, &pd, &dt))
        LLSV_HIP(launch_dense_mfma(s->stream, s->amps, pd, dt, reinterpret_cast<const v2d*>(s->params)));
    else
        LLSV_HIP(launch_dense(s->stream, s->amps, s->n, k, b, reinterpret_cast<const v2d*>(s->params),
                              reinterpret_cast<const uint64_t*>(s->params + ubytes)));
    return LLSV_OK;
}

int llsv_apply_program(llsv_state* s, const llsv_op* ops, uint64_t n_ops) {
    LLSV_HANDLE(s);
    if (n_ops == 0) return LLSV_OK;
    if (!ops) return fail(LLSV_ERR_ARG, "llsv_apply_program: ops is null");
    return run_program(s, ops, n_ops);
}

int llsv_norm2(llsv_state* s, double* out) {
    LLSV_HANDLE(s);
    if (!out) return fail(LLSV_ERR_ARG, "llsv_norm2: out is null");
    LLSV_HIP(launch_reduce_prob(s->stream, s->amps, s->n, 0, 0, s->partials, s->result_dev));
    LLSV_HIP(hipMemcpyAsync(s->result_host, s->result_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    LLSV_HIP(hipStreamSynchronize(s->stream));
    *out = s->result_host[0];
    return LLSV_OK;
}

int llsv_expect_z(llsv_state* s, int q, double* out) {
    LLSV_HANDLE(s);
    if (!out) return fail(LLSV_ERR_ARG, "llsv_expect_z: out is null");
    if (int rc = check_qubit(s, q, "llsv_expect_z")) return rc;
    LLSV_HIP(launch_reduce_prob(s->stream, s->amps, s->n, 1, q, s->partials, s->result_dev));
    LLSV_HIP(hipMemcpyAsync(s->result_host, s->result_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    LLSV_HIP(hipStreamSynchronize(s->stream));
    *out = s->result_host[0];
    return LLSV_OK;
}

int llsv_probabilities(llsv_state* s, double* out, uint64_t n) {
    LLSV_HANDLE(s);
    if (!out) return fail(LLSV_ERR_ARG, "llsv_probabilities: out is null");
    if (n != s->size) return fail(LLSV_ERR_ARG, "llsv_probabilities: buffer holds %llu values, state has %llu",
                                  (unsigned long long)n, (unsigned long long)s->size);
    // |psi|^2 in chunks through one device buffer kept by the state (no allocation per call)
    if (!s->prob_chunk) {
        if (dev_alloc(reinterpret_cast<void**>(&s->prob_chunk), kProbChunk * sizeof(double)) != hipSuccess) {
            s->prob_chunk = nullptr;
            (void)hipGetLastError();
            return fail(LLSV_ERR_NOMEM, "llsv_probabilities: cannot allocate %llu bytes",
                        (unsigned long long)(kProbChunk * sizeof(double)));
        }
    }
    hipError_t e = hipSuccess;
    for (uint64_t off = 0; off < n && e == hipSuccess; off += kProbChunk) {
        const uint64_t cnt = std::min(kProbChunk, n - off);
        e = launch_probabilities(s->stream, s->amps + off, cnt, s->prob_chunk);
        if (e == hipSuccess) e = hipMemcpyAsync(out + off, s->prob_chunk, cnt * sizeof(double), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    }
    if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_probabilities: %s", hipGetErrorString(e));
    return LLSV_OK;
}

int llsv_copy_to_host(llsv_state* s, double* amps, uint64_t n_amps) {
    LLSV_HANDLE(s);
    if (!amps) return fail(LLSV_ERR_ARG, "llsv_copy_to_host: buffer is null");
    if (n_amps != s->size) return fail(LLSV_ERR_ARG, "llsv_copy_to_host: buffer holds %llu amplitudes, state has %llu",
                                       (unsigned long long)n_amps, (unsigned long long)s->size);
    LLSV_HIP(hipMemcpyAsync(amps, s->amps, n_amps * sizeof(v2d), hipMemcpyDeviceToHost, s->stream));
    LLSV_HIP(hipStreamSynchronize(s->stream));
    return LLSV_OK;
}

int llsv_copy_from_host(llsv_state* s, const double* amps, uint64_t n_amps) {
    LLSV_HANDLE(s);
    if (!amps) return fail(LLSV_ERR_ARG, "llsv_copy_from_host: buffer is null");
    if (n_amps != s->size)
        return fail(LLSV_ERR_ARG, "llsv_copy_from_host: buffer holds %llu amplitudes, state has %llu",
                    (unsigned long long)n_amps, (unsigned long long)s->size);
    LLSV_HIP(hipMemcpyAsync(s->amps, amps, n_amps * sizeof(v2d), hipMemcpyHostToDevice, s->stream));
    LLSV_HIP(hipStreamSynchronize(s->stream));
    return LLSV_OK;
}

int llsv_sync(llsv_state* s) {
    LLSV_HANDLE(s);
    LLSV_HIP(hipStreamSynchronize(s->stream));
    return LLSV_OK;
}

int llsv_copy_state(llsv_state* dst, llsv_state* src) {
    LLSV_HANDLE(dst);
    LLSV_HANDLE(src);
    if (dst->n != src->n) return fail(LLSV_ERR_ARG, "llsv_copy_state: %d-qubit source, %d-qubit destination", src->n, dst->n);
    if (dst == src) return LLSV_OK;
    LLSV_HIP(hipEventRecord(src->ev1, src->stream));
    LLSV_HIP(hipStreamWaitEvent(dst->stream, src->ev1, 0));
    LLSV_HIP(launch_copy(dst->stream, dst->amps, src->amps, dst->n));
    // the source must not be changed before the copy has read it
    LLSV_HIP(hipEventRecord(dst->ev1, dst->stream));
    LLSV_HIP(hipStreamWaitEvent(src->stream, dst->ev1, 0));
    return LLSV_OK;
}

int llsv_stream_sync(void* stream) {
    LLSV_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return LLSV_OK;
}

int llsv_mark(llsv_state* s, int slot) {
    LLSV_HANDLE(s);
    if (slot < 0 || slot >= LLSV_MARK_SLOTS) return fail(LLSV_ERR_ARG, "llsv_mark: slot %d outside [0, %d)", slot, LLSV_MARK_SLOTS);
    if (!s->marks[slot]) LLSV_HIP(hipEventCreate(&s->marks[slot]));
    LLSV_HIP(hipEventRecord(s->marks[slot], s->stream));
    return LLSV_OK;
}

int llsv_mark_elapsed(llsv_state* s, int a, int b, float* ms) {
    LLSV_HANDLE(s);
    if (!ms) return fail(LLSV_ERR_ARG, "llsv_mark_elapsed: ms is null");
    if (a < 0 || a >= LLSV_MARK_SLOTS || b < 0 || b >= LLSV_MARK_SLOTS || !s->marks[a] || !s->marks[b])
        return fail(LLSV_ERR_ARG, "llsv_mark_elapsed: slots %d, %d not recorded", a, b);
    LLSV_HIP(hipEventSynchronize(s->marks[b]));
    LLSV_HIP(hipEventElapsedTime(ms, s->marks[a], s->marks[b]));
    return LLSV_OK;
}

void* llsv_device_ptr(llsv_state* s) { return s ? static_cast<void*>(s->amps) : nullptr; }
void* llsv_stream(llsv_state* s) { return s ? static_cast<void*>(s->stream) : nullptr; }

int llsv_timer_start(llsv_state* s) {
    LLSV_HANDLE(s);
    LLSV_HIP(hipEventRecord(s->ev0, s->stream));
    return LLSV_OK;
}

int llsv_timer_stop(llsv_state* s, float* ms) {
    LLSV_HANDLE(s);
    if (!ms) return fail(LLSV_ERR_ARG, "llsv_timer_stop: ms is null");
    LLSV_HIP(hipEventRecord(s->ev1, s->stream));
    LLSV_HIP(hipEventSynchronize(s->ev1));
    LLSV_HIP(hipEventElapsedTime(ms, s->ev0, s->ev1));
    return LLSV_OK;
}

}  // extern "C"


// sv_common.h -- internal types and helpers shared by the gfx950 kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/llsv.h"

namespace llsv {

// One complex128 amplitude as a 16-byte vector: x = re, y = im.  Loads and stores of
// this type are single 128-bit global_load/store_dwordx4 instructions.
typedef double v2d __attribute__((ext_vector_type(2)));

// A one-qubit gate, row-major {g00, g01, g10, g11} as (re, im) pairs.  Passed to kernels
// by value, so it lands in the kernel-argument segment and is read into SGPRs.
struct Gate2 {
    double r[8];
};

__device__ __forceinline__ v2d cmul(double ar, double ai, v2d b) {
    v2d r;
    r.x = ar * b.x - ai * b.y;
    r.y = ar * b.y + ai * b.x;
    return r;
}

// g0 * a0 + g1 * a1 with g0 = (r[0], r[1]), g1 = (r[2], r[3]).
__device__ __forceinline__ v2d row2(const double* r, v2d a0, v2d a1) {
    v2d o;
    o.x = r[0] * a0.x - r[1] * a0.y + r[2] * a1.x - r[3] * a1.y;
    o.y = r[0] * a0.y + r[1] * a0.x + r[2] * a1.y + r[3] * a1.x;
    return o;
}

// Insert a zero bit at position pos of x.
__device__ __host__ __forceinline__ uint64_t insert0(uint64_t x, int pos) {
    const uint64_t lo = x & ((uint64_t(1) << pos) - 1);
    return ((x >> pos) << (pos + 1)) | lo;
}

// Launch grid for `chunks` logical workgroups: kernels loop over chunks with a grid stride, so the
// grid stays at or below 2^20 workgroups (a dispatch counts work-items in 32 bits).
inline unsigned grid_for(uint64_t chunks) {
    return unsigned(chunks < (uint64_t(1) << 20) ? (chunks ? chunks : 1) : (uint64_t(1) << 20));
}

// Tile of workgroup b out of nb: workgroups are dealt round-robin to the 8 XCDs; with a multiple
// of 8 tiles, XCD k takes the k-th eighth of them in order (measured on a 30-qubit register-tile
// pass whose tiles are 128 runs of 512 bytes 2 MiB apart: 5.4 against 4.6 TB/s).
__device__ __forceinline__ uint64_t xcd_tile(uint64_t b, uint64_t nb) {
    return (nb & 7u) ? b : (b & 7u) * (nb >> 3) + (b >> 3);
}

__device__ __forceinline__ v2d ld_nt(const v2d* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_nt(v2d* p, v2d v) { __builtin_nontemporal_store(v, p); }

// Every device and pinned-host allocation of the library goes through these two, so a test can
// make the k-th one fail (llsv_test_fail_allocation) and reach each out-of-memory path.
hipError_t dev_alloc(void** p, size_t bytes);
// LLSV_ERR_NOMEM for an out-of-memory error, else LLSV_ERR_DEVICE.
inline int hip_code(hipError_t e) { return e == hipErrorOutOfMemory ? LLSV_ERR_NOMEM : LLSV_ERR_DEVICE; }
hipError_t dev_host_alloc(void** p, size_t bytes);

// Thread-local error message for llsv_last_error().
void set_error(const std::string& msg);
// Formats the message, stores it for llsv_last_error() and returns code.
int fail(int code, const char* fmt, ...);

}  // namespace llsv

// sv_readout.hip -- readout of a state: Pauli-sum expectation values and seeded sampling.
//
// Expectation: <psi| sum_t c_t P_t |psi> with P_t a Pauli string (X on xmask bits, Z on zmask
// bits, Y where both are set).  (P psi)_i = i^{nY} (-1)^{popcount(j & z)} psi_j with j = i ^ x,
// so every term is one stream over psi with a permuted partner read; up to kTermsPerPass terms
// share a pass.  Two-stage fixed-shape reduction, bitwise reproducible run to run.
//
// Sampling: shot s draws u_s from Philox4x32-10 (key = seed, counter = (s lo, s hi, 'SAMP', 0),
// u = (x >> 11) 2^-53 from the first 64 output bits) and returns the smallest basis index i
// with CDF(i) > u_s * CDF(2^n - 1), CDF(i) = sum_{k <= i} |psi_k|^2.  The CDF is built in
// chunks of 4096 amplitudes, each summed by one wavefront as 64 sub-blocks of 64 (a wave
// reduction per sub-block, sub-blocks in order); the chunk totals are prefix-summed; a shot
// binary-searches the chunk prefixes, then walks its chunk's sub-blocks in the same order and
// finishes with a wave scan inside the sub-block, so the search reproduces exactly the sums
// the prefixes were built from.  Zero-probability amplitudes are never returned.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/llsv.h"
#include "sv_common.h"
#include "sv_kernels.h"
#include "sv_state.h"

namespace llsv {
namespace {

constexpr int kTermsPerPass = 16;
constexpr int kChunkBits = 12;  // 4096 amplitudes per CDF chunk

struct PauliTerm {
    double coeff;
    uint64_t x, z;
};

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// partials[t * gridDim.x + block] = partial sum of c_t Re(conj(psi_i) (P_t psi)_i).
__global__ __launch_bounds__(256) void k_expect_pauli(const v2d* __restrict__ psi, uint64_t size,
                                                      const PauliTerm* __restrict__ terms, int nterms,
                                                      double* __restrict__ partials) {
    __shared__ double lds[kTermsPerPass][4];
    double acc[kTermsPerPass];
#pragma unroll
    for (int t = 0; t < kTermsPerPass; ++t) acc[t] = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < size; i += uint64_t(gridDim.x) * 256) {
        const v2d p = psi[i];
#pragma unroll
        for (int t = 0; t < kTermsPerPass; ++t) {
            if (t < nterms) {
                const PauliTerm T = terms[t];
                const uint64_t j = i ^ T.x;
                const v2d a = psi[j];
                const int ny = __popcll(T.x & T.z) & 3;
                double qr = a.x, qi = a.y;
                if (ny == 1) { const double t0 = qr; qr = -qi; qi = t0; }
                else if (ny == 2) { qr = -qr; qi = -qi; }
                else if (ny == 3) { const double t0 = qr; qr = qi; qi = -t0; }
                const double re = p.x * qr + p.y * qi;
                acc[t] += (__popcll(j & T.z) & 1) ? -re : re;
            }
        }
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < kTermsPerPass; ++t) {
        const double v = wave_sum64(acc[t]);
        if ((threadIdx.x & 63) == 0) lds[t][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < kTermsPerPass && int(threadIdx.x) < nterms) {
        const int t = threadIdx.x;
        partials[size_t(t) * gridDim.x + blockIdx.x] = terms[t].coeff * ((lds[t][0] + lds[t][1]) + (lds[t][2] + lds[t][3]));
    }
}

// out[t] = sum over blocks of partials[t * nblk + b], fixed tree; one block per term.
__global__ __launch_bounds__(256) void k_expect_final(const double* __restrict__ partials, int nblk, double* __restrict__ out) {
    __shared__ double lds[4];
    const int t = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) acc += partials[size_t(t) * nblk + b];
    acc = wave_sum64(acc);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[t] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

__device__ __forceinline__ double prob(const v2d* psi, uint64_t i, uint64_t size) {
    if (i >= size) return 0.0;
    const v2d a = psi[i];
    return a.x * a.x + a.y * a.y;
}

// One wavefront per chunk: sum of its sub-block sums, sub-blocks in order.
__global__ __launch_bounds__(256) void k_chunk_sums(const v2d* __restrict__ psi, uint64_t size, uint64_t chunk,
                                                    uint64_t nchunks, double* __restrict__ sums) {
    const int lane = threadIdx.x & 63;
    const uint64_t nsub = (chunk + 63) / 64;
    for (uint64_t c = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); c < nchunks; c += uint64_t(gridDim.x) * 4) {
        double acc = 0.0;
        for (uint64_t j = 0; j < nsub; ++j) acc += wave_sum64(prob(psi, c * chunk + j * 64 + lane, size));
        if (lane == 0) sums[c] = acc;
    }
}

// Exclusive prefix of the chunk sums, one block of 1024 threads: each thread scans a contiguous
// segment, thread 0 scans the segment totals, every thread adds its offset.  total[0] = sum.
__global__ __launch_bounds__(1024) void k_chunk_prefix(const double* __restrict__ sums, uint64_t nchunks,
                                                       double* __restrict__ prefix, double* __restrict__ total) {
    __shared__ double seg[1024];
    const uint64_t per = (nchunks + 1023) / 1024;
    const uint64_t b0 = uint64_t(threadIdx.x) * per, b1 = std::min<uint64_t>(nchunks, b0 + per);
    double run = 0.0;
    for (uint64_t b = b0; b < b1; ++b) {
        prefix[b] = run;
        run += sums[b];
    }
    seg[threadIdx.x] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int t = 0; t < 1024; ++t) {
            const double v = seg[t];
            seg[t] = s;
            s += v;
        }
        total[0] = s;
    }
    __syncthreads();
    const double off = seg[threadIdx.x];
    for (uint64_t b = b0; b < b1; ++b) prefix[b] += off;
}

__device__ inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
        c[1] = uint32_t(p1);
        c[3] = uint32_t(p0);
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

constexpr uint32_t kSampleTag = 0x53414D50u;  // "SAMP"

// One wavefront per shot (grid-stride): the search described at the top of the file.
__global__ __launch_bounds__(256) void k_sample(const v2d* __restrict__ psi, uint64_t size, uint64_t chunk,
                                                uint64_t nchunks, const double* __restrict__ prefix,
                                                const double* __restrict__ sums, const double* __restrict__ total,
                                                uint64_t seed, uint64_t shots, uint64_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const uint64_t nsub = (chunk + 63) / 64;
    const double tot = total[0];
    for (uint64_t s = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); s < shots; s += uint64_t(gridDim.x) * 4) {
        uint32_t c[4] = {uint32_t(s), uint32_t(s >> 32), kSampleTag, 0u};
        philox10(c, uint32_t(seed), uint32_t(seed >> 32));
        const uint64_t x = uint64_t(c[0]) | (uint64_t(c[1]) << 32);
        const double target = double(x >> 11) * (1.0 / 9007199254740992.0) * tot;
        // last chunk whose prefix is <= target
        uint64_t lo = 0, hi = nchunks;  // invariant: prefix[lo] <= target, answer in [lo, hi)
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (prefix[mid] <= target) lo = mid;
            else hi = mid;
        }
        // the chunk's total can fall short of the target only by rounding: move back to the
        // last chunk with probability in it
        while (lo > 0 && sums[lo] == 0.0) --lo;
        const double r = target - prefix[lo];
        double acc = 0.0;
        uint64_t found = ~uint64_t(0), last_nonzero = ~uint64_t(0);
        for (uint64_t j = 0; j < nsub && found == ~uint64_t(0); ++j) {
            const uint64_t i = lo * chunk + j * 64 + lane;
            const double v = prob(psi, i, size);
            double incl = v;  // inclusive scan over the wave
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const double y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            const double ssum = wave_sum64(v);
            const uint64_t hit = __ballot(acc + incl > r && v > 0.0);
            const uint64_t nz = __ballot(v > 0.0);
            if (nz) last_nonzero = lo * chunk + j * 64 + (63 - __builtin_clzll(nz));
            if (hit) found = lo * chunk + j * 64 + __builtin_ctzll(hit);
            acc += ssum;
        }
        if (found == ~uint64_t(0)) found = last_nonzero;  // rounding at the very end of the chunk
        if (lane == 0) out[s] = found;
    }
}

}  // namespace
}  // namespace llsv

using namespace llsv;

extern "C" int llsv_expect_pauli(llsv_state* s, const llsv_pauli* terms, int n_terms, double* per_term, double* total) {
    if (!s) return fail(LLSV_ERR_HANDLE, "llsv_expect_pauli: null state handle");
    if (n_terms < 0 || (n_terms > 0 && !terms) || !total) return fail(LLSV_ERR_ARG, "llsv_expect_pauli: bad arguments");
    const uint64_t all = s->n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << s->n) - 1);
    for (int t = 0; t < n_terms; ++t)
        if ((terms[t].xmask | terms[t].zmask) & ~all)
            return fail(LLSV_ERR_ARG, "llsv_expect_pauli: term %d acts on a qubit outside [0, %d)", t, s->n);
    std::vector<double> host(std::max(1, n_terms), 0.0);
    const uint64_t size = s->size;
    const int grid = int(std::min<uint64_t>((size + 255) / 256, 1024));
    double* d = nullptr;  // [terms][kTermsPerPass] + partials[kTermsPerPass * grid] + out[kTermsPerPass]
    const size_t bytes = sizeof(PauliTerm) * kTermsPerPass + sizeof(double) * (size_t(kTermsPerPass) * grid + kTermsPerPass);
    hipError_t e = n_terms ? dev_alloc(reinterpret_cast<void**>(&d), bytes) : hipSuccess;
    PauliTerm* dterms = reinterpret_cast<PauliTerm*>(d);
    double* dpart = d ? d + 3 * kTermsPerPass : nullptr;
    double* dout = dpart ? dpart + size_t(kTermsPerPass) * grid : nullptr;
    for (int t0 = 0; t0 < n_terms && e == hipSuccess; t0 += kTermsPerPass) {
        const int nt = std::min(kTermsPerPass, n_terms - t0);
        PauliTerm tt[kTermsPerPass];
        for (int t = 0; t < nt; ++t) tt[t] = PauliTerm{terms[t0 + t].coeff, terms[t0 + t].xmask, terms[t0 + t].zmask};
        e = hipMemcpyAsync(dterms, tt, sizeof(PauliTerm) * nt, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_expect_pauli, dim3(grid), dim3(256), 0, s->stream, s->amps, size, dterms, nt, dpart);
            hipLaunchKernelGGL(k_expect_final, dim3(nt), dim3(256), 0, s->stream, dpart, grid, dout);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(host.data() + t0, dout, sizeof(double) * nt, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    }
    if (d) (void)hipFree(d);
    if (e != hipSuccess) return fail(hip_code(e), "llsv_expect_pauli: %s", hipGetErrorString(e));
    double sum = 0.0;
    for (int t = 0; t < n_terms; ++t) sum += host[t];
    if (per_term) std::memcpy(per_term, host.data(), sizeof(double) * n_terms);
    *total = sum;
    return LLSV_OK;
}

// Scratch of llsv_sample_noalloc: chunk sums, prefixes, the total and the sampled indices, one
// copy per device for the life of the process.  Host code never passes `scratch` itself to a
// kernel (on the host it is a stand-in, not a device address): it uses the address that
// hipGetSymbolAddress returns for the device (scratch_on below).
static __device__ __align__(256) unsigned char scratch[LLSV_SAMPLE_SCRATCH_BYTES];

namespace {

// Device bytes a sampling call needs: chunk sums and prefixes, the total, the sampled indices.
uint64_t sample_bytes(uint64_t size, uint64_t shots) {
    const uint64_t chunk = std::min<uint64_t>(size, uint64_t(1) << kChunkBits);
    return sizeof(double) * (2 * (size / chunk) + 8) + sizeof(uint64_t) * shots;
}

// The sampling of llsv_sample with its device buffer d (sample_bytes(size, shots) bytes) given.
int sample_into(llsv_state* s, uint64_t shots, uint64_t seed, uint64_t* out, unsigned char* d, const char* who) {
    const uint64_t size = s->size;
    const uint64_t chunk = std::min<uint64_t>(size, uint64_t(1) << kChunkBits);
    const uint64_t nchunks = size / chunk;
    double* sums = reinterpret_cast<double*>(d);
    double* prefix = sums + nchunks;
    double* total = prefix + nchunks;
    uint64_t* dout = reinterpret_cast<uint64_t*>(total + 8);
    hipLaunchKernelGGL(k_chunk_sums, dim3(unsigned(std::min<uint64_t>((nchunks + 3) / 4, 1u << 16))), dim3(256), 0,
                       s->stream, s->amps, size, chunk, nchunks, sums);
    hipLaunchKernelGGL(k_chunk_prefix, dim3(1), dim3(1024), 0, s->stream, sums, nchunks, prefix, total);
    hipLaunchKernelGGL(k_sample, dim3(unsigned(std::min<uint64_t>((shots + 3) / 4, 1u << 16))), dim3(256), 0, s->stream,
                       s->amps, size, chunk, nchunks, prefix, sums, total, seed, shots, dout);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(uint64_t) * shots, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->result_host, total, sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return fail(LLSV_ERR_DEVICE, "%s: %s", who, hipGetErrorString(e));
    const double tot = s->result_host[0];
    if (!(tot > 0.0) || !std::isfinite(tot))
        return fail(LLSV_ERR_NUMERIC, "%s: the state has no probability to sample from (norm %g)", who, tot);
    return LLSV_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with that device current.
int scratch_on(int device, unsigned char** out, const char* who) {
    if (device < 0 || device >= kMaxDevices) return fail(LLSV_ERR_DEVICE, "%s: device %d out of range", who, device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(LLSV_ERR_DEVICE, "%s: scratch address on device %d: %s", who, device, hipGetErrorString(e));
        g_scratch_dev[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_dev[device]);
    return LLSV_OK;
}

}  // namespace

extern "C" int llsv_sample(llsv_state* s, uint64_t shots, uint64_t seed, uint64_t* out) {
    if (!s) return fail(LLSV_ERR_HANDLE, "llsv_sample: null state handle");
    if (shots && !out) return fail(LLSV_ERR_ARG, "llsv_sample: out is null");
    if (shots == 0) return LLSV_OK;
    const uint64_t bytes = sample_bytes(s->size, shots);
    unsigned char* d = nullptr;
    hipError_t e = dev_alloc(reinterpret_cast<void**>(&d), bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(LLSV_ERR_NOMEM, "llsv_sample: cannot allocate %llu bytes", (unsigned long long)bytes);
    }
    const int rc = sample_into(s, shots, seed, out, d, "llsv_sample");
    (void)hipFree(d);
    return rc;
}

extern "C" int llsv_sample_noalloc(llsv_state* s, uint64_t shots, uint64_t seed, uint64_t* out) {
    if (!s) return fail(LLSV_ERR_HANDLE, "llsv_sample_noalloc: null state handle");
    if (shots && !out) return fail(LLSV_ERR_ARG, "llsv_sample_noalloc: out is null");
    if (shots == 0) return LLSV_OK;
    const uint64_t bytes = sample_bytes(s->size, shots);
    if (shots > LLSV_SAMPLE_SCRATCH_BYTES / 8 || bytes > LLSV_SAMPLE_SCRATCH_BYTES)
        return fail(LLSV_ERR_ARG,
                    "llsv_sample_noalloc: %llu shots on %d qubits need %llu bytes of scratch, more than the %llu "
                    "available (use llsv_sample)",
                    (unsigned long long)shots, s->n, (unsigned long long)bytes,
                    (unsigned long long)LLSV_SAMPLE_SCRATCH_BYTES);
    unsigned char* d = nullptr;
    if (int rc = scratch_on(s->device, &d, "llsv_sample_noalloc")) return rc;
    return sample_into(s, shots, seed, out, d, "llsv_sample_noalloc");
}

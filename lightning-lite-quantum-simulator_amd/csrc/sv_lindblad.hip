// sv_lindblad.hip -- batched Lindblad quantum trajectories with dense operators (<= 10 qubits):
// the propagation prop . psi of every trajectory in a step is one complex GEMM on the FP64 matrix
// cores, Psi_next^T = Psi^T prop^T (trajectories as rows).
//
// k_zgemm: C = A B for complex128 row-major matrices, M % 64 == 0, N % 64 == 0, K % 16 == 0 (callers
// pad with zeros).  A 256-thread workgroup computes a 64 x 64 block of C; each wavefront a 32 x 32
// quarter as 2 x 2 blocks of v_mfma_f64_16x16x4f64, complex through the real embedding
//     Cr += Ar Br - Ai Bi,   Ci += Ar Bi + Ai Br          (four MFMAs per block and K step).
// A and B tiles (64 x 16 and 16 x 64) are staged in LDS as separate real and imaginary planes,
// K-major, so every MFMA operand is one conflict-free 8-byte LDS read per lane.
//
// The jump algorithm (src/TrajectorySolver.h:379-428) per step, for all trajectories at once:
//   k_zgemm       tmp = prop psi                         (or the RK4 stages for a time-dependent H)
//   k_jump_decide one wavefront per trajectory: survive = ||tmp||^2; jump iff u >= survive (u the
//                 trajectory's next uniform); otherwise psi = tmp / sqrt(survive)
//   k_zgemm<true> the jumping trajectories' tmp rows (gathered by a device-side list) times
//                 [C_1^T | .. | C_n^T]: every channel's C_k tmp in one GEMM
//   k_jump_finish one wavefront per jumping trajectory: w_k = ||C_k tmp||^2, channel by cumulative
//                 weight (a second uniform), psi = C_k tmp / sqrt(w_k)
// The uniforms are the reference's: per trajectory std::mt19937_64 seeded with
// splitmix64(seed ^ splitmix64(i + 1)) and std::uniform_real_distribution<double>(0, 1), drawn on
// the host (at most two per step) -- or counter-based Philox draws on the device.
// The diffusive unraveling (src/TrajectorySolver.h:432-460) stacks every C_k and C_k^+ C_k into
// one GEMM per step and updates each trajectory in one wavefront (k_diffuse).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <thread>
#include <vector>

#include "../../include/llsv.h"
#include "sv_common.h"

namespace llsv {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64, kBN = 64, kBK = 16;

// GATHER: row r of A is row idx[r0 + r] of the A array when r0 + r < *count, zero otherwise, and
// workgroups whose first row is at or past *count return at once (the count lives on the device).
// BM: rows of C per workgroup, 64 (2 x 2 wavefronts of 32 x 32) or 16 (1 x 4 wavefronts of one
// 16 x 16 block each: few rows, as in the gathered jump products).
template <bool GATHER, int BM>
__global__ __launch_bounds__(256) void k_zgemm(const v2d* __restrict__ A, int lda, const v2d* __restrict__ B,
                                               int ldb, v2d* __restrict__ C, int ldc, int K,
                                               const uint32_t* __restrict__ idx, const uint32_t* __restrict__ count,
                                               uint32_t r0) {
    static_assert(BM == 64 || BM == 16, "tile rows");
    constexpr int WN = BM == 64 ? 2 : 4;         // wavefronts along N
    constexpr int MB = BM == 64 ? 2 : 1;         // 16-row blocks per wavefront
    constexpr int NB = BM == 64 ? 2 : 1;         // 16-column blocks per wavefront
    constexpr int EA = BM * kBK / 256;           // A elements per thread per tile (4 or 1)
    __shared__ double as_r[kBK][BM], as_i[kBK][BM], bs_r[kBK][kBN], bs_i[kBK][kBN];
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int m0 = int(blockIdx.y) * BM, n0 = int(blockIdx.x) * kBN;
    uint32_t rows = 0;
    if (GATHER) {
        const uint32_t cnt = *count;
        if (r0 + uint32_t(m0) >= cnt) return;  // uniform across the workgroup
        rows = cnt - r0 - uint32_t(m0);        // valid rows of this block (may exceed BM)
    }
    d4 cr[MB][NB], ci[MB][NB];
    for (int a = 0; a < MB; ++a)
        for (int b = 0; b < NB; ++b) {
            cr[a][b] = d4{0.0, 0.0, 0.0, 0.0};
            ci[a][b] = d4{0.0, 0.0, 0.0, 0.0};
        }
    // tile loads: A row ar, columns ac .. ac + EA - 1; B row br = tid / 16, columns 4 (tid % 16) .. +3
    const int ar = tid / (kBK / EA), ac = (tid % (kBK / EA)) * EA, br = tid >> 4, bc = (tid & 15) * 4;
    const bool a_live = !GATHER || uint32_t(ar) < rows;
    const v2d* arow = GATHER ? (a_live ? A + size_t(idx[r0 + uint32_t(m0 + ar)]) * lda : A) : A + size_t(m0 + ar) * lda;
    // software pipeline: the next tile's global loads are in flight while this tile is multiplied
    v2d va[EA], vb[4];
    auto load = [&](int k0) {
        const v2d* pa = arow + k0 + ac;
        const v2d* pb = B + size_t(k0 + br) * ldb + n0 + bc;
#pragma unroll
        for (int j = 0; j < EA; ++j) va[j] = a_live ? pa[j] : v2d{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 4; ++j) vb[j] = pb[j];
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += kBK) {
        __syncthreads();  // the previous tile is no longer read
#pragma unroll
        for (int j = 0; j < EA; ++j) {
            as_r[ac + j][ar] = va[j].x;
            as_i[ac + j][ar] = va[j].y;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bs_r[br][bc + j] = vb[j].x;
            bs_i[br][bc + j] = vb[j].y;
        }
        __syncthreads();
        if (k0 + kBK < K) load(k0 + kBK);
#pragma unroll
        for (int ks = 0; ks < kBK / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            double a_r[MB], a_i[MB], a_n[MB], b_r[NB], b_i[NB];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                a_r[mb] = as_r[kk][wm * MB * 16 + mb * 16 + (lane & 15)];
                a_i[mb] = as_i[kk][wm * MB * 16 + mb * 16 + (lane & 15)];
                a_n[mb] = -a_i[mb];
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                b_r[nb] = bs_r[kk][wn * NB * 16 + nb * 16 + (lane & 15)];
                b_i[nb] = bs_i[kk][wn * NB * 16 + nb * 16 + (lane & 15)];
            }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    cr[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[mb], b_r[nb], cr[mb][nb], 0, 0, 0);
                    ci[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[mb], b_i[nb], ci[mb][nb], 0, 0, 0);
                }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    cr[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_n[mb], b_i[nb], cr[mb][nb], 0, 0, 0);
                    ci[mb][nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i[mb], b_r[nb], ci[mb][nb], 0, 0, 0);
                }
        }
    }
    // D[i][j] of a 16 x 16 block: lane j + 16 (i mod 4), element i / 4
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm * MB * 16 + mb * 16 + 4 * v + (lane >> 4);
                const int col = n0 + wn * NB * 16 + nb * 16 + (lane & 15);
                v2d y;
                y.x = cr[mb][nb][v];
                y.y = ci[mb][nb][v];
                C[size_t(row) * ldc + col] = y;
            }
}

// ---- small elementwise kernels -------------------------------------------------------------

// X[i] = X[i] / d (both components), n elements.
__global__ void k_div(v2d* __restrict__ X, uint64_t n, double d) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) {
        v2d v = X[i];
        v.x = v.x / d;
        v.y = v.y / d;
        X[i] = v;
    }
}

// Y[i] += X[i].
__global__ void k_add(v2d* __restrict__ Y, const v2d* __restrict__ X, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) Y[i] = Y[i] + X[i];
}

// Y = X + c Z (real c).
__global__ void k_axpy(v2d* __restrict__ Y, const v2d* __restrict__ X, const v2d* __restrict__ Z, double c, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) {
        const v2d x = X[i], z = Z[i];
        v2d y;
        y.x = x.x + c * z.x;
        y.y = x.y + c * z.y;
        Y[i] = y;
    }
}

// RK4 combination: T = P + (h / 6) (K1 + 2 K2 + 2 K3 + K4), in the reference's order of operations.
__global__ void k_rk4(v2d* __restrict__ T, const v2d* __restrict__ P, const v2d* __restrict__ K1,
                      const v2d* __restrict__ K2, const v2d* __restrict__ K3, const v2d* __restrict__ K4, double h6,
                      uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256) {
        const v2d s = ((K1[i] + 2.0 * K2[i]) + 2.0 * K3[i]) + K4[i];
        T[i] = P[i] + h6 * s;
    }
}

// AT = X^T (D x T) and BC = conj(X) (T x D) for X of T rows of D.
__global__ void k_transpose_conj(const v2d* __restrict__ X, v2d* __restrict__ AT, v2d* __restrict__ BC, uint64_t T, int D) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= T * uint64_t(D)) return;
    const uint64_t t = i / D, c = i % D;
    const v2d v = X[i];
    AT[c * T + t] = v;
    BC[i] = v2d{v.x, -v.y};
}

__global__ void k_identity(v2d* __restrict__ X, int D) {
    const uint64_t n = uint64_t(D) * D;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256)
        X[i] = v2d{(i / D == i % D) ? 1.0 : 0.0, 0.0};
}

// ---- plan assembly on the device ----------------------------------------------------------------

// n row-major dim x dim operators C_k (raw) into BT = [C_1^T | .. | C_n^T] (row stride W, column
// block k at k D) and CJ = [conj(C_1); ..; conj(C_n)] (n D rows of D).  Padding entries are not
// written (the caller clears both).
__global__ void k_stack_ops(const v2d* __restrict__ raw, int n, int dim, int D, int W, v2d* __restrict__ BT,
                            v2d* __restrict__ CJ) {
    const uint64_t per = uint64_t(dim) * dim, total = per * uint64_t(n);
    for (uint64_t e = uint64_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += uint64_t(gridDim.x) * 256) {
        const uint64_t k = e / per, r = e % per, i = r / dim, j = r % dim;
        const v2d v = raw[e];
        BT[j * W + k * D + i] = v;
        CJ[(k * D + i) * D + j] = v2d{v.x, -v.y};
    }
}

// out = (-i s H + g G)^T as a D x D block: out[i][j] = -i s H[j][i] + g GT[i][j] for i, j < dim,
// zero in the padding.  H row-major dim x dim; GT (D x D) may be null (g G = 0).
__global__ void k_node_t(const v2d* __restrict__ H, int dim, const v2d* __restrict__ GT, double s, double g,
                         v2d* __restrict__ out, int D) {
    const uint64_t nn = uint64_t(D) * D;
    for (uint64_t e = uint64_t(blockIdx.x) * 256 + threadIdx.x; e < nn; e += uint64_t(gridDim.x) * 256) {
        const uint64_t i = e / D, j = e % D;
        v2d v{0.0, 0.0};
        if (i < uint64_t(dim) && j < uint64_t(dim)) {
            const v2d x = H[j * dim + i];
            v = v2d{s * x.y, -(s * x.x)};  // -i s (x + i y)
            if (GT) v = v + g * GT[e];
        }
        out[e] = v;
    }
}

// Column sums of |X| for a D x D matrix X (the row sums of the matrix whose transpose X holds: the
// infinity norm's terms), in chunks of kColRows rows: sums[y][j] = sum over rows i of chunk y of
// |X[i][j]|, for grid (ceil(D / 256), D / kColRows).  The caller adds the chunks.
constexpr int kColRows = 64;
__global__ void k_colabs(const v2d* __restrict__ X, int D, double* __restrict__ sums) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= D) return;
    double r = 0.0;
    const int i0 = blockIdx.y * kColRows;
    for (int i = i0; i < i0 + kColRows; ++i) {
        const v2d v = X[size_t(i) * D + j];
        r += hypot(v.x, v.y);
    }
    sums[size_t(blockIdx.y) * D + j] = r;
}

// Entries of a D x D block outside the leading dim x dim corner = 0.
__global__ void k_zero_pad(v2d* __restrict__ X, int dim, int D) {
    const uint64_t nn = uint64_t(D) * D;
    for (uint64_t e = uint64_t(blockIdx.x) * 256 + threadIdx.x; e < nn; e += uint64_t(gridDim.x) * 256)
        if (e / D >= uint64_t(dim) || e % D >= uint64_t(dim)) X[e] = v2d{0.0, 0.0};
}

// ---- random numbers ---------------------------------------------------------------------------

__device__ __host__ inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
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

// The c-th uniform of trajectory t: a host-drawn table (stride per trajectory) or Philox.
struct Draws {
    const double* table;  // null: Philox
    uint64_t stride;
    uint64_t seed;
    uint64_t first;  // index of trajectory 0 of this launch in the ensemble
    __device__ double u(uint64_t t, uint64_t c) const {
        if (table) return table[t * stride + c];
        uint32_t x[4] = {uint32_t(first + t), uint32_t((first + t) >> 32), uint32_t(c), uint32_t(c >> 32) ^ 0x4C494E44u};
        philox10(x, uint32_t(seed), uint32_t(seed >> 32));
        const uint64_t v = uint64_t(x[0]) | (uint64_t(x[1]) << 32);
        return double(v >> 11) * (1.0 / 9007199254740992.0);
    }
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---- jump steps -----------------------------------------------------------------------------

// One wavefront per trajectory.  state[t]: 0 running, 1 jump pending (set here), 2 norm vanished.
__global__ __launch_bounds__(256) void k_jump_decide(const v2d* __restrict__ tmp, v2d* __restrict__ psi, int D,
                                                     uint64_t T, int nch, Draws dr, uint64_t* __restrict__ cursor,
                                                     double* __restrict__ survive_out, int* __restrict__ flag,
                                                     uint32_t* __restrict__ jlist, uint32_t* __restrict__ jcount) {
    const uint64_t t = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = int(threadIdx.x & 63);
    if (t >= T) return;
    const v2d* x = tmp + t * D;
    double s = 0.0;
    for (int i = lane; i < D; i += 64) s += x[i].x * x[i].x + x[i].y * x[i].y;
    s = wave_sum(s);
    bool jump = false;
    if (!(s > 0.0) || !(s <= 1.7976931348623157e308)) {
        if (lane == 0) atomicOr(flag, 1);
        s = 1.0;
    } else if (nch > 0) {
        const uint64_t c = cursor[t];
        jump = dr.u(t, c) >= s;
        if (lane == 0) cursor[t] = c + 1;
    }
    if (lane == 0) survive_out[t] = s;
    if (jump) {
        if (lane == 0) jlist[atomicAdd(jcount, 1u)] = uint32_t(t);
        return;
    }
    const double r = sqrt(s);
    v2d* y = psi + t * D;
    for (int i = lane; i < D; i += 64) {
        v2d v = x[i];
        v.x = v.x / r;
        v.y = v.y / r;
        y[i] = v;
    }
}

// One wavefront per jumping trajectory r0 + r (r < rows of this round): Y row r holds
// [C_1 tmp | .. | C_n tmp] (n D entries); w_k = ||C_k tmp||^2, the channel by cumulative weight
// (src/TrajectorySolver.h:395-417), psi = C_k tmp / sqrt(w_k), or tmp / sqrt(survive) when no
// channel can fire.
__global__ __launch_bounds__(256) void k_jump_finish(const v2d* __restrict__ Y, const v2d* __restrict__ tmp,
                                                     v2d* __restrict__ psi, int D, int nch,
                                                     const uint32_t* __restrict__ jlist,
                                                     const uint32_t* __restrict__ jcount, uint32_t r0, uint32_t rmax,
                                                     Draws dr, uint64_t* __restrict__ cursor,
                                                     const double* __restrict__ survive) {
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = int(threadIdx.x & 63);
    const uint32_t cnt = *jcount;
    if (r >= rmax || r0 + r >= cnt) return;
    const uint64_t t = jlist[r0 + r];
    const v2d* y = Y + size_t(r) * size_t(nch) * D;
    double wk[16];
    double total = 0.0;
    for (int k = 0; k < nch; ++k) {
        double acc = 0.0;
        for (int i = lane; i < D; i += 64) {
            const v2d v = y[size_t(k) * D + i];
            acc += v.x * v.x + v.y * v.y;
        }
        wk[k] = wave_sum(acc);
        total += wk[k];
    }
    v2d* out = psi + t * D;
    if (!(total > 0.0)) {
        const double rs = sqrt(survive[t]);
        for (int i = lane; i < D; i += 64) {
            v2d v = tmp[t * D + i];
            v.x = v.x / rs;
            v.y = v.y / rs;
            out[i] = v;
        }
        return;
    }
    const uint64_t c = cursor[t];
    const double rr = dr.u(t, c) * total;
    int pick = nch;
    double acc = 0.0;
    for (int k = 0; k < nch; ++k) {
        acc += wk[k];
        if (rr < acc && wk[k] > 0.0) {
            pick = k;
            break;
        }
    }
    if (pick == nch)
        for (int k = nch; k-- > 0;)
            if (wk[k] > 0.0) {
                pick = k;
                break;
            }
    if (lane == 0) cursor[t] = c + 1;
    const double nrm = sqrt(wk[pick]);
    for (int i = lane; i < D; i += 64) {
        v2d v = y[size_t(pick) * D + i];
        v.x = v.x / nrm;
        v.y = v.y / nrm;
        out[i] = v;
    }
}

// ---- diffusive steps ------------------------------------------------------------------------

// One wavefront per trajectory.  cp = [C_1 psi | .. | C_n psi | D_1 psi | .. | D_n psi] (row t,
// 2 n D entries, D_k = C_k^+ C_k); drift (D entries, or null: zero); psi updated in place.
// normals: n per step per trajectory (host table; Philox mode draws Box-Muller pairs).
__global__ __launch_bounds__(256) void k_diffuse(v2d* __restrict__ psi, const v2d* __restrict__ cp,
                                                 const v2d* __restrict__ drift_in, int D, uint64_t T, int nch,
                                                 double h, double sq_h, const double* __restrict__ normals,
                                                 uint64_t nstride, uint64_t step, Draws dr, int* __restrict__ flag) {
    const uint64_t t = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = int(threadIdx.x & 63);
    if (t >= T) return;
    v2d* p = psi + t * D;
    const v2d* c = cp + t * size_t(2 * nch) * D;
    constexpr int kMaxCh = 16;
    double x[kMaxCh], g[kMaxCh];
    for (int k = 0; k < nch; ++k) {
        // x_k = 2 Re <psi | C_k psi>
        double s = 0.0;
        for (int i = lane; i < D; i += 64) {
            const v2d a = p[i], b = c[size_t(k) * D + i];
            s += a.x * b.x + a.y * b.y;
        }
        x[k] = 2.0 * wave_sum(s);
        if (normals) {
            g[k] = normals[t * nstride + step * nch + k];
        } else {  // Box-Muller on two Philox uniforms, scaled by sqrt(h)
            const double u1 = dr.u(t, 2 * (step * nch + k)), u2 = dr.u(t, 2 * (step * nch + k) + 1);
            g[k] = sq_h * sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
        }
    }
    double nn = 0.0;
    for (int i = lane; i < D; i += 64) {
        const v2d a = p[i];
        v2d drift = drift_in ? drift_in[t * D + i] : v2d{0.0, 0.0};
        v2d noise = v2d{0.0, 0.0};
        for (int k = 0; k < nch; ++k) {
            const v2d cpk = c[size_t(k) * D + i], t2 = c[size_t(nch + k) * D + i];
            // drift += -0.5 (t2 - x cp + 0.25 x^2 psi); noise += (cp - 0.5 x psi) g
            const v2d inner = (t2 - x[k] * cpk) + (0.25 * x[k] * x[k]) * a;
            drift = drift + (-0.5) * inner;
            noise = noise + (cpk - (0.5 * x[k]) * a) * g[k];
        }
        const v2d v = a + (h * drift + noise);  // psi += h drift + noise
        p[i] = v;
        nn += v.x * v.x + v.y * v.y;
    }
    nn = sqrt(wave_sum(nn));
    if (!(nn > 0.0) || !(nn <= 1.7976931348623157e308)) {
        if (lane == 0) atomicOr(flag, 2);
        return;
    }
    for (int i = lane; i < D; i += 64) {
        v2d v = p[i];
        v.x = v.x / nn;
        v.y = v.y / nn;
        p[i] = v;
    }
}

// ---- host ----------------------------------------------------------------------------------

int pad_dim(int dim) { return dim < kBN ? kBN : dim; }

hipError_t gemm(hipStream_t st, const v2d* A, int lda, const v2d* B, int ldb, v2d* C, int ldc, int M, int N, int K) {
    hipLaunchKernelGGL((k_zgemm<false, 64>), dim3(unsigned(N / kBN), unsigned(M / kBM)), dim3(256), 0, st, A, lda, B, ldb,
                       C, ldc, K, nullptr, nullptr, 0u);
    return hipGetLastError();
}

// C = A[idx[r0 ..]] B over at most M gathered rows; blocks past the device-side count return.
hipError_t gemm_gather(hipStream_t st, const v2d* A, int lda, const uint32_t* idx, const uint32_t* count, uint32_t r0,
                       const v2d* B, int ldb, v2d* C, int ldc, int M, int N, int K) {
    hipLaunchKernelGGL((k_zgemm<true, 16>), dim3(unsigned(N / kBN), unsigned(M / 16)), dim3(256), 0, st, A, lda, B, ldb,
                       C, ldc, K, idx, count, r0);
    return hipGetLastError();
}

unsigned grid_of(uint64_t n) { return unsigned(std::min<uint64_t>((n + 255) / 256, 65536)); }

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Host upload of a dim x dim complex matrix (row-major) into a D x D device block, transposed
// when `transpose` (the GEMM's B operand is prop^T), zero padding.
std::vector<double> padded(const double* m, int dim, int D, bool transpose) {
    std::vector<double> out(size_t(2) * D * D, 0.0);
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) {
            const size_t src = 2 * (size_t(i) * dim + j);
            const size_t dst = transpose ? 2 * (size_t(j) * D + i) : 2 * (size_t(i) * D + j);
            out[dst] = m[src];
            out[dst + 1] = m[src + 1];
        }
    return out;
}

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const hipError_t e = dev_alloc(&p, bytes ? bytes : 16);
        if (e == hipSuccess) cap = bytes ? bytes : 16;
        return e;
    }
    // grow-only: keeps the allocation when it is large enough
    hipError_t need(size_t bytes) { return bytes <= cap && p ? hipSuccess : alloc(bytes); }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Scaling power of exp(A): the least s >= 0 with ||A||_inf / 2^s <= 1/2 (src/TrajectorySolver.h:97-111).
int expm_power(double nrm) { return nrm > 0.5 ? std::max(0, int(std::ceil(std::log2(nrm / 0.5)))) : 0; }

// res = exp(A) for the D x D matrix in `a` (scaled in place by 2^-s): 18-term Taylor series of
// A / 2^s, then s squarings, all GEMMs on the device.  term and tmp are scratch; every buffer holds
// D x D and the result may end up in any of them but is swapped into `res`.
hipError_t expm_dev(hipStream_t st, DevBuf& a, DevBuf& res, DevBuf& term, DevBuf& tmp, int D, int s) {
    const uint64_t nn = uint64_t(D) * D;
    const unsigned g = grid_of(nn);
    hipLaunchKernelGGL(k_div, dim3(g), dim3(256), 0, st, a.as<v2d>(), nn, std::ldexp(1.0, s));
    hipLaunchKernelGGL(k_identity, dim3(g), dim3(256), 0, st, term.as<v2d>(), D);
    hipLaunchKernelGGL(k_identity, dim3(g), dim3(256), 0, st, res.as<v2d>(), D);
    hipError_t e = hipGetLastError();
    for (int k = 1; k <= 18 && e == hipSuccess; ++k) {
        e = gemm(st, term.as<v2d>(), D, a.as<v2d>(), D, tmp.as<v2d>(), D, D, D, D);
        hipLaunchKernelGGL(k_div, dim3(g), dim3(256), 0, st, tmp.as<v2d>(), nn, double(k));
        term.swap(tmp);
        hipLaunchKernelGGL(k_add, dim3(g), dim3(256), 0, st, res.as<v2d>(), term.as<const v2d>(), nn);
        if (e == hipSuccess) e = hipGetLastError();
    }
    for (int i = 0; i < s && e == hipSuccess; ++i) {
        e = gemm(st, res.as<v2d>(), D, res.as<v2d>(), D, tmp.as<v2d>(), D, D, D, D);
        res.swap(tmp);
    }
    return e;
}

// Makes `dev` the current device for a scope (a plan's buffers live on the device it was made on).
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != dev && hipSetDevice(dev) == hipSuccess) prev = cur;
    }
    ~DeviceScope() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    DeviceScope(const DeviceScope&) = delete;
    DeviceScope& operator=(const DeviceScope&) = delete;
};

// Rows 0 .. T-1 of X (Tp rows of D) = src, rows T .. Tp-1 = 0.
__global__ void k_fill_rows(v2d* __restrict__ X, const v2d* __restrict__ src, uint64_t T, uint64_t Tp, int D) {
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < Tp * uint64_t(D); i += uint64_t(gridDim.x) * 256)
        X[i] = (i / D < T) ? src[i % D] : v2d{0.0, 0.0};
}

}  // namespace
}  // namespace llsv

struct llsv_lplan {
    int device = 0;  // the device every buffer below lives on (current at creation)
    int dim = 0, D = 0, kind = 0, nch = 0, n_nodes = 0;
    int64_t steps = 0;
    double h = 0.0;
    llsv::DevBuf nodes;  // propagator or RK4 / drift nodes, transposed and padded (D x D each)
    llsv::DevBuf ops;    // jump operators (padded, row-major), or for ANALOG the stacked
                         // [C_1^T .. C_n^T | D_1^T .. D_n^T] (D x 2nD)
    // working buffers of llsv_lplan_run, kept (grow-only) across runs of the plan
    llsv::DevBuf psi, tmp, cur, surv, flag, dtab, psi0, w1, k1, k2, k3, k4, cpb, jl, jc, ybuf, at, bc, rs;
};

extern "C" {

int llsv_expm(int dim, const double* a, double* out) {
    using namespace llsv;
    if (dim < 1 || dim > 4096 || !a || !out) return fail(LLSV_ERR_ARG, "llsv_expm: bad arguments");
    // exp(A) by scaling and squaring: ||A / 2^s||_inf <= 1/2, 18-term Taylor series
    // (src/TrajectorySolver.h:97-111)
    double nrm = 0.0;
    bool finite = true;
    for (int i = 0; i < dim; ++i) {
        double r = 0.0;
        for (int j = 0; j < dim; ++j) r += std::hypot(a[2 * (size_t(i) * dim + j)], a[2 * (size_t(i) * dim + j) + 1]);
        finite = finite && std::isfinite(r);
        nrm = std::max(nrm, r);
    }
    if (!finite) return fail(LLSV_ERR_ARG, "llsv_expm: non-finite matrix");
    const int s = expm_power(nrm);
    const int D = ((std::max(dim, kBN) + kBN - 1) / kBN) * kBN;
    const uint64_t nn = uint64_t(D) * D;
    DevBuf b, term, res, tmp;
#define XK(expr)                                                                                      \
    do {                                                                                              \
        const hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(hip_code(e_), "llsv_expm: %s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
    XK(b.alloc(nn * 16));
    XK(term.alloc(nn * 16));
    XK(res.alloc(nn * 16));
    XK(tmp.alloc(nn * 16));
    const std::vector<double> hb = padded(a, dim, D, false);
    XK(hipMemcpy(b.p, hb.data(), nn * 16, hipMemcpyHostToDevice));
    XK(expm_dev(nullptr, b, res, term, tmp, D, s));
    std::vector<double> hr(2 * nn);
    XK(hipMemcpy(hr.data(), res.p, nn * 16, hipMemcpyDeviceToHost));
#undef XK
    for (int i = 0; i < dim; ++i)
        std::memcpy(out + 2 * size_t(i) * dim, hr.data() + 2 * size_t(i) * D, sizeof(double) * 2 * dim);
    return LLSV_OK;
}

int llsv_zgemm(int M, int N, int K, const double* a, const double* b, double* c) {
    using namespace llsv;
    if (M < 1 || N < 1 || K < 1 || !a || !b || !c) return fail(LLSV_ERR_ARG, "llsv_zgemm: bad arguments");
    if (M % kBM || N % kBN || K % kBK)
        return fail(LLSV_ERR_ARG, "llsv_zgemm: M, N must be multiples of 64 and K of 16");
    DevBuf da, db, dc;
#define GK(expr)                                                                                      \
    do {                                                                                              \
        const hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(hip_code(e_), "llsv_zgemm: %s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
    GK(da.alloc(size_t(M) * K * 16));
    GK(db.alloc(size_t(K) * N * 16));
    GK(dc.alloc(size_t(M) * N * 16));
    GK(hipMemcpy(da.p, a, size_t(M) * K * 16, hipMemcpyHostToDevice));
    GK(hipMemcpy(db.p, b, size_t(K) * N * 16, hipMemcpyHostToDevice));
    GK(gemm(nullptr, da.as<v2d>(), K, db.as<v2d>(), N, dc.as<v2d>(), N, M, N, K));
    GK(hipMemcpy(c, dc.p, size_t(M) * N * 16, hipMemcpyDeviceToHost));
    return LLSV_OK;
}

int llsv_zgemm_time(int M, int N, int K, int reps, float* ms) {
    using namespace llsv;
    if (M < 1 || N < 1 || K < 1 || reps < 1 || !ms || M % kBM || N % kBN || K % kBK)
        return fail(LLSV_ERR_ARG, "llsv_zgemm_time: bad arguments");
    DevBuf da, db, dc;
    GK(da.alloc(size_t(M) * K * 16));
    GK(db.alloc(size_t(K) * N * 16));
    GK(dc.alloc(size_t(M) * N * 16));
    GK(hipMemset(da.p, 0, size_t(M) * K * 16));
    GK(hipMemset(db.p, 0, size_t(K) * N * 16));
    hipEvent_t e0, e1;
    GK(hipEventCreate(&e0));
    GK(hipEventCreate(&e1));
    GK(gemm(nullptr, da.as<v2d>(), K, db.as<v2d>(), N, dc.as<v2d>(), N, M, N, K));  // warm-up
    GK(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) GK(gemm(nullptr, da.as<v2d>(), K, db.as<v2d>(), N, dc.as<v2d>(), N, M, N, K));
    GK(hipEventRecord(e1, nullptr));
    GK(hipEventSynchronize(e1));
    float t = 0.0f;
    GK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
#undef GK
    *ms = t / float(reps);
    return LLSV_OK;
}

int llsv_lplan_create(int dim, int kind, int64_t steps, double h, int n_nodes, const double* nodes, int n_ops,
                      const double* ops, llsv_lplan** out) {
    using namespace llsv;
    if (!out) return fail(LLSV_ERR_ARG, "llsv_lplan_create: out is null");
    *out = nullptr;
    if (dim < 1 || dim > 1024 || (dim & (dim - 1))) return fail(LLSV_ERR_ARG, "llsv_lplan_create: dim must be 2^n, n <= 10");
    if (kind < 0 || kind > 1) return fail(LLSV_ERR_ARG, "llsv_lplan_create: kind must be 0 (jumps) or 1 (diffusive)");
    if (steps < 1 || !(h > 0.0)) return fail(LLSV_ERR_ARG, "llsv_lplan_create: need steps >= 1 and h > 0");
    if (n_ops < 0 || n_ops > 16 || (n_ops && !ops)) return fail(LLSV_ERR_ARG, "llsv_lplan_create: 0..16 channels");
    const int64_t want = n_nodes == 1 ? 1 : (kind == 0 ? 2 * steps + 1 : steps);
    if (!nodes || (n_nodes != 1 && n_nodes != want))
        return fail(LLSV_ERR_ARG, "llsv_lplan_create: %d nodes given, 1 or %lld expected", n_nodes, (long long)want);
    auto p = std::make_unique<llsv_lplan>();
    if (hipGetDevice(&p->device) != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_lplan_create: no device");
    p->dim = dim;
    p->D = pad_dim(dim);
    p->kind = kind;
    p->nch = n_ops;
    p->n_nodes = n_nodes;
    p->steps = steps;
    p->h = h;
    const int D = p->D;
    const size_t blk = size_t(D) * D * 16;
#define PK(expr)                                                                                      \
    do {                                                                                              \
        const hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(hip_code(e_), "llsv_lplan_create: %s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
    PK(p->nodes.alloc(blk * size_t(n_nodes)));
    for (int k = 0; k < n_nodes; ++k) {
        const std::vector<double> m = padded(nodes + 2 * size_t(k) * dim * dim, dim, D, true);
        PK(hipMemcpy(p->nodes.as<char>() + blk * k, m.data(), blk, hipMemcpyHostToDevice));
    }
    if (kind == 0) {
        // B = [C_1^T | .. | C_n^T], D x nD: one gathered GEMM gives every channel's C_k tmp
        const int W = n_ops * D;
        std::vector<double> B(size_t(2) * D * std::max(W, 1), 0.0);
        for (int k = 0; k < n_ops; ++k) {
            const double* c = ops + 2 * size_t(k) * dim * dim;
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j) {
                    B[2 * (size_t(j) * W + size_t(k) * D + i)] = c[2 * (size_t(i) * dim + j)];
                    B[2 * (size_t(j) * W + size_t(k) * D + i) + 1] = c[2 * (size_t(i) * dim + j) + 1];
                }
        }
        PK(p->ops.alloc(B.size() * sizeof(double)));
        PK(hipMemcpy(p->ops.p, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        // B = [C_1^T | .. | C_n^T | D_1^T | .. | D_n^T], D x 2nD, D_k = C_k^+ C_k; the products
        // D_k^T = C_k^T conj(C_k) are GEMMs on the device, written into their column blocks
        const int W = 2 * n_ops * D;
        std::vector<double> B(size_t(2) * D * std::max(W, 1), 0.0);
        std::vector<double> cj(size_t(2) * D * D * std::max(n_ops, 1), 0.0);  // conj(C_k), padded
        for (int k = 0; k < n_ops; ++k) {
            const double* c = ops + 2 * size_t(k) * dim * dim;
            double* q = cj.data() + 2 * size_t(k) * D * D;
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j) {
                    // C^T[j][i] = C[i][j] in column block k
                    B[2 * (size_t(j) * W + size_t(k) * D + i)] = c[2 * (size_t(i) * dim + j)];
                    B[2 * (size_t(j) * W + size_t(k) * D + i) + 1] = c[2 * (size_t(i) * dim + j) + 1];
                    q[2 * (size_t(i) * D + j)] = c[2 * (size_t(i) * dim + j)];
                    q[2 * (size_t(i) * D + j) + 1] = -c[2 * (size_t(i) * dim + j) + 1];
                }
        }
        PK(p->ops.alloc(B.size() * sizeof(double)));
        PK(hipMemcpy(p->ops.p, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice));
        if (n_ops > 0) {
            DevBuf dcj;
            PK(dcj.alloc(cj.size() * sizeof(double)));
            PK(hipMemcpy(dcj.p, cj.data(), cj.size() * sizeof(double), hipMemcpyHostToDevice));
            v2d* stack = p->ops.as<v2d>();
            for (int k = 0; k < n_ops; ++k)
                PK(gemm(nullptr, stack + size_t(k) * D, W, dcj.as<const v2d>() + size_t(k) * D * D, D,
                        stack + size_t(n_ops + k) * D, W, D, D, D));
            PK(hipDeviceSynchronize());
        }
    }
#undef PK
    *out = p.release();
    return LLSV_OK;
}

int llsv_lplan_create_h(int dim, int kind, int64_t steps, double h, int n_h, const double* H, int n_ops,
                        const double* ops, llsv_lplan** out) {
    using namespace llsv;
    if (!out) return fail(LLSV_ERR_ARG, "llsv_lplan_create_h: out is null");
    *out = nullptr;
    if (dim < 1 || dim > 1024 || (dim & (dim - 1)))
        return fail(LLSV_ERR_ARG, "llsv_lplan_create_h: dim must be 2^n, n <= 10");
    if (kind < 0 || kind > 1) return fail(LLSV_ERR_ARG, "llsv_lplan_create_h: kind must be 0 (jumps) or 1 (diffusive)");
    if (steps < 1 || !(h > 0.0) || !std::isfinite(h))
        return fail(LLSV_ERR_ARG, "llsv_lplan_create_h: need steps >= 1 and finite h > 0");
    if (n_ops < 0 || n_ops > 16 || (n_ops && !ops))
        return fail(LLSV_ERR_ARG, "llsv_lplan_create_h: 0..16 channels");
    const int64_t want = kind == 0 ? 2 * steps + 1 : steps;
    if (!H || (n_h != 1 && n_h != want))
        return fail(LLSV_ERR_ARG, "llsv_lplan_create_h: %d Hamiltonians given, 1 or %lld expected", n_h,
                    (long long)want);
    auto p = std::make_unique<llsv_lplan>();
    if (hipGetDevice(&p->device) != hipSuccess) return fail(LLSV_ERR_DEVICE, "llsv_lplan_create_h: no device");
    p->dim = dim;
    p->D = pad_dim(dim);
    p->kind = kind;
    p->nch = n_ops;
    p->n_nodes = n_h;
    p->steps = steps;
    p->h = h;
    const int D = p->D, n = n_ops;
    const uint64_t nn = uint64_t(D) * D;
    const size_t blk = nn * 16, mat = size_t(dim) * dim * 16;
    const hipStream_t st = nullptr;
    const unsigned g = grid_of(nn);
#define PK(expr)                                                                                      \
    do {                                                                                              \
        const hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(hip_code(e_), "llsv_lplan_create_h: %s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
    // channels: BT = [C_1^T | .. | C_n^T] (jumps) or [C^T .. | (C^+ C)^T ..] (diffusive), D x W
    const int W = (kind == 0 ? 1 : 2) * n * D;
    DevBuf raw, cj, gt;
    PK(p->ops.alloc(size_t(D) * std::max(W, 1) * 16));
    if (n > 0) {
        PK(raw.alloc(mat * n));
        PK(cj.alloc(blk * n));
        PK(hipMemcpy(raw.p, ops, mat * n, hipMemcpyHostToDevice));
        if (dim < D) {
            PK(hipMemsetAsync(p->ops.p, 0, size_t(D) * W * 16, st));
            PK(hipMemsetAsync(cj.p, 0, blk * n, st));
        }
        hipLaunchKernelGGL(k_stack_ops, dim3(grid_of(uint64_t(dim) * dim * n)), dim3(256), 0, st, raw.as<const v2d>(),
                           n, dim, D, W, p->ops.as<v2d>(), cj.as<v2d>());
        PK(hipGetLastError());
        v2d* stack = p->ops.as<v2d>();
        if (kind == 1) {
            // (C_k^+ C_k)^T = C_k^T conj(C_k) into column block n + k
            for (int k = 0; k < n; ++k)
                PK(gemm(st, stack + size_t(k) * D, W, cj.as<const v2d>() + size_t(k) * nn, D, stack + size_t(n + k) * D,
                        W, D, D, D));
        } else {
            // G^T = sum_k C_k^T conj(C_k): one GEMM over the stacked channels (K = n D)
            PK(gt.alloc(blk));
            PK(gemm(st, stack, W, cj.as<const v2d>(), D, gt.as<v2d>(), D, D, D, n * D));
        }
    }
    // nodes, transposed: the jump kinds add H_nh = -(i/2) G, so -i (H + H_nh) = -i H - G / 2
    const v2d* GT = (kind == 0 && n > 0) ? gt.as<const v2d>() : nullptr;
    DevBuf hraw;
    PK(hraw.alloc(mat));
    if (n_h == 1) {
        // propagator exp(-i (H + H_nh) h)^T = exp(A^T), A = -i h H - (h / 2) G
        DevBuf a, res, term, tmp, sums;
        PK(a.alloc(blk));
        PK(res.alloc(blk));
        PK(term.alloc(blk));
        PK(tmp.alloc(blk));
        const int chunks = D / kColRows;  // D is a multiple of 64
        PK(sums.alloc(size_t(D) * chunks * sizeof(double)));
        PK(hipMemcpy(hraw.p, H, mat, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_node_t, dim3(g), dim3(256), 0, st, hraw.as<const v2d>(), dim, GT, h, -0.5 * h,
                           a.as<v2d>(), D);
        hipLaunchKernelGGL(k_colabs, dim3(unsigned((D + 255) / 256), unsigned(chunks)), dim3(256), 0, st,
                           a.as<const v2d>(), D, sums.as<double>());
        PK(hipGetLastError());
        std::vector<double> hs(size_t(D) * chunks);
        PK(hipMemcpy(hs.data(), sums.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
        double nrm = 0.0;
        bool finite = true;
        for (int j = 0; j < D; ++j) {
            double r = 0.0;
            for (int y = 0; y < chunks; ++y) r += hs[size_t(y) * D + j];
            finite = finite && std::isfinite(r);
            nrm = std::max(nrm, r);
        }
        if (!finite) return fail(LLSV_ERR_ARG, "llsv_lplan_create_h: non-finite H or channels");
        PK(expm_dev(st, a, res, term, tmp, D, expm_power(nrm)));
        if (dim < D) hipLaunchKernelGGL(k_zero_pad, dim3(g), dim3(256), 0, st, res.as<v2d>(), dim, D);
        PK(hipGetLastError());
        p->nodes.swap(res);
    } else {
        // RK4 / Euler nodes -i (H(t_j) + H_nh), transposed
        PK(p->nodes.alloc(blk * size_t(n_h)));
        for (int j = 0; j < n_h; ++j) {
            PK(hipMemcpy(hraw.p, H + 2 * size_t(j) * dim * dim, mat, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_node_t, dim3(g), dim3(256), 0, st, hraw.as<const v2d>(), dim, GT, 1.0, -0.5,
                               p->nodes.as<v2d>() + size_t(j) * nn, D);
            PK(hipGetLastError());
        }
    }
    PK(hipDeviceSynchronize());
#undef PK
    *out = p.release();
    return LLSV_OK;
}

void llsv_lplan_destroy(llsv_lplan* p) {
    if (!p) return;
    llsv::DeviceScope on(p->device);
    delete p;
}

int llsv_lplan_run(llsv_lplan* p, const double* psi0, int64_t n_traj, int64_t first_traj, uint64_t seed, int rng,
                   double* states, double* rho_sum) {
    using namespace llsv;
    if (!p) return fail(LLSV_ERR_HANDLE, "llsv_lplan_run: null plan");
    DeviceScope on(p->device);
    if (!psi0 || (!states && !rho_sum) || n_traj < 1 || first_traj < 0 || (rng != 0 && rng != 1))
        return fail(LLSV_ERR_ARG, "llsv_lplan_run: bad arguments");
    const int dim = p->dim, D = p->D, nch = p->nch;
    const uint64_t T = uint64_t(n_traj), Tp = (T + kBM - 1) / kBM * kBM;
    const int64_t S = p->steps;
    const hipStream_t st = nullptr;
#define RK(expr)                                                                                      \
    do {                                                                                              \
        const hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(hip_code(e_), "llsv_lplan_run: %s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
    // the reference's per-trajectory streams, drawn on the host by worker threads
    std::vector<double> table;
    uint64_t stride = 0;
    if (rng == 0 && nch > 0) {
        stride = p->kind == 0 ? uint64_t(2 * S) : uint64_t(S) * nch;
        if (double(stride) * double(T) * 8.0 > 2.0e9)
            return fail(LLSV_ERR_ARG, "llsv_lplan_run: random-number table over 2 GB; run fewer trajectories per call");
        table.resize(size_t(stride * T));
        const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned w = 0; w < nth; ++w)
            pool.emplace_back([&, w] {
                for (uint64_t t = w; t < T; t += nth) {
                    std::mt19937_64 mt(splitmix64(seed ^ splitmix64(uint64_t(first_traj) + t + 1)));
                    double* o = table.data() + t * stride;
                    if (p->kind == 0) {
                        std::uniform_real_distribution<double> uni(0.0, 1.0);
                        for (uint64_t c = 0; c < stride; ++c) o[c] = uni(mt);
                    } else {
                        std::normal_distribution<double> gauss(0.0, std::sqrt(p->h));
                        for (uint64_t c = 0; c < stride; ++c) o[c] = gauss(mt);
                    }
                }
            });
        for (auto& th : pool) th.join();
    }
    const size_t sbytes = size_t(Tp) * D * 16;
    DevBuf &psi = p->psi, &tmp = p->tmp, &cur = p->cur, &surv = p->surv, &flag = p->flag, &dtab = p->dtab,
           &w1 = p->w1, &k1 = p->k1, &k2 = p->k2, &k3 = p->k3, &k4 = p->k4, &cpb = p->cpb, &jl = p->jl, &jc = p->jc,
           &ybuf = p->ybuf;
    RK(psi.need(sbytes));
    RK(tmp.need(sbytes));
    RK(cur.need(sizeof(uint64_t) * T));
    RK(surv.need(sizeof(double) * T));
    RK(flag.need(sizeof(int)));
    RK(p->psi0.need(size_t(D) * 16));
    RK(hipMemset(cur.p, 0, sizeof(uint64_t) * T));
    RK(hipMemset(flag.p, 0, sizeof(int)));
    {
        std::vector<double> h0(size_t(2) * D, 0.0);  // psi0 padded to D amplitudes, broadcast on the device
        std::memcpy(h0.data(), psi0, sizeof(double) * 2 * dim);
        RK(hipMemcpy(p->psi0.p, h0.data(), h0.size() * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_fill_rows, dim3(grid_of(Tp * D)), dim3(256), 0, st, psi.as<v2d>(), p->psi0.as<const v2d>(), T,
                           Tp, D);
        RK(hipGetLastError());
    }
    if (!table.empty()) {
        RK(dtab.need(table.size() * sizeof(double)));
        RK(hipMemcpy(dtab.p, table.data(), table.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    Draws dr{table.empty() ? nullptr : dtab.as<const double>(), stride, seed, uint64_t(first_traj)};
    // jumps: the list of this step's jumping trajectories and the gathered products, processed in
    // rounds of at most R rows (Y of R x nD entries, at most 512 MiB)
    uint32_t R = 0;
    if (p->kind == 0 && nch > 0) {
        RK(jl.need(sizeof(uint32_t) * T));
        RK(jc.need(sizeof(uint32_t)));
        const uint64_t row_bytes = uint64_t(nch) * D * 16;
        R = uint32_t(std::min<uint64_t>(Tp, std::max<uint64_t>(kBM, ((uint64_t(512) << 20) / row_bytes) / kBM * kBM)));
        RK(ybuf.need(size_t(R) * row_bytes));
    }
    const uint64_t nn = Tp * D;
    const size_t blk = size_t(D) * D;
    const v2d* nodes = p->nodes.as<const v2d>();
    const bool constant = p->n_nodes == 1;
    if (p->kind == 0 && !constant) {
        RK(w1.need(sbytes));
        RK(k1.need(sbytes));
        RK(k2.need(sbytes));
        RK(k3.need(sbytes));
        RK(k4.need(sbytes));
    }
    if (p->kind == 1) {
        RK(cpb.need(size_t(Tp) * std::max(1, 2 * nch) * D * 16));
        RK(w1.need(sbytes));
    }
    const double h = p->h;
    for (int64_t s = 0; s < S; ++s) {
        if (p->kind == 0) {
            if (constant) {
                RK(gemm(st, psi.as<v2d>(), D, nodes, D, tmp.as<v2d>(), D, int(Tp), D, D));
            } else {
                // RK4 over H(t) sampled at the step's endpoints and midpoint (src/TrajectorySolver.h:386-396)
                const v2d *a0 = nodes + blk * size_t(2 * s), *a1 = nodes + blk * size_t(2 * s + 1),
                          *a2 = nodes + blk * size_t(2 * s + 2);
                v2d* P = psi.as<v2d>();
                RK(gemm(st, P, D, a0, D, k1.as<v2d>(), D, int(Tp), D, D));
                hipLaunchKernelGGL(k_axpy, dim3(grid_of(nn)), dim3(256), 0, st, w1.as<v2d>(), P, k1.as<const v2d>(), 0.5 * h, nn);
                RK(gemm(st, w1.as<v2d>(), D, a1, D, k2.as<v2d>(), D, int(Tp), D, D));
                hipLaunchKernelGGL(k_axpy, dim3(grid_of(nn)), dim3(256), 0, st, w1.as<v2d>(), P, k2.as<const v2d>(), 0.5 * h, nn);
                RK(gemm(st, w1.as<v2d>(), D, a1, D, k3.as<v2d>(), D, int(Tp), D, D));
                hipLaunchKernelGGL(k_axpy, dim3(grid_of(nn)), dim3(256), 0, st, w1.as<v2d>(), P, k3.as<const v2d>(), h, nn);
                RK(gemm(st, w1.as<v2d>(), D, a2, D, k4.as<v2d>(), D, int(Tp), D, D));
                hipLaunchKernelGGL(k_rk4, dim3(grid_of(nn)), dim3(256), 0, st, tmp.as<v2d>(), P, k1.as<const v2d>(),
                                   k2.as<const v2d>(), k3.as<const v2d>(), k4.as<const v2d>(), h / 6.0, nn);
                RK(hipGetLastError());
            }
            if (nch > 0) RK(hipMemsetAsync(jc.p, 0, sizeof(uint32_t), st));
            hipLaunchKernelGGL(k_jump_decide, dim3(unsigned((T + 3) / 4)), dim3(256), 0, st, tmp.as<const v2d>(),
                               psi.as<v2d>(), D, T, nch, dr, cur.as<uint64_t>(), surv.as<double>(), flag.as<int>(),
                               jl.as<uint32_t>(), jc.as<uint32_t>());
            RK(hipGetLastError());
            for (uint32_t r0 = 0; nch > 0 && uint64_t(r0) < T; r0 += R) {
                // rounds past the device-side count return at once
                RK(gemm_gather(st, tmp.as<const v2d>(), D, jl.as<const uint32_t>(), jc.as<const uint32_t>(), r0,
                               p->ops.as<const v2d>(), nch * D, ybuf.as<v2d>(), nch * D, int(R), nch * D, D));
                hipLaunchKernelGGL(k_jump_finish, dim3((R + 3) / 4), dim3(256), 0, st, ybuf.as<const v2d>(),
                                   tmp.as<const v2d>(), psi.as<v2d>(), D, nch, jl.as<const uint32_t>(),
                                   jc.as<const uint32_t>(), r0, R, dr, cur.as<uint64_t>(), surv.as<const double>());
                RK(hipGetLastError());
            }
        } else {
            const v2d* drift = nullptr;
            if (constant) {
                // exact Hamiltonian step, zero drift (src/TrajectorySolver.h:437-441)
                RK(gemm(st, psi.as<v2d>(), D, nodes, D, tmp.as<v2d>(), D, int(Tp), D, D));
                psi.swap(tmp);
            } else {
                RK(gemm(st, psi.as<v2d>(), D, nodes + blk * size_t(s), D, w1.as<v2d>(), D, int(Tp), D, D));
                drift = w1.as<const v2d>();
            }
            if (nch > 0) RK(gemm(st, psi.as<v2d>(), D, p->ops.as<const v2d>(), 2 * nch * D, cpb.as<v2d>(), 2 * nch * D,
                                 int(Tp), 2 * nch * D, D));
            hipLaunchKernelGGL(k_diffuse, dim3(unsigned((T + 3) / 4)), dim3(256), 0, st, psi.as<v2d>(),
                               cpb.as<const v2d>(), drift, D, T, nch, h, std::sqrt(h),
                               table.empty() ? nullptr : dtab.as<const double>(), stride, uint64_t(s), dr, flag.as<int>());
            RK(hipGetLastError());
        }
    }
    int hflag = 0;
    RK(hipMemcpy(&hflag, flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (hflag & 1) return fail(LLSV_ERR_NUMERIC, "TrajectorySolver: state norm vanished; reduce dt");
    if (hflag & 2) return fail(LLSV_ERR_NUMERIC, "TrajectorySolver: diffusion state diverged; reduce dt");
    if (states) {
        std::vector<double> hs(size_t(2) * D * T);  // trajectories as rows of D amplitudes
        RK(hipMemcpy(hs.data(), psi.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (uint64_t t = 0; t < T; ++t) std::memcpy(states + 2 * t * dim, hs.data() + 2 * t * D, sizeof(double) * 2 * dim);
    }
    if (rho_sum) {
        // sum_t psi_t psi_t^+ = Psi^T conj(Psi) (trajectories as rows; padded rows are zero)
        DevBuf &at = p->at, &bc = p->bc, &rs = p->rs;
        RK(at.need(sbytes));
        RK(bc.need(sbytes));
        RK(rs.need(size_t(D) * D * 16));
        hipLaunchKernelGGL(k_transpose_conj, dim3(unsigned((Tp * D + 255) / 256)), dim3(256), 0, st, psi.as<const v2d>(),
                           at.as<v2d>(), bc.as<v2d>(), Tp, D);
        RK(hipGetLastError());
        RK(gemm(st, at.as<const v2d>(), int(Tp), bc.as<const v2d>(), D, rs.as<v2d>(), D, D, D, int(Tp)));
        std::vector<double> hr(size_t(2) * D * D);
        RK(hipMemcpy(hr.data(), rs.p, hr.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < dim; ++i) std::memcpy(rho_sum + 2 * size_t(i) * dim, hr.data() + 2 * size_t(i) * D, sizeof(double) * 2 * dim);
    }
#undef RK
    return LLSV_OK;
}

}  // extern "C"

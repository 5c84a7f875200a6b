/*
 * llsv.h -- C ABI of the MI355X (gfx950) complex128 state-vector engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, no torch or HIP types in
 * the signatures (streams are passed as void*).  Every function returns
 * LLSV_OK (0) on success or a negative LLSV_ERR_* code; the message of the last
 * failure on the calling thread is llsv_last_error().  No C++ exception crosses
 * this boundary.
 *
 * Conventions (the reference's, src/Backend.h:3-8):
 *   - a state of n qubits is 2^n complex128 amplitudes, interleaved (re, im);
 *   - qubit q is bit q of the basis index (little endian);
 *   - a one-qubit gate is the row-major 2x2 g = {g00, g01, g10, g11}, passed as
 *     8 doubles {g00re, g00im, g01re, g01im, g10re, g10im, g11re, g11im}:
 *     a0' = g00 a0 + g01 a1, a1' = g10 a0 + g11 a1 (src/StateVectorOptimized.cpp:71-75);
 *   - a dense k-qubit block is 2^k x 2^k complex row-major; operator bit j <->
 *     targets[j] (src/Noise.h:58-65).
 *
 * Work is asynchronous on the state's own HIP stream; calls that return host
 * values (norm, expectations, copies) synchronise that stream.  A state must not
 * be used from two threads at once (as in the reference, which holds the GIL).
 */
#ifndef LLSV_H
#define LLSV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LLSV_OK 0
#define LLSV_ERR_ARG (-1)    /* bad argument: the reference's std::invalid_argument / runtime_error */
#define LLSV_ERR_NOMEM (-2)  /* device or pinned-host allocation failed */
#define LLSV_ERR_DEVICE (-3) /* HIP runtime error */
#define LLSV_ERR_HANDLE (-4) /* null handle */
#define LLSV_ERR_NUMERIC (-5) /* numerical breakdown (the reference's std::runtime_error, e.g. a vanishing norm) */

#define LLSV_MAX_QUBITS 34   /* 2^34 x 16 B = 256 GiB, the most one 288 GB HBM3E device holds */
#define LLSV_MAX_DENSE_K 5   /* largest fused block: 32 x 32 */

typedef struct llsv_state llsv_state;

/* ---- lifetime ------------------------------------------------------------ */

/* New state |0...0> on the current HIP device.
 * Replaces StateVectorOptimized(int n), src/StateVectorOptimized.cpp:17-24. */
int llsv_create(int n_qubits, llsv_state** out);

/* New state holding a copy of n_amps host amplitudes; n_amps must be a power of two.
 * Replaces StateVectorOptimized(py::array_t<complex<double>>), src/StateVectorOptimized.cpp:26-47
 * (the reference borrows the buffer; here the device copy is the state and the caller
 * keeps host and device coherent with llsv_copy_from_host / llsv_copy_to_host). */
int llsv_create_from_host(const double* amps, uint64_t n_amps, llsv_state** out);

/* Replaces ~StateVectorOptimized, src/StateVectorOptimized.cpp:49-53.  Null is a no-op. */
int llsv_destroy(llsv_state* s);

/* Back to |0...0>.  Replaces ll::Backend::reset (src/Backend.h:30) for the same n. */
int llsv_reset(llsv_state* s);

/* get_num_qubits / get_size, src/StateVectorOptimized.cpp:108-109.  -1 / 0 on a null handle. */
int llsv_num_qubits(const llsv_state* s);
uint64_t llsv_size(const llsv_state* s);

/* ---- gates ---------------------------------------------------------------- */

/* Dense one-qubit gate on `target` in [0, n).
 * Replaces StateVectorOptimized::apply_gate_cache_optimized (src/StateVectorOptimized.cpp:56-78),
 * apply_gate_standard (:81-97) and the v2 apply_gate_optimized (src/bindings_v2.cpp:56-72). */
int llsv_apply_1q(llsv_state* s, int target, const double g[8]);

/* CNOT: swap amplitudes i <-> i|t for control bit 1, target bit 0.
 * Replaces ll::Backend::cnot (src/Backend.h:90-97, src/StateVectorBackend.h:109-135);
 * documented for v3 at README:158 but never bound there. */
int llsv_apply_cnot(llsv_state* s, int control, int target);

/* CZ = mcz({a, b}) (src/Backend.h:7). */
int llsv_apply_cz(llsv_state* s, int a, int b);

/* Negate every amplitude whose bits on `qubits` are all 1.
 * Replaces ll::Backend::mcz (src/Backend.h:99-110, src/StateVectorBackend.h:137-151). */
int llsv_apply_mcz(llsv_state* s, const int* qubits, int n_qubits);

/* Dense k-qubit block, 1 <= k <= LLSV_MAX_DENSE_K, m = 2^k x 2^k complex row-major
 * (2 * 4^k doubles).  The apply step for the fused blocks of the fusion pre-pass
 * (python/gate_fusion.py:15-37 emits only 2x2; k > 1 blocks are new). */
int llsv_apply_dense(llsv_state* s, int k, const int* targets, const double* m);

/* One op of a gate program (a whole circuit handed over at once). */
typedef enum {
    LLSV_OP_1Q = 0,    /* q[0] = target, m[8] = the 2x2 gate */
    LLSV_OP_CNOT = 1,  /* q[0] = control, q[1] = target */
    LLSV_OP_CZ = 2,    /* q[0], q[1] */
    LLSV_OP_MCZ = 3,   /* q[0..nq-1] */
    LLSV_OP_DENSE = 4  /* q[0..nq-1] = targets, dense = 2 * 4^nq doubles */
} llsv_op_kind;

typedef struct {
    int32_t kind;      /* llsv_op_kind */
    int32_t nq;        /* qubits used in q[] */
    int32_t q[8];
    double m[8];       /* LLSV_OP_1Q matrix */
    const double* dense; /* LLSV_OP_DENSE matrix (host memory), else NULL */
} llsv_op;

/* Apply ops[0..n_ops) in order.  The engine groups ops into HBM passes: gates whose
 * qubits fit one on-chip tile are applied together in LDS during a single read and
 * write of the state.  Results equal op-by-op application up to rounding.
 * Replaces the per-gate Python loop of benchmarks/benchmark_fusion.py:49-50. */
int llsv_apply_program(llsv_state* s, const llsv_op* ops, uint64_t n_ops);

/* The pass plan llsv_apply_program would run for this circuit with tiles of 2^tile_bits
 * amplitudes (host only, no device needed): *n_passes gets the number of HBM passes;
 * order[0..n_ops), if not null, the op indices in execution order; pass_end[p], if not
 * null, the end of pass p in order[] (size n_ops is always enough).  Ops are only moved
 * past ops they share no qubit with. */
int llsv_plan_passes(int n_qubits, const llsv_op* ops, uint64_t n_ops, int tile_bits, uint64_t* n_passes,
                     uint64_t* order, uint64_t* pass_end);

/* Tile bits llsv_apply_program uses on this state's device (from the LDS per workgroup). */
int llsv_program_tile_bits(llsv_state* s);

/* Pass compiler of llsv_apply_program (no counterpart in the reference, which applies gates
 * one call at a time).  A register-resident pass (12-bit tiles) can run on a kernel generated
 * for its op sequence and compiled once per process with hipRTC (about a second per distinct
 * pass structure; coefficients stay run-time data, so changed angles reuse the kernel), or on
 * the built-in interpreter kernel.  Both give the same amplitudes.  mode 0 = interpreter only,
 * 1 = passes of states with >= 20 qubits are compiled in a background thread and interpreted
 * until their kernel is ready (the default; also set by the environment variable
 * LLSV_PROGRAM_JIT = 0, 1, 2 before first use), 2 = every eligible pass is compiled before its
 * first launch.
 * Passes holding dense blocks or more than 256 ops stay on the interpreter.  Returns the
 * previous mode. */
int llsv_set_program_jit(int mode);

/* Process-wide counters of the pass compiler: distinct pass kernels compiled, launches of
 * compiled kernels, compilations that failed (those passes ran on the interpreter). */
void llsv_program_jit_stats(uint64_t* compiled, uint64_t* launches, uint64_t* failures);

/* Compiles, without a device, the pass kernels llsv_apply_program would use for this program
 * on an n-qubit state with 12-bit tiles under the current mode, so a later call does not wait
 * for them.  Returns the number of passes that have a compiled kernel, or an error code. */
int llsv_program_jit_precompile(int n_qubits, const llsv_op* ops, uint64_t n_ops);

/* Blocks until the pass compiler's background thread (automatic mode) has compiled every
 * pass requested so far. */
void llsv_program_jit_wait(void);

/* ---- readout ---------------------------------------------------------------- */

/* sum_i |psi_i|^2 (fixed-order reduction: bitwise reproducible run to run). */
int llsv_norm2(llsv_state* s, double* out);

/* <Z_q> = P(bit q = 0) - P(bit q = 1). */
int llsv_expect_z(llsv_state* s, int q, double* out);

/* |psi_i|^2 into host out[0..n) (n = 2^n_qubits).
 * Replaces StateVectorBackend::probabilities, src/StateVectorBackend.h:153-161. */
int llsv_probabilities(llsv_state* s, double* out, uint64_t n);

/* Copies of the amplitudes (2 * n_amps doubles).  n_amps must equal the state size.
 * copy_to_host replaces ll::Backend::state (src/Backend.h:44) and backs get_numpy_view
 * (src/StateVectorOptimized.cpp:99-106). */
int llsv_copy_to_host(llsv_state* s, double* amps, uint64_t n_amps);
int llsv_copy_from_host(llsv_state* s, const double* amps, uint64_t n_amps);

/* Wait for all queued work on the state's stream. */
int llsv_sync(llsv_state* s);

/* dst := src (same qubit count), device to device; ordered after work queued on both
 * streams.  No reference counterpart: ll::Backend::state() (src/Backend.h:44) is the only
 * way the reference copies a state, through the host. */
int llsv_copy_state(llsv_state* dst, llsv_state* src);

/* ---- batched parameterised circuits: energies and adjoint gradients ---------- */

/* One gate of a parameterised circuit.  RX/RY/RZ take their angle from the parameter
 * vector (params[b * n_params + param], b = batch index); LLSV_PG_1Q is a fixed 2x2 in m. */
typedef enum {
    LLSV_PG_RX = 0,
    LLSV_PG_RY = 1,
    LLSV_PG_RZ = 2,
    LLSV_PG_1Q = 3,   /* q0 = target, m = row-major 2x2 */
    LLSV_PG_CNOT = 4, /* q0 = control, q1 = target */
    LLSV_PG_CZ = 5    /* q0, q1 */
} llsv_pgate_kind;

typedef struct {
    int32_t kind;
    int32_t q0, q1;
    int32_t param; /* parameter index for RX/RY/RZ, else ignored */
    double m[8];
} llsv_pgate;

/* Term coeff * P of a Hamiltonian H = sum of terms; P = prod over qubits of X (x bit only),
 * Y (x and z bits) or Z (z bit only). */
typedef struct {
    double coeff;
    uint64_t xmask, zmask;
} llsv_pauli;

/* Device workspace llsv_pc_adjoint needs for n qubits, `batch` parameter vectors and a circuit
 * of n_gates gates: two batches of states (psi, lambda) plus reduction partials.
 * llsv_pc_energy needs less (one batch of states) and accepts this size too. */
uint64_t llsv_pc_workspace_bytes(int n_qubits, int batch, int n_gates);

/* energies[b] = <psi_b|H|psi_b>, psi_b = U(params_b)|0...0>, for b < batch.  params (batch x
 * n_params doubles), energies (batch doubles) and workspace are device pointers; work is
 * queued on `stream` (a hipStream_t, NULL = default stream).  The energy of the reference's
 * VQE loop, E(theta) (src/Algorithms.h:205-213), evaluated for a whole batch at once. */
int llsv_pc_energy(int n_qubits, int batch, const llsv_pgate* gates, int n_gates, int n_params,
                   const double* params, const llsv_pauli* terms, int n_terms, double* energies,
                   void* workspace, uint64_t workspace_bytes, void* stream);

/* Energies and gradients grads[b * n_params + p] = dE_b / dparam_p by the adjoint method
 * (one forward sweep, lambda = H psi, one backward sweep applying U_k^dagger to psi and
 * lambda; dE/dtheta_k = Im <lambda_k| G_k |psi_k> for U_k = exp(-i theta_k G_k / 2)).
 * Replaces the parameter-shift loop of VQE::optimize (src/Algorithms.h:277-292), which costs
 * 2 n_params circuit evaluations per gradient. */
int llsv_pc_adjoint(int n_qubits, int batch, const llsv_pgate* gates, int n_gates, int n_params,
                    const double* params, const llsv_pauli* terms, int n_terms, double* energies,
                    double* grads, void* workspace, uint64_t workspace_bytes, void* stream);

/* Compiles, without a device, the tiled sweep kernels llsv_pc_energy / llsv_pc_adjoint use for
 * this gate list (n >= 20 in the automatic mode of llsv_set_program_jit, n >= 12 when it is 2):
 * forward and backward passes of 12-qubit register tiles, so the first batched call does not
 * wait for them.  Returns the number of compiled passes (0 when the sweeps run gate by gate),
 * or an error code; *forward_passes (may be NULL) receives the forward sweep's share.
 * No counterpart in the reference. */
int llsv_pc_precompile(int n_qubits, const llsv_pgate* gates, int n_gates, int* forward_passes);

/* ---- readout ------------------------------------------------------------------ */

/* <psi| sum_t terms[t].coeff P_t |psi> for Pauli strings P_t (llsv_pauli masks), in one pass
 * per 16 terms with a fixed-shape reduction (bitwise reproducible).  per_term (may be NULL)
 * receives the n_terms contributions, *total their sum.  Generalises the reference's
 * Z-only readout (probabilities + NumPy, README:162) to any Pauli sum; no GPU counterpart
 * in the reference. */
int llsv_expect_pauli(llsv_state* s, const llsv_pauli* terms, int n_terms, double* per_term, double* total);

/* Seeded measurement sampling in the computational basis: out[k] (host, shots entries) is the
 * smallest basis index i with CDF(i) > u_k * CDF(last), CDF(i) = sum_{j <= i} |psi_j|^2 and
 * u_k = (x >> 11) 2^-53, x the first 64 bits of Philox4x32-10(key = seed, counter = (k low,
 * k high, 0x53414D50, 0)).  Indices with zero probability are never returned; a state whose total
 * probability is zero or not finite is LLSV_ERR_NUMERIC (out is then unspecified).  The state is
 * unchanged.  No counterpart in the reference (its scripts sample from probabilities() in
 * NumPy). */
int llsv_sample(llsv_state* s, uint64_t shots, uint64_t seed, uint64_t* out);

/* Device scratch held by llsv_sample_noalloc on each GPU it runs on, until the process ends. */
#define LLSV_SAMPLE_SCRATCH_BYTES (32ull << 20)

/* llsv_sample without any allocation: the same draws and the same indices, with the chunk sums,
 * prefixes and sampled indices kept in a fixed device array of LLSV_SAMPLE_SCRATCH_BYTES per GPU
 * instead of a buffer allocated per call.  Needs 16 * max(1, 2^n / 4096) + 64 + 8 * shots <=
 * LLSV_SAMPLE_SCRATCH_BYTES (so n <= 32; at 30 qubits up to 3.6 M shots per call), else
 * LLSV_ERR_ARG.  Calls no hipMalloc, hipFree or other allocating routine.  All calls on one GPU
 * share its scratch, which keeps the last call's contents: calls that overlap on one GPU (two
 * threads, or states with different streams) race and must be serialised by the caller.  A
 * zero or non-finite total probability is LLSV_ERR_NUMERIC (as for llsv_sample).  No counterpart
 * in the reference; for callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation. */
int llsv_sample_noalloc(llsv_state* s, uint64_t shots, uint64_t seed, uint64_t* out);

/* ---- batched noisy trajectories ---------------------------------------------- */

/* A gate of a noisy circuit: after it, every qubit it touches goes through amplitude damping
 * AD(gamma_q) and then phase damping PD(lambda_q) (Kraus sets of src/Noise.h:114-122),
 * unravelled by sampling: jump iff u >= ||K0 psi||^2 (src/TrajectorySolver.h:425). */
typedef struct {
    int32_t kind; /* 0 = one-qubit gate m (row-major 2x2), 1 = CNOT(q0 control, q1 target) */
    int32_t q0, q1, pad;
    double m[8];
} llsv_tgate;

/* One trajectory: a circuit (gates[first_gate .. first_gate + n_gates)) on n_qubits qubits with
 * noise rates gamma[first_rate + q], lambda[first_rate + q], and its Philox4x32-10 stream:
 * key = seed, counter = (traj low, traj high, gate index, site), site 0/1 = AD/PD on the gate's
 * first qubit, 2/3 on its second; u = (x >> 11) 2^-53 with x = out0 | out1 << 32. */
typedef struct {
    int32_t n_qubits; /* 1..20 */
    int32_t first_gate;
    int32_t n_gates;
    int32_t first_rate;
    uint64_t traj;
} llsv_traj_item;

#define LLSV_TRAJ_ZSTRIDE 32

/* Runs every item in one persistent launch (workgroups take items from a work counter, the
 * largest first) and returns, in host memory: z[i * LLSV_TRAJ_ZSTRIDE + q] = <Z_q> of item i's
 * final state; jumps[i] (may be NULL) the number of jumps; hashes[i] (may be NULL) an FNV-1a
 * hash of (gate, site) of every jump, for checking decisions against a CPU run; state (may be
 * NULL) the final 2^n amplitudes of item 0.  Synchronous.
 * Batches TrajectorySolver::estimate_observable (src/TrajectorySolver.h:239-280) at the gate
 * level, beyond its 10-qubit cap. */
int llsv_traj_run(const llsv_traj_item* items, int n_items, const llsv_tgate* gates, int n_gates,
                  const double* gamma, const double* lambda, int n_rates, uint64_t seed, double* z,
                  uint32_t* jumps, uint64_t* hashes, double* state);

/* A batch prepared for repeated runs (no counterpart in the reference, whose TrajectorySolver
 * samples one trajectory per call; for callers that re-run a batch with new seeds).  Uploads
 * items, gates and rates once and chooses the launch configuration: threads per trajectory
 * (256, 512 or 1024), persistent workgroups per CU (1, 2 or 4) and the fused mode: 0, every
 * gate and noise operator applied in a pass of its own; 1, operators of sites that cannot jump
 * deferred and multiplied into the next pass on their qubit, CNOTs with no possible jump run two
 * per pass; 2, the same with such CNOTs held in lists that run in one pass through LDS tiles of
 * the state (at most two workgroups per CU).  tune = 1: the autotuner times every candidate on a sample of this batch (about
 * 2048 items) and keeps the fastest, cached per batch shape (qubit range, log2 of the item
 * count and of the gates per item) for the process, so llsv_traj_run uses it too; tune = 0:
 * the cached choice or the default; tune = -1: the default (llsv_traj_default_config).  One run
 * (seed upload, work-counter reset, the persistent kernel, result copies) is captured in a
 * hipGraph. */
typedef struct llsv_traj_plan llsv_traj_plan;
int llsv_traj_plan_create(const llsv_traj_item* items, int n_items, const llsv_tgate* gates, int n_gates,
                          const double* gamma, const double* lambda, int n_rates, int tune,
                          llsv_traj_plan** out);

/* Replays the plan's graph with this seed; results as llsv_traj_run.  Synchronous. */
int llsv_traj_plan_run(llsv_traj_plan* p, uint64_t seed, double* z, uint32_t* jumps, uint64_t* hashes);

/* The plan's configuration: threads per trajectory, workgroups in the persistent launch,
 * fused mode (0 / 1 / 2) and the time the autotuner took in ms (0 when the choice came from the cache). */
int llsv_traj_plan_config(const llsv_traj_plan* p, int* threads, int* workgroups, int* fused, double* tune_ms);
int llsv_traj_plan_destroy(llsv_traj_plan* p);

/* Sets the process-wide default configuration (initially 256 threads, 4 workgroups per CU,
 * fused mode 1): threads in {256, 512, 1024}, per_cu in {1, 2, 4}, threads * per_cu <= 2048,
 * fused in {0, 1, 2}. */
int llsv_traj_default_config(int threads, int per_cu, int fused);

/* The tuner's choices (batch shape -> threads, workgroups per CU, fused) as a JSON file, so a
 * later process skips the tuning runs: load adds a file's entries to the process cache (entries
 * for the same shape are replaced; a malformed file adds nothing and is LLSV_ERR_ARG); save writes
 * the cache; entries counts it.  With LLSV_TUNE_CACHE=<file> in the environment the library loads
 * that file on first use and rewrites it after every new choice; LLSV_TRAJ_CONFIG=t,p,f sets the
 * default configuration and makes plans use it instead of tuning.  No counterpart in the
 * reference: its CPU loops have no launch configuration, only compile-time constants and
 * OMP_NUM_THREADS (src/StateVectorBackend.h). */
int llsv_traj_tune_load(const char* path);
int llsv_traj_tune_save(const char* path);
int llsv_traj_tune_entries(void);

/* ---- zero-noise extrapolation: folded circuits under sampled depolarizing noise ---- */
/* The noise model of ll::NoisyBackend (src/Mitigation.h:196-303): every logical gate U runs as
 * U (U^+ U)^k for scale 2k + 1, and after each application the touched qubits get depolarizing
 * noise of total Pauli probability p, sampled per site: one channel for a 1-qubit gate, one
 * 2-qubit channel for cnot and a 2-qubit mcz (Pauli index 1..15, low two bits on the first
 * qubit), independent 1-qubit channels on each qubit of a wider mcz.  A Pauli Y is applied as
 * Z then X, like the reference.  Trajectories run in one launch, each state in LDS. */
#define LLSV_ZG_H 0
#define LLSV_ZG_X 1
#define LLSV_ZG_RX 2
#define LLSV_ZG_RY 3
#define LLSV_ZG_RZ 4
#define LLSV_ZG_CNOT 5 /* q[0] = control, q[1] = target */
#define LLSV_ZG_MCZ 6  /* q[0..nq) */
#define LLSV_ZNE_MAX_QUBITS 20 /* up to 12 qubits a trajectory's state lives in LDS, above in a device-memory slot */

typedef struct llsv_zgate {
    int32_t kind; /* LLSV_ZG_* */
    int32_t nq;   /* qubits in q (1 for one-qubit gates, 2 for cnot) */
    int8_t q[16];
    double theta; /* rx / ry / rz angle */
} llsv_zgate;

/* The sampler of one NoisyBackend: p, the random stream and its position.
 * rng 0: the reference's stream -- std::mt19937_64 seeded with `seed`, u = (x >> 11) 2^-53 per
 *        draw, one draw per channel and one more for the Pauli when u < p, consumed in order
 *        (trajectory by trajectory, site by site), so the same seed and circuit give the same
 *        Pauli insertions as ll::NoisyBackend; the draws are made on the host.
 * rng 1: counter-based Philox4x32-10 on the device, counter (trajectory, site, slot); runs
 *        continue the trajectory count, so successive runs draw fresh noise.
 * Replaces ll::NoisyBackend's constructor (src/Mitigation.h:118-121): p outside [0, 1] (NaN
 * included) is LLSV_ERR_ARG. */
typedef struct llsv_noisy llsv_noisy;
int llsv_noisy_create(double p, uint64_t seed, int rng, llsv_noisy** out);
void llsv_noisy_destroy(llsv_noisy* z);
/* NoisyBackend::reseed (src/Mitigation.h:123): restarts the stream (rng 1: the trajectory count). */
int llsv_noisy_reseed(llsv_noisy* z, uint64_t seed);

/* Draws one noise site of the reference stream (rng 0; NoisyBackend::noise,
 * src/Mitigation.h:254-262), for a gate on nq qubits: *code gets two bits per channel slot
 * (0 none, 1 X, 2 Y, 3 Z; slot i = i-th qubit of the gate).  For callers that apply gates one
 * at a time. */
int llsv_noisy_site(llsv_noisy* z, int nq, uint32_t* code);

/* Runs n_traj trajectories of the logical circuit `gates` at noise scale `scale` (odd, 1..99)
 * from |0...0> on n <= LLSV_ZNE_MAX_QUBITS qubits, n_gates * scale <= 2^22 gate applications (a
 * state of up to 12 qubits lives in a workgroup's LDS, a larger one in a device-memory slot of the
 * library: at most 4 GiB of slots, kept on the device for later calls until the process ends).
 * The folded circuit's tables take up to 104 bytes of device memory per gate application (0.44 GB
 * at the cap); they are kept for the next call while they hold at most 64 MiB, and freed at the
 * end of a call above that.  Every gate's qubits number at most 16 (the noise code of a site).
 * weights (host, 2^n doubles, or NULL): values[t]
 * = sum_i weights[i] |psi_t[i]|^2 (host, n_traj doubles).  amps (host, n_traj * 2^n complex128
 * interleaved, or NULL): the final states.  With p = 0 no draws are made.  Replaces the
 * trajectory loop of ZNE::extrapolate (src/Mitigation.h:355-376) for one scale. */
int llsv_noisy_run(llsv_noisy* z, int n, const llsv_zgate* gates, int n_gates, int scale, int64_t n_traj,
                   const double* weights, double* values, double* amps);

/* ---- dense Lindblad trajectories (<= 10 qubits) ------------------------------------ */
/* The trajectory solver of src/TrajectorySolver.h: many trajectories of a dim = 2^n state
 * (n <= 10) under a dense propagator, stepped together -- each step's propagation of all
 * trajectories is one complex GEMM on the FP64 matrix cores.  Matrices are dim x dim complex128,
 * row-major, interleaved (re, im). */

/* exp(A) by scaling and squaring with an 18-term Taylor series, ||A / 2^s||_inf <= 1/2
 * (traj_detail::expm, src/TrajectorySolver.h:97-111), the products on the GPU.  Host in/out. */
int llsv_expm(int dim, const double* a, double* out);

/* C = A B for complex128 row-major matrices (host in/out), M and N multiples of 64, K of 16: the
 * GEMM the trajectory steps run on the FP64 matrix cores.  llsv_zgemm_time: average milliseconds of
 * `reps` back-to-back launches on device-resident operands (HIP events).  No counterpart in the
 * reference. */
int llsv_zgemm(int M, int N, int K, const double* a, const double* b, double* c);
int llsv_zgemm_time(int M, int N, int K, int reps, float* ms);

/* A plan (TrajectorySolver::Plan, src/TrajectorySolver.h:291-301) on the device.
 * kind 0: jumps (STANDARD / PROJECTOR; src/TrajectorySolver.h:379-428): nodes = 1 propagator
 *         exp(-i H_eff h), or 2 steps + 1 matrices -i (H(t_j) + H_nh) at t_j = j h / 2 (RK4);
 *         ops = the n_ops jump operators.
 * kind 1: diffusive (ANALOG; src/TrajectorySolver.h:432-460): nodes = 1 propagator exp(-i H h),
 *         or `steps` matrices -i H(t_j) at t_j = j h (Euler drift); ops = the C_k.
 * n_ops <= 16. */
typedef struct llsv_lplan llsv_lplan;
int llsv_lplan_create(int dim, int kind, int64_t steps, double h, int n_nodes, const double* nodes, int n_ops,
                      const double* ops, llsv_lplan** out);

/* The same plan assembled on the device from the model itself, as TrajectorySolver::make_plan
 * does on the host (src/TrajectorySolver.h:268-290): H = 1 Hamiltonian (constant: the propagator
 * exp(-i (H + H_nh) h), scaling and squaring on the device) or 2 steps + 1 (jumps) / `steps`
 * (diffusive) samples H(t_j); ops = the n_ops channel operators as the unraveling applies them
 * (sqrt(gamma) L, or sqrt(c gamma) (1 - P) for PROJECTOR); for jumps H_nh = -(i/2) sum C_k^+ C_k
 * is one GEMM on the device.  Only H and the operators cross PCIe. */
int llsv_lplan_create_h(int dim, int kind, int64_t steps, double h, int n_h, const double* H, int n_ops,
                        const double* ops, llsv_lplan** out);
void llsv_lplan_destroy(llsv_lplan* p);

/* Runs trajectories first_traj .. first_traj + n_traj - 1 of an ensemble from psi0 (dim complex,
 * normalised): states (host, n_traj x dim complex128, or NULL) gets the final normalised states,
 * rho_sum (host, dim x dim complex128, or NULL) the sum of their |psi><psi| (a GEMM on the device).
 * rng 0: trajectory i draws from std::mt19937_64(splitmix64(seed ^ splitmix64(i + 1))) with
 * std::uniform_real_distribution<double>(0, 1) (jumps) or std::normal_distribution<double>(0,
 * sqrt(h)) (diffusive), like TrajectorySolver::make_rng (src/TrajectorySolver.h:311-313), so an
 * ensemble reproduces the reference's draws; rng 1: Philox4x32-10 on the device.  A vanishing or
 * diverging norm is LLSV_ERR_NUMERIC with the reference's message.  A plan keeps its working
 * buffers between runs, so one plan must not run on two threads at once (two plans may); it runs
 * on the device that was current when it was made, whatever device is current at the call. */
int llsv_lplan_run(llsv_lplan* p, const double* psi0, int64_t n_traj, int64_t first_traj, uint64_t seed, int rng,
                   double* states, double* rho_sum);

/* ---- interop and measurement ------------------------------------------------ */

/* Device address of the amplitudes and the state's stream (hipStream_t as void*). */
void* llsv_device_ptr(llsv_state* s);
void* llsv_stream(llsv_state* s);

/* HIP-event timer on the state's stream: start records an event, stop records a
 * second one, waits for it and returns the elapsed milliseconds. */
int llsv_timer_start(llsv_state* s);
int llsv_timer_stop(llsv_state* s, float* ms);

/* Per-kernel timing: llsv_mark records event `slot` (0 <= slot < LLSV_MARK_SLOTS) on the
 * state's stream; llsv_mark_elapsed waits for event b and returns the milliseconds
 * between events a and b. */
#define LLSV_MARK_SLOTS 256
int llsv_mark(llsv_state* s, int slot);
int llsv_mark_elapsed(llsv_state* s, int a, int b, float* ms);

/* Wait for all work queued on `stream` (a hipStream_t; NULL = the default stream). */
int llsv_stream_sync(void* stream);

/* ---- testing ------------------------------------------------------------------ */

/* Makes one device or pinned-host allocation of the library fail as if memory had run out: the
 * one `after` allocations from now (0 = the next one); the hook then disarms.  -1 disarms it.
 * Returns the previous setting: the allocations that were still to pass before the armed
 * failure, or -1 if none was armed (so a caller can tell whether the failure was used).
 * Process-wide.  For tests of the LLSV_ERR_NOMEM paths (state creation, program buffers,
 * staging, readout, trajectory, ZNE and Lindblad workspaces); no counterpart in the reference. */
long llsv_test_fail_allocation(long after);

/* Message of the last failure on this thread ("" if none). */
const char* llsv_last_error(void);

/* Library version string. */
const char* llsv_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LLSV_H */

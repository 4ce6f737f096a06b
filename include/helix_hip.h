/*
 * helix_hip.h -- C-ABI between the C++17 host of helix_amd and its gfx950 HIP kernels.
 *
 * The ABI replaces the reference's internal hot-path boundary
 *   initialize()        src/initialize.h:8, src/initialize.cu:342-352
 *   develop()           src/liouville.h:27, src/liouville.cu:502-591
 *   getdRhoSparse()     src/liouville.h:33, src/liouville.cu:749-869
 *   clearLiouvilleStorage() / clearMatrixStorage()   src/liouville.cu:871, src/matrix_storage.cu:32
 *   ResultExtractor::final_reduced_density()         src/library/result_extractor.cu:35-89
 * which src/library/legacy_runtime_session.cu:53-99 drives.  Plain pointers and sizes only.
 *
 * Conventions
 *   - every function returns int: 0 ok, < 0 contract error (HELIX_HIP_E*), > 0 a hipError_t;
 *     helix_hip_last_error() gives the message of the calling thread's last failure.
 *   - no exception crosses the ABI; host buffers are caller-owned; device memory and the
 *     HIP stream are owned by the plan.  One plan per host thread at a time; not re-entrant.
 *   - complex numbers are interleaved (re, im) doubles; matrices are row-major N x N.
 *
 * The HEOM right-hand side evaluated for ADO n (multi-bath form of src/liouville.cu:783-858):
 *   L(rho)_n = -i[H, rho_n]
 *            + sum_k (-i) [V_{b(k)}, rho_{up(n,k)}]
 *            + sum_k ( down_comm[n][k] [V_{b(k)}, rho_{down(n,k)}] + down_anti[n][k] {V_{b(k)}, rho_{down(n,k)}} )
 *            + self_coef[n] rho_n
 *            - sum_b xi[b] [V_b, [V_b, rho_n]] - sum_b xi_counter[b] {V_b, [V_b, rho_n]}
 * and one step is the Taylor-m recurrence rho <- sum_{j=0..m} (dt L)^j / j! rho (src/liouville.cu:537-590).
 */
#ifndef HELIX_HIP_H
#define HELIX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HELIX_HIP_ABI_VERSION 2

/* contract error codes (< 0), cf. src/backend/transpose_backend.h:14-16 */
#define HELIX_HIP_EINVAL (-1)
#define HELIX_HIP_EUNSUPPORTED (-2)
#define HELIX_HIP_ENOMEM (-3)
#define HELIX_HIP_ESTATE (-4)

enum helix_hip_precision {
	HELIX_HIP_FP64 = 0,       /* complex128 everywhere (the plain form) */
	HELIX_HIP_FP32 = 1,       /* complex64 ADOs; tier-0 accumulation in FP64 */
	HELIX_HIP_SPLIT_FP32 = 2  /* as FP32, with operator values split into two FP32 terms */
};

/* Storage of each ADO block in device memory.  When every exponent is real (Drude-Lorentz with Pade
 * or Matsubara poles), down_comm is imaginary, down_anti real, the self and Xi coefficients real and
 * KXiCounter imaginary, the Hamiltonian and couplings Hermitian, then L maps Hermitian ADOs to
 * Hermitian ADOs and a Hermitian initial state keeps every ADO Hermitian.  The Hermitian layout
 * stores N(N+1)/2 elements per ADO instead of N^2 (about half the HBM traffic).  N <= 64:
 *   main [N][N/2]: element (x, k) = rho[k][k] for x = 0, else rho[i][i ^ x] with i the k-th index whose
 *                  bit h is clear, h the highest set bit of x (so i < i ^ x)
 *   tail [N/2]:    element t = rho[N/2 + t][N/2 + t]
 * N > 64:
 *   main [N][N/2]: element (i, d) = rho[i][(i + d) mod N] for 0 <= d < N/2
 *   tail [N/2]:    element t = rho[t][t + N/2]
 * every other element is the conjugate of a stored one.  States passed in must then be Hermitian
 * (to 1e-12 relative), else the call returns HELIX_HIP_EINVAL; states passed out are full blocks. */
enum helix_hip_layout {
	HELIX_HIP_LAYOUT_AUTO = 0,      /* Hermitian where the problem and the stage kernel allow it */
	HELIX_HIP_LAYOUT_FULL = 1,      /* N^2 elements per ADO (any state; needed by the learned closure) */
	HELIX_HIP_LAYOUT_HERMITIAN = 2  /* Hermitian layout or HELIX_HIP_EUNSUPPORTED */
};

/* A "spin-structured" operator: O = diag(d) + sum_f c_f X_f, (X_f)_{i, i^mask_f} = 1.
 * The legacy V (src/initialize.cu:95-123), the transverse-field H and sigma^z couplings all fit. */
typedef struct helix_hip_spin_operator {
	const double* diag;          /* [n] real diagonal */
	int32_t nflips;              /* number of bit-flip terms (<= 16) */
	const int32_t* flip_masks;   /* [nflips], 0 < mask < n */
	const double* flip_coefs;    /* [nflips] real coefficient of each flip */
} helix_hip_spin_operator;

/* A general sparse Hermitian operator in CSR form (the reference's SparseOperator,
 * include/helix/types.h:100-106, with int32 indices); values are complex [nnz][2]. */
typedef struct helix_hip_csr_operator {
	int32_t nnz;
	const int32_t* row_offsets;  /* [n+1] */
	const int32_t* columns;      /* [nnz], 0 <= column < n */
	const double* values;        /* [nnz][2] */
} helix_hip_csr_operator;

typedef struct helix_hip_problem {
	int32_t abi_version;         /* HELIX_HIP_ABI_VERSION */
	int32_t n;                   /* Hilbert dimension N, a power of two, 8 <= N <= 4096 */
	int32_t n_ado;               /* hierarchy size H */
	int32_t n_modes;             /* M exponential modes over all baths */
	int32_t n_baths;             /* B coupling operators */
	const int32_t* up;           /* [H][M] index of n+e_k, H when absent  (src/initialize_detail.h:234-269) */
	const int32_t* down;         /* [H][M] index of n-e_k, H when absent */
	const int32_t* mode_bath;    /* [M] bath of each mode */
	const int32_t* tier;         /* [H] sum_k n_k of each ADO (ADO 0 is tier 0) */
	const double* down_comm;     /* [H][M][2] */
	const double* down_anti;     /* [H][M][2] */
	const double* self_coef;     /* [H][2] */
	const double* xi;            /* [B][2] KXi_b */
	const double* xi_counter;    /* [B][2] KXiCounter_b */
	helix_hip_spin_operator hamiltonian;
	const double* hamiltonian_dense; /* optional [N][N][2]; non-null selects the dense MFMA commutator */
	const helix_hip_spin_operator* couplings; /* [B] */
	int32_t precision;           /* enum helix_hip_precision */
	int32_t order;               /* Taylor order m, 1 <= m <= 8 (ContextOptions::integrationOrder) */
	double dt;                   /* time step (ContextOptions::timeStep) */
	int32_t layout;              /* enum helix_hip_layout */
	/* optional general sparse operators (each must be Hermitian): a non-null hamiltonian_csr replaces
	 * `hamiltonian`, a non-null couplings_csr [B] replaces `couplings` (whose diagonals must still be given,
	 * e.g. the CSR diagonal).  Such problems run on the generic two-pass stage kernels. */
	const helix_hip_csr_operator* hamiltonian_csr;
	const helix_hip_csr_operator* couplings_csr;
} helix_hip_problem;

typedef struct helix_hip_plan helix_hip_plan;

/* Uploads tables and operators, allocates the state (rho plus m-1 stage buffers) on `device`.
 * Replaces initialize() + the first-call initSparse() (src/liouville.cu:685-747).
 * Size limit: n_ado * N^2 < 2^31 elements (n_ado * N^2 * n_baths with bit-flip or sparse couplings, whose
 * commutator pass keeps one buffer per bath); HELIX_HIP_EINVAL beyond. */
int helix_hip_plan_create(const helix_hip_problem* problem, int device, helix_hip_plan** out);

/* rho_0 <- rho0 (tier 0, [N][N][2]); every other ADO <- 0 (src/initialize.cu:287-307). */
int helix_hip_set_initial(helix_hip_plan* plan, const double* rho0);

/* Full state in/out, [H][N][N][2] complex128 regardless of the plan precision. */
int helix_hip_set_state(helix_hip_plan* plan, const double* state);
int helix_hip_get_state(helix_hip_plan* plan, double* state);

/* out <- L(state), [H][N][N][2] (getdRhoSparse, src/liouville.cu:749-869). */
int helix_hip_rhs(helix_hip_plan* plan, double* out);

/* Advances nsteps Taylor-m steps (develop() x nsteps, src/library/legacy_runtime_session.cu:79-99).
 * When record_every > 0, after every record_every-th step the tier-0 block is recorded:
 *   rho_trace  [nsteps/record_every][N][N][2] (may be NULL)
 *   diag_trace [nsteps/record_every][N] real diagonal (may be NULL).
 * Returns after the device finished. */
int helix_hip_step(helix_hip_plan* plan, uint64_t nsteps, uint64_t record_every, double* rho_trace,
	double* diag_trace);

/* helix_hip_step without any device allocation.  No counterpart in the reference: it serves callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, steps and records as helix_hip_step, bit for bit; every stage is launched eagerly (no hipGraph
 * is captured), and the trace records are staged in a 32 MiB device array of the library (one copy per GPU,
 * held until the process ends) and copied out whenever it fills, so nsteps is unbounded but one record
 * (2 N^2 doubles with rho_trace, plus N with diag_trace) must fit in 32 MiB: N <= 1024 with rho_trace, else
 * HELIX_HIP_EINVAL.  Every call on one GPU uses that GPU's array: calls that overlap on one GPU (several host
 * threads) race. */
int helix_hip_step_noalloc(helix_hip_plan* plan, uint64_t nsteps, uint64_t record_every, double* rho_trace,
	double* diag_trace);

/* Tier-0 block (the reduced density rho_S) as [N][N][2] row-major complex128
 * (ResultExtractor::final_reduced_density, src/library/result_extractor.cu:35-89). */
int helix_hip_get_tier0(helix_hip_plan* plan, double* out);

/* Selected ADO blocks out, out[q] = ADO ados[q] (hierarchy order, 0 <= ados[q] < H) as [N][N][2] complex128, unpacked
 * from the stored layout (FP32 forms: tier 0 from its FP64 master copy).  Generalises the reference's extraction of
 * block 0 (ResultExtractor::final_reduced_density, src/library/result_extractor.cu:35-89) to any ADO; gathered on the
 * device, so only the listed blocks cross PCIe. */
int helix_hip_get_ados(helix_hip_plan* plan, const int32_t* ados, int count, double* out);

/* Device-timed steps: enqueues nsteps steps with HIP events around every stage launch on the plan's
 * stream; stage_ms[j] (j < order, may be NULL) gets the mean duration of stage j+1 launches and
 * total_ms the event time of the whole sequence. */
int helix_hip_time_steps(helix_hip_plan* plan, uint64_t nsteps, double* stage_ms, double* total_ms);

/* Plan facts: bytes of device memory held, launches per step, kernel names. */
int helix_hip_plan_info(const helix_hip_plan* plan, size_t* device_bytes, int* launches_per_step,
	int* uses_flip_coupling, int* uses_dense_hamiltonian);

/* The layout the plan stores its ADOs in (HELIX_HIP_LAYOUT_FULL or HELIX_HIP_LAYOUT_HERMITIAN). */
int helix_hip_plan_layout(const helix_hip_plan* plan, int32_t* layout);

void helix_hip_plan_destroy(helix_hip_plan* plan);

/* Learned closure of the deepest kept tier (the network of include/helix_mlp.h).  No counterpart in the
 * reference, whose deepest tier simply loses its up-neighbour terms (src/liouville.cu:818-823).  The
 * missing term of each listed ADO n = ados[a] is modelled as an elementwise complex rate,
 *   L(rho)_n[ij] += kappa_a[ij] rho_n[ij],   (Re, Im) kappa_a[ij] = f(x),
 *   x = [2 n_b(i) - 1 for b < log2 N, 2 n_b(j) - 1 for b < log2 N, extra[a][0 .. extra_dim)]
 * (the form of a Markovian terminator for sigma^z couplings).  The network runs once here, on the GPU,
 * and kappa stays resident; every stage kernel then applies it (linear, so exact in every Taylor stage).
 * mlp == NULL switches the closure off (plain truncation).  The MLP may be destroyed afterwards.
 * Needs a plan in HELIX_HIP_LAYOUT_FULL (the rates need not be conjugate-symmetric). */
struct helix_hip_mlp;
/* Same closure with the rates given directly: kappa [count][N][N][2] (e.g. an analytic terminator);
 * kappa == NULL switches it off. */
int helix_hip_plan_set_closure_rates(helix_hip_plan* plan, const int32_t* ados, int count, const double* kappa);
int helix_hip_plan_set_closure(helix_hip_plan* plan, struct helix_hip_mlp* mlp, const int32_t* ados, int count,
	const float* extra, int extra_dim);

/* Physics terminator of the deepest kept tier L.  No counterpart in the reference, whose deepest tier simply loses
 * its up-neighbour terms (src/liouville.cu:818-823).  In every Taylor stage each of the `count` virtual tier-(L+1)
 * ADOs u is solved from the stationary condition of its own equation, its own up-neighbours dropped:
 *   0 = -i[H, X_u] + (rate[u] + S) X_u + sum_s ( comm[s] [V_b(s), rho_src[s]] + anti[s] {V_b(s), rho_src[s]} )
 * (s over the sources of u: its tier-L down-neighbours, b(s) the bath of the edge's mode; S the plan's elementwise
 * Xi self term of sigma^z couplings).  The solve works in the eigenbasis of H = Q diag(E) Q^dagger,
 *   X~_ab = -(Q^dagger (D + S X_prev) Q)_ab / (rate[u] - i (E_a - E_b)),
 * `iterations` fixed-point rounds for S (X_prev = 0 first; S does not commute with H).  Every listed kept tier-L
 * ADO up_ado[e] then gains (-i)[V_b(up_mode[e]), X_up_virtual[e]], the term a real up-neighbour would give.  Linear
 * in the state, so exact inside every Taylor stage; combines with the closure rates above.  Needs a plan in
 * HELIX_HIP_LAYOUT_FULL with sigma^z couplings, N <= 32, on the tiled stage kernel.  NULL switches it off. */
typedef struct helix_hip_terminator {
	int32_t count;               /* virtual ADOs U */
	const double* rate;          /* [U][2] self coefficient of each virtual ADO (-sum_k nu_k n_k) */
	const int32_t* src_offsets;  /* [U+1] sources of u: src_offsets[u] .. src_offsets[u+1]-1 */
	const int32_t* src_ado;      /* [S] kept tier-L ADO (hierarchy order) */
	const int32_t* src_mode;     /* [S] mode k of the edge (u = src + e_k) */
	const double* src_comm;      /* [S][2] down_comm of u for mode k */
	const double* src_anti;      /* [S][2] down_anti of u for mode k */
	int32_t n_up;
	const int32_t* up_ado;       /* [n_up] kept tier-L ADO n (hierarchy order) */
	const int32_t* up_mode;      /* [n_up] mode k with n + e_k = virtual ADO up_virtual */
	const int32_t* up_virtual;   /* [n_up] u */
	const double* h_vectors;     /* [N][N][2] unitary Q whose columns are eigenvectors of H */
	const double* h_values;      /* [N] eigenvalues E */
	int32_t iterations;          /* fixed-point rounds for S, 1..16 */
} helix_hip_terminator;
int helix_hip_plan_set_terminator(helix_hip_plan* plan, const helix_hip_terminator* terminator);
/* Learned correction of the installed terminator (no counterpart in the reference): corr [U][N][N][2] multiplies
 * the response of every eigen-element, X~_ab = -corr[u][a][b] (Q^dagger (D + S X_prev) Q)_ab / (rate[u] - i (E_a - E_b)),
 * so corr = 1 is the physics terminator above.  NULL restores it.  EINVAL without a terminator; installing another
 * terminator drops the correction. */
int helix_hip_plan_set_terminator_gain(helix_hip_plan* plan, const double* corr);

/* ---- batched form: many independent problems advanced by one launch per Taylor stage ----
 * No counterpart in the reference (one hard-coded problem per process, src/helix.cpp:73-104).
 * Every problem must have N <= 64, sigma^z couplings, a spin-structured H, and the same precision
 * and Taylor order (>= 2); dt and everything else may differ.  The batch stores every problem in the
 * Hermitian layout when all of them allow it and none asks for HELIX_HIP_LAYOUT_FULL.  Flags: */
#define HELIX_HIP_BATCH_GRAPH 1u     /* capture the m stage launches of a step in a hipGraph */
#define HELIX_HIP_BATCH_AUTOTUNE 2u  /* time the workgroup shapes at creation and keep the fastest */
#define HELIX_HIP_BATCH_PER_SIZE 4u  /* one launch per block size and stage only (the autotuner does not try the persistent launch) */
#define HELIX_HIP_BATCH_PERSISTENT 8u  /* one persistent launch per stage from per-XCD task queues */

typedef struct helix_hip_batch helix_hip_batch;

int helix_hip_batch_create(const helix_hip_problem* const* problems, int count, int device, unsigned flags,
	helix_hip_batch** out);
/* index < 0: every problem gets rho_0 = |N-1><N-1| (rho0 ignored); else problem `index` gets rho0. */
int helix_hip_batch_set_initial(helix_hip_batch* batch, int index, const double* rho0);
int helix_hip_batch_step(helix_hip_batch* batch, uint64_t nsteps);
int helix_hip_batch_get_tier0(helix_hip_batch* batch, int index, double* out);
int helix_hip_batch_get_state(helix_hip_batch* batch, int index, double* state);
/* helix_hip_get_ados for problem `index` of the batch */
int helix_hip_batch_get_ados(helix_hip_batch* batch, int index, const int32_t* ados, int count, double* out);
int helix_hip_batch_time_steps(helix_hip_batch* batch, uint64_t nsteps, double* total_ms);
/* total ADOs, device bytes, workgroup threads chosen, graph in use */
int helix_hip_batch_info(const helix_hip_batch* batch, int64_t* total_ado, size_t* device_bytes, int* threads,
	int* graph);
int helix_hip_batch_layout(const helix_hip_batch* batch, int32_t* layout);
/* Kernel launches per step: the order m when every stage is one persistent launch (HELIX_HIP_BATCH_PERSISTENT, or
 * the autotuner's pick: workgroups dealt to the block sizes in proportion to their work take tiles from per-XCD task
 * queues), else m times the number of block sizes present (the default). */
int helix_hip_batch_launches(const helix_hip_batch* batch, int32_t* launches_per_step);
void helix_hip_batch_destroy(helix_hip_batch* batch);

/* Attainable peaks of one device, measured by the library's own kernels (no counterpart in the reference: the
 * roofline fractions are quoted against datasheet peaks, and these put what the GPU sustains beside them).
 * hbm_read_gbs: a streaming read of `bytes`; hbm_copy_gbs: a copy of `bytes`, counted as bytes read plus bytes
 * written (both in one pass of short-lived workgroups, 8 KiB read or 4 KiB copied each); mfma_f64_tflops / mfma_f32_tflops: dense v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 issue (2048
 * flops each) from 8 waves per CU, 8 independent accumulators per wave.  Each figure is the best of `reps` timings of 8
 * back-to-back launches, after one untimed launch.  1 MiB <= bytes <= 1 TiB (rounded down to 16 B); allocates 2 x bytes of device memory
 * for the call and synchronises the device. */
typedef struct helix_hip_peaks {
	double hbm_read_gbs;
	double hbm_copy_gbs;
	double mfma_f64_tflops;
	double mfma_f32_tflops;
	size_t bytes;        /* bytes each bandwidth kernel read */
	int compute_units;   /* CUs the launches were sized for */
} helix_hip_peaks;
int helix_hip_measure_peaks(int device, size_t bytes, int reps, helix_hip_peaks* out);

/* Message of the calling thread's last error ("" when none). */
const char* helix_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

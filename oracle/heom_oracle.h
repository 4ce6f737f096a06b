/*
 * heom_oracle.h -- CPU FP64 restatement of the HELIX HEOM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * helix_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Complex numbers are stored as interleaved (re, im) doubles; matrices are
 * row-major; an ADO state vector is [H][N][N] complex.
 */
#ifndef HEOM_ORACLE_H
#define HEOM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Pade spectrum decomposition (src/psd/psd.cu:12-113, eigenvalues.cu:92-128) ---- */
int hx_psd(int K, double* poles, double* residues);

/* ---- hierarchy (src/initialize_detail.h:159-269) ---- */
long hx_hier_size(int modes, int tiers);
long hx_hier_rows(int modes, int tiers, int* rows);
int hx_hier_edges(const int* rows, int H, int M, int* edges);

/* ---- bath coefficients (src/initialize.cu:160-218) ----
 * out[0..K]        nu_k
 * out[K+1]         KXi (real)
 * out[K+2..K+3]    KXiCounter (complex)
 * out[K+4..K+5]    KThetaCommutate (complex)
 * out[K+6..K+7]    KThetaAntiCommutate (complex)
 * out[K+8 ..]      KPhi_k, k=1..K (complex, 2K doubles)
 */
int hx_bath_coefficients(double beta, double gamma, double zeta, double omega0, int K, double* out);

/* ---- operators ---- */
typedef struct {
	int n;
	int nnz;
	const int* offsets; /* [n+1] */
	const int* cols;    /* [nnz] */
	const double* vals; /* [nnz][2] */
} hx_csr;

/* legacy spin glass (src/initialize.cu:95-157, src/initialize_detail.h:12-157).
 * h_diag: [N] real diagonal of H; v_*: CSR of V with capacity N*(spins+1). */
int hx_legacy_operators(int N, double omega0, double dlong, double dtrans, double jn_aver, int is_square,
	double* h_diag, int* v_offsets, int* v_cols, double* v_vals);
int hx_legacy_spin_count(int N);
double hx_legacy_hamiltonian_diagonal(int state, int spins, double jn_aver, int is_square, double omega0);
int hx_legacy_coupling_matrix(int spins, double value, int is_square, double* out);

/* seeded transverse-field spin glass: J_ij ~ N(0,1) (i<j) from SplitMix64 + Box-Muller */
unsigned long long hx_splitmix64(unsigned long long* state);
double hx_uniform(unsigned long long* state);
double hx_normal(unsigned long long* state);
int hx_tfim_couplings(int spins, unsigned long long seed, double* J /* [spins][spins], upper triangle */);
/* H as CSR with (spins+1) entries per row; column order ascending. */
int hx_tfim_hamiltonian(int spins, const double* J, double h, int* offsets, int* cols, double* vals);

/* ---- the HEOM problem (multi-bath generalisation of src/liouville.cu:685-869) ---- */
typedef struct {
	int N, H, M, B;
	const int* rows;      /* [H][M] ADO multi-indices */
	const int* edges;     /* [H][2M] up (n+e_k) then down (n-e_k); sentinel H */
	const int* mode_bath; /* [M] bath of mode k */
	const int* mode_kind; /* [M] 0 = Drude mode (theta terms), >0 = Pade mode index */
	const double* nu;     /* [M] */
	const double* kphi;   /* [M][2] KPhi of Pade modes */
	const double* kxi;    /* [B][2] */
	const double* kxic;   /* [B][2] */
	const double* kthc;   /* [B][2] */
	const double* ktha;   /* [B][2] */
	int h_diagonal;       /* 1: H given by h_diag (reference H_DIAGONAL kernel) */
	const double* h_diag; /* [N][2] */
	hx_csr h;             /* general H when h_diagonal == 0 */
	const hx_csr* v;      /* [B] coupling operators (real symmetric in the reference) */
	int threads;          /* OpenMP threads, <=0: default */
} hx_problem;

int hx_rhs(const hx_problem* p, const double* rho, double* drho);
int hx_taylor_step(const hx_problem* p, int order, double dt, long nsteps, double* rho);
int hx_rk4_step(const hx_problem* p, double dt, long nsteps, double* rho);

/* observables */
double hx_energy_diagonal(const double* h_dense_or_null, const double* h_diag, int N, const double* rho0);
int hx_trace_h_rho(const hx_csr* h, const double* rho0, double* out /* [2] */);

#ifdef __cplusplus
}
#endif
#endif

/*
 * heom_oracle.c -- CPU FP64 restatement of the HELIX HEOM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of the
 * product.  It follows the reference algorithm term by term:
 *   - RHS L(rho):           src/liouville.cu:749-869 (term order :783-858),
 *                           coefficient slots :685-747, commutator helpers
 *                           :275-366 (+ transpose trick, src/matrix_util.cu:27-81),
 *                           diagonal-H kernel :368-386
 *   - Taylor-m integrator:  src/liouville.cu:502-591
 *   - operators / bath:     src/initialize.cu:95-218, src/initialize_detail.h:12-157
 *   - hierarchy:            src/initialize_detail.h:159-269
 *   - PSD:                  src/psd/psd.cu:12-113, src/psd/eigenvalues.cu:92-128
 *   - energy:               src/library/energy_observable.cpp:8-39
 * Multi-bath generalisation (one bath per coupling operator V_b, modes
 * concatenated bath by bath) keeps the single-bath term order per bath.
 */
#include "heom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ PSD */

/* Jacobi eigenvalues of a symmetric matrix, largest-pivot rotations until the
 * strict-lower off-diagonal L1 norm is <= 1e-14 (src/psd/eigenvalues.cu:92-128). */
static int jacobi_eigenvalues(int n, double* a /* n*n, symmetric, overwritten */, double* eig)
{
	const double tol = 1.0e-14;
	long max_sweeps = 100L * n * n;
	if(max_sweeps < 16) max_sweeps = 16;
	for(long sweep = 0; sweep < max_sweeps; sweep++)
	{
		double off = 0.0;
		int p = 1, q = 0;
		double best = -1.0;
		for(int c = 0; c < n; c++)
			for(int r = c + 1; r < n; r++)
			{
				const double v = fabs(a[r * n + c]);
				off += v;
				if(v > best) { best = v; p = r; q = c; }
			}
		if(off <= tol || n < 2) break;
		const double app = a[q * n + q], aqq = a[p * n + p], apq = a[p * n + q];
		if(apq == 0.0) break;
		/* rotation that annihilates a[p][q] (index q plays "p" of the textbook form) */
		const double theta = 0.5 * atan2(2.0 * apq, aqq - app);
		const double c = cos(theta), s = sin(theta);
		for(int k = 0; k < n; k++)
		{
			if(k == p || k == q) continue;
			const double akq = a[k * n + q], akp = a[k * n + p];
			const double nq = c * akq - s * akp;
			const double np = s * akq + c * akp;
			a[k * n + q] = a[q * n + k] = nq;
			a[k * n + p] = a[p * n + k] = np;
		}
		a[q * n + q] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
		a[p * n + p] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
		a[p * n + q] = a[q * n + p] = 0.0;
	}
	double off = 0.0;
	for(int c = 0; c < n; c++)
		for(int r = c + 1; r < n; r++) off += fabs(a[r * n + c]);
	if(off > tol) return -1;
	for(int i = 0; i < n; i++) eig[i] = a[i * n + i];
	return 0;
}

static int cmp_double(const void* x, const void* y)
{
	const double a = *(const double*)x, b = *(const double*)y;
	return (a > b) - (a < b);
}

/* bosonic [K-1/K] Pade poles xi_k and residues eta_k (src/psd/psd.cu:12-113) */
int hx_psd(int K, double* poles, double* residues)
{
	if(K <= 0) return -1;
	const int n1 = 2 * K;
	double* m = (double*)calloc((size_t)n1 * n1, sizeof(double));
	double* eig = (double*)calloc((size_t)n1, sizeof(double));
	double* xi = (double*)calloc((size_t)K, sizeof(double));
	double* zeta = (double*)calloc((size_t)K + 1, sizeof(double));
	int rc = 0;
	/* b(m) = 2m+1 for bosons; tridiagonal 1/sqrt(b(x+1) b(y+1)) */
	for(int x = 0; x < n1; x++)
		for(int y = 0; y < n1; y++)
			m[x * n1 + y] = (abs(x - y) == 1) ? 1.0 / sqrt((2.0 * (x + 1) + 1.0) * (2.0 * (y + 1) + 1.0)) : 0.0;
	if(jacobi_eigenvalues(n1, m, eig) != 0) { rc = -2; goto done; }
	qsort(eig, (size_t)n1, sizeof(double), cmp_double);
	for(int i = 0; i < K; i++) poles[i] = 2.0 / eig[n1 - 1 - i];

	{
		const int n2 = n1 - 1;
		memset(m, 0, sizeof(double) * (size_t)n1 * n1);
		for(int x = 0; x < n2; x++)
			for(int y = 0; y < n2; y++)
				m[x * n2 + y] = (abs(x - y) == 1) ? 1.0 / sqrt((2.0 * (x + 2) + 1.0) * (2.0 * (y + 2) + 1.0)) : 0.0;
		if(jacobi_eigenvalues(n2, m, eig) != 0) { rc = -3; goto done; }
		qsort(eig, (size_t)n2, sizeof(double), cmp_double);
		const int mm = n2 / 2;
		for(int i = 0; i < K; i++) xi[i] = poles[i] * poles[i];
		for(int i = 0; i < mm; i++) zeta[i] = 4.0 / (eig[n2 - 1 - i] * eig[n2 - 1 - i]);
		const double scaling = (double)K * (2.0 * (K + 1) + 1.0);
		for(int i = 0; i < K; i++)
		{
			double t = (i == K - 1) ? 0.5 * scaling
			                        : 0.5 * scaling * (zeta[i] - xi[i]) / (xi[K - 1] - xi[i]);
			for(int j = 0; j < mm; j++)
			{
				if(i == j) continue;
				t *= (zeta[j] - xi[i]) / (xi[j] - xi[i]);
			}
			residues[i] = t;
		}
	}
done:
	free(m); free(eig); free(xi); free(zeta);
	return rc;
}

/* ------------------------------------------------------------ hierarchy */

/* number of multi-indices over `modes` modes with sum < tiers
 * (src/initialize_detail.h:159-172: modes==1 -> tiers, else a sum over the first index) */
long hx_hier_size(int modes, int tiers)
{
	if(modes <= 0 || tiers <= 0) return 0;
	if(modes == 1) return tiers;
	long total = 0;
	for(int i = 0; i < tiers; i++) total += hx_hier_size(modes - 1, tiers - i);
	return total;
}

static void hier_rec(int depth, int modes, int remaining, int* prefix, int* rows, long* count)
{
	if(depth == modes)
	{
		memcpy(rows + (*count) * modes, prefix, sizeof(int) * (size_t)modes);
		(*count)++;
		return;
	}
	for(int i = 0; i < remaining; i++)
	{
		prefix[depth] = i;
		hier_rec(depth + 1, modes, remaining - i, prefix, rows, count);
	}
}

/* rows in the reference's lexicographic order, mode 0 most significant
 * (src/initialize_detail.h:174-214; the depth-1 pruning slip for two modes is not reproduced) */
long hx_hier_rows(int modes, int tiers, int* rows)
{
	int prefix[256];
	if(modes <= 0 || modes > 256) return -1;
	long count = 0;
	hier_rec(0, modes, tiers, prefix, rows, &count);
	return count;
}

/* up/down edge table by exhaustive search, sentinel H (src/initialize_detail.h:234-269) */
int hx_hier_edges(const int* rows, int H, int M, int* edges)
{
	for(long i = 0; i < (long)H * 2 * M; i++) edges[i] = H;
	int* target = (int*)malloc(sizeof(int) * (size_t)M);
	for(int i = 0; i < H; i++)
	{
		for(int j = 0; j < M; j++)
		{
			for(int k = 0; k < M; k++) target[k] = rows[(long)i * M + k] + (j == k ? 1 : 0);
			for(int k = 0; k < H; k++)
			{
				const int* r = rows + (long)k * M;
				int eq = 1;
				for(int q = 0; q < M; q++)
					if(r[q] != target[q]) { eq = 0; break; }
				if(eq)
				{
					edges[(long)i * 2 * M + j] = k;
					edges[(long)k * 2 * M + j + M] = i;
					break;
				}
			}
		}
	}
	free(target);
	return 0;
}

/* ------------------------------------------------------ bath coefficients */

/* src/initialize.cu:160-218 (nu, KXi, KXiCounter, KPhi, KTheta*) */
int hx_bath_coefficients(double beta, double gamma, double zeta, double omega0, int K, double* out)
{
	double poles[64], residues[64];
	if(K < 0 || K > 60) return -1;
	if(K > 0 && hx_psd(K, poles, residues) != 0) return -2;
	double* nu = out;
	nu[0] = gamma;
	for(int i = 1; i <= K; i++) nu[i] = poles[i - 1] / beta;
	const double bgh = 0.5 * beta * gamma;
	const double bo = beta * omega0;
	double sigma = 1.0 - bgh / tan(bgh);
	for(int i = 1; i <= K; i++)
		sigma = sigma - 2.0 * residues[i - 1] * gamma * gamma / (nu[i] * nu[i] - gamma * gamma);
	out[K + 1] = sigma * (zeta / bo);
	out[K + 2] = 0.0;
	out[K + 3] = zeta * bgh / bo;
	out[K + 4] = 0.0;
	out[K + 5] = beta * gamma * 0.5 / tan(beta * gamma * 0.5) * zeta / bo;
	out[K + 6] = beta * gamma * 0.5 * zeta / bo;
	out[K + 7] = 0.0;
	for(int i = 1; i <= K; i++)
	{
		out[K + 8 + 2 * (i - 1)] = 0.0;
		out[K + 9 + 2 * (i - 1)] = (zeta / bo) * 2.0 * residues[i - 1] * gamma * gamma / (nu[i] * nu[i] - gamma * gamma);
	}
	return 0;
}

/* --------------------------------------------------------- legacy model */

int hx_legacy_spin_count(int N)
{
	return (int)((log((double)N) / log(2.0)) + 0.5);
}

static int is_sparse_flip_mask(int mask)
{
	/* src/initialize_detail.h:38-45: 1..256 (the 512 mask is omitted) */
	return mask == 1 || mask == 2 || mask == 4 || mask == 8 || mask == 16 || mask == 32 || mask == 64 ||
	       mask == 128 || mask == 256;
}

int hx_legacy_coupling_matrix(int spins, double value, int is_square, double* c)
{
	/* src/initialize_detail.h:52-114 */
	static const int tri[32][2] = {{0, 1}, {0, 7}, {1, 0}, {1, 7}, {1, 2}, {1, 8}, {2, 1}, {2, 3}, {2, 8}, {3, 2}, {3, 4},
		{3, 5}, {3, 8}, {4, 3}, {4, 5}, {5, 4}, {5, 6}, {5, 3}, {5, 8}, {6, 5}, {6, 7}, {6, 8}, {7, 6}, {7, 0}, {7, 1},
		{7, 8}, {8, 1}, {8, 2}, {8, 3}, {8, 5}, {8, 6}, {8, 7}};
	static const int sq[24][2] = {{0, 1}, {0, 7}, {1, 0}, {1, 2}, {1, 8}, {2, 1}, {2, 3}, {3, 2}, {3, 4}, {3, 8}, {4, 3},
		{4, 5}, {5, 4}, {5, 6}, {5, 8}, {6, 5}, {6, 7}, {7, 6}, {7, 0}, {7, 8}, {8, 1}, {8, 3}, {8, 5}, {8, 7}};
	for(int i = 0; i < spins * spins; i++) c[i] = 0.0;
	if(spins != 9)
	{
		for(int i = 0; i < spins; i++)
			for(int j = 0; j < spins; j++) c[i * spins + j] = (i - j == -1 || i - j == 1) ? value : 0.0;
	}
	else if(is_square)
	{
		for(int e = 0; e < 24; e++) c[sq[e][0] * 9 + sq[e][1]] = value;
	}
	else
	{
		for(int e = 0; e < 32; e++) c[tri[e][0] * 9 + tri[e][1]] = value;
	}
	return 0;
}

static double projection(int state, int spin, double omega0)
{
	return (((state >> spin) & 1) == 1) ? 0.5 * omega0 : -0.5 * omega0;
}

static double hamiltonian_diag_from(const double* c, int state, int spins, double omega0)
{
	/* src/initialize_detail.h:143-157: sum of projections, then sum_{j} sum_{k<j} c_jk p_j p_k */
	double total = 0.0;
	for(int s = 0; s < spins; s++) total += projection(state, s, omega0);
	for(int j = 0; j < spins; j++)
		for(int k = 0; k < j; k++)
			total += c[j * spins + k] * projection(state, j, omega0) * projection(state, k, omega0);
	return total;
}

double hx_legacy_hamiltonian_diagonal(int state, int spins, double jn_aver, int is_square, double omega0)
{
	double* c = (double*)malloc(sizeof(double) * (size_t)spins * spins);
	hx_legacy_coupling_matrix(spins, jn_aver, is_square, c);
	const double v = hamiltonian_diag_from(c, state, spins, omega0);
	free(c);
	return v;
}

/* src/initialize.cu:95-157 (sparse H diagonal and CSR V) */
int hx_legacy_operators(int N, double omega0, double dlong, double dtrans, double jn_aver, int is_square,
	double* h_diag, int* v_offsets, int* v_cols, double* v_vals)
{
	const int spins = hx_legacy_spin_count(N);
	double* c = (double*)malloc(sizeof(double) * (size_t)spins * spins);
	hx_legacy_coupling_matrix(spins, jn_aver, is_square, c);
	int nnz = 0;
	v_offsets[0] = 0;
	for(int row = 0; row < N; row++)
	{
		double psum = 0.0;
		for(int s = 0; s < spins; s++) psum += projection(row, s, omega0);
		h_diag[row] = hamiltonian_diag_from(c, row, spins, omega0);
		for(int col = 0; col < N; col++)
		{
			if(row == col)
			{
				v_cols[nnz] = row;
				v_vals[2 * nnz] = dlong * psum;
				v_vals[2 * nnz + 1] = 0.0;
				nnz++;
			}
			if(is_sparse_flip_mask(row ^ col))
			{
				v_cols[nnz] = col;
				v_vals[2 * nnz] = dtrans;
				v_vals[2 * nnz + 1] = 0.0;
				nnz++;
			}
		}
		v_offsets[row + 1] = nnz;
	}
	free(c);
	return nnz;
}

/* ----------------------------------------- seeded spin glass (TFIM) model */

unsigned long long hx_splitmix64(unsigned long long* state)
{
	/* SplitMix64 (Steele, Lea, Flood 2014; constants as published by S. Vigna) */
	unsigned long long z = (*state += 0x9E3779B97F4A7C15ULL);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

double hx_uniform(unsigned long long* state)
{
	return (double)(hx_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

double hx_normal(unsigned long long* state)
{
	/* Box-Muller, cosine branch only */
	const double u1 = 1.0 - hx_uniform(state);
	const double u2 = hx_uniform(state);
	return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

int hx_tfim_couplings(int spins, unsigned long long seed, double* J)
{
	unsigned long long st = seed;
	for(int i = 0; i < spins * spins; i++) J[i] = 0.0;
	for(int i = 0; i < spins; i++)
		for(int j = i + 1; j < spins; j++) J[i * spins + j] = hx_normal(&st);
	return 0;
}

int hx_tfim_hamiltonian(int spins, const double* J, double h, int* offsets, int* cols, double* vals)
{
	const int N = 1 << spins;
	int nnz = 0;
	offsets[0] = 0;
	for(int row = 0; row < N; row++)
	{
		double e = 0.0;
		for(int i = 0; i < spins; i++)
			for(int j = i + 1; j < spins; j++)
			{
				const double si = ((row >> i) & 1) ? 1.0 : -1.0;
				const double sj = ((row >> j) & 1) ? 1.0 : -1.0;
				e += J[i * spins + j] * si * sj;
			}
		for(int col = 0; col < N; col++)
		{
			const int x = row ^ col;
			if(col == row)
			{
				cols[nnz] = col; vals[2 * nnz] = e; vals[2 * nnz + 1] = 0.0; nnz++;
			}
			else if((x & (x - 1)) == 0)
			{
				cols[nnz] = col; vals[2 * nnz] = h; vals[2 * nnz + 1] = 0.0; nnz++;
			}
		}
		offsets[row + 1] = nnz;
	}
	return nnz;
}

/* ---------------------------------------------------------------- RHS */

static inline void cmul(double ar, double ai, double br, double bi, double* cr, double* ci)
{
	*cr = ar * br - ai * bi;
	*ci = ar * bi + ai * br;
}

/* out += alpha * (A X), A in CSR, X dense row-major (SpMM, src/cuda_sparse_backend_plan.cu:262-272) */
static void csr_left_acc(const hx_csr* A, const double* X, double ar, double ai, double* out)
{
	const int n = A->n;
	for(int i = 0; i < n; i++)
	{
		double* o = out + 2L * i * n;
		for(int z = A->offsets[i]; z < A->offsets[i + 1]; z++)
		{
			double wr, wi;
			cmul(ar, ai, A->vals[2 * z], A->vals[2 * z + 1], &wr, &wi);
			const double* x = X + 2L * A->cols[z] * n;
			for(int j = 0; j < n; j++)
			{
				o[2 * j] += wr * x[2 * j] - wi * x[2 * j + 1];
				o[2 * j + 1] += wr * x[2 * j + 1] + wi * x[2 * j];
			}
		}
	}
}

/* out += alpha * (X A) */
static void csr_right_acc(const hx_csr* A, const double* X, double ar, double ai, double* out)
{
	const int n = A->n;
	for(int i = 0; i < n; i++)
	{
		const double* x = X + 2L * i * n;
		double* o = out + 2L * i * n;
		for(int k = 0; k < n; k++)
		{
			double xr, xi;
			cmul(ar, ai, x[2 * k], x[2 * k + 1], &xr, &xi);
			if(xr == 0.0 && xi == 0.0) continue;
			for(int z = A->offsets[k]; z < A->offsets[k + 1]; z++)
			{
				const double vr = A->vals[2 * z], vi = A->vals[2 * z + 1];
				const int j = A->cols[z];
				o[2 * j] += xr * vr - xi * vi;
				o[2 * j + 1] += xr * vi + xi * vr;
			}
		}
	}
}

/* out += alpha * X */
static void axpy(long n2, double ar, double ai, const double* x, double* out)
{
	for(long e = 0; e < n2; e++)
	{
		const double xr = x[2 * e], xi = x[2 * e + 1];
		out[2 * e] += ar * xr - ai * xi;
		out[2 * e + 1] += ar * xi + ai * xr;
	}
}

static int drude_mode_of(const hx_problem* p, int b)
{
	for(int k = 0; k < p->M; k++)
		if(p->mode_bath[k] == b && p->mode_kind[k] == 0) return k;
	return -1;
}

int hx_rhs(const hx_problem* p, const double* rho, double* drho)
{
	const int N = p->N, H = p->H, M = p->M, B = p->B;
	const long n2 = (long)N * N;
	double* buf = (double*)calloc((size_t)B * H * n2 * 2, sizeof(double));
	if(!buf) return -1;
	int* drude = (int*)malloc(sizeof(int) * (size_t)(B > 0 ? B : 1));
	for(int b = 0; b < B; b++) drude[b] = drude_mode_of(p, b);
#ifdef _OPENMP
	if(p->threads > 0) omp_set_num_threads(p->threads);
#endif

	/* stage 1: buf_{b,n} = [V_b, rho_n]  (CommutateSparse with k = 1, src/liouville.cu:792-801) */
#pragma omp parallel for schedule(dynamic, 1)
	for(long t = 0; t < (long)B * H; t++)
	{
		const int b = (int)(t / H), n = (int)(t % H);
		double* o = buf + (t * n2) * 2;
		const double* x = rho + (long)n * n2 * 2;
		csr_left_acc(&p->v[b], x, 1.0, 0.0, o);
		csr_right_acc(&p->v[b], x, -1.0, 0.0, o);
	}

	/* stage 2 per ADO in the reference's accumulation order (src/liouville.cu:783-858) */
#pragma omp parallel for schedule(dynamic, 1)
	for(int n = 0; n < H; n++)
	{
		double* d = drho + (long)n * n2 * 2;
		const double* x = rho + (long)n * n2 * 2;
		const int* row = p->rows + (long)n * M;
		const int* edge = p->edges + (long)n * 2 * M;
		memset(d, 0, sizeof(double) * (size_t)n2 * 2);

		/* -i[H, rho_n] */
		if(p->h_diagonal)
		{
			for(int r = 0; r < N; r++)
				for(int c = 0; c < N; c++)
				{
					const double dr = p->h_diag[2 * r] - p->h_diag[2 * c];
					const double di = p->h_diag[2 * r + 1] - p->h_diag[2 * c + 1];
					double kr, ki;
					cmul(0.0, -1.0, dr, di, &kr, &ki);
					const long e = (long)r * N + c;
					cmul(kr, ki, x[2 * e], x[2 * e + 1], &d[2 * e], &d[2 * e + 1]);
				}
		}
		else
		{
			csr_left_acc(&p->h, x, 0.0, -1.0, d);
			csr_right_acc(&p->h, x, 0.0, 1.0, d);
		}

		/* phi: -i buf_{n+e_k} */
		for(int k = 0; k < M; k++)
		{
			const int up = edge[k];
			if(up >= H) continue;
			axpy(n2, 0.0, -1.0, buf + ((long)p->mode_bath[k] * H + up) * n2 * 2, d);
		}
		/* psi: c_psi_k(n) buf_{n-e_k}, c = -nu_k n_k KPhi_k */
		for(int k = 0; k < M; k++)
		{
			if(p->mode_kind[k] == 0) continue;
			const int dn = edge[M + k];
			if(dn >= H) continue;
			const double j = -p->nu[k] * row[k];
			axpy(n2, p->kphi[2 * k] * j, p->kphi[2 * k + 1] * j, buf + ((long)p->mode_bath[k] * H + dn) * n2 * 2, d);
		}
		/* theta and theta2 per bath */
		for(int b = 0; b < B; b++)
		{
			const int k0 = drude[b];
			if(k0 < 0) continue;
			const int dn = edge[M + k0];
			if(dn >= H) continue;
			const double g = -p->nu[k0] * row[k0];
			axpy(n2, g * p->kthc[2 * b], g * p->kthc[2 * b + 1], buf + ((long)b * H + dn) * n2 * 2, d);
			const double* xd = rho + (long)dn * n2 * 2;
			const double ar = g * p->ktha[2 * b], ai = g * p->ktha[2 * b + 1];
			csr_right_acc(&p->v[b], xd, ar, ai, d);
			csr_left_acc(&p->v[b], xd, ar, ai, d);
		}
		/* Sigma: -(sum_k nu_k n_k) rho_n */
		{
			double s = 0.0;
			for(int k = 0; k < M; k++) s += p->nu[k] * row[k];
			axpy(n2, -s, 0.0, x, d);
		}
		/* Xi: R += (-KXi) V buf + KXi buf V  (addCommutateSparse, src/liouville.cu:847-851) */
		for(int b = 0; b < B; b++)
		{
			const double* bn = buf + ((long)b * H + n) * n2 * 2;
			csr_left_acc(&p->v[b], bn, -p->kxi[2 * b], -p->kxi[2 * b + 1], d);
			csr_right_acc(&p->v[b], bn, p->kxi[2 * b], p->kxi[2 * b + 1], d);
		}
		/* counter: R += (-KXiCounter) {V, buf}  (src/liouville.cu:853-858) */
		for(int b = 0; b < B; b++)
		{
			const double* bn = buf + ((long)b * H + n) * n2 * 2;
			csr_right_acc(&p->v[b], bn, -p->kxic[2 * b], -p->kxic[2 * b + 1], d);
			csr_left_acc(&p->v[b], bn, -p->kxic[2 * b], -p->kxic[2 * b + 1], d);
		}
	}
	free(drude);
	free(buf);
	return 0;
}

/* Taylor-m recurrence: F = rho; cur = rho; next = (dt/j) L(cur); F += next (src/liouville.cu:502-591) */
int hx_taylor_step(const hx_problem* p, int order, double dt, long nsteps, double* rho)
{
	const long total = (long)p->H * p->N * p->N * 2;
	double* F = (double*)malloc(sizeof(double) * (size_t)total);
	double* cur = (double*)malloc(sizeof(double) * (size_t)total);
	double* next = (double*)malloc(sizeof(double) * (size_t)total);
	if(!F || !cur || !next) { free(F); free(cur); free(next); return -1; }
	for(long s = 0; s < nsteps; s++)
	{
		memcpy(F, rho, sizeof(double) * (size_t)total);
		memcpy(cur, rho, sizeof(double) * (size_t)total);
		for(int j = 1; j <= order; j++)
		{
			if(hx_rhs(p, cur, next) != 0) { free(F); free(cur); free(next); return -2; }
			const double tj = dt / j;
#pragma omp parallel for
			for(long e = 0; e < total; e++)
			{
				next[e] *= tj;
				F[e] += next[e];
			}
			double* t = cur; cur = next; next = t;
		}
		memcpy(rho, F, sizeof(double) * (size_t)total);
	}
	free(F); free(cur); free(next);
	return 0;
}

/* classical RK4 -- a second integrator used to check the Taylor recurrence (equal for linear L) */
int hx_rk4_step(const hx_problem* p, double dt, long nsteps, double* rho)
{
	const long total = (long)p->H * p->N * p->N * 2;
	double* k = (double*)malloc(sizeof(double) * (size_t)total * 5);
	if(!k) return -1;
	double *k1 = k, *k2 = k + total, *k3 = k + 2 * total, *k4 = k + 3 * total, *tmp = k + 4 * total;
	for(long s = 0; s < nsteps; s++)
	{
		hx_rhs(p, rho, k1);
		for(long e = 0; e < total; e++) tmp[e] = rho[e] + 0.5 * dt * k1[e];
		hx_rhs(p, tmp, k2);
		for(long e = 0; e < total; e++) tmp[e] = rho[e] + 0.5 * dt * k2[e];
		hx_rhs(p, tmp, k3);
		for(long e = 0; e < total; e++) tmp[e] = rho[e] + dt * k3[e];
		hx_rhs(p, tmp, k4);
		for(long e = 0; e < total; e++) rho[e] += dt / 6.0 * (k1[e] + 2.0 * k2[e] + 2.0 * k3[e] + k4[e]);
	}
	free(k);
	return 0;
}

/* ---------------------------------------------------------- observables */

/* sum_i Re H_ii Re rho_ii (src/library/energy_observable.cpp:8-39) */
double hx_energy_diagonal(const double* h_dense, const double* h_diag, int N, const double* rho0)
{
	double e = 0.0;
	for(int i = 0; i < N; i++)
	{
		const double hii = h_dense ? h_dense[2 * ((long)i * N + i)] : h_diag[i];
		e += hii * rho0[2 * ((long)i * N + i)];
	}
	return e;
}

/* Tr(H rho) = sum_ij H_ij rho_ji */
int hx_trace_h_rho(const hx_csr* h, const double* rho0, double* out)
{
	const int N = h->n;
	double sr = 0.0, si = 0.0;
	for(int i = 0; i < N; i++)
		for(int z = h->offsets[i]; z < h->offsets[i + 1]; z++)
		{
			const int j = h->cols[z];
			const double* r = rho0 + 2 * ((long)j * N + i);
			sr += h->vals[2 * z] * r[0] - h->vals[2 * z + 1] * r[1];
			si += h->vals[2 * z] * r[1] + h->vals[2 * z + 1] * r[0];
		}
	out[0] = sr;
	out[1] = si;
	return 0;
}

// heom_kernels.hpp -- gfx950 kernels of the HEOM Liouville action and its Taylor-m step.
//
// Layout in HBM: a state vector is H blocks of complex, one per ADO: N x N row-major (full layout) or the
// N(N+1)/2 stored elements of the Hermitian layout (include/helix_hip.h).  On the device the ADOs are stored
// in colexicographic order of their multi-indices; the C-ABI converts from and to the hierarchy's
// lexicographic order (src/initialize_detail.h:174-214) at the boundary.  No sentinel block is stored:
// absent neighbours are skipped by index (the reference reads a zero block, src/initialize.cu:289,306).
// A Taylor stage j of order m is one fused launch (diagonal couplings) or two launches (couplings
// with bit flips): it evaluates L(cur) and applies the recurrence in its epilogue,
//   j <  m : t_j = (dt/j) L(t_{j-1})                          (t_0 = rho)
//   j == m : rho = ((rho + t_1) + ... + t_{m-1}) + (dt/m) L(t_{m-1})
// which keeps the reference's summation order F = rho; F += t_j (src/liouville.cu:525-590) without
// storing F, and replaces cublasScal/cublasAxpy and both D2D copies.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace helix_amd {

constexpr int kMaxFlips = 16;
constexpr int kMaxOrder = 8;
constexpr int kBlock = 256;

// Everything a stage launch needs; passed by value as the kernel argument block.
struct StageArgs {
	const void* cur;          // t_{j-1} (rho for j = 1)
	void* out;                // t_j, or the new rho for the last stage
	const void* rho;          // last stage: the old rho (may alias out)
	const void* tprev[kMaxOrder]; // last stage: t_1 .. t_{m-1}
	int ntprev;
	double* rho0_f64;         // FP32 forms: tier-0 master copy in FP64 (last stage), else null
	const void* buf;          // flip couplings: [B][H][N][N] = [V_b, cur_n]
	const int32_t* up;        // [H][M]
	const int32_t* down;      // [H][M]
	const int32_t* mode_bath; // [M]
	const double2* dcomm;     // [H][M]
	const double2* danti;     // [H][M]
	const double2* selfc;     // [H]
	const double2* xi;        // [B]
	const double2* xic;       // [B]
	const double* vdiag;      // [B][N]
	const int32_t* vflip_n;   // [B] flips of V_b (flip couplings only)
	const int32_t* vflip_mask;// [B][kMaxFlips]
	const double* vflip_coef; // [B][kMaxFlips]
	const double* hdiag;      // [N]
	int hflips;
	int hmask[kMaxFlips];
	double hcoef[kMaxFlips];
	int n_ado, M, B, logN;
	double scale;             // dt / j
	// sigma^z-structured couplings (every V_b = alpha_b (2 n_{s_b} - 1), no flips): the bit mask of the
	// spin s_b and alpha_b per bath, and the per-(i XOR j) self-term table
	//   selfx[x] = -sum_b xi_b (v_b(i) - v_b(j))^2 = -4 sum_{b: x has bit s_b} alpha_b^2 xi_b
	// (the counter term vanishes because (v_i + v_j)(v_i - v_j) = alpha^2 (1 - 1) = 0).
	int sz;
	int sz_mask[16];
	double sz_alpha[16];
	int sz_shift[16];         // per mode k < 16: 2 s for the spin s of its bath
	double sz_ta[16];         // per mode k < 16: 2 alpha of its bath
	int sz_anti;              // bit k (k < 16): some ADO has a non-zero down_anti for mode k
	const double2* selfx;     // [N]
	// dense Hamiltonian (MFMA commutator path): [N][N] complex in the state precision; for the
	// split-FP32 form hdense holds the FP32 high part and hdense_lo the FP32 remainder
	const void* hdense;
	const void* hdense_lo;
	int hreal;                // every Im H_ij == 0: the imaginary-H products are skipped
	// learned closure of the deepest kept tier: L(rho)_n += kappa[kslot[n]][e] rho_n[e] where kslot[n] >= 0
	const double2* kappa;     // [count][N^2] or null
	const int32_t* kslot;     // [H]
	// general sparse Hermitian operators (CSR): H, and every V_b with its row offsets [B][N+1] absolute
	// into the concatenated column/value arrays
	const int32_t* hcsr_off;  // [N+1]
	const int32_t* hcsr_col;
	const double2* hcsr_val;
	const int32_t* vcsr_off;  // [B][N+1]
	const int32_t* vcsr_col;
	const double2* vcsr_val;
};

}  // namespace helix_amd

// heom_plan.hip -- gfx950 HEOM propagator behind the C-ABI of include/helix_hip.h.
//
// Replaces the reference's per-ADO stream/cuBLAS/cuSPARSE orchestration (src/liouville.cu:685-869),
// its transpose kernel (src/matrix_util.cu:27-81) and the develop() scal/axpy/D2D sequence
// (src/liouville.cu:502-591) by one launch per Taylor stage (two with bit-flip couplings).
#include "heom_kernels.hpp"

#include "helix_hip.h"
#include "helix_mlp.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

namespace helix_amd {

namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
	g_last_error = msg;
	return code;
}

int hip_fail(hipError_t e, const char* what)
{
	g_last_error = std::string(what) + ": " + hipGetErrorString(e);
	return static_cast<int>(e);
}
}  // namespace

// for the library's other translation units (peaks.hip)
void set_last_error(const std::string& msg)
{
	g_last_error = msg;
}

#define HX_HIP(call)                                         \
	do {                                                     \
		hipError_t hx_e_ = (call);                           \
		if(hx_e_ != hipSuccess) return hip_fail(hx_e_, #call); \
	} while(0)

template <typename T>
struct C2t;
template <>
struct C2t<double> {
	using type = double2;
};
template <>
struct C2t<float> {
	using type = float2;
};

// Streaming accesses of the Taylor epilogue: a stage's output and the single-use Taylor terms are
// neither re-read within the stage nor by a neighbour, so they are written and read non-temporally and
// leave the L2 and the memory-side cache to the blocks other workgroups still gather as neighbours.
template <typename T>
__device__ __forceinline__ void st_stream(typename C2t<T>::type* p, T re, T im)
{
	typedef T ev2 __attribute__((ext_vector_type(2)));
	ev2 t;
	t.x = re;
	t.y = im;
	__builtin_nontemporal_store(t, reinterpret_cast<ev2*>(p));
}
template <typename T>
__device__ __forceinline__ typename C2t<T>::type ld_stream(const typename C2t<T>::type* p)
{
	typedef T ev2 __attribute__((ext_vector_type(2)));
	const ev2 t = __builtin_nontemporal_load(reinterpret_cast<const ev2*>(p));
	return typename C2t<T>::type{t.x, t.y};
}

// two consecutive complex64 elements (16-byte aligned), streamed as above
__device__ __forceinline__ float4 ld_stream4(const float2* p)
{
	typedef float ev4 __attribute__((ext_vector_type(4)));
	const ev4 t = __builtin_nontemporal_load(reinterpret_cast<const ev4*>(p));
	return float4{t.x, t.y, t.z, t.w};
}
__device__ __forceinline__ void st_stream4(float2* p, const float4& v)
{
	typedef float ev4 __attribute__((ext_vector_type(4)));
	ev4 t;
	t.x = v.x;
	t.y = v.y;
	t.z = v.z;
	t.w = v.w;
	__builtin_nontemporal_store(t, reinterpret_cast<ev4*>(p));
}

// Hermitian layout (include/helix_hip.h): stored index of full element (i, j) of an N x N block, and
// whether the stored value must be conjugated.
// N <= 64 (whole blocks in a workgroup): ordered by x = i XOR j.  Main part [N][N/2], element (x, k):
// x = 0 is the diagonal rho[k][k]; for x > 0 it is rho[i][i ^ x] with i the k-th index whose bit h(x)
// (the highest set bit of x) is clear, so i < i ^ x.  Tail [N/2]: rho[N/2 + t][N/2 + t].  A sigma^z
// coupling of spin s has a zero commutator weight on every element whose x has bit s clear, and each row of
// the main part has one x, so a commutator-only neighbour term needs only the rows with bit s set.
// N > 64 (row bands): stored elements rho[i][(i + d) mod N] for d < N/2 as [N][N/2] element (i, d) (a band
// of rows is one contiguous slice, as the row tiles read it) and the tail rho[t][t + N/2] for t < N/2.
__host__ __device__ inline int herm_insert0(int k, int h)  // k with a zero bit inserted at position h
{
	return ((k >> h) << (h + 1)) | (k & ((1 << h) - 1));
}

__host__ __device__ inline int herm_high_bit(int x)  // x > 0
{
	return 31 - __builtin_clz(static_cast<unsigned>(x));
}

__host__ __device__ inline int herm_index(int i, int j, int logN, bool& conj)
{
	const int N = 1 << logN;
	if(logN <= 6) {
		const int x = i ^ j;
		if(x == 0) {
			conj = false;
			return (i < N / 2) ? i : N * N / 2 + (i - N / 2);
		}
		const int lo = (i < j) ? i : j;
		conj = i > j;
		const int h = herm_high_bit(x);
		const int k = ((lo >> (h + 1)) << h) | (lo & ((1 << h) - 1));
		return x * (N / 2) + k;
	}
	const int d = (j - i) & (N - 1);
	if(d < N / 2) {
		conj = false;
		return (i << (logN - 1)) | d;
	}
	if(d > N / 2) {
		conj = true;
		return (j << (logN - 1)) | (N - d);
	}
	conj = i >= N / 2;
	return N * N / 2 + (conj ? j : i);
}

// acc += w * x for a complex weight w given in FP64.  SPLIT (FP32 only) applies w as the two-term
// sum w_hi + w_lo of FP32 values so operator values keep ~FP64 accuracy while products and
// accumulation stay in FP32.
template <typename T, bool SPLIT>
__device__ __forceinline__ void cmadd(T& ar, T& ai, double wr, double wi, T xr, T xi)
{
	if constexpr(SPLIT) {
		const float rh = static_cast<float>(wr), ih = static_cast<float>(wi);
		const float rl = static_cast<float>(wr - static_cast<double>(rh));
		const float il = static_cast<float>(wi - static_cast<double>(ih));
		ar += rl * xr - il * xi;
		ai += rl * xi + il * xr;
		ar += rh * xr - ih * xi;
		ai += rh * xi + ih * xr;
	} else {
		const T r = static_cast<T>(wr), m = static_cast<T>(wi);
		ar += r * xr - m * xi;
		ai += r * xi + m * xr;
	}
}

// [V_b, X]_e and {V_b, X}_e for a bit-flip coupling V_b = diag(v) + sum_f c_f X_f:
//   [V,X]_ij = (v_i - v_j) X_ij + sum_f c_f (X_{i^m,j} - X_{i,j^m})
//   {V,X}_ij = (v_i + v_j) X_ij + sum_f c_f (X_{i^m,j} + X_{i,j^m})
// (the row-major result of CommutateSparse / addAntiCommutateSparse, src/liouville.cu:275-366).
template <typename T, bool SPLIT, bool ANTI>
__device__ __forceinline__ void flip_comm_anti(const typename C2t<T>::type* __restrict__ X, const StageArgs& a,
	int b, uint32_t i, uint32_t j, uint32_t e, T& cr, T& ci, T& sr, T& si)
{
	const uint32_t N = 1u << a.logN;
	const double vi = a.vdiag[b * N + i], vj = a.vdiag[b * N + j];
	const auto x = X[e];
	cr = ci = sr = si = T(0);
	cmadd<T, SPLIT>(cr, ci, vi - vj, 0.0, x.x, x.y);
	if constexpr(ANTI) cmadd<T, SPLIT>(sr, si, vi + vj, 0.0, x.x, x.y);
	const int nf = a.vflip_n[b];
	for(int f = 0; f < nf; ++f) {
		const uint32_t m = static_cast<uint32_t>(a.vflip_mask[b * kMaxFlips + f]);
		const double cf = a.vflip_coef[b * kMaxFlips + f];
		const auto x1 = X[e ^ (m << a.logN)];
		const auto x2 = X[e ^ m];
		cmadd<T, SPLIT>(cr, ci, cf, 0.0, x1.x - x2.x, x1.y - x2.y);
		if constexpr(ANTI) cmadd<T, SPLIT>(sr, si, cf, 0.0, x1.x + x2.x, x1.y + x2.y);
	}
}

// [O,X]_ij and {O,X}_ij for a general sparse Hermitian O in CSR form (row offsets `off`):
//   (O X)_ij = sum_{p in row i} o_p X[col_p][j],   (X O)_ij = sum_k X_ik O_kj = sum_{p in row j} X[i][col_p] conj(o_p)
// (Hermiticity turns the column walk of O into a row walk).  The reference computes the same with two
// column-major cuSPARSE SpMMs and transposes (src/liouville.cu:275-366).
template <typename T, bool SPLIT, bool ANTI>
__device__ __forceinline__ void csr_comm_anti(const typename C2t<T>::type* __restrict__ X, const int32_t* __restrict__ off,
	const int32_t* __restrict__ col, const double2* __restrict__ val, int logN, uint32_t i, uint32_t j, T& cr, T& ci,
	T& sr, T& si)
{
	cr = ci = sr = si = T(0);
	for(int p = off[i]; p < off[i + 1]; ++p) {
		const double2 o = val[p];
		const auto x = X[(static_cast<uint32_t>(col[p]) << logN) | j];
		cmadd<T, SPLIT>(cr, ci, o.x, o.y, x.x, x.y);
		if constexpr(ANTI) cmadd<T, SPLIT>(sr, si, o.x, o.y, x.x, x.y);
	}
	for(int p = off[j]; p < off[j + 1]; ++p) {
		const double2 o = val[p];  // X O picks conj(o): subtract it from the commutator, add it to the anticommutator
		const auto x = X[(i << logN) | static_cast<uint32_t>(col[p])];
		cmadd<T, SPLIT>(cr, ci, -o.x, o.y, x.x, x.y);
		if constexpr(ANTI) cmadd<T, SPLIT>(sr, si, o.x, -o.y, x.x, x.y);
	}
}

// coupling operator kinds of the generic stage kernels
constexpr int kVDiag = 0;  // every V_b diagonal: all V terms elementwise
constexpr int kVFlip = 1;  // V_b = diag + bit flips
constexpr int kVCsr = 2;   // V_b general sparse (CSR)

template <typename T, bool SPLIT, bool ANTI, int VOP>
__device__ __forceinline__ void coupling_comm_anti(const typename C2t<T>::type* __restrict__ X, const StageArgs& a, int b,
	uint32_t i, uint32_t j, uint32_t e, T& cr, T& ci, T& sr, T& si)
{
	if constexpr(VOP == kVCsr)
		csr_comm_anti<T, SPLIT, ANTI>(X, a.vcsr_off + b * ((1 << a.logN) + 1), a.vcsr_col, a.vcsr_val, a.logN, i, j, cr, ci,
			sr, si);
	else
		flip_comm_anti<T, SPLIT, ANTI>(X, a, b, i, j, e, cr, ci, sr, si);
}

// Pass A of a stage with non-diagonal couplings: buf[b][n] = [V_b, cur_n]  (src/liouville.cu:792-801).
template <typename T, bool SPLIT, int VOP>
__global__ __launch_bounds__(kBlock) void k_flip_commutator(const StageArgs a)
{
	using V2 = typename C2t<T>::type;
	const V2* __restrict__ cur = static_cast<const V2*>(a.cur);
	V2* __restrict__ buf = static_cast<V2*>(const_cast<void*>(a.buf));
	const uint32_t logE = 2u * a.logN;
	const uint32_t E = 1u << logE;
	const uint32_t N = 1u << a.logN;
	const uint32_t total = static_cast<uint32_t>(a.B * a.n_ado) << logE;
	for(uint32_t g = blockIdx.x * kBlock + threadIdx.x; g < total; g += gridDim.x * kBlock) {
		const uint32_t bn = g >> logE;
		const uint32_t e = g & (E - 1);
		const int b = static_cast<int>(bn / a.n_ado);
		const uint32_t n = bn - static_cast<uint32_t>(b) * a.n_ado;
		T cr, ci, sr, si;
		coupling_comm_anti<T, SPLIT, false, VOP>(cur + (static_cast<size_t>(n) << logE), a, b, e >> a.logN, e & (N - 1), e,
			cr, ci, sr, si);
		buf[g] = V2{cr, ci};
	}
}

// One Taylor stage: L(cur) for every ADO element with the recurrence epilogue.
// VOP = kVDiag: every V_b diagonal, so [V_b, X] and {V_b, X} are elementwise and the whole RHS is one pass.
// otherwise   : V_b with bit flips or general sparse; [V_b, cur] comes from pass A (buf), {V_b, .} terms
//               are gathered here.
// HCSR: H is a general sparse operator (CSR) instead of diag + bit flips.
template <typename T, bool SPLIT, bool LAST, int VOP, bool HCSR>
__global__ __launch_bounds__(kBlock) void k_stage(const StageArgs a)
{
	constexpr bool FLIP = VOP != kVDiag;
	using V2 = typename C2t<T>::type;
	const V2* __restrict__ cur = static_cast<const V2*>(a.cur);
	const V2* __restrict__ buf = static_cast<const V2*>(a.buf);
	const int logN = a.logN;
	const uint32_t N = 1u << logN;
	const uint32_t logE = 2u * logN;
	const uint32_t E = 1u << logE;
	const uint32_t total = static_cast<uint32_t>(a.n_ado) << logE;
	const int M = a.M;
	for(uint32_t g = blockIdx.x * kBlock + threadIdx.x; g < total; g += gridDim.x * kBlock) {
		const uint32_t n = g >> logE;
		const uint32_t e = g & (E - 1);
		const uint32_t i = e >> logN, j = e & (N - 1);
		const V2* __restrict__ cn = cur + (static_cast<size_t>(n) << logE);
		const V2 c = cn[e];
		T ar = T(0), ai = T(0);

		// -i[H, rho_n] (diagonal part as in src/liouville.cu:368-386, plus bit-flip terms, or a general CSR H)
		{
			T hr = T(0), hi = T(0);
			if constexpr(HCSR) {
				T sr, si;
				csr_comm_anti<T, SPLIT, false>(cn, a.hcsr_off, a.hcsr_col, a.hcsr_val, logN, i, j, hr, hi, sr, si);
			} else {
				cmadd<T, SPLIT>(hr, hi, a.hdiag[i] - a.hdiag[j], 0.0, c.x, c.y);
				for(int f = 0; f < a.hflips; ++f) {
					const uint32_t m = static_cast<uint32_t>(a.hmask[f]);
					const V2 x1 = cn[e ^ (m << logN)];
					const V2 x2 = cn[e ^ m];
					cmadd<T, SPLIT>(hr, hi, a.hcoef[f], 0.0, x1.x - x2.x, x1.y - x2.y);
				}
			}
			ar += hi;
			ai -= hr;
		}

		const int32_t* __restrict__ upn = a.up + static_cast<size_t>(n) * M;
		const int32_t* __restrict__ dnn = a.down + static_cast<size_t>(n) * M;

		// phi: (-i) [V_{b(k)}, rho_{n+e_k}]
		for(int k = 0; k < M; ++k) {
			const int u = upn[k];
			if(u >= a.n_ado) continue;
			const int b = a.mode_bath[k];
			if constexpr(!FLIP) {
				const double d = a.vdiag[b * N + i] - a.vdiag[b * N + j];
				const V2 x = cur[(static_cast<size_t>(u) << logE) + e];
				cmadd<T, SPLIT>(ar, ai, 0.0, -d, x.x, x.y);
			} else {
				const V2 x = buf[(static_cast<size_t>(b * a.n_ado + u) << logE) + e];
				ar += x.y;
				ai -= x.x;
			}
		}

		// psi / theta / theta2: down_comm [V_b, rho_{n-e_k}] + down_anti {V_b, rho_{n-e_k}}
		for(int k = 0; k < M; ++k) {
			const int dn = dnn[k];
			if(dn >= a.n_ado) continue;
			const int b = a.mode_bath[k];
			const double2 wc = a.dcomm[static_cast<size_t>(n) * M + k];
			const double2 wa = a.danti[static_cast<size_t>(n) * M + k];
			if constexpr(!FLIP) {
				const double vi = a.vdiag[b * N + i], vj = a.vdiag[b * N + j];
				const double d = vi - vj, s = vi + vj;
				const V2 x = cur[(static_cast<size_t>(dn) << logE) + e];
				cmadd<T, SPLIT>(ar, ai, wc.x * d + wa.x * s, wc.y * d + wa.y * s, x.x, x.y);
			} else {
				const V2 x = buf[(static_cast<size_t>(b * a.n_ado + dn) << logE) + e];
				cmadd<T, SPLIT>(ar, ai, wc.x, wc.y, x.x, x.y);
				if(wa.x != 0.0 || wa.y != 0.0) {
					T cr, ci, sr, si;
					coupling_comm_anti<T, SPLIT, true, VOP>(cur + (static_cast<size_t>(dn) << logE), a, b, i, j, e, cr, ci,
						sr, si);
					cmadd<T, SPLIT>(ar, ai, wa.x, wa.y, sr, si);
				}
			}
		}

		// Sigma, Xi and counter terms (+ the learned closure rate of the deepest kept tier)
		{
			double2 sc = a.selfc[n];
			if(a.kappa) {
				const int sl = a.kslot[n];
				if(sl >= 0) {
					const double2 kk = a.kappa[(static_cast<size_t>(sl) << logE) + e];
					sc.x += kk.x;
					sc.y += kk.y;
				}
			}
			if constexpr(!FLIP) {
				double wr = sc.x, wi = sc.y;
				for(int b = 0; b < a.B; ++b) {
					const double vi = a.vdiag[b * N + i], vj = a.vdiag[b * N + j];
					const double d = vi - vj, s = vi + vj;
					wr -= a.xi[b].x * d * d + a.xic[b].x * s * d;
					wi -= a.xi[b].y * d * d + a.xic[b].y * s * d;
				}
				cmadd<T, SPLIT>(ar, ai, wr, wi, c.x, c.y);
			} else {
				cmadd<T, SPLIT>(ar, ai, sc.x, sc.y, c.x, c.y);
				for(int b = 0; b < a.B; ++b) {
					T cr, ci, sr, si;
					coupling_comm_anti<T, SPLIT, true, VOP>(buf + (static_cast<size_t>(b * a.n_ado + n) << logE), a, b, i, j,
						e, cr, ci, sr, si);
					cmadd<T, SPLIT>(ar, ai, -a.xi[b].x, -a.xi[b].y, cr, ci);
					cmadd<T, SPLIT>(ar, ai, -a.xic[b].x, -a.xic[b].y, sr, si);
				}
			}
		}

		// Taylor epilogue
		const T s = static_cast<T>(a.scale);
		const T tr = ar * s, ti = ai * s;
		V2* __restrict__ out = static_cast<V2*>(a.out);
		if constexpr(!LAST) {
			st_stream<T>(out + g, tr, ti);
		} else {
			if(a.rho0_f64 != nullptr && n == 0) {
				double2* r0 = reinterpret_cast<double2*>(a.rho0_f64);
				double2 r = r0[e];
				for(int q = 0; q < a.ntprev; ++q) {
					const V2 t = ld_stream<T>(static_cast<const V2*>(a.tprev[q]) + g);
					r.x += static_cast<double>(t.x);
					r.y += static_cast<double>(t.y);
				}
				r.x += static_cast<double>(tr);
				r.y += static_cast<double>(ti);
				r0[e] = r;
				out[g] = V2{static_cast<T>(r.x), static_cast<T>(r.y)};
			} else {
				V2 r = ld_stream<T>(static_cast<const V2*>(a.rho) + g);
				for(int q = 0; q < a.ntprev; ++q) {
					const V2 t = ld_stream<T>(static_cast<const V2*>(a.tprev[q]) + g);
					r.x += t.x;
					r.y += t.y;
				}
				r.x += tr;
				r.y += ti;
				st_stream<T>(out + g, r.x, r.y);
			}
		}
	}
}

// LDS-tiled stage for diagonal couplings, N <= 1024 (LOGN 3..10).
// A workgroup owns a tile of 4096 elements = A = 4096/N^2 whole ADO blocks (N <= 64) or one row band of
// one block (N > 64, TPA bands per block); each ADO is handled by TA = THREADS/A threads, element
// e = lane_in_ado + TA r, so every global access is a contiguous row of lanes.  The tile's own blocks are
// staged once in LDS, so every Hamiltonian bit-flip gather (rows i^m, columns j^m) is served from LDS.
// Neighbours (n+e_k, n-e_k) are walked a batch at a time, every load of a batch issued before any
// arithmetic: tiles of several ADOs keep a compacted list per ADO in LDS; sigma^z tiles of one ADO (TABLE)
// build the list per wave from a ballot and table each neighbour's weights per bit pattern in LDS.
// Workgroups are renumbered so each XCD (round-robin dispatch over 8 XCDs) sweeps one contiguous range of
// the device's ADO order (colexicographic), so the near neighbours hit that XCD's L2.
// SZ: every coupling is alpha_b (2 n_{s_b} - 1): commutator / anticommutator weights come from bits
// (d_b = v_i - v_j and s_b = v_i + v_j, one of them zero) and the Sigma + Xi weight from selfx[i^j].
// HERM: the Hermitian layout of include/helix_hip.h (N(N+1)/2 stored elements per ADO).  The stored elements
// are loaded once and written to both (i, j) and (j, i) of the block in LDS (a row band completes its rows
// from the transposed elements of the other rows); the tile computes only the stored elements (every term
// of L is elementwise in the neighbours, and the Hamiltonian gathers come from the LDS block), so loads,
// stores and neighbour reads all halve.
constexpr int kTile = 4096;  // elements per workgroup
constexpr int kTileMaxModes = 16;
constexpr int kXcds = 8;

// acc += w * x for a real weight w (two FMAs; cmadd with a zero imaginary part would spend four)
template <typename T, bool SPLIT>
__device__ __forceinline__ void rmadd(T& ar, T& ai, double w, T xr, T xi)
{
	if constexpr(SPLIT) {
		const float h = static_cast<float>(w);
		const float l = static_cast<float>(w - static_cast<double>(h));
		ar += l * xr;
		ai += l * xi;
		ar += h * xr;
		ai += h * xi;
	} else {
		const T r = static_cast<T>(w);
		ar += r * xr;
		ai += r * xi;
	}
}

// A real operator weight in the form the FMAs consume: FP64 as is, FP32 rounded, split-FP32 as the two-term
// sum hi + lo.
template <typename T, bool SPLIT>
struct Weight {
	T hi;
	T lo;
	__device__ __forceinline__ static Weight make(double w)
	{
		Weight r;
		r.hi = static_cast<T>(w);
		r.lo = SPLIT ? static_cast<T>(w - static_cast<double>(r.hi)) : T(0);
		return r;
	}
	__device__ __forceinline__ Weight select(bool take_other, const Weight& other, bool negate) const
	{
		Weight r{take_other ? other.hi : hi, take_other ? other.lo : lo};
		if(negate) {
			r.hi = -r.hi;
			r.lo = -r.lo;
		}
		return r;
	}
};

// acc += w * x with a prepared weight (same operation order as rmadd)
template <typename T, bool SPLIT>
__device__ __forceinline__ void wmadd(T& ar, T& ai, const Weight<T, SPLIT>& w, T xr, T xi)
{
	if constexpr(SPLIT) {
		ar += w.lo * xr;
		ai += w.lo * xi;
	}
	ar += w.hi * xr;
	ai += w.hi * xi;
}

// LDS a tile body needs (sized for the largest tile that uses it)
template <typename T>
struct TileLds {
	using V2 = typename C2t<T>::type;
	V2* blk;        // [kTile]
	double* v;      // [16 N]   general diagonal couplings only
	double2* selfx; // [N]      sigma^z couplings only
	int* nb;        // [A][2 kTileMaxModes] valid neighbours of each ADO, packed (index << 5 | slot q)
	int* cnt;       // [A] number of valid neighbours
	int* mb;        // [kTileMaxModes] per mode: the bath (general couplings) or its sigma^z bit mask (SZ)
	double* ka;     // [kTileMaxModes] per mode, SZ: 2 alpha of its bath
	double* hd;     // [N] diagonal of H
	void* wt;       // [2 kTileMaxModes][4] PatternWeight, sigma^z tiles of one ADO: each neighbour's weights
	int* spread;    // [N] sigma^z tiles of one ADO: bit s of x moved to bit 2s
};

// bit s of x (x < 1024) moved to bit 2s
__device__ __forceinline__ int spread_bits(int x)
{
	x = (x | (x << 8)) & 0x00FF00FF;
	x = (x | (x << 4)) & 0x0F0F0F0F;
	x = (x | (x << 2)) & 0x33333333;
	x = (x | (x << 1)) & 0x55555555;
	return x;
}

// Complex neighbour weight of one sigma^z bit pattern in the form the FMAs consume: FP64 or FP32 as (re, im),
// split-FP32 as (hi re, hi im, lo re, lo im).
template <typename T, bool SPLIT>
struct PatternWeight {
	using type = typename C2t<T>::type;
	__device__ __forceinline__ static type make(double wr, double wi) { return type{static_cast<T>(wr), static_cast<T>(wi)}; }
};
template <>
struct PatternWeight<float, true> {
	using type = float4;
	__device__ __forceinline__ static float4 make(double wr, double wi)
	{
		const float rh = static_cast<float>(wr), ih = static_cast<float>(wi);
		return float4{rh, ih, static_cast<float>(wr - static_cast<double>(rh)), static_cast<float>(wi - static_cast<double>(ih))};
	}
};

__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }

// acc += w * x with a prepared complex weight: four FMAs per term (split: the FP32 remainder first)
template <typename T>
__device__ __forceinline__ void pmadd(T& ar, T& ai, const typename C2t<T>::type& w, T xr, T xi)
{
	ar = fmx(w.x, xr, ar);
	ai = fmx(w.x, xi, ai);
	ar = fmx(-w.y, xi, ar);
	ai = fmx(w.y, xr, ai);
}
__device__ __forceinline__ void pmadd(float& ar, float& ai, const float4& w, float xr, float xi)
{
	ar = __builtin_fmaf(w.z, xr, ar);
	ai = __builtin_fmaf(w.z, xi, ai);
	ar = __builtin_fmaf(-w.w, xi, ar);
	ai = __builtin_fmaf(w.w, xr, ai);
	ar = __builtin_fmaf(w.x, xr, ar);
	ai = __builtin_fmaf(w.x, xi, ai);
	ar = __builtin_fmaf(-w.y, xi, ar);
	ai = __builtin_fmaf(w.y, xr, ai);
}

// Register slot r of a tile thread (lt = lane within its ADO): the stored element q within the ADO and
// its full-block row i and column j.  Full layout: q = ebase + lt + TA r = i N + j.  Hermitian layout,
// whole blocks: slots r < EPT are main elements q = lt + TA r = x N/2 + k (row x = i XOR j, see herm_index);
// slot EPT is the tail element q = N^2/2 + lt (the diagonal element i = j = N/2 + lt), held by threads
// lt < N/2.  Row bands: the band's slice of the [N][N/2] main part and its rows' tail elements.
template <int LOGN, int TA, int EPT, bool HERM, bool PAIR = false>
__device__ __forceinline__ void tile_elem(int r, int lt, int ebase, int& q, int& i, int& j)
{
	constexpr int N = 1 << LOGN;
	if constexpr(HERM && LOGN <= 6) {
		if(r < EPT) {
			q = PAIR ? 2 * (lt + TA * (r >> 1)) + (r & 1) : lt + TA * r;
			const int x = q >> (LOGN - 1), k = q & (N / 2 - 1);
			i = (x == 0) ? k : herm_insert0(k, herm_high_bit(x));
			j = i ^ x;
		} else {
			q = N * N / 2 + lt;
			i = N / 2 + lt;
			j = i;
		}
	} else if constexpr(HERM) {
		// row band of R = 4096/N rows from r0 = ebase / N: its stored main part is one contiguous slice
		const int r0 = ebase >> LOGN;
		if(r < EPT) {
			const int qq = PAIR ? 2 * (lt + TA * (r >> 1)) + (r & 1) : lt + TA * r;  // 0 .. 2047
			q = (r0 << (LOGN - 1)) + qq;
			i = r0 + (qq >> (LOGN - 1));
			j = (i + (qq & (N / 2 - 1))) & (N - 1);
		} else {
			i = r0 + lt;
			q = N * N / 2 + i;
			j = i + N / 2;
		}
	} else {
		q = ebase + lt + TA * r;
		i = q >> LOGN;
		j = q & (N - 1);
	}
}

template <typename T, bool SPLIT, bool LAST, int LOGN, int THREADS, bool SZ, bool HERM, bool SINGLE = true>
__device__ __forceinline__ void tile_body(const StageArgs& a, const int group, const TileLds<T>& lds,
	const int tid = static_cast<int>(threadIdx.x))
{
	using V2 = typename C2t<T>::type;
	constexpr int N = 1 << LOGN;
	constexpr int LOGE = 2 * LOGN;
	constexpr int E = 1 << LOGE;
	constexpr int TPA = (E > kTile) ? E / kTile : 1;  // tiles per ADO (N > 64: row tiles of one block)
	constexpr int R = kTile >> LOGN;                  // rows of a row tile (N > 64)
	constexpr int A = (E >= kTile) ? 1 : kTile / E;   // ADOs per workgroup (N < 64)
	constexpr int TE = (E >= kTile) ? kTile : E;      // full-block elements of one ADO held in LDS
	constexpr int TA = THREADS / A;                   // threads per ADO
	constexpr int EPT = (HERM ? TE / 2 : TE) / TA;    // main elements per thread
	constexpr int RT = HERM ? EPT + 1 : EPT;          // register slots per thread (+ the tail element)
	constexpr size_t PE = HERM ? (static_cast<size_t>(E) + N) / 2 : static_cast<size_t>(E);  // stored per ADO
	// sigma^z tiles of one ADO: each neighbour's weights for the four bit patterns of an element, tabled in LDS
	constexpr bool TABLE = SZ && A == 1;
	constexpr bool SPREAD_LDS = LOGN <= 8;  // N >= 512: computed, the LDS of two complex128 workgroups is full
	using PW = PatternWeight<T, SPLIT>;
	using WT = typename PW::type;
	// complex64 Hermitian tiles (whole blocks of one or several ADOs, row bands of N >= 128): a thread's main slots
	// 2p, 2p + 1 are consecutive stored elements of one stored row (N/2 elements: one x = i XOR j, or one row i of a
	// band), so each global load and store of a pair moves 16 bytes per lane
	constexpr bool PAIR = sizeof(T) == 4 && HERM && EPT % 2 == 0 && (A > 1 || TABLE);
	static_assert(!HERM || TA >= (TPA == 1 ? N / 2 : R), "every tail element needs a thread");
	V2* __restrict__ s_blk = lds.blk;
	double* __restrict__ s_v = lds.v;
	double2* __restrict__ s_selfx = lds.selfx;
	int (*s_nb)[2 * kTileMaxModes] = reinterpret_cast<int (*)[2 * kTileMaxModes]>(lds.nb);
	int* __restrict__ s_cnt = lds.cnt;
	int* __restrict__ s_mb = lds.mb;
	double* __restrict__ s_ka = lds.ka;
	double* __restrict__ s_hd = lds.hd;
	WT* __restrict__ s_wt = static_cast<WT*>(lds.wt);
	int* __restrict__ s_spread = lds.spread;

	const V2* __restrict__ cur = static_cast<const V2*>(a.cur);
	const int H = a.n_ado, M = a.M, B = a.B;
	const int slot = tid / TA;
	const int ebase = (group % TPA) * kTile;  // first element of this row tile
	const int lt = tid - slot * TA;
	const int n = (group / TPA) * A + slot;  // this thread's ADO (uniform over TA threads)
	// its first stored element; with one ADO per workgroup it is wave-uniform and stays in scalar registers (as a
	// per-lane 64-bit value the complex64 N = 64 tile spilled it, 25 MB of scratch writes per launch)
	const size_t nbase = static_cast<size_t>(A == 1 ? __builtin_amdgcn_readfirstlane(n) : n) * PE;
	const bool live = n < H;
	// the tail slot holds an element for lt < N/2 (whole blocks), or for lt < R in row tiles of the upper half
	const bool tail_ok = HERM && (TPA == 1 ? lt < N / 2 : (lt < R && ebase < E / 2));
#define HX_SLOT_LIVE(r) (!HERM || (r) < EPT || tail_ok)
	// whole-block Hermitian tiles: the tile's own elements are loaded first, ahead of every table read
	V2 ownv[RT];
	if constexpr(HERM && TPA == 1) {
#pragma unroll
		for(int r = 0; r < RT; ++r) ownv[r] = V2{T(0), T(0)};
#pragma unroll
		for(int r = 0; r < RT; ++r) {
			if(!live || !HX_SLOT_LIVE(r)) continue;
			int q, i, j;
			tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, q, i, j);
			if constexpr(PAIR) {
				if(r < EPT) {
					if(r & 1) continue;  // loaded with slot r - 1
					const float4 v = *reinterpret_cast<const float4*>(cur + nbase + q);
					ownv[r] = V2{v.x, v.y};
					ownv[r + 1] = V2{v.z, v.w};
					continue;
				}
			}
			ownv[r] = cur[nbase + q];
		}
	}

	if constexpr(SZ) {
		for(int q = tid; q < N; q += THREADS) s_selfx[q] = a.selfx[q];
	} else {
		for(int q = tid; q < B * N; q += THREADS) s_v[q] = a.vdiag[q];
	}
	if constexpr(!TABLE) {
		for(int q = tid; q < M; q += THREADS) {
			const int b = a.mode_bath[q];
			if constexpr(SZ) {
				s_mb[q] = a.sz_mask[b];
				s_ka[q] = 2.0 * a.sz_alpha[b];
			} else {
				s_mb[q] = b;
			}
		}
	}
	for(int q = tid; q < N; q += THREADS) s_hd[q] = a.hdiag[q];
	if constexpr(TABLE && SPREAD_LDS)
		for(int q = tid; q < N; q += THREADS) s_spread[q] = spread_bits(q);
	// neighbour lists hold valid entries only (the deepest tier, most of the hierarchy, has no up-neighbours):
	// a ballot over the 2M <= 32 candidate slots of an ADO, held by 32 lanes of a wave, gives each valid
	// neighbour its position
	static_assert(2 * kTileMaxModes == 32 && THREADS % 64 == 0, "one half-wave per ADO slot list");
	// neighbours are walked NB entries at a time: every load of a batch is issued before any arithmetic, so a
	// thread waits for one memory round trip per NB neighbours (sized to what each variant's register budget
	// holds without spills: 128 VGPRs at 512 threads -- 64 for the complex64 N = 64 sigma^z tile --, 64 at 1024)
	constexpr int NB = TABLE ? (sizeof(T) == 8 ? 2 : (THREADS <= 512 ? 2 : 4))
	                         : HERM ? (sizeof(T) == 8 ? (THREADS <= 512 ? 2 : 1) : ((THREADS <= 512 || LOGN >= 6) ? 4 : 2))
	                                : (sizeof(T) == 8 ? 1 : (THREADS <= 512 ? 1 : 2));
	unsigned rem = 0;  // TABLE: the wave's valid candidate lanes not yet visited
	int cand = -1;     // TABLE: lane k < 2M holds candidate slot k as an entry (see below)
	int nvalid = 0;
	if constexpr(TABLE) {
		// every wave reads the ADO's 2M candidate slots (up n+e_k, then down n-e_k) and keeps the valid ones as a
		// lane mask, so it knows its neighbours without waiting for the workgroup; an entry is
		// (index << 6) | (anticommutator weight present << 5) | 2 s, s the spin of the mode's bath.  Wave 0 tables the weights of each bit pattern p = 2 bit_s(i) + bit_s(j) of
		// an element (i, j) by the entry's position: with v = alpha (2 bit - 1) the pair (v_i - v_j, v_i + v_j) is
		// (0, -2a), (-2a, 0), (2a, 0), (0, 2a) and the weight is wc (v_i - v_j) + wa (v_i + v_j), wc = -i (up) or
		// down_comm, wa = 0 (up) or down_anti.
		const int lane = tid & 63;
		const int km = (lane < M) ? lane : lane - M;
		double2 wc = make_double2(0.0, -1.0), wa = make_double2(0.0, 0.0);
		if(lane < 2 * M && live) {
			const int v = (lane < M) ? a.up[n * M + km] : a.down[n * M + km];
			if(tid >= M && tid < 2 * M) {  // wave 0's down slots: coefficients for the weight table
				wc = a.dcomm[n * M + km];
				wa = a.danti[n * M + km];
			}
			const int anti = (lane >= M) ? ((a.sz_anti >> km) & 1) : 0;
			if(v < H) cand = (v << 6) | (anti << 5) | a.sz_shift[km];
		}
		rem = static_cast<unsigned>(__ballot(cand >= 0));
		nvalid = __popc(rem);
		if(tid < 2 * M && cand >= 0) {
			const int pos = __popc(rem & ((1u << lane) - 1u));
			const double ta = a.sz_ta[km];
			const double dd[4] = {0.0, -ta, ta, 0.0}, ss[4] = {-ta, 0.0, 0.0, ta};
#pragma unroll
			for(int pt = 0; pt < 4; ++pt)
				s_wt[pos * 4 + pt] = PW::make(wc.x * dd[pt] + wa.x * ss[pt], wc.y * dd[pt] + wa.y * ss[pt]);
		}
	} else {
		for(int base = 0; base < A * 32; base += THREADS) {
			const int t = base + tid;
			const int s = t >> 5, k = t & 31;
			const int nn = (group / TPA) * A + s;
			int idx = -1;
			if(s < A && k < 2 * M && nn < H) {
				const int v = (k < M) ? a.up[nn * M + k] : a.down[nn * M + (k - M)];
				idx = (v < H) ? v : -1;
			}
			const unsigned long long bal = __ballot(idx >= 0);
			const int lane = tid & 63;
			const unsigned long long half = (lane < 32) ? 0xffffffffull : 0xffffffff00000000ull;
			if(idx >= 0) s_nb[s][__popcll(bal & half & ((1ull << lane) - 1ull))] = (idx << 5) | k;
			if(k == 0 && s < A) s_cnt[s] = __popcll(bal & half);
		}
	}
	// TABLE: the first batch of neighbour elements is loaded before the tile is staged, so its round trip
	// overlaps the staging, the barrier and -i[H, rho_n]; later batches follow in the neighbour loop
	// (neighbour loads: a wave-uniform block base plus each element's 32-bit byte offset)
	unsigned qb[RT];
	int shb[NB];
	V2 xb[NB][RT];
#pragma unroll
	for(int r = 0; r < RT; ++r) {
		qb[r] = 0;
		if(!HX_SLOT_LIVE(r)) continue;
		int qe, i, j;
		tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
		qb[r] = static_cast<unsigned>(qe) * sizeof(V2);
	}
	// Whole Hermitian blocks (rows of one x = i XOR j): a neighbour without an anticommutator weight has a
	// zero weight on every element whose x has bit s clear (bits s of i and j equal), so those elements -- half
	// of each such block, whole rows -- are not loaded (the tail is the diagonal, x = 0).
	constexpr bool SKIP = TABLE && HERM && TPA == 1;
	constexpr int QSH = (sizeof(V2) == 16 ? 4 : 3) + LOGN - 1;  // byte offset of a main element >> QSH = x
	auto load_batch = [&]() {
#pragma unroll
		for(int u = 0; u < NB; ++u) {
#pragma unroll
			for(int r = 0; r < RT; ++r) xb[u][r] = V2{T(0), T(0)};
			shb[u] = 0;
			if(rem == 0) continue;
			const int pos = __builtin_ctz(rem);
			rem &= rem - 1u;
			const int packed = __builtin_amdgcn_readlane(cand, pos);
			shb[u] = packed & 31;
			const bool anti = (packed >> 5) & 1;
			const char* __restrict__ src = reinterpret_cast<const char*>(cur + static_cast<size_t>(packed >> 6) * PE);
#pragma unroll
			for(int r = 0; r < RT; ++r) {
				if(!HX_SLOT_LIVE(r)) continue;
				if constexpr(SKIP)
					if(!anti && (r == EPT || !((qb[r] >> (QSH + (shb[u] >> 1))) & 1u))) continue;
				if constexpr(PAIR) {
					if(r < EPT) {
						if(r & 1) continue;  // loaded with slot r - 1 (same row x, same skip decision)
						const float4 v = *reinterpret_cast<const float4*>(src + qb[r]);
						xb[u][r] = V2{v.x, v.y};
						xb[u][r + 1] = V2{v.z, v.w};
						continue;
					}
				}
				xb[u][r] = *reinterpret_cast<const V2*>(src + qb[r]);
			}
		}
	};
	if constexpr(TABLE) load_batch();
	// -i[H, rho_n] accumulates in (hr, hi); row bands add their far row flips' partners first (below)
	T hr[RT], hi[RT];
#pragma unroll
	for(int r = 0; r < RT; ++r) {
		hr[r] = T(0);
		hi[r] = T(0);
	}
	if constexpr(HERM && TPA > 1) {
		// Row flips by a mask m >= R leave the band: the partner of stored element (i, j) is (i ^ m, j).  Where
		// (j - (i ^ m)) mod N < N/2 it is stored in row i ^ m, and the owners' own lanes (consecutive d) read it
		// contiguously.  Otherwise it is the conjugate of an element stored in row j; for those, lanes walk the band
		// rows at a fixed column j (their partners i ^ m are consecutive offsets of row j), so each load is a whole
		// run of row j, and park the values in the LDS tile (not staged yet) at their owner's slot: owner (i, j),
		// d = (j - i) mod N, sits at (i - r0) N/2 + d, a tail owner (d = N/2, i < N/2) at R N/2 + (i - r0).  Each
		// owner then adds h_f x (partner); the column-flip part of these flips follows from the staged tile.
		const int r0 = ebase >> LOGN;
		const V2* __restrict__ ownh = cur + nbase;
		for(int f = 0; f < a.hflips; ++f) {
			const int m = a.hmask[f];
			if((m << LOGN) < kTile) continue;  // flip within the band: from the staged tile
			if(live) {
				for(int t = tid; t < kTile; t += THREADS) {
					const int c = t / R, il = t - (t / R) * R;
					const int i = r0 + il;
					const int d = (c - i) & (N - 1);
					int ql;
					if(d < N / 2)
						ql = (il << (LOGN - 1)) | d;
					else if(d == N / 2 && i < N / 2)
						ql = R * (N / 2) + il;
					else
						continue;  // (i, c) is not stored in this band
					const int pa = i ^ m;
					if(((c - pa) & (N - 1)) < N / 2) continue;  // stored in row i ^ m: its owner reads it
					bool cj;
					V2 v = ownh[herm_index(pa, c, LOGN, cj)];
					if(cj) v.y = -v.y;
					s_blk[ql] = v;
				}
			}
			__syncthreads();
			if(live) {
				const Weight<T, SPLIT> w = Weight<T, SPLIT>::make(a.hcoef[f]);
#pragma unroll
				for(int r = 0; r < RT; ++r) {
					if(!HX_SLOT_LIVE(r)) continue;
					int q, i, j;
					tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, q, i, j);
					const int pa = i ^ m, dp = (j - pa) & (N - 1);
					const V2 x1 = (dp < N / 2) ? ownh[(pa << (LOGN - 1)) | dp] : s_blk[(r < EPT) ? q - (r0 << (LOGN - 1)) : R * (N / 2) + lt];
					wmadd<T, SPLIT>(hr[r], hi[r], w, x1.x, x1.y);
				}
			}
			__syncthreads();  // the next far flip (or the staging below) overwrites the parked partners
		}
	}
	if constexpr(HERM && TPA > 1) {
		// row band: pass 1 stores the band's own elements (one contiguous slice); pass 2 fills the rest of its
		// rows with conjugates of elements stored in other rows, lanes walking the band rows at a fixed
		// column so that the loads of one column are contiguous
		if(live) {
			const int r0 = ebase >> LOGN;
			const V2* __restrict__ ownh = cur + nbase;
#pragma unroll
			for(int r = 0; r < RT; ++r) {
				if(!HX_SLOT_LIVE(r)) continue;
				int q, i, j;
				tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, q, i, j);
				if constexpr(PAIR) {
					if(r < EPT) {
						if(r & 1) continue;  // staged with slot r - 1: same row i, column j + 1 (mod N)
						const float4 v = *reinterpret_cast<const float4*>(ownh + q);
						s_blk[((i - r0) << LOGN) | j] = V2{v.x, v.y};
						s_blk[((i - r0) << LOGN) | ((j + 1) & (N - 1))] = V2{v.z, v.w};
						continue;
					}
				}
				s_blk[((i - r0) << LOGN) | j] = ownh[q];
			}
			for(int t = tid; t < kTile; t += THREADS) {
				const int c = t / R, il = t - (t / R) * R;
				const int i = r0 + il;
				const int d = (c - i) & (N - 1);
				if(d < N / 2 || (d == N / 2 && i < N / 2)) continue;  // stored in row i: pass 1
				bool cj;
				const int src = herm_index(i, c, LOGN, cj);
				V2 v = ownh[src];
				if(cj) v.y = -v.y;
				s_blk[(il << LOGN) | c] = v;
			}
		}
	} else if constexpr(HERM) {
		if(live) {
#pragma unroll
			for(int r = 0; r < RT; ++r) {
				if(!HX_SLOT_LIVE(r)) continue;
				int q, i, j;
				tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, q, i, j);
				const V2 v = ownv[r];
				s_blk[slot * TE + ((i << LOGN) | j)] = v;
				if(i != j) s_blk[slot * TE + ((j << LOGN) | i)] = V2{v.x, -v.y};
			}
		}
	} else {
#pragma unroll
		for(int r = 0; r < EPT; ++r) {
			const int le = lt + TA * r;
			s_blk[slot * TE + le] = live ? cur[(static_cast<size_t>(n) << LOGE) + ebase + le] : V2{T(0), T(0)};
		}
	}
	__syncthreads();
	if(!live) return;

	T ar[RT], ai[RT];
	const V2* __restrict__ blk = s_blk + slot * TE;
	const V2* __restrict__ own = cur + (static_cast<size_t>(n) << LOGE);  // full layout, N > 64 only
	// -i[H, rho_n] from LDS: the diagonal, then flip by flip (each coefficient prepared once for all elements)
#pragma unroll
	for(int r = 0; r < RT; ++r) {
		if(!HX_SLOT_LIVE(r)) continue;
		int q, i, j;
		tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, q, i, j);
		const V2 c = blk[((i << LOGN) | j) - ebase];
		rmadd<T, SPLIT>(hr[r], hi[r], s_hd[i] - s_hd[j], c.x, c.y);
	}
	for(int f = 0; f < a.hflips; ++f) {
		const int m = a.hmask[f];
		const int rf = m << LOGN;  // row flip stays in the tile unless it crosses tiles
		const Weight<T, SPLIT> w = Weight<T, SPLIT>::make(a.hcoef[f]);
#pragma unroll
		for(int r = 0; r < RT; ++r) {
			if(!HX_SLOT_LIVE(r)) continue;
			int q, i, j;
			tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, q, i, j);
			const int e = (i << LOGN) | j;
			const int le = e - ebase;
			// LDS gathers by byte offset: (le ^ f) * size == (le * size) ^ (f * size), one XOR per gather
			const char* __restrict__ bb = reinterpret_cast<const char*>(blk);
			const unsigned lb = static_cast<unsigned>(le) * sizeof(V2);
			V2 x1;
			if(TPA == 1 || rf < kTile) {
				x1 = *reinterpret_cast<const V2*>(bb + (lb ^ (static_cast<unsigned>(rf) * sizeof(V2))));
			} else if constexpr(HERM) {  // row i^m lies outside the band: its part was added before the staging
				x1 = V2{T(0), T(0)};
			} else {
				x1 = own[e ^ rf];
			}
			const V2 x2 = *reinterpret_cast<const V2*>(bb + (lb ^ (static_cast<unsigned>(m) * sizeof(V2))));
			wmadd<T, SPLIT>(hr[r], hi[r], w, x1.x - x2.x, x1.y - x2.y);
		}
	}
#pragma unroll
	for(int r = 0; r < RT; ++r) {
		ar[r] = hi[r];
		ai[r] = -hr[r];
	}

	// neighbours: (-i)[V_b, rho_{n+e_k}] and down_comm [V_b, rho_{n-e_k}] + down_anti {V_b, rho_{n-e_k}}
	if constexpr(TABLE) {
		// each element's pattern word (bits 2s+1, 2s = bit s of i, bit s of j); an entry holds 2s, so its weight
		// for the element is a two-bit field of the word and one LDS read
		unsigned pw[RT];
#pragma unroll
		for(int r = 0; r < RT; ++r) {
			pw[r] = 0;
			if(!HX_SLOT_LIVE(r)) continue;
			int qe, i, j;
			tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
			pw[r] = SPREAD_LDS ? (s_spread[i] << 1) | s_spread[j] : (spread_bits(i) << 1) | spread_bits(j);
		}
		for(int c0 = 0;; c0 += NB) {
#pragma unroll
			for(int u = 0; u < NB; ++u) {
				if(c0 + u >= nvalid) continue;
				const WT* __restrict__ w4 = s_wt + (c0 + u) * 4;
#pragma unroll
				for(int r = 0; r < RT; ++r) {
					if(!HX_SLOT_LIVE(r)) continue;
					pmadd(ar[r], ai[r], w4[(pw[r] >> shb[u]) & 3u], xb[u][r].x, xb[u][r].y);
				}
			}
			if(c0 + NB >= nvalid) break;
			load_batch();
		}
	} else {
		nvalid = s_cnt[slot];
		for(int c0 = 0; c0 < nvalid; c0 += NB) {
			int nid[NB], nq[NB];
			double2 wc[NB], wa[NB];
			V2 x[NB][RT];
#pragma unroll
			for(int u = 0; u < NB; ++u) {
				const int packed = (c0 + u < nvalid) ? s_nb[slot][c0 + u] : -1;
				nid[u] = packed >> 5;  // -1 when past the list
				nq[u] = packed & 31;
				const int q = nq[u];
				wc[u] = make_double2(0.0, -1.0);  // up: weight -i on the commutator
				wa[u] = make_double2(0.0, 0.0);
#pragma unroll
				for(int r = 0; r < RT; ++r) x[u][r] = V2{T(0), T(0)};
				if(nid[u] < 0) continue;
				if(q >= M) {
					wc[u] = a.dcomm[n * M + (q - M)];
					wa[u] = a.danti[n * M + (q - M)];
				}
				const V2* __restrict__ src = cur + static_cast<size_t>(nid[u]) * PE;
#pragma unroll
				for(int r = 0; r < RT; ++r) {
					if(!HX_SLOT_LIVE(r)) continue;
					int qe, i, j;
					tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
					if constexpr(PAIR) {
						if(r < EPT) {
							if(r & 1) continue;  // loaded with slot r - 1
							const float4 v = *reinterpret_cast<const float4*>(src + qe);
							x[u][r] = V2{v.x, v.y};
							x[u][r + 1] = V2{v.z, v.w};
							continue;
						}
					}
					x[u][r] = src[qe];
				}
			}
#pragma unroll
			for(int u = 0; u < NB; ++u) {
				if(nid[u] < 0) continue;
				const int q = nq[u];
				const int k = q >= M ? q - M : q;
				if constexpr(SZ) {
					// d = v_i - v_j and s = v_i + v_j are +-2 alpha or 0, exactly one of them non-zero; with
					// sg = +-2 alpha from bit i: bits differ -> d = sg, s = 0; bits equal -> d = 0, s = sg.
					// Up: (-i) d x.  Down: i wc.y d x + wa.x s x (down_comm imaginary, down_anti real).
					// The two weights 2 alpha wc.y and 2 alpha wa.x are formed once; bit i only flips the sign.
					const int mask = s_mb[k];
					const double two_alpha = s_ka[k];
					const Weight<T, SPLIT> wd = Weight<T, SPLIT>::make(two_alpha * wc[u].y);
					const Weight<T, SPLIT> ws = Weight<T, SPLIT>::make(two_alpha * wa[u].x);
#pragma unroll
					for(int r = 0; r < RT; ++r) {
						if(!HX_SLOT_LIVE(r)) continue;
						int qe, i, j;
						tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
						const int bi = i & mask, bj = j & mask;
						const bool differ = bi != bj;
						const Weight<T, SPLIT> w = ws.select(differ, wd, bi == 0);
						// commutator part multiplies by i: (x.x, x.y) -> (-x.y, x.x)
						const T re = differ ? -x[u][r].y : x[u][r].x;
						const T im = differ ? x[u][r].x : x[u][r].y;
						wmadd<T, SPLIT>(ar[r], ai[r], w, re, im);
					}
				} else {
					const int b = s_mb[k];
					const double* __restrict__ vb = s_v + b * N;
#pragma unroll
					for(int r = 0; r < RT; ++r) {
						if(!HX_SLOT_LIVE(r)) continue;
						int qe, i, j;
						tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
						const double vi = vb[i], vj = vb[j];
						const double d = vi - vj, sm = vi + vj;
						cmadd<T, SPLIT>(ar[r], ai[r], wc[u].x * d + wa[u].x * sm, wc[u].y * d + wa[u].y * sm, x[u][r].x,
							x[u][r].y);
					}
				}
			}
		}
	}

	// Sigma, Xi, counter; Taylor epilogue.  In the last stage t_{m-1} is this tile's own LDS block
	// (tprev[ntprev-1] == cur), and for order 1 so is rho.
	// BATCHED: the last stage's sum rho + t_1 + ... + t_m issues the loads of all slots together, first rho, then each
	// earlier Taylor term, and stores last, so no load waits behind another slot's store or sum (per slot, one load
	// and one wait per term).  Kept for the complex64 row bands (three slots; N = 256 last stage 5.5% faster) and the
	// paired complex64 N = 64 blocks of single problems (last stage 0.125 -> 0.122 ms, split-FP32 0.133 -> 0.123 ms;
	// SINGLE: in the batched kernels the same sum spilled 28 bytes and their N = 64 last stage ran 13% slower) and the
	// complex128 N = 64 blocks (single problems 0.234 -> 0.228 ms, 512-instance batch +0.8%; it spilled, and ran 5%
	// slower, until the epilogue re-read its own element from LDS).
	constexpr bool BATCHED = LAST && HERM && ((sizeof(T) == 4 && ((RT == 3 && TPA > 1) || (PAIR && TPA == 1 && A == 1 && SINGLE))) || (sizeof(T) == 8 && TPA == 1 && A == 1));
	const T sc = static_cast<T>(a.scale);
	const double2 sf = a.selfc[n];
	V2* __restrict__ out = static_cast<V2*>(a.out);
	constexpr int EPI_UNROLL = PAIR ? RT : 2;  // PAIR: whole unroll, the odd slot stores its pair
#pragma unroll EPI_UNROLL
	for(int r = 0; r < RT; ++r) {
		if(!HX_SLOT_LIVE(r)) continue;
		int qe, i, j;
		tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
		const int e = (i << LOGN) | j;
		double wr = sf.x, wi = sf.y;
		if constexpr(SZ) {
			const double2 x = s_selfx[i ^ j];
			wr += x.x;
			wi += x.y;
		} else {
			for(int b = 0; b < B; ++b) {
				const double vi = s_v[b * N + i], vj = s_v[b * N + j];
				const double d = vi - vj, s = vi + vj;
				wr -= a.xi[b].x * d * d + a.xic[b].x * s * d;
				wi -= a.xi[b].y * d * d + a.xic[b].y * s * d;
			}
		}
		if constexpr(!HERM) {
			if(a.kappa) {
				const int sl = a.kslot[n];
				if(sl >= 0) {
					const double2 kk = a.kappa[(static_cast<size_t>(sl) << LOGE) + e];
					wr += kk.x;
					wi += kk.y;
				}
			}
		}
		// complex128 tiles and last stages re-read their own element from the LDS tile (index laundered through an
		// empty asm statement): merged with the H phase's read of it, the value stays live through the neighbour loop
		// (the complex64 last stage spilled it; complex128 N = 64 stage 0.1735 -> 0.169 ms re-reading; the complex64
		// inner stages ran 1.5-3.5% slower re-reading and keep the merged value)
		int le = e - ebase;
		if constexpr(sizeof(T) == 8 || LAST) asm volatile("" : "+v"(le));
		const V2 c = blk[le];
		cmadd<T, SPLIT>(ar[r], ai[r], wr, wi, c.x, c.y);
		const T tr = ar[r] * sc, ti = ai[r] * sc;
		const size_t g = nbase + qe;
		if constexpr(!LAST) {
			// complex64 inner stages: ordinary stores measured 3% faster than non-temporal ones on the N = 64
			// Hermitian tile (their output is the next stage's input); complex128 keeps the streaming store (3.5% the
			// other way)
			if constexpr(PAIR) {
				if(r < EPT) {
					ar[r] = tr;  // the pair (r - 1, r) is stored by its odd slot
					ai[r] = ti;
					if(r & 1)
						*reinterpret_cast<float4*>(out + g - 1) = float4{ar[r - 1], ai[r - 1], tr, ti};
				} else {
					out[g] = V2{tr, ti};
				}
			} else if constexpr(sizeof(T) == 4) {
				out[g] = V2{tr, ti};
			} else {
				st_stream<T>(out + g, tr, ti);
			}
		} else {
			const int nt = a.ntprev;
			if(a.rho0_f64 != nullptr && n == 0) {
				double2* r0 = reinterpret_cast<double2*>(a.rho0_f64);
				double2 rv = r0[qe];
				for(int q = 0; q + 1 < nt; ++q) {
					const V2 t = ld_stream<T>(static_cast<const V2*>(a.tprev[q]) + g);
					rv.x += static_cast<double>(t.x);
					rv.y += static_cast<double>(t.y);
				}
				if(nt > 0) {
					rv.x += static_cast<double>(c.x);
					rv.y += static_cast<double>(c.y);
				}
				rv.x += static_cast<double>(tr);
				rv.y += static_cast<double>(ti);
				r0[qe] = rv;
				out[g] = V2{static_cast<T>(rv.x), static_cast<T>(rv.y)};
			} else if constexpr(BATCHED) {
				ar[r] = tr;  // t_m of the slot, summed below
				ai[r] = ti;
			} else {
				bool paired = false;
				if constexpr(PAIR) {
					if(r < EPT) {
						paired = true;
						// the odd slot sums and stores the pair (r - 1, r) with 16-byte accesses, each element's terms
						// added in the order of the single-element sum below
						ar[r] = tr;
						ai[r] = ti;
						if(r & 1) {
							int q0, i0, j0;
							tile_elem<LOGN, TA, EPT, HERM, PAIR>(r - 1, lt, ebase, q0, i0, j0);
							const V2 c0 = blk[((i0 << LOGN) | j0) - ebase];
							const size_t g0 = g - 1;
							float4 rv = (nt == 0) ? float4{c0.x, c0.y, c.x, c.y} : ld_stream4(static_cast<const V2*>(a.rho) + g0);
							for(int q = 0; q + 1 < nt; ++q) {
								const float4 t = ld_stream4(static_cast<const V2*>(a.tprev[q]) + g0);
								rv.x += t.x;
								rv.y += t.y;
								rv.z += t.z;
								rv.w += t.w;
							}
							if(nt > 0) {
								rv.x += c0.x;
								rv.y += c0.y;
								rv.z += c.x;
								rv.w += c.y;
							}
							rv.x += ar[r - 1];
							rv.y += ai[r - 1];
							rv.z += tr;
							rv.w += ti;
							st_stream4(out + g0, rv);
						}
					}
				}
				if(!paired) {
					V2 rv = (nt == 0) ? c : ld_stream<T>(static_cast<const V2*>(a.rho) + g);
					for(int q = 0; q + 1 < nt; ++q) {
						const V2 t = ld_stream<T>(static_cast<const V2*>(a.tprev[q]) + g);
						rv.x += t.x;
						rv.y += t.y;
					}
					if(nt > 0) {
						rv.x += c.x;
						rv.y += c.y;
					}
					rv.x += tr;
					rv.y += ti;
					st_stream<T>(out + g, rv.x, rv.y);
				}
			}
		}
	}
	if constexpr(BATCHED) {
		const int nt = a.ntprev;
		if(!(a.rho0_f64 != nullptr && n == 0)) {
			unsigned g[RT];  // element offsets (a state holds fewer than 2^31 elements)
			V2 rv[RT];
			// slots' loads of one term: PAIR slots (2p, 2p + 1) as one 16-byte load
			auto load_term = [&](const V2* __restrict__ base, V2 (&dst)[RT]) {
#pragma unroll
				for(int r = 0; r < RT; ++r) dst[r] = V2{T(0), T(0)};  // before the loads: a pair fills two slots
#pragma unroll
				for(int r = 0; r < RT; ++r) {
					if(!HX_SLOT_LIVE(r)) continue;
					if constexpr(PAIR) {
						if(r < EPT) {
							if(r & 1) continue;  // loaded with slot r - 1
							const float4 v = ld_stream4(base + g[r]);
							dst[r] = V2{v.x, v.y};
							dst[r + 1] = V2{v.z, v.w};
							continue;
						}
					}
					dst[r] = ld_stream<T>(base + g[r]);
				}
			};
#pragma unroll
			for(int r = 0; r < RT; ++r) {
				g[r] = 0;
				if(!HX_SLOT_LIVE(r)) continue;
				int qe, i, j;
				tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
				g[r] = static_cast<unsigned>(n) * static_cast<unsigned>(PE) + static_cast<unsigned>(qe);
			}
			if(nt == 0) {
#pragma unroll
				for(int r = 0; r < RT; ++r) {
					rv[r] = V2{T(0), T(0)};
					if(!HX_SLOT_LIVE(r)) continue;
					int qe, i, j;
					tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
					rv[r] = blk[((i << LOGN) | j) - ebase];
				}
			} else {
				load_term(static_cast<const V2*>(a.rho), rv);
			}
			for(int q = 0; q + 1 < nt; ++q) {
				V2 t[RT];
				load_term(static_cast<const V2*>(a.tprev[q]), t);
#pragma unroll
				for(int r = 0; r < RT; ++r) {
					rv[r].x += t[r].x;
					rv[r].y += t[r].y;
				}
			}
#pragma unroll
			for(int r = 0; r < RT; ++r) {
				if(!HX_SLOT_LIVE(r)) continue;
				int qe, i, j;
				tile_elem<LOGN, TA, EPT, HERM, PAIR>(r, lt, ebase, qe, i, j);
				if(nt > 0) {
					int le = ((i << LOGN) | j) - ebase;
					asm volatile("" : "+v"(le));
					const V2 c = blk[le];
					rv[r].x += c.x;
					rv[r].y += c.y;
				}
				if constexpr(PAIR) {
					if(r < EPT) {
						if(r & 1)  // the pair (r - 1, r), both complete
							st_stream4(out + g[r - 1], float4{rv[r - 1].x + ar[r - 1], rv[r - 1].y + ai[r - 1],
								rv[r].x + ar[r], rv[r].y + ai[r]});
						continue;
					}
				}
				st_stream<T>(out + g[r], rv[r].x + ar[r], rv[r].y + ai[r]);
			}
		}
	}
#undef HX_SLOT_LIVE
}

// waves per SIMD the tile kernel is compiled for: two workgroups per CU, one when the complex128 tile and the
// N = 1024 self table (80 KiB of LDS) leave room for only one.  The 512-thread complex64 Hermitian sigma^z tile of
// one N = 64 ADO (36 KiB of LDS) is held to 64 VGPRs so that four workgroups -- four ADOs -- share a CU: its
// stage is bound by each workgroup's memory round trips, and twice the ADOs in flight hide twice as many.
// (Row tiles of N = 128, 256 held the same way spill and run slower than their 1024-thread shape.)
template <typename T, int LOGN, int THREADS, bool SZ, bool HERM>
constexpr int tile_waves()
{
	return (sizeof(T) == 8 && LOGN >= 10) ? THREADS / 256
	     : (sizeof(T) == 4 && THREADS == 512 && SZ && HERM && LOGN == 6) ? THREADS / 64
	                                                                      : THREADS / 128;
}

template <typename T, bool SPLIT, bool LAST, int LOGN, int THREADS, bool SZ, bool HERM>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(tile_waves<T, LOGN, THREADS, SZ, HERM>(),
	tile_waves<T, LOGN, THREADS, SZ, HERM>()))) void
k_stage_tile(const StageArgs a, int ngroups)
{
	using V2 = typename C2t<T>::type;
	constexpr int N = 1 << LOGN;
	constexpr int E = 1 << (2 * LOGN);
	constexpr int A = (E >= kTile) ? 1 : kTile / E;
	__shared__ V2 s_blk[kTile];
	__shared__ double s_v[SZ ? 1 : kTileMaxModes * N];
	__shared__ double2 s_selfx[SZ ? N : 1];
	__shared__ int s_nb[A * 2 * kTileMaxModes];
	__shared__ int s_cnt[A];
	__shared__ int s_mb[kTileMaxModes];
	__shared__ double s_ka[kTileMaxModes];
	__shared__ double s_hd[N];
	constexpr bool TABLE = SZ && A == 1;
	__shared__ typename PatternWeight<T, SPLIT>::type s_wt[TABLE ? 2 * kTileMaxModes * 4 : 1];
	__shared__ int s_spread[(TABLE && LOGN <= 8) ? N : 1];
	// XCD-aware renumbering: dispatch slot w runs on XCD w % 8
	const int per_xcd = (ngroups + kXcds - 1) / kXcds;
	const int group = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
	if(group >= ngroups) return;
	tile_body<T, SPLIT, LAST, LOGN, THREADS, SZ, HERM>(a, group,
		TileLds<T>{s_blk, s_v, s_selfx, s_nb, s_cnt, s_mb, s_ka, s_hd, s_wt, s_spread});
}

// Batched stage: one launch per block size N covers every (instance, tile) task of the batch's instances of that
// size (independent problems with N <= 64 and sigma^z couplings); table[q] is instance q's argument block for
// this stage, tasks[t] = (instance, tile group).  Tasks are numbered instance by instance, so the XCD
// renumbering keeps one instance's tiles on one XCD.  Specialising the launch by N sizes its registers and LDS
// to that tile alone, so the complex64 N = 64 tile gets its four-per-CU shape here too.
constexpr int kBatchMinLogN = 3;  // block sizes N = 8 .. 64: four groups

template <typename T, bool SPLIT, bool LAST, int THREADS, bool HERM, int LOGN>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(tile_waves<T, LOGN, THREADS, true, HERM>(),
	tile_waves<T, LOGN, THREADS, true, HERM>()))) void
k_batch_stage(const StageArgs* __restrict__ table, const int2* __restrict__ tasks, int ntasks)
{
	using V2 = typename C2t<T>::type;
	constexpr int N = 1 << LOGN;
	constexpr int E = 1 << (2 * LOGN);
	constexpr int A = (E >= kTile) ? 1 : kTile / E;
	constexpr bool TABLE = A == 1;
	__shared__ V2 s_blk[kTile];
	__shared__ double s_v[1];
	__shared__ double2 s_selfx[N];
	__shared__ int s_nb[A * 2 * kTileMaxModes];
	__shared__ int s_cnt[A];
	__shared__ int s_mb[kTileMaxModes];
	__shared__ double s_ka[kTileMaxModes];
	__shared__ double s_hd[N];
	__shared__ typename PatternWeight<T, SPLIT>::type s_wt[TABLE ? 2 * kTileMaxModes * 4 : 1];
	__shared__ int s_spread[TABLE ? N : 1];
	const int per_xcd = (ntasks + kXcds - 1) / kXcds;
	const int t = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
	if(t >= ntasks) return;
	const int2 tk = tasks[t];
	tile_body<T, SPLIT, LAST, LOGN, THREADS, true, HERM, false>(table[tk.x], tk.y,
		TileLds<T>{s_blk, s_v, s_selfx, s_nb, s_cnt, s_mb, s_ka, s_hd, s_wt, s_spread});
}

// Persistent batched stage: one launch per Taylor stage covers the tiles of every block size.  The grid is the
// number of workgroups the GPU holds at once; workgroups are dealt to the block sizes in proportion to their stored
// bytes (slots [wg_off[g], wg_off[g+1]) serve block size N = 2^(3+g), each a multiple of 8), and within its block
// size workgroup w serves the queue of XCD w % 8 (dispatch slot w runs on XCD w % 8), taking that queue's next tile
// from an atomic counter until the queue is empty.  Each workgroup runs one block size's loop only, so the four
// specialised tile bodies are separate code paths.  The thread index reaches the tile body through an empty asm
// statement in every iteration, so the compiler recomputes each task's per-slot element indices instead of keeping
// them live across tasks (hoisted, they made the kernel spill 100-180 VGPRs; each body alone needs 54-60).
// Instances are assigned whole to a queue, balanced by stored bytes, so an instance's neighbour blocks are read
// through one L2.  Task = (instance, tile group).  The workgroups of slots 0..31 reset the next stage's counters
// (that launch starts after this one ends), so every launch finds its own counters at zero.
struct BatchQueues {
	const int2* tasks;    // queue (g, x) is tasks[qoff[8 g + x] .. qoff[8 g + x + 1])
	const int* qoff;      // [4 * 8 + 1]
	int wg_off[5];        // workgroup slots of each block size
};

template <typename T, bool SPLIT, bool LAST, int THREADS, bool HERM, int LOGN>
__device__ __forceinline__ void run_queue(const StageArgs* __restrict__ table, const BatchQueues& bq, int x,
	unsigned* __restrict__ counter, int* s_task, const TileLds<T>& lds)
{
	constexpr int g = LOGN - kBatchMinLogN;
	const int begin = bq.qoff[g * kXcds + x], count = bq.qoff[g * kXcds + x + 1] - begin;
	for(;;) {
		if(threadIdx.x == 0) *s_task = static_cast<int>(atomicAdd(counter, 1u));
		__syncthreads();
		const int t = *s_task;
		if(t >= count) break;
		const int2 tk = bq.tasks[begin + t];
		int tid = static_cast<int>(threadIdx.x);
		asm volatile("" : "+v"(tid));
		tile_body<T, SPLIT, LAST, LOGN, THREADS, true, HERM, false>(table[tk.x], tk.y, lds, tid);
		__syncthreads();  // the next tile stages into the LDS this one read (and thread 0 rewrites s_task)
	}
}

// two workgroups per CU: the one-ADO N = 64 complex64 body alone fits four (tile_waves), but with the other three
// tile bodies in the same kernel the registers allow two
template <typename T, bool SPLIT, bool LAST, int THREADS, bool HERM>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(THREADS / 128, THREADS / 128))) void
k_batch_persistent(const StageArgs* __restrict__ table, const BatchQueues bq, unsigned* __restrict__ counters,
	unsigned* __restrict__ next_counters)
{
	using V2 = typename C2t<T>::type;
	constexpr int kNb = 64 * 2 * kTileMaxModes;  // neighbour-list entries of the tile with the most ADOs (N = 8)
	__shared__ V2 s_blk[kTile];
	__shared__ double s_v[1];
	__shared__ double2 s_selfx[64];
	// tiles of several ADOs keep their neighbour lists here, tiles of one ADO their weight table
	__shared__ double2 s_nbwt[kNb / 4];
	__shared__ int s_cnt[64];
	__shared__ int s_mb[kTileMaxModes];
	__shared__ double s_ka[kTileMaxModes];
	__shared__ double s_hd[64];
	__shared__ int s_spread[64];
	__shared__ int s_task;
	static_assert(2 * kTileMaxModes * 4 * sizeof(typename PatternWeight<T, SPLIT>::type) <= sizeof(s_nbwt), "weight table");
	const int w = blockIdx.x;
	if(w < 4 * kXcds && threadIdx.x == 0) next_counters[w] = 0u;
	const int x = w % kXcds;
	const TileLds<T> lds{s_blk, s_v, s_selfx, reinterpret_cast<int*>(s_nbwt), s_cnt, s_mb, s_ka, s_hd, s_nbwt, s_spread};
	if(w >= bq.wg_off[3])
		run_queue<T, SPLIT, LAST, THREADS, HERM, 6>(table, bq, x, counters + 3 * kXcds + x, &s_task, lds);
	else if(w >= bq.wg_off[2])
		run_queue<T, SPLIT, LAST, THREADS, HERM, 5>(table, bq, x, counters + 2 * kXcds + x, &s_task, lds);
	else if(w >= bq.wg_off[1])
		run_queue<T, SPLIT, LAST, THREADS, HERM, 4>(table, bq, x, counters + 1 * kXcds + x, &s_task, lds);
	else
		run_queue<T, SPLIT, LAST, THREADS, HERM, 3>(table, bq, x, counters + x, &s_task, lds);
}

// Dense-Hamiltonian stage for N in {16, 32, 64}: one ADO per 256-thread workgroup.
// rho_n and H sit in LDS with a padded row stride (N+1) so the 16 rows an MFMA operand read touches
// fall in different banks.  G = H rho - rho H is computed as real 16x16x4 MFMA products of the real
// and imaginary planes (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32; the split-FP32 form adds the
// products with the FP32 remainder of H), each wave owning whole 16x16 output tiles.  The result
// -i G + (Sigma + Xi) rho replaces rho_n in LDS; the neighbour terms and the Taylor epilogue follow as
// in the tiled kernel.  MFMA operand layout (16x16x4, measured on gfx950): A[i][k] and B[k][j] from
// lane l with i (or j) = l % 16, k = l / 16; D[i][j] in register v of lane l with j = l % 16 and
// i = 4 (l / 16) + v for f32 but i = l / 16 + 4 v for f64.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
	using acc = double __attribute__((ext_vector_type(4)));
	__device__ static acc mma(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
	__device__ static int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Mfma<float> {
	using acc = float __attribute__((ext_vector_type(4)));
	__device__ static acc mma(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
	__device__ static int row(int lk, int v) { return 4 * lk + v; }
};

constexpr int kDenseThreads = 256;

// Whole-block dense stage: at N = 64 its two padded matrices fill the LDS, so a CU holds one workgroup, and 1024
// threads (16 waves, one 16 x 16 output tile each, four elements each in the neighbour terms) keep 4 waves on every
// SIMD to hide the MFMA chains and the neighbour round trips (1.47x the 256-thread shape in FP64, 1.50x complex64).
// Smaller blocks keep 256 threads (N = 32 at 1024 threads: one element each, 12 waves without a tile, 1.2x slower).
template <int LOGN>
constexpr int dense_threads()
{
	return LOGN >= 6 ? 1024 : kDenseThreads;
}

template <typename T, bool SPLIT, bool LAST, int LOGN, bool SZ>
__global__ __launch_bounds__(dense_threads<LOGN>()) void k_stage_dense(const StageArgs a)
{
	using V2 = typename C2t<T>::type;
	using MF = Mfma<T>;
	using acc4 = typename MF::acc;
	constexpr int N = 1 << LOGN;
	constexpr int LOGE = 2 * LOGN;
	constexpr int E = 1 << LOGE;
	constexpr int P = N + 1;
	constexpr int DT = dense_threads<LOGN>();
	constexpr int NW = DT / 64;  // waves
	constexpr int EPT = E / DT;
	constexpr int NT = N / 16;
	constexpr int TILES = NT * NT;
	constexpr int MAXT = (TILES + NW - 1) / NW;
	__shared__ V2 s_rho[N * P];
	__shared__ V2 s_h[N * P];
	__shared__ V2 s_hlo[SPLIT ? N * P : 1];
	__shared__ double s_v[SZ ? 1 : kTileMaxModes * N];
	__shared__ double2 s_selfx[SZ ? N : 1];
	__shared__ int s_nb[2 * kTileMaxModes];

	const int H = a.n_ado, M = a.M, B = a.B;
	const int per_xcd = (H + kXcds - 1) / kXcds;
	const int n = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
	if(n >= H) return;
	const int tid = threadIdx.x;
	const V2* __restrict__ cur = static_cast<const V2*>(a.cur);
	const V2* __restrict__ own = cur + (static_cast<size_t>(n) << LOGE);
	const V2* __restrict__ hd = static_cast<const V2*>(a.hdense);
	const V2* __restrict__ hl = static_cast<const V2*>(a.hdense_lo);

	if constexpr(SZ) {
		for(int q = tid; q < N; q += DT) s_selfx[q] = a.selfx[q];
	} else {
		for(int q = tid; q < B * N; q += DT) s_v[q] = a.vdiag[q];
	}
	for(int q = tid; q < 2 * M; q += DT) {
		const int v = (q < M) ? a.up[n * M + q] : a.down[n * M + (q - M)];
		s_nb[q] = (v < H) ? v : -1;
	}
	for(int e = tid; e < E; e += DT) {
		const int i = e >> LOGN, j = e & (N - 1);
		s_rho[i * P + j] = own[e];
		s_h[i * P + j] = hd[e];
		if constexpr(SPLIT) s_hlo[i * P + j] = hl[e];
	}
	__syncthreads();

	// G = H rho - rho H on MFMA
	const int wave = tid >> 6, lane = tid & 63;
	const int li = lane & 15, lk = lane >> 4;
	// at N = 64 each tile accumulates alternate K steps in two register pairs (four independent MFMA chains per
	// wave), summed after; smaller blocks, whose stages are not bound by the chains, keep one pair
	constexpr int KS = LOGN >= 6 ? 2 : 1;
	acc4 gr[MAXT], gi[MAXT];
#pragma unroll
	for(int tt = 0; tt < MAXT; ++tt) {
		const int t = wave + NW * tt;
		gr[tt] = acc4{};
		gi[tt] = acc4{};
		if(t >= TILES) continue;
		const int I0 = (t / NT) * 16, J0 = (t % NT) * 16;
		acc4 cr[KS], ci[KS];
#pragma unroll
		for(int h = 0; h < KS; ++h) {
			cr[h] = acc4{};
			ci[h] = acc4{};
		}
		for(int k0 = 0; k0 < N; k0 += 4 * KS) {
#pragma unroll
			for(int h = 0; h < KS; ++h) {
				const int kk = k0 + 4 * h + lk;
				const V2 rb = s_rho[kk * P + J0 + li];
				const V2 ra = s_rho[(I0 + li) * P + kk];
#pragma unroll
				for(int part = 0; part < (SPLIT ? 2 : 1); ++part) {
					const V2* __restrict__ hs = part ? s_hlo : s_h;
					const V2 ha = hs[(I0 + li) * P + kk];
					const V2 hb = hs[kk * P + J0 + li];
					// H rho
					cr[h] = MF::mma(ha.x, rb.x, cr[h]);
					ci[h] = MF::mma(ha.x, rb.y, ci[h]);
					// - rho H
					cr[h] = MF::mma(-ra.x, hb.x, cr[h]);
					ci[h] = MF::mma(-ra.y, hb.x, ci[h]);
					if(!a.hreal) {
						cr[h] = MF::mma(-ha.y, rb.y, cr[h]);
						ci[h] = MF::mma(ha.y, rb.x, ci[h]);
						cr[h] = MF::mma(ra.y, hb.y, cr[h]);
						ci[h] = MF::mma(-ra.x, hb.y, ci[h]);
					}
				}
			}
		}
		gr[tt] = cr[0];
		gi[tt] = ci[0];
#pragma unroll
		for(int h = 1; h < KS; ++h) {
			gr[tt] += cr[h];
			gi[tt] += ci[h];
		}
	}
	__syncthreads();  // every MFMA operand read of rho is done

	// rho_n <- -i G + (Sigma + Xi + counter) rho_n, at each wave's own tile positions
	const double2 sf = a.selfc[n];
#pragma unroll
	for(int tt = 0; tt < MAXT; ++tt) {
		const int t = wave + NW * tt;
		if(t >= TILES) continue;
		const int I0 = (t / NT) * 16, J0 = (t % NT) * 16;
#pragma unroll
		for(int v = 0; v < 4; ++v) {
			const int i = I0 + MF::row(lk, v), j = J0 + li;
			double wr = sf.x, wi = sf.y;
			if constexpr(SZ) {
				const double2 x = s_selfx[i ^ j];
				wr += x.x;
				wi += x.y;
			} else {
				for(int b = 0; b < B; ++b) {
					const double vi = s_v[b * N + i], vj = s_v[b * N + j];
					const double d = vi - vj, s = vi + vj;
					wr -= a.xi[b].x * d * d + a.xic[b].x * s * d;
					wi -= a.xi[b].y * d * d + a.xic[b].y * s * d;
				}
			}
			if(a.kappa) {
				const int sl = a.kslot[n];
				if(sl >= 0) {
					const double2 kk = a.kappa[(static_cast<size_t>(sl) << LOGE) + i * N + j];
					wr += kk.x;
					wi += kk.y;
				}
			}
			const V2 c = s_rho[i * P + j];
			T vr = gi[tt][v], vi_ = -gr[tt][v];
			cmadd<T, SPLIT>(vr, vi_, wr, wi, c.x, c.y);
			s_rho[i * P + j] = V2{vr, vi_};
		}
	}
	__syncthreads();

	T ar[EPT], ai[EPT];
#pragma unroll
	for(int r = 0; r < EPT; ++r) {
		const int e = tid + DT * r;
		const V2 c = s_rho[(e >> LOGN) * P + (e & (N - 1))];
		ar[r] = c.x;
		ai[r] = c.y;
	}

	for(int q = 0; q < 2 * M; ++q) {
		const int idx = s_nb[q];
		if(idx < 0) continue;
		const bool down = q >= M;
		const int k = down ? q - M : q;
		const int b = a.mode_bath[k];
		const V2* __restrict__ src = cur + (static_cast<size_t>(idx) << LOGE);
		V2 x[EPT];
#pragma unroll
		for(int r = 0; r < EPT; ++r) x[r] = src[tid + DT * r];
		if constexpr(SZ) {
			const int mask = a.sz_mask[b];
			const double two_alpha = 2.0 * a.sz_alpha[b];
			double wcomm = -1.0, wanti = 0.0;
			if(down) {
				wcomm = a.dcomm[n * M + k].y;
				wanti = a.danti[n * M + k].x;
			}
#pragma unroll
			for(int r = 0; r < EPT; ++r) {
				const int e = tid + DT * r;
				const int bi = (e >> LOGN) & mask, bj = e & mask;
				const bool differ = bi != bj;
				const double sg = bi ? two_alpha : -two_alpha;
				const double w = sg * (differ ? wcomm : wanti);
				const T u = differ ? -x[r].y : x[r].x;
				const T v = differ ? x[r].x : x[r].y;
				rmadd<T, SPLIT>(ar[r], ai[r], w, u, v);
			}
		} else {
			double2 wc = make_double2(0.0, -1.0), wa = make_double2(0.0, 0.0);
			if(down) {
				wc = a.dcomm[n * M + k];
				wa = a.danti[n * M + k];
			}
			const double* __restrict__ vb = s_v + b * N;
#pragma unroll
			for(int r = 0; r < EPT; ++r) {
				const int e = tid + DT * r;
				const double vi = vb[e >> LOGN], vj = vb[e & (N - 1)];
				const double d = vi - vj, s = vi + vj;
				cmadd<T, SPLIT>(ar[r], ai[r], wc.x * d + wa.x * s, wc.y * d + wa.y * s, x[r].x, x[r].y);
			}
		}
	}

	const T sc = static_cast<T>(a.scale);
	V2* __restrict__ out = static_cast<V2*>(a.out);
	if constexpr(LAST) {
		if(!(a.rho0_f64 != nullptr && n == 0)) {
			// rho_new = rho + t_1 + ... + t_m with the loads of all the thread's elements issued together, first rho,
			// then each earlier Taylor term, and the stores last (per element, one load and one wait per term made
			// the last stage 1.35x an inner one)
			V2 rv[EPT];
#pragma unroll
			for(int r = 0; r < EPT; ++r)
				rv[r] = ld_stream<T>(static_cast<const V2*>(a.rho) + (static_cast<size_t>(n) << LOGE) + tid + DT * r);
			for(int q = 0; q < a.ntprev; ++q) {
				const V2* __restrict__ tp = static_cast<const V2*>(a.tprev[q]) + (static_cast<size_t>(n) << LOGE);
				V2 t[EPT];
#pragma unroll
				for(int r = 0; r < EPT; ++r) t[r] = ld_stream<T>(tp + tid + DT * r);
#pragma unroll
				for(int r = 0; r < EPT; ++r) {
					rv[r].x += t[r].x;
					rv[r].y += t[r].y;
				}
			}
#pragma unroll
			for(int r = 0; r < EPT; ++r)
				st_stream<T>(out + (static_cast<size_t>(n) << LOGE) + tid + DT * r, rv[r].x + ar[r] * sc, rv[r].y + ai[r] * sc);
			return;
		}
	}
#pragma unroll
	for(int r = 0; r < EPT; ++r) {
		const int e = tid + DT * r;
		const T tr = ar[r] * sc, ti = ai[r] * sc;
		const size_t g = (static_cast<size_t>(n) << LOGE) + e;
		if constexpr(!LAST) {
			st_stream<T>(out + g, tr, ti);
		} else {
			if(a.rho0_f64 != nullptr && n == 0) {
				double2* r0 = reinterpret_cast<double2*>(a.rho0_f64);
				double2 rv = r0[e];
				for(int q = 0; q < a.ntprev; ++q) {
					const V2 t = ld_stream<T>(static_cast<const V2*>(a.tprev[q]) + g);
					rv.x += static_cast<double>(t.x);
					rv.y += static_cast<double>(t.y);
				}
				rv.x += static_cast<double>(tr);
				rv.y += static_cast<double>(ti);
				r0[e] = rv;
				out[g] = V2{static_cast<T>(rv.x), static_cast<T>(rv.y)};
			} else {
				V2 rv = ld_stream<T>(static_cast<const V2*>(a.rho) + g);
				for(int q = 0; q < a.ntprev; ++q) {
					const V2 t = ld_stream<T>(static_cast<const V2*>(a.tprev[q]) + g);
					rv.x += t.x;
					rv.y += t.y;
				}
				rv.x += tr;
				rv.y += ti;
				st_stream<T>(out + g, rv.x, rv.y);
			}
		}
	}
}

// Dense-Hamiltonian stage for N = 128 .. 1024 (full row-major blocks): one 64 x 64 output tile of one ADO per
// 256-thread workgroup, the tiles of one ADO on consecutive dispatch slots of one XCD (its block stays in that
// L2).  G = H rho - rho H of the tile is accumulated on 16x16x4 MFMA over K = N in chunks of 16 staged in LDS:
// the tile's 64 rows of H and of rho (A operands) and its 64 columns of rho and of H (B operands), each LDS row
// padded by one element so that the 16 rows an operand read touches fall in different banks.  Wave w owns the
// 32 x 32 quarter (w / 2, w % 2) as 2 x 2 MFMA tiles; the split-FP32 form adds the products with the FP32
// remainder of H, a real H skips the imaginary-H products.  The epilogue runs on the accumulator registers:
// register v of lane l in MFMA tile (ta, tb) holds element (I0 + 32 (w / 2) + 16 ta + row(l / 16, v),
// J0 + 32 (w % 2) + 16 tb + l % 16); it adds (Sigma + Xi + closure) rho_n, the neighbour terms (elementwise in
// the neighbour blocks) and applies the Taylor recurrence, as the whole-block kernel does.
constexpr int kDT = 64;  // output tile edge
constexpr int kDK = 16;  // K chunk staged per LDS round

template <typename T, bool SPLIT, bool LAST, int LOGN, bool SZ>
__global__ __launch_bounds__(kDenseThreads) void k_stage_dense_tiled(const StageArgs a, int ngroups)
{
	using V2 = typename C2t<T>::type;
	using MF = Mfma<T>;
	using acc4 = typename MF::acc;
	constexpr int N = 1 << LOGN;
	constexpr int LOGE = 2 * LOGN;
	constexpr int NTT = N / kDT;
	constexpr int TPA = NTT * NTT;
	static_assert(N >= 2 * kDT && N % kDK == 0, "tiled dense path: N >= 128");
	__shared__ V2 s_hr[kDT][kDK + 1];   // H[I0 + r][k0 + c]
	__shared__ V2 s_rr[kDT][kDK + 1];   // rho[I0 + r][k0 + c]
	__shared__ V2 s_rc[kDK][kDT + 1];   // rho[k0 + r][J0 + c]
	__shared__ V2 s_hc[kDK][kDT + 1];   // H[k0 + r][J0 + c]
	__shared__ V2 s_hrl[SPLIT ? kDT : 1][kDK + 1];
	__shared__ V2 s_hcl[SPLIT ? kDK : 1][kDT + 1];
	__shared__ double2 s_selfx[SZ ? N : 1];
	__shared__ int s_nb[2 * kTileMaxModes];

	const int per_xcd = (ngroups + kXcds - 1) / kXcds;
	const int group = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
	if(group >= ngroups) return;
	const int H = a.n_ado, M = a.M, B = a.B;
	const int n = group / TPA;
	const int tile = group - n * TPA;
	const int I0 = (tile / NTT) * kDT, J0 = (tile % NTT) * kDT;
	const int tid = threadIdx.x;
	const V2* __restrict__ cur = static_cast<const V2*>(a.cur);
	const V2* __restrict__ own = cur + (static_cast<size_t>(n) << LOGE);
	const V2* __restrict__ hd = static_cast<const V2*>(a.hdense);
	const V2* __restrict__ hl = static_cast<const V2*>(a.hdense_lo);

	if constexpr(SZ)
		for(int q = tid; q < N; q += kDenseThreads) s_selfx[q] = a.selfx[q];
	for(int q = tid; q < 2 * M; q += kDenseThreads) {
		const int v = (q < M) ? a.up[n * M + q] : a.down[n * M + (q - M)];
		s_nb[q] = (v < H) ? v : -1;
	}

	const int wave = tid >> 6, lane = tid & 63;
	const int li = lane & 15, lk = lane >> 4;
	const int WR = (wave >> 1) * 32, WC = (wave & 1) * 32;
	acc4 gr[2][2], gi[2][2];
#pragma unroll
	for(int ta = 0; ta < 2; ++ta)
#pragma unroll
		for(int tb = 0; tb < 2; ++tb) {
			gr[ta][tb] = acc4{};
			gi[ta][tb] = acc4{};
		}
	// staging: the A panels are 64 rows x 16 (thread: row tid / 4, four consecutive k), the B panels 16 rows x 64
	// (thread: row tid / 16, four consecutive columns)
	const int ar_ = tid >> 2, ac = (tid & 3) * 4;
	const int br = tid >> 4, bc = (tid & 15) * 4;
	for(int k0 = 0; k0 < N; k0 += kDK) {
		__syncthreads();  // the previous chunk's operand reads are done
#pragma unroll
		for(int u = 0; u < 4; ++u) {
			const size_t ea = static_cast<size_t>(I0 + ar_) * N + k0 + ac + u;
			const size_t eb = static_cast<size_t>(k0 + br) * N + J0 + bc + u;
			s_hr[ar_][ac + u] = hd[ea];
			s_rr[ar_][ac + u] = own[ea];
			s_rc[br][bc + u] = own[eb];
			s_hc[br][bc + u] = hd[eb];
			if constexpr(SPLIT) {
				s_hrl[ar_][ac + u] = hl[ea];
				s_hcl[br][bc + u] = hl[eb];
			}
		}
		__syncthreads();
#pragma unroll 1
		for(int kk = 0; kk < kDK; kk += 4) {
			const int k = kk + lk;
			V2 ha[2], ra[2], hal[2], rb[2], hb[2], hbl[2];
#pragma unroll
			for(int t = 0; t < 2; ++t) {
				ha[t] = s_hr[WR + 16 * t + li][k];
				ra[t] = s_rr[WR + 16 * t + li][k];
				rb[t] = s_rc[k][WC + 16 * t + li];
				hb[t] = s_hc[k][WC + 16 * t + li];
				if constexpr(SPLIT) {
					hal[t] = s_hrl[WR + 16 * t + li][k];
					hbl[t] = s_hcl[k][WC + 16 * t + li];
				}
			}
#pragma unroll
			for(int ta = 0; ta < 2; ++ta)
#pragma unroll
				for(int tb = 0; tb < 2; ++tb) {
#pragma unroll
					for(int part = 0; part < (SPLIT ? 2 : 1); ++part) {
						const V2 h_a = part ? hal[ta] : ha[ta];
						const V2 h_b = part ? hbl[tb] : hb[tb];
						// H rho
						gr[ta][tb] = MF::mma(h_a.x, rb[tb].x, gr[ta][tb]);
						gi[ta][tb] = MF::mma(h_a.x, rb[tb].y, gi[ta][tb]);
						// - rho H
						gr[ta][tb] = MF::mma(-ra[ta].x, h_b.x, gr[ta][tb]);
						gi[ta][tb] = MF::mma(-ra[ta].y, h_b.x, gi[ta][tb]);
						if(!a.hreal) {
							gr[ta][tb] = MF::mma(-h_a.y, rb[tb].y, gr[ta][tb]);
							gi[ta][tb] = MF::mma(h_a.y, rb[tb].x, gi[ta][tb]);
							gr[ta][tb] = MF::mma(ra[ta].y, h_b.y, gr[ta][tb]);
							gi[ta][tb] = MF::mma(-ra[ta].x, h_b.y, gi[ta][tb]);
						}
					}
				}
		}
	}

	// the lane's 16 tile elements, four at a time (one MFMA tile): L = -i G + (Sigma + Xi + closure) rho_n, the
	// neighbour terms, the Taylor recurrence
	const double2 sf = a.selfc[n];
	const int ks = a.kappa ? a.kslot[n] : -1;
	const T sc = static_cast<T>(a.scale);
	V2* __restrict__ out = static_cast<V2*>(a.out);
#pragma unroll
	for(int ta = 0; ta < 2; ++ta)
#pragma unroll
		for(int tb = 0; tb < 2; ++tb) {
			T ar[4], ai[4];
			int ei[4];
#pragma unroll
			for(int v = 0; v < 4; ++v) {
				const int i = I0 + WR + 16 * ta + MF::row(lk, v), j = J0 + WC + 16 * tb + li;
				const int e = (i << LOGN) | j;
				ei[v] = e;
				double wr = sf.x, wi = sf.y;
				if constexpr(SZ) {
					const double2 x = s_selfx[i ^ j];
					wr += x.x;
					wi += x.y;
				} else {
					for(int b = 0; b < B; ++b) {
						const double vi = a.vdiag[b * N + i], vj = a.vdiag[b * N + j];
						const double d = vi - vj, sm = vi + vj;
						wr -= a.xi[b].x * d * d + a.xic[b].x * sm * d;
						wi -= a.xi[b].y * d * d + a.xic[b].y * sm * d;
					}
				}
				if(ks >= 0) {
					const double2 kk = a.kappa[(static_cast<size_t>(ks) << LOGE) + e];
					wr += kk.x;
					wi += kk.y;
				}
				const V2 c = own[e];
				T vr = gi[ta][tb][v], vim = -gr[ta][tb][v];
				cmadd<T, SPLIT>(vr, vim, wr, wi, c.x, c.y);
				ar[v] = vr;
				ai[v] = vim;
			}
			for(int q = 0; q < 2 * M; ++q) {
				const int idx = s_nb[q];
				if(idx < 0) continue;
				const bool down = q >= M;
				const int k = down ? q - M : q;
				const int b = a.mode_bath[k];
				const V2* __restrict__ src = cur + (static_cast<size_t>(idx) << LOGE);
				V2 x[4];
#pragma unroll
				for(int v = 0; v < 4; ++v) x[v] = src[ei[v]];
				if constexpr(SZ) {
					const int mask = a.sz_mask[b];
					const double two_alpha = 2.0 * a.sz_alpha[b];
					double wcomm = -1.0, wanti = 0.0;
					if(down) {
						wcomm = a.dcomm[n * M + k].y;
						wanti = a.danti[n * M + k].x;
					}
#pragma unroll
					for(int v = 0; v < 4; ++v) {
						const int bi = (ei[v] >> LOGN) & mask, bj = ei[v] & mask;
						const bool differ = bi != bj;
						const double sg = bi ? two_alpha : -two_alpha;
						const double w = sg * (differ ? wcomm : wanti);
						const T u = differ ? -x[v].y : x[v].x;
						const T vv = differ ? x[v].x : x[v].y;
						rmadd<T, SPLIT>(ar[v], ai[v], w, u, vv);
					}
				} else {
					double2 wc = make_double2(0.0, -1.0), wa = make_double2(0.0, 0.0);
					if(down) {
						wc = a.dcomm[n * M + k];
						wa = a.danti[n * M + k];
					}
					const double* __restrict__ vb = a.vdiag + b * N;
#pragma unroll
					for(int v = 0; v < 4; ++v) {
						const double vi = vb[ei[v] >> LOGN], vj = vb[ei[v] & (N - 1)];
						const double d = vi - vj, sm = vi + vj;
						cmadd<T, SPLIT>(ar[v], ai[v], wc.x * d + wa.x * sm, wc.y * d + wa.y * sm, x[v].x, x[v].y);
					}
				}
			}
#pragma unroll
			for(int v = 0; v < 4; ++v) {
				const int e = ei[v];
				const T tr = ar[v] * sc, ti = ai[v] * sc;
				const size_t g = (static_cast<size_t>(n) << LOGE) + e;
				if constexpr(!LAST) {
					st_stream<T>(out + g, tr, ti);
				} else {
					if(a.rho0_f64 != nullptr && n == 0) {
						double2* r0 = reinterpret_cast<double2*>(a.rho0_f64);
						double2 rv = r0[e];
						for(int q = 0; q < a.ntprev; ++q) {
							const V2 t = ld_stream<T>(static_cast<const V2*>(a.tprev[q]) + g);
							rv.x += static_cast<double>(t.x);
							rv.y += static_cast<double>(t.y);
						}
						rv.x += static_cast<double>(tr);
						rv.y += static_cast<double>(ti);
						r0[e] = rv;
						out[g] = V2{static_cast<T>(rv.x), static_cast<T>(rv.y)};
					} else {
						V2 rv = ld_stream<T>(static_cast<const V2*>(a.rho) + g);
						for(int q = 0; q < a.ntprev; ++q) {
							const V2 t = ld_stream<T>(static_cast<const V2*>(a.tprev[q]) + g);
							rv.x += t.x;
							rv.y += t.y;
						}
						rv.x += tr;
						rv.y += ti;
						st_stream<T>(out + g, rv.x, rv.y);
					}
				}
			}
		}
}

// Physics terminator of the deepest kept tier (helix_hip_plan_set_terminator): virtual tier-(L+1) ADO u (block
// n_real + u of the stage input) solved from the stationary condition of its own equation, in the eigenbasis of H
// (include/helix_hip.h gives the equations).  One workgroup of N^2 threads per u, one element each; the four N x N
// products per round go through LDS.  Arithmetic in FP64 whatever the state precision.
struct TermArgs {
	void* cur;                // stage input; virtual ADO u is block n_real + u
	int n_real;
	int iterations;
	const double2* rate;      // [U]
	const int32_t* src_off;   // [U+1]
	const int32_t* src_ado;   // [S] stored index of the source
	const int32_t* src_mask;  // [S] bit mask of the spin of the edge's bath
	const double* src_alpha;  // [S] alpha of that bath (V_b = alpha (2 n_s - 1))
	const double2* src_comm;  // [S]
	const double2* src_anti;  // [S]
	const double2* q;         // [N][N] eigenvectors of H (columns)
	const double* ev;         // [N] eigenvalues
	const double2* selfx;     // [N] the plan's sigma^z Xi self table, indexed by i XOR j
	const double2* corr;      // [U][N][N] complex factor on each eigen-element's response, or null (factor 1)
};

// The four N x N products of each round are loops over k unrolled in groups of up to 16 terms (N is a template
// constant), so a thread issues a group's LDS reads before the FMAs that consume them and pays the LDS latency
// once per group rather than once per term.
template <int N>
__device__ __forceinline__ double2 row_times_col(const double2* __restrict__ a, int lda, const double2* __restrict__ b,
	int ldb, bool conj_a, bool conj_b)
{
	constexpr int G = N <= 16 ? N : 8;  // N = 32 runs 1024 threads: 8-term groups keep it within its registers
	double re = 0.0, im = 0.0;
#pragma unroll 1
	for(int k0 = 0; k0 < N; k0 += G) {
		double2 av[G], bv[G];
#pragma unroll
		for(int k = 0; k < G; ++k) {
			av[k] = a[(k0 + k) * lda];
			bv[k] = b[(k0 + k) * ldb];
		}
#pragma unroll
		for(int k = 0; k < G; ++k) {
			const double ar = av[k].x, ai = conj_a ? -av[k].y : av[k].y;
			const double br = bv[k].x, bi = conj_b ? -bv[k].y : bv[k].y;
			re += ar * br - ai * bi;
			im += ar * bi + ai * br;
		}
	}
	return make_double2(re, im);
}

template <typename T, int LOGN>
__global__ __launch_bounds__(1 << (2 * LOGN)) void k_terminator(const TermArgs t)
{
	using V2 = typename C2t<T>::type;
	constexpr int N = 1 << LOGN;
	constexpr int E = N * N;
	__shared__ double2 s_q[E];
	__shared__ double2 s_a[E];
	__shared__ double s_ev[N];
	const int e = threadIdx.x;
	const int i = e >> LOGN, j = e & (N - 1);
	const int u = blockIdx.x;
	s_q[e] = t.q[e];
	if(e < N) s_ev[e] = t.ev[e];
	V2* __restrict__ cur = static_cast<V2*>(t.cur);
	// D_u[ij] = sum_s (comm (v_i - v_j) + anti (v_i + v_j)) rho_src[ij]
	double dr = 0.0, di = 0.0;
	for(int s = t.src_off[u]; s < t.src_off[u + 1]; ++s) {
		const V2 x = cur[static_cast<size_t>(t.src_ado[s]) * E + e];
		const int m = t.src_mask[s];
		const double al = t.src_alpha[s];
		const double vi = (i & m) ? al : -al, vj = (j & m) ? al : -al;
		const double d = vi - vj, sm = vi + vj;
		const double2 c = t.src_comm[s], an = t.src_anti[s];
		const double wr = c.x * d + an.x * sm, wi = c.y * d + an.y * sm;
		const double xr = static_cast<double>(x.x), xi = static_cast<double>(x.y);
		dr += wr * xr - wi * xi;
		di += wr * xi + wi * xr;
	}
	const double2 rate = t.rate[u];
	const double2 sw = t.selfx[i ^ j];
	double2 corr = make_double2(1.0, 0.0);
	if(t.corr) corr = t.corr[static_cast<size_t>(u) * E + e];
	__syncthreads();  // s_q, s_ev
	// response of eigen-element (a, b) = (i, j): -corr / (rate - i (E_a - E_b)), the same every round
	double gr, gi;
	{
		const double den_r = rate.x, den_i = rate.y - (s_ev[i] - s_ev[j]);
		const double inv = 1.0 / (den_r * den_r + den_i * den_i);
		gr = -(corr.x * den_r + corr.y * den_i) * inv;
		gi = -(corr.y * den_r - corr.x * den_i) * inv;
	}
	double xr = 0.0, xi = 0.0;
	for(int it = 0; it < t.iterations; ++it) {
		// R = D + S X_prev, then R Q
		s_a[e] = make_double2(dr + sw.x * xr - sw.y * xi, di + sw.x * xi + sw.y * xr);
		__syncthreads();
		const double2 rq = row_times_col<N>(s_a + i * N, 1, s_q + j, N, false, false);
		__syncthreads();
		s_a[e] = rq;
		__syncthreads();
		// R~ = Q^dagger (R Q); X~ = response * R~
		const double2 tt = row_times_col<N>(s_q + i, N, s_a + j, N, true, false);
		const double yr = gr * tt.x - gi * tt.y, yi = gr * tt.y + gi * tt.x;
		__syncthreads();
		// X = Q X~ Q^dagger
		s_a[e] = make_double2(yr, yi);
		__syncthreads();
		const double2 yq = row_times_col<N>(s_a + i * N, 1, s_q + j * N, 1, false, true);
		__syncthreads();
		s_a[e] = yq;
		__syncthreads();
		const double2 xx = row_times_col<N>(s_q + i * N, 1, s_a + j, N, false, false);
		xr = xx.x;
		xi = xx.y;
		__syncthreads();
	}
	cur[static_cast<size_t>(t.n_real + u) * E + e] = V2{static_cast<T>(xr), static_cast<T>(xi)};
}

// N = 16: the same terminator on the FP64 matrix cores, one wave per virtual ADO.  Lane l owns the four elements
// (i, j) = (l / 16 + 4 v, l % 16), v = 0..3, of every 16 x 16 matrix -- the output layout of v_mfma_f64_16x16x4f64,
// whose operands are A[i][k] from lane l with i = l % 16, k = 4 kk + l / 16, and B[k][j] with j = l % 16 (kk the
// step of 4 in K).  A complex product is 16 MFMAs (real and imaginary planes, 4 steps each); between products the
// result passes through LDS once to reach the operand layout.  Q and Q^dagger stay in registers in both operand
// layouts (Q^dagger's are the conjugates of Q's in the other layout).
typedef double f64x4_t __attribute__((ext_vector_type(4)));

struct CMat16 {  // one lane's four elements of a 16 x 16 complex matrix, output layout
	f64x4_t re, im;
};

// C = A B for A, B given per lane in operand layout (a[kk], b[kk] complex, kk = 0..3)
__device__ __forceinline__ CMat16 cmma16(const double2 (&a)[4], const double2 (&b)[4])
{
	f64x4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
	for(int kk = 0; kk < 4; ++kk) {
		cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk].x, b[kk].x, cr, 0, 0, 0);
		cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[kk].y, b[kk].y, cr, 0, 0, 0);
		ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk].x, b[kk].y, ci, 0, 0, 0);
		ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk].y, b[kk].x, ci, 0, 0, 0);
	}
	return CMat16{cr, ci};
}

// writes a lane's output-layout elements to the row-major LDS matrix and reads them back in operand layout: as A
// (row l % 16, columns 4 kk + l / 16) or as B (rows 4 kk + l / 16, column l % 16)
__device__ __forceinline__ void relayout16(double2* __restrict__ s, const CMat16& m, int lane, bool as_a, double2 (&out)[4])
{
	const int r = lane >> 4, c = lane & 15;
#pragma unroll
	for(int v = 0; v < 4; ++v) s[(r + 4 * v) * 16 + c] = make_double2(m.re[v], m.im[v]);
	__syncthreads();
#pragma unroll
	for(int kk = 0; kk < 4; ++kk) out[kk] = as_a ? s[c * 16 + 4 * kk + r] : s[(4 * kk + r) * 16 + c];
	__syncthreads();
}

template <typename T>
__global__ __launch_bounds__(64) void k_terminator16(const TermArgs t)
{
	using V2 = typename C2t<T>::type;
	constexpr int N = 16, E = 256;
	__shared__ double2 s_m[E];
	const int lane = threadIdx.x;
	const int r = lane >> 4, c = lane & 15;
	const int u = blockIdx.x;
	V2* __restrict__ cur = static_cast<V2*>(t.cur);
	// Q in both operand layouts: qa[kk] = Q[c][4kk + r] (as A), qb[kk] = Q[4kk + r][c] (as B); Q^dagger as A is
	// conj(qb), as B conj(qa)
	double2 qa[4], qb[4], qha[4], qhb[4];
#pragma unroll
	for(int kk = 0; kk < 4; ++kk) {
		qa[kk] = t.q[c * N + 4 * kk + r];
		qb[kk] = t.q[(4 * kk + r) * N + c];
		qha[kk] = make_double2(qb[kk].x, -qb[kk].y);
		qhb[kk] = make_double2(qa[kk].x, -qa[kk].y);
	}
	// D_u, the response and the Xi self weight of the lane's four elements
	f64x4_t dr, di, gr, gi, swr, swi;
	const double2 rate = t.rate[u];
#pragma unroll
	for(int v = 0; v < 4; ++v) {
		const int i = r + 4 * v, j = c, e = i * N + j;
		double ar = 0.0, ai = 0.0;
		for(int q = t.src_off[u]; q < t.src_off[u + 1]; ++q) {
			const V2 x = cur[static_cast<size_t>(t.src_ado[q]) * E + e];
			const int m = t.src_mask[q];
			const double al = t.src_alpha[q];
			const double vi = (i & m) ? al : -al, vj = (j & m) ? al : -al;
			const double d = vi - vj, sm = vi + vj;
			const double2 cc = t.src_comm[q], an = t.src_anti[q];
			const double wr = cc.x * d + an.x * sm, wi = cc.y * d + an.y * sm;
			const double xr = static_cast<double>(x.x), xi = static_cast<double>(x.y);
			ar += wr * xr - wi * xi;
			ai += wr * xi + wi * xr;
		}
		dr[v] = ar;
		di[v] = ai;
		double2 corr = make_double2(1.0, 0.0);
		if(t.corr) corr = t.corr[static_cast<size_t>(u) * E + e];
		const double den_r = rate.x, den_i = rate.y - (t.ev[i] - t.ev[j]);
		const double inv = 1.0 / (den_r * den_r + den_i * den_i);
		gr[v] = -(corr.x * den_r + corr.y * den_i) * inv;
		gi[v] = -(corr.y * den_r - corr.x * den_i) * inv;
		const double2 sw = t.selfx[i ^ j];
		swr[v] = sw.x;
		swi[v] = sw.y;
	}
	CMat16 x{{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
	double2 op[4];
	for(int it = 0; it < t.iterations; ++it) {
		CMat16 m;  // R = D + S X
#pragma unroll
		for(int v = 0; v < 4; ++v) {
			m.re[v] = dr[v] + swr[v] * x.re[v] - swi[v] * x.im[v];
			m.im[v] = di[v] + swr[v] * x.im[v] + swi[v] * x.re[v];
		}
		relayout16(s_m, m, lane, true, op);
		m = cmma16(op, qb);              // R Q
		relayout16(s_m, m, lane, false, op);
		m = cmma16(qha, op);             // Q^dagger (R Q)
#pragma unroll
		for(int v = 0; v < 4; ++v) {  // X~ = response * R~
			const double yr = gr[v] * m.re[v] - gi[v] * m.im[v], yi = gr[v] * m.im[v] + gi[v] * m.re[v];
			m.re[v] = yr;
			m.im[v] = yi;
		}
		relayout16(s_m, m, lane, true, op);
		m = cmma16(op, qhb);             // X~ Q^dagger
		relayout16(s_m, m, lane, false, op);
		x = cmma16(qa, op);              // Q (X~ Q^dagger)
	}
#pragma unroll
	for(int v = 0; v < 4; ++v)
		cur[static_cast<size_t>(t.n_real + u) * E + (r + 4 * v) * N + c] = V2{static_cast<T>(x.re[v]), static_cast<T>(x.im[v])};
}

// Learned closure: feature rows for every (listed ADO a, element e = (i, j)):
//   x = [n_b(i) for the logN spins (as +-1), n_b(j) likewise, extra[a][0 .. extra_dim)]
__global__ __launch_bounds__(kBlock) void k_kappa_features(int count, int logN, const float* __restrict__ extra,
	int extra_dim, float* __restrict__ x)
{
	const int E = 1 << (2 * logN);
	const int din = 2 * logN + extra_dim;
	const size_t rows = static_cast<size_t>(count) * E;
	for(size_t r = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; r < rows; r += gridDim.x * static_cast<size_t>(kBlock)) {
		const int a = static_cast<int>(r >> (2 * logN));
		const int e = static_cast<int>(r & (E - 1));
		const int i = e >> logN, j = e & ((1 << logN) - 1);
		float* row = x + r * din;
		for(int b = 0; b < logN; ++b) {
			row[b] = ((i >> b) & 1) ? 1.f : -1.f;
			row[logN + b] = ((j >> b) & 1) ? 1.f : -1.f;
		}
		for(int q = 0; q < extra_dim; ++q) row[2 * logN + q] = extra[static_cast<size_t>(a) * extra_dim + q];
	}
}

// kappa[a][e] = y[a E + e][0] + i y[a E + e][1]
__global__ __launch_bounds__(kBlock) void k_kappa_store(size_t rows, const float* __restrict__ y, double2* __restrict__ kappa)
{
	for(size_t r = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; r < rows; r += gridDim.x * static_cast<size_t>(kBlock))
		kappa[r] = make_double2(static_cast<double>(y[2 * r]), static_cast<double>(y[2 * r + 1]));
}

// tier-0 block -> trace buffers (FP64): full block and/or real diagonal; herm: the block is stored in
// the Hermitian layout
template <typename T>
__global__ __launch_bounds__(kBlock) void k_record(const void* tier0, const double* rho0_f64, int logN, int herm,
	double* rho_out, double* diag_out)
{
	using V2 = typename C2t<T>::type;
	const uint32_t N = 1u << logN;
	const uint32_t E = N * N;
	for(uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < E; e += gridDim.x * kBlock) {
		const int i = static_cast<int>(e >> logN), j = static_cast<int>(e & (N - 1));
		bool conj = false;
		const uint32_t src = herm ? static_cast<uint32_t>(herm_index(i, j, logN, conj)) : e;
		double re, im;
		if(rho0_f64) {
			re = rho0_f64[2 * src];
			im = rho0_f64[2 * src + 1];
		} else {
			const V2 v = static_cast<const V2*>(tier0)[src];
			re = static_cast<double>(v.x);
			im = static_cast<double>(v.y);
		}
		if(conj) im = -im;
		if(rho_out) {
			rho_out[2 * e] = re;
			rho_out[2 * e + 1] = im;
		}
		if(diag_out && i == j) diag_out[i] = re;
	}
}

}  // namespace helix_amd

using namespace helix_amd;

// Copies `count` stored ADO blocks of `words` 8-byte words each, block q from base + offs[q] bytes, into out
// back to back (helix_hip_get_ados, helix_hip_batch_get_ados).
__global__ __launch_bounds__(kBlock) void k_gather_blocks(const char* __restrict__ base, const int64_t* __restrict__ offs,
	int count, int words, uint2* __restrict__ out)
{
	const size_t total = static_cast<size_t>(count) * words;
	for(size_t t = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; t < total; t += gridDim.x * static_cast<size_t>(kBlock)) {
		const size_t q = t / words, w = t - q * words;
		out[t] = reinterpret_cast<const uint2*>(base + offs[q])[w];
	}
}

// Device scratch of helix_hip_step_noalloc (include/helix_hip.h): the staging buffer of its trace records.  Each GPU
// the library runs on holds its own copy of this array until the process ends; host code never passes the
// array's name to a kernel or a copy (on the host it only stands in for the symbol), it uses the device address
// that resolve_scratch() obtained with hipGetSymbolAddress on that device.
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

struct helix_hip_plan {
	int device = 0;
	hipStream_t stream = nullptr;
	int N = 0, logN = 0, H = 0, M = 0, B = 0, order = 4, precision = 0;
	double dt = 0.0;
	bool flip = false;        // some coupling has bit flips
	bool csr_h = false;       // general sparse H (CSR)
	bool csr_v = false;       // general sparse couplings (CSR)
	bool tiled = false;       // LDS-tiled stage kernel (diagonal couplings, N <= 64)
	bool dense = false;       // dense Hamiltonian on MFMA (N = 16 .. 1024)
	bool herm = false;        // Hermitian layout: N(N+1)/2 stored elements per ADO
	int threads = 0;          // tile workgroup size (512 or 1024; N <= 64 sigma^z tiles are autotuned)
	size_t esize = 16;        // bytes per complex element
	size_t pe = 0;            // stored elements per ADO (N^2, or N(N+1)/2 in the Hermitian layout)
	std::vector<int32_t> perm; // stored index of each reference ADO (colex_order)
	size_t elems = 0;         // H * pe
	void* rho = nullptr;
	void* t[kMaxOrder + 1] = {};
	double* rho0_f64 = nullptr;
	void* buf = nullptr;
	std::vector<void*> allocations;
	size_t device_bytes = 0;
	StageArgs base{};
	std::vector<hipEvent_t> events;
	// kChunkSteps steps captured once in a hipGraph and replayed by helix_hip_step (order >= 2); rebuilt
	// after any change of the stage arguments
	hipGraph_t chunk_graph = nullptr;
	hipGraphExec_t chunk_exec = nullptr;
	// learned closure of the deepest kept tier (helix_hip_plan_set_closure)
	double2* kappa = nullptr;
	int32_t* kslot = nullptr;
	int kappa_count = 0;
	// trace staging of helix_hip_step, allocated on the first recording call and reused while it suffices
	double* trace = nullptr;
	size_t trace_doubles = 0;
	// this device's copy of `scratch` (helix_hip_step_noalloc)
	double* scratch = nullptr;
	// physics terminator (helix_hip_plan_set_terminator): U virtual ADOs after the H real ones in every state vector
	// (capacity virt_cap), the tables extended to H + U rows while it is on, the original tables kept
	int term_count = 0;
	int virt_cap = 0;
	TermArgs term{};
	const int32_t* orig_up = nullptr;
	const int32_t* orig_down = nullptr;
	const double2* orig_dcomm = nullptr;
	const double2* orig_danti = nullptr;
	const double2* orig_selfc = nullptr;
	std::vector<void*> term_allocs;  // the extended tables and terminator arguments of the current terminator
	void* term_corr = nullptr;       // its learned correction table (one of term_allocs), or null
	std::vector<int32_t> kslot_host;  // closure slot of every stored ADO (-1: none), H + virt_cap entries on the device
	std::vector<int32_t> mode_bath_host;
};

namespace helix_amd_mlp {
float* input_rows(helix_hip_mlp* m, int n);
const float* forward_on(helix_hip_mlp* m, int n, hipStream_t stream);
int dims(const helix_hip_mlp* m, int* in, int* out);
}  // namespace helix_amd_mlp

namespace {

int plan_alloc(helix_hip_plan* p, void** ptr, size_t bytes)
{
	*ptr = nullptr;
	if(bytes == 0) return 0;
	hipError_t e = hipMalloc(ptr, bytes);
	if(e != hipSuccess) return hip_fail(e, "hipMalloc");
	p->allocations.push_back(*ptr);
	p->device_bytes += bytes;
	return 0;
}

template <typename X>
int upload(helix_hip_plan* p, const X* host, size_t count, const X** dev)
{
	void* d = nullptr;
	int rc = plan_alloc(p, &d, count * sizeof(X));
	if(rc) return rc;
	if(count) HX_HIP(hipMemcpy(d, host, count * sizeof(X), hipMemcpyHostToDevice));
	*dev = static_cast<const X*>(d);
	return 0;
}

// count CSR operators (each [n+1] offsets) concatenated: offsets made absolute into one column / value array
int upload_csr(helix_hip_plan* p, const helix_hip_csr_operator* ops, int count, int n, const int32_t** off,
	const int32_t** col, const double2** val)
{
	std::vector<int32_t> o, c;
	std::vector<double> v;
	for(int q = 0; q < count; ++q) {
		const int32_t base = static_cast<int32_t>(c.size());
		for(int i = 0; i <= n; ++i) o.push_back(base + ops[q].row_offsets[i]);
		c.insert(c.end(), ops[q].columns, ops[q].columns + ops[q].nnz);
		v.insert(v.end(), ops[q].values, ops[q].values + 2 * static_cast<size_t>(ops[q].nnz));
	}
	if(c.empty()) {  // keep valid pointers for an all-zero operator
		c.push_back(0);
		v.push_back(0.0);
		v.push_back(0.0);
	}
	int rc = upload(p, o.data(), o.size(), off);
	if(!rc) rc = upload(p, c.data(), c.size(), col);
	if(!rc) rc = upload(p, reinterpret_cast<const double2*>(v.data()), v.size() / 2, val);
	return rc;
}

// sigma^z couplings: per mode (the first 16), 2 s for the spin s of its bath, 2 alpha of the bath, and whether
// any ADO has an anticommutator (down_anti) weight for it
void set_mode_spins(StageArgs& a, const helix_hip_problem* pr)
{
	a.sz_anti = 0;
	for(int k = 0; k < pr->n_modes && k < 16; ++k) {
		const int b = pr->mode_bath[k];
		a.sz_shift[k] = 2 * __builtin_ctz(static_cast<unsigned>(a.sz_mask[b]));
		a.sz_ta[k] = 2.0 * a.sz_alpha[b];
		for(int n = 0; n < pr->n_ado; ++n) {
			const double* w = pr->down_anti + 2 * (static_cast<size_t>(n) * pr->n_modes + k);
			if(w[0] != 0.0 || w[1] != 0.0) {
				a.sz_anti |= 1 << k;
				break;
			}
		}
	}
}

int ilog2_exact(int v)
{
	int l = 0;
	while((1 << l) < v) ++l;
	return ((1 << l) == v) ? l : -1;
}

// A CSR operator: well-formed offsets and columns, and Hermitian (O_ji = conj O_ij to 1e-12 of max|O|,
// duplicates summed) -- the X O product of the commutators walks rows of O.
int check_csr(const helix_hip_csr_operator& op, int n, const std::string& name)
{
	if(!op.row_offsets || (op.nnz > 0 && (!op.columns || !op.values)) || op.nnz < 0)
		return fail(HELIX_HIP_EINVAL, name + ": null arrays");
	if(op.row_offsets[0] != 0 || op.row_offsets[n] != op.nnz) return fail(HELIX_HIP_EINVAL, name + ": bad row offsets");
	std::map<std::pair<int, int>, std::pair<double, double>> entries;
	double mx = 0.0;
	for(int i = 0; i < n; ++i) {
		if(op.row_offsets[i + 1] < op.row_offsets[i]) return fail(HELIX_HIP_EINVAL, name + ": decreasing row offsets");
		for(int z = op.row_offsets[i]; z < op.row_offsets[i + 1]; ++z) {
			const int j = op.columns[z];
			if(j < 0 || j >= n) return fail(HELIX_HIP_EINVAL, name + ": column out of range");
			auto& v = entries[{i, j}];
			v.first += op.values[2 * z];
			v.second += op.values[2 * z + 1];
		}
	}
	for(const auto& e : entries) mx = std::max(mx, std::hypot(e.second.first, e.second.second));
	for(const auto& e : entries) {
		const auto it = entries.find({e.first.second, e.first.first});
		const double tr = (it == entries.end()) ? 0.0 : it->second.first;
		const double ti = (it == entries.end()) ? 0.0 : -it->second.second;
		if(std::fabs(tr - e.second.first) > 1e-12 * mx || std::fabs(ti - e.second.second) > 1e-12 * mx)
			return fail(HELIX_HIP_EINVAL, name + ": operator is not Hermitian");
	}
	return 0;
}

int validate(const helix_hip_problem* pr)
{
	if(!pr) return fail(HELIX_HIP_EINVAL, "problem is null");
	if(pr->abi_version != HELIX_HIP_ABI_VERSION) return fail(HELIX_HIP_EINVAL, "abi_version mismatch");
	const int logN = ilog2_exact(pr->n);
	if(logN < 3 || logN > 12) return fail(HELIX_HIP_EINVAL, "n must be a power of two in [8, 4096]");
	if(pr->n_ado <= 0 || pr->n_modes <= 0 || pr->n_baths <= 0 || pr->n_modes > 64 || pr->n_baths > 64)
		return fail(HELIX_HIP_EINVAL, "hierarchy sizes out of range");
	if(!pr->up || !pr->down || !pr->mode_bath || !pr->tier || !pr->down_comm || !pr->down_anti || !pr->self_coef ||
		!pr->xi || !pr->xi_counter || !pr->couplings || !pr->hamiltonian.diag)
		return fail(HELIX_HIP_EINVAL, "problem table pointer is null");
	if(pr->order < 1 || pr->order > kMaxOrder) return fail(HELIX_HIP_EINVAL, "order must be in [1, 8]");
	if(!(pr->dt > 0.0)) return fail(HELIX_HIP_EINVAL, "dt must be positive");
	if(pr->precision < HELIX_HIP_FP64 || pr->precision > HELIX_HIP_SPLIT_FP32)
		return fail(HELIX_HIP_EINVAL, "unknown precision");
	if(pr->layout < HELIX_HIP_LAYOUT_AUTO || pr->layout > HELIX_HIP_LAYOUT_HERMITIAN)
		return fail(HELIX_HIP_EINVAL, "unknown layout");
	if(pr->hamiltonian_dense) {
		if(pr->couplings_csr) return fail(HELIX_HIP_EUNSUPPORTED, "dense Hamiltonian path needs diagonal couplings");
		if(pr->n < 16 || pr->n > 1024)
			return fail(HELIX_HIP_EUNSUPPORTED, "dense Hamiltonian path supports N = 16 .. 1024");
		for(int b = 0; b < pr->n_baths; ++b)
			if(pr->couplings[b].nflips != 0)
				return fail(HELIX_HIP_EUNSUPPORTED, "dense Hamiltonian path needs diagonal couplings");
		if(pr->n_baths > 16 || pr->n_modes > 16)
			return fail(HELIX_HIP_EUNSUPPORTED, "dense Hamiltonian path supports up to 16 baths and modes");
	}
	auto check_op = [&](const helix_hip_spin_operator& op, const char* name) -> int {
		if(op.nflips < 0 || op.nflips > kMaxFlips) return fail(HELIX_HIP_EINVAL, std::string(name) + ": nflips > 16");
		for(int f = 0; f < op.nflips; ++f)
			if(op.flip_masks[f] <= 0 || op.flip_masks[f] >= pr->n)
				return fail(HELIX_HIP_EINVAL, std::string(name) + ": flip mask outside (0, n)");
		return 0;
	};
	if(int rc = check_op(pr->hamiltonian, "hamiltonian")) return rc;
	if(pr->hamiltonian_csr) {
		if(pr->hamiltonian_dense) return fail(HELIX_HIP_EINVAL, "give either a dense or a CSR Hamiltonian");
		if(int rc = check_csr(pr->hamiltonian_csr[0], pr->n, "hamiltonian_csr")) return rc;
	}
	if(pr->couplings_csr)
		for(int b = 0; b < pr->n_baths; ++b)
			if(int rc = check_csr(pr->couplings_csr[b], pr->n, "couplings_csr[" + std::to_string(b) + "]")) return rc;
	for(int b = 0; b < pr->n_baths; ++b) {
		if(!pr->couplings[b].diag) return fail(HELIX_HIP_EINVAL, "coupling diag is null");
		if(int rc = check_op(pr->couplings[b], "coupling")) return rc;
	}
	// element counters are 32-bit: over the state, and over the [B][H][N][N] buffer of the commutator pass that
	// bit-flip and general sparse couplings use
	bool coupling_pass = pr->couplings_csr != nullptr;
	for(int b = 0; b < pr->n_baths; ++b) coupling_pass = coupling_pass || pr->couplings[b].nflips > 0;
	const unsigned long long elems = static_cast<unsigned long long>(pr->n_ado) << (2 * logN);
	if(elems * (coupling_pass ? static_cast<unsigned long long>(pr->n_baths) : 1ull) >= (1ull << 31))
		return fail(HELIX_HIP_EINVAL, "state too large for 32-bit element indexing");
	for(long q = 0; q < static_cast<long>(pr->n_ado) * pr->n_modes; ++q) {
		if(pr->up[q] < 0 || pr->up[q] > pr->n_ado || pr->down[q] < 0 || pr->down[q] > pr->n_ado)
			return fail(HELIX_HIP_EINVAL, "up/down index outside [0, n_ado]");
	}
	for(int k = 0; k < pr->n_modes; ++k)
		if(pr->mode_bath[k] < 0 || pr->mode_bath[k] >= pr->n_baths) return fail(HELIX_HIP_EINVAL, "mode_bath out of range");
	if(pr->tier[0] != 0) return fail(HELIX_HIP_EINVAL, "ADO 0 must be the tier-0 ADO");
	return 0;
}

unsigned grid_for(size_t total)
{
	size_t g = (total + kBlock - 1) / kBlock;
	if(g > (1u << 30)) g = 1u << 30;
	return static_cast<unsigned>(g ? g : 1);
}

template <typename T, bool SPLIT, int LOGN, int THREADS, bool SZ, bool HERM>
void launch_tile_v(helix_hip_plan* p, const StageArgs& a, bool last)
{
	constexpr int E = 1 << (2 * LOGN);
	constexpr int TPA = (E > kTile) ? E / kTile : 1;
	constexpr int A = (E >= kTile) ? 1 : kTile / E;
	const int ngroups = ((p->H + A - 1) / A) * TPA;
	const unsigned grid = static_cast<unsigned>(((ngroups + kXcds - 1) / kXcds) * kXcds);
	if(last)
		hipLaunchKernelGGL((k_stage_tile<T, SPLIT, true, LOGN, THREADS, SZ, HERM>), dim3(grid), dim3(THREADS), 0,
			p->stream, a, ngroups);
	else
		hipLaunchKernelGGL((k_stage_tile<T, SPLIT, false, LOGN, THREADS, SZ, HERM>), dim3(grid), dim3(THREADS), 0,
			p->stream, a, ngroups);
}

// 512 threads for complex128 tiles (64 KiB LDS, 4 waves/SIMD), 1024 for complex64 sigma^z tiles (32 KiB,
// 8 waves/SIMD; general diagonal couplings need more registers and stay at 512); whole-block sigma^z tiles
// (N <= 64) also have the other shape, which the plan's autotuner may pick
template <typename T, bool SPLIT, int LOGN, bool SZ>
void launch_tile_sz(helix_hip_plan* p, const StageArgs& a, bool last)
{
	constexpr int THREADS = (sizeof(T) == 8 || !SZ) ? 512 : 1024;
	constexpr int ALT = (sizeof(T) == 8) ? 1024 : 512;
	if constexpr(LOGN <= 6 && SZ) {
		if(p->threads == ALT) {
			if(p->herm)
				launch_tile_v<T, SPLIT, LOGN, ALT, SZ, true>(p, a, last);
			else
				launch_tile_v<T, SPLIT, LOGN, ALT, SZ, false>(p, a, last);
			return;
		}
	}
	if constexpr(LOGN <= 6 || SZ) {
		if(p->herm) {
			launch_tile_v<T, SPLIT, LOGN, THREADS, SZ, true>(p, a, last);
			return;
		}
	}
	launch_tile_v<T, SPLIT, LOGN, THREADS, SZ, false>(p, a, last);
}

template <typename T, bool SPLIT, int LOGN>
void launch_tile(helix_hip_plan* p, const StageArgs& a, bool last)
{
	if(a.sz) {
		launch_tile_sz<T, SPLIT, LOGN, true>(p, a, last);
	} else {
		// general diagonal couplings keep their B x N diagonals in LDS: N <= 64 only
		if constexpr(LOGN <= 6) launch_tile_sz<T, SPLIT, LOGN, false>(p, a, last);
	}
}

template <typename T, bool SPLIT, int LOGN, bool SZ>
void launch_dense_v(helix_hip_plan* p, const StageArgs& a, bool last)
{
	if constexpr(LOGN <= 6) {
		const unsigned grid = static_cast<unsigned>(((p->H + kXcds - 1) / kXcds) * kXcds);
		if(last)
			hipLaunchKernelGGL((k_stage_dense<T, SPLIT, true, LOGN, SZ>), dim3(grid), dim3(dense_threads<LOGN>()), 0, p->stream, a);
		else
			hipLaunchKernelGGL((k_stage_dense<T, SPLIT, false, LOGN, SZ>), dim3(grid), dim3(dense_threads<LOGN>()), 0, p->stream, a);
	} else {
		constexpr int NTT = (1 << LOGN) / kDT;
		const int ngroups = p->H * NTT * NTT;
		const unsigned grid = static_cast<unsigned>(((ngroups + kXcds - 1) / kXcds) * kXcds);
		if(last)
			hipLaunchKernelGGL((k_stage_dense_tiled<T, SPLIT, true, LOGN, SZ>), dim3(grid), dim3(kDenseThreads), 0, p->stream,
				a, ngroups);
		else
			hipLaunchKernelGGL((k_stage_dense_tiled<T, SPLIT, false, LOGN, SZ>), dim3(grid), dim3(kDenseThreads), 0,
				p->stream, a, ngroups);
	}
}

template <typename T, bool SPLIT, int LOGN>
void launch_dense(helix_hip_plan* p, const StageArgs& a, bool last)
{
	if(a.sz)
		launch_dense_v<T, SPLIT, LOGN, true>(p, a, last);
	else
		launch_dense_v<T, SPLIT, LOGN, false>(p, a, last);
}

// untiled stage: pass A ([V_b, cur] for non-diagonal couplings) and the fused RHS + epilogue
template <typename T, bool SPLIT, int VOP, bool HCSR>
void launch_generic(helix_hip_plan* p, const StageArgs& a, bool last)
{
	if constexpr(VOP != kVDiag)
		hipLaunchKernelGGL((k_flip_commutator<T, SPLIT, VOP>), dim3(grid_for(p->elems * p->B)), dim3(kBlock), 0, p->stream,
			a);
	const unsigned grid = grid_for(p->elems);
	if(last)
		hipLaunchKernelGGL((k_stage<T, SPLIT, true, VOP, HCSR>), dim3(grid), dim3(kBlock), 0, p->stream, a);
	else
		hipLaunchKernelGGL((k_stage<T, SPLIT, false, VOP, HCSR>), dim3(grid), dim3(kBlock), 0, p->stream, a);
}

template <typename T, bool SPLIT, int VOP>
void launch_generic_h(helix_hip_plan* p, const StageArgs& a, bool last)
{
	if(p->csr_h)
		launch_generic<T, SPLIT, VOP, true>(p, a, last);
	else
		launch_generic<T, SPLIT, VOP, false>(p, a, last);
}

template <typename T, bool SPLIT>
void launch_stage(helix_hip_plan* p, const StageArgs& a, bool last)
{
	const unsigned grid = grid_for(p->elems);
	if(p->dense) {
		switch(p->logN) {
		case 4: launch_dense<T, SPLIT, 4>(p, a, last); return;
		case 5: launch_dense<T, SPLIT, 5>(p, a, last); return;
		case 6: launch_dense<T, SPLIT, 6>(p, a, last); return;
		case 7: launch_dense<T, SPLIT, 7>(p, a, last); return;
		case 8: launch_dense<T, SPLIT, 8>(p, a, last); return;
		case 9: launch_dense<T, SPLIT, 9>(p, a, last); return;
		default: launch_dense<T, SPLIT, 10>(p, a, last); return;
		}
	}
	if(p->tiled) {
		switch(p->logN) {
		case 3: launch_tile<T, SPLIT, 3>(p, a, last); return;
		case 4: launch_tile<T, SPLIT, 4>(p, a, last); return;
		case 5: launch_tile<T, SPLIT, 5>(p, a, last); return;
		case 6: launch_tile<T, SPLIT, 6>(p, a, last); return;
		case 7: launch_tile<T, SPLIT, 7>(p, a, last); return;
		case 8: launch_tile<T, SPLIT, 8>(p, a, last); return;
		case 9: launch_tile<T, SPLIT, 9>(p, a, last); return;
		default: launch_tile<T, SPLIT, 10>(p, a, last); return;
		}
	}
	(void)grid;
	const int vop = p->csr_v ? kVCsr : (p->flip ? kVFlip : kVDiag);
	switch(vop) {
	case kVCsr: launch_generic_h<T, SPLIT, kVCsr>(p, a, last); break;
	case kVFlip: launch_generic_h<T, SPLIT, kVFlip>(p, a, last); break;
	default: launch_generic_h<T, SPLIT, kVDiag>(p, a, last); break;
	}
}

template <typename T>
void launch_terminator(helix_hip_plan* p, const void* cur)
{
	TermArgs t = p->term;
	t.cur = const_cast<void*>(cur);
	const dim3 grid(static_cast<unsigned>(p->term_count));
	switch(p->logN) {
	case 3: hipLaunchKernelGGL((k_terminator<T, 3>), grid, dim3(64), 0, p->stream, t); break;
	case 4: hipLaunchKernelGGL((k_terminator16<T>), grid, dim3(64), 0, p->stream, t); break;
	default: hipLaunchKernelGGL((k_terminator<T, 5>), grid, dim3(1024), 0, p->stream, t); break;
	}
}

void dispatch_stage(helix_hip_plan* p, const StageArgs& a, bool last)
{
	if(p->term_count > 0) {
		if(p->precision == HELIX_HIP_FP64)
			launch_terminator<double>(p, a.cur);
		else
			launch_terminator<float>(p, a.cur);
	}
	switch(p->precision) {
	case HELIX_HIP_FP64: launch_stage<double, false>(p, a, last); break;
	case HELIX_HIP_FP32: launch_stage<float, false>(p, a, last); break;
	default: launch_stage<float, true>(p, a, last); break;
	}
}

// Enqueues stage j (1-based) of the current step.
void enqueue_stage(helix_hip_plan* p, int j)
{
	StageArgs a = p->base;
	a.scale = p->dt / j;
	a.cur = (j == 1) ? p->rho : p->t[j - 1];
	const bool last = (j == p->order);
	if(!last) {
		a.out = p->t[j];
	} else {
		a.rho = p->rho;
		a.ntprev = p->order - 1;
		for(int q = 0; q < a.ntprev; ++q) a.tprev[q] = p->t[q + 1];
		a.out = (p->order == 1) ? p->t[1] : p->rho;
		a.rho0_f64 = p->rho0_f64;
	}
	dispatch_stage(p, a, last);
	if(last && p->order == 1) std::swap(p->rho, p->t[1]);
}

void enqueue_step(helix_hip_plan* p)
{
	for(int j = 1; j <= p->order; ++j) enqueue_stage(p, j);
}

// Steps replayed from one hipGraph: small problems are bound by launch latency (a 210-ADO N=8 stage
// runs in a few microseconds), which the replay of a captured chunk removes.
constexpr uint64_t kChunkSteps = 16;

void drop_chunk_graph(helix_hip_plan* p)
{
	if(p->chunk_exec) (void)hipGraphExecDestroy(p->chunk_exec);
	if(p->chunk_graph) (void)hipGraphDestroy(p->chunk_graph);
	p->chunk_exec = nullptr;
	p->chunk_graph = nullptr;
}

// Captures and instantiates the chunk graph unless it exists (at plan creation, and again after a change of the
// stage arguments), so that no step pays for it.
int capture_chunk(helix_hip_plan* p)
{
	if(p->chunk_exec) return 0;
	HX_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
	for(uint64_t s = 0; s < kChunkSteps; ++s) enqueue_step(p);
	hipError_t e = hipStreamEndCapture(p->stream, &p->chunk_graph);
	if(e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
	e = hipGraphInstantiate(&p->chunk_exec, p->chunk_graph, nullptr, nullptr, 0);
	if(e != hipSuccess) {
		drop_chunk_graph(p);
		return hip_fail(e, "hipGraphInstantiate");
	}
	// upload the executable graph now, so that its first replay (possibly inside a caller's timed region) does not
	// pay for it
	e = hipGraphUpload(p->chunk_exec, p->stream);
	if(e == hipSuccess) e = hipStreamSynchronize(p->stream);
	if(e != hipSuccess) {
		drop_chunk_graph(p);
		return hip_fail(e, "hipGraphUpload");
	}
	return 0;
}

int enqueue_chunk(helix_hip_plan* p)
{
	if(int rc = capture_chunk(p)) return rc;
	HX_HIP(hipGraphLaunch(p->chunk_exec, p->stream));
	return 0;
}

// Event time of `steps` eager steps on the plan's stream.
int time_eager_steps(helix_hip_plan* p, int steps, double* ms)
{
	hipEvent_t e0, e1;
	HX_HIP(hipEventCreate(&e0));
	if(hipError_t e = hipEventCreate(&e1); e != hipSuccess) {
		(void)hipEventDestroy(e0);
		return hip_fail(e, "hipEventCreate");
	}
	hipError_t e = hipEventRecord(e0, p->stream);
	for(int s = 0; s < steps && e == hipSuccess; ++s) enqueue_step(p);
	if(e == hipSuccess) e = hipGetLastError();
	if(e == hipSuccess) e = hipEventRecord(e1, p->stream);
	if(e == hipSuccess) e = hipEventSynchronize(e1);
	float f = 0.f;
	if(e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if(e != hipSuccess) return hip_fail(e, "autotune timing");
	*ms = f;
	return 0;
}

// Whole-block sigma^z tiles (N <= 64) come in two workgroup shapes: time both
// on the all-zero state (stepping zero leaves it zero, and the kernels move the same bytes whatever the values)
// and keep the faster.  HELIX_HIP_TILE_THREADS=512 or 1024 forces one shape instead (diagnostics and tests).
int autotune_threads(helix_hip_plan* p)
{
	const int def = (p->precision == HELIX_HIP_FP64) ? 512 : 1024;
	p->threads = def;
	if(!p->tiled || !p->base.sz || p->logN > 6) return 0;
	if(const char* force = std::getenv("HELIX_HIP_TILE_THREADS")) {
		const int th = std::atoi(force);
		if(th == 512 || th == 1024) {
			p->threads = th;
			return 0;
		}
	}
	double best = 1e300;
	int pick = def;
	for(int th : {def, def == 512 ? 1024 : 512}) {
		p->threads = th;
		double ms = 0.0;
		if(int rc = time_eager_steps(p, 1, &ms)) return rc;  // first launch of this shape
		if(int rc = time_eager_steps(p, 3, &ms)) return rc;
		if(ms < best) {
			best = ms;
			pick = th;
		}
	}
	p->threads = pick;
	return 0;
}

int launch_error(const char* what)
{
	hipError_t e = hipGetLastError();
	if(e != hipSuccess) return hip_fail(e, what);
	return 0;
}

// Device address of `scratch` on `device`, resolved once per device with that device current (each GPU holds its
// own copy of the array at its own address).
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};

int resolve_scratch(int device, double** out)
{
	*out = nullptr;
	if(device < 0 || device >= kMaxScratchDevices) return fail(HELIX_HIP_EINVAL, "device index out of range");
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if(!g_scratch_ptr[device]) {
		int cur = -1;
		HX_HIP(hipGetDevice(&cur));
		if(cur != device) return fail(HELIX_HIP_ESTATE, "resolve_scratch: the plan's device is not current");
		void* ptr = nullptr;
		HX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
		g_scratch_ptr[device] = static_cast<double*>(ptr);
	}
	*out = g_scratch_ptr[device];
	return 0;
}

// Trace records of one helix_hip_step call: k_record writes each record into a device staging buffer of `cap`
// records (cap full blocks [N][N][2], then cap diagonals [N]), which is copied to the caller's arrays whenever it
// fills and at the end, so a call of any length needs only a bounded buffer.
struct TraceStaging {
	uint64_t nrec = 0;      // records of the call
	uint64_t cap = 0;       // records the staging buffer holds
	double* dev = nullptr;  // staging buffer
	bool rho = false;
	bool diag = false;
};

constexpr size_t kTraceStagingBytes = size_t(64) << 20;  // plan-owned staging of helix_hip_step

size_t record_doubles(const helix_hip_plan* p, bool rho, bool diag)
{
	const size_t n2 = static_cast<size_t>(p->N) * p->N;
	return (rho ? 2 * n2 : 0) + (diag ? static_cast<size_t>(p->N) : 0);
}

int trace_staging_on(const helix_hip_plan* p, uint64_t nsteps, uint64_t record_every, const double* rho_trace,
	const double* diag_trace, double* dev, size_t dev_doubles, TraceStaging& st)
{
	st = TraceStaging{};
	if(record_every == 0 || (!rho_trace && !diag_trace)) return 0;
	st.nrec = nsteps / record_every;
	if(st.nrec == 0) return 0;
	st.rho = rho_trace != nullptr;
	st.diag = diag_trace != nullptr;
	const size_t per = record_doubles(p, st.rho, st.diag);
	st.cap = std::min<uint64_t>(st.nrec, dev_doubles / per);
	if(st.cap == 0)
		return fail(HELIX_HIP_EINVAL, "one trace record of N = " + std::to_string(p->N) + " needs " +
			std::to_string(per * sizeof(double)) + " bytes, more than the " + std::to_string(dev_doubles * sizeof(double)) +
			"-byte staging buffer");
	st.dev = dev;
	return 0;
}

int plan_trace_staging(helix_hip_plan* p, uint64_t nsteps, uint64_t record_every, const double* rho_trace,
	const double* diag_trace, TraceStaging& st)
{
	st = TraceStaging{};
	if(record_every == 0 || (!rho_trace && !diag_trace) || nsteps / record_every == 0) return 0;
	const size_t per = record_doubles(p, rho_trace != nullptr, diag_trace != nullptr);
	const uint64_t want = std::min<uint64_t>(nsteps / record_every, std::max<size_t>(1, kTraceStagingBytes / (per * 8)));
	if(p->trace_doubles < want * per) {
		if(p->trace) {
			HX_HIP(hipStreamSynchronize(p->stream));
			(void)hipFree(p->trace);
			p->device_bytes -= p->trace_doubles * sizeof(double);
			p->trace = nullptr;
			p->trace_doubles = 0;
		}
		void* d = nullptr;
		if(hipMalloc(&d, want * per * sizeof(double)) != hipSuccess) return fail(HELIX_HIP_ENOMEM, "trace staging buffer");
		p->trace = static_cast<double*>(d);
		p->trace_doubles = want * per;
		p->device_bytes += p->trace_doubles * sizeof(double);
	}
	return trace_staging_on(p, nsteps, record_every, rho_trace, diag_trace, p->trace, p->trace_doubles, st);
}

int flush_records(helix_hip_plan* p, const TraceStaging& st, uint64_t first, uint64_t count, double* rho_trace,
	double* diag_trace)
{
	if(count == 0) return 0;
	if(hipError_t e = hipStreamSynchronize(p->stream); e != hipSuccess) return hip_fail(e, "step");
	const size_t n2 = static_cast<size_t>(p->N) * p->N;
	if(st.rho) HX_HIP(hipMemcpy(rho_trace + first * n2 * 2, st.dev, count * n2 * 16, hipMemcpyDeviceToHost));
	if(st.diag)
		HX_HIP(hipMemcpy(diag_trace + first * p->N, st.dev + (st.rho ? st.cap * n2 * 2 : 0), count * p->N * sizeof(double),
			hipMemcpyDeviceToHost));
	return 0;
}

// nsteps steps with the records of `st`; graphs: replay 16-step chunks from the plan's hipGraph where they fit
// (capturing it on first use), else launch every stage eagerly.  Returns after the device finished.
int step_impl(helix_hip_plan* p, uint64_t nsteps, uint64_t record_every, double* rho_trace, double* diag_trace,
	const TraceStaging& st, bool graphs)
{
	const size_t n2 = static_cast<size_t>(p->N) * p->N;
	uint64_t r = 0, s = 0, flushed = 0;
	while(s < nsteps) {
		// steps up to the next record point (or the end), in graph chunks where they fit
		const uint64_t next = (r < st.nrec) ? (r + 1) * record_every : nsteps;
		while(s < next) {
			if(graphs && p->order >= 2 && next - s >= kChunkSteps) {
				if(int rc = enqueue_chunk(p)) return rc;
				s += kChunkSteps;
			} else {
				enqueue_step(p);
				++s;
			}
		}
		if(r < st.nrec && s == (r + 1) * record_every) {
			const uint64_t slot = r - flushed;
			double* ro = st.rho ? st.dev + slot * n2 * 2 : nullptr;
			double* dg = st.diag ? st.dev + (st.rho ? st.cap * n2 * 2 : 0) + slot * p->N : nullptr;
			const unsigned g = grid_for(n2);
			const int hm = p->herm ? 1 : 0;
			if(p->precision == HELIX_HIP_FP64)
				hipLaunchKernelGGL((k_record<double>), dim3(g), dim3(kBlock), 0, p->stream, p->rho, nullptr, p->logN, hm, ro, dg);
			else
				hipLaunchKernelGGL((k_record<float>), dim3(g), dim3(kBlock), 0, p->stream, p->rho, p->rho0_f64, p->logN, hm,
					ro, dg);
			++r;
			if(r - flushed == st.cap) {
				if(int rc = launch_error("record launch")) return rc;
				if(int rc = flush_records(p, st, flushed, st.cap, rho_trace, diag_trace)) return rc;
				flushed = r;
			}
		}
		if(int rc = launch_error("step launch")) return rc;
	}
	if(int rc = launch_error("step launch")) return rc;
	if(hipError_t e = hipStreamSynchronize(p->stream); e != hipSuccess) return hip_fail(e, "step");
	return flush_records(p, st, flushed, r - flushed, rho_trace, diag_trace);
}

// One N x N complex block (interleaved doubles) into the Hermitian layout; false when some
// |rho[j][i] - conj(rho[i][j])| exceeds tol.
bool pack_block(const double* full, int logN, double* packed, double tol)
{
	const int N = 1 << logN;
	for(int i = 0; i < N; ++i)
		for(int j = i + 1; j < N; ++j) {
			const double* a = full + 2 * (static_cast<size_t>(i) * N + j);
			const double* b = full + 2 * (static_cast<size_t>(j) * N + i);
			if(std::fabs(a[0] - b[0]) > tol || std::fabs(a[1] + b[1]) > tol) return false;
		}
	for(int i = 0; i < N; ++i)
		for(int j = 0; j < N; ++j) {
			bool conj = false;
			const int q = herm_index(i, j, logN, conj);
			if(conj) continue;
			packed[2 * q] = full[2 * (static_cast<size_t>(i) * N + j)];
			packed[2 * q + 1] = full[2 * (static_cast<size_t>(i) * N + j) + 1];
		}
	return true;
}

void unpack_block(const double* packed, int logN, double* full)
{
	const int N = 1 << logN;
	for(int i = 0; i < N; ++i)
		for(int j = 0; j < N; ++j) {
			bool conj = false;
			const int q = herm_index(i, j, logN, conj);
			full[2 * (static_cast<size_t>(i) * N + j)] = packed[2 * q];
			full[2 * (static_cast<size_t>(i) * N + j) + 1] = conj ? -packed[2 * q + 1] : packed[2 * q + 1];
		}
}

// Device storage order of the ADOs: colexicographic (the last mode most significant) instead of the
// reference's lexicographic order.  Each XCD sweeps one contiguous range of stored ADOs and the neighbour
// blocks n +- e_k it reads stay in its L2 only while they are close in that order; an LRU model of the
// 4 MB per-XCD L2 over the 6188-ADO, 12-mode hierarchy counts 10% fewer block fetches in colexicographic
// order (lexicographic 5.6, colexicographic 5.0 per ADO and stage).  perm[h] = stored index of reference
// ADO h; ADO 0 (tier 0) stays first.  The multi-indices are rebuilt from the up table.
std::vector<int32_t> colex_order(const int32_t* up, int H, int M)
{
	std::vector<int32_t> rows(static_cast<size_t>(H) * M, 0);
	std::vector<char> seen(H, 0);
	std::vector<int32_t> queue{0};
	seen[0] = 1;
	for(size_t q = 0; q < queue.size(); ++q) {
		const int n = queue[q];
		for(int k = 0; k < M; ++k) {
			const int u = up[static_cast<size_t>(n) * M + k];
			if(u >= H || seen[u]) continue;
			seen[u] = 1;
			std::copy(&rows[static_cast<size_t>(n) * M], &rows[static_cast<size_t>(n) * M] + M, &rows[static_cast<size_t>(u) * M]);
			++rows[static_cast<size_t>(u) * M + k];
			queue.push_back(u);
		}
	}
	std::vector<int32_t> order(H);
	for(int h = 0; h < H; ++h) order[h] = h;
	if(static_cast<int>(queue.size()) != H) return order;  // not a full hierarchy: keep the given order
	std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
		for(int k = M - 1; k >= 0; --k) {
			const int32_t a = rows[static_cast<size_t>(x) * M + k], b = rows[static_cast<size_t>(y) * M + k];
			if(a != b) return a < b;
		}
		return false;
	});
	std::vector<int32_t> perm(H);
	for(int r = 0; r < H; ++r) perm[order[r]] = r;
	return perm;
}

// The per-ADO tables of a problem in stored order.
struct StoredTables {
	std::vector<int32_t> up, down, tier;
	std::vector<double> dcomm, danti, selfc;
};

StoredTables permute_tables(const helix_hip_problem* pr, const std::vector<int32_t>& perm)
{
	const int H = pr->n_ado, M = pr->n_modes;
	const size_t HM = static_cast<size_t>(H) * M;
	StoredTables t;
	t.up.resize(HM);
	t.down.resize(HM);
	t.tier.resize(H);
	t.dcomm.resize(2 * HM);
	t.danti.resize(2 * HM);
	t.selfc.resize(2 * static_cast<size_t>(H));
	for(int n = 0; n < H; ++n) {
		const size_t to = static_cast<size_t>(perm[n]);
		for(int k = 0; k < M; ++k) {
			const size_t src = static_cast<size_t>(n) * M + k, dst = to * M + k;
			const int u = pr->up[src], d = pr->down[src];
			t.up[dst] = (u < H) ? perm[u] : H;
			t.down[dst] = (d < H) ? perm[d] : H;
			t.dcomm[2 * dst] = pr->down_comm[2 * src];
			t.dcomm[2 * dst + 1] = pr->down_comm[2 * src + 1];
			t.danti[2 * dst] = pr->down_anti[2 * src];
			t.danti[2 * dst + 1] = pr->down_anti[2 * src + 1];
		}
		t.tier[to] = pr->tier[n];
		t.selfc[2 * to] = pr->self_coef[2 * static_cast<size_t>(n)];
		t.selfc[2 * to + 1] = pr->self_coef[2 * static_cast<size_t>(n) + 1];
	}
	return t;
}

// A full [H][N][N] state (reference ADO order) into the stored order and layout (FP64, interleaved); EINVAL
// for a non-Hermitian state in the Hermitian layout.  Tolerance: 1e-12 of the state's largest component.
int to_stored(const double* state, int H, int logN, bool herm, const std::vector<int32_t>& perm, std::vector<double>& st)
{
	const size_t n2 = static_cast<size_t>(1) << (2 * logN);
	const size_t pe = herm ? (n2 + (static_cast<size_t>(1) << logN)) / 2 : n2;
	st.assign(static_cast<size_t>(H) * pe * 2, 0.0);
	double mx = 0.0;
	if(herm)
		for(size_t q = 0; q < static_cast<size_t>(H) * n2 * 2; ++q) mx = std::max(mx, std::fabs(state[q]));
	for(int h = 0; h < H; ++h) {
		const double* src = state + static_cast<size_t>(h) * n2 * 2;
		double* dst = st.data() + static_cast<size_t>(perm.empty() ? h : perm[h]) * pe * 2;
		if(!herm) {
			std::memcpy(dst, src, n2 * 16);
		} else if(!pack_block(src, logN, dst, 1e-12 * mx)) {
			return fail(HELIX_HIP_EINVAL, "state is not Hermitian (ADO " + std::to_string(h) +
				"); the Hermitian layout needs Hermitian ADOs -- create the plan with HELIX_HIP_LAYOUT_FULL");
		}
	}
	return 0;
}

void from_stored(const double* st, int H, int logN, bool herm, const std::vector<int32_t>& perm, double* state)
{
	const size_t n2 = static_cast<size_t>(1) << (2 * logN);
	const size_t pe = herm ? (n2 + (static_cast<size_t>(1) << logN)) / 2 : n2;
	for(int h = 0; h < H; ++h) {
		const double* src = st + static_cast<size_t>(perm.empty() ? h : perm[h]) * pe * 2;
		double* dst = state + static_cast<size_t>(h) * n2 * 2;
		if(herm)
			unpack_block(src, logN, dst);
		else
			std::memcpy(dst, src, n2 * 16);
	}
}

// device stored elements (state precision) -> FP64 interleaved
int download_stored(const void* dev, size_t count, bool fp64, double* out)
{
	if(fp64) {
		HX_HIP(hipMemcpy(out, dev, count * 16, hipMemcpyDeviceToHost));
	} else {
		std::vector<float> f(count * 2);
		HX_HIP(hipMemcpy(f.data(), dev, count * 8, hipMemcpyDeviceToHost));
		for(size_t q = 0; q < count * 2; ++q) out[q] = static_cast<double>(f[q]);
	}
	return 0;
}

int upload_stored(void* dev, const double* st, size_t count, bool fp64)
{
	if(fp64) {
		HX_HIP(hipMemcpy(dev, st, count * 16, hipMemcpyHostToDevice));
	} else {
		std::vector<float> f(count * 2);
		for(size_t q = 0; q < count * 2; ++q) f[q] = static_cast<float>(st[q]);
		HX_HIP(hipMemcpy(dev, f.data(), count * 8, hipMemcpyHostToDevice));
	}
	return 0;
}

int host_to_state(helix_hip_plan* p, const double* state)
{
	std::vector<double> st;
	if(int rc = to_stored(state, p->H, p->logN, p->herm, p->perm, st)) return rc;
	const bool fp64 = p->precision == HELIX_HIP_FP64;
	if(int rc = upload_stored(p->rho, st.data(), p->elems, fp64)) return rc;
	if(!fp64) HX_HIP(hipMemcpy(p->rho0_f64, st.data(), p->pe * 16, hipMemcpyHostToDevice));
	return 0;
}

int state_to_host(helix_hip_plan* p, const void* dev, double* out, bool tier0_master)
{
	std::vector<double> st(p->elems * 2);
	const bool fp64 = p->precision == HELIX_HIP_FP64;
	if(int rc = download_stored(dev, p->elems, fp64, st.data())) return rc;
	if(!fp64 && tier0_master) HX_HIP(hipMemcpy(st.data(), p->rho0_f64, p->pe * 16, hipMemcpyDeviceToHost));
	from_stored(st.data(), p->H, p->logN, p->herm, p->perm, out);
	return 0;
}

// owns a device buffer for the length of one host call
struct DeviceBuffer {
	void* ptr = nullptr;
	~DeviceBuffer()
	{
		if(ptr) (void)hipFree(ptr);
	}
};

// Stored blocks at the given byte offsets from base (pe elements of esize bytes each) gathered on the device and
// downloaded as FP64 interleaved, block after block.
int gather_stored(hipStream_t stream, const void* base, const std::vector<int64_t>& offs, size_t pe, size_t esize,
	std::vector<double>& out)
{
	const int count = static_cast<int>(offs.size());
	const size_t block_bytes = pe * esize;
	const size_t off_bytes = (offs.size() * sizeof(int64_t) + 255) & ~size_t(255);
	DeviceBuffer d;
	HX_HIP(hipMalloc(&d.ptr, off_bytes + count * block_bytes));
	char* dev = static_cast<char*>(d.ptr);
	HX_HIP(hipMemcpyAsync(dev, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
	const int words = static_cast<int>(block_bytes / 8);
	const size_t total = static_cast<size_t>(count) * words;
	const unsigned grid = static_cast<unsigned>(std::min<size_t>((total + kBlock - 1) / kBlock, 4096));
	hipLaunchKernelGGL(k_gather_blocks, dim3(grid), dim3(kBlock), 0, stream, static_cast<const char*>(base),
		reinterpret_cast<const int64_t*>(dev), count, words, reinterpret_cast<uint2*>(dev + off_bytes));
	if(int rc = launch_error("gather blocks")) return rc;
	out.resize(static_cast<size_t>(count) * pe * 2);
	if(esize == 16) {
		HX_HIP(hipMemcpyAsync(out.data(), dev + off_bytes, count * block_bytes, hipMemcpyDeviceToHost, stream));
		HX_HIP(hipStreamSynchronize(stream));
	} else {
		std::vector<float> f(static_cast<size_t>(count) * pe * 2);
		HX_HIP(hipMemcpyAsync(f.data(), dev + off_bytes, count * block_bytes, hipMemcpyDeviceToHost, stream));
		HX_HIP(hipStreamSynchronize(stream));
		for(size_t q = 0; q < f.size(); ++q) out[q] = static_cast<double>(f[q]);
	}
	return 0;
}

// problem coefficients under which L maps Hermitian ADOs to Hermitian ADOs (see helix_hip_layout)
bool preserves_hermiticity(const helix_hip_problem* pr)
{
	const size_t HM = static_cast<size_t>(pr->n_ado) * pr->n_modes;
	for(size_t q = 0; q < HM; ++q)
		if(pr->down_comm[2 * q] != 0.0 || pr->down_anti[2 * q + 1] != 0.0) return false;
	for(int h = 0; h < pr->n_ado; ++h)
		if(pr->self_coef[2 * h + 1] != 0.0) return false;
	for(int b = 0; b < pr->n_baths; ++b)
		if(pr->xi[2 * b + 1] != 0.0 || pr->xi_counter[2 * b] != 0.0) return false;
	if(pr->hamiltonian_dense) {
		const int N = pr->n;
		for(int i = 0; i < N; ++i)
			for(int j = 0; j < N; ++j) {
				const double* a = pr->hamiltonian_dense + 2 * (static_cast<size_t>(i) * N + j);
				const double* b = pr->hamiltonian_dense + 2 * (static_cast<size_t>(j) * N + i);
				if(a[0] != b[0] || a[1] != -b[1]) return false;
			}
	}
	return true;
}

}  // namespace

extern "C" {

const char* helix_hip_last_error(void)
{
	return g_last_error.c_str();
}

int helix_hip_plan_create(const helix_hip_problem* pr, int device, helix_hip_plan** out)
{
	if(!out) return fail(HELIX_HIP_EINVAL, "out is null");
	*out = nullptr;
	if(int rc = validate(pr)) return rc;
	HX_HIP(hipSetDevice(device));
	helix_hip_plan* p = new(std::nothrow) helix_hip_plan();
	if(!p) return fail(HELIX_HIP_ENOMEM, "host allocation failed");
	auto bail = [&](int rc) {
		helix_hip_plan_destroy(p);
		return rc;
	};
	p->device = device;
	if(int rc = resolve_scratch(device, &p->scratch)) return bail(rc);
	p->N = pr->n;
	p->logN = ilog2_exact(pr->n);
	p->H = pr->n_ado;
	p->M = pr->n_modes;
	p->B = pr->n_baths;
	p->order = pr->order;
	p->precision = pr->precision;
	p->dt = pr->dt;
	p->esize = (pr->precision == HELIX_HIP_FP64) ? 16 : 8;
	p->csr_h = pr->hamiltonian_csr != nullptr;
	p->csr_v = pr->couplings_csr != nullptr;
	if(!p->csr_v)
		for(int b = 0; b < p->B; ++b)
			if(pr->couplings[b].nflips > 0) p->flip = true;
	{
		const char* force = std::getenv("HELIX_HIP_GENERIC_STAGE");
		p->tiled = !p->flip && !p->csr_h && !p->csr_v && p->B <= 16 && p->M <= kTileMaxModes &&
			!(force && force[0] == '1');
	}
	if(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
		return bail(fail(HELIX_HIP_ENOMEM, "hipStreamCreate failed"));

	int rc = 0;
	StageArgs& a = p->base;
	std::memset(&a, 0, sizeof(a));
	const size_t HM = static_cast<size_t>(p->H) * p->M;
	p->perm = colex_order(pr->up, p->H, p->M);
	p->mode_bath_host.assign(pr->mode_bath, pr->mode_bath + p->M);
	const StoredTables tabs = permute_tables(pr, p->perm);
	if((rc = upload(p, tabs.up.data(), HM, &a.up)) || (rc = upload(p, tabs.down.data(), HM, &a.down)) ||
		(rc = upload(p, pr->mode_bath, static_cast<size_t>(p->M), &a.mode_bath)))
		return bail(rc);
	const double2* d2 = nullptr;
	if((rc = upload(p, reinterpret_cast<const double2*>(tabs.dcomm.data()), HM, &d2))) return bail(rc);
	a.dcomm = d2;
	if((rc = upload(p, reinterpret_cast<const double2*>(tabs.danti.data()), HM, &d2))) return bail(rc);
	a.danti = d2;
	if((rc = upload(p, reinterpret_cast<const double2*>(tabs.selfc.data()), static_cast<size_t>(p->H), &d2))) return bail(rc);
	a.selfc = d2;
	if((rc = upload(p, reinterpret_cast<const double2*>(pr->xi), static_cast<size_t>(p->B), &d2))) return bail(rc);
	a.xi = d2;
	if((rc = upload(p, reinterpret_cast<const double2*>(pr->xi_counter), static_cast<size_t>(p->B), &d2))) return bail(rc);
	a.xic = d2;
	std::vector<double> vdiag(static_cast<size_t>(p->B) * p->N);
	std::vector<int32_t> vfn(p->B), vfm(static_cast<size_t>(p->B) * kMaxFlips, 0);
	std::vector<double> vfc(static_cast<size_t>(p->B) * kMaxFlips, 0.0);
	for(int b = 0; b < p->B; ++b) {
		std::memcpy(&vdiag[static_cast<size_t>(b) * p->N], pr->couplings[b].diag, sizeof(double) * p->N);
		vfn[b] = pr->couplings[b].nflips;
		for(int f = 0; f < vfn[b]; ++f) {
			vfm[b * kMaxFlips + f] = pr->couplings[b].flip_masks[f];
			vfc[b * kMaxFlips + f] = pr->couplings[b].flip_coefs[f];
		}
	}
	if((rc = upload(p, vdiag.data(), vdiag.size(), &a.vdiag)) || (rc = upload(p, vfn.data(), vfn.size(), &a.vflip_n)) ||
		(rc = upload(p, vfm.data(), vfm.size(), &a.vflip_mask)) || (rc = upload(p, vfc.data(), vfc.size(), &a.vflip_coef)) ||
		(rc = upload(p, pr->hamiltonian.diag, static_cast<size_t>(p->N), &a.hdiag)))
		return bail(rc);
	if(pr->hamiltonian_dense) {
		const size_t n2 = static_cast<size_t>(p->N) * p->N;
		p->dense = true;
		a.hreal = 1;
		for(size_t q = 0; q < n2; ++q)
			if(pr->hamiltonian_dense[2 * q + 1] != 0.0) a.hreal = 0;
		if(p->precision == HELIX_HIP_FP64) {
			const double2* hd2 = nullptr;
			if((rc = upload(p, reinterpret_cast<const double2*>(pr->hamiltonian_dense), n2, &hd2))) return bail(rc);
			a.hdense = hd2;
		} else {
			std::vector<float> hi(2 * n2), lo(2 * n2);
			for(size_t q = 0; q < 2 * n2; ++q) {
				hi[q] = static_cast<float>(pr->hamiltonian_dense[q]);
				lo[q] = static_cast<float>(pr->hamiltonian_dense[q] - static_cast<double>(hi[q]));
			}
			const float2* f2 = nullptr;
			if((rc = upload(p, reinterpret_cast<const float2*>(hi.data()), n2, &f2))) return bail(rc);
			a.hdense = f2;
			if((rc = upload(p, reinterpret_cast<const float2*>(lo.data()), n2, &f2))) return bail(rc);
			a.hdense_lo = f2;
		}
	}
	a.hflips = pr->hamiltonian.nflips;
	for(int f = 0; f < a.hflips; ++f) {
		a.hmask[f] = pr->hamiltonian.flip_masks[f];
		a.hcoef[f] = pr->hamiltonian.flip_coefs[f];
	}
	// sigma^z structure: every coupling diagonal equals alpha_b (2 n_s - 1) for one spin s, no flips
	a.sz = (p->B <= 16 && !p->csr_v) ? 1 : 0;
	for(int b = 0; b < p->B && a.sz; ++b) {
		const helix_hip_spin_operator& op = pr->couplings[b];
		int found = -1;
		for(int s = 0; s < p->logN && found < 0 && op.nflips == 0; ++s) {
			const double alpha = op.diag[1 << s];
			bool ok = alpha != 0.0;
			for(int i = 0; i < p->N && ok; ++i) ok = op.diag[i] == (((i >> s) & 1) ? alpha : -alpha);
			if(ok) found = s;
		}
		if(found < 0) {
			a.sz = 0;
		} else {
			a.sz_mask[b] = 1 << found;
			a.sz_alpha[b] = op.diag[1 << found];
		}
	}
	for(size_t q = 0; q < HM && a.sz; ++q)
		if(pr->down_comm[2 * q] != 0.0 || pr->down_anti[2 * q + 1] != 0.0) a.sz = 0;
	if(a.sz) set_mode_spins(a, pr);
	if(a.sz) {
		std::vector<double> selfx(static_cast<size_t>(p->N) * 2, 0.0);
		for(int x = 0; x < p->N; ++x)
			for(int b = 0; b < p->B; ++b)
				if(x & a.sz_mask[b]) {
					const double d2 = 4.0 * a.sz_alpha[b] * a.sz_alpha[b];
					selfx[2 * x] -= pr->xi[2 * b] * d2;
					selfx[2 * x + 1] -= pr->xi[2 * b + 1] * d2;
				}
		if((rc = upload(p, reinterpret_cast<const double2*>(selfx.data()), static_cast<size_t>(p->N), &d2))) return bail(rc);
		a.selfx = d2;
	}
	if(p->logN > 6 && !a.sz) p->tiled = false;  // see launch_tile
	if(p->dense) p->tiled = false;
	if(p->logN > 10) p->tiled = false;

	// storage layout: Hermitian wherever the tile kernel runs and L preserves it
	const bool can_herm = p->tiled && preserves_hermiticity(pr);
	if(pr->layout == HELIX_HIP_LAYOUT_HERMITIAN && !can_herm)
		return bail(fail(HELIX_HIP_EUNSUPPORTED, "Hermitian layout needs the tiled stage (diagonal couplings) "
			"and Hermiticity-preserving coefficients"));
	p->herm = can_herm && pr->layout != HELIX_HIP_LAYOUT_FULL;
	const size_t n2 = static_cast<size_t>(p->N) * p->N;
	p->pe = p->herm ? (n2 + p->N) / 2 : n2;
	p->elems = static_cast<size_t>(p->H) * p->pe;
	if((rc = plan_alloc(p, &p->rho, p->elems * p->esize))) return bail(rc);
	const int nt = std::max(1, p->order - 1);
	for(int q = 1; q <= nt; ++q)
		if((rc = plan_alloc(p, &p->t[q], p->elems * p->esize))) return bail(rc);
	if(p->precision != HELIX_HIP_FP64) {
		void* r0 = nullptr;
		if((rc = plan_alloc(p, &r0, p->pe * 16))) return bail(rc);
		p->rho0_f64 = static_cast<double*>(r0);
	}
	if((p->flip || p->csr_v) && (rc = plan_alloc(p, &p->buf, p->elems * p->esize * p->B))) return bail(rc);
	if(p->csr_h && (rc = upload_csr(p, pr->hamiltonian_csr, 1, p->N, &a.hcsr_off, &a.hcsr_col, &a.hcsr_val)))
		return bail(rc);
	if(p->csr_v && (rc = upload_csr(p, pr->couplings_csr, p->B, p->N, &a.vcsr_off, &a.vcsr_col, &a.vcsr_val)))
		return bail(rc);
	a.buf = p->buf;
	a.n_ado = p->H;
	a.M = p->M;
	a.B = p->B;
	a.logN = p->logN;
	HX_HIP(hipMemsetAsync(p->rho, 0, p->elems * p->esize, p->stream));
	if(p->rho0_f64) HX_HIP(hipMemsetAsync(p->rho0_f64, 0, p->pe * 16, p->stream));
	HX_HIP(hipStreamSynchronize(p->stream));
	if((rc = autotune_threads(p))) return bail(rc);
	if(p->order >= 2 && (rc = capture_chunk(p))) return bail(rc);
	*out = p;
	return 0;
}

void helix_hip_plan_destroy(helix_hip_plan* p)
{
	if(!p) return;
	if(p->stream) (void)hipStreamSynchronize(p->stream);
	drop_chunk_graph(p);
	for(hipEvent_t e : p->events) (void)hipEventDestroy(e);
	for(void* q : p->allocations) (void)hipFree(q);
	for(void* q : p->term_allocs) (void)hipFree(q);
	if(p->trace) (void)hipFree(p->trace);
	if(p->stream) (void)hipStreamDestroy(p->stream);
	delete p;
}

int helix_hip_set_initial(helix_hip_plan* p, const double* rho0)
{
	if(!p || !rho0) return fail(HELIX_HIP_EINVAL, "null argument");
	// every ADO but tier 0 is zero: the state is cleared on the device and only tier 0's block crosses the bus (a
	// host-built state of all H blocks kept the GPU idle for a large fraction of a second)
	std::vector<double> st;
	if(int rc = to_stored(rho0, 1, p->logN, p->herm, {}, st)) return rc;
	const bool fp64 = p->precision == HELIX_HIP_FP64;
	const size_t esize = fp64 ? 16 : 8;
	HX_HIP(hipStreamSynchronize(p->stream));
	HX_HIP(hipMemset(p->rho, 0, p->elems * esize));
	char* block0 = static_cast<char*>(p->rho) + static_cast<size_t>(p->perm.empty() ? 0 : p->perm[0]) * p->pe * esize;
	if(int rc = upload_stored(block0, st.data(), p->pe, fp64)) return rc;
	if(!fp64) HX_HIP(hipMemcpy(p->rho0_f64, st.data(), p->pe * 16, hipMemcpyHostToDevice));
	HX_HIP(hipDeviceSynchronize());
	return 0;
}

int helix_hip_set_state(helix_hip_plan* p, const double* state)
{
	if(!p || !state) return fail(HELIX_HIP_EINVAL, "null argument");
	return host_to_state(p, state);
}

int helix_hip_get_state(helix_hip_plan* p, double* state)
{
	if(!p || !state) return fail(HELIX_HIP_EINVAL, "null argument");
	HX_HIP(hipStreamSynchronize(p->stream));
	return state_to_host(p, p->rho, state, true);
}

int helix_hip_rhs(helix_hip_plan* p, double* out)
{
	if(!p || !out) return fail(HELIX_HIP_EINVAL, "null argument");
	StageArgs a = p->base;
	a.scale = 1.0;
	a.cur = p->rho;
	a.out = p->t[1];
	dispatch_stage(p, a, false);
	if(int rc = launch_error("rhs launch")) return rc;
	HX_HIP(hipStreamSynchronize(p->stream));
	return state_to_host(p, p->t[1], out, false);
}

int helix_hip_step(helix_hip_plan* p, uint64_t nsteps, uint64_t record_every, double* rho_trace, double* diag_trace)
{
	if(!p) return fail(HELIX_HIP_EINVAL, "plan is null");
	TraceStaging st;
	if(int rc = plan_trace_staging(p, nsteps, record_every, rho_trace, diag_trace, st)) return rc;
	return step_impl(p, nsteps, record_every, rho_trace, diag_trace, st, true);
}

int helix_hip_step_noalloc(helix_hip_plan* p, uint64_t nsteps, uint64_t record_every, double* rho_trace,
	double* diag_trace)
{
	if(!p) return fail(HELIX_HIP_EINVAL, "plan is null");
	if(!p->scratch) return fail(HELIX_HIP_ESTATE, "the device scratch of this plan's device was not resolved");
	TraceStaging st;
	if(int rc = trace_staging_on(p, nsteps, record_every, rho_trace, diag_trace, p->scratch, kScratchDoubles, st))
		return rc;
	return step_impl(p, nsteps, record_every, rho_trace, diag_trace, st, false);
}

int helix_hip_get_tier0(helix_hip_plan* p, double* out)
{
	if(!p || !out) return fail(HELIX_HIP_EINVAL, "null argument");
	HX_HIP(hipStreamSynchronize(p->stream));
	std::vector<double> st(p->pe * 2);
	if(p->precision == HELIX_HIP_FP64) {
		HX_HIP(hipMemcpy(st.data(), p->rho, p->pe * 16, hipMemcpyDeviceToHost));
	} else {
		HX_HIP(hipMemcpy(st.data(), p->rho0_f64, p->pe * 16, hipMemcpyDeviceToHost));
	}
	from_stored(st.data(), 1, p->logN, p->herm, {}, out);
	return 0;
}

int helix_hip_get_ados(helix_hip_plan* p, const int32_t* ados, int count, double* out)
{
	if(!p || !ados || !out || count <= 0) return fail(HELIX_HIP_EINVAL, "bad arguments");
	std::vector<int64_t> offs(static_cast<size_t>(count));
	for(int q = 0; q < count; ++q) {
		if(ados[q] < 0 || ados[q] >= p->H) return fail(HELIX_HIP_EINVAL, "ADO index out of range");
		const int64_t s = p->perm.empty() ? ados[q] : p->perm[ados[q]];
		offs[q] = s * static_cast<int64_t>(p->pe * p->esize);
	}
	HX_HIP(hipStreamSynchronize(p->stream));
	std::vector<double> st;
	if(int rc = gather_stored(p->stream, p->rho, offs, p->pe, p->esize, st)) return rc;
	const size_t n2 = static_cast<size_t>(p->N) * p->N;
	for(int q = 0; q < count; ++q) {
		double* blk = st.data() + static_cast<size_t>(q) * p->pe * 2;
		if(ados[q] == 0 && p->precision != HELIX_HIP_FP64)  // the FP64 master of the tier-0 block
			HX_HIP(hipMemcpy(blk, p->rho0_f64, p->pe * 16, hipMemcpyDeviceToHost));
		from_stored(blk, 1, p->logN, p->herm, {}, out + static_cast<size_t>(q) * n2 * 2);
	}
	return 0;
}

int helix_hip_time_steps(helix_hip_plan* p, uint64_t nsteps, double* stage_ms, double* total_ms)
{
	if(!p || nsteps == 0) return fail(HELIX_HIP_EINVAL, "bad arguments");
	const size_t need = 2 * static_cast<size_t>(p->order) * nsteps;
	while(p->events.size() < need) {
		hipEvent_t e;
		HX_HIP(hipEventCreate(&e));
		p->events.push_back(e);
	}
	size_t q = 0;
	for(uint64_t s = 0; s < nsteps; ++s)
		for(int j = 1; j <= p->order; ++j) {
			HX_HIP(hipEventRecord(p->events[q++], p->stream));
			enqueue_stage(p, j);
			HX_HIP(hipEventRecord(p->events[q++], p->stream));
		}
	if(int rc = launch_error("timed step launch")) return rc;
	HX_HIP(hipEventSynchronize(p->events[q - 1]));
	std::vector<double> acc(p->order, 0.0);
	q = 0;
	for(uint64_t s = 0; s < nsteps; ++s)
		for(int j = 0; j < p->order; ++j) {
			float ms = 0.f;
			HX_HIP(hipEventElapsedTime(&ms, p->events[q], p->events[q + 1]));
			acc[j] += ms;
			q += 2;
		}
	if(stage_ms)
		for(int j = 0; j < p->order; ++j) stage_ms[j] = acc[j] / static_cast<double>(nsteps);
	if(total_ms) {
		float ms = 0.f;
		HX_HIP(hipEventElapsedTime(&ms, p->events[0], p->events[q - 1]));
		*total_ms = ms;
	}
	return 0;
}

int helix_hip_plan_set_closure(helix_hip_plan* p, helix_hip_mlp* mlp, const int32_t* ados, int count, const float* extra,
	int extra_dim)
{
	if(!p) return fail(HELIX_HIP_EINVAL, "plan is null");
	HX_HIP(hipStreamSynchronize(p->stream));
	drop_chunk_graph(p);
	if(!mlp) {
		p->base.kappa = nullptr;
		p->base.kslot = nullptr;
		return p->order >= 2 ? capture_chunk(p) : 0;
	}
	if(!ados || count <= 0 || extra_dim < 0 || (extra_dim > 0 && !extra)) return fail(HELIX_HIP_EINVAL, "bad closure");
	if(p->herm) return fail(HELIX_HIP_EUNSUPPORTED, "the learned closure needs a plan in HELIX_HIP_LAYOUT_FULL");
	int din = 0, dout = 0;
	helix_amd_mlp::dims(mlp, &din, &dout);
	if(din != 2 * p->logN + extra_dim || dout != 2) return fail(HELIX_HIP_EINVAL, "MLP sizes must be (2 log2 N + extra, 2)");
	std::vector<int32_t> slot(static_cast<size_t>(p->H) + p->virt_cap, -1);
	for(int a = 0; a < count; ++a) {
		if(ados[a] < 0 || ados[a] >= p->H) return fail(HELIX_HIP_EINVAL, "closure ADO index out of range");
		slot[p->perm[ados[a]]] = a;
	}
	const size_t E = static_cast<size_t>(p->N) * p->N;
	const size_t rows = static_cast<size_t>(count) * E;
	int rc = 0;
	void* q = nullptr;
	if(!p->kslot || p->kappa_count < count) {
		if((rc = plan_alloc(p, &q, sizeof(int32_t) * (static_cast<size_t>(p->H) + p->virt_cap)))) return rc;
		p->kslot = static_cast<int32_t*>(q);
		if((rc = plan_alloc(p, &q, sizeof(double2) * rows))) return rc;
		p->kappa = static_cast<double2*>(q);
		p->kappa_count = count;
	}
	HX_HIP(hipMemcpy(p->kslot, slot.data(), sizeof(int32_t) * slot.size(), hipMemcpyHostToDevice));
	p->kslot_host = slot;
	float* dextra = nullptr;
	if(extra_dim > 0) {
		HX_HIP(hipMalloc(&dextra, sizeof(float) * count * extra_dim));
		HX_HIP(hipMemcpy(dextra, extra, sizeof(float) * count * extra_dim, hipMemcpyHostToDevice));
	}
	float* x = helix_amd_mlp::input_rows(mlp, static_cast<int>(rows));
	if(!x) return fail(HELIX_HIP_ENOMEM, "closure workspace");
	hipLaunchKernelGGL(k_kappa_features, dim3(grid_for(rows)), dim3(kBlock), 0, p->stream, count, p->logN, dextra, extra_dim, x);
	const float* y = helix_amd_mlp::forward_on(mlp, static_cast<int>(rows), p->stream);
	hipLaunchKernelGGL(k_kappa_store, dim3(grid_for(rows)), dim3(kBlock), 0, p->stream, rows, y, p->kappa);
	if(int e = launch_error("closure launch")) return e;
	HX_HIP(hipStreamSynchronize(p->stream));
	if(dextra) (void)hipFree(dextra);
	p->base.kappa = p->kappa;
	p->base.kslot = p->kslot;
	return p->order >= 2 ? capture_chunk(p) : 0;
}

int helix_hip_plan_set_closure_rates(helix_hip_plan* p, const int32_t* ados, int count, const double* kappa)
{
	if(!p) return fail(HELIX_HIP_EINVAL, "plan is null");
	HX_HIP(hipStreamSynchronize(p->stream));
	drop_chunk_graph(p);
	if(!kappa) {
		p->base.kappa = nullptr;
		p->base.kslot = nullptr;
		return p->order >= 2 ? capture_chunk(p) : 0;
	}
	if(!ados || count <= 0) return fail(HELIX_HIP_EINVAL, "bad closure rates");
	if(p->herm) return fail(HELIX_HIP_EUNSUPPORTED, "closure rates need a plan in HELIX_HIP_LAYOUT_FULL");
	std::vector<int32_t> slot(static_cast<size_t>(p->H) + p->virt_cap, -1);
	for(int a = 0; a < count; ++a) {
		if(ados[a] < 0 || ados[a] >= p->H) return fail(HELIX_HIP_EINVAL, "closure ADO index out of range");
		slot[p->perm[ados[a]]] = a;
	}
	const size_t rows = static_cast<size_t>(count) * p->N * p->N;
	int rc = 0;
	void* q = nullptr;
	if(!p->kslot || p->kappa_count < count) {
		if((rc = plan_alloc(p, &q, sizeof(int32_t) * (static_cast<size_t>(p->H) + p->virt_cap)))) return rc;
		p->kslot = static_cast<int32_t*>(q);
		if((rc = plan_alloc(p, &q, sizeof(double2) * rows))) return rc;
		p->kappa = static_cast<double2*>(q);
		p->kappa_count = count;
	}
	HX_HIP(hipMemcpy(p->kslot, slot.data(), sizeof(int32_t) * slot.size(), hipMemcpyHostToDevice));
	p->kslot_host = slot;
	HX_HIP(hipMemcpy(p->kappa, kappa, sizeof(double2) * rows, hipMemcpyHostToDevice));
	p->base.kappa = p->kappa;
	p->base.kslot = p->kslot;
	return p->order >= 2 ? capture_chunk(p) : 0;
}

int helix_hip_plan_set_terminator(helix_hip_plan* p, const helix_hip_terminator* tm)
{
	if(!p) return fail(HELIX_HIP_EINVAL, "plan is null");
	HX_HIP(hipStreamSynchronize(p->stream));
	drop_chunk_graph(p);
	StageArgs& a = p->base;
	if(!p->orig_up) {
		p->orig_up = a.up;
		p->orig_down = a.down;
		p->orig_dcomm = a.dcomm;
		p->orig_danti = a.danti;
		p->orig_selfc = a.selfc;
	}
	auto free_term = [&]() {
		for(void* q : p->term_allocs) (void)hipFree(q);
		p->term_allocs.clear();
		p->term_corr = nullptr;
	};
	if(!tm) {
		a.up = p->orig_up;
		a.down = p->orig_down;
		a.dcomm = p->orig_dcomm;
		a.danti = p->orig_danti;
		a.selfc = p->orig_selfc;
		a.n_ado = p->H;
		p->term_count = 0;
		free_term();
		return p->order >= 2 ? capture_chunk(p) : 0;
	}
	const int U = tm->count, N = p->N, H = p->H, M = p->M;
	if(U <= 0 || !tm->rate || !tm->src_offsets || !tm->h_vectors || !tm->h_values || tm->n_up < 0 ||
		(tm->n_up > 0 && (!tm->up_ado || !tm->up_mode || !tm->up_virtual)) || tm->iterations < 1 || tm->iterations > 16)
		return fail(HELIX_HIP_EINVAL, "bad terminator");
	if(p->herm || !a.sz || !p->tiled || p->logN > 5)
		return fail(HELIX_HIP_EUNSUPPORTED, "the terminator needs a full-layout plan with sigma^z couplings and N <= 32");
	const int S = tm->src_offsets[U];
	if(tm->src_offsets[0] != 0 || S < 0 || (S > 0 && (!tm->src_ado || !tm->src_mode || !tm->src_comm || !tm->src_anti)))
		return fail(HELIX_HIP_EINVAL, "bad terminator sources");
	for(int u = 0; u < U; ++u)
		if(tm->src_offsets[u + 1] < tm->src_offsets[u]) return fail(HELIX_HIP_EINVAL, "terminator source offsets decrease");
	for(int q = 0; q < S; ++q)
		if(tm->src_ado[q] < 0 || tm->src_ado[q] >= H || tm->src_mode[q] < 0 || tm->src_mode[q] >= M)
			return fail(HELIX_HIP_EINVAL, "terminator source out of range");
	for(int q = 0; q < tm->n_up; ++q)
		if(tm->up_ado[q] < 0 || tm->up_ado[q] >= H || tm->up_mode[q] < 0 || tm->up_mode[q] >= M || tm->up_virtual[q] < 0 ||
			tm->up_virtual[q] >= U)
			return fail(HELIX_HIP_EINVAL, "terminator up entry out of range");
	int rc = 0;
	// state vectors with room for the virtual ADOs (rho keeps its H real blocks)
	if(U > p->virt_cap) {
		const size_t blocks = static_cast<size_t>(H) + U;
		const size_t bytes = blocks * p->pe * p->esize;
		auto replace = [&](void*& buf, bool keep) -> int {
			void* nb = nullptr;
			if(int r = plan_alloc(p, &nb, bytes)) return r;
			HX_HIP(hipMemset(nb, 0, bytes));
			if(keep && buf) HX_HIP(hipMemcpy(nb, buf, p->elems * p->esize, hipMemcpyDeviceToDevice));
			if(buf) {
				auto it = std::find(p->allocations.begin(), p->allocations.end(), buf);
				if(it != p->allocations.end()) p->allocations.erase(it);
				(void)hipFree(buf);
				p->device_bytes -= p->elems * p->esize + static_cast<size_t>(p->virt_cap) * p->pe * p->esize;
			}
			buf = nb;
			return 0;
		};
		if((rc = replace(p->rho, true))) return rc;
		for(int q = 1; q <= std::max(1, p->order - 1); ++q)
			if((rc = replace(p->t[q], false))) return rc;
		p->virt_cap = U;
		if(p->kslot) {  // the closure slots of the virtual rows
			std::vector<int32_t> slot = p->kslot_host;
			slot.resize(static_cast<size_t>(H) + U, -1);
			void* q = nullptr;
			if((rc = plan_alloc(p, &q, sizeof(int32_t) * slot.size()))) return rc;
			HX_HIP(hipMemcpy(q, slot.data(), sizeof(int32_t) * slot.size(), hipMemcpyHostToDevice));
			p->kslot = static_cast<int32_t*>(q);
			p->kslot_host = slot;
			if(a.kslot) a.kslot = p->kslot;
		}
	}
	// tables extended by the U virtual rows: no neighbours, no coefficients; absent entries point past them
	const size_t HM = static_cast<size_t>(H) * M, HUM = (static_cast<size_t>(H) + U) * M;
	std::vector<int32_t> up(HUM, H + U), down(HUM, H + U);
	std::vector<double2> dcomm(HUM, make_double2(0.0, 0.0)), danti(HUM, make_double2(0.0, 0.0));
	std::vector<double2> selfc(static_cast<size_t>(H) + U, make_double2(0.0, 0.0));
	HX_HIP(hipMemcpy(up.data(), p->orig_up, HM * sizeof(int32_t), hipMemcpyDeviceToHost));
	HX_HIP(hipMemcpy(down.data(), p->orig_down, HM * sizeof(int32_t), hipMemcpyDeviceToHost));
	HX_HIP(hipMemcpy(dcomm.data(), p->orig_dcomm, HM * sizeof(double2), hipMemcpyDeviceToHost));
	HX_HIP(hipMemcpy(danti.data(), p->orig_danti, HM * sizeof(double2), hipMemcpyDeviceToHost));
	HX_HIP(hipMemcpy(selfc.data(), p->orig_selfc, static_cast<size_t>(H) * sizeof(double2), hipMemcpyDeviceToHost));
	for(size_t q = 0; q < HM; ++q) {
		if(up[q] >= H) up[q] = H + U;
		if(down[q] >= H) down[q] = H + U;
	}
	for(int q = 0; q < tm->n_up; ++q) {
		const size_t slot = static_cast<size_t>(p->perm[tm->up_ado[q]]) * M + tm->up_mode[q];
		if(up[slot] != H + U) return fail(HELIX_HIP_EINVAL, "terminator up entry where the ADO has an up-neighbour");
		up[slot] = H + tm->up_virtual[q];
	}
	// terminator arguments
	std::vector<int32_t> sado(std::max(S, 1)), smask(std::max(S, 1));
	std::vector<double> salpha(std::max(S, 1));
	for(int q = 0; q < S; ++q) {
		const int b = p->mode_bath_host[tm->src_mode[q]];
		sado[q] = p->perm[tm->src_ado[q]];
		smask[q] = a.sz_mask[b];
		salpha[q] = a.sz_alpha[b];
	}
	free_term();
	auto up_t = [&](const auto* host, size_t count, auto** dev) -> int {
		using X = std::remove_const_t<std::remove_pointer_t<decltype(host)>>;
		void* d = nullptr;
		if(count == 0) count = 1;
		HX_HIP(hipMalloc(&d, count * sizeof(X)));
		p->term_allocs.push_back(d);
		HX_HIP(hipMemcpy(d, host, count * sizeof(X), hipMemcpyHostToDevice));
		*dev = static_cast<const X*>(d);
		return 0;
	};
	const int32_t *d_up = nullptr, *d_down = nullptr;
	const double2 *d_dcomm = nullptr, *d_danti = nullptr, *d_selfc = nullptr;
	TermArgs& t = p->term;
	t = TermArgs{};
	const int32_t* d_off = nullptr;
	const double* d_ev = nullptr;
	if((rc = up_t(up.data(), HUM, &d_up)) || (rc = up_t(down.data(), HUM, &d_down)) || (rc = up_t(dcomm.data(), HUM, &d_dcomm)) ||
		(rc = up_t(danti.data(), HUM, &d_danti)) || (rc = up_t(selfc.data(), selfc.size(), &d_selfc)) ||
		(rc = up_t(reinterpret_cast<const double2*>(tm->rate), static_cast<size_t>(U), &t.rate)) ||
		(rc = up_t(tm->src_offsets, static_cast<size_t>(U) + 1, &d_off)) || (rc = up_t(sado.data(), sado.size(), &t.src_ado)) ||
		(rc = up_t(smask.data(), smask.size(), &t.src_mask)) || (rc = up_t(salpha.data(), salpha.size(), &t.src_alpha)) ||
		(rc = up_t(reinterpret_cast<const double2*>(S ? tm->src_comm : tm->rate), std::max<size_t>(S, 1), &t.src_comm)) ||
		(rc = up_t(reinterpret_cast<const double2*>(S ? tm->src_anti : tm->rate), std::max<size_t>(S, 1), &t.src_anti)) ||
		(rc = up_t(reinterpret_cast<const double2*>(tm->h_vectors), static_cast<size_t>(N) * N, &t.q)) ||
		(rc = up_t(tm->h_values, static_cast<size_t>(N), &d_ev))) {
		free_term();
		return rc;
	}
	t.src_off = d_off;
	t.ev = d_ev;
	t.n_real = H;
	t.iterations = tm->iterations;
	t.selfx = a.selfx;
	a.up = d_up;
	a.down = d_down;
	a.dcomm = d_dcomm;
	a.danti = d_danti;
	a.selfc = d_selfc;
	a.n_ado = H + U;
	p->term_count = U;
	return p->order >= 2 ? capture_chunk(p) : 0;
}

int helix_hip_plan_set_terminator_gain(helix_hip_plan* p, const double* corr)
{
	if(!p) return fail(HELIX_HIP_EINVAL, "plan is null");
	if(p->term_count == 0) return fail(HELIX_HIP_EINVAL, "no terminator is installed on this plan");
	HX_HIP(hipStreamSynchronize(p->stream));
	drop_chunk_graph(p);
	if(p->term_corr) {
		auto it = std::find(p->term_allocs.begin(), p->term_allocs.end(), p->term_corr);
		if(it != p->term_allocs.end()) p->term_allocs.erase(it);
		(void)hipFree(p->term_corr);
		p->term_corr = nullptr;
	}
	p->term.corr = nullptr;
	if(corr) {
		const size_t bytes = static_cast<size_t>(p->term_count) * p->N * p->N * sizeof(double2);
		void* d = nullptr;
		HX_HIP(hipMalloc(&d, bytes));
		p->term_allocs.push_back(d);
		p->term_corr = d;
		HX_HIP(hipMemcpy(d, corr, bytes, hipMemcpyHostToDevice));
		p->term.corr = static_cast<const double2*>(d);
	}
	return p->order >= 2 ? capture_chunk(p) : 0;
}

int helix_hip_plan_info(const helix_hip_plan* p, size_t* device_bytes, int* launches_per_step, int* uses_flip_coupling,
	int* uses_dense_hamiltonian)
{
	if(!p) return fail(HELIX_HIP_EINVAL, "plan is null");
	if(device_bytes) *device_bytes = p->device_bytes;
	if(launches_per_step) *launches_per_step = p->order * ((p->flip || p->csr_v) ? 2 : 1);
	if(uses_flip_coupling) *uses_flip_coupling = p->flip ? 1 : 0;
	if(uses_dense_hamiltonian) *uses_dense_hamiltonian = p->dense ? 1 : 0;
	return 0;
}

int helix_hip_plan_layout(const helix_hip_plan* p, int32_t* layout)
{
	if(!p || !layout) return fail(HELIX_HIP_EINVAL, "null argument");
	*layout = p->herm ? HELIX_HIP_LAYOUT_HERMITIAN : HELIX_HIP_LAYOUT_FULL;
	return 0;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// Batched form
// ------------------------------------------------------------------------------------------------

namespace helix_amd {
namespace {

// sigma^z structure of a problem's couplings (see StageArgs::sz); fills masks/alpha and the
// per-(i XOR j) self table.  Returns false when any coupling is not alpha (2 n_s - 1).
bool sigma_z_structure(const helix_hip_problem* pr, int logN, StageArgs& a, std::vector<double>& selfx)
{
	const int N = pr->n;
	if(pr->n_baths > 16) return false;
	for(int b = 0; b < pr->n_baths; ++b) {
		const helix_hip_spin_operator& op = pr->couplings[b];
		if(op.nflips != 0) return false;
		int found = -1;
		for(int s = 0; s < logN && found < 0; ++s) {
			const double alpha = op.diag[1 << s];
			bool ok = alpha != 0.0;
			for(int i = 0; i < N && ok; ++i) ok = op.diag[i] == (((i >> s) & 1) ? alpha : -alpha);
			if(ok) found = s;
		}
		if(found < 0) return false;
		a.sz_mask[b] = 1 << found;
		a.sz_alpha[b] = op.diag[1 << found];
	}
	const size_t HM = static_cast<size_t>(pr->n_ado) * pr->n_modes;
	for(size_t q = 0; q < HM; ++q)
		if(pr->down_comm[2 * q] != 0.0 || pr->down_anti[2 * q + 1] != 0.0) return false;
	selfx.assign(static_cast<size_t>(N) * 2, 0.0);
	for(int x = 0; x < N; ++x)
		for(int b = 0; b < pr->n_baths; ++b)
			if(x & a.sz_mask[b]) {
				const double d2 = 4.0 * a.sz_alpha[b] * a.sz_alpha[b];
				selfx[2 * x] -= pr->xi[2 * b] * d2;
				selfx[2 * x + 1] -= pr->xi[2 * b + 1] * d2;
			}
	set_mode_spins(a, pr);
	a.sz = 1;
	return true;
}

// host staging of many small tables uploaded as one device allocation
struct Arena {
	std::vector<char> bytes;
	template <typename X>
	size_t add(const X* p, size_t count)
	{
		const size_t off = (bytes.size() + 255) & ~size_t(255);
		bytes.resize(off + count * sizeof(X));
		if(count) std::memcpy(bytes.data() + off, p, count * sizeof(X));
		return off;
	}
};

}  // namespace
}  // namespace helix_amd

struct helix_hip_batch {
	int device = 0;
	hipStream_t stream = nullptr;
	int Q = 0, order = 4, precision = 0;
	bool herm = false;  // every instance stored in the Hermitian layout
	size_t esize = 16;
	std::vector<int> N, logN, H;
	std::vector<size_t> pe;  // stored elements per ADO of each instance
	std::vector<std::vector<int32_t>> perm;  // stored index of each reference ADO, per instance (colex_order)
	std::vector<size_t> elem_off, r0_off;
	size_t total_elems = 0;
	int64_t total_ado = 0;
	void* rho = nullptr;
	void* t[kMaxOrder + 1] = {};
	double* rho0_f64 = nullptr;
	StageArgs* table = nullptr;  // [order][Q]
	int2* tasks = nullptr;  // grouped by block size: N = 2^(kBatchMinLogN + g) from task_off[g], task_cnt[g] long
	int task_off[4] = {};
	int task_cnt[4] = {};
	int group_threads[4] = {};  // workgroup shape of each group (512 or 1024)
	int threads = 0;            // shape of the group with the most tasks (reported by helix_hip_batch_info)
	// persistent form: one launch per stage from per-(block size, XCD) task queues (k_batch_persistent)
	bool persistent = false;
	int pthreads = 0;           // its workgroup shape
	int pgrid[2] = {};          // resident workgroups at 512 / 1024 threads (a multiple of 8)
	int2* ptasks = nullptr;     // queue of block size g and XCD x: ptasks[pqoff[8 g + x] .. pqoff[8 g + x + 1])
	int* pqoff = nullptr;       // [33]
	double pcost[4] = {};       // stored bytes of each block size's instances (workgroups dealt in proportion)
	unsigned* pcounters = nullptr;  // [order][4][8]
	bool use_graph = false;
	hipGraph_t graph = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	std::vector<void*> allocations;
	size_t device_bytes = 0;
	std::vector<hipEvent_t> events;
};

namespace helix_amd {
namespace {

int batch_alloc(helix_hip_batch* b, void** ptr, size_t bytes)
{
	*ptr = nullptr;
	if(bytes == 0) return 0;
	hipError_t e = hipMalloc(ptr, bytes);
	if(e != hipSuccess) return hip_fail(e, "hipMalloc");
	b->allocations.push_back(*ptr);
	b->device_bytes += bytes;
	return 0;
}

template <typename T, bool SPLIT, int THREADS, bool HERM, int LOGN>
void batch_launch_t(helix_hip_batch* b, int j)
{
	constexpr int g = LOGN - kBatchMinLogN;
	const int n = b->task_cnt[g];
	if(n == 0) return;
	const unsigned grid = static_cast<unsigned>(((n + kXcds - 1) / kXcds) * kXcds);
	const StageArgs* tab = b->table + static_cast<size_t>(j - 1) * b->Q;
	const int2* tasks = b->tasks + b->task_off[g];
	if(j == b->order)
		hipLaunchKernelGGL((k_batch_stage<T, SPLIT, true, THREADS, HERM, LOGN>), dim3(grid), dim3(THREADS), 0, b->stream,
			tab, tasks, n);
	else
		hipLaunchKernelGGL((k_batch_stage<T, SPLIT, false, THREADS, HERM, LOGN>), dim3(grid), dim3(THREADS), 0, b->stream,
			tab, tasks, n);
}

template <typename T, bool SPLIT, bool HERM, int LOGN>
void batch_launch_n(helix_hip_batch* b, int j)
{
	if(b->group_threads[LOGN - kBatchMinLogN] == 1024)
		batch_launch_t<T, SPLIT, 1024, HERM, LOGN>(b, j);
	else
		batch_launch_t<T, SPLIT, 512, HERM, LOGN>(b, j);
}

template <typename T, bool SPLIT, bool HERM>
void batch_launch_h(helix_hip_batch* b, int j)
{
	batch_launch_n<T, SPLIT, HERM, 6>(b, j);
	batch_launch_n<T, SPLIT, HERM, 5>(b, j);
	batch_launch_n<T, SPLIT, HERM, 4>(b, j);
	batch_launch_n<T, SPLIT, HERM, 3>(b, j);
}

template <typename T, bool SPLIT, int THREADS, bool HERM>
void batch_launch_persistent(helix_hip_batch* b, int j)
{
	const StageArgs* tab = b->table + static_cast<size_t>(j - 1) * b->Q;
	unsigned* cnt = b->pcounters + static_cast<size_t>(j - 1) * 4 * kXcds;
	unsigned* next = b->pcounters + static_cast<size_t>(j % b->order) * 4 * kXcds;
	const int total = b->pgrid[THREADS == 1024 ? 1 : 0];
	// workgroups dealt to the block sizes in proportion to their stored bytes, at least 8 for a size with work
	BatchQueues bq{b->ptasks, b->pqoff, {}};
	double all = 0.0;
	int present = 0;
	for(int g = 0; g < 4; ++g) {
		all += b->pcost[g];
		present += b->pcost[g] > 0.0 ? 1 : 0;
	}
	const int spare = std::max(0, total / kXcds - present);
	bq.wg_off[0] = 0;
	for(int g = 0; g < 4; ++g) {
		int units = 0;
		if(b->pcost[g] > 0.0) units = 1 + static_cast<int>(spare * (b->pcost[g] / all));
		bq.wg_off[g + 1] = bq.wg_off[g] + units * kXcds;
	}
	const dim3 grid(static_cast<unsigned>(bq.wg_off[4]));
	if(j == b->order)
		hipLaunchKernelGGL((k_batch_persistent<T, SPLIT, true, THREADS, HERM>), grid, dim3(THREADS), 0, b->stream, tab, bq,
			cnt, next);
	else
		hipLaunchKernelGGL((k_batch_persistent<T, SPLIT, false, THREADS, HERM>), grid, dim3(THREADS), 0, b->stream, tab, bq,
			cnt, next);
}

template <typename T, bool SPLIT, bool HERM>
void batch_launch_ph(helix_hip_batch* b, int j)
{
	if(b->pthreads == 1024)
		batch_launch_persistent<T, SPLIT, 1024, HERM>(b, j);
	else
		batch_launch_persistent<T, SPLIT, 512, HERM>(b, j);
}

template <typename T, bool SPLIT>
void batch_launch_p(helix_hip_batch* b, int j)
{
	if(b->persistent) {
		if(b->herm)
			batch_launch_ph<T, SPLIT, true>(b, j);
		else
			batch_launch_ph<T, SPLIT, false>(b, j);
		return;
	}
	if(b->herm)
		batch_launch_h<T, SPLIT, true>(b, j);
	else
		batch_launch_h<T, SPLIT, false>(b, j);
}

// resident workgroups of the persistent stage kernel at `threads` (the inner-stage instance; the last-stage one has
// the same resources), a multiple of the 8 XCDs
template <typename T, bool SPLIT, bool HERM>
int persistent_grid(int device, int threads, int* grid)
{
	int cus = 0, per_cu = 0;
	HX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
	if(threads == 1024)
		HX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persistent<T, SPLIT, false, 1024, HERM>, 1024, 0));
	else
		HX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persistent<T, SPLIT, false, 512, HERM>, 512, 0));
	*grid = std::max(kXcds, ((std::max(per_cu, 1) * cus) / kXcds) * kXcds);
	return 0;
}

int batch_persistent_grids(helix_hip_batch* b)
{
	for(int k = 0; k < 2; ++k) {
		const int th = k ? 1024 : 512;
		int rc = 0;
		switch(b->precision) {
		case HELIX_HIP_FP64:
			rc = b->herm ? persistent_grid<double, false, true>(b->device, th, &b->pgrid[k])
			             : persistent_grid<double, false, false>(b->device, th, &b->pgrid[k]);
			break;
		case HELIX_HIP_FP32:
			rc = b->herm ? persistent_grid<float, false, true>(b->device, th, &b->pgrid[k])
			             : persistent_grid<float, false, false>(b->device, th, &b->pgrid[k]);
			break;
		default:
			rc = b->herm ? persistent_grid<float, true, true>(b->device, th, &b->pgrid[k])
			             : persistent_grid<float, true, false>(b->device, th, &b->pgrid[k]);
			break;
		}
		if(rc) return rc;
	}
	return 0;
}

void batch_enqueue_step(helix_hip_batch* b)
{
	for(int j = 1; j <= b->order; ++j) {
		switch(b->precision) {
		case HELIX_HIP_FP64: batch_launch_p<double, false>(b, j); break;
		case HELIX_HIP_FP32: batch_launch_p<float, false>(b, j); break;
		default: batch_launch_p<float, true>(b, j); break;
		}
	}
}

int batch_time(helix_hip_batch* b, uint64_t nsteps, double* ms)
{
	hipEvent_t e0, e1;
	HX_HIP(hipEventCreate(&e0));
	HX_HIP(hipEventCreate(&e1));
	HX_HIP(hipEventRecord(e0, b->stream));
	for(uint64_t s = 0; s < nsteps; ++s) {
		if(b->graph_exec)
			HX_HIP(hipGraphLaunch(b->graph_exec, b->stream));
		else
			batch_enqueue_step(b);
	}
	HX_HIP(hipEventRecord(e1, b->stream));
	HX_HIP(hipEventSynchronize(e1));
	float f = 0.f;
	HX_HIP(hipEventElapsedTime(&f, e0, e1));
	*ms = f;
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	hipError_t e = hipGetLastError();
	if(e != hipSuccess) return hip_fail(e, "batch launch");
	return 0;
}

}  // namespace
}  // namespace helix_amd

extern "C" {

void helix_hip_batch_destroy(helix_hip_batch* b)
{
	if(!b) return;
	if(b->stream) (void)hipStreamSynchronize(b->stream);
	if(b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
	if(b->graph) (void)hipGraphDestroy(b->graph);
	for(void* q : b->allocations) (void)hipFree(q);
	if(b->stream) (void)hipStreamDestroy(b->stream);
	delete b;
}

int helix_hip_batch_create(const helix_hip_problem* const* problems, int count, int device, unsigned flags,
	helix_hip_batch** out)
{
	if(!out || !problems || count <= 0) return fail(HELIX_HIP_EINVAL, "bad batch arguments");
	*out = nullptr;
	for(int q = 0; q < count; ++q) {
		if(int rc = validate(problems[q])) return rc;
		const helix_hip_problem* pr = problems[q];
		if(pr->precision != problems[0]->precision || pr->order != problems[0]->order)
			return fail(HELIX_HIP_EINVAL, "batched problems must share precision and order");
		if(pr->order < 2) return fail(HELIX_HIP_EUNSUPPORTED, "batched form needs Taylor order >= 2");
		if(pr->n > 64) return fail(HELIX_HIP_EUNSUPPORTED, "batched form supports N <= 64");
		if(pr->hamiltonian_dense || pr->hamiltonian_csr || pr->couplings_csr)
			return fail(HELIX_HIP_EUNSUPPORTED, "batched form uses spin-structured H and sigma^z couplings");
		if(pr->n_modes > kTileMaxModes) return fail(HELIX_HIP_EUNSUPPORTED, "batched form supports up to 16 modes");
	}
	bool herm = true;
	for(int q = 0; q < count; ++q)
		if(problems[q]->layout == HELIX_HIP_LAYOUT_FULL || !preserves_hermiticity(problems[q])) herm = false;
	for(int q = 0; q < count; ++q)
		if(problems[q]->layout == HELIX_HIP_LAYOUT_HERMITIAN && !herm)
			return fail(HELIX_HIP_EUNSUPPORTED, "Hermitian layout asked for a batch that cannot use it");
	HX_HIP(hipSetDevice(device));
	helix_hip_batch* b = new(std::nothrow) helix_hip_batch();
	if(!b) return fail(HELIX_HIP_ENOMEM, "host allocation failed");
	auto bail = [&](int rc) {
		helix_hip_batch_destroy(b);
		return rc;
	};
	b->device = device;
	b->Q = count;
	b->order = problems[0]->order;
	b->precision = problems[0]->precision;
	b->esize = (b->precision == HELIX_HIP_FP64) ? 16 : 8;
	b->herm = herm;
	if(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
		return bail(fail(HELIX_HIP_ENOMEM, "hipStreamCreate failed"));

	// instance layout
	size_t r0 = 0;
	for(int q = 0; q < count; ++q) {
		const helix_hip_problem* pr = problems[q];
		const int lg = ilog2_exact(pr->n);
		const size_t n2 = static_cast<size_t>(pr->n) * pr->n;
		const size_t pe = herm ? (n2 + pr->n) / 2 : n2;
		b->N.push_back(pr->n);
		b->logN.push_back(lg);
		b->H.push_back(pr->n_ado);
		b->pe.push_back(pe);
		b->total_elems += b->total_elems & 1;  // instance states start 16-byte aligned (paired complex64 accesses)
		b->elem_off.push_back(b->total_elems);
		b->total_elems += static_cast<size_t>(pr->n_ado) * pe;
		b->total_ado += pr->n_ado;
		b->r0_off.push_back(r0);
		r0 += pe;
	}
	int rc = 0;
	if((rc = batch_alloc(b, &b->rho, b->total_elems * b->esize))) return bail(rc);
	for(int q = 1; q < b->order; ++q)
		if((rc = batch_alloc(b, &b->t[q], b->total_elems * b->esize))) return bail(rc);
	if(b->precision != HELIX_HIP_FP64) {
		void* p0 = nullptr;
		if((rc = batch_alloc(b, &p0, r0 * 16))) return bail(rc);
		b->rho0_f64 = static_cast<double*>(p0);
	}

	// per-instance tables in one arena; argument blocks hold arena offsets until the upload
	Arena arena;
	struct Offs {
		size_t up, down, mb, dcomm, danti, selfc, xi, xic, hdiag, selfx;
	};
	std::vector<StageArgs> base(count);
	std::vector<Offs> offs(count);
	std::vector<int2> tasks;
	for(int q = 0; q < count; ++q) {
		const helix_hip_problem* pr = problems[q];
		StageArgs& a = base[q];
		std::memset(&a, 0, sizeof(a));
		std::vector<double> selfx;
		if(!sigma_z_structure(pr, b->logN[q], a, selfx))
			return bail(fail(HELIX_HIP_EUNSUPPORTED, "batched form needs sigma^z couplings"));
		const size_t HM = static_cast<size_t>(pr->n_ado) * pr->n_modes;
		b->perm.push_back(colex_order(pr->up, pr->n_ado, pr->n_modes));
		const StoredTables tabs = permute_tables(pr, b->perm.back());
		Offs& o = offs[q];
		o.up = arena.add(tabs.up.data(), HM);
		o.down = arena.add(tabs.down.data(), HM);
		o.mb = arena.add(pr->mode_bath, static_cast<size_t>(pr->n_modes));
		o.dcomm = arena.add(tabs.dcomm.data(), 2 * HM);
		o.danti = arena.add(tabs.danti.data(), 2 * HM);
		o.selfc = arena.add(tabs.selfc.data(), 2 * static_cast<size_t>(pr->n_ado));
		o.xi = arena.add(pr->xi, 2 * static_cast<size_t>(pr->n_baths));
		o.xic = arena.add(pr->xi_counter, 2 * static_cast<size_t>(pr->n_baths));
		o.hdiag = arena.add(pr->hamiltonian.diag, static_cast<size_t>(pr->n));
		o.selfx = arena.add(selfx.data(), selfx.size());
		a.hflips = pr->hamiltonian.nflips;
		for(int f = 0; f < a.hflips; ++f) {
			a.hmask[f] = pr->hamiltonian.flip_masks[f];
			a.hcoef[f] = pr->hamiltonian.flip_coefs[f];
		}
		a.n_ado = pr->n_ado;
		a.M = pr->n_modes;
		a.B = pr->n_baths;
		a.logN = b->logN[q];
	}
	for(int g = 0; g < 4; ++g) {
		b->task_off[g] = static_cast<int>(tasks.size());
		for(int q = 0; q < count; ++q) {
			if(b->logN[q] != kBatchMinLogN + g) continue;
			const int E = 1 << (2 * b->logN[q]);
			const int A = (E >= kTile) ? 1 : kTile / E;
			const int ngroups = (problems[q]->n_ado + A - 1) / A;
			for(int t = 0; t < ngroups; ++t) tasks.push_back(make_int2(q, t));
		}
		b->task_cnt[g] = static_cast<int>(tasks.size()) - b->task_off[g];
	}
	void* dar = nullptr;
	if((rc = batch_alloc(b, &dar, arena.bytes.size()))) return bail(rc);
	HX_HIP(hipMemcpy(dar, arena.bytes.data(), arena.bytes.size(), hipMemcpyHostToDevice));
	const char* ab = static_cast<const char*>(dar);
	std::vector<StageArgs> table(static_cast<size_t>(b->order) * count);
	for(int q = 0; q < count; ++q) {
		StageArgs a = base[q];
		const Offs& o = offs[q];
		a.up = reinterpret_cast<const int32_t*>(ab + o.up);
		a.down = reinterpret_cast<const int32_t*>(ab + o.down);
		a.mode_bath = reinterpret_cast<const int32_t*>(ab + o.mb);
		a.dcomm = reinterpret_cast<const double2*>(ab + o.dcomm);
		a.danti = reinterpret_cast<const double2*>(ab + o.danti);
		a.selfc = reinterpret_cast<const double2*>(ab + o.selfc);
		a.xi = reinterpret_cast<const double2*>(ab + o.xi);
		a.xic = reinterpret_cast<const double2*>(ab + o.xic);
		a.hdiag = reinterpret_cast<const double*>(ab + o.hdiag);
		a.selfx = reinterpret_cast<const double2*>(ab + o.selfx);
		const size_t off = b->elem_off[q] * b->esize;
		auto at = [&](void* base_ptr) { return static_cast<void*>(static_cast<char*>(base_ptr) + off); };
		for(int j = 1; j <= b->order; ++j) {
			StageArgs s = a;
			s.scale = problems[q]->dt / j;
			s.cur = (j == 1) ? at(b->rho) : at(b->t[j - 1]);
			if(j < b->order) {
				s.out = at(b->t[j]);
			} else {
				s.out = at(b->rho);
				s.rho = at(b->rho);
				s.ntprev = b->order - 1;
				for(int k = 0; k < s.ntprev; ++k) s.tprev[k] = at(b->t[k + 1]);
				s.rho0_f64 = b->rho0_f64 ? b->rho0_f64 + 2 * b->r0_off[q] : nullptr;
			}
			table[static_cast<size_t>(j - 1) * count + q] = s;
		}
	}
	void* dt = nullptr;
	if((rc = batch_alloc(b, &dt, table.size() * sizeof(StageArgs)))) return bail(rc);
	HX_HIP(hipMemcpy(dt, table.data(), table.size() * sizeof(StageArgs), hipMemcpyHostToDevice));
	b->table = static_cast<StageArgs*>(dt);
	void* dk = nullptr;
	if((rc = batch_alloc(b, &dk, tasks.size() * sizeof(int2)))) return bail(rc);
	HX_HIP(hipMemcpy(dk, tasks.data(), tasks.size() * sizeof(int2), hipMemcpyHostToDevice));
	b->tasks = static_cast<int2*>(dk);
	{
		// persistent form: per block size, instances to its 8 XCD queues, largest first to the least loaded queue
		// (by stored bytes)
		std::vector<int2> pt;
		int qoff[4 * kXcds + 1];
		for(int g = 0; g < 4; ++g) {
			std::vector<int> qs;
			for(int q = 0; q < count; ++q)
				if(b->logN[q] == kBatchMinLogN + g) qs.push_back(q);
			std::stable_sort(qs.begin(), qs.end(), [&](int u, int v) {
				return static_cast<double>(b->H[u]) * b->pe[u] > static_cast<double>(b->H[v]) * b->pe[v];
			});
			std::vector<std::vector<int>> queue(kXcds);
			double load[kXcds] = {};
			for(int q : qs) {
				const int x = static_cast<int>(std::min_element(load, load + kXcds) - load);
				queue[x].push_back(q);
				load[x] += static_cast<double>(b->H[q]) * b->pe[q];
				b->pcost[g] += static_cast<double>(b->H[q]) * b->pe[q];
			}
			for(int x = 0; x < kXcds; ++x) {
				qoff[g * kXcds + x] = static_cast<int>(pt.size());
				for(int q : queue[x]) {
					const int E = 1 << (2 * b->logN[q]);
					const int A = (E >= kTile) ? 1 : kTile / E;
					const int ngroups = (b->H[q] + A - 1) / A;
					for(int t = 0; t < ngroups; ++t) pt.push_back(make_int2(q, t));
				}
			}
		}
		qoff[4 * kXcds] = static_cast<int>(pt.size());
		void* d = nullptr;
		if((rc = batch_alloc(b, &d, pt.size() * sizeof(int2)))) return bail(rc);
		HX_HIP(hipMemcpy(d, pt.data(), pt.size() * sizeof(int2), hipMemcpyHostToDevice));
		b->ptasks = static_cast<int2*>(d);
		if((rc = batch_alloc(b, &d, sizeof(qoff)))) return bail(rc);
		HX_HIP(hipMemcpy(d, qoff, sizeof(qoff), hipMemcpyHostToDevice));
		b->pqoff = static_cast<int*>(d);
		const size_t cbytes = static_cast<size_t>(b->order) * 4 * kXcds * sizeof(unsigned);
		if((rc = batch_alloc(b, &d, cbytes))) return bail(rc);
		HX_HIP(hipMemset(d, 0, cbytes));
		b->pcounters = static_cast<unsigned*>(d);
		if((rc = batch_persistent_grids(b))) return bail(rc);
	}
	HX_HIP(hipMemsetAsync(b->rho, 0, b->total_elems * b->esize, b->stream));
	for(int q = 1; q < b->order; ++q) HX_HIP(hipMemsetAsync(b->t[q], 0, b->total_elems * b->esize, b->stream));
	if(b->rho0_f64) HX_HIP(hipMemsetAsync(b->rho0_f64, 0, r0 * 16, b->stream));
	HX_HIP(hipStreamSynchronize(b->stream));

	// workgroup shape per block size: complex128 tiles default to 512 threads, complex64 to 1024; the autotuner
	// times each group's two shapes on the all-zero state (stepping zero leaves it zero), the other groups
	// fixed, and keeps the faster
	const int def = (b->precision == HELIX_HIP_FP64) ? 512 : 1024;
	for(int g = 0; g < 4; ++g) b->group_threads[g] = def;
	if(flags & HELIX_HIP_BATCH_AUTOTUNE) {
		for(int g = 0; g < 4; ++g) {
			if(b->task_cnt[g] == 0) continue;
			double best = 1e300;
			int pick = def;
			for(int th : {512, 1024}) {
				b->group_threads[g] = th;
				double ms = 0.0;
				if((rc = batch_time(b, 1, &ms)) || (rc = batch_time(b, 3, &ms))) return bail(rc);
				if(ms < best) {
					best = ms;
					pick = th;
				}
			}
			b->group_threads[g] = pick;
		}
	}
	{
		int most = 0;
		for(int g = 1; g < 4; ++g)
			if(b->task_cnt[g] > b->task_cnt[most]) most = g;
		b->threads = b->group_threads[most];
	}
	// one launch per block size and stage by default (measured faster on the 512-instance workload: 13.0 M against
	// 11.4 M ADO-steps/s FP64, 19.9 M against 14.9 M complex64); HELIX_HIP_BATCH_PERSISTENT asks for one persistent
	// launch per stage, and the autotuner times the persistent launch at both shapes against the per-size launches
	// tuned above and keeps the fastest of the three (HELIX_HIP_BATCH_PER_SIZE rules the persistent launch out)
	b->persistent = (flags & HELIX_HIP_BATCH_PERSISTENT) && !(flags & HELIX_HIP_BATCH_PER_SIZE);
	b->pthreads = (b->precision == HELIX_HIP_FP64) ? 1024 : 512;
	if((flags & HELIX_HIP_BATCH_AUTOTUNE) && !(flags & HELIX_HIP_BATCH_PER_SIZE)) {
		double best = 1e300;
		bool pick_p = true;
		int pick_t = b->pthreads;
		for(int mode = 0; mode < 3; ++mode) {
			b->persistent = mode > 0;
			b->pthreads = mode == 2 ? 1024 : 512;
			double ms = 0.0;
			if((rc = batch_time(b, 1, &ms)) || (rc = batch_time(b, 3, &ms))) return bail(rc);
			if(ms < best) {
				best = ms;
				pick_p = b->persistent;
				pick_t = b->pthreads;
			}
		}
		b->persistent = pick_p;
		b->pthreads = pick_t;
	}
	if(b->persistent) b->threads = b->pthreads;
	if(flags & HELIX_HIP_BATCH_GRAPH) {
		HX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
		batch_enqueue_step(b);
		hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
		if(e != hipSuccess) return bail(hip_fail(e, "hipStreamEndCapture"));
		e = hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0);
		if(e != hipSuccess) return bail(hip_fail(e, "hipGraphInstantiate"));
		e = hipGraphUpload(b->graph_exec, b->stream);
		if(e == hipSuccess) e = hipStreamSynchronize(b->stream);
		if(e != hipSuccess) return bail(hip_fail(e, "hipGraphUpload"));
		b->use_graph = true;
	}
	*out = b;
	return 0;
}

int helix_hip_batch_set_initial(helix_hip_batch* b, int index, const double* rho0)
{
	if(!b || index >= b->Q || (index >= 0 && !rho0)) return fail(HELIX_HIP_EINVAL, "bad arguments");
	HX_HIP(hipStreamSynchronize(b->stream));
	const int q0 = index < 0 ? 0 : index, q1 = index < 0 ? b->Q : index + 1;
	for(int q = q0; q < q1; ++q) {
		const size_t n2 = static_cast<size_t>(b->N[q]) * b->N[q];
		const size_t pe = b->pe[q];
		std::vector<double> r(n2 * 2, 0.0);
		if(index < 0)
			r[2 * (n2 - 1)] = 1.0;  // |N-1><N-1|
		else
			std::memcpy(r.data(), rho0, n2 * 16);
		std::vector<double> st;
		if(int rc = to_stored(r.data(), 1, b->logN[q], b->herm, {}, st)) return rc;
		char* base = static_cast<char*>(b->rho) + b->elem_off[q] * b->esize;
		HX_HIP(hipMemset(base, 0, static_cast<size_t>(b->H[q]) * pe * b->esize));
		if(int rc = upload_stored(base, st.data(), pe, b->precision == HELIX_HIP_FP64)) return rc;
		if(b->precision != HELIX_HIP_FP64)
			HX_HIP(hipMemcpy(b->rho0_f64 + 2 * b->r0_off[q], st.data(), pe * 16, hipMemcpyHostToDevice));
	}
	return 0;
}

int helix_hip_batch_step(helix_hip_batch* b, uint64_t nsteps)
{
	if(!b) return fail(HELIX_HIP_EINVAL, "batch is null");
	for(uint64_t s = 0; s < nsteps; ++s) {
		if(b->graph_exec)
			HX_HIP(hipGraphLaunch(b->graph_exec, b->stream));
		else
			batch_enqueue_step(b);
	}
	if(int rc = launch_error("batch step")) return rc;
	HX_HIP(hipStreamSynchronize(b->stream));
	return 0;
}

int helix_hip_batch_get_tier0(helix_hip_batch* b, int index, double* out)
{
	if(!b || !out || index < 0 || index >= b->Q) return fail(HELIX_HIP_EINVAL, "bad arguments");
	HX_HIP(hipStreamSynchronize(b->stream));
	const size_t pe = b->pe[index];
	std::vector<double> st(pe * 2);
	if(b->precision == HELIX_HIP_FP64)
		HX_HIP(hipMemcpy(st.data(), static_cast<char*>(b->rho) + b->elem_off[index] * 16, pe * 16, hipMemcpyDeviceToHost));
	else
		HX_HIP(hipMemcpy(st.data(), b->rho0_f64 + 2 * b->r0_off[index], pe * 16, hipMemcpyDeviceToHost));
	from_stored(st.data(), 1, b->logN[index], b->herm, {}, out);
	return 0;
}

int helix_hip_batch_get_state(helix_hip_batch* b, int index, double* state)
{
	if(!b || !state || index < 0 || index >= b->Q) return fail(HELIX_HIP_EINVAL, "bad arguments");
	HX_HIP(hipStreamSynchronize(b->stream));
	const size_t pe = b->pe[index];
	const size_t elems = static_cast<size_t>(b->H[index]) * pe;
	const char* base = static_cast<const char*>(b->rho) + b->elem_off[index] * b->esize;
	std::vector<double> st(elems * 2);
	if(int rc = download_stored(base, elems, b->precision == HELIX_HIP_FP64, st.data())) return rc;
	if(b->precision != HELIX_HIP_FP64)
		HX_HIP(hipMemcpy(st.data(), b->rho0_f64 + 2 * b->r0_off[index], pe * 16, hipMemcpyDeviceToHost));
	from_stored(st.data(), b->H[index], b->logN[index], b->herm, b->perm[index], state);
	return 0;
}

int helix_hip_batch_get_ados(helix_hip_batch* b, int index, const int32_t* ados, int count, double* out)
{
	if(!b || !ados || !out || count <= 0 || index < 0 || index >= b->Q) return fail(HELIX_HIP_EINVAL, "bad arguments");
	const size_t pe = b->pe[index];
	std::vector<int64_t> offs(static_cast<size_t>(count));
	for(int q = 0; q < count; ++q) {
		if(ados[q] < 0 || ados[q] >= b->H[index]) return fail(HELIX_HIP_EINVAL, "ADO index out of range");
		offs[q] = static_cast<int64_t>(b->perm[index][ados[q]]) * static_cast<int64_t>(pe * b->esize);
	}
	HX_HIP(hipStreamSynchronize(b->stream));
	const char* base = static_cast<const char*>(b->rho) + b->elem_off[index] * b->esize;
	std::vector<double> st;
	if(int rc = gather_stored(b->stream, base, offs, pe, b->esize, st)) return rc;
	const size_t n2 = static_cast<size_t>(b->N[index]) * b->N[index];
	for(int q = 0; q < count; ++q) {
		double* blk = st.data() + static_cast<size_t>(q) * pe * 2;
		if(ados[q] == 0 && b->precision != HELIX_HIP_FP64)
			HX_HIP(hipMemcpy(blk, b->rho0_f64 + 2 * b->r0_off[index], pe * 16, hipMemcpyDeviceToHost));
		from_stored(blk, 1, b->logN[index], b->herm, {}, out + static_cast<size_t>(q) * n2 * 2);
	}
	return 0;
}

int helix_hip_batch_time_steps(helix_hip_batch* b, uint64_t nsteps, double* total_ms)
{
	if(!b || !total_ms) return fail(HELIX_HIP_EINVAL, "bad arguments");
	return batch_time(b, nsteps, total_ms);
}

int helix_hip_batch_info(const helix_hip_batch* b, int64_t* total_ado, size_t* device_bytes, int* threads, int* graph)
{
	if(!b) return fail(HELIX_HIP_EINVAL, "batch is null");
	if(total_ado) *total_ado = b->total_ado;
	if(device_bytes) *device_bytes = b->device_bytes;
	if(threads) *threads = b->threads;
	if(graph) *graph = b->use_graph ? 1 : 0;
	return 0;
}

int helix_hip_batch_launches(const helix_hip_batch* b, int32_t* launches_per_step)
{
	if(!b || !launches_per_step) return fail(HELIX_HIP_EINVAL, "null argument");
	int per_stage = 1;
	if(!b->persistent) {
		per_stage = 0;
		for(int g = 0; g < 4; ++g) per_stage += b->task_cnt[g] > 0 ? 1 : 0;
	}
	*launches_per_step = per_stage * b->order;
	return 0;
}

int helix_hip_batch_layout(const helix_hip_batch* b, int32_t* layout)
{
	if(!b || !layout) return fail(HELIX_HIP_EINVAL, "null argument");
	*layout = b->herm ? HELIX_HIP_LAYOUT_HERMITIAN : HELIX_HIP_LAYOUT_FULL;
	return 0;
}

}  // extern "C"

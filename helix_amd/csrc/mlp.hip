// mlp.hip -- the learned-closure network: a 3-layer perceptron (in -> H1 -> H2 -> out, tanh on the
// hidden layers) with BF16 weights, FP32 activations and FP32 accumulation, trained on the GPU by the
// kernels below (forward, backward, Adam on FP32 master weights).  No counterpart in the reference.
//
// Weights are stored [out][in] (row-major, one row per output unit).  Every GEMM of the forward and backward
// passes runs on the gfx950 matrix cores (k_gemm_mfma: v_mfma_f32_16x16x16_bf16, FP32 accumulation):
// C[m][n] = sum_k bf16(A(m,k)) bf16(B(k,n)), A and B read through strides so transposes cost nothing, FP32
// operands (activations, back-propagated errors) rounded to BF16 (nearest even) as they are staged in LDS.
#include "helix_mlp.h"

#include <hip/hip_bfloat16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace helix_amd_mlp {

namespace {
thread_local std::string g_err;
int fail(int code, const std::string& m)
{
	g_err = m;
	return code;
}
int hfail(hipError_t e, const char* what)
{
	g_err = std::string(what) + ": " + hipGetErrorString(e);
	return static_cast<int>(e);
}
}  // namespace

#define MLP_HIP(call)                                           \
	do {                                                        \
		hipError_t e_ = (call);                                 \
		if(e_ != hipSuccess) return hfail(e_, #call);           \
	} while(0)

// ---- BF16 MFMA GEMM --------------------------------------------------------------------------------
// One 256-thread workgroup (4 waves) per 64 x 64 tile of C; wave w owns the 32 x 32 quarter (w / 2, w % 2) as
// 2 x 2 MFMA tiles of 16 x 16.  K advances 32 at a time: the A tile [64][32] and the B tile, stored transposed as
// [64 n][32 k], are staged in LDS as BF16 (rows padded to 40 elements), then each wave issues two
// v_mfma_f32_16x16x16_bf16 per 16 x 16 tile.  Operand layout of that instruction: lane l supplies A[i][k] for
// i = l % 16, k = 4 (l / 16) .. +3 and B[k][j] for j = l % 16, the same k; it returns D[i][j] for j = l % 16,
// i = 4 (l / 16) + v in register v.  Rows of M tiles run along grid x (any M), N tiles along grid y.
constexpr int kGT = 64;      // C tile
constexpr int kGK = 32;      // K per LDS stage
constexpr int kGLd = kGK + 8;  // padded LDS row (BF16 elements)

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_bits(float v)  // round to nearest even (finite inputs)
{
	const unsigned u = __float_as_uint(v);
	return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ unsigned short bf16_bits(hip_bfloat16 v)
{
	unsigned short b;
	__builtin_memcpy(&b, &v, sizeof(b));
	return b;
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_gemm_mfma(int M, int N, int K, const TA* __restrict__ A, long sam, long sak,
	const TB* __restrict__ B, long sbk, long sbn, float* __restrict__ C, long ldc, const float* __restrict__ bias,
	int accumulate)
{
	__shared__ unsigned short sa[kGT * kGLd];  // [m][k]
	__shared__ unsigned short sb[kGT * kGLd];  // [n][k]
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int m0 = blockIdx.x * kGT, n0 = blockIdx.y * kGT;
	const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
	f32x4_t acc[2][2];
#pragma unroll
	for(int a = 0; a < 2; ++a)
#pragma unroll
		for(int b = 0; b < 2; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
	// staging: 64 x 32 elements per operand, 8 per thread; the unit-stride index runs fastest over the threads
	const bool a_k_fast = sak == 1, b_k_fast = sbk == 1;
	for(int k0 = 0; k0 < K; k0 += kGK) {
#pragma unroll
		for(int r = 0; r < 8; ++r) {
			const int e = tid + 256 * r;  // 0 .. 2047
			const int mi = a_k_fast ? e / kGK : e % kGT, ki = a_k_fast ? e % kGK : e / kGT;
			const int gm = m0 + mi, gk = k0 + ki;
			unsigned short v = 0;
			if(gm < M && gk < K) v = bf16_bits(A[static_cast<size_t>(gm) * sam + static_cast<size_t>(gk) * sak]);
			sa[mi * kGLd + ki] = v;
			const int ni = b_k_fast ? e / kGK : e % kGT, kj = b_k_fast ? e % kGK : e / kGT;
			const int gn = n0 + ni, gk2 = k0 + kj;
			unsigned short w = 0;
			if(gn < N && gk2 < K) w = bf16_bits(B[static_cast<size_t>(gk2) * sbk + static_cast<size_t>(gn) * sbn]);
			sb[ni * kGLd + kj] = w;
		}
		__syncthreads();
#pragma unroll
		for(int kk = 0; kk < kGK; kk += 16) {
			const int kq = kk + 4 * (lane >> 4);
			bf16x4_t av[2], bv[2];
#pragma unroll
			for(int a = 0; a < 2; ++a) {
				const unsigned short* pa = sa + (wm + 16 * a + (lane & 15)) * kGLd + kq;
				const unsigned short* pb = sb + (wn + 16 * a + (lane & 15)) * kGLd + kq;
				av[a] = bf16x4_t{static_cast<short>(pa[0]), static_cast<short>(pa[1]), static_cast<short>(pa[2]),
					static_cast<short>(pa[3])};
				bv[a] = bf16x4_t{static_cast<short>(pb[0]), static_cast<short>(pb[1]), static_cast<short>(pb[2]),
					static_cast<short>(pb[3])};
			}
#pragma unroll
			for(int a = 0; a < 2; ++a)
#pragma unroll
				for(int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[a], bv[b], acc[a][b], 0, 0, 0);
		}
		__syncthreads();
	}
#pragma unroll
	for(int a = 0; a < 2; ++a)
#pragma unroll
		for(int b = 0; b < 2; ++b)
#pragma unroll
			for(int v = 0; v < 4; ++v) {
				const int m = m0 + wm + 16 * a + 4 * (lane >> 4) + v;
				const int n = n0 + wn + 16 * b + (lane & 15);
				if(m < M && n < N) {
					float out = acc[a][b][v] + (bias ? bias[n] : 0.f);
					float* c = C + static_cast<size_t>(m) * ldc + n;
					*c = accumulate ? *c + out : out;
				}
			}
}

__global__ void k_tanh(float* x, size_t n)
{
	const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
	if(i < n) x[i] = tanhf(x[i]);
}

// dz = da * (1 - a^2)
__global__ void k_tanh_back(float* da, const float* a, size_t n)
{
	const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
	if(i < n) da[i] = da[i] * (1.f - a[i] * a[i]);
}

// d = (y - t) * scale (MSE gradient) and per-block partial loss sum_{i} (y - t)^2
__global__ void k_mse(const float* y, const float* t, float* d, size_t n, float scale, float* partial)
{
	__shared__ float red[256];
	const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
	float e = 0.f;
	if(i < n) {
		const float r = y[i] - t[i];
		d[i] = r * scale;
		e = r * r;
	}
	red[threadIdx.x] = e;
	__syncthreads();
	for(int s = blockDim.x / 2; s > 0; s >>= 1) {
		if(threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
		__syncthreads();
	}
	if(threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// db[j] = sum_m dz[m][j]
__global__ void k_colsum(const float* dz, int M, int N, float* db)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= N) return;
	float s = 0.f;
	for(int m = 0; m < M; ++m) s += dz[static_cast<size_t>(m) * N + j];
	db[j] = s;
}

// Adam on FP32 master parameters, then the BF16 copy used by the forward pass
__global__ void k_adam(float* w, float* g, float* mo, float* ve, hip_bfloat16* wb, size_t n, float lr, float b1, float b2,
	float eps, float c1, float c2)
{
	const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
	if(i >= n) return;
	const float gi = g[i];
	const float m = b1 * mo[i] + (1.f - b1) * gi;
	const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
	mo[i] = m;
	ve[i] = v;
	const float upd = lr * (m / c1) / (sqrtf(v / c2) + eps);
	w[i] -= upd;
	if(wb) wb[i] = static_cast<hip_bfloat16>(w[i]);
}

// dst[r][:] = src[idx[r]][:]
__global__ void k_gather_rows(float* dst, const float* src, const int* idx, int rows, int cols)
{
	const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
	if(i >= static_cast<size_t>(rows) * cols) return;
	const int r = static_cast<int>(i / cols), c = static_cast<int>(i % cols);
	dst[i] = src[static_cast<size_t>(idx[r]) * cols + c];
}

__global__ void k_to_bf16(const float* w, hip_bfloat16* wb, size_t n)
{
	const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
	if(i < n) wb[i] = static_cast<hip_bfloat16>(w[i]);
}

unsigned blocks(size_t n, unsigned b = 256)
{
	return static_cast<unsigned>((n + b - 1) / b);
}

}  // namespace helix_amd_mlp

using namespace helix_amd_mlp;

struct helix_hip_mlp {
	int in = 0, h1 = 0, h2 = 0, out = 0;
	hipStream_t stream = nullptr;
	// parameters: FP32 master [W1 b1 W2 b2 W3 b3] in one block; BF16 copies of W1 W2 W3
	size_t nW1 = 0, nW2 = 0, nW3 = 0, nparam = 0;
	float* p = nullptr;    // master
	float* g = nullptr;    // gradient
	float* am = nullptr;   // Adam first moment
	float* av = nullptr;   // Adam second moment
	hip_bfloat16* wb = nullptr;
	long step = 0;
	// activations / workspace for up to `cap` samples
	int cap = 0;
	float *x = nullptr, *a1 = nullptr, *a2 = nullptr, *y = nullptr, *t = nullptr, *d3 = nullptr, *d2 = nullptr,
		  *d1 = nullptr, *partial = nullptr;
	std::vector<void*> allocs;

	float* W1() { return p; }
	float* b1() { return p + nW1; }
	float* W2() { return p + nW1 + h1; }
	float* b2() { return p + nW1 + h1 + nW2; }
	float* W3() { return p + nW1 + h1 + nW2 + h2; }
	float* b3() { return p + nW1 + h1 + nW2 + h2 + nW3; }
};

namespace {

int mlp_alloc(helix_hip_mlp* m, void** ptr, size_t bytes)
{
	MLP_HIP(hipMalloc(ptr, bytes));
	m->allocs.push_back(*ptr);
	return 0;
}

int ensure_cap(helix_hip_mlp* m, int n)
{
	if(n <= m->cap) return 0;
	const size_t N = static_cast<size_t>(n);
	void* q;
	int rc;
#define MLP_A(field, count)                                   \
	if((rc = mlp_alloc(m, &q, (count) * sizeof(float)))) return rc; \
	m->field = static_cast<float*>(q);
	MLP_A(x, N * m->in)
	MLP_A(a1, N * m->h1)
	MLP_A(a2, N * m->h2)
	MLP_A(y, N * m->out)
	MLP_A(t, N * m->out)
	MLP_A(d3, N * m->out)
	MLP_A(d2, N * m->h2)
	MLP_A(d1, N * m->h1)
	MLP_A(partial, blocks(N * m->out) + 1)
#undef MLP_A
	m->cap = n;
	return 0;
}

template <typename TA, typename TB>
void gemm(hipStream_t s, int M, int N, int K, const TA* A, long sam, long sak, const TB* B, long sbk, long sbn, float* C,
	long ldc, const float* bias, int acc)
{
	dim3 grid((M + kGT - 1) / kGT, (N + kGT - 1) / kGT);
	hipLaunchKernelGGL((k_gemm_mfma<TA, TB>), grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias,
		acc);
}

// forward on m->x (n rows) -> m->y, keeping a1, a2 for the backward pass
void forward(helix_hip_mlp* m, int n)
{
	const hip_bfloat16* W1 = m->wb;
	const hip_bfloat16* W2 = m->wb + m->nW1;
	const hip_bfloat16* W3 = m->wb + m->nW1 + m->nW2;
	// a1 = tanh(x W1^T + b1): A = x [n][in], B(k, j) = W1[j][k]
	gemm(m->stream, n, m->h1, m->in, m->x, m->in, 1, W1, 1, m->in, m->a1, m->h1, m->b1(), 0);
	hipLaunchKernelGGL(k_tanh, dim3(blocks(static_cast<size_t>(n) * m->h1)), dim3(256), 0, m->stream, m->a1,
		static_cast<size_t>(n) * m->h1);
	gemm(m->stream, n, m->h2, m->h1, m->a1, m->h1, 1, W2, 1, m->h1, m->a2, m->h2, m->b2(), 0);
	hipLaunchKernelGGL(k_tanh, dim3(blocks(static_cast<size_t>(n) * m->h2)), dim3(256), 0, m->stream, m->a2,
		static_cast<size_t>(n) * m->h2);
	gemm(m->stream, n, m->out, m->h2, m->a2, m->h2, 1, W3, 1, m->h2, m->y, m->out, m->b3(), 0);
}

// gradients of loss = sum (y - t)^2 / (n * out) into m->g; returns the partial sums in m->partial
void backward(helix_hip_mlp* m, int n)
{
	const size_t no = static_cast<size_t>(n) * m->out;
	const float scale = 2.f / static_cast<float>(no);
	hipLaunchKernelGGL(k_mse, dim3(blocks(no)), dim3(256), 0, m->stream, m->y, m->t, m->d3, no, scale, m->partial);
	const hip_bfloat16* W2 = m->wb + m->nW1;
	const hip_bfloat16* W3 = m->wb + m->nW1 + m->nW2;
	float* gW1 = m->g;
	float* gb1 = m->g + m->nW1;
	float* gW2 = gb1 + m->h1;
	float* gb2 = gW2 + m->nW2;
	float* gW3 = gb2 + m->h2;
	float* gb3 = gW3 + m->nW3;
	// layer 3: gW3[o][k] = sum_n d3[n][o] a2[n][k]; gb3 = colsum d3; d2 = d3 W3
	gemm(m->stream, m->out, m->h2, n, m->d3, 1, m->out, m->a2, m->h2, 1, gW3, m->h2, nullptr, 0);
	hipLaunchKernelGGL(k_colsum, dim3(blocks(m->out)), dim3(256), 0, m->stream, m->d3, n, m->out, gb3);
	gemm(m->stream, n, m->h2, m->out, m->d3, m->out, 1, W3, m->h2, 1, m->d2, m->h2, nullptr, 0);
	hipLaunchKernelGGL(k_tanh_back, dim3(blocks(static_cast<size_t>(n) * m->h2)), dim3(256), 0, m->stream, m->d2, m->a2,
		static_cast<size_t>(n) * m->h2);
	// layer 2
	gemm(m->stream, m->h2, m->h1, n, m->d2, 1, m->h2, m->a1, m->h1, 1, gW2, m->h1, nullptr, 0);
	hipLaunchKernelGGL(k_colsum, dim3(blocks(m->h2)), dim3(256), 0, m->stream, m->d2, n, m->h2, gb2);
	gemm(m->stream, n, m->h1, m->h2, m->d2, m->h2, 1, W2, m->h1, 1, m->d1, m->h1, nullptr, 0);
	hipLaunchKernelGGL(k_tanh_back, dim3(blocks(static_cast<size_t>(n) * m->h1)), dim3(256), 0, m->stream, m->d1, m->a1,
		static_cast<size_t>(n) * m->h1);
	// layer 1
	gemm(m->stream, m->h1, m->in, n, m->d1, 1, m->h1, m->x, m->in, 1, gW1, m->in, nullptr, 0);
	hipLaunchKernelGGL(k_colsum, dim3(blocks(m->h1)), dim3(256), 0, m->stream, m->d1, n, m->h1, gb1);
}

int sync_err(helix_hip_mlp* m, const char* what)
{
	hipError_t e = hipGetLastError();
	if(e != hipSuccess) return hfail(e, what);
	MLP_HIP(hipStreamSynchronize(m->stream));
	return 0;
}

}  // namespace

namespace helix_amd_mlp {

// Device-side inference for callers inside this library: input rows in input_rows(m, n) -> output rows
// (returned pointer), enqueued on `stream` without synchronising.
float* input_rows(helix_hip_mlp* m, int n)
{
	return ensure_cap(m, n) == 0 ? m->x : nullptr;
}

const float* forward_on(helix_hip_mlp* m, int n, hipStream_t stream)
{
	hipStream_t keep = m->stream;
	m->stream = stream;
	forward(m, n);
	m->stream = keep;
	return m->y;
}

int dims(const helix_hip_mlp* m, int* in, int* out)
{
	*in = m->in;
	*out = m->out;
	return 0;
}

}  // namespace helix_amd_mlp

extern "C" {

const char* helix_hip_mlp_last_error(void)
{
	return g_err.c_str();
}

void helix_hip_mlp_destroy(helix_hip_mlp* m)
{
	if(!m) return;
	if(m->stream) (void)hipStreamSynchronize(m->stream);
	for(void* q : m->allocs) (void)hipFree(q);
	if(m->stream) (void)hipStreamDestroy(m->stream);
	delete m;
}

int helix_hip_mlp_create(int in, int hidden1, int hidden2, int out, uint64_t seed, int device, helix_hip_mlp** outp)
{
	if(!outp || in <= 0 || hidden1 <= 0 || hidden2 <= 0 || out <= 0) return fail(-1, "bad MLP sizes");
	*outp = nullptr;
	MLP_HIP(hipSetDevice(device));
	helix_hip_mlp* m = new(std::nothrow) helix_hip_mlp();
	if(!m) return fail(-3, "host allocation failed");
	m->in = in;
	m->h1 = hidden1;
	m->h2 = hidden2;
	m->out = out;
	m->nW1 = static_cast<size_t>(in) * hidden1;
	m->nW2 = static_cast<size_t>(hidden1) * hidden2;
	m->nW3 = static_cast<size_t>(hidden2) * out;
	m->nparam = m->nW1 + hidden1 + m->nW2 + hidden2 + m->nW3 + out;
	int rc = 0;
	void* q;
	if(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
		helix_hip_mlp_destroy(m);
		return fail(-3, "stream");
	}
	for(float** f : {&m->p, &m->g, &m->am, &m->av}) {
		if((rc = mlp_alloc(m, &q, m->nparam * sizeof(float)))) {
			helix_hip_mlp_destroy(m);
			return rc;
		}
		*f = static_cast<float*>(q);
	}
	if((rc = mlp_alloc(m, &q, (m->nW1 + m->nW2 + m->nW3) * sizeof(hip_bfloat16)))) {
		helix_hip_mlp_destroy(m);
		return rc;
	}
	m->wb = static_cast<hip_bfloat16*>(q);
	// Glorot-uniform weights from SplitMix64(seed), zero biases
	std::vector<float> host(m->nparam, 0.f);
	uint64_t st = seed;
	auto uni = [&]() {
		uint64_t z = (st += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		z ^= z >> 31;
		return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
	};
	auto fill = [&](size_t off, int fan_in, int fan_out) {
		const double lim = std::sqrt(6.0 / (fan_in + fan_out));
		for(size_t i = 0; i < static_cast<size_t>(fan_in) * fan_out; ++i)
			host[off + i] = static_cast<float>((2.0 * uni() - 1.0) * lim);
	};
	fill(0, in, hidden1);
	fill(m->nW1 + hidden1, hidden1, hidden2);
	fill(m->nW1 + hidden1 + m->nW2 + hidden2, hidden2, out);
	if(helix_hip_mlp_set_params(m, host.data()) != 0) {
		helix_hip_mlp_destroy(m);
		return -1;
	}
	*outp = m;
	return 0;
}

size_t helix_hip_mlp_param_count(const helix_hip_mlp* m)
{
	return m ? m->nparam : 0;
}

int helix_hip_mlp_set_params(helix_hip_mlp* m, const float* params)
{
	if(!m || !params) return fail(-1, "null argument");
	MLP_HIP(hipMemcpy(m->p, params, m->nparam * sizeof(float), hipMemcpyHostToDevice));
	MLP_HIP(hipMemset(m->am, 0, m->nparam * sizeof(float)));
	MLP_HIP(hipMemset(m->av, 0, m->nparam * sizeof(float)));
	m->step = 0;
	// BF16 copies of W1, W2, W3 (biases stay FP32)
	const size_t o2 = m->nW1 + m->h1, o3 = o2 + m->nW2 + m->h2;
	hipLaunchKernelGGL(k_to_bf16, dim3(blocks(m->nW1)), dim3(256), 0, m->stream, m->p, m->wb, m->nW1);
	hipLaunchKernelGGL(k_to_bf16, dim3(blocks(m->nW2)), dim3(256), 0, m->stream, m->p + o2, m->wb + m->nW1, m->nW2);
	hipLaunchKernelGGL(k_to_bf16, dim3(blocks(m->nW3)), dim3(256), 0, m->stream, m->p + o3, m->wb + m->nW1 + m->nW2, m->nW3);
	return sync_err(m, "set_params");
}

int helix_hip_mlp_get_params(helix_hip_mlp* m, float* params)
{
	if(!m || !params) return fail(-1, "null argument");
	MLP_HIP(hipStreamSynchronize(m->stream));
	MLP_HIP(hipMemcpy(params, m->p, m->nparam * sizeof(float), hipMemcpyDeviceToHost));
	return 0;
}

int helix_hip_mlp_forward(helix_hip_mlp* m, const float* x, int n, float* y)
{
	if(!m || !x || !y || n <= 0) return fail(-1, "bad arguments");
	if(int rc = ensure_cap(m, n)) return rc;
	MLP_HIP(hipMemcpyAsync(m->x, x, static_cast<size_t>(n) * m->in * sizeof(float), hipMemcpyHostToDevice, m->stream));
	forward(m, n);
	MLP_HIP(hipMemcpyAsync(y, m->y, static_cast<size_t>(n) * m->out * sizeof(float), hipMemcpyDeviceToHost, m->stream));
	return sync_err(m, "forward");
}

int helix_hip_mlp_gradient(helix_hip_mlp* m, const float* x, const float* t, int n, float* grad, float* loss)
{
	if(!m || !x || !t || n <= 0) return fail(-1, "bad arguments");
	if(int rc = ensure_cap(m, n)) return rc;
	MLP_HIP(hipMemcpyAsync(m->x, x, static_cast<size_t>(n) * m->in * sizeof(float), hipMemcpyHostToDevice, m->stream));
	MLP_HIP(hipMemcpyAsync(m->t, t, static_cast<size_t>(n) * m->out * sizeof(float), hipMemcpyHostToDevice, m->stream));
	forward(m, n);
	backward(m, n);
	if(grad) MLP_HIP(hipMemcpyAsync(grad, m->g, m->nparam * sizeof(float), hipMemcpyDeviceToHost, m->stream));
	const size_t no = static_cast<size_t>(n) * m->out;
	std::vector<float> part(blocks(no));
	MLP_HIP(hipMemcpyAsync(part.data(), m->partial, part.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
	if(int rc = sync_err(m, "gradient")) return rc;
	if(loss) {
		double s = 0.0;
		for(float v : part) s += v;
		*loss = static_cast<float>(s / static_cast<double>(no));
	}
	return 0;
}

int helix_hip_mlp_train(helix_hip_mlp* m, const float* x, const float* t, int n, int epochs, int batch, float lr,
	uint64_t shuffle_seed, float* epoch_loss)
{
	if(!m || !x || !t || n <= 0 || epochs <= 0 || batch <= 0) return fail(-1, "bad arguments");
	const int bsz = batch < n ? batch : n;
	if(int rc = ensure_cap(m, bsz)) return rc;
	// the whole data set stays resident on the device for the run
	float *dx = nullptr, *dt = nullptr;
	int* didx = nullptr;
	MLP_HIP(hipMalloc(&didx, static_cast<size_t>(n) * sizeof(int)));
	MLP_HIP(hipMalloc(&dx, static_cast<size_t>(n) * m->in * sizeof(float)));
	MLP_HIP(hipMalloc(&dt, static_cast<size_t>(n) * m->out * sizeof(float)));
	MLP_HIP(hipMemcpy(dx, x, static_cast<size_t>(n) * m->in * sizeof(float), hipMemcpyHostToDevice));
	MLP_HIP(hipMemcpy(dt, t, static_cast<size_t>(n) * m->out * sizeof(float), hipMemcpyHostToDevice));
	std::vector<int> order(n);
	for(int i = 0; i < n; ++i) order[i] = i;
	uint64_t st = shuffle_seed;
	auto next = [&]() {
		uint64_t z = (st += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	};
	int rc = 0;
	std::vector<float> part;
	for(int ep = 0; ep < epochs && rc == 0; ++ep) {
		for(int i = n - 1; i > 0; --i) std::swap(order[i], order[next() % static_cast<uint64_t>(i + 1)]);
		if(hipMemcpyAsync(didx, order.data(), static_cast<size_t>(n) * sizeof(int), hipMemcpyHostToDevice, m->stream) !=
			hipSuccess) {
			rc = hfail(hipGetLastError(), "upload order");
			break;
		}
		double esum = 0.0;
		size_t ecount = 0;
		for(int b0 = 0; b0 + bsz <= n; b0 += bsz) {
			// gather the minibatch rows (device to device)
			hipLaunchKernelGGL(k_gather_rows, dim3(blocks(static_cast<size_t>(bsz) * m->in)), dim3(256), 0, m->stream,
				m->x, dx, didx + b0, bsz, m->in);
			hipLaunchKernelGGL(k_gather_rows, dim3(blocks(static_cast<size_t>(bsz) * m->out)), dim3(256), 0, m->stream,
				m->t, dt, didx + b0, bsz, m->out);
			forward(m, bsz);
			backward(m, bsz);
			++m->step;
			const float b1 = 0.9f, b2 = 0.999f;
			const float c1 = 1.f - std::pow(b1, static_cast<float>(m->step));
			const float c2 = 1.f - std::pow(b2, static_cast<float>(m->step));
			const size_t o2 = m->nW1 + m->h1, o3 = o2 + m->nW2 + m->h2;
			// one Adam launch per parameter segment so each weight segment refreshes its BF16 copy
			struct Seg {
				size_t off, n;
				hip_bfloat16* wb;
			} segs[] = {{0, m->nW1, m->wb}, {m->nW1, static_cast<size_t>(m->h1), nullptr},
				{o2, m->nW2, m->wb + m->nW1}, {o2 + m->nW2, static_cast<size_t>(m->h2), nullptr},
				{o3, m->nW3, m->wb + m->nW1 + m->nW2}, {o3 + m->nW3, static_cast<size_t>(m->out), nullptr}};
			for(const Seg& s : segs)
				hipLaunchKernelGGL(k_adam, dim3(blocks(s.n)), dim3(256), 0, m->stream, m->p + s.off, m->g + s.off,
					m->am + s.off, m->av + s.off, s.wb, s.n, lr, b1, b2, 1e-8f, c1, c2);
			const size_t no = static_cast<size_t>(bsz) * m->out;
			part.resize(blocks(no));
			(void)hipMemcpyAsync(part.data(), m->partial, part.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream);
			if((rc = sync_err(m, "train step"))) break;
			for(float v : part) esum += v;
			ecount += no;
		}
		if(epoch_loss) epoch_loss[ep] = static_cast<float>(esum / static_cast<double>(ecount ? ecount : 1));
	}
	(void)hipFree(dx);
	(void)hipFree(dt);
	(void)hipFree(didx);
	return rc;
}

}  // extern "C"

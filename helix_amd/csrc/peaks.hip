// peaks.hip -- attainable peaks of the device, measured by this library's own kernels: HBM read bandwidth, HBM
// copy bandwidth, and the dense throughput of the two matrix-core instructions the stage kernels use
// (v_mfma_f64_16x16x4f64 for complex128 states, v_mfma_f32_16x16x4f32 for complex64).  They put the roofline
// fractions computed against the datasheet peaks beside what this GPU sustains.  No counterpart in the reference.
#include "helix_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

namespace helix_amd {

void set_last_error(const std::string& msg);  // heom_plan.hip: the text helix_hip_last_error returns

namespace {

constexpr int kPeakBlock = 256;
constexpr int kChains = 8;   // independent MFMA accumulators per wave
constexpr int kMfmaIters = 2048;

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// HBM streams in one pass: every workgroup reads (or copies) one contiguous chunk of U x kPeakBlock 16-byte
// elements and exits.  Many short-lived workgroups sustain more than a grid-stride loop over a resident grid on
// MI355X (copy 6.7 against 4.4-5.2 TB/s for grid-stride shapes, read 7.1 against 6.0), and more than larger chunks
// per workgroup (copy 5.7-6.2 TB/s at 32 KiB).  The read's XOR of every word reaches memory only through a
// comparison that zero-filled data never meets, so no load is dead.
constexpr int kReadPerThread = 2;

__global__ __launch_bounds__(kPeakBlock) void k_peak_read(const u32x4* __restrict__ src, size_t n, unsigned* __restrict__ sink)
{
	const size_t base = static_cast<size_t>(blockIdx.x) * kReadPerThread * kPeakBlock + threadIdx.x;
	unsigned acc = 0;
#pragma unroll
	for(int u = 0; u < kReadPerThread; ++u) {
		if(base + u * kPeakBlock < n) {
			const u32x4 v = __builtin_nontemporal_load(src + base + u * kPeakBlock);
			acc ^= v.x ^ v.y ^ v.z ^ v.w;
		}
	}
	if(acc == 0x9E3779B9u) sink[0] = acc;
}

// copy: one element per thread, non-temporal like the stage kernels' output stores
__global__ __launch_bounds__(kPeakBlock) void k_peak_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n)
{
	const size_t i = static_cast<size_t>(blockIdx.x) * kPeakBlock + threadIdx.x;
	if(i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Matrix cores: every wave runs kChains independent accumulation chains of kMfmaIters instructions each
// (2 * 16 * 16 * 4 = 2048 flops per instruction), two waves per SIMD, so 16 instructions are independent at any
// time; one value per thread is stored so no chain is dead.
__global__ __launch_bounds__(kPeakBlock) void k_peak_mfma_f64(double* __restrict__ sink)
{
	const double a = 1.0 + 1e-9 * threadIdx.x, b = 1.0 - 1e-9 * threadIdx.x;
	f64x4 c[kChains];
#pragma unroll
	for(int ch = 0; ch < kChains; ++ch) c[ch] = f64x4{0.0, 0.0, 0.0, 0.0};
	for(int it = 0; it < kMfmaIters; ++it) {
#pragma unroll
		for(int ch = 0; ch < kChains; ++ch) c[ch] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c[ch], 0, 0, 0);
	}
	double s = 0.0;
#pragma unroll
	for(int ch = 0; ch < kChains; ++ch) s += c[ch][0] + c[ch][1] + c[ch][2] + c[ch][3];
	sink[static_cast<size_t>(blockIdx.x) * kPeakBlock + threadIdx.x] = s;
}

__global__ __launch_bounds__(kPeakBlock) void k_peak_mfma_f32(float* __restrict__ sink)
{
	const float a = 1.0f + 1e-6f * threadIdx.x, b = 1.0f - 1e-6f * threadIdx.x;
	f32x4 c[kChains];
#pragma unroll
	for(int ch = 0; ch < kChains; ++ch) c[ch] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
	for(int it = 0; it < kMfmaIters; ++it) {
#pragma unroll
		for(int ch = 0; ch < kChains; ++ch) c[ch] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c[ch], 0, 0, 0);
	}
	float s = 0.0f;
#pragma unroll
	for(int ch = 0; ch < kChains; ++ch) s += c[ch][0] + c[ch][1] + c[ch][2] + c[ch][3];
	sink[static_cast<size_t>(blockIdx.x) * kPeakBlock + threadIdx.x] = s;
}

struct DevBuf {
	void* p = nullptr;
	~DevBuf()
	{
		if(p) (void)hipFree(p);
	}
};

int hip_err(hipError_t e, const char* what)
{
	set_last_error(std::string(what) + ": " + hipGetErrorString(e));
	return static_cast<int>(e);
}

#define PK_HIP(call)                                          \
	do {                                                      \
		const hipError_t pk_e_ = (call);                      \
		if(pk_e_ != hipSuccess) return hip_err(pk_e_, #call); \
	} while(0)

constexpr int kPerTiming = 8;  // back-to-back launches between one pair of events, so the gap before the first
                               // launch is a small part of what is timed

// best over reps timings of kPerTiming launches each (after one untimed launch), in ms per launch
template <typename Launch>
int best_ms(int reps, hipStream_t s, Launch launch, float* out)
{
	hipEvent_t e0, e1;
	PK_HIP(hipEventCreate(&e0));
	if(hipEventCreate(&e1) != hipSuccess) {
		(void)hipEventDestroy(e0);
		return hip_err(hipErrorOutOfMemory, "hipEventCreate");
	}
	float best = 0.0f;
	hipError_t e = hipSuccess;
	launch();
	for(int r = 0; r < reps && e == hipSuccess; ++r) {
		e = hipEventRecord(e0, s);
		for(int k = 0; k < kPerTiming && e == hipSuccess; ++k) {
			launch();
			e = hipGetLastError();
		}
		if(e == hipSuccess) e = hipEventRecord(e1, s);
		if(e == hipSuccess) e = hipEventSynchronize(e1);
		float ms = 0.0f;
		if(e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
		ms /= kPerTiming;
		if(e == hipSuccess && (r == 0 || ms < best)) best = ms;
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if(e != hipSuccess) return hip_err(e, "timed launch");
	*out = best;
	return 0;
}

}  // namespace
}  // namespace helix_amd

extern "C" int helix_hip_measure_peaks(int device, size_t bytes, int reps, helix_hip_peaks* out)
{
	using namespace helix_amd;
	if(!out || reps <= 0 || bytes < (static_cast<size_t>(1) << 20) || bytes > (static_cast<size_t>(1) << 40)) {
		set_last_error("helix_hip_measure_peaks: out must be non-null, reps > 0 and 1 MiB <= bytes <= 1 TiB");
		return HELIX_HIP_EINVAL;
	}
	*out = helix_hip_peaks{};
	int count = 0;
	PK_HIP(hipGetDeviceCount(&count));
	if(device < 0 || device >= count) {
		set_last_error("helix_hip_measure_peaks: no such device");
		return HELIX_HIP_EINVAL;
	}
	PK_HIP(hipSetDevice(device));
	hipDeviceProp_t prop;
	PK_HIP(hipGetDeviceProperties(&prop, device));
	const int cus = std::max(1, prop.multiProcessorCount);
	const size_t n = bytes / sizeof(u32x4);
	const size_t read_blocks = (n + kReadPerThread * kPeakBlock - 1) / (kReadPerThread * kPeakBlock);
	const size_t copy_blocks = (n + kPeakBlock - 1) / kPeakBlock;
	const int mfma_blocks = cus * 2;  // 8 waves per CU, 2 per SIMD (the FP64 loop holds 136 registers: 3 fit)
	DevBuf src, dst, sink;
	PK_HIP(hipMalloc(&src.p, n * sizeof(u32x4)));
	PK_HIP(hipMalloc(&dst.p, n * sizeof(u32x4)));
	PK_HIP(hipMalloc(&sink.p, static_cast<size_t>(mfma_blocks) * kPeakBlock * sizeof(double)));
	PK_HIP(hipMemset(src.p, 0, n * sizeof(u32x4)));
	PK_HIP(hipDeviceSynchronize());
	const hipStream_t s = nullptr;
	float ms_read = 0.0f, ms_copy = 0.0f, ms_f64 = 0.0f, ms_f32 = 0.0f;
	int rc = best_ms(reps, s, [&] {
		hipLaunchKernelGGL(k_peak_read, dim3(static_cast<unsigned>(read_blocks)), dim3(kPeakBlock), 0, s, static_cast<const u32x4*>(src.p), n,
			static_cast<unsigned*>(sink.p));
	}, &ms_read);
	if(!rc)
		rc = best_ms(reps, s, [&] {
			hipLaunchKernelGGL(k_peak_copy, dim3(static_cast<unsigned>(copy_blocks)), dim3(kPeakBlock), 0, s, static_cast<const u32x4*>(src.p),
				static_cast<u32x4*>(dst.p), n);
		}, &ms_copy);
	if(!rc)
		rc = best_ms(reps, s, [&] {
			hipLaunchKernelGGL(k_peak_mfma_f64, dim3(mfma_blocks), dim3(kPeakBlock), 0, s, static_cast<double*>(sink.p));
		}, &ms_f64);
	if(!rc)
		rc = best_ms(reps, s, [&] {
			hipLaunchKernelGGL(k_peak_mfma_f32, dim3(mfma_blocks), dim3(kPeakBlock), 0, s, static_cast<float*>(sink.p));
		}, &ms_f32);
	if(rc) return rc;
	const double moved = static_cast<double>(n) * sizeof(u32x4);
	const double mfma_flops = static_cast<double>(mfma_blocks) * (kPeakBlock / 64) * kChains * kMfmaIters * 2048.0;
	out->hbm_read_gbs = moved / (ms_read * 1e-3) / 1e9;
	out->hbm_copy_gbs = 2.0 * moved / (ms_copy * 1e-3) / 1e9;
	out->mfma_f64_tflops = mfma_flops / (ms_f64 * 1e-3) / 1e12;
	out->mfma_f32_tflops = mfma_flops / (ms_f32 * 1e-3) / 1e12;
	out->bytes = n * sizeof(u32x4);
	out->compute_units = cus;
	return 0;
}

// Bandwidth reference kernels: the measured ceilings the SpMV kernels are
// compared with, and the known-size traffic that calibrates the hardware
// byte counters (FETCH_SIZE / WRITE_SIZE) of the profiler.
//   stream_read:  reads n doubles once (16-byte loads), writes one partial
//                 sum per workgroup
//   stream_copy:  dst[i] = src[i] over n doubles (16-byte loads and stores)
//   gather_cost:  the vector-memory cost of one 64-lane load by width and
//                 address pattern, on a cache-resident buffer (what the SpMV
//                 kernels' x gathers pay per instruction)

#include "format.hpp"

namespace spmvhip {

typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(kBlock) stream_read_kernel(const dbl2 *__restrict__ src, long n2,
		double *__restrict__ partial)
{
	__shared__ double red[kWavesPerBlock];
	double acc = 0;
	for (long i = (long) blockIdx.x * kBlock + threadIdx.x; i < n2; i += (long) gridDim.x * kBlock) {
		dbl2 v = __builtin_nontemporal_load(src + i);
		acc += v.x + v.y;
	}
	acc = wave_sum(acc);
	if (lane_id() == 0)
		red[threadIdx.x / kWave] = acc;
	__syncthreads();
	if (threadIdx.x == 0)
		partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(kBlock) stream_copy_kernel(const dbl2 *__restrict__ src, dbl2 *__restrict__ dst,
		long n2)
{
	for (long i = (long) blockIdx.x * kBlock + threadIdx.x; i < n2; i += (long) gridDim.x * kBlock)
		__builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Stream read through direct global-to-LDS loads (16 bytes per lane, no
// VGPR staging): per round a wave requests 8 KiB into its LDS buffer, waits,
// and reads the round back from LDS. Compared with stream_read_kernel it
// shows what the load path costs when the data lands in LDS instead of VGPRs.
typedef unsigned int u32x4p __attribute__((ext_vector_type(4)));
constexpr int kLdsRound = 8; // 64-lane loads per round

__global__ void __launch_bounds__(kBlock) stream_read_lds_kernel(const u32x4p *__restrict__ src, long n16,
		double *__restrict__ partial)
{
	__shared__ u32x4p buf[kWavesPerBlock][kLdsRound * kWave];
	__shared__ double red[kWavesPerBlock];
	const int lane = lane_id(), wv = threadIdx.x / kWave;
	u32x4p *b = buf[wv];
	uint32_t acc = 0;
	const long W = (long) gridDim.x * kWavesPerBlock;
	for (long c = (long) blockIdx.x * kWavesPerBlock + wv; c * kLdsRound * kWave < n16; c += W) {
		const long base = c * kLdsRound * kWave;
#pragma unroll
		for (int k = 0; k < kLdsRound; k++) {
			const long i = base + k * kWave + lane;
			if (i < n16)
				__builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *) (src + i),
						(__attribute__((address_space(3))) void *) (b + k * kWave), 16, 0, 2);
		}
		__builtin_amdgcn_s_waitcnt(0);
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		__builtin_amdgcn_wave_barrier();
#pragma unroll
		for (int k = 0; k < kLdsRound; k++) {
			const u32x4p v = b[k * kWave + lane];
			acc += v.x ^ v.w;
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
	}
	const double a = wave_sum((double) acc);
	if (lane == 0)
		red[wv] = a;
	__syncthreads();
	if (threadIdx.x == 0)
		partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// One wave's load pattern per iteration, every index masked into the buffer
// (a power-of-two element count `mask + 1`): 0 contiguous 4-byte, 1 contiguous
// 8-byte, 2 contiguous 16-byte, 3 8-byte gather over 64 lines, 4 the same
// with half the lanes, 5 16-byte gather over 64 lines, 6 8-byte gather over
// 16 lines (4 lanes per line), 7 8-byte gather over 4 lines (16 lanes per
// 128-byte line); 8..13 the 8-byte gather over 64 lines of mode 3 as buffer
// loads with cache-policy bits none, sc0, nt, sc1, sc0 sc1, sc0 sc1 nt (the
// L1 and L2 treatment of a gathered line).
__device__ __forceinline__ uint32_t mix32(uint32_t a)
{
	a ^= a >> 16;
	a *= 0x7feb352dU;
	a ^= a >> 15;
	a *= 0x846ca68bU;
	a ^= a >> 16;
	return a;
}

template <int MODE>
__global__ void __launch_bounds__(kBlock) gather_cost_kernel(const double *__restrict__ buf, uint32_t mask,
		int iters, double *__restrict__ sink)
{
	const uint32_t lane = lane_id(), w = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
	double acc = 0;
	uint32_t s = mix32(w * kWave + lane);
	// buffer descriptor over the mask + 1 doubles of buf (modes 8..13)
	const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void *) buf, 0,
			(int) ((mask + 1) * 8u), 0x00020000);
#pragma unroll 8
	for (int it = 0; it < iters; it++) {
		const uint32_t base = (w * 131u + (uint32_t) it) * kWave;
		if (MODE == 0) {
			acc += ((const float *) buf)[(base + lane) & (2 * mask + 1)];
		} else if (MODE == 1) {
			acc += buf[(base + lane) & mask];
		} else if (MODE == 2) {
			const dbl2 v = ((const dbl2 *) buf)[(base + lane) & (mask >> 1)];
			acc += v.x + v.y;
		} else if (MODE == 3 || MODE == 4) {
			s = s * 1664525u + 1013904223u;
			if (MODE == 3 || (lane & 1))
				acc += buf[(s >> 8) & mask];
		} else if (MODE == 5) {
			s = s * 1664525u + 1013904223u;
			const dbl2 v = ((const dbl2 *) buf)[(s >> 8) & (mask >> 1)];
			acc += v.x + v.y;
		} else if (MODE == 6) {
			acc += buf[(mix32(base + (lane >> 2)) * 16 + (lane & 3)) & mask];
		} else if (MODE == 7) {
			acc += buf[(mix32(base + (lane >> 4)) * 16 + (lane & 15)) & mask];
		} else {
			// cache-policy bits: sc0 = 1, nt = 2, sc1 = 16
			constexpr int kPol = MODE == 8 ? 0 : MODE == 9 ? 1 : MODE == 10 ? 2 : MODE == 11 ? 16 : MODE == 12 ? 17 : 19;
			s = s * 1664525u + 1013904223u;
			const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, ((s >> 8) & mask) * 8, 0, kPol);
			double d;
			__builtin_memcpy(&d, &v, 8);
			acc += d;
		}
	}
	if (acc == 12345.678) // never true for the probe's data; keeps the loads
		sink[w] = acc;
}

} // namespace spmvhip

using namespace spmvhip;

// mode 0..13 (gather_cost_kernel), over the first n doubles of buf (n a power
// of two, at least 1024): 8 workgroups of 4 waves per CU, each wave `iters`
// loads; *cycles_per_load = kernel time x clock_ghz / loads per CU (after one
// warm-up launch). sink: at least (8 x CUs x 4) doubles, never written for
// the probe's data (buf filled with values in [1, 8]).
extern "C" int spmvhip_gather_probe(int mode, const double *buf, long n, int iters, double clock_ghz, double *sink,
		double *cycles_per_load)
{
	if (!buf || !sink || !cycles_per_load || n < 1024 || (n & (n - 1)) || n > (1L << 30) || iters < 1 ||
			mode < 0 || mode > 13 || !(clock_ghz > 0) || (mode >= 8 && n > (1L << 27)))
		return set_error(SPMVHIP_EINVAL, "gather_probe: bad arguments");
	int dev = 0, cus = 0;
	SPMVHIP_CHECK(hipGetDevice(&dev));
	SPMVHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
	const int grid = cus * 8;
	const uint32_t mask = (uint32_t) (n - 1);
	hipEvent_t e0, e1;
	SPMVHIP_CHECK(hipEventCreate(&e0));
	SPMVHIP_CHECK(hipEventCreate(&e1));
	float ms = 0;
	hipError_t he = hipSuccess;
	for (int rep = 0; rep < 2 && he == hipSuccess; rep++) {
		he = hipEventRecord(e0, nullptr);
		switch (mode) {
		case 0: gather_cost_kernel<0><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 1: gather_cost_kernel<1><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 2: gather_cost_kernel<2><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 3: gather_cost_kernel<3><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 4: gather_cost_kernel<4><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 5: gather_cost_kernel<5><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 6: gather_cost_kernel<6><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 7: gather_cost_kernel<7><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 8: gather_cost_kernel<8><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 9: gather_cost_kernel<9><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 10: gather_cost_kernel<10><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 11: gather_cost_kernel<11><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		case 12: gather_cost_kernel<12><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		default: gather_cost_kernel<13><<<grid, kBlock>>>(buf, mask, iters, sink); break;
		}
		if (he == hipSuccess)
			he = hipGetLastError();
		if (he == hipSuccess)
			he = hipEventRecord(e1, nullptr);
		if (he == hipSuccess)
			he = hipEventSynchronize(e1);
		if (he == hipSuccess)
			he = hipEventElapsedTime(&ms, e0, e1);
	}
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	if (he != hipSuccess)
		return set_error(SPMVHIP_EHIP, "gather_probe: %s", hipGetErrorString(he));
	const double loads_per_cu = (double) grid * kWavesPerBlock * iters / cus;
	*cycles_per_load = ms * 1e-3 * clock_ghz * 1e9 / loads_per_cu;
	return SPMVHIP_OK;
}

// mode 0: stream read of src; mode 1: copy src -> dst; mode 2: stream read
// through direct global-to-LDS loads. n doubles (even).
// ms[i] = time of iteration i (HIP events).
extern "C" int spmvhip_bandwidth_probe(int mode, const double *src, double *dst, long n, int iters, float *ms)
{
	if (!src || n < 2 || (n & 1) || iters < 1 || (mode == 1 && !dst) || mode < 0 || mode > 2)
		return set_error(SPMVHIP_EINVAL, "bandwidth_probe: bad arguments");
	int dev = 0;
	hipDeviceProp_t prop;
	SPMVHIP_CHECK(hipGetDevice(&dev));
	SPMVHIP_CHECK(hipGetDeviceProperties(&prop, dev));
	const int grid = prop.multiProcessorCount * 8;
	DevBuf partial;
	SPMVHIP_TRY(partial.alloc(grid * sizeof(double)));
	hipEvent_t e0, e1;
	SPMVHIP_CHECK(hipEventCreate(&e0));
	SPMVHIP_CHECK(hipEventCreate(&e1));
	int rc = SPMVHIP_OK;
	for (int it = -1; it < iters && rc == SPMVHIP_OK; it++) { // iteration -1 warms up
		hipError_t he = hipEventRecord(e0, nullptr);
		if (mode == 0)
			stream_read_kernel<<<grid, kBlock>>>((const dbl2 *) src, n / 2, partial.as<double>());
		else if (mode == 2)
			stream_read_lds_kernel<<<grid, kBlock>>>((const u32x4p *) src, n / 2, partial.as<double>());
		else
			stream_copy_kernel<<<grid, kBlock>>>((const dbl2 *) src, (dbl2 *) dst, n / 2);
		if (he == hipSuccess)
			he = hipGetLastError();
		if (he == hipSuccess)
			he = hipEventRecord(e1, nullptr);
		if (he == hipSuccess)
			he = hipEventSynchronize(e1);
		float t = 0;
		if (he == hipSuccess)
			he = hipEventElapsedTime(&t, e0, e1);
		if (he != hipSuccess)
			rc = set_error(SPMVHIP_EHIP, "bandwidth_probe: %s", hipGetErrorString(he));
		else if (it >= 0 && ms)
			ms[it] = t;
	}
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	return rc;
}

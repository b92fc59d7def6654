// Bandwidth reference kernels: the measured ceilings the SpMV kernels are
// compared with, and the known-size traffic that calibrates the hardware
// byte counters (FETCH_SIZE / WRITE_SIZE) of the profiler.
//   stream_read:  reads n doubles once (16-byte loads), writes one partial
//                 sum per workgroup
//   stream_copy:  dst[i] = src[i] over n doubles (16-byte loads and stores)

#include "format.hpp"

namespace spmvhip {

typedef double dbl2 __attribute__((ext_vector_type(2)));

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

} // namespace spmvhip

using namespace spmvhip;

// mode 0: stream read of src; mode 1: copy src -> dst. n doubles (even).
// ms[i] = time of iteration i (HIP events).
extern "C" int spmvhip_bandwidth_probe(int mode, const double *src, double *dst, long n, int iters, float *ms)
{
	if (!src || n < 2 || (n & 1) || iters < 1 || (mode == 1 && !dst) || mode < 0 || mode > 1)
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

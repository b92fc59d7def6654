// Shared definitions of the gfx950 SpMV library: status handling, launch
// geometry and wave64 helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "spmvhip.h"

namespace spmvhip {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // threads per workgroup (4 waves)
constexpr int kWavesPerBlock = kBlock / kWave;

// Records a message for spmvhip_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...);

// Status propagation: every HIP call is checked; a failure is turned into a
// SPMVHIP_EHIP status with the HIP error string.
#define SPMVHIP_CHECK(call)                                                         \
	do {                                                                        \
		hipError_t e_ = (call);                                             \
		if (e_ != hipSuccess)                                               \
			return ::spmvhip::set_error(SPMVHIP_EHIP, "%s:%d %s: %s", __FILE__, \
					__LINE__, #call, hipGetErrorString(e_));          \
	} while (0)

#define SPMVHIP_TRY(expr)                  \
	do {                               \
		int rc_ = (expr);          \
		if (rc_ != SPMVHIP_OK)     \
			return rc_;        \
	} while (0)

inline unsigned grid_for(long items, int per_block)
{
	return (unsigned) ((items + per_block - 1) / per_block);
}

// Sum over the 64 lanes of a wavefront; every lane receives the total.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off, kWave);
	return v;
}

// Sum over aligned groups of `width` lanes (width a power of two <= 64).
__device__ __forceinline__ double group_sum(double v, int width)
{
	for (int off = width >> 1; off > 0; off >>= 1)
		v += __shfl_xor(v, off, kWave);
	return v;
}

__device__ __forceinline__ int lane_id()
{
	return (int) (threadIdx.x & (kWave - 1));
}

// Owned device allocation (freed with the owning handle).
struct DevBuf {
	void *p = nullptr;
	size_t bytes = 0;
	DevBuf() = default;
	DevBuf(const DevBuf &) = delete;
	DevBuf &operator=(const DevBuf &) = delete;
	~DevBuf() { release(); }
	int alloc(size_t nbytes)
	{
		release();
		if (nbytes == 0)
			nbytes = 8;
		hipError_t e = hipMalloc(&p, nbytes);
		if (e != hipSuccess) {
			p = nullptr;
			return set_error(e == hipErrorOutOfMemory ? SPMVHIP_ENOMEM : SPMVHIP_EHIP,
					"hipMalloc(%zu): %s", nbytes, hipGetErrorString(e));
		}
		bytes = nbytes;
		return SPMVHIP_OK;
	}
	void release()
	{
		if (p)
			(void) hipFree(p);
		p = nullptr;
		bytes = 0;
	}
	template <class T> T *as() const { return static_cast<T *>(p); }
};

} // namespace spmvhip

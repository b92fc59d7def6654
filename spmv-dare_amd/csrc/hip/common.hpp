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
// zeroed entries after the last stored column index / value of the device
// CSR arrays (the paired stream loads read up to one entry past nnz)
constexpr int kCsrPad = 16;

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

// A device pointer kept in a table (not passed as a kernel argument) is
// stored as its integer value and materialised here directly in the global
// address space: accesses through a generic pointer loaded from memory would
// compile to flat instructions, which also count against the LDS wait
// counter and so serialise LDS work behind outstanding memory loads.
template <typename T>
__device__ __forceinline__ T *global_ptr(uint64_t bits)
{
	return (T *) (__attribute__((address_space(1))) T *) bits;
}

// The same for a read-only table, in the constant address space: a load at a
// wave-uniform address compiles to a scalar load (counted on the LDS/scalar
// wait counter), a divergent one to a global load. Through a global pointer a
// uniform load is a vector load whenever the compiler cannot prove that no
// store of the kernel reaches it (any persistent loop that writes y), and a
// vector load's wait also waits for every earlier store of the wave, so the
// next work unit's bounds would only arrive after the previous unit's y
// stores were acknowledged.
template <typename T>
__device__ __forceinline__ const T *const_ptr(uint64_t bits)
{
	return (const T *) (__attribute__((address_space(4))) const T *) bits;
}

__host__ __device__ inline uint64_t ptr_bits(const void *p)
{
	return (uint64_t) (uintptr_t) p;
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

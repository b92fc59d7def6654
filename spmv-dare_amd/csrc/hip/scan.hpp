// Device-wide scans used by the on-device csr_to_format conversions
// (slice pointers, row-block positions, segment pointers).
//
// Three phases per level: per-tile reduction, recursive scan of the tile
// totals, per-tile scan plus offset. A tile is 256 threads x 8 items.
#pragma once

#include <climits>
#include <limits>
#include "common.hpp"

namespace spmvhip {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

struct OpSum {
	template <class T> __device__ __host__ static T apply(T a, T b) { return a + b; }
	template <class T> __device__ __host__ static T identity() { return T(0); }
};
struct OpMax {
	template <class T> __device__ __host__ static T apply(T a, T b) { return a > b ? a : b; }
	template <class T> __device__ __host__ static T identity() { return std::numeric_limits<T>::lowest(); }
};

// Workgroup-wide scan of one value per thread: exclusive prefix and total.
template <class T, class Op>
__device__ void block_scan(T v, T &excl, T &total)
{
	__shared__ T wave_tot[kWavesPerBlock];
	const int lane = lane_id(), wave = threadIdx.x / kWave;
	T incl = v;
#pragma unroll
	for (int off = 1; off < kWave; off <<= 1) {
		T o = __shfl_up(incl, off, kWave);
		if (lane >= off)
			incl = Op::apply(o, incl);
	}
	T ex = __shfl_up(incl, 1, kWave);
	if (lane == 0)
		ex = Op::template identity<T>();
	if (lane == kWave - 1)
		wave_tot[wave] = incl;
	__syncthreads();
	T pre = Op::template identity<T>(), tot = Op::template identity<T>();
	for (int w = 0; w < kWavesPerBlock; w++) {
		if (w < wave)
			pre = Op::apply(pre, wave_tot[w]);
		tot = Op::apply(tot, wave_tot[w]);
	}
	__syncthreads();
	excl = Op::apply(pre, ex);
	total = tot;
}

template <class T, class Op>
__global__ void __launch_bounds__(kBlock) scan_tile_reduce(const T *__restrict__ in, long n, T *__restrict__ tile_tot)
{
	long base = (long) blockIdx.x * kScanTile + (long) threadIdx.x * kScanItems;
	T acc = Op::template identity<T>();
#pragma unroll
	for (int k = 0; k < kScanItems; k++)
		if (base + k < n)
			acc = Op::apply(acc, in[base + k]);
	T ex, tot;
	block_scan<T, Op>(acc, ex, tot);
	if (threadIdx.x == 0)
		tile_tot[blockIdx.x] = tot;
}

// Per-tile scan plus the tile's offset. Exclusive scans with write_total put
// the grand total in out[n].
template <class T, class Op>
__global__ void __launch_bounds__(kBlock) scan_tile_apply(const T *__restrict__ in, long n, const T *__restrict__ tile_off,
		T *__restrict__ out, int inclusive, int write_total)
{
	long base = (long) blockIdx.x * kScanTile + (long) threadIdx.x * kScanItems;
	T vals[kScanItems];
	T acc = Op::template identity<T>();
#pragma unroll
	for (int k = 0; k < kScanItems; k++) {
		vals[k] = (base + k < n) ? in[base + k] : Op::template identity<T>();
		acc = Op::apply(acc, vals[k]);
	}
	T ex, tot;
	block_scan<T, Op>(acc, ex, tot);
	T off = tile_off ? tile_off[blockIdx.x] : Op::template identity<T>();
	T cur = Op::apply(off, ex);
#pragma unroll
	for (int k = 0; k < kScanItems; k++) {
		if (base + k < n) {
			if (inclusive) {
				cur = Op::apply(cur, vals[k]);
				out[base + k] = cur;
			} else {
				out[base + k] = cur;
				cur = Op::apply(cur, vals[k]);
			}
		}
	}
	if (write_total && !inclusive && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
		out[n] = Op::apply(off, tot);
}

// Exclusive (inclusive = 0) or inclusive scan of in[0..n) into out; with an
// exclusive scan and write_total, out must hold n + 1 elements and out[n] is
// the total. Temporaries are allocated per call (conversion time only) and
// the call returns after the scan has completed on stream s.
template <class T, class Op>
int device_scan(const T *in, T *out, long n, int inclusive, int write_total, hipStream_t s)
{
	if (n <= 0) {
		if (write_total && !inclusive) {
			T z = Op::template identity<T>();
			SPMVHIP_CHECK(hipMemcpyAsync(out, &z, sizeof(T), hipMemcpyHostToDevice, s));
			SPMVHIP_CHECK(hipStreamSynchronize(s));
		}
		return SPMVHIP_OK;
	}
	long tiles = (n + kScanTile - 1) / kScanTile;
	if (tiles == 1) {
		scan_tile_apply<T, Op><<<1, kBlock, 0, s>>>(in, n, nullptr, out, inclusive, write_total);
		SPMVHIP_CHECK(hipGetLastError());
		SPMVHIP_CHECK(hipStreamSynchronize(s));
		return SPMVHIP_OK;
	}
	DevBuf tot, off;
	SPMVHIP_TRY(tot.alloc(tiles * sizeof(T)));
	SPMVHIP_TRY(off.alloc(tiles * sizeof(T)));
	scan_tile_reduce<T, Op><<<(unsigned) tiles, kBlock, 0, s>>>(in, n, tot.as<T>());
	SPMVHIP_CHECK(hipGetLastError());
	SPMVHIP_TRY((device_scan<T, Op>(tot.as<T>(), off.as<T>(), tiles, 0, 0, s)));
	scan_tile_apply<T, Op><<<(unsigned) tiles, kBlock, 0, s>>>(in, n, off.as<T>(), out, inclusive, write_total);
	SPMVHIP_CHECK(hipGetLastError());
	// the temporaries are freed at scope exit: wait until the kernels are done
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	return SPMVHIP_OK;
}

} // namespace spmvhip

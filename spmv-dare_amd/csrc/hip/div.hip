// DIV-RF on gfx950: host encoding (csrc/div/div_codec.cpp, the reference's
// packet layout) and a device decoder that computes y = A x straight from the
// packet stream (the CPU kernel it replaces: decompress_and_compute_kernel_div,
// src/spmv_kernels/DIV/div_kernels_rf.h:948-1162, driven by compute_div,
// src/spmv_kernels/DIV/div.cpp:1201-1316).
//
// Device layout: the packets of the reference stream, each starting at a
// 16-byte aligned offset (the bytes of a packet are unchanged), plus
//   pk_off[p]   byte offset of packet p       pk_nnz[p]  its first nnz index
//   grp_pk[g]   first packet of packet group g (a group starts at a row start)
//   grp_row[g]  first row owned by group g: group g writes every y row in
//               [grp_row[g], grp_row[g+1]), empty rows included, so no
//               separate zeroing pass is needed.
//
// Decoder: one wavefront per packet group. The L <= 64 value streams of a
// packet map to lanes 0..L-1; a wave-level exclusive scan of the header's
// lane sizes gives each lane the start of its stream. Per replication-factor
// group a lane walks its values in order (the reference's interleave puts
// value p of lane s at position p*L + s), adds each decoded delta
// (bits << trailing_zeros) to its running value, and for each of the rf
// coordinates of that value (element-interleaved, so the lanes read them
// coalesced) accumulates value * x[col] into an LDS copy of the packet's
// <= 255 rows. At the end of the packet the LDS rows are added to y.
//
// Three kernels implement this: div_spmv_lds_kernel (the packet copied to
// LDS with coalesced 16-byte loads, the decode reading LDS only; used when
// the largest packet fits), div_spmv_kernel (the same decode reading the
// stream from global memory, for larger packets) and div_export_kernel, the
// byte-at-a-time decode that exports the decoded (row, col,
// value) sequence for the parity checks.

#include "format.hpp"
#include "kernels.hpp"
#include "../div/div_codec.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace spmvhip {

constexpr int kDivWavesPerBlock = kWavesPerBlock;
constexpr long kDivLdsMax = 65536; // LDS of the staged decoder without raising the per-kernel limit

__device__ __forceinline__ uint64_t load_le(const uint8_t *p, int nbytes)
{
	uint64_t v = 0;
	for (int b = 0; b < nbytes; b++)
		v |= (uint64_t) p[b] << (8 * b);
	return v;
}

// Byte b of lane `lane`'s value stream (the byte-at-a-time decode of the
// export kernels): interleaved (row bytes lrow) or contiguous from sbase.
__device__ __forceinline__ uint64_t stream_le(const uint8_t *pk, long sbase, int lane, long b, int nbytes, bool il,
		int lrow)
{
	uint64_t v = 0;
	for (int k = 0; k < nbytes; k++) {
		const long q = b + k;
		const uint8_t byte = il ? pk[sbase + (q >> 2) * lrow + 4 * lane + (q & 3)] : pk[sbase + q];
		v |= (uint64_t) byte << (8 * k);
	}
	return v;
}

__device__ __forceinline__ void wave_fence()
{
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
	__builtin_amdgcn_wave_barrier();
}

// Export of the decoded (row, col, value) sequence in packet order
// (rows/cols/vals at pk_nnz + index): the decode stated byte by byte, for the
// parity checks of the device decoder.
__global__ void __launch_bounds__(kBlock) div_export_kernel(const uint8_t *__restrict__ bytes,
		const int64_t *__restrict__ pk_off, const int64_t *__restrict__ pk_nnz, const int32_t *__restrict__ grp_pk,
		long ngroups, int L, int il, int32_t *__restrict__ rows_out, int32_t *__restrict__ cols_out,
		double *__restrict__ vals_out)
{
	const int lane = lane_id(), wv = threadIdx.x / kWave;
	const long g = (long) blockIdx.x * kDivWavesPerBlock + wv;
	if (g >= ngroups)
		return; // waves do not synchronise with each other
	for (int p = grp_pk[g]; p < grp_pk[g + 1]; p++) {
		const uint8_t *pk = bytes + pk_off[p];
		const uint32_t *h = (const uint32_t *) pk;
		const uint32_t nv = h[0], nu = h[1], row_min = h[3], col_min = h[4];
		const int row_bits = pk[20], col_bits = pk[21];
		const uint32_t nrfs = h[6];
		const int cb = (row_bits + col_bits + 7) >> 3;
		const uint64_t row_mask = row_bits ? ((1ull << row_bits) - 1) : 0;
		// start of this lane's value stream: wave exclusive scan of the lane sizes
		long my_size = lane < L ? (long) h[7 + lane] : 0;
		long incl = my_size;
		for (int off = 1; off < kWave; off <<= 1) {
			long o = __shfl_up(incl, off, kWave);
			if (lane >= off)
				incl += o;
		}
		const long lane_base = incl - my_size;
		const uint32_t *rfs = h + 7 + L;
		const uint32_t *per_rf = rfs + nrfs;
		const uint8_t *coords = pk + 28 + 4 * L + 8 * (long) nrfs;
		const uint8_t *lens = coords + (long) nv * cb;
		const long sbase = il ? ((lens - pk) + nu + 3) & ~3L : (lens - pk) + nu + lane_base;
		const long nnz0 = pk_nnz[p];
		long cursor = 0, ks = 0, cs = 0;
		for (uint32_t gi = 0; gi < nrfs; gi++) {
			const long rf = rfs[gi], n = per_rf[gi];
			const long dv = n / L, md = n % L, mul_c = dv * rf * L;
			if (lane < L) {
				uint64_t val = 0;
				const long cnt = dv + (lane < md ? 1 : 0);
				for (long pp = 0; pp < cnt; pp++) {
					const uint8_t lb = lens[ks + pp * L + lane];
					const int len = lb & 15, tz = (lb >> 4) << 2;
					if (len) {
						val += stream_le(pk, sbase, lane, cursor, len, il, 4 * L) << tz;
						cursor += len;
					}
					double d;
					memcpy(&d, &val, 8);
					for (long t = 0; t < rf; t++) {
						const long cpos = pp < dv ? (pp * rf + t) * L + lane : mul_c + lane * rf + t;
						const uint64_t c = load_le(coords + (cs + cpos) * cb, cb);
						const long o = nnz0 + cs + cpos;
						rows_out[o] = (int32_t) (row_min + (uint32_t) (c & row_mask));
						cols_out[o] = (int32_t) (col_min + (uint32_t) (c >> row_bits));
						vals_out[o] = d;
					}
				}
			}
			ks += n;
			cs += n * rf;
		}
	}
}

// nbytes (0..8) little-endian bytes at byte offset pos of an 8-byte aligned
// base: two aligned 8-byte loads and a funnel shift instead of a byte loop.
// The device stream carries 16 bytes of padding, so base + pos + 16 stays in
// the buffer.
__device__ __forceinline__ uint64_t load_le_fast(const uint8_t *__restrict__ base, long pos, int nbytes)
{
	const uint64_t *w = (const uint64_t *) (base + (pos & ~7L));
	const int sh = (int) (pos & 7) * 8;
	const uint64_t w0 = w[0], w1 = w[1];
	const uint64_t v = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;
	return nbytes >= 8 ? v : v & ((1ull << (8 * nbytes)) - 1);
}

// y = A x from the packet stream, the decode restructured for latency: each
// lane takes its values in chunks of K; the K length bytes are loaded
// together (coalesced across lanes), the K stream offsets follow from a
// running sum of the lengths in registers, then the K deltas and their
// coordinates are requested together and the value prefix sum runs in
// registers. Products go to the packet's rows in LDS (atomic adds: the
// lanes of one instruction that hit the same row are applied by the LDS in a
// fixed hardware order, not the sequential order of csr_naive, so y agrees
// with csr_naive to rounding and repeated launches give the same bits),
// which are added to y at the end of the packet.
template <int K>
__global__ void __launch_bounds__(kBlock) div_spmv_kernel(const uint8_t *__restrict__ bytes,
		const int64_t *__restrict__ pk_off, const int32_t *__restrict__ grp_pk, const int32_t *__restrict__ grp_row,
		long ngroups, const int32_t *__restrict__ grp_order, int L, const double *__restrict__ x,
		double *__restrict__ y)
{
	__shared__ double yl[kDivWavesPerBlock][256];
	const int lane = lane_id(), wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
	const long k = (long) blockIdx.x * kDivWavesPerBlock + wv;
	if (k >= ngroups)
		return; // waves do not synchronise with each other
	const long g = grp_order ? grp_order[k] : k;
	for (long r = grp_row[g] + lane; r < grp_row[g + 1]; r += kWave)
		y[r] = 0.0;
	wave_fence();
	double *yw = yl[wv];
	for (int p = grp_pk[g]; p < grp_pk[g + 1]; p++) {
		const uint8_t *pk = bytes + pk_off[p];
		const uint32_t *h = (const uint32_t *) pk;
		const uint32_t nv = h[0], nu = h[1], nrows = h[2], row_min = h[3], col_min = h[4];
		const int row_bits = pk[20], col_bits = pk[21];
		const uint32_t nrfs = h[6];
		const int cb = (row_bits + col_bits + 7) >> 3;
		const uint64_t row_mask = row_bits ? ((1ull << row_bits) - 1) : 0;
		long my_size = lane < L ? (long) h[7 + lane] : 0;
		long incl = my_size;
		for (int off = 1; off < kWave; off <<= 1) {
			long o = __shfl_up(incl, off, kWave);
			if (lane >= off)
				incl += o;
		}
		const uint32_t *rfs = h + 7 + L;
		const uint32_t *per_rf = rfs + nrfs;
		const long coords_at = 28 + 4 * (long) L + 8 * (long) nrfs; // byte offsets in the packet
		const long lens_at = coords_at + (long) nv * cb;
		const long stream_at = lens_at + nu + (incl - my_size); // this lane's value stream
		for (uint32_t r = lane; r < nrows; r += kWave)
			yw[r] = 0.0;
		wave_fence();
		long cursor = 0, ks = 0, cs = 0;
		for (uint32_t gi = 0; gi < nrfs; gi++) {
			const long rf = rfs[gi], n = per_rf[gi];
			const long dv = n / L, md = n % L, mul_c = dv * rf * L;
			const long cnt = lane < L ? dv + (lane < md ? 1 : 0) : 0;
			const long steps = dv + (md ? 1 : 0);
			uint64_t val = 0;
			for (long p0 = 0; p0 < steps; p0 += K) {
				int len[K], tz[K];
				long off[K];
#pragma unroll
				for (int k = 0; k < K; k++) {
					const long pp = p0 + k;
					const uint8_t lb = pp < cnt ? pk[lens_at + ks + pp * L + lane] : 0;
					len[k] = lb & 15;
					tz[k] = (lb >> 4) << 2;
				}
#pragma unroll
				for (int k = 0; k < K; k++) {
					off[k] = cursor;
					cursor += len[k];
				}
				uint64_t dlt[K];
#pragma unroll
				for (int k = 0; k < K; k++)
					dlt[k] = len[k] ? load_le_fast(pk, stream_at + off[k], len[k]) << tz[k] : 0;
				if (rf == 1) {
					uint64_t c[K];
#pragma unroll
					for (int k = 0; k < K; k++) {
						const long pp = p0 + k;
						c[k] = pp < cnt ? load_le_fast(pk, coords_at + (cs + pp * L + lane) * cb, cb) : 0;
					}
					double xv[K];
#pragma unroll
					for (int k = 0; k < K; k++)
						xv[k] = x[col_min + (uint32_t) (c[k] >> row_bits)];
#pragma unroll
					for (int k = 0; k < K; k++) {
						val += dlt[k];
						if (p0 + k < cnt) {
							double d;
							__builtin_memcpy(&d, &val, 8);
							atomicAdd(&yw[(uint32_t) (c[k] & row_mask)], d * xv[k]);
						}
					}
				} else {
#pragma unroll
					for (int k = 0; k < K; k++) {
						const long pp = p0 + k;
						val += dlt[k];
						if (pp >= cnt)
							continue;
						double d;
						__builtin_memcpy(&d, &val, 8);
						for (long t = 0; t < rf; t++) {
							const long cpos = pp < dv ? (pp * rf + t) * L + lane : mul_c + lane * rf + t;
							const uint64_t c = load_le_fast(pk, coords_at + (cs + cpos) * cb, cb);
							atomicAdd(&yw[(uint32_t) (c & row_mask)], d * x[col_min + (uint32_t) (c >> row_bits)]);
						}
					}
				}
			}
			ks += n;
			cs += n * rf;
		}
		wave_fence();
		for (uint32_t r = lane; r < nrows; r += kWave)
			y[row_min + r] += yw[r];
		wave_fence();
	}
}

// The decoder with every packet staged in LDS (one wavefront per workgroup,
// dynamic LDS = the largest packet rounded up + 16 bytes of padding, plus
// row accumulators for the largest packet's rows, so that as many
// workgroups fit per CU as the packets allow). A packet is copied to LDS by
// direct global-to-LDS 16-byte loads, all requested at once; the other
// resident waves hide the latency. All stream, length and coordinate reads
// of the decode are LDS reads; only the x gathers go to global memory. The
// arithmetic is div_spmv_kernel's.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t lds_le(const uint8_t *base, long pos, int nbytes)
{
	const uint64_t *w = (const uint64_t *) (base + (pos & ~7L));
	const int sh = (int) (pos & 7) * 8;
	const uint64_t w0 = w[0], w1 = w[1];
	const uint64_t v = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;
	return nbytes >= 8 ? v : v & ((1ull << (8 * nbytes)) - 1);
}

// The same read from three aligned dwords and two byte-aligns (v_alignbyte
// takes the 8 bytes {hi, lo} shifted right by a byte count): fewer
// instructions than 64-bit shifts. pos is a packet offset below 2^31.
__device__ __forceinline__ uint64_t lds_le_b32(const uint8_t *base, int pos, int nbytes)
{
	const uint32_t *w = (const uint32_t *) (base + (pos & ~3));
	const uint32_t sh = (uint32_t) (pos & 3);
	const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
	uint32_t lo = __builtin_amdgcn_alignbyte(w1, w0, sh);
	uint32_t hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
	lo &= nbytes >= 4 ? 0xffffffffu : (1u << (8 * nbytes)) - 1;
	hi &= nbytes >= 8 ? 0xffffffffu : nbytes > 4 ? (1u << (8 * (nbytes - 4))) - 1 : 0u;
	return ((uint64_t) hi << 32) | lo;
}

// At most 4 bytes (a coordinate of a packet with row_bits + col_bits <= 32).
__device__ __forceinline__ uint32_t lds_le_b32x(const uint8_t *base, int pos, uint32_t mask)
{
	const uint32_t *w = (const uint32_t *) (base + (pos & ~3));
	return __builtin_amdgcn_alignbyte(w[1], w[0], (uint32_t) (pos & 3)) & mask;
}

// A lane's next delta: nbytes (0..8) at byte `cursor` of its value stream,
// which starts at sbase. IL: the streams are interleaved by dwords (lane l's
// dword w at sbase + w * lrow, sbase = stream section + 4 l, lrow = 4 L;
// LROW is lrow when known at compile time), so the lanes' three dword reads
// fall in distinct banks; the reads past a stream's end stay inside the
// packet buffer's two rows of slack. Otherwise the stream is contiguous.
//
// The low nbytes bytes are kept with one 64-bit mask, ~((~0 << 4n) << 4n)
// (two shifts of at most 32, so n = 0 and n = 8 need no special case):
// about half the instructions of separate 32-bit masks.
template <bool IL, int LROW>
__device__ __forceinline__ uint64_t stream_read(const uint8_t *pk, int sbase, int cursor, int nbytes, int lrow)
{
	uint32_t lo, hi;
	if (IL) {
		const int row = LROW ? LROW : lrow;
		const uint32_t *w = (const uint32_t *) (pk + sbase + (cursor >> 2) * row);
		const uint32_t sh = (uint32_t) (cursor & 3);
		const uint32_t w0 = w[0], w1 = w[row / 4], w2 = w[row / 2];
		lo = __builtin_amdgcn_alignbyte(w1, w0, sh);
		hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
	} else {
		const int pos = sbase + cursor;
		const uint32_t *w = (const uint32_t *) (pk + (pos & ~3));
		const uint32_t sh = (uint32_t) (pos & 3);
		const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
		lo = __builtin_amdgcn_alignbyte(w1, w0, sh);
		hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
	}
	const uint64_t v = ((uint64_t) hi << 32) | lo;
	const uint32_t s4 = 4u * (uint32_t) nbytes;
	return v & ~((~0ull << s4) << s4);
}


// One replication-factor-1 group of a packet staged in LDS (the common case:
// every value distinct). C4: the coordinates fit 4 bytes and are read with
// one byte-align.
//
// A step of K values runs in two passes: the first decodes the K columns
// and requests all K gathers of x (one memory round trip per step), the
// second reads the K length bytes and deltas and accumulates. Full steps
// (div_full_steps) keep the K rows in registers between the passes, and a
// lane with exactly two full steps requests both steps' gathers at once.
// The steps past the full ones test each value: their loads are
// unconditional (a value past the lane's last reads the packet's first
// coordinate and length byte, positions inside the packet, gathers
// x[col_min], a valid column, and decodes a zero delta), only the LDS
// accumulation is skipped, and the second pass re-reads the coordinate from
// LDS (cheap) so that only the K gathered values stay in registers.
template <int K, bool C4>
__device__ __forceinline__ uint32_t div_coord(const uint8_t *pk, int at, int cb, uint32_t cmask, int row_bits,
		uint32_t &row)
{
	const uint32_t rmask = (1u << row_bits) - 1; // row_bits <= 8 (at most 255 rows per packet)
	if (C4) {
		const uint32_t c = lds_le_b32x(pk, at, cmask);
		row = c & rmask;
		return c >> row_bits;
	}
	const uint64_t c = lds_le_b32(pk, at, cb);
	row = (uint32_t) c & rmask;
	return (uint32_t) (c >> row_bits);
}

// S full steps of K values of one lane (every lane below L has them all):
// first the S * K coordinates and their x gathers, then per step its K
// length bytes, its K deltas and its K accumulations. The step's length
// bytes, then its deltas, are read before its first accumulation (the
// scheduling barriers keep them together): the compiler does not move loads
// across the LDS atomics, so reads between the atomics would put one LDS
// round trip per value on the critical path.
template <int K, int S, bool C4, bool A4L64, bool IL>
__device__ __forceinline__ void div_full_steps(const uint8_t *__restrict__ pk, int &at, int &la, int cstride,
		int lstride, int cb, uint32_t cmask, int row_bits, uint32_t col_min, int stream_at,
		const double *__restrict__ x, double *__restrict__ yw, int &cursor, uint64_t &val)
{
	uint32_t row[S * K];
	double xv[S * K];
#pragma unroll
	for (int k = 0; k < S * K; k++) {
		uint32_t col;
		if (A4L64) {
			const uint32_t c = *(const uint32_t *) (pk + at + k * cstride);
			row[k] = c & ((1u << row_bits) - 1);
			col = c >> row_bits;
		} else {
			col = div_coord<K, C4>(pk, at + k * cstride, cb, cmask, row_bits, row[k]);
		}
		xv[k] = x[col_min + col];
	}
	at += S * K * cstride;
#pragma unroll
	for (int s = 0; s < S; s++) {
		int lb[K];
#pragma unroll
		for (int k = 0; k < K; k++)
			lb[k] = pk[la + k * lstride];
		__builtin_amdgcn_sched_barrier(0); // all length bytes requested together
		uint64_t dlt[K];
#pragma unroll
		for (int k = 0; k < K; k++) {
			const int len = lb[k] & 15, tz = (lb[k] >> 4) << 2;
			dlt[k] = stream_read<IL, A4L64 ? 4 * kWave : 0>(pk, stream_at, cursor, len, 4 * lstride) << tz;
			cursor += len;
		}
		__builtin_amdgcn_sched_barrier(0); // then all deltas, before the first accumulation
#pragma unroll
		for (int k = 0; k < K; k++) {
			val += dlt[k];
			double d;
			__builtin_memcpy(&d, &val, 8);
			atomicAdd(&yw[row[s * K + k]], d * xv[s * K + k]);
		}
		la += K * lstride;
	}
}

//
// A4L64: the packet's coordinates are exactly 4 bytes (so, from the header's
// 4-byte alignment, every coordinate is dword-aligned) and it has 64 lanes.
// The full steps then read each coordinate with one aligned LDS dword read
// and each length byte at a compile-time stride, with immediate offsets.
template <int K, bool C4, bool A4L64, bool IL>
__device__ __forceinline__ void div_rf1_group(const uint8_t *__restrict__ pk, int coords_at, int lens_at, int lens0,
		int crd0_at, int stream_at, int steps, int cnt, int dv, int L, int cb, uint32_t col_min, int row_bits,
		const double *__restrict__ x, double *__restrict__ yw, int &cursor)
{
	const uint32_t cmask = cb >= 4 ? 0xffffffffu : (1u << (8 * cb)) - 1;
	const int cstride = A4L64 ? kWave * 4 : L * cb;
	const int lstride = A4L64 ? kWave : L;
	uint64_t val = 0;
	// full steps: every lane below L has at least dv values, so the first
	// dv / K * K of them need no validity test; addresses advance by
	// constant strides and the K rows stay in registers between the passes
	const int nfull = dv / K * K;
	if (nfull > 0 && lane_id() < L) {
		int at = crd0_at, la = lens0, p0 = 0;
		if constexpr (K <= 8) {
			// exactly two full steps (packets of 1024 values in 64 lanes at
			// K = 8): both steps' gathers are requested before either is
			// decoded, so the second step's gathers wait behind no decode
			if (nfull == 2 * K) {
				div_full_steps<K, 2, C4, A4L64, IL>(pk, at, la, cstride, lstride, cb, cmask, row_bits, col_min,
						stream_at, x, yw, cursor, val);
				p0 = nfull;
			}
		}
		for (; p0 < nfull; p0 += K)
			div_full_steps<K, 1, C4, A4L64, IL>(pk, at, la, cstride, lstride, cb, cmask, row_bits, col_min,
					stream_at, x, yw, cursor, val);
	}
	// the rest (at most K + 1 values per lane), each value tested
	for (int p0 = nfull; p0 < steps; p0 += K) {
		double xv[K];
#pragma unroll
		for (int k = 0; k < K; k++) {
			const bool ok = p0 + k < cnt;
			uint32_t row;
			const uint32_t col = div_coord<K, C4>(pk, ok ? crd0_at + (p0 + k) * cstride : coords_at, cb, cmask,
					row_bits, row);
			xv[k] = x[col_min + (ok ? col : 0u)];
		}
#pragma unroll
		for (int k = 0; k < K; k++) {
			const bool ok = p0 + k < cnt;
			uint32_t row;
			(void) div_coord<K, C4>(pk, ok ? crd0_at + (p0 + k) * cstride : coords_at, cb, cmask, row_bits, row);
			const int lb = pk[ok ? lens0 + (p0 + k) * L : lens_at];
			const int len = ok ? lb & 15 : 0, tz = (lb >> 4) << 2;
			val += stream_read<IL, 0>(pk, stream_at, cursor, len, 4 * L) << tz;
			cursor += len;
			double d;
			__builtin_memcpy(&d, &val, 8);
			if (ok) // (an unconditional add of 0.0 would pile every idle lane onto one LDS address)
				atomicAdd(&yw[row], d * xv[k]);
		}
	}
}

// The automatic device layout's order of preference (div_layout codes).
constexpr int kDivAutoLayouts[] = {3, 1};
constexpr int kDivAccCopiesMax = 64;    // copies of the LDS row accumulators (a power of two)
// The automatic choice: up to 4 copies within 4 KiB. Measured on the 8M-row
// matrix (packets of at most 48 rows): 1 copy 0.89 ms, 4 copies 0.747 ms,
// 8 copies 0.767 ms, 16 copies 0.89 ms, 32 copies 1.29 ms (more LDS per
// workgroup, fewer resident waves, more zeroing and summing).
constexpr int kDivAccCopiesAuto = 4;
constexpr long kDivAccBudget = 4096;

// row r summed over the nc accumulator copies, in a fixed order
__device__ __forceinline__ double div_acc_sum(const double *yw, int stride, int nc, int r)
{
	double s = yw[r];
	for (int c = 1; c < nc; c++)
		s += yw[c * stride + r];
	return s;
}

// Packet group g of a matrix, decoded by the calling wavefront with its
// packets staged in the LDS at smem (the standalone kernel below and the
// batched launch's DIV kernel).
template <int K, bool IL>
__device__ __forceinline__ void div_group_lds(const uint8_t *__restrict__ bytes, const int64_t *__restrict__ pk_off,
		const int32_t *__restrict__ grp_pk, const int32_t *__restrict__ grp_row, long g, int L, int buf_bytes,
		int acc_stride, int acc_copies, const double *__restrict__ x, double *__restrict__ y, u32x4 *smem)
{
	uint8_t *pk = (uint8_t *) smem;                   // buf_bytes (a multiple of 16, padding included)
	double *yw = (double *) (pk + buf_bytes);          // acc_copies copies of the row accumulators
	const int lane = lane_id();
	// lane l adds into copy l % acc_copies: lanes whose values fall in the
	// same row collide on one LDS address (and serialise) acc_copies times
	// less often -- with one copy those collisions bound the decoder
	double *ywl = yw + (lane & (acc_copies - 1)) * acc_stride;
	const int p_end = grp_pk[g + 1];
	int p = grp_pk[g];
	// a group of one packet (the common case) stores its rows once at the
	// end; a longer group zeroes them first and adds each packet's rows
	const bool single = p_end - p == 1;
	const long gr0 = grp_row[g], gr1 = grp_row[g + 1];
	if (!single)
		for (long r = gr0 + lane; r < gr1; r += kWave)
			y[r] = 0.0;
	for (; p < p_end; p++) {
		// copy the packet to LDS with direct global-to-LDS 16-byte loads (lane i
		// of a round writes bytes [16 i, 16 i + 16) of the round's 1 KiB; lanes
		// past the packet are masked off, so no write leaves the packet
		// buffer), all rounds in flight at once, no VGPR staging. The loads are
		// non-temporal (cache policy 2 = nt): packets are read once, and the L2
		// is left to the x lines the gathers reuse
		const u32x4 *src = (const u32x4 *) (bytes + pk_off[p]);
		const int n16 = (int) ((pk_off[p + 1] - pk_off[p] + 15) >> 4);
		for (int j0 = 0; j0 < n16; j0 += kWave) {
			if (j0 + lane < n16)
				__builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *) (src + j0 + lane),
						(__attribute__((address_space(3))) void *) (smem + j0), 16, 0, 2);
		}
		__builtin_amdgcn_s_waitcnt(0); // the copies have landed in LDS
		if (lane == 0) // zero padding after the packet: the decode reads whole words
			smem[n16] = u32x4{0, 0, 0, 0};
		wave_fence();
		const uint32_t *h = (const uint32_t *) pk;
		const uint32_t nv = h[0], nu = h[1], nrows = h[2], row_min = h[3], col_min = h[4];
		const int row_bits = pk[20], col_bits = pk[21];
		const uint32_t nrfs = h[6];
		const int cb = (row_bits + col_bits + 7) >> 3;
		const uint64_t row_mask = row_bits ? ((1ull << row_bits) - 1) : 0;
		const uint32_t *rfs = h + 7 + L;
		const uint32_t *per_rf = rfs + nrfs;
		const long coords_at = 28 + 4 * (long) L + 8 * (long) nrfs;
		const long lens_at = coords_at + (long) nv * cb;
		long stream_at;
		if (IL) {
			stream_at = ((lens_at + nu + 3) & ~3L) + 4 * lane; // the lane's column of the interleaved streams
		} else {
			// the lane's stream start: exclusive scan of the lane sizes
			long my_size = lane < L ? (long) h[7 + lane] : 0;
			long incl = my_size;
			for (int off = 1; off < kWave; off <<= 1) {
				long o = __shfl_up(incl, off, kWave);
				if (lane >= off)
					incl += o;
			}
			stream_at = lens_at + nu + (incl - my_size);
		}
		for (int r = lane; r < acc_copies * acc_stride; r += kWave)
			yw[r] = 0.0;
		wave_fence();
		long cursor = 0, ks = 0, cs = 0;
		for (uint32_t gi = 0; gi < nrfs; gi++) {
			const long rf = rfs[gi], n = per_rf[gi];
			const long dv = n / L, md = n % L, mul_c = dv * rf * L;
			const long cnt = lane < L ? dv + (lane < md ? 1 : 0) : 0;
			const long steps = dv + (md ? 1 : 0);
			uint64_t val = 0;
			if (rf == 1) {
				int cur = (int) cursor;
				const int lens0 = (int) (lens_at + ks) + lane, crd0 = (int) cs + lane;
				if (cb == 4 && L == kWave)
					div_rf1_group<K, true, true, IL>(pk, (int) coords_at, (int) lens_at, lens0,
							(int) coords_at + crd0 * cb, (int) stream_at, (int) steps, (int) cnt, (int) dv, L, cb,
							col_min, row_bits, x, ywl, cur);
				else if (cb <= 4)
					div_rf1_group<K, true, false, IL>(pk, (int) coords_at, (int) lens_at, lens0,
							(int) coords_at + crd0 * cb, (int) stream_at, (int) steps, (int) cnt, (int) dv, L, cb,
							col_min, row_bits, x, ywl, cur);
				else
					div_rf1_group<K, false, false, IL>(pk, (int) coords_at, (int) lens_at, lens0,
							(int) coords_at + crd0 * cb, (int) stream_at, (int) steps, (int) cnt, (int) dv, L, cb,
							col_min, row_bits, x, ywl, cur);
				cursor = cur;
				ks += n;
				cs += n;
				continue;
			}
			for (long p0 = 0; p0 < steps; p0 += K) {
#pragma unroll
				for (int k = 0; k < K; k++) {
					const long pp = p0 + k;
					if (pp >= cnt)
						continue;
					const uint8_t lb = pk[lens_at + ks + pp * L + lane];
					const int len = lb & 15, tz = (lb >> 4) << 2;
					if (len) {
						val += stream_read<IL, 0>(pk, (int) stream_at, (int) cursor, len, 4 * L) << tz;
						cursor += len;
					}
					double d;
					__builtin_memcpy(&d, &val, 8);
					{
						for (long t = 0; t < rf; t++) {
							const long cpos = pp < dv ? (pp * rf + t) * L + lane : mul_c + lane * rf + t;
							const uint64_t cc = lds_le(pk, coords_at + (cs + cpos) * cb, cb);
							atomicAdd(&ywl[(uint32_t) (cc & row_mask)],
									d * x[col_min + (uint32_t) (cc >> row_bits)]);
						}
					}
				}
			}
			ks += n;
			cs += n * rf;
		}
		wave_fence();
		if (single) {
			// the group's rows outside the packet are empty rows
			for (long r = gr0 + lane; r < gr1; r += kWave) {
				const long i = r - (long) row_min;
				y[r] = i >= 0 && i < (long) nrows ? div_acc_sum(yw, acc_stride, acc_copies, (int) i) : 0.0;
			}
		} else {
			for (uint32_t r = lane; r < nrows; r += kWave)
				y[row_min + r] += div_acc_sum(yw, acc_stride, acc_copies, (int) r);
		}
		wave_fence(); // the packet buffer and accumulators are reused
	}
}

template <int K, bool IL>
__global__ void __launch_bounds__(kWave) div_spmv_lds_kernel(const uint8_t *__restrict__ bytes,
		const int64_t *__restrict__ pk_off, const int32_t *__restrict__ grp_pk, const int32_t *__restrict__ grp_row,
		long ngroups, const int32_t *__restrict__ grp_order, int L, int buf_bytes, int acc_stride, int acc_copies,
		int xcd_k, const double *__restrict__ x, double *__restrict__ y)
{
	extern __shared__ u32x4 smem[];
	const long k = xcd_logical_block(xcd_k); // groups dealt to XCDs in chunks of xcd_k (0: in order)
	if (k >= ngroups)
		return;
	const long g = grp_order ? grp_order[k] : k;
	div_group_lds<K, IL>(bytes, pk_off, grp_pk, grp_row, g, L, buf_bytes, acc_stride, acc_copies, x, y, smem);
}

// ------------------------------------------------------------ x-slot layout
//
// The decoder of the x-slot device layout (div_codec.hpp: to_slot_layout).
// Per packet: (1) the packet is copied to LDS as above; (2) x is staged:
// slot s of the packet (its s-th smallest distinct column) gets xs[s] =
// x[column], lane l of a 64-slot chunk taking slot 64 c + l, so the loads of
// a chunk follow the packet's runs of consecutive columns (a few cache lines
// per load instead of one per lane); (3) the values are decoded as in
// div_spmv_lds_kernel, each coordinate giving a row and a slot, x read from
// xs in LDS.
//
// The run of slot s: the bit set of run starts (one u64 word per chunk) is
// counted up to s. Runs started before chunk c and the start of the run open
// at the chunk's first slot are carried from chunk to chunk in scalar
// registers.

struct SlotHeader {
	uint32_t nv, nu, nrows, row_min, nslots, nrfs, nruns;
	int rb, cb;
	int lanes_at, rfs_at, runs_at, bits_at, coords_at, lens_at, streams_at;
};

// The header of a staged x-slot packet (wave-uniform values, scalar registers).
__device__ __forceinline__ SlotHeader slot_header(const uint8_t *pk, int L)
{
	const uint32_t *h = (const uint32_t *) pk;
	SlotHeader z;
	z.nv = __builtin_amdgcn_readfirstlane(h[0]);
	z.nu = __builtin_amdgcn_readfirstlane(h[1]);
	z.nrows = __builtin_amdgcn_readfirstlane(h[2]);
	z.row_min = __builtin_amdgcn_readfirstlane(h[3]);
	z.nslots = __builtin_amdgcn_readfirstlane(h[4]);
	const uint32_t b = __builtin_amdgcn_readfirstlane(h[5]);
	z.rb = (int) (b & 0xff);
	z.cb = (int) ((b >> 16) & 0xff);
	z.nrfs = __builtin_amdgcn_readfirstlane(h[6]);
	z.nruns = __builtin_amdgcn_readfirstlane(h[7]);
	z.lanes_at = spmvdiv::kSlotHeaderFixed;
	z.rfs_at = spmvdiv::kSlotHeaderFixed + 4 * L;
	z.runs_at = z.rfs_at + 8 * (int) z.nrfs;
	z.bits_at = (z.runs_at + 4 * (int) z.nruns + 7) & ~7;
	z.coords_at = z.bits_at + 8 * (((int) z.nslots + 63) >> 6);
	z.lens_at = z.coords_at + (int) z.nv * z.cb;
	z.streams_at = z.lens_at + (int) z.nu;
	return z;
}

// xs[s] = x[column of slot s] for every slot of the staged packet. Eight
// chunks of 64 slots per round: their run-start words and run columns are
// read from LDS, then all eight loads of x are requested before the LDS
// stores. Slots past the last are given the last slot's column (a valid
// column) and not stored.
__device__ __forceinline__ void slot_stage_x(const uint8_t *pk, const SlotHeader &z, const double *__restrict__ x,
		double *xs)
{
	const int lane = lane_id();
	const int nsl = (int) z.nslots;
	const uint32_t *run_col = (const uint32_t *) (pk + z.runs_at);
	const uint32_t *bits = (const uint32_t *) (pk + z.bits_at);
	int run_base = 0, carry = 0; // runs started before the chunk; first slot of the run open at its start
	for (int c0 = 0; c0 * kWave < nsl; c0 += 8) {
		double xv[8];
#pragma unroll
		for (int i = 0; i < 8; i++) {
			const int c = c0 + i;
			xv[i] = 0.0;
			if (c * kWave < nsl) { // wave-uniform
				const uint32_t wlo = __builtin_amdgcn_readfirstlane(bits[2 * c]);
				const uint32_t whi = __builtin_amdgcn_readfirstlane(bits[2 * c + 1]);
				const uint64_t w = ((uint64_t) whi << 32) | wlo;
				const int s = min(c * kWave + lane, nsl - 1);
				const int l = s - c * kWave;
				const uint64_t mk = w & ((2ull << l) - 1); // run starts at slots <= s of this chunk
				const int k = __popcll(mk);
				const int r = run_base + k - 1;
				const int start = k ? c * kWave + 63 - __clzll(mk) : carry;
				xv[i] = x[run_col[r] + (uint32_t) (s - start)];
				run_base += __popcll(w);
				if (w)
					carry = c * kWave + 63 - __clzll(w);
			}
		}
#pragma unroll
		for (int i = 0; i < 8; i++) {
			const int s = (c0 + i) * kWave + lane;
			if (s < nsl)
				xs[s] = xv[i];
		}
	}
}

// One replication-factor-1 group of a staged x-slot packet, K values per
// lane and step: the step's K coordinates and length bytes, then its K
// staged x values and deltas, then its K accumulations. Steps where every
// lane below L has values need no validity tests; later values are tested
// (their reads are redirected to positions inside the packet).
template <int K, bool CB2, bool IL>
__device__ __forceinline__ void slot_rf1_group(const uint8_t *__restrict__ pk, const double *xs, int coords_at,
		int lens_at, int crd, int la, int cstride, int lstride, int stream_at, int steps, int cnt, int dv, int L,
		int cb, int rb, double *yw, int &cursor)
{
	const uint32_t cmask = cb >= 4 ? 0xffffffffu : (1u << (8 * cb)) - 1;
	const uint32_t rmask = (1u << rb) - 1;
	uint64_t val = 0;
	const int nfull = dv / K * K;
	int p0 = 0;
	if (lane_id() < L) {
		for (; p0 < nfull; p0 += K) {
			uint32_t c[K];
			int lb[K];
#pragma unroll
			for (int k = 0; k < K; k++) {
				c[k] = CB2 ? (uint32_t) * (const uint16_t *) (pk + crd + k * cstride)
				           : lds_le_b32x(pk, crd + k * cstride, cmask);
				lb[k] = pk[la + k * lstride];
			}
			crd += K * cstride;
			la += K * lstride;
			__builtin_amdgcn_sched_barrier(0);
			double xv[K];
			uint64_t dlt[K];
#pragma unroll
			for (int k = 0; k < K; k++)
				xv[k] = xs[c[k] >> rb];
#pragma unroll
			for (int k = 0; k < K; k++) {
				const int len = lb[k] & 15, tz = (lb[k] >> 4) << 2;
				dlt[k] = stream_read<IL, 0>(pk, stream_at, cursor, len, 4 * L) << tz;
				cursor += len;
			}
			__builtin_amdgcn_sched_barrier(0);
#pragma unroll
			for (int k = 0; k < K; k++) {
				val += dlt[k];
				double d;
				__builtin_memcpy(&d, &val, 8);
				atomicAdd(&yw[c[k] & rmask], d * xv[k]);
			}
		}
	}
	for (; p0 < steps; p0 += K) {
		uint32_t c[K];
		int lb[K];
#pragma unroll
		for (int k = 0; k < K; k++) {
			const bool ok = p0 + k < cnt;
			const int at = ok ? crd + k * cstride : coords_at;
			c[k] = CB2 ? (uint32_t) * (const uint16_t *) (pk + at) : lds_le_b32x(pk, at, cmask);
			lb[k] = ok ? pk[la + k * lstride] : 0;
		}
		crd += K * cstride;
		la += K * lstride;
#pragma unroll
		for (int k = 0; k < K; k++) {
			const double xv = xs[c[k] >> rb];
			const int len = lb[k] & 15, tz = (lb[k] >> 4) << 2;
			val += stream_read<IL, 0>(pk, stream_at, cursor, len, 4 * L) << tz;
			cursor += len;
			double d;
			__builtin_memcpy(&d, &val, 8);
			if (p0 + k < cnt)
				atomicAdd(&yw[c[k] & rmask], d * xv);
		}
	}
	(void) lens_at;
}

template <int K, bool IL>
__global__ void __launch_bounds__(kWave) div_slot_kernel(const uint8_t *__restrict__ bytes,
		const int64_t *__restrict__ pk_off, const int32_t *__restrict__ grp_pk, const int32_t *__restrict__ grp_row,
		long ngroups, const int32_t *__restrict__ grp_order, int L, int buf_bytes, int xs_bytes, int acc_stride,
		int acc_copies, int xcd_k, const double *__restrict__ x, double *__restrict__ y)
{
	extern __shared__ u32x4 smem[];
	uint8_t *pk = (uint8_t *) smem;                    // buf_bytes (a multiple of 16, padding included)
	double *xs = (double *) (pk + buf_bytes);          // xs_bytes: the staged x slots
	double *yw = (double *) (pk + buf_bytes + xs_bytes); // acc_copies copies of the row accumulators
	const int lane = lane_id();
	double *ywl = yw + (lane & (acc_copies - 1)) * acc_stride;
	const long k = xcd_logical_block(xcd_k);
	if (k >= ngroups)
		return;
	const long g = grp_order ? grp_order[k] : k;
	const int p_end = grp_pk[g + 1];
	int p = grp_pk[g];
	const bool single = p_end - p == 1;
	const long gr0 = grp_row[g], gr1 = grp_row[g + 1];
	if (!single)
		for (long r = gr0 + lane; r < gr1; r += kWave)
			y[r] = 0.0;
	for (; p < p_end; p++) {
		// the packet copy, as in div_spmv_lds_kernel
		const u32x4 *src = (const u32x4 *) (bytes + pk_off[p]);
		const int n16 = (int) ((pk_off[p + 1] - pk_off[p] + 15) >> 4);
		for (int j0 = 0; j0 < n16; j0 += kWave) {
			if (j0 + lane < n16)
				__builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *) (src + j0 + lane),
						(__attribute__((address_space(3))) void *) (smem + j0), 16, 0, 2);
		}
		__builtin_amdgcn_s_waitcnt(0);
		if (lane == 0)
			smem[n16] = u32x4{0, 0, 0, 0};
		wave_fence();
		const SlotHeader z = slot_header(pk, L);
		for (int r = lane; r < acc_copies * acc_stride; r += kWave)
			yw[r] = 0.0;
		slot_stage_x(pk, z, x, xs);
		int stream_at = ((z.streams_at + 3) & ~3) + 4 * lane; // the lane's column of the interleaved streams
		if (!IL) {
			// this lane's value stream: exclusive scan of the lane sizes
			const uint32_t *lsz = (const uint32_t *) (pk + z.lanes_at);
			int my_size = lane < L ? (int) lsz[lane] : 0;
			int incl = my_size;
			for (int off = 1; off < kWave; off <<= 1) {
				const int o = __shfl_up(incl, off, kWave);
				if (lane >= off)
					incl += o;
			}
			stream_at = z.streams_at + incl - my_size;
		}
		wave_fence(); // the staged x and the zeroed accumulators
		const uint32_t *rfs = (const uint32_t *) (pk + z.rfs_at);
		const uint32_t *per_rf = rfs + z.nrfs;
		const uint32_t cmask = z.cb >= 4 ? 0xffffffffu : (1u << (8 * z.cb)) - 1;
		const uint32_t rmask = (1u << z.rb) - 1;
		int cursor = 0, ks = 0, cs = 0;
		for (uint32_t gi = 0; gi < z.nrfs; gi++) {
			const int rf = (int) __builtin_amdgcn_readfirstlane(rfs[gi]);
			const int n = (int) __builtin_amdgcn_readfirstlane(per_rf[gi]);
			const int dv = n / L, md = n % L, mul_c = dv * rf * L;
			const int cnt = lane < L ? dv + (lane < md ? 1 : 0) : 0;
			const int steps = dv + (md ? 1 : 0);
			if (rf == 1) {
				const int crd = z.coords_at + (cs + lane) * z.cb, la = z.lens_at + ks + lane;
				if (z.cb == 2)
					slot_rf1_group<K, true, IL>(pk, xs, z.coords_at, z.lens_at, crd, la, L * 2, L, stream_at, steps,
							cnt, dv, L, 2, z.rb, ywl, cursor);
				else
					slot_rf1_group<K, false, IL>(pk, xs, z.coords_at, z.lens_at, crd, la, L * z.cb, L, stream_at,
							steps, cnt, dv, L, z.cb, z.rb, ywl, cursor);
				ks += n;
				cs += n;
				continue;
			}
			uint64_t val = 0;
			for (int pp = 0; pp < cnt; pp++) {
				const int lb = pk[z.lens_at + ks + pp * L + lane];
				const int len = lb & 15, tz = (lb >> 4) << 2;
				val += stream_read<IL, 0>(pk, stream_at, cursor, len, 4 * L) << tz;
				cursor += len;
				double d;
				__builtin_memcpy(&d, &val, 8);
				for (int t = 0; t < rf; t++) {
					const int cpos = pp < dv ? (pp * rf + t) * L + lane : mul_c + lane * rf + t;
					const uint32_t c = lds_le_b32x(pk, z.coords_at + (cs + cpos) * z.cb, cmask);
					atomicAdd(&ywl[c & rmask], d * xs[c >> z.rb]);
				}
			}
			ks += n;
			cs += n * rf;
		}
		wave_fence();
		if (single) {
			for (long r = gr0 + lane; r < gr1; r += kWave) {
				const long i = r - (long) z.row_min;
				y[r] = i >= 0 && i < (long) z.nrows ? div_acc_sum(yw, acc_stride, acc_copies, (int) i) : 0.0;
			}
		} else {
			for (uint32_t r = lane; r < z.nrows; r += kWave)
				y[z.row_min + r] += div_acc_sum(yw, acc_stride, acc_copies, (int) r);
		}
		wave_fence(); // the packet buffer, slots and accumulators are reused
	}
}

// Export of the decoded (row, col, value) sequence of the x-slot layout, in
// packet order (the same sequence as div_export_kernel's on the reference
// bytes). scratch: nnz int32, the slot -> column table of each packet at
// its pk_nnz offset.
__global__ void __launch_bounds__(kBlock) div_slot_export_kernel(const uint8_t *__restrict__ bytes,
		const int64_t *__restrict__ pk_off, const int64_t *__restrict__ pk_nnz, const int32_t *__restrict__ grp_pk,
		long ngroups, int L, int il, int32_t *__restrict__ scratch, int32_t *__restrict__ rows_out,
		int32_t *__restrict__ cols_out, double *__restrict__ vals_out)
{
	const int lane = lane_id(), wv = threadIdx.x / kWave;
	const long g = (long) blockIdx.x * kDivWavesPerBlock + wv;
	if (g >= ngroups)
		return;
	for (int p = grp_pk[g]; p < grp_pk[g + 1]; p++) {
		const uint8_t *pk = bytes + pk_off[p];
		const uint32_t *h = (const uint32_t *) pk;
		const uint32_t nv = h[0], nu = h[1], row_min = h[3], nslots = h[4];
		const int rb = pk[20], cb = pk[22];
		const uint32_t nrfs = h[6], nruns = h[7];
		const long runs_at = spmvdiv::kSlotHeaderFixed + 4 * (long) L + 8 * (long) nrfs;
		const long bits_at = (runs_at + 4 * (long) nruns + 7) & ~7L;
		const long coords_at = bits_at + 8 * (((long) nslots + 63) >> 6);
		const long lens_at = coords_at + (long) nv * cb;
		const uint32_t *run_col = (const uint32_t *) (pk + runs_at);
		const uint64_t *bits = (const uint64_t *) (pk + bits_at);
		int32_t *slot_col = scratch + pk_nnz[p];
		if (lane == 0) { // the table, slot by slot
			long r = -1;
			uint32_t start = 0;
			for (uint32_t s = 0; s < nslots; s++) {
				if (bits[s >> 6] >> (s & 63) & 1) {
					r++;
					start = s;
				}
				slot_col[s] = (int32_t) (run_col[r] + (s - start));
			}
		}
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
		__builtin_amdgcn_wave_barrier();
		long my_size = lane < L ? (long) h[8 + lane] : 0;
		long incl = my_size;
		for (int off = 1; off < kWave; off <<= 1) {
			long o = __shfl_up(incl, off, kWave);
			if (lane >= off)
				incl += o;
		}
		const uint32_t *rfs = (const uint32_t *) (pk + spmvdiv::kSlotHeaderFixed + 4 * L);
		const uint32_t *per_rf = rfs + nrfs;
		const uint8_t *lens = pk + lens_at;
		const long sbase = il ? (lens_at + nu + 3) & ~3L : lens_at + nu + (incl - my_size);
		const uint64_t rmask = rb ? (1ull << rb) - 1 : 0;
		const long nnz0 = pk_nnz[p];
		long cursor = 0, ks = 0, cs = 0;
		for (uint32_t gi = 0; gi < nrfs; gi++) {
			const long rf = rfs[gi], n = per_rf[gi];
			const long dv = n / L, md = n % L, mul_c = dv * rf * L;
			if (lane < L) {
				uint64_t val = 0;
				const long cnt = dv + (lane < md ? 1 : 0);
				for (long pp = 0; pp < cnt; pp++) {
					const uint8_t lb = lens[ks + pp * L + lane];
					const int len = lb & 15, tz = (lb >> 4) << 2;
					if (len) {
						val += stream_le(pk, sbase, lane, cursor, len, il, 4 * L) << tz;
						cursor += len;
					}
					double d;
					memcpy(&d, &val, 8);
					for (long t = 0; t < rf; t++) {
						const long cpos = pp < dv ? (pp * rf + t) * L + lane : mul_c + lane * rf + t;
						const uint64_t c = load_le(pk + coords_at + (cs + cpos) * cb, cb);
						const long o = nnz0 + cs + cpos;
						rows_out[o] = (int32_t) (row_min + (uint32_t) (c & rmask));
						cols_out[o] = slot_col[c >> rb];
						vals_out[o] = d;
					}
				}
			}
			ks += n;
			cs += n * rf;
		}
	}
}

// ------------------------------------------------------------ batched launch
//
// The DIV-RF members of a batch: their packet groups form one list, drawn by
// the resident wavefronts (one per workgroup) by ticket, unit g from counter
// g % nc, each wave requesting its next ticket before it decodes the current
// group (the schedule of batch.hip's kernel). A group is decoded by
// div_group_lds with the member's own lanes, accumulator copies and stream
// layout; the LDS holds the largest member's packet buffer and accumulators.

__global__ void div_batch_reset(int *counter, int nc, int stride)
{
	if ((int) threadIdx.x < nc)
		counter[threadIdx.x * stride] = 0;
}

template <int K>
__global__ void __launch_bounds__(kWave) div_batch_kernel(const BatchDivMat *__restrict__ mats,
		const BatchDivUnit *__restrict__ units, int nunits, int nc, int stride, int *__restrict__ counter,
		int buf_bytes)
{
	extern __shared__ u32x4 smem[];
	const int c = (int) (blockIdx.x & (unsigned) (nc - 1));
	int *ctr = counter + c * stride;
	const BatchDivMat *cm = const_ptr<BatchDivMat>(ptr_bits(mats));
	const BatchDivUnit *cu = const_ptr<BatchDivUnit>(ptr_bits(units));
	int tk = 0;
	if (lane_id() == 0)
		tk = atomicAdd(ctr, 1);
	int t = __builtin_amdgcn_readfirstlane(__shfl(tk, 0, kWave));
	for (;;) {
		const long g = (long) t * nc + c;
		if (g >= nunits)
			break; // wave-uniform: tickets only grow, every wave reaches it
		int tn = 0;
		if (lane_id() == 0)
			tn = atomicAdd(ctr, 1); // the next ticket, in flight during this group
		const BatchDivUnit u = cu[g];
		const BatchDivMat M = cm[u.mat];
		if (M.il)
			div_group_lds<K, true>(global_ptr<const uint8_t>(M.bytes), global_ptr<const int64_t>(M.pk_off),
					global_ptr<const int32_t>(M.grp_pk), global_ptr<const int32_t>(M.grp_row), u.group, M.L,
					buf_bytes, M.acc_stride, M.acc_copies, global_ptr<const double>(M.x), global_ptr<double>(M.y),
					smem);
		else
			div_group_lds<K, false>(global_ptr<const uint8_t>(M.bytes), global_ptr<const int64_t>(M.pk_off),
					global_ptr<const int32_t>(M.grp_pk), global_ptr<const int32_t>(M.grp_row), u.group, M.L,
					buf_bytes, M.acc_stride, M.acc_copies, global_ptr<const double>(M.x), global_ptr<double>(M.y),
					smem);
		t = __builtin_amdgcn_readfirstlane(__shfl(tn, 0, kWave));
	}
}

int div_batch_member(const spmvhip_format *f, const double *d_x, double *d_y, BatchDivMat *out, int *buf_bytes,
		int *acc_bytes)
{
	if (f->div_slots || f->div_buf_bytes <= 0)
		return set_error(SPMVHIP_EINVAL, "batch: div_rf member needs the LDS-staged decoder without x slots "
				"(div_layout 1 or 3, packets that fit the LDS)");
	BatchDivMat M{};
	M.bytes = ptr_bits(f->div_stream.p);
	M.pk_off = ptr_bits(f->div_pk_off.p);
	M.grp_pk = ptr_bits(f->div_groups.p);
	M.grp_row = ptr_bits(f->div_group_rows.p);
	M.x = ptr_bits(d_x);
	M.y = ptr_bits(d_y);
	M.L = f->div_lanes;
	M.acc_stride = f->div_acc_stride;
	M.acc_copies = f->p.div_acc_copies;
	M.il = f->div_il ? 1 : 0;
	*out = M;
	*buf_bytes = f->div_buf_bytes;
	*acc_bytes = f->div_acc_bytes;
	return SPMVHIP_OK;
}

int div_batch_occupancy(int lds_bytes, int *per_cu)
{
	SPMVHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, div_batch_kernel<8>, kWave, lds_bytes));
	return SPMVHIP_OK;
}

int div_batch_launch(const BatchDivMat *d_mats, const BatchDivUnit *d_units, int nunits, int nc, int stride,
		int *d_counter, int grid, int buf_bytes, int acc_bytes, hipStream_t s)
{
	div_batch_reset<<<1, 1024, 0, s>>>(d_counter, nc, stride);
	SPMVHIP_CHECK(hipGetLastError());
	if (nunits > 0) {
		div_batch_kernel<8><<<grid, kWave, (size_t) buf_bytes + (size_t) acc_bytes, s>>>(d_mats, d_units, nunits, nc,
				stride, d_counter, buf_bytes);
		SPMVHIP_CHECK(hipGetLastError());
	}
	return SPMVHIP_OK;
}

__global__ void fill_zero(double *y, long m)
{
	long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i < m)
		y[i] = 0.0;
}

// The LDS row accumulators of the staged decoders for packets of up to
// max_rows rows: an odd stride puts the copies of one row in different LDS
// banks; the automatic copy count is the largest power of two up to
// kDivAccCopiesAuto within kDivAccBudget bytes.
static void div_acc_setup(spmvhip_format *f, uint32_t max_rows)
{
	f->div_acc_stride = (int) (max_rows | 1);
	int nc = f->p.div_acc_copies;
	if (nc == 0) {
		nc = 1;
		while (nc < kDivAccCopiesAuto && 2 * nc * f->div_acc_stride * (long) sizeof(double) <= kDivAccBudget)
			nc *= 2;
	}
	f->p.div_acc_copies = nc;
	f->div_acc_bytes = (int) ((nc * f->div_acc_stride * sizeof(double) + 15) / 16 * 16);
}

// div_rf conversion with the packets encoded on the device (div_enc.hip):
// the packet plan from the host row pointers, the bytes from the device CSR,
// in the automatic layout (interleaved lane streams when a packet fits the
// LDS, else the reference bytes) or the requested one (1 or 3).
static int div_convert_device(spmvhip_format *f, const int32_t *h_rp, int T, int L, long P)
{
	spmvdiv::PacketPlan plan;
	spmvdiv::plan_packets(h_rp, f->m, f->nnz, T, P, &plan);
	const long np = (long) plan.i0.size(), ng = (long) plan.grp_pk.size() - 1;
	f->div_order_host.clear();
	f->div_npackets = np;
	f->div_ngroups = ng;
	DivDeviceEncoder enc;
	SPMVHIP_TRY(enc.prepare(f, plan, L));
	const int want = f->p.div_layout;
	bool il = want == 0 || want == 3;
	int64_t mx = 0;
	uint32_t max_rows = 1;
	SPMVHIP_TRY(enc.size(il, &mx, &max_rows));
	div_acc_setup(f, max_rows);
	if (il && mx + 16 + 8L * L + f->div_acc_bytes > kDivLdsMax) {
		if (want == 3)
			return set_error(SPMVHIP_EINVAL, "div_rf: layout 3 needs %ld bytes of LDS per packet (at most %ld)",
					(long) (mx + 16 + 8L * L + f->div_acc_bytes), kDivLdsMax);
		il = false;
		SPMVHIP_TRY(enc.size(il, &mx, &max_rows));
	}
	SPMVHIP_TRY(enc.write());
	f->div_slots = false;
	f->div_il = il;
	f->div_xs_bytes = 0;
	const int64_t max_pk = mx + (il ? 8L * L : 0);
	f->div_buf_bytes = max_pk + 16 + f->div_acc_bytes <= kDivLdsMax ? (int) (max_pk + 16) : 0;
	std::vector<int64_t> pk_nnz(plan.j0.begin(), plan.j0.end());
	pk_nnz.push_back(f->nnz);
	SPMVHIP_TRY(f->div_pk_nnz.alloc((np + 1) * sizeof(int64_t)));
	SPMVHIP_TRY(f->div_groups.alloc((ng + 1) * sizeof(int32_t)));
	SPMVHIP_TRY(f->div_group_rows.alloc((ng + 1) * sizeof(int32_t)));
	hipStream_t s = f->stream;
	SPMVHIP_CHECK(hipMemcpyAsync(f->div_pk_nnz.p, pk_nnz.data(), (np + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipMemcpyAsync(f->div_groups.p, plan.grp_pk.data(), (ng + 1) * sizeof(int32_t),
			hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipMemcpyAsync(f->div_group_rows.p, plan.grp_row.data(), (ng + 1) * sizeof(int32_t),
			hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	f->row_ptr.release();
	f->col.release();
	f->val.release();
	f->mem_footprint = (double) f->div_bytes + (double) (np + 1) * 16 + (double) (ng + 1) * 8;
	return SPMVHIP_OK;
}

int div_convert(spmvhip_format *f, const int32_t *h_rp, const int32_t *h_col, const double *h_val)
{
	// the packet plan reads the row pointers on the host (the CSR of a
	// device-array conversion is brought over as far as it is needed)
	std::vector<int32_t> hc_rp;
	if (!h_rp) {
		hc_rp.resize(f->m + 1);
		SPMVHIP_CHECK(hipMemcpy(hc_rp.data(), f->row_ptr.p, (f->m + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
		h_rp = hc_rp.data();
	}
	const int T = f->p.div_threads > 0 ? f->p.div_threads : 1;
	const int L = f->p.div_lanes > 0 ? f->p.div_lanes : 64;
	const long P = f->p.div_packet_vals > 0 ? f->p.div_packet_vals : 16384;
	if (L > kWave)
		return set_error(SPMVHIP_EINVAL, "div_rf: %d lanes > the 64-lane wavefront", L);
	const int ac = f->p.div_acc_copies;
	if (ac < 0 || ac > kDivAccCopiesMax || (ac & (ac - 1)))
		return set_error(SPMVHIP_EINVAL, "div_rf: div_acc_copies %d (0, or a power of two up to %d)", ac,
				kDivAccCopiesMax);
	if (T > 4096)
		return set_error(SPMVHIP_EINVAL, "div_rf: %d partitions", T);
	f->p.div_threads = T;
	f->p.div_lanes = L;
	f->p.div_packet_vals = P;
	f->div_lanes = L;
	if (f->p.div_reorder < 0 || f->p.div_reorder > 1)
		return set_error(SPMVHIP_EINVAL, "div_rf: div_reorder %d (0 stream order, 1 by median column)",
				f->p.div_reorder);
	if (f->p.div_layout < 0 || f->p.div_layout > 4)
		return set_error(SPMVHIP_EINVAL, "div_rf: div_layout %d (0 automatic, 1 reference bytes, 2 x slots, "
				"3 interleaved lane streams, 4 x slots and interleaved lane streams)", f->p.div_layout);
	const int enc = f->p.div_encoder;
	if (enc < 0 || enc > 2)
		return set_error(SPMVHIP_EINVAL, "div_rf: div_encoder %d (0 automatic, 1 host, 2 device)", enc);
	const bool dev_ok = div_device_encodable(P, L) && f->p.div_reorder == 0 &&
			(f->p.div_layout == 0 || f->p.div_layout == 1 || f->p.div_layout == 3);
	if (enc == 2 && !dev_ok)
		return set_error(SPMVHIP_EINVAL, "div_rf: the device encoder takes packets of at most 2048 values, at "
				"most 64 lanes, layouts 1 and 3 and no packet reordering");
	if (enc != 1 && dev_ok && f->nnz > 0)
		return div_convert_device(f, h_rp, T, L, P);
	// the host encoder reads the CSR on the host
	std::vector<int32_t> hc_col;
	std::vector<double> hc_val;
	if (!h_col && f->nnz > 0) {
		hc_col.resize(f->nnz);
		hc_val.resize(f->nnz);
		SPMVHIP_CHECK(hipMemcpy(hc_col.data(), f->col.p, f->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
		SPMVHIP_CHECK(hipMemcpy(hc_val.data(), f->val.p, f->nnz * sizeof(double), hipMemcpyDeviceToHost));
		h_col = hc_col.data();
		h_val = hc_val.data();
	}
	spmvdiv::Stream st;
	const char *err = nullptr;
	std::vector<double> lossy;
	if (f->p.div_tolerance != 0) {
		lossy.resize(f->nnz);
		if (spmvdiv::reduce_precision(h_val, f->nnz, f->p.div_tolerance, lossy.data(), &err))
			return set_error(SPMVHIP_EINVAL, "div_rf: %s (tolerance %g)", err, f->p.div_tolerance);
		h_val = lossy.data();
	}
	if (spmvdiv::encode(h_rp, h_col, h_val, f->m, f->n, f->nnz, T, L, P, &st, &err))
		return set_error(SPMVHIP_EFORMAT, "div_rf: %s (the reference rejects this matrix too)", err);
	const long np = st.npackets(), ng = st.ngroups();
	f->div_order_host.clear();
	if (f->p.div_reorder && ng > 0)
		f->div_order_host = spmvdiv::group_order_by_median(st, h_col);
	f->div_npackets = np;
	f->div_ngroups = ng;
	// the most rows of any packet (header word 2, at most 255)
	uint32_t max_rows = 1;
	for (long p = 0; p < np; p++) {
		uint32_t nr;
		memcpy(&nr, &st.bytes[st.pk_off[p] + 8], sizeof(nr));
		max_rows = std::max(max_rows, nr);
	}
	div_acc_setup(f, max_rows);
	// the device layout: the requested one, or (automatic) the first of
	// kDivAutoLayouts whose packet (+ 16 of zero padding, + 2 rows of stream
	// slack when interleaved), staged x slots and row accumulators fit the LDS
	// of one workgroup; the reference bytes when none does
	std::vector<uint8_t> dev;
	std::vector<int64_t> doff;
	int64_t max_pk = 0;
	f->div_slots = false;
	f->div_il = false;
	f->div_xs_bytes = 0;
	const int want = f->p.div_layout;
	const int *tries = want == 0 ? kDivAutoLayouts : &want;
	const int ntries = want == 0 ? (int) (sizeof(kDivAutoLayouts) / sizeof(int)) : 1;
	bool placed = false;
	for (int t = 0; t < ntries && np > 0 && !placed; t++) {
		const int lay = tries[t];
		const bool slots = lay == 2 || lay == 4, il = lay == 3 || lay == 4;
		spmvdiv::DeviceLayout dl;
		if (spmvdiv::to_device_layout(st, slots, il, &dl, &err)) {
			if (want != 0)
				return set_error(SPMVHIP_EFORMAT, "div_rf: %s", err);
			continue;
		}
		const long xs_bytes = slots ? ((long) dl.max_slots * 8 + 15) / 16 * 16 : 0;
		const long need = dl.max_packet + 16 + (il ? 8L * L : 0) + xs_bytes + f->div_acc_bytes;
		if (lay != 1 && need > kDivLdsMax) {
			if (want != 0)
				return set_error(SPMVHIP_EINVAL, "div_rf: layout %d needs %ld bytes of LDS per packet (at most %ld)",
						lay, need, kDivLdsMax);
			continue;
		}
		f->div_slots = slots;
		f->div_il = il;
		f->div_xs_bytes = (int) xs_bytes;
		max_pk = dl.max_packet + (il ? 8L * L : 0);
		dev.swap(dl.bytes);
		doff.swap(dl.pk_off);
		placed = true;
	}
	if (!placed) {
		// the reference's packet bytes, each starting at a 16-byte aligned offset
		doff.assign(np + 1, 0);
		int64_t at = 0;
		for (long p = 0; p < np; p++) {
			doff[p] = at;
			at += (st.pk_off[p + 1] - st.pk_off[p] + 15) / 16 * 16;
		}
		doff[np] = at;
		dev.assign(at + 16, 0); // +16: the decoder reads whole aligned 8-byte words
		for (long p = 0; p < np; p++)
			memcpy(&dev[doff[p]], &st.bytes[st.pk_off[p]], st.pk_off[p + 1] - st.pk_off[p]);
		for (long p = 0; p < np; p++)
			max_pk = std::max<int64_t>(max_pk, doff[p + 1] - doff[p]);
	}
	f->div_bytes = (long) doff[np > 0 ? np : 0];
	// the staged decoders keep a packet (16-byte rounded, + 16 of zero padding,
	// + stream slack), the x slots and one accumulator per row of the packet in
	// LDS: used while that fits 64 KiB (else the reference bytes are decoded
	// from global memory)
	f->div_buf_bytes = max_pk + 16 + f->div_xs_bytes + f->div_acc_bytes <= kDivLdsMax ? (int) (max_pk + 16) : 0;
	if (dev.empty())
		dev.assign(16, 0);
	if (doff.empty())
		doff.assign(1, 0);
	SPMVHIP_TRY(f->div_stream.alloc(dev.size()));
	SPMVHIP_TRY(f->div_pk_off.alloc((np + 1) * sizeof(int64_t)));
	SPMVHIP_TRY(f->div_pk_nnz.alloc((np + 1) * sizeof(int64_t)));
	SPMVHIP_TRY(f->div_groups.alloc((ng + 1) * sizeof(int32_t)));
	SPMVHIP_TRY(f->div_group_rows.alloc((ng + 1) * sizeof(int32_t)));
	hipStream_t s = f->stream;
	SPMVHIP_CHECK(hipMemcpyAsync(f->div_stream.p, dev.data(), dev.size(), hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipMemcpyAsync(f->div_pk_off.p, doff.data(), (np + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipMemcpyAsync(f->div_pk_nnz.p, st.pk_nnz.data(), (np + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipMemcpyAsync(f->div_groups.p, st.grp_pk.data(), (ng + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipMemcpyAsync(f->div_group_rows.p, st.grp_row.data(), (ng + 1) * sizeof(int32_t),
			hipMemcpyHostToDevice, s));
	if (!f->div_order_host.empty()) {
		SPMVHIP_TRY(f->div_order.alloc(ng * sizeof(int32_t)));
		SPMVHIP_CHECK(hipMemcpyAsync(f->div_order.p, f->div_order_host.data(), ng * sizeof(int32_t),
				hipMemcpyHostToDevice, s));
	}
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	f->row_ptr.release();
	f->col.release();
	f->val.release();
	f->mem_footprint = (double) f->div_bytes + (double) (np + 1) * 16 + (double) (ng + 1) * 8;
	return SPMVHIP_OK;
}

int div_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s)
{
	if (f->m == 0)
		return SPMVHIP_OK;
	if (f->div_ngroups == 0) {
		fill_zero<<<grid_for(f->m, kBlock), kBlock, 0, s>>>(y, f->m);
		SPMVHIP_CHECK(hipGetLastError());
		return SPMVHIP_OK;
	}
	const int32_t *order = f->div_order.p ? f->div_order.as<int32_t>() : nullptr;
	if (f->div_slots) {
		const size_t lds = (size_t) f->div_buf_bytes + (size_t) f->div_xs_bytes + (size_t) f->div_acc_bytes;
		const int xk = xcd_chunk(f->p.xcd_remap);
		const unsigned grid = xcd_padded_grid((unsigned) f->div_ngroups, xk);
#define LAUNCH_SLOT(K, IL)                                                                                        \
	div_slot_kernel<K, IL><<<grid, kWave, lds, s>>>(f->div_stream.as<uint8_t>(), f->div_pk_off.as<int64_t>(),  \
			f->div_groups.as<int32_t>(), f->div_group_rows.as<int32_t>(), f->div_ngroups, order, f->div_lanes, \
			f->div_buf_bytes, f->div_xs_bytes, f->div_acc_stride, f->p.div_acc_copies, xk, x, y)
#define LAUNCH_SLOT_K(IL)          \
	if (f->p.unroll >= 16)     \
		LAUNCH_SLOT(16, IL); \
	else if (f->p.unroll >= 8) \
		LAUNCH_SLOT(8, IL);  \
	else                       \
		LAUNCH_SLOT(4, IL)
		if (f->div_il) {
			LAUNCH_SLOT_K(true);
		} else {
			LAUNCH_SLOT_K(false);
		}
#undef LAUNCH_SLOT_K
#undef LAUNCH_SLOT
	} else if (f->div_buf_bytes > 0) {
		const size_t lds = (size_t) f->div_buf_bytes + (size_t) f->div_acc_bytes;
		const int xk = xcd_chunk(f->p.xcd_remap);
		const unsigned grid = xcd_padded_grid((unsigned) f->div_ngroups, xk);
#define LAUNCH_LDS(K, IL)                                                                                       \
	div_spmv_lds_kernel<K, IL><<<grid, kWave, lds, s>>>(f->div_stream.as<uint8_t>(), f->div_pk_off.as<int64_t>(), \
			f->div_groups.as<int32_t>(), f->div_group_rows.as<int32_t>(), f->div_ngroups, order, f->div_lanes, \
			f->div_buf_bytes, f->div_acc_stride, f->p.div_acc_copies, xk, x, y)
#define LAUNCH_LDS_K(IL)           \
	if (f->p.unroll >= 16)     \
		LAUNCH_LDS(16, IL);  \
	else if (f->p.unroll >= 8) \
		LAUNCH_LDS(8, IL);   \
	else                       \
		LAUNCH_LDS(4, IL)
		if (f->div_il) {
			LAUNCH_LDS_K(true);
		} else {
			LAUNCH_LDS_K(false);
		}
#undef LAUNCH_LDS_K
#undef LAUNCH_LDS
	} else {
		div_spmv_kernel<8><<<grid_for(f->div_ngroups, kDivWavesPerBlock), kBlock, 0, s>>>(f->div_stream.as<uint8_t>(),
				f->div_pk_off.as<int64_t>(), f->div_groups.as<int32_t>(), f->div_group_rows.as<int32_t>(),
				f->div_ngroups, order, f->div_lanes, x, y);
	}
	SPMVHIP_CHECK(hipGetLastError());
	return SPMVHIP_OK;
}

int div_export_decoded(const spmvhip_format *f, int which, void *buf, size_t *bytes)
{
	size_t need = f->nnz * (which == 8 ? sizeof(double) : sizeof(int32_t));
	if (!buf) {
		*bytes = need;
		return SPMVHIP_OK;
	}
	if (*bytes < need)
		return set_error(SPMVHIP_EINVAL, "export %d needs %zu bytes", which, need);
	if (f->nnz == 0) {
		*bytes = 0;
		return SPMVHIP_OK;
	}
	DevBuf rows, cols, vals, scratch;
	SPMVHIP_TRY(rows.alloc(f->nnz * sizeof(int32_t)));
	SPMVHIP_TRY(cols.alloc(f->nnz * sizeof(int32_t)));
	SPMVHIP_TRY(vals.alloc(f->nnz * sizeof(double)));
	if (f->div_slots) {
		SPMVHIP_TRY(scratch.alloc(f->nnz * sizeof(int32_t)));
		div_slot_export_kernel<<<grid_for(f->div_ngroups, kDivWavesPerBlock), kBlock, 0, f->stream>>>(
				f->div_stream.as<uint8_t>(), f->div_pk_off.as<int64_t>(), f->div_pk_nnz.as<int64_t>(),
				f->div_groups.as<int32_t>(), f->div_ngroups, f->div_lanes, f->div_il ? 1 : 0, scratch.as<int32_t>(),
				rows.as<int32_t>(), cols.as<int32_t>(), vals.as<double>());
	} else {
		div_export_kernel<<<grid_for(f->div_ngroups, kDivWavesPerBlock), kBlock, 0, f->stream>>>(
				f->div_stream.as<uint8_t>(), f->div_pk_off.as<int64_t>(), f->div_pk_nnz.as<int64_t>(),
				f->div_groups.as<int32_t>(), f->div_ngroups, f->div_lanes, f->div_il ? 1 : 0, rows.as<int32_t>(),
				cols.as<int32_t>(), vals.as<double>());
	}
	SPMVHIP_CHECK(hipGetLastError());
	SPMVHIP_CHECK(hipStreamSynchronize(f->stream));
	const void *src = which == 6 ? rows.p : which == 7 ? cols.p : vals.p;
	SPMVHIP_CHECK(hipMemcpy(buf, src, need, hipMemcpyDeviceToHost));
	*bytes = need;
	return SPMVHIP_OK;
}

} // namespace spmvhip

extern "C" int spmvhip_div_encode(const int32_t *row_ptr, const int32_t *col_ind, const double *values, long m,
		long n, long nnz, int T, int L, long P, uint8_t *bytes, size_t *nbytes, int64_t *pk_off, long *npackets,
		int32_t *grp_pk, long *ngroups)
{
	using spmvhip::set_error;
	if (!row_ptr || (nnz > 0 && (!col_ind || !values)) || !nbytes || !npackets || !ngroups)
		return set_error(SPMVHIP_EINVAL, "NULL argument");
	if (T < 1 || L < 1 || P < 1 || m < 0)
		return set_error(SPMVHIP_EINVAL, "T, L, P must be positive");
	spmvdiv::Stream st;
	const char *err = nullptr;
	if (spmvdiv::encode(row_ptr, col_ind, values, m, n, nnz, T, L, P, &st, &err))
		return set_error(SPMVHIP_EFORMAT, "div_rf: %s", err);
	if (bytes) {
		if (*nbytes < st.bytes.size() || *npackets < st.npackets() || *ngroups < st.ngroups())
			return set_error(SPMVHIP_EINVAL, "output buffers too small");
		memcpy(bytes, st.bytes.data(), st.bytes.size());
		if (pk_off)
			memcpy(pk_off, st.pk_off.data(), st.pk_off.size() * sizeof(int64_t));
		if (grp_pk)
			memcpy(grp_pk, st.grp_pk.data(), st.grp_pk.size() * sizeof(int32_t));
	}
	*nbytes = st.bytes.size();
	*npackets = st.npackets();
	*ngroups = st.ngroups();
	return SPMVHIP_OK;
}

extern "C" int spmvhip_div_reduce_precision(const double *in, long n, double tolerance, double *out)
{
	using spmvhip::set_error;
	if ((n > 0 && (!in || !out)) || n < 0)
		return set_error(SPMVHIP_EINVAL, "NULL argument or negative length");
	const char *err = nullptr;
	if (spmvdiv::reduce_precision(in, n, tolerance, out, &err))
		return set_error(SPMVHIP_EINVAL, "div_rf: %s (tolerance %g)", err, tolerance);
	return SPMVHIP_OK;
}

extern "C" int spmvhip_div_group_order(const int32_t *row_ptr, const int32_t *col_ind, const double *values, long m,
		long n, long nnz, int T, int L, long P, int32_t *order, long *ngroups)
{
	using spmvhip::set_error;
	if (!row_ptr || (nnz > 0 && (!col_ind || !values)) || !ngroups)
		return set_error(SPMVHIP_EINVAL, "NULL argument");
	if (T < 1 || L < 1 || P < 1 || m < 0)
		return set_error(SPMVHIP_EINVAL, "T, L, P must be positive");
	spmvdiv::Stream st;
	const char *err = nullptr;
	if (spmvdiv::encode(row_ptr, col_ind, values, m, n, nnz, T, L, P, &st, &err))
		return set_error(SPMVHIP_EFORMAT, "div_rf: %s", err);
	if (order) {
		if (*ngroups < st.ngroups())
			return set_error(SPMVHIP_EINVAL, "output buffer too small");
		const std::vector<int32_t> o = spmvdiv::group_order_by_median(st, col_ind);
		if (!o.empty())
			memcpy(order, o.data(), o.size() * sizeof(int32_t));
	}
	*ngroups = st.ngroups();
	return SPMVHIP_OK;
}

extern "C" int spmvhip_div_device_encode(const int32_t *row_ptr, const int32_t *col_ind, const double *values,
		long m, long n, long nnz, int T, int L, long P, int layout, uint8_t *bytes, size_t *nbytes, int64_t *pk_off,
		long *npackets)
{
	using spmvhip::set_error;
	if (!row_ptr || (nnz > 0 && (!col_ind || !values)) || !nbytes || !npackets)
		return set_error(SPMVHIP_EINVAL, "NULL argument");
	if (T < 1 || L < 1 || P < 1 || m < 0)
		return set_error(SPMVHIP_EINVAL, "T, L, P must be positive");
	if (layout < 1 || layout > 4)
		return set_error(SPMVHIP_EINVAL, "layout %d (1..4)", layout);
	spmvdiv::Stream st;
	const char *err = nullptr;
	if (spmvdiv::encode(row_ptr, col_ind, values, m, n, nnz, T, L, P, &st, &err))
		return set_error(SPMVHIP_EFORMAT, "div_rf: %s", err);
	spmvdiv::DeviceLayout dl;
	if (spmvdiv::to_device_layout(st, layout == 2 || layout == 4, layout == 3 || layout == 4, &dl, &err))
		return set_error(SPMVHIP_EFORMAT, "div_rf: %s", err);
	const size_t total = (size_t) dl.pk_off.back();
	if (bytes) {
		if (*nbytes < total || *npackets < st.npackets())
			return set_error(SPMVHIP_EINVAL, "output buffers too small");
		if (total)
			memcpy(bytes, dl.bytes.data(), total);
		if (pk_off)
			memcpy(pk_off, dl.pk_off.data(), dl.pk_off.size() * sizeof(int64_t));
	}
	*nbytes = total;
	*npackets = st.npackets();
	return SPMVHIP_OK;
}

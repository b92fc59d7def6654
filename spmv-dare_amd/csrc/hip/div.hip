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
		long ngroups, int L, int32_t *__restrict__ rows_out, int32_t *__restrict__ cols_out,
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
		const uint8_t *stream = lens + nu + lane_base;
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
						val += load_le(stream + cursor, len) << tz;
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
// registers. Products go to the packet's rows in LDS (atomic adds; the
// order of the additions is not fixed), which are added to y at the end of
// the packet.
template <int K>
__global__ void __launch_bounds__(kBlock) div_spmv_kernel(const uint8_t *__restrict__ bytes,
		const int64_t *__restrict__ pk_off, const int32_t *__restrict__ grp_pk, const int32_t *__restrict__ grp_row,
		long ngroups, int L, const double *__restrict__ x, double *__restrict__ y)
{
	__shared__ double yl[kDivWavesPerBlock][256];
	const int lane = lane_id(), wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
	const long g = (long) blockIdx.x * kDivWavesPerBlock + wv;
	if (g >= ngroups)
		return; // waves do not synchronise with each other
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
template <int K, int S, bool C4, bool A4L64>
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
			dlt[k] = lds_le_b32(pk, stream_at + cursor, len) << tz;
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
template <int K, bool C4, bool A4L64 = false>
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
				div_full_steps<K, 2, C4, A4L64>(pk, at, la, cstride, lstride, cb, cmask, row_bits, col_min,
						stream_at, x, yw, cursor, val);
				p0 = nfull;
			}
		}
		for (; p0 < nfull; p0 += K)
			div_full_steps<K, 1, C4, A4L64>(pk, at, la, cstride, lstride, cb, cmask, row_bits, col_min, stream_at,
					x, yw, cursor, val);
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
			val += lds_le_b32(pk, stream_at + cursor, len) << tz;
			cursor += len;
			double d;
			__builtin_memcpy(&d, &val, 8);
			if (ok) // (an unconditional add of 0.0 would pile every idle lane onto one LDS address)
				atomicAdd(&yw[row], d * xv[k]);
		}
	}
}

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

template <int K>
__global__ void __launch_bounds__(kWave) div_spmv_lds_kernel(const uint8_t *__restrict__ bytes,
		const int64_t *__restrict__ pk_off, const int32_t *__restrict__ grp_pk, const int32_t *__restrict__ grp_row,
		long ngroups, int L, int buf_bytes, int acc_stride, int acc_copies, int xcd_k, const double *__restrict__ x,
		double *__restrict__ y)
{
	extern __shared__ u32x4 smem[];
	uint8_t *pk = (uint8_t *) smem;                   // buf_bytes (a multiple of 16, padding included)
	double *yw = (double *) (pk + buf_bytes);          // acc_copies copies of the row accumulators
	const int lane = lane_id();
	// lane l adds into copy l % acc_copies: lanes whose values fall in the
	// same row collide on one LDS address (and serialise) acc_copies times
	// less often -- with one copy those collisions bound the decoder
	double *ywl = yw + (lane & (acc_copies - 1)) * acc_stride;
	const long g = xcd_logical_block(xcd_k); // groups dealt to XCDs in chunks of xcd_k (0: in order)
	if (g >= ngroups)
		return;
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
		long my_size = lane < L ? (long) h[7 + lane] : 0;
		long incl = my_size;
		for (int off = 1; off < kWave; off <<= 1) {
			long o = __shfl_up(incl, off, kWave);
			if (lane >= off)
				incl += o;
		}
		const uint32_t *rfs = h + 7 + L;
		const uint32_t *per_rf = rfs + nrfs;
		const long coords_at = 28 + 4 * (long) L + 8 * (long) nrfs;
		const long lens_at = coords_at + (long) nv * cb;
		const long stream_at = lens_at + nu + (incl - my_size);
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
					div_rf1_group<K, true, true>(pk, (int) coords_at, (int) lens_at, lens0,
							(int) coords_at + crd0 * cb, (int) stream_at, (int) steps, (int) cnt, (int) dv, L, cb,
							col_min, row_bits, x, ywl, cur);
				else if (cb <= 4)
					div_rf1_group<K, true>(pk, (int) coords_at, (int) lens_at, lens0, (int) coords_at + crd0 * cb,
							(int) stream_at, (int) steps, (int) cnt, (int) dv, L, cb, col_min, row_bits, x, ywl, cur);
				else
					div_rf1_group<K, false>(pk, (int) coords_at, (int) lens_at, lens0, (int) coords_at + crd0 * cb,
							(int) stream_at, (int) steps, (int) cnt, (int) dv, L, cb, col_min, row_bits, x, ywl, cur);
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
						val += lds_le(pk, stream_at + cursor, len) << tz;
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

__global__ void fill_zero(double *y, long m)
{
	long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i < m)
		y[i] = 0.0;
}

int div_convert(spmvhip_format *f, const int32_t *h_rp, const int32_t *h_col, const double *h_val)
{
	if (!h_rp)
		return set_error(SPMVHIP_EINVAL, "div_rf: host CSR arrays are required");
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
	spmvdiv::Stream st;
	const char *err = nullptr;
	if (spmvdiv::encode(h_rp, h_col, h_val, f->m, f->n, f->nnz, T, L, P, &st, &err))
		return set_error(SPMVHIP_EFORMAT, "div_rf: %s (the reference rejects this matrix too)", err);
	const long np = st.npackets(), ng = st.ngroups();
	// 16-byte aligned packet starts on the device
	std::vector<int64_t> doff(np + 1);
	int64_t at = 0;
	for (long p = 0; p < np; p++) {
		doff[p] = at;
		at += (st.pk_off[p + 1] - st.pk_off[p] + 15) / 16 * 16;
	}
	doff[np] = at;
	std::vector<uint8_t> dev(at + 16, 0); // +16: the decoder reads whole aligned 8-byte words
	for (long p = 0; p < np; p++)
		memcpy(&dev[doff[p]], &st.bytes[st.pk_off[p]], st.pk_off[p + 1] - st.pk_off[p]);
	f->div_npackets = np;
	f->div_ngroups = ng;
	f->div_bytes = (long) st.bytes.size();
	int64_t max_pk = 0;
	for (long p = 0; p < np; p++)
		max_pk = std::max<int64_t>(max_pk, doff[p + 1] - doff[p]);
	// the staged decoder keeps a packet (16-byte rounded, + 16 of zero padding)
	// and one accumulator per row of the packet in LDS: used while that fits
	// 64 KiB. Header word 2 is the packet's row count (at most 255).
	uint32_t max_rows = 1;
	for (long p = 0; p < np; p++) {
		uint32_t nr;
		memcpy(&nr, &st.bytes[st.pk_off[p] + 8], sizeof(nr));
		max_rows = std::max(max_rows, nr);
	}
	// an odd stride puts the copies of one row in different LDS banks
	f->div_acc_stride = (int) (max_rows | 1);
	int nc = f->p.div_acc_copies;
	if (nc == 0) {
		nc = 1;
		while (nc < kDivAccCopiesAuto && 2 * nc * f->div_acc_stride * (long) sizeof(double) <= kDivAccBudget)
			nc *= 2;
	}
	f->p.div_acc_copies = nc;
	f->div_acc_bytes = (int) ((nc * f->div_acc_stride * sizeof(double) + 15) / 16 * 16);
	f->div_buf_bytes = max_pk + 16 + f->div_acc_bytes <= kDivLdsMax ? (int) (max_pk + 16) : 0;
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
	if (f->div_buf_bytes > 0) {
		const size_t lds = (size_t) f->div_buf_bytes + (size_t) f->div_acc_bytes;
		const int xk = xcd_chunk(f->p.xcd_remap);
		const unsigned grid = xcd_padded_grid((unsigned) f->div_ngroups, xk);
		if (f->p.unroll >= 16)
			div_spmv_lds_kernel<16><<<grid, kWave, lds, s>>>(f->div_stream.as<uint8_t>(),
					f->div_pk_off.as<int64_t>(), f->div_groups.as<int32_t>(), f->div_group_rows.as<int32_t>(),
					f->div_ngroups, f->div_lanes, f->div_buf_bytes, f->div_acc_stride, f->p.div_acc_copies, xk, x, y);
		else if (f->p.unroll >= 8)
			div_spmv_lds_kernel<8><<<grid, kWave, lds, s>>>(f->div_stream.as<uint8_t>(),
					f->div_pk_off.as<int64_t>(), f->div_groups.as<int32_t>(), f->div_group_rows.as<int32_t>(),
					f->div_ngroups, f->div_lanes, f->div_buf_bytes, f->div_acc_stride, f->p.div_acc_copies, xk, x, y);
		else
			div_spmv_lds_kernel<4><<<grid, kWave, lds, s>>>(f->div_stream.as<uint8_t>(),
					f->div_pk_off.as<int64_t>(), f->div_groups.as<int32_t>(), f->div_group_rows.as<int32_t>(),
					f->div_ngroups, f->div_lanes, f->div_buf_bytes, f->div_acc_stride, f->p.div_acc_copies, xk, x, y);
	} else {
		div_spmv_kernel<8><<<grid_for(f->div_ngroups, kDivWavesPerBlock), kBlock, 0, s>>>(f->div_stream.as<uint8_t>(),
				f->div_pk_off.as<int64_t>(), f->div_groups.as<int32_t>(), f->div_group_rows.as<int32_t>(),
				f->div_ngroups, f->div_lanes, x, y);
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
	DevBuf rows, cols, vals;
	SPMVHIP_TRY(rows.alloc(f->nnz * sizeof(int32_t)));
	SPMVHIP_TRY(cols.alloc(f->nnz * sizeof(int32_t)));
	SPMVHIP_TRY(vals.alloc(f->nnz * sizeof(double)));
	div_export_kernel<<<grid_for(f->div_ngroups, kDivWavesPerBlock), kBlock, 0, f->stream>>>(
			f->div_stream.as<uint8_t>(), f->div_pk_off.as<int64_t>(), f->div_pk_nnz.as<int64_t>(),
			f->div_groups.as<int32_t>(), f->div_ngroups, f->div_lanes, rows.as<int32_t>(), cols.as<int32_t>(),
			vals.as<double>());
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

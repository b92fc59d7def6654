// DIV-RF encoding on the device: the packets of div_codec.cpp's encoder
// (the reference's lossless DIV-RF packets, src/spmv_kernels/DIV/
// div_kernels_rf.h:538-919, lossy values by :467-510), written straight into
// a device layout (div_codec.hpp: the reference bytes or interleaved lane
// streams, each packet at a 16-byte aligned offset), byte for byte what the
// host encoder followed by to_device_layout produces.
//
// The packet boundaries chain (each packet starts where the previous one
// ended, div.cpp:236-513), so they are planned on the host from the row
// pointers alone (spmvdiv::plan_packets, O(rows)); everything that reads the
// values and columns runs here. One workgroup encodes one packet of at most
// kEncMax values, in two launches: the first computes every packet's size,
// a device scan turns the sizes into offsets, the second writes the bytes.
//
// Per packet, in LDS:
//   1. keys (value bits, row, column) of the packet's nonzeros, the row by a
//      binary search in the packet's row pointers; column range;
//   2. bitonic sort by (value bits, row, column), the host's std::sort order
//      (the triples are distinct, so any correct sort gives the same order);
//   3. runs of equal value bits: a run of length r is a value of replication
//      factor r; the runs stably sorted by r (the reference's counting sort
//      by replication factor), as keys r << 11 | run;
//   4. per replication factor group: values interleaved over the L lanes in
//      units of 1, coordinates in units of r (lib/macros/permutation.h:
//      47-111), each output position mapped back to its source;
//   5. one thread per lane walks its values: delta to the lane's previous
//      value in the group, the byte length and trailing-zero nibbles of the
//      delta, then the payload bytes.

#include "format.hpp"
#include "scan.hpp"
#include "../div/div_codec.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace spmvhip {

constexpr int kEncMax = 2048;     // values per packet encoded on the device
constexpr int kEncThreads = 256;
constexpr int kEncMaxGroups = 64; // distinct replication factors of kEncMax values: d (d + 1) / 2 <= 2048
constexpr int kRunBits = 11;      // run index bits of a run key (runs < kEncMax)

// per-packet results of the size pass
struct EncInfo {
	int64_t bytes;    // device bytes of the packet (16-byte rounded)
	int32_t rows;     // header num_rows
	int32_t err;      // 0, 1 = row span above 255, 2 = column span above 32 bits
};

struct EncShared {
	uint64_t hi[kEncMax];          // value bits
	uint64_t lo[kEncMax];          // row offset << 32 | column
	uint32_t rkey[kEncMax];        // run keys rf << kRunBits | run, sorted
	uint16_t rstart[kEncMax + 1];  // first sorted position of each run, + nv
	uint16_t skpos[kEncMax + 1];   // first coordinate of each run in replication-factor order
	int32_t rps[257];              // rp[i0 .. i0 + 256] (INT32_MAX past m)
	uint32_t gq[kEncMaxGroups + 1], grf[kEncMaxGroups];
	uint32_t lane_bytes[kWave];
	uint32_t lane_off[kWave];
	uint32_t red[kEncThreads / kWave][2];
	uint32_t scan[kEncThreads];
	int nruns, ngroups, row_last;
};

__device__ __forceinline__ uint32_t req_bits_dev(uint64_t v)
{
	return v == 0 ? 0u : 64u - (uint32_t) __builtin_clzll(v);
}

// exclusive scan of v over the workgroup (every thread calls it); *total
// receives the sum
__device__ uint32_t block_excl_scan(EncShared &sh, uint32_t v, uint32_t *total)
{
	const int t = threadIdx.x;
	sh.scan[t] = v;
	__syncthreads();
	for (int off = 1; off < kEncThreads; off <<= 1) {
		const uint32_t o = t >= off ? sh.scan[t - off] : 0u;
		__syncthreads();
		sh.scan[t] += o;
		__syncthreads();
	}
	const uint32_t incl = sh.scan[t];
	*total = sh.scan[kEncThreads - 1];
	__syncthreads();
	return incl - v;
}

// Source position of output position j when N elements in units of U are
// interleaved over S segments (rem = 0: N is a multiple of U in every use):
// element-wise over the full rounds, then one remaining unit per segment.
__device__ __forceinline__ uint32_t interleave_src(uint32_t j, uint32_t N, uint32_t S, uint32_t U)
{
	const uint32_t units = N / U, div = units / S, mod = units % S, mul = div * U * S;
	uint32_t s, off;
	if (j < mul) {
		s = j % S;
		off = j / S;
	} else {
		s = ((j - mul) / U) % S;
		off = U * div + j % U;
	}
	return U * (s * div + (s < mod ? s : mod)) + off;
}

// The encoding of one delta (div_kernels_rf.h:835-893, compress_kernel_div):
// byte length, trailing-zero nibbles and the payload shifted down.
__device__ __forceinline__ void delta_code(uint64_t diff, uint32_t *len, uint32_t *tz4, uint64_t *payload)
{
	if (diff == 0) {
		*len = 0;
		*tz4 = 0;
		*payload = 0;
		return;
	}
	uint64_t lz = (uint64_t) __builtin_clzll(diff), tz = (uint64_t) __builtin_ctzll(diff);
	const uint64_t rem_tz = tz & 3;
	uint64_t len_bits = 64 - lz - tz;
	uint64_t ln = (len_bits + 7) >> 3;
	uint64_t rem = (ln << 3) - len_bits;
	if (rem >= rem_tz) {
		rem -= rem_tz;
		tz -= rem_tz;
	}
	if (rem <= lz) {
		lz -= rem;
		rem = 0;
	} else {
		rem -= lz;
		lz = 0;
	}
	tz -= rem;
	const uint64_t t4 = tz >> 2;
	tz = t4 << 2;
	len_bits = 64 - lz - tz;
	ln = (len_bits + 7) >> 3;
	*len = (uint32_t) ln;
	*tz4 = (uint32_t) t4;
	*payload = diff >> tz;
}

__device__ __forceinline__ void put_u32(uint8_t *o, long at, uint32_t v)
{
	o[at] = (uint8_t) v;
	o[at + 1] = (uint8_t) (v >> 8);
	o[at + 2] = (uint8_t) (v >> 16);
	o[at + 3] = (uint8_t) (v >> 24);
}

__device__ __forceinline__ bool key_gt(uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl)
{
	return ah > bh || (ah == bh && al > bl);
}

// Bitonic sort of n (a power of two) 128-bit keys in LDS, ascending.
__device__ void sort_keys(EncShared &sh, int n)
{
	for (int k = 2; k <= n; k <<= 1) {
		for (int j = k >> 1; j > 0; j >>= 1) {
			for (int i = threadIdx.x; i < n; i += kEncThreads) {
				const int ixj = i ^ j;
				if (ixj > i) {
					const uint64_t ah = sh.hi[i], al = sh.lo[i], bh = sh.hi[ixj], bl = sh.lo[ixj];
					if (key_gt(ah, al, bh, bl) == ((i & k) == 0)) {
						sh.hi[i] = bh;
						sh.lo[i] = bl;
						sh.hi[ixj] = ah;
						sh.lo[ixj] = al;
					}
				}
			}
			__syncthreads();
		}
	}
}

__device__ void sort_runs(EncShared &sh, int n)
{
	for (int k = 2; k <= n; k <<= 1) {
		for (int j = k >> 1; j > 0; j >>= 1) {
			for (int i = threadIdx.x; i < n; i += kEncThreads) {
				const int ixj = i ^ j;
				if (ixj > i) {
					const uint32_t a = sh.rkey[i], b = sh.rkey[ixj];
					if ((a > b) == ((i & k) == 0)) {
						sh.rkey[i] = b;
						sh.rkey[ixj] = a;
					}
				}
			}
			__syncthreads();
		}
	}
}

__device__ __forceinline__ uint32_t pow2_ceil(uint32_t v)
{
	uint32_t n = 1;
	while (n < v)
		n <<= 1;
	return n;
}

// Encodes packet blockIdx.x. WRITE = false: fills info[p]; WRITE = true:
// writes the packet's bytes at out + off[p] (the buffer zeroed beforehand).
// IL: interleaved lane streams (device layout 3), else the reference bytes.
template <bool WRITE, bool IL>
__global__ void __launch_bounds__(kEncThreads) div_encode_kernel(const int32_t *__restrict__ rp,
		const int32_t *__restrict__ col, const double *__restrict__ val, long m, const int32_t *__restrict__ pi0,
		const int64_t *__restrict__ pj0, const int32_t *__restrict__ pnv, int L, EncInfo *__restrict__ info,
		int64_t *__restrict__ size_out, const int64_t *__restrict__ off, uint8_t *__restrict__ out)
{
	__shared__ EncShared sh;
	const int t = threadIdx.x;
	const long p = blockIdx.x;
	const long i0 = pi0[p], j0 = pj0[p];
	const int nv = pnv[p];

	// 1. keys
	for (int r = t; r < 257; r += kEncThreads)
		sh.rps[r] = i0 + r <= m ? rp[i0 + r] : INT32_MAX;
	__syncthreads();
	const uint32_t N2 = pow2_ceil((uint32_t) nv);
	uint32_t cmin = 0xffffffffu, cmax = 0;
	for (int k = t; k < (int) N2; k += kEncThreads) {
		if (k < nv) {
			const long j = j0 + k;
			int lo_r = 0, hi_r = 256; // the last r with rps[r] <= j (rps[0] <= j0 <= j)
			while (lo_r < hi_r) {
				const int mid = (lo_r + hi_r + 1) >> 1;
				if ((long) sh.rps[mid] <= j)
					lo_r = mid;
				else
					hi_r = mid - 1;
			}
			const uint32_t c = (uint32_t) col[j];
			uint64_t u;
			const double v = val[j];
			__builtin_memcpy(&u, &v, 8);
			sh.hi[k] = u;
			sh.lo[k] = ((uint64_t) lo_r << 32) | c;
			cmin = c < cmin ? c : cmin;
			cmax = c > cmax ? c : cmax;
			if (k == nv - 1)
				sh.row_last = lo_r;
		} else {
			sh.hi[k] = ~0ull; // padding sorts last: no real key has row offset 2^32 - 1
			sh.lo[k] = ~0ull;
		}
	}
	for (int o = 32; o > 0; o >>= 1) {
		const uint32_t a = __shfl_xor(cmin, o, kWave), b = __shfl_xor(cmax, o, kWave);
		cmin = a < cmin ? a : cmin;
		cmax = b > cmax ? b : cmax;
	}
	if ((t & (kWave - 1)) == 0) {
		sh.red[t / kWave][0] = cmin;
		sh.red[t / kWave][1] = cmax;
	}
	__syncthreads();
	cmin = sh.red[0][0];
	cmax = sh.red[0][1];
	for (int w = 1; w < kEncThreads / kWave; w++) {
		cmin = sh.red[w][0] < cmin ? sh.red[w][0] : cmin;
		cmax = sh.red[w][1] > cmax ? sh.red[w][1] : cmax;
	}
	const int row_last = sh.row_last;
	// coordinate widths (div_kernels_rf.h:557-607): the row field padded to a
	// byte boundary when that costs no coordinate byte
	uint32_t row_bits = req_bits_dev((uint64_t) row_last);
	uint32_t col_bits = req_bits_dev((uint64_t) (cmax - cmin));
	const uint32_t cb = (row_bits + col_bits + 7) >> 3;
	if (row_bits % 8) {
		const uint32_t cmpl = 8 - row_bits % 8;
		if ((((row_bits + cmpl) + col_bits + 7) >> 3) == cb)
			row_bits += cmpl;
	}
	col_bits = (cb << 3) - row_bits;
	if (row_bits > 8 || col_bits > 32) {
		if (!WRITE && t == 0) {
			info[p] = EncInfo{0, 0, row_bits > 8 ? 1 : 2};
			size_out[p] = 0;
		}
		return; // uniform over the workgroup
	}

	// 2. sort
	sort_keys(sh, (int) N2);

	// 3. runs of equal value bits, then the runs by replication factor
	{
		uint32_t cnt = 0;
		const int per = (nv + kEncThreads - 1) / kEncThreads, k0 = t * per;
		for (int k = k0; k < k0 + per && k < nv; k++)
			cnt += (k == 0 || sh.hi[k] != sh.hi[k - 1]) ? 1u : 0u;
		uint32_t total;
		uint32_t r = block_excl_scan(sh, cnt, &total);
		for (int k = k0; k < k0 + per && k < nv; k++)
			if (k == 0 || sh.hi[k] != sh.hi[k - 1])
				sh.rstart[r++] = (uint16_t) k;
		if (t == 0) {
			sh.nruns = (int) total;
			sh.rstart[total] = (uint16_t) nv;
		}
	}
	__syncthreads();
	const int nruns = sh.nruns;
	const uint32_t R2 = pow2_ceil((uint32_t) nruns);
	for (int r = t; r < (int) R2; r += kEncThreads)
		sh.rkey[r] = r < nruns ? ((uint32_t) (sh.rstart[r + 1] - sh.rstart[r]) << kRunBits) | (uint32_t) r
				: 0xffffffffu;
	__syncthreads();
	sort_runs(sh, (int) R2);
	{
		// coordinate start of every run in replication-factor order, and the
		// groups of equal replication factor
		const int per = (nruns + kEncThreads - 1) / kEncThreads, q0 = t * per;
		uint32_t sum = 0, ng = 0;
		for (int q = q0; q < q0 + per && q < nruns; q++) {
			sum += sh.rkey[q] >> kRunBits;
			ng += (q == 0 || (sh.rkey[q] >> kRunBits) != (sh.rkey[q - 1] >> kRunBits)) ? 1u : 0u;
		}
		uint32_t tot_s, tot_g;
		uint32_t s0 = block_excl_scan(sh, sum, &tot_s);
		uint32_t g0 = block_excl_scan(sh, ng, &tot_g);
		for (int q = q0; q < q0 + per && q < nruns; q++) {
			const uint32_t rf = sh.rkey[q] >> kRunBits;
			sh.skpos[q] = (uint16_t) s0;
			s0 += rf;
			if (q == 0 || rf != (sh.rkey[q - 1] >> kRunBits)) {
				sh.gq[g0] = (uint32_t) q;
				sh.grf[g0] = rf;
				g0++;
			}
		}
		if (t == 0) {
			sh.ngroups = (int) tot_g;
			sh.gq[tot_g] = (uint32_t) nruns;
			sh.skpos[nruns] = (uint16_t) nv;
		}
	}
	__syncthreads();
	const int ngroups = sh.ngroups;
	const long hb = spmvdiv::kHeaderFixed + 4L * L;
	const long coords_at = hb + 8L * ngroups;
	const long lens_at = coords_at + (long) nv * cb;
	const long streams_at = lens_at + nruns;
	const long il_at = (streams_at + 3) & ~3L;
	uint8_t *o = WRITE ? out + off[p] : nullptr;

	// 5. lane streams (thread l < L is lane l): sizes first
	uint32_t my_bytes = 0;
	if (t < L) {
		for (int g = 0; g < ngroups; g++) {
			const uint32_t u0 = sh.gq[g], n = sh.gq[g + 1] - u0;
			uint64_t prev = 0;
			for (uint32_t q = t; q < n; q += L) {
				const uint32_t run = sh.rkey[u0 + interleave_src(q, n, L, 1)] & ((1u << kRunBits) - 1);
				const uint64_t v = sh.hi[sh.rstart[run]];
				uint32_t len, tz4;
				uint64_t pay;
				delta_code(v - prev, &len, &tz4, &pay);
				prev = v;
				my_bytes += len;
				if (WRITE) {
					o[lens_at + u0 + q] = (uint8_t) (len | (tz4 << 4));
				}
			}
		}
		sh.lane_bytes[t] = my_bytes;
	}
	__syncthreads();
	uint32_t total_bytes = 0, maxw = 0;
	for (int l = 0; l < L; l++) {
		total_bytes += sh.lane_bytes[l];
		maxw = max(maxw, (sh.lane_bytes[l] + 3) / 4);
	}
	const long end = IL ? il_at + (long) maxw * 4 * L : streams_at + total_bytes;
	if (!WRITE) {
		if (t == 0) {
			info[p] = EncInfo{(end + 15) / 16 * 16, row_last + 1, 0};
			size_out[p] = (end + 15) / 16 * 16;
		}
		return;
	}

	// header (the reference's words; byte 22 carries the interleaved flag)
	if (t == 0) {
		put_u32(o, 0, (uint32_t) nv);
		put_u32(o, 4, (uint32_t) nruns);
		put_u32(o, 8, (uint32_t) (row_last + 1));
		put_u32(o, 12, (uint32_t) i0);
		put_u32(o, 16, cmin);
		o[20] = (uint8_t) row_bits;
		o[21] = (uint8_t) col_bits;
		o[22] = IL ? 1 : 0;
		put_u32(o, 24, (uint32_t) ngroups);
	}
	if (t < L)
		put_u32(o, spmvdiv::kHeaderFixed + 4L * t, sh.lane_bytes[t]);
	for (int g = t; g < ngroups; g += kEncThreads) {
		put_u32(o, hb + 4L * g, sh.grf[g]);
		put_u32(o, hb + 4L * ngroups + 4L * g, sh.gq[g + 1] - sh.gq[g]);
	}
	// coordinates: coordinate e of group g (units of rf) maps back to element
	// (source unit, offset) of the group's runs in replication-factor order
	for (int e = t; e < nv; e += kEncThreads) {
		int g = 0;
		while (g + 1 < ngroups && sh.skpos[sh.gq[g + 1]] <= (uint32_t) e)
			g++;
		const uint32_t u0 = sh.gq[g], n = sh.gq[g + 1] - u0, rf = sh.grf[g];
		const uint32_t ls = sh.skpos[u0];
		const uint32_t src = interleave_src((uint32_t) e - ls, n * rf, L, rf);
		const uint32_t run = sh.rkey[u0 + src / rf] & ((1u << kRunBits) - 1);
		const uint64_t lo = sh.lo[sh.rstart[run] + src % rf];
		const uint64_t c = (lo >> 32) | ((uint64_t) ((uint32_t) lo - cmin) << row_bits);
		for (uint32_t b = 0; b < cb; b++)
			o[coords_at + (long) e * cb + b] = (uint8_t) (c >> (8 * b));
	}
	// lane streams: the payload bytes
	if (!IL && t < kWave) {
		uint32_t incl = t < L ? sh.lane_bytes[t] : 0u;
		for (int d = 1; d < kWave; d <<= 1) {
			const uint32_t v = __shfl_up(incl, d, kWave);
			if (t >= d)
				incl += v;
		}
		sh.lane_off[t] = incl - (t < L ? sh.lane_bytes[t] : 0u);
	}
	__syncthreads();
	if (t < L) {
		long pos = 0;
		for (int g = 0; g < ngroups; g++) {
			const uint32_t u0 = sh.gq[g], n = sh.gq[g + 1] - u0;
			uint64_t prev = 0;
			for (uint32_t q = t; q < n; q += L) {
				const uint32_t run = sh.rkey[u0 + interleave_src(q, n, L, 1)] & ((1u << kRunBits) - 1);
				const uint64_t v = sh.hi[sh.rstart[run]];
				uint32_t len, tz4;
				uint64_t pay;
				delta_code(v - prev, &len, &tz4, &pay);
				prev = v;
				for (uint32_t b = 0; b < len; b++, pos++) {
					const long at = IL ? il_at + (pos >> 2) * 4L * L + 4L * t + (pos & 3)
							: streams_at + sh.lane_off[t] + pos;
					o[at] = (uint8_t) (pay >> (8 * b));
				}
			}
		}
	}
}

// Lossy values (div_kernels_rf.h:467-510, div_codec.cpp: reduce_precision):
// keep = ceil(|log2 tolerance|) fraction bits; *bad set when a rounded
// value's relative error exceeds the tolerance.
__global__ void div_reduce_precision_kernel(const double *__restrict__ in, long n, long keep, double tolerance,
		double *__restrict__ out, int *__restrict__ bad)
{
	const long k = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const double x = in[k];
	uint64_t u;
	__builtin_memcpy(&u, &x, 8);
	double r = x;
	if (u != 0 && __builtin_ctzll(u) < 53 && keep < 52) {
		const uint64_t low = (1ull << (52 - keep)) - 1;
		const uint64_t dn = u & ~low, up = (u + low) & ~low;
		double vd, vu;
		__builtin_memcpy(&vd, &dn, 8);
		__builtin_memcpy(&vu, &up, 8);
		if (fabs(x - vd) / fabs(x) > tolerance || fabs(x - vu) / fabs(x) > tolerance)
			atomicOr(bad, 1);
		const int tz_dn = dn ? __builtin_ctzll(dn) : 64, tz_up = up ? __builtin_ctzll(up) : 64;
		r = tz_dn > tz_up ? vd : vu;
	}
	out[k] = r;
}

bool div_device_encodable(long P, int L)
{
	return P >= 1 && P <= kEncMax && L >= 1 && L <= kWave;
}

DivDeviceEncoder::~DivDeviceEncoder() = default;

int DivDeviceEncoder::prepare(spmvhip_format *f, const spmvdiv::PacketPlan &plan, int L)
{
	f_ = f;
	L_ = L;
	np_ = (long) plan.i0.size();
	hipStream_t s = f->stream;
	vals_ = f->val.as<double>();
	if (f->p.div_tolerance != 0) {
		const double tol = f->p.div_tolerance;
		if (tol >= 1.0)
			return set_error(SPMVHIP_EINVAL, "div_rf: DIV_VC_TOLERANCE must be < 1 (tolerance %g)", tol);
		if (tol > 0 && f->nnz > 0) {
			DevBuf bad;
			SPMVHIP_TRY(lossy_.alloc(f->nnz * sizeof(double)));
			SPMVHIP_TRY(bad.alloc(sizeof(int)));
			SPMVHIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
			const long keep = (long) ceil(fabs(log2(tol)));
			div_reduce_precision_kernel<<<grid_for(f->nnz, kBlock), kBlock, 0, s>>>(vals_, f->nnz, keep, tol,
					lossy_.as<double>(), bad.as<int>());
			SPMVHIP_CHECK(hipGetLastError());
			int hbad = 0;
			SPMVHIP_CHECK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
			SPMVHIP_CHECK(hipStreamSynchronize(s));
			if (hbad)
				return set_error(SPMVHIP_EINVAL, "div_rf: a rounded value exceeds DIV_VC_TOLERANCE (tolerance %g)",
						tol);
			vals_ = lossy_.as<double>();
		}
	}
	SPMVHIP_TRY(i0_.alloc(np_ * sizeof(int32_t)));
	SPMVHIP_TRY(j0_.alloc(np_ * sizeof(int64_t)));
	SPMVHIP_TRY(nv_.alloc(np_ * sizeof(int32_t)));
	SPMVHIP_TRY(info_.alloc(np_ * sizeof(EncInfo)));
	SPMVHIP_TRY(size_.alloc((np_ + 1) * sizeof(int64_t)));
	SPMVHIP_CHECK(hipMemcpyAsync(i0_.p, plan.i0.data(), np_ * sizeof(int32_t), hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipMemcpyAsync(j0_.p, plan.j0.data(), np_ * sizeof(int64_t), hipMemcpyHostToDevice, s));
	SPMVHIP_CHECK(hipMemcpyAsync(nv_.p, plan.nv.data(), np_ * sizeof(int32_t), hipMemcpyHostToDevice, s));
	return SPMVHIP_OK;
}

int DivDeviceEncoder::launch(bool write, bool il, const int64_t *off, uint8_t *out)
{
	spmvhip_format *f = f_;
#define ENC_LAUNCH(W, I)                                                                                         \
	div_encode_kernel<W, I><<<(unsigned) np_, kEncThreads, 0, f->stream>>>(f->row_ptr.as<int32_t>(),            \
			f->col.as<int32_t>(), vals_, f->m, i0_.as<int32_t>(), j0_.as<int64_t>(), nv_.as<int32_t>(), L_, \
			info_.as<EncInfo>(), size_.as<int64_t>(), off, out)
	if (write) {
		if (il)
			ENC_LAUNCH(true, true);
		else
			ENC_LAUNCH(true, false);
	} else {
		if (il)
			ENC_LAUNCH(false, true);
		else
			ENC_LAUNCH(false, false);
	}
#undef ENC_LAUNCH
	SPMVHIP_CHECK(hipGetLastError());
	return SPMVHIP_OK;
}

int DivDeviceEncoder::size(bool il, int64_t *max_packet, uint32_t *max_rows)
{
	SPMVHIP_TRY(launch(false, il, nullptr, nullptr));
	std::vector<EncInfo> info(np_);
	SPMVHIP_CHECK(hipMemcpyAsync(info.data(), info_.p, np_ * sizeof(EncInfo), hipMemcpyDeviceToHost, f_->stream));
	SPMVHIP_CHECK(hipStreamSynchronize(f_->stream));
	int64_t mx = 0;
	uint32_t mr = 1;
	for (long p = 0; p < np_; p++) {
		if (info[p].err)
			return set_error(SPMVHIP_EFORMAT, "div_rf: %s (the reference rejects this matrix too)",
					info[p].err == 1 ? "row_bits > 8" : "col_bits > 32");
		mx = std::max<int64_t>(mx, info[p].bytes);
		mr = std::max<uint32_t>(mr, (uint32_t) info[p].rows);
	}
	sized_il_ = il;
	*max_packet = mx;
	*max_rows = mr;
	return SPMVHIP_OK;
}

int DivDeviceEncoder::write()
{
	spmvhip_format *f = f_;
	hipStream_t s = f->stream;
	// packet offsets: an exclusive scan of the sizes (with the total) on the device
	SPMVHIP_TRY(f->div_pk_off.alloc((np_ + 1) * sizeof(int64_t)));
	SPMVHIP_TRY((device_scan<int64_t, OpSum>(size_.as<int64_t>(), f->div_pk_off.as<int64_t>(), np_, 0, 1, s)));
	int64_t total = 0;
	SPMVHIP_CHECK(hipMemcpyAsync(&total, f->div_pk_off.as<int64_t>() + np_, sizeof(total), hipMemcpyDeviceToHost, s));
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	SPMVHIP_TRY(f->div_stream.alloc((size_t) total + 16)); // + 16: the decoder reads whole aligned words
	SPMVHIP_CHECK(hipMemsetAsync(f->div_stream.p, 0, (size_t) total + 16, s));
	SPMVHIP_TRY(launch(true, sized_il_, f->div_pk_off.as<int64_t>(), f->div_stream.as<uint8_t>()));
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	f->div_bytes = (long) total;
	return SPMVHIP_OK;
}

} // namespace spmvhip

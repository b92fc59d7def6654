// Per-wavefront SpMV work units shared by the standalone kernels (csr.hip,
// sell.hip) and the batched persistent kernel (batch.hip).
#pragma once

#include "common.hpp"

namespace spmvhip {

// Workgroups are handed to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8). With a chunk size K > 0 the work list is cut into chunks of K
// workgroups and XCD x takes chunks x, x + 8, x + 16, ...: each XCD's L2
// then holds the x lines of one contiguous range of rows at a time while
// the chunks keep the XCDs' loads even. The grid must be padded to a
// multiple of 8 K; K = 0 is the identity.
constexpr int kXcds = 8;

__device__ __forceinline__ long xcd_logical_block(int K)
{
	if (K <= 0)
		return blockIdx.x;
	const long i = blockIdx.x / kXcds, x = blockIdx.x % kXcds;
	return (i / K) * (long) kXcds * K + x * (long) K + (i % K);
}

inline unsigned xcd_padded_grid(unsigned grid, int K)
{
	if (K <= 0)
		return grid;
	const unsigned q = (unsigned) kXcds * (unsigned) K;
	return (grid + q - 1) / q * q;
}

// the remap parameter: 0 off, 1 the default chunk, n > 1 chunks of n workgroups
inline int xcd_chunk(int remap)
{
	return remap <= 0 ? 0 : (remap == 1 ? 64 : remap);
}

__device__ __forceinline__ void wave_lds_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// p[e] + p[e + L] + ... for e < b, with four LDS reads in flight per step.
__device__ __forceinline__ double lds_strided_sum(const double *p, int e, int b, int L)
{
	double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
	for (; e + 3 * L < b; e += 4 * L) {
		s0 += p[e];
		s1 += p[e + L];
		s2 += p[e + 2 * L];
		s3 += p[e + 3 * L];
	}
	for (; e < b; e += L)
		s0 += p[e];
	return (s0 + s1) + (s2 + s3);
}

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef int i32x2 __attribute__((ext_vector_type(2)));

// Two adjacent stored entries [j, j + 1] (j even) with one 8-byte column
// load and one 16-byte value load. The CSR arrays of a converted matrix carry
// a zeroed tail of kCsrPad entries, so j + 1 may equal nnz.

__device__ __forceinline__ void load_pair(const int32_t *__restrict__ col, const double *__restrict__ val, int j,
		i32x2 &c, f64x2 &v)
{
	c = __builtin_nontemporal_load((const i32x2 *) (col + j));
	v = __builtin_nontemporal_load((const f64x2 *) (val + j));
}

// y[r] = sum of row r over [j0, j1), one wave. Lane l takes the entry pairs
// starting at the even index jb + 2 l, jb + 2 l + 128, ... (jb = j0 rounded
// down to even), entries outside [j0, j1) masked; two pairs in flight.
__device__ __forceinline__ double wave_row_sum_pairs(const int32_t *__restrict__ col,
		const double *__restrict__ val, const double *__restrict__ x, int j0, int j1)
{
	const int lane = lane_id();
	double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
	int j = (j0 & ~1) + 2 * lane;
	for (; j + 2 * kWave + 1 < j1; j += 4 * kWave) {
		i32x2 c0, c1;
		f64x2 v0, v1;
		load_pair(col, val, j, c0, v0);
		load_pair(col, val, j + 2 * kWave, c1, v1);
		const double x0 = x[c0.x], x1 = x[c0.y], x2 = x[c1.x], x3 = x[c1.y];
		a0 = j >= j0 ? fma(v0.x, x0, a0) : a0;
		a1 = fma(v0.y, x1, a1);
		a2 = fma(v1.x, x2, a2);
		a3 = fma(v1.y, x3, a3);
	}
	for (; j < j1; j += 2 * kWave) {
		i32x2 c;
		f64x2 v;
		load_pair(col, val, j, c, v);
		const double x0 = x[c.x], x1 = x[c.y];
		a0 = j >= j0 ? fma(v.x, x0, a0) : a0;
		a1 = j + 1 < j1 ? fma(v.y, x1, a1) : a1;
	}
	return wave_sum((a0 + a1) + (a2 + a3));
}

// y[r] = sum of row r over [j0, j1), one wave, lanes striding the row.
__device__ __forceinline__ double wave_row_sum(const int32_t *__restrict__ col, const double *__restrict__ val,
		const double *__restrict__ x, int j0, int j1)
{
	const int lane = lane_id();
	double a0 = 0, a1 = 0;
	int j = j0 + lane;
	for (; j + kWave < j1; j += 2 * kWave) {
		a0 = fma(val[j], x[col[j]], a0);
		a1 = fma(val[j + kWave], x[col[j + kWave]], a1);
	}
	if (j < j1)
		a0 = fma(val[j], x[col[j]], a0);
	return wave_sum(a0 + a1);
}

// Second phase of a CSR row block: p[0, nz) holds the block's products in
// entry order, lane i's `ra` the start of row i relative to the block
// (rows past the block hold nz); yb[r] = sum of row r for r < nrows.
// Lanes per row L = 2^lg: the largest power of two <= 64 / nrows (or the
// configured value); groups of L lanes take rows g, g + 64/L, ...
//
// The first row's bounds are shuffled before the (divergent) row loop, so
// that every path through the block consumes the row-start load: a load
// left pending on some path makes the compiler wait for all of the wave's
// memory operations, its y stores included, at the next reuse of the
// load's register -- in a persistent loop, at the start of the next block.
__device__ __forceinline__ void block_rows_reduce(const double *p, int ra, int nrows, int nz, int lanes_cfg,
		double *__restrict__ yb)
{
	const int lane = lane_id();
	int lg = 6;
	while (lg > 0 && (nrows << lg) > kWave)
		lg--;
	if (lanes_cfg > 0 && lanes_cfg < (1 << lg))
		lg = 31 - __builtin_clz(lanes_cfg);
	const int L = 1 << lg, g = lane >> lg, sub = lane & (L - 1), gshift = 6 - lg;
	const int rows_up = (nrows + (1 << gshift) - 1) >> gshift << gshift;
	int an = __shfl(ra, g, kWave), bnn = __shfl(ra, (g + 1) & (kWave - 1), kWave);
	for (int r = g; r < rows_up; r += 1 << gshift) {
		const int a = an, bn = bnn;
		const int rn = r + (1 << gshift);
		an = __shfl(ra, rn & (kWave - 1), kWave);
		bnn = __shfl(ra, (rn + 1) & (kWave - 1), kWave);
		const int b = r + 1 < kWave ? bn : nz;
		double s = r < nrows ? lds_strided_sum(p, a + sub, b, L) : 0.0;
		s = group_sum(s, L);
		if (r < nrows && sub == 0)
			yb[r] = s;
	}
	wave_lds_sync(); // the buffer is reused by the wave's next block
}

// One stream step of csr_block_at's paired layout: U halves of 128 entries
// from the even index b (lane l takes the pairs at b + 2 l and, for U = 2,
// b + 2 l + 128); lanes past the block re-read its last pair jl; products
// of entries in [j0, j1) go to p[j - j0].
template <int U>
__device__ __forceinline__ void csr_pair_step(const int32_t *__restrict__ col, const double *__restrict__ val,
		const double *__restrict__ x, int b, int j0, int j1, int jl, double *p)
{
	const int lane = lane_id();
	i32x2 c[U];
	f64x2 v[U];
#pragma unroll
	for (int u = 0; u < U; u++) {
		const int j = b + u * 2 * kWave + 2 * lane;
		load_pair(col, val, j < jl ? j : jl, c[u], v[u]);
	}
	double xv[2 * U];
#pragma unroll
	for (int u = 0; u < U; u++) {
		xv[2 * u] = x[c[u].x];
		xv[2 * u + 1] = x[c[u].y];
	}
#pragma unroll
	for (int u = 0; u < U; u++) {
		const int j = b + u * 2 * kWave + 2 * lane;
		if (j >= j0 && j < j1)
			p[j - j0] = v[u].x * xv[2 * u];
		if (j + 1 >= j0 && j + 1 < j1)
			p[j + 1 - j0] = v[u].y * xv[2 * u + 1];
	}
}

// One row block [r0, r1) of the binned CSR format (see csr.hip), processed
// by the calling wavefront. `p` is the wave's LDS buffer of 2 * BNNZ_MAX
// doubles. A single row longer than BNNZ_MAX is summed by the whole wave;
// with skip_split, rows longer than split_nnz are left to the split kernels.
//
// The block bounds are wave-uniform (scalar registers), so the control below
// runs on the scalar unit. Lane i loads the bounds of row r0 + i before the
// stream phase, so the reduction does not wait for a second round of
// row-pointer loads. The stream phase has no per-lane branches: 4 x 64
// entries are requested per step (in the paired layout, a last step with at
// most 128 entries left requests only those), lanes past the block's end
// re-read its last entry, so a block of <= 256 nonzeros costs one index
// round trip and one gather round trip.
//
// csr_block_at takes the block's nonzero range [j0, j1) = [rp[r0], rp[r1])
// already known (a precomputed work-unit record); csr_block loads it.
//
// Stream layouts: NARROW, one entry per lane and load (entry e0 + 64 u +
// lane); otherwise two adjacent entries per lane and load, lane l taking the
// pairs at b + 2 l and b + 2 l + 128 of each 256-entry step from the even
// index b = j0 rounded down (a 16-byte value load and an 8-byte column load
// per pair: 8 vector-memory instructions per 256 entries instead of 12, the
// address unit's work per nonzero cut by a third). Lanes past the block
// re-read its last pair; products outside [j0, j1) are not stored.
template <int BNNZ_MAX, bool NARROW = false>
__device__ __forceinline__ void csr_block_at(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
		const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y, int r0, int r1,
		int j0, int j1, int split_nnz, bool skip_split, int lanes_cfg, double *p)
{
	static_assert(BNNZ_MAX % 128 == 0, "the stream phase writes p in steps of 256 entries");
	const int lane = lane_id();
	r0 = __builtin_amdgcn_readfirstlane(r0);
	r1 = __builtin_amdgcn_readfirstlane(r1);
	j0 = __builtin_amdgcn_readfirstlane(j0);
	j1 = __builtin_amdgcn_readfirstlane(j1);
	const int nrows = r1 - r0, nz = j1 - j0;
	if (nrows == 1 && nz > BNNZ_MAX) {
		if (skip_split && nz > split_nnz)
			return;
		double sum = NARROW ? wave_row_sum(col, val, x, j0, j1) : wave_row_sum_pairs(col, val, x, j0, j1);
		if (lane == 0)
			y[r0] = sum;
		return;
	}
	// start of row r0 + lane relative to the block start (nrows <= 64; lanes
	// past the block read its end): one unconditional load, in flight during
	// the stream phase; row r ends where row r + 1 starts
	const int ra = rp[r0 + (lane < nrows ? lane : nrows)] - j0;
	if (!NARROW) {
		// stream phase, entry pairs: p[j - j0] for j in [j0, j1) (nz < 2 *
		// BNNZ_MAX for a multi-row block, so p stays inside the buffer).
		// Steps of 256 entries (both halves) while more than 128 remain,
		// then one step of 128: a block of about 256 entries that spills
		// past a 256-entry step costs 4 more vector-memory instructions,
		// not 8 (the branch is wave-uniform)
		const int jl = (j1 - 1) & ~1; // the block's last pair (used only when j1 > b >= 0)
		for (int b = j0 & ~1; b < j1;) {
			if (j1 - b > 2 * kWave) {
				csr_pair_step<2>(col, val, x, b, j0, j1, jl, p);
				b += 4 * kWave;
			} else {
				csr_pair_step<1>(col, val, x, b, j0, j1, jl, p);
				b += 2 * kWave;
			}
		}
		wave_lds_sync();
		block_rows_reduce(p, ra, nrows, nz, lanes_cfg, y + r0);
		return;
	}
	// stream phase: products to LDS, coalesced over the block's nnz (nz < 2 *
	// BNNZ_MAX for a multi-row block, so p[e] stays inside the buffer)
	for (int e0 = 0; e0 < nz; e0 += 4 * kWave) {
		double v[4];
		int32_t c[4];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int e = e0 + u * kWave + lane;
			const int ec = e < nz ? e : nz - 1;
			v[u] = __builtin_nontemporal_load(val + j0 + ec);
			c[u] = __builtin_nontemporal_load(col + j0 + ec);
		}
		double xv[4];
#pragma unroll
		for (int u = 0; u < 4; u++)
			xv[u] = x[c[u]];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int e = e0 + u * kWave + lane;
			p[e] = e < nz ? v[u] * xv[u] : 0.0;
		}
	}
	wave_lds_sync();
	block_rows_reduce(p, ra, nrows, nz, lanes_cfg, y + r0);
}

template <int BNNZ_MAX, bool NARROW = false>
__device__ __forceinline__ void csr_block(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
		const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y, int r0, int r1,
		int split_nnz, bool skip_split, int lanes_cfg, double *p)
{
	r0 = __builtin_amdgcn_readfirstlane(r0);
	r1 = __builtin_amdgcn_readfirstlane(r1);
	csr_block_at<BNNZ_MAX, NARROW>(rp, col, val, x, y, r0, r1, rp[r0], rp[r1], split_nnz, skip_split, lanes_cfg,
			p);
}

// The same row block split in three stages, so that a persistent wave can
// request block i + 1's indices and values (csr_stage_load) after block i's
// gathers (csr_stage_gather) and before block i's products, LDS reduction
// and stores (csr_stage_finish). Loads complete in issue order, so this
// order lets the wave wait for block i's gathers while block i + 1's stream
// is still in flight. The arithmetic is csr_block's, entry for entry.
struct CsrStage {
	int r0, nrows, j0, nz, ra;
	bool long_row; // one row longer than BNNZ_MAX: summed by wave_row_sum
	double v[4];
	int32_t c[4];
};

template <int BNNZ_MAX>
__device__ __forceinline__ void csr_stage_load(CsrStage &st, const int32_t *__restrict__ rp,
		const int32_t *__restrict__ col, const double *__restrict__ val, int r0, int r1)
{
	const int lane = lane_id();
	st.r0 = __builtin_amdgcn_readfirstlane(r0);
	r1 = __builtin_amdgcn_readfirstlane(r1);
	st.j0 = __builtin_amdgcn_readfirstlane(rp[st.r0]);
	st.nz = __builtin_amdgcn_readfirstlane(rp[r1]) - st.j0;
	st.nrows = r1 - st.r0;
	st.long_row = st.nrows == 1 && st.nz > BNNZ_MAX;
	if (st.long_row)
		return;
	st.ra = rp[st.r0 + (lane < st.nrows ? lane : st.nrows)] - st.j0;
#pragma unroll
	for (int u = 0; u < 4; u++) {
		const int e = u * kWave + lane;
		const int ec = e < st.nz ? e : st.nz - 1;
		st.v[u] = __builtin_nontemporal_load(val + st.j0 + ec);
		st.c[u] = __builtin_nontemporal_load(col + st.j0 + ec);
	}
}

__device__ __forceinline__ void csr_stage_gather(const CsrStage &st, const double *__restrict__ x, double xv[4])
{
	if (st.long_row)
		return;
#pragma unroll
	for (int u = 0; u < 4; u++)
		xv[u] = x[st.c[u]];
}

template <int BNNZ_MAX>
__device__ __forceinline__ void csr_stage_finish(const CsrStage &st, const double xv[4],
		const int32_t *__restrict__ col, const double *__restrict__ val, const double *__restrict__ x,
		double *__restrict__ y, int split_nnz, bool skip_split, int lanes_cfg, double *p)
{
	const int lane = lane_id();
	const int nz = st.nz, nrows = st.nrows, j0 = st.j0, r0 = st.r0;
	if (st.long_row) {
		if (skip_split && nz > split_nnz)
			return;
		double sum = wave_row_sum(col, val, x, j0, j0 + nz);
		if (lane == 0)
			y[r0] = sum;
		return;
	}
#pragma unroll
	for (int u = 0; u < 4; u++) {
		const int e = u * kWave + lane;
		p[e] = e < nz ? st.v[u] * xv[u] : 0.0;
	}
	// entries past the first 256 (blocks hold fewer than 2 * BNNZ_MAX)
	for (int e0 = 4 * kWave; e0 < nz; e0 += 4 * kWave) {
		double v[4], xw[4];
		int32_t c[4];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int e = e0 + u * kWave + lane;
			const int ec = e < nz ? e : nz - 1;
			v[u] = __builtin_nontemporal_load(val + j0 + ec);
			c[u] = __builtin_nontemporal_load(col + j0 + ec);
		}
#pragma unroll
		for (int u = 0; u < 4; u++)
			xw[u] = x[c[u]];
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int e = e0 + u * kWave + lane;
			p[e] = e < nz ? v[u] * xw[u] : 0.0;
		}
	}
	wave_lds_sync();
	block_rows_reduce(p, st.ra, nrows, nz, lanes_cfg, y + r0);
}

// x[c] for a SELL entry. With XW the workgroup has copied x[lo, lo + W) to
// xs (LDS): columns in that window are read from LDS, any other column (a
// padding entry's, or all of them when W = 0) from x.
template <bool XW>
__device__ __forceinline__ double gather_x(const double *__restrict__ x, const double *xs, int lo, unsigned W, int c)
{
	if (XW) {
		const unsigned i = (unsigned) (c - lo);
		if (i < W)
			return xs[i];
	}
	return x[c];
}

// N consecutive wave steps k..k+N-1 of a SELL slice (v, c point at the
// lane's entry of step 0): all loads issued before the gathers, the FMAs in
// entry order.
template <int N, bool XW>
__device__ __forceinline__ double sell_group(const double *__restrict__ v, const int32_t *__restrict__ c,
		const double *__restrict__ x, const double *xs, int lo, unsigned W, int k, double acc)
{
	double vv[N], xx[N];
	int32_t cc[N];
#pragma unroll
	for (int u = 0; u < N; u++) {
		vv[u] = __builtin_nontemporal_load(v + (long) (k + u) * kWave);
		cc[u] = __builtin_nontemporal_load(c + (long) (k + u) * kWave);
	}
#pragma unroll
	for (int u = 0; u < N; u++)
		xx[u] = gather_x<XW>(x, xs, lo, W, cc[u]);
#pragma unroll
	for (int u = 0; u < N; u++)
		acc = fma(vv[u], xx[u], acc);
	return acc;
}

// One SELL slice of 64 rows with T lanes per row (T a power of two; the
// slice is processed in T rounds of 64 / T rows). Every wave step reads 64
// consecutive stored entries: with T = 1 these are entry k of all 64 rows
// (the reference's column-major slice), with T > 1 entries jT..jT+T-1 of
// 64 / T rows, so x is gathered in runs of T along each row. Each lane keeps
// one fused multiply-add chain in entry order; UNROLL entries are loaded
// ahead of their FMAs; the T partial sums of a row are added at the end.
// With ystage != nullptr the results go to ystage[target - row0] (LDS)
// instead of y[target]: the caller writes the staged rows out coalesced.
//
// With PIPE the column indices and values of the next UNROLL entries are
// requested before the x gathers of the current ones, so each step waits for
// one memory round trip (the gathers) instead of two (indices, then x).
// With XW, x is read through gather_x from the workgroup's LDS window.
template <int UNROLL, bool PIPE = false, bool XW = false>
__device__ __forceinline__ void sell_slice(const int64_t *__restrict__ sptr, const int32_t *__restrict__ width,
		const int32_t *__restrict__ scol, const double *__restrict__ sval, const int32_t *__restrict__ order,
		const double *__restrict__ x, double *__restrict__ y, long m, long s, int T, double *ystage = nullptr,
		long row0 = 0, const double *xs = nullptr, int lo = 0, unsigned W = 0)
{
	constexpr int C = kWave;
	const int lane = lane_id();
	const int w = width[s];
	const int wp = (w + T - 1) / T * T;
	const int steps = wp / T, R = C / T;
	for (int q = 0; q < T; q++) {
		const long off = sptr[s] + (long) q * R * wp + lane;
		const double *v = sval + off;
		const int32_t *c = scol + off;
		double acc = 0;
		int k = 0;
		if (PIPE && steps >= UNROLL) {
			double vv[UNROLL];
			int32_t cc[UNROLL];
#pragma unroll
			for (int u = 0; u < UNROLL; u++) {
				cc[u] = __builtin_nontemporal_load(c + (long) u * C);
				vv[u] = __builtin_nontemporal_load(v + (long) u * C);
			}
			for (k = UNROLL; k + UNROLL <= steps; k += UNROLL) {
				double xx[UNROLL];
#pragma unroll
				for (int u = 0; u < UNROLL; u++)
					xx[u] = gather_x<XW>(x, xs, lo, W, cc[u]);
				double vn[UNROLL];
				int32_t cn[UNROLL];
#pragma unroll
				for (int u = 0; u < UNROLL; u++) {
					cn[u] = __builtin_nontemporal_load(c + (long) (k + u) * C);
					vn[u] = __builtin_nontemporal_load(v + (long) (k + u) * C);
				}
#pragma unroll
				for (int u = 0; u < UNROLL; u++) {
					acc = fma(vv[u], xx[u], acc);
					cc[u] = cn[u];
					vv[u] = vn[u];
				}
			}
#pragma unroll
			for (int u = 0; u < UNROLL; u++)
				acc = fma(vv[u], gather_x<XW>(x, xs, lo, W, cc[u]), acc);
		}
		for (; !PIPE && k + UNROLL <= steps; k += UNROLL) {
			double vv[UNROLL];
			int32_t cc[UNROLL];
#pragma unroll
			for (int u = 0; u < UNROLL; u++) {
				vv[u] = __builtin_nontemporal_load(v + (long) (k + u) * C);
				cc[u] = __builtin_nontemporal_load(c + (long) (k + u) * C);
			}
			double xx[UNROLL];
#pragma unroll
			for (int u = 0; u < UNROLL; u++)
				xx[u] = gather_x<XW>(x, xs, lo, W, cc[u]);
#pragma unroll
			for (int u = 0; u < UNROLL; u++)
				acc = fma(vv[u], xx[u], acc);
		}
		// the last steps - k < UNROLL entries in groups of UNROLL/2, UNROLL/4,
		// ..., 1 (steps is uniform over the wave): each group's loads are in
		// flight together
		if (UNROLL >= 8 && k + 4 <= steps) {
			acc = sell_group<4, XW>(v, c, x, xs, lo, W, k, acc);
			k += 4;
		}
		if (UNROLL >= 4 && k + 2 <= steps) {
			acc = sell_group<2, XW>(v, c, x, xs, lo, W, k, acc);
			k += 2;
		}
		if (k < steps)
			acc = sell_group<1, XW>(v, c, x, xs, lo, W, k, acc);
		if (T > 1)
			acc = group_sum(acc, T);
		const long r = s * C + (long) q * R + lane / T;
		if ((lane & (T - 1)) == 0 && r < m) {
			const long tgt = order ? order[r] : r;
			if (ystage)
				ystage[tgt - row0] = acc;
			else
				y[tgt] = acc;
		}
	}
}

} // namespace spmvhip

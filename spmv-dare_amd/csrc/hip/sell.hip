// SELL, SELL-C-sigma and SELL_SORTED on gfx950, slice height C = 64 = one
// wavefront.
//
// Layout in HBM (column-major inside a slice, as the reference's
// csr_to_sellcs writes it, src/spmv_kernels/sell-C-s/sellcs_format.c:173-186):
//   entry k of slice row l lives at slice_ptr[s] + k*64 + l, so step k of a
//   slice is one 512-byte coalesced load of values and a 256-byte load of
//   column indices for the whole wave.
//
// csr_to_format runs on the device:
//   SELL-C-sigma (sell_c_s.cpp:39-77 -> radix_sort.c:103-122, sellcs_format.c:137-200)
//     1. row lengths;
//     2. per sigma-window stable sort by descending length (ties keep the
//        ascending row id) -- an LSD radix sort on 4-bit digits in LDS, one
//        workgroup per window; the permutation equals the reference's;
//     3. slice widths = longest row of the slice (the reference takes the
//        first sorted row, which is the longest whenever sigma % 64 == 0);
//     4. slice pointers = exclusive scan of width * 64;
//     5. fill: one wave per slice, padding value 0.0 / column 0 (the
//        reference memsets both arrays).
//   SELL (sell.cpp:105-199, slice = 64 rows, no sort): width = longest row of
//   the slice; padding value 0 with the column of the previous entry of the
//   row, i.e. the row's last column, or for an empty row the last column of
//   the nearest non-empty row above (0 if none) -- the reference with one
//   OpenMP thread (its per-thread `col` restarts at 0 for each thread).
//   SELL_SORTED (sell_sorted.cpp:119-316): inside each window of sigma rows
//   (a multiple of 64; the reference's windows are its threads' nnz-balanced
//   ranges of slices) the rows are stable-sorted by ascending length -- the
//   same LDS radix sort with the digit order reversed; width = longest row
//   of the slice (its last); padding value 0 with the column of the previous
//   entry written in the window: the row's own last column, 0 for the empty
//   rows that open a window, and for the missing rows of the last slice the
//   last column of the last sorted row.
//
// spmv: one wave per slice, lane = row, one fused multiply-add per entry in
// entry order (the reference's vfmacc, sellcs_mv_kernels_epi.c:292-297), then
// y[row_order[row]] = acc for the sorted formats and y[row] = acc for SELL. Only
// rows < m are written.

#include "format.hpp"
#include "kernels.hpp"
#include "scan.hpp"

namespace spmvhip {

constexpr int kSliceC = 64;
constexpr int kMaxSigma = 16384;
constexpr int kSortThreads = 256;
// x window staged in LDS per workgroup (doubles): default and largest
constexpr int kXWindowDefault = 2048;
constexpr int kXWindowMax = 16384;

__global__ void row_lengths(const int32_t *__restrict__ rp, long m, int32_t *__restrict__ len)
{
	long r = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (r < m)
		len[r] = rp[r + 1] - rp[r];
}

// Stable sort of one window of row lengths, descending (SELL-C-sigma) or
// ascending (SELL_SORTED); writes the row ids. Ties keep the ascending row id.
// LDS: two id arrays of `sigma` ints plus a 16 x 256 digit-count table.
__global__ void __launch_bounds__(kSortThreads) sigma_window_sort(const int32_t *__restrict__ len, long m, int sigma,
		int ascending, int32_t *__restrict__ order)
{
	extern __shared__ int32_t lds[];
	int32_t *ids = lds;
	int32_t *tmp = lds + sigma;
	int32_t *cnt = lds + 2 * sigma; // [16][kSortThreads]
	__shared__ int32_t red[kSortThreads / kWave];
	const long w0 = (long) blockIdx.x * sigma;
	const int wl = (int) (m - w0 < sigma ? m - w0 : sigma);
	const int t = threadIdx.x;
	const int q = (wl + kSortThreads - 1) / kSortThreads;
	const int i0 = t * q < wl ? t * q : wl;
	const int i1 = (t + 1) * q < wl ? (t + 1) * q : wl;

	int32_t mx = 0;
	for (int i = t; i < wl; i += kSortThreads) {
		ids[i] = (int32_t) (w0 + i);
		int32_t k = len[w0 + i];
		mx = k > mx ? k : mx;
	}
	for (int off = 32; off > 0; off >>= 1) {
		int32_t o = __shfl_xor(mx, off, kWave);
		mx = o > mx ? o : mx;
	}
	if (lane_id() == 0)
		red[t / kWave] = mx;
	__syncthreads();
	mx = 0;
	for (int wv = 0; wv < kSortThreads / kWave; wv++)
		mx = red[wv] > mx ? red[wv] : mx;
	int passes = 0;
	for (uint32_t v = (uint32_t) mx; v; v >>= 4)
		passes++;

	for (int pass = 0, shift = 0; pass < passes; pass++, shift += 4) {
		int c[16];
#pragma unroll
		for (int d = 0; d < 16; d++)
			c[d] = 0;
		for (int i = i0; i < i1; i++)
			c[(len[ids[i]] >> shift) & 15]++;
		// table in scan order: digit 15 first (descending) or digit 0 first
		// (ascending), then thread id
#pragma unroll
		for (int d = 0; d < 16; d++)
			cnt[(ascending ? d : 15 - d) * kSortThreads + t] = c[d];
		__syncthreads();
		// exclusive scan of the 16*256 table: thread t scans entries [16t, 16t+16)
		int run = 0;
		int loc[16];
#pragma unroll
		for (int k = 0; k < 16; k++) {
			loc[k] = run;
			run += cnt[16 * t + k];
		}
		int ex, tot;
		block_scan<int, OpSum>(run, ex, tot);
#pragma unroll
		for (int k = 0; k < 16; k++)
			cnt[16 * t + k] = ex + loc[k];
		__syncthreads();
#pragma unroll
		for (int d = 0; d < 16; d++)
			c[d] = cnt[(ascending ? d : 15 - d) * kSortThreads + t];
		for (int i = i0; i < i1; i++) {
			int32_t id = ids[i];
			int d = (len[id] >> shift) & 15;
			tmp[c[d]++] = id;
		}
		__syncthreads();
		int32_t *sw = ids;
		ids = tmp;
		tmp = sw;
	}
	for (int i = t; i < wl; i += kSortThreads)
		order[w0 + i] = ids[i];
}

// Slice widths (longest row) and slice sizes: width rounded up to a multiple
// of the lanes per row T, times 64.
__global__ void slice_widths(const int32_t *__restrict__ len, const int32_t *__restrict__ order, long m, long nslices,
		int T, int32_t *__restrict__ width, int64_t *__restrict__ size)
{
	long s = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= nslices)
		return;
	int w = 0;
	for (int l = 0; l < kSliceC; l++) {
		long r = s * kSliceC + l;
		if (r >= m)
			break;
		int k = len[order ? order[r] : r];
		w = k > w ? k : w;
	}
	width[s] = w;
	size[s] = (int64_t) ((w + T - 1) / T * T) * kSliceC;
}

// last_ne[r] = largest r' <= r with a non-empty row r' (or -1): input of the
// SELL padding rule. Produced by an inclusive max-scan of these keys.
__global__ void nonempty_keys(const int32_t *__restrict__ len, long m, int64_t *__restrict__ key)
{
	long r = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (r < m)
		key[r] = len[r] > 0 ? r : -1;
}

__global__ void __launch_bounds__(kBlock) sell_fill(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
		const double *__restrict__ val, const int32_t *__restrict__ order, const int64_t *__restrict__ last_ne,
		const int32_t *__restrict__ width, const int64_t *__restrict__ sptr, long m, long nslices, int T,
		int ascending, int32_t *__restrict__ scol, double *__restrict__ sval)
{
	const long s = (long) blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
	if (s >= nslices)
		return;
	const int lane = lane_id();          // the slice row this lane fills
	const long r = s * kSliceC + lane;
	const int w = width[s];
	const int wp = (w + T - 1) / T * T; // width padded to the lanes per row
	const int R = kSliceC / T;          // rows per round
	// entry k of slice row `lane`: round lane / R, lane group lane % R, step k / T
	const int64_t base = sptr[s] + (int64_t) (lane / R) * R * wp + (lane % R) * T;
	int j0 = 0, len = 0, pad_col = 0;
	if (r < m) {
		long src = order ? order[r] : r;
		j0 = rp[src];
		len = rp[src + 1] - j0;
	}
	if (!order) {
		// SELL padding column: last column of this row or of the nearest non-empty row above
		long rr = r < m ? r : m - 1;
		int64_t ne = m > 0 ? last_ne[rr] : -1;
		pad_col = ne >= 0 ? col[rp[ne + 1] - 1] : 0;
	} else if (ascending) {
		if (len > 0) {
			pad_col = col[j0 + len - 1];
		} else if (r >= m) { // missing rows of the last slice: the last sorted row's last column
			const long last = order[m - 1];
			if (rp[last + 1] > rp[last])
				pad_col = col[rp[last + 1] - 1];
		}
	} else if (T > 1 && len > 0) {
		pad_col = col[j0 + len - 1]; // keeps the padding's x reads on the row's own lines
	}
	for (int k = 0; k < wp; k++) {
		int64_t idx = base + (int64_t) (k / T) * kSliceC + (k % T);
		if (k < len) {
			scol[idx] = col[j0 + k];
			sval[idx] = val[j0 + k];
		} else {
			scol[idx] = pad_col;
			sval[idx] = 0.0;
		}
	}
}

// inv[order[i]] = i: the sorted position of every row
__global__ void invert_order(const int32_t *__restrict__ order, long m, int32_t *__restrict__ inv)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i < m)
		inv[order[i]] = (int32_t) i;
}

// y[r] = ys[inv[r]]: results stored in sorted order back to row order (the
// reads stay inside each row's sigma window, so they mostly hit L2)
__global__ void unpermute_y(const double *__restrict__ ys, const int32_t *__restrict__ inv, long m,
		double *__restrict__ y)
{
	const long r = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (r < m)
		y[r] = ys[inv[r]];
}

// x window of each workgroup's 256 stored rows (4 slices): the smallest and
// largest column of their entries. win[2b], win[2b + 1] = (lo, width) when
// width <= cap, else (0, 0) and workgroup b reads x from global memory.
__global__ void __launch_bounds__(kBlock) window_ranges(const int32_t *__restrict__ rp,
		const int32_t *__restrict__ col, const int32_t *__restrict__ order, long m, int cap,
		int32_t *__restrict__ win)
{
	__shared__ int lo_w[kWavesPerBlock], hi_w[kWavesPerBlock];
	const long r = (long) blockIdx.x * kBlock + threadIdx.x;
	int lo = INT_MAX, hi = -1;
	if (r < m) {
		const long src = order ? order[r] : r;
		for (int j = rp[src]; j < rp[src + 1]; j++) {
			const int c = col[j];
			lo = c < lo ? c : lo;
			hi = c > hi ? c : hi;
		}
	}
	for (int off = kWave / 2; off > 0; off >>= 1) {
		const int l2 = __shfl_xor(lo, off, kWave), h2 = __shfl_xor(hi, off, kWave);
		lo = l2 < lo ? l2 : lo;
		hi = h2 > hi ? h2 : hi;
	}
	if (lane_id() == 0) {
		lo_w[threadIdx.x / kWave] = lo;
		hi_w[threadIdx.x / kWave] = hi;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < kWavesPerBlock; w++) {
			lo = lo_w[w] < lo ? lo_w[w] : lo;
			hi = hi_w[w] > hi ? hi_w[w] : hi;
		}
		const long width = (long) hi - lo + 1;
		const bool ok = hi >= 0 && width <= cap;
		win[2 * blockIdx.x] = ok ? lo : 0;
		win[2 * blockIdx.x + 1] = ok ? (int) width : 0;
	}
}

int sell_convert(spmvhip_format *f, SellOrder ord)
{
	const bool sorted = ord != kSellUnsorted;
	hipStream_t s = f->stream;
	const long m = f->m;
	if (f->p.C != 0 && f->p.C != kSliceC)
		return set_error(SPMVHIP_EINVAL, "sell: C = %d unsupported (slice height is one 64-lane wavefront)", f->p.C);
	long sigma = f->p.sigma > 0 ? f->p.sigma : 4096;
	if (sorted && sigma > kMaxSigma)
		return set_error(SPMVHIP_EINVAL, "%s: sigma = %ld > %d (window must fit one workgroup's LDS)",
				f->name, sigma, kMaxSigma);
	if (ord == kSellAscending && sigma % kSliceC != 0)
		return set_error(SPMVHIP_EINVAL, "sell_sorted: the row window sigma = %ld must be a multiple of %d "
				"(whole slices, as the reference's thread ranges)", sigma, kSliceC);
	const int T = f->p.lanes_per_row > 0 ? f->p.lanes_per_row : 1;
	if (T > kSliceC || (T & (T - 1)))
		return set_error(SPMVHIP_EINVAL, "sell: lanes_per_row %d is not a power of two in [1, 64]", T);
	f->lanes = T;
	f->p.C = kSliceC;
	f->p.sigma = sorted ? sigma : 1;
	f->nslices = (m + kSliceC - 1) / kSliceC;
	const int ym = f->p.y_mode;
	if (ym < 0 || ym > 3)
		return set_error(SPMVHIP_EINVAL, "sell: y_mode %d (0 auto, 1 direct, 2 LDS-staged, 3 sorted buffer)", ym);
	if (ym == 2 && sorted && (kWavesPerBlock * kSliceC) % sigma != 0)
		return set_error(SPMVHIP_EINVAL, "sell_c_s: y_mode 2 needs sigma dividing 256 (got %ld)", sigma);

	DevBuf len, sizes;
	SPMVHIP_TRY(len.alloc((m + 1) * sizeof(int32_t)));
	SPMVHIP_TRY(sizes.alloc((f->nslices + 1) * sizeof(int64_t)));
	row_lengths<<<grid_for(m, kBlock), kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), m, len.as<int32_t>());
	SPMVHIP_CHECK(hipGetLastError());

	const int32_t *order = nullptr;
	if (sorted) {
		SPMVHIP_TRY(f->row_order.alloc(m * sizeof(int32_t)));
		long nwin = (m + sigma - 1) / sigma;
		size_t lds = (size_t) (2 * sigma + 16 * kSortThreads) * sizeof(int32_t);
		SPMVHIP_CHECK(hipFuncSetAttribute((const void *) sigma_window_sort,
				hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
		sigma_window_sort<<<(unsigned) nwin, kSortThreads, lds, s>>>(len.as<int32_t>(), m, (int) sigma,
				ord == kSellAscending, f->row_order.as<int32_t>());
		SPMVHIP_CHECK(hipGetLastError());
		order = f->row_order.as<int32_t>();
		if (ym == 3) {
			SPMVHIP_TRY(f->row_inv.alloc(m * sizeof(int32_t)));
			SPMVHIP_TRY(f->ysorted.alloc(f->nslices * kSliceC * sizeof(double)));
			invert_order<<<grid_for(m, kBlock), kBlock, 0, s>>>(order, m, f->row_inv.as<int32_t>());
			SPMVHIP_CHECK(hipGetLastError());
		}
	}
	SPMVHIP_TRY(f->slice_w.alloc(f->nslices * sizeof(int32_t)));
	SPMVHIP_TRY(f->slice_ptr.alloc((f->nslices + 1) * sizeof(int64_t)));
	slice_widths<<<grid_for(f->nslices, kBlock), kBlock, 0, s>>>(len.as<int32_t>(), order, m, f->nslices, T,
			f->slice_w.as<int32_t>(), sizes.as<int64_t>());
	SPMVHIP_CHECK(hipGetLastError());
	SPMVHIP_TRY((device_scan<int64_t, OpSum>(sizes.as<int64_t>(), f->slice_ptr.as<int64_t>(), f->nslices, 0, 1, s)));
	int64_t total = 0;
	SPMVHIP_CHECK(hipMemcpyAsync(&total, f->slice_ptr.as<int64_t>() + f->nslices, sizeof(total),
			hipMemcpyDeviceToHost, s));
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	f->padded = total;

	DevBuf last_ne;
	if (!sorted) {
		DevBuf keys;
		SPMVHIP_TRY(keys.alloc(m * sizeof(int64_t)));
		SPMVHIP_TRY(last_ne.alloc(m * sizeof(int64_t)));
		nonempty_keys<<<grid_for(m, kBlock), kBlock, 0, s>>>(len.as<int32_t>(), m, keys.as<int64_t>());
		SPMVHIP_CHECK(hipGetLastError());
		SPMVHIP_TRY((device_scan<int64_t, OpMax>(keys.as<int64_t>(), last_ne.as<int64_t>(), m, 1, 0, s)));
	}
	SPMVHIP_TRY(f->s_cols.alloc(total * sizeof(int32_t)));
	SPMVHIP_TRY(f->s_vals.alloc(total * sizeof(double)));
	sell_fill<<<grid_for(f->nslices, kWavesPerBlock), kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), f->col.as<int32_t>(),
			f->val.as<double>(), order, sorted ? nullptr : last_ne.as<int64_t>(), f->slice_w.as<int32_t>(),
			f->slice_ptr.as<int64_t>(), m, f->nslices, T, ord == kSellAscending, f->s_cols.as<int32_t>(),
			f->s_vals.as<double>());
	SPMVHIP_CHECK(hipGetLastError());
	f->x_cap = 0;
	if (f->p.x_window > 0) {
		const int cap = f->p.x_window == 1 ? kXWindowDefault : f->p.x_window;
		if (cap > kXWindowMax)
			return set_error(SPMVHIP_EINVAL, "sell: x_window %d doubles exceeds the LDS limit of %d", cap,
					kXWindowMax);
		const long nwg = (f->nslices + kWavesPerBlock - 1) / kWavesPerBlock;
		SPMVHIP_TRY(f->x_win.alloc(2 * nwg * sizeof(int32_t)));
		window_ranges<<<(unsigned) nwg, kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), f->col.as<int32_t>(), order, m,
				cap, f->x_win.as<int32_t>());
		SPMVHIP_CHECK(hipGetLastError());
		f->x_cap = cap;
	}
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	// the CSR copy is no longer needed once the slices exist
	f->row_ptr.release();
	f->col.release();
	f->val.release();
	f->mem_footprint = (double) total * 12 + (double) f->nslices * 12 + 8 + (sorted ? (double) m * 4 : 0);
	return SPMVHIP_OK;
}

// One wave per slice. UNROLL entries are loaded before their FMAs so that
// several loads are in flight per lane; the FMAs still run in entry order.
// With XW the workgroup first copies its x window (win, from window_ranges)
// into dynamic LDS.
template <int UNROLL, bool PIPE, bool XW>
__global__ void __launch_bounds__(kBlock) sell_kernel(const int64_t *__restrict__ sptr, const int32_t *__restrict__ width,
		const int32_t *__restrict__ scol, const double *__restrict__ sval, const int32_t *__restrict__ order,
		const double *__restrict__ x, double *__restrict__ y, long m, long nslices, int T, int remap, int stage,
		const int32_t *__restrict__ win)
{
	__shared__ double ybuf[kWavesPerBlock * kSliceC];
	extern __shared__ double xs[];
	const long blk = xcd_logical_block(remap);
	const long nwg = (nslices + kWavesPerBlock - 1) / kWavesPerBlock;
	if (blk >= nwg) // padding of an XCD-remapped grid
		return;
	const long s = blk * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
	int lo = 0;
	unsigned W = 0;
	if (XW) {
		lo = win[2 * blk];
		W = (unsigned) win[2 * blk + 1];
		for (unsigned i = threadIdx.x; i < W; i += kBlock)
			xs[i] = x[lo + i];
		__syncthreads();
	}
	if (!stage) {
		if (s < nslices)
			sell_slice<UNROLL, PIPE, XW>(sptr, width, scol, sval, order, x, y, m, s, T, nullptr, 0, xs, lo, W);
		return;
	}
	// The workgroup's 4 slices are rows [row0, row0 + 256) and every target
	// row of the permutation stays inside them (sigma divides 256): stage the
	// results in LDS and store them as one contiguous run.
	const long row0 = blk * kWavesPerBlock * kSliceC;
	if (s < nslices)
		sell_slice<UNROLL, PIPE, XW>(sptr, width, scol, sval, order, x, y, m, s, T, ybuf, row0, xs, lo, W);
	__syncthreads();
	const long r = row0 + threadIdx.x;
	if (r < m)
		y[r] = ybuf[threadIdx.x];
}

template <int U, bool P>
static void launch_sell(spmvhip_format *f, unsigned grid, const int32_t *order, const double *x, double *y,
		int remap, int stage, hipStream_t s)
{
	if (f->x_cap > 0) {
		const size_t lds = (size_t) f->x_cap * sizeof(double);
		if (!f->x_attr_set) {
			(void) hipFuncSetAttribute((const void *) sell_kernel<U, P, true>,
					hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
			f->x_attr_set = true;
		}
		sell_kernel<U, P, true><<<grid, kBlock, lds, s>>>(f->slice_ptr.as<int64_t>(), f->slice_w.as<int32_t>(),
				f->s_cols.as<int32_t>(), f->s_vals.as<double>(), order, x, y, f->m, f->nslices, f->lanes, remap,
				stage, f->x_win.as<int32_t>());
	} else {
		sell_kernel<U, P, false><<<grid, kBlock, 0, s>>>(f->slice_ptr.as<int64_t>(), f->slice_w.as<int32_t>(),
				f->s_cols.as<int32_t>(), f->s_vals.as<double>(), order, x, y, f->m, f->nslices, f->lanes, remap,
				stage, nullptr);
	}
}

int sell_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s)
{
	if (f->m == 0)
		return SPMVHIP_OK;
	const int remap = xcd_chunk(f->p.xcd_remap);
	const unsigned grid = xcd_padded_grid(grid_for(f->nslices, kWavesPerBlock), remap);
	const int32_t *order = f->row_order.p ? f->row_order.as<int32_t>() : nullptr;
	int u = f->p.unroll > 0 ? f->p.unroll : 8;
	// y_mode 0 (auto): through LDS when the permutation cannot leave a
	// workgroup's 256 rows (sigma divides 256) or several lanes per row would
	// store it piecewise, else direct stores
	const long wg_rows = (long) kWavesPerBlock * kSliceC;
	int stage = 0;
	double *yk = y;
	switch (f->p.y_mode) {
	case 0: stage = (order && wg_rows % f->p.sigma == 0) || (!order && f->lanes > 1) ? 1 : 0; break;
	case 1: stage = 0; break;
	case 2: stage = 1; break;
	case 3:
		if (order) { // results in sorted order, no permutation inside the kernel
			yk = f->ysorted.as<double>();
			order = nullptr;
		}
		stage = 0;
		break;
	}
	const bool pipe = f->p.pipeline != 0;
	if (u >= 8)
		pipe ? launch_sell<8, true>(f, grid, order, x, yk, remap, stage, s)
		     : launch_sell<8, false>(f, grid, order, x, yk, remap, stage, s);
	else if (u >= 4)
		pipe ? launch_sell<4, true>(f, grid, order, x, yk, remap, stage, s)
		     : launch_sell<4, false>(f, grid, order, x, yk, remap, stage, s);
	else if (u >= 2)
		pipe ? launch_sell<2, true>(f, grid, order, x, yk, remap, stage, s)
		     : launch_sell<2, false>(f, grid, order, x, yk, remap, stage, s);
	else
		pipe ? launch_sell<1, true>(f, grid, order, x, yk, remap, stage, s)
		     : launch_sell<1, false>(f, grid, order, x, yk, remap, stage, s);
	SPMVHIP_CHECK(hipGetLastError());
	if (yk != y) {
		unpermute_y<<<grid_for(f->m, kBlock), kBlock, 0, s>>>(yk, f->row_inv.as<int32_t>(), f->m, y);
		SPMVHIP_CHECK(hipGetLastError());
	}
	return SPMVHIP_OK;
}

} // namespace spmvhip

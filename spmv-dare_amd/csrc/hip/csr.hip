// CSR formats on gfx950.
//
// csr_naive (src/spmv_kernels/csr.cpp:166-182): one thread per row, a fresh
// sum per row accumulated left to right as a separate multiply and add (no
// fused multiply-add), so y is bit-identical to a sequential CPU loop.
//
// csr_vec (src/spmv_kernels/csr_vec.cpp:183-243 is the CPU kernel it
// replaces): row-length-binned wave64 CSR. csr_to_format builds, on the
// device, a list of row blocks, one per wavefront:
//   * stream blocks: consecutive rows whose starts fall in the same
//     block_nnz-sized nnz window, at most 64 rows. The wave streams the
//     block's values/columns with coalesced loads, stages the products in LDS
//     and reduces each row with a group of L lanes (L = 64 / rows, so short
//     rows get one lane -- thread per row -- and a lone row gets all 64);
//   * long rows (block_nnz < len <= split_nnz): one row per wave, each lane
//     strides the row and the wave reduces (wave per row);
//   * split rows (len > split_nnz): cut into split_nnz segments, one wave per
//     segment writes a partial, and a second pass adds each row's partials in
//     segment order (several waves per row, deterministic y).

#include <algorithm>

#include "format.hpp"
#include "kernels.hpp"
#include "scan.hpp"

namespace spmvhip {

// ------------------------------------------------------------------ naive

__global__ void __launch_bounds__(kBlock) csr_naive_kernel(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
		const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y, long m)
{
#pragma clang fp contract(off)
	long i = (long) blockIdx.x * kBlock + threadIdx.x;
	if (i >= m)
		return;
	double sum = 0;
	for (int j = rp[i]; j < rp[i + 1]; j++) {
		double prod = val[j] * x[col[j]];
		sum = sum + prod;
	}
	y[i] = sum;
}

int csr_naive_convert(spmvhip_format *f)
{
	f->mem_footprint = f->csr_mem_footprint;
	return SPMVHIP_OK;
}

int csr_naive_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s)
{
	if (f->m == 0)
		return SPMVHIP_OK;
	csr_naive_kernel<<<grid_for(f->m, kBlock), kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), f->col.as<int32_t>(),
			f->val.as<double>(), x, y, f->m);
	SPMVHIP_CHECK(hipGetLastError());
	return SPMVHIP_OK;
}

// ------------------------------------------------------------------ binned conversion

constexpr int kMaxBlockRows = 64;

// A row block's rows [r0, r1) and nonzeros [j0, j1).
struct alignas(16) BlockRec {
	int32_t r0, r1, j0, j1;
};

__global__ void csr_fill_records(const int32_t *__restrict__ blocks, const int32_t *__restrict__ rp, long nb,
		BlockRec *__restrict__ out)
{
	const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nb)
		return;
	const int32_t r0 = blocks[i], r1 = blocks[i + 1];
	out[i] = BlockRec{r0, r1, rp[r0], rp[r1]};
}

// flag[r] = 1 where row r starts a row block; split[r] = 1 for split rows.
__global__ void csr_mark_blocks(const int32_t *__restrict__ rp, long m, int bnnz, int split_nnz,
		int64_t *__restrict__ flag, int64_t *__restrict__ split)
{
	long r = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m)
		return;
	int s = rp[r], len = rp[r + 1] - s;
	int start = 0;
	if (r == 0 || (r % kMaxBlockRows) == 0 || len > bnnz) {
		start = 1;
	} else {
		int sp = rp[r - 1], lp = s - sp;
		if (lp > bnnz || (s / bnnz) != (sp / bnnz))
			start = 1;
	}
	flag[r] = start;
	split[r] = len > split_nnz ? 1 : 0;
}

__global__ void csr_scatter_blocks(const int64_t *__restrict__ flag, const int64_t *__restrict__ pos, long m,
		int32_t *__restrict__ blocks, const int64_t *__restrict__ split, const int64_t *__restrict__ spos,
		int32_t *__restrict__ split_rows)
{
	long r = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= m)
		return;
	if (flag[r])
		blocks[pos[r]] = (int32_t) r;
	if (split_rows && split[r])
		split_rows[spos[r]] = (int32_t) r;
	if (r == m - 1)
		blocks[pos[m]] = (int32_t) m;
}

__global__ void csr_split_counts(const int32_t *__restrict__ rp, const int32_t *__restrict__ split_rows, long nsplit,
		int split_nnz, int64_t *__restrict__ cnt)
{
	long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nsplit)
		return;
	int r = split_rows[i];
	long len = rp[r + 1] - rp[r];
	cnt[i] = (len + split_nnz - 1) / split_nnz;
}

__global__ void csr_split_desc(const int32_t *__restrict__ rp, const int32_t *__restrict__ split_rows, long nsplit,
		int split_nnz, const int64_t *__restrict__ seg_pos, int32_t *__restrict__ seg_ptr32,
		int32_t *__restrict__ desc)
{
	long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nsplit)
		return;
	int r = split_rows[i];
	int j0 = rp[r], j1 = rp[r + 1];
	long s0 = seg_pos[i];
	seg_ptr32[i] = (int32_t) s0;
	if (i == nsplit - 1)
		seg_ptr32[nsplit] = (int32_t) seg_pos[nsplit];
	long k = 0;
	for (int j = j0; j < j1; j += split_nnz, k++) {
		int je = j1 - j > split_nnz ? j + split_nnz : j1;
		desc[3 * (s0 + k) + 0] = r;
		desc[3 * (s0 + k) + 1] = j;
		desc[3 * (s0 + k) + 2] = je;
	}
}

int csr_vec_convert(spmvhip_format *f)
{
	hipStream_t s = f->stream;
	const long m = f->m;
	f->block_nnz = f->p.block_nnz > 0 ? f->p.block_nnz : 256;
	f->split_nnz = f->p.split_nnz > 0 ? f->p.split_nnz : 8192;
	if (f->block_nnz < 64 || f->block_nnz > 1024)
		return set_error(SPMVHIP_EINVAL, "csr_vec: block_nnz %d outside [64, 1024]", f->block_nnz);
	if (f->split_nnz < f->block_nnz)
		return set_error(SPMVHIP_EINVAL, "csr_vec: split_nnz %d < block_nnz %d", f->split_nnz, f->block_nnz);
	int L = f->p.lanes_per_row;
	if (L != 0 && (L < 1 || L > 64 || (L & (L - 1))))
		return set_error(SPMVHIP_EINVAL, "csr_vec: lanes_per_row %d is not a power of two in [1, 64]", L);
	f->lanes = L;

	DevBuf flag, pos, split, spos;
	SPMVHIP_TRY(flag.alloc((m + 1) * sizeof(int64_t)));
	SPMVHIP_TRY(pos.alloc((m + 1) * sizeof(int64_t)));
	SPMVHIP_TRY(split.alloc((m + 1) * sizeof(int64_t)));
	SPMVHIP_TRY(spos.alloc((m + 1) * sizeof(int64_t)));
	csr_mark_blocks<<<grid_for(m, kBlock), kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), m, f->block_nnz, f->split_nnz,
			flag.as<int64_t>(), split.as<int64_t>());
	SPMVHIP_CHECK(hipGetLastError());
	SPMVHIP_TRY((device_scan<int64_t, OpSum>(flag.as<int64_t>(), pos.as<int64_t>(), m, 0, 1, s)));
	SPMVHIP_TRY((device_scan<int64_t, OpSum>(split.as<int64_t>(), spos.as<int64_t>(), m, 0, 1, s)));
	int64_t nb = 0, ns = 0;
	SPMVHIP_CHECK(hipMemcpyAsync(&nb, pos.as<int64_t>() + m, sizeof(nb), hipMemcpyDeviceToHost, s));
	SPMVHIP_CHECK(hipMemcpyAsync(&ns, spos.as<int64_t>() + m, sizeof(ns), hipMemcpyDeviceToHost, s));
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	f->nblocks = nb;
	f->nsplit = ns;
	SPMVHIP_TRY(f->blocks.alloc((nb + 1) * sizeof(int32_t)));
	if (ns > 0)
		SPMVHIP_TRY(f->split_rows.alloc(ns * sizeof(int32_t)));
	csr_scatter_blocks<<<grid_for(m, kBlock), kBlock, 0, s>>>(flag.as<int64_t>(), pos.as<int64_t>(), m,
			f->blocks.as<int32_t>(), split.as<int64_t>(), spos.as<int64_t>(),
			ns > 0 ? f->split_rows.as<int32_t>() : nullptr);
	SPMVHIP_CHECK(hipGetLastError());
	// one 16-byte record per block: a wave reads its block's rows and
	// nonzero range with one scalar load
	SPMVHIP_TRY(f->brec.alloc((nb > 0 ? nb : 1) * sizeof(BlockRec)));
	if (nb > 0) {
		csr_fill_records<<<grid_for(nb, kBlock), kBlock, 0, s>>>(f->blocks.as<int32_t>(), f->row_ptr.as<int32_t>(),
				nb, f->brec.as<BlockRec>());
		SPMVHIP_CHECK(hipGetLastError());
	}

	if (ns > 0) {
		DevBuf cnt, segpos;
		SPMVHIP_TRY(cnt.alloc(ns * sizeof(int64_t)));
		SPMVHIP_TRY(segpos.alloc((ns + 1) * sizeof(int64_t)));
		csr_split_counts<<<grid_for(ns, kBlock), kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), f->split_rows.as<int32_t>(),
				ns, f->split_nnz, cnt.as<int64_t>());
		SPMVHIP_CHECK(hipGetLastError());
		SPMVHIP_TRY((device_scan<int64_t, OpSum>(cnt.as<int64_t>(), segpos.as<int64_t>(), ns, 0, 1, s)));
		int64_t nsegs = 0;
		SPMVHIP_CHECK(hipMemcpyAsync(&nsegs, segpos.as<int64_t>() + ns, sizeof(nsegs), hipMemcpyDeviceToHost, s));
		SPMVHIP_CHECK(hipStreamSynchronize(s));
		f->nsegs = nsegs;
		SPMVHIP_TRY(f->seg_desc.alloc(3 * nsegs * sizeof(int32_t)));
		SPMVHIP_TRY(f->seg_ptr.alloc((ns + 1) * sizeof(int32_t)));
		SPMVHIP_TRY(f->partials.alloc(nsegs * sizeof(double)));
		csr_split_desc<<<grid_for(ns, kBlock), kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), f->split_rows.as<int32_t>(),
				ns, f->split_nnz, segpos.as<int64_t>(), f->seg_ptr.as<int32_t>(), f->seg_desc.as<int32_t>());
		SPMVHIP_CHECK(hipGetLastError());
		// cnt / segpos are released at the end of this block
		SPMVHIP_CHECK(hipStreamSynchronize(s));
	}
	SPMVHIP_CHECK(hipStreamSynchronize(s));
	f->mem_footprint = f->csr_mem_footprint + (double) (nb + 1) * 4 + (double) nb * sizeof(BlockRec) +
			(double) f->nsegs * 20;
	return SPMVHIP_OK;
}

// ------------------------------------------------------------------ binned kernels

// One wave per row block. The block's bounds come from its 16-byte record
// with one scalar load (through the constant address space), so the wave's
// stream is requested one memory round trip after it starts, not two (block
// table, then row pointers).
template <int BNNZ_MAX, bool NARROW>
__global__ void __launch_bounds__(kBlock) csr_binned_kernel(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
		const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
		const BlockRec *__restrict__ brec, long nblocks, int split_nnz, int lanes_cfg, int remap)
{
	__shared__ double prod[kWavesPerBlock][2 * BNNZ_MAX];
	const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave); // wave-uniform: scalar loads below
	const long w = xcd_logical_block(remap) * kWavesPerBlock + wv;
	if (w >= nblocks)
		return; // waves never synchronise with each other: early exit is safe
	const BlockRec r = const_ptr<BlockRec>(ptr_bits(brec))[w];
	csr_block_at<BNNZ_MAX, NARROW>(rp, col, val, x, y, r.r0, r.r1, r.j0, r.j1, split_nnz, true, lanes_cfg,
			prod[wv]);
}

// Persistent variant: a grid of resident workgroups whose waves take blocks// Persistent variant: a grid of resident workgroups whose waves take blocks
// w, w + W, w + 2W, ... (W waves in all, so at any time the waves work on W
// consecutive blocks), with block w + W's stream requested while block w is
// reduced (kernels.hpp: csr_stage_*).
template <int BNNZ_MAX>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8)))
csr_binned_pipe_kernel(const int32_t *__restrict__ rp,
		const int32_t *__restrict__ col, const double *__restrict__ val, const double *__restrict__ x,
		double *__restrict__ y, const int32_t *__restrict__ blocks, long nblocks, int split_nnz, int lanes_cfg)
{
	__shared__ double prod[kWavesPerBlock][2 * BNNZ_MAX];
	const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
	const long W = (long) gridDim.x * kWavesPerBlock;
	long w = (long) blockIdx.x * kWavesPerBlock + wv;
	if (w >= nblocks)
		return; // waves never synchronise with each other: early exit is safe
	// block bounds and row starts through the constant address space: scalar
	// loads, which do not wait for the previous block's gathers and stores
	rp = const_ptr<int32_t>(ptr_bits(rp));
	blocks = const_ptr<int32_t>(ptr_bits(blocks));
	CsrStage cur, nxt;
	csr_stage_load<BNNZ_MAX>(cur, rp, col, val, blocks[w], blocks[w + 1]);
	for (;;) {
		double xv[4];
		csr_stage_gather(cur, x, xv);
		const long wn = w + W;
		if (wn < nblocks)
			csr_stage_load<BNNZ_MAX>(nxt, rp, col, val, blocks[wn], blocks[wn + 1]);
		csr_stage_finish<BNNZ_MAX>(cur, xv, col, val, x, y, split_nnz, true, lanes_cfg, prod[wv]);
		if (wn >= nblocks)
			break;
		cur = nxt;
		w = wn;
	}
}

template <bool NARROW>
__global__ void __launch_bounds__(kBlock) csr_split_kernel(const int32_t *__restrict__ col, const double *__restrict__ val,
		const double *__restrict__ x, const int32_t *__restrict__ desc, long nsegs, double *__restrict__ partials)
{
	const int lane = lane_id();
	const long w = (long) blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
	if (w >= nsegs)
		return;
	const int j0 = desc[3 * w + 1], j1 = desc[3 * w + 2];
	const double sum = NARROW ? wave_row_sum(col, val, x, j0, j1) : wave_row_sum_pairs(col, val, x, j0, j1);
	if (lane == 0)
		partials[w] = sum;
}

__global__ void csr_split_finalize(const int32_t *__restrict__ split_rows, const int32_t *__restrict__ seg_ptr, long nsplit,
		const double *__restrict__ partials, double *__restrict__ y)
{
	long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nsplit)
		return;
	double s = 0;
	for (int k = seg_ptr[i]; k < seg_ptr[i + 1]; k++)
		s += partials[k];
	y[split_rows[i]] = s;
}

int csr_vec_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s)
{
	if (f->m == 0)
		return SPMVHIP_OK;
	const bool narrow = f->p.narrow_loads != 0;
	if (f->nsegs > 0) {
		if (narrow)
			csr_split_kernel<true><<<grid_for(f->nsegs, kWavesPerBlock), kBlock, 0, s>>>(f->col.as<int32_t>(),
					f->val.as<double>(), x, f->seg_desc.as<int32_t>(), f->nsegs, f->partials.as<double>());
		else
			csr_split_kernel<false><<<grid_for(f->nsegs, kWavesPerBlock), kBlock, 0, s>>>(f->col.as<int32_t>(),
					f->val.as<double>(), x, f->seg_desc.as<int32_t>(), f->nsegs, f->partials.as<double>());
		SPMVHIP_CHECK(hipGetLastError());
	}
	const int remap = xcd_chunk(f->p.xcd_remap);
	const unsigned grid = xcd_padded_grid(grid_for(f->nblocks, kWavesPerBlock), remap);
	if (f->p.pipeline) {
		// persistent grid: the resident workgroups of the kernel instance
		if (f->pipe_grid == 0) {
			int dev = 0, cus = 0, per_cu = 0;
			SPMVHIP_CHECK(hipGetDevice(&dev));
			SPMVHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
			const void *k = f->block_nnz <= 256 ? (const void *) csr_binned_pipe_kernel<256>
					: f->block_nnz <= 512 ? (const void *) csr_binned_pipe_kernel<512>
					: (const void *) csr_binned_pipe_kernel<1024>;
			SPMVHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kBlock, 0));
			f->pipe_grid = (unsigned) std::max(1, cus * std::max(per_cu, 1));
		}
		const unsigned pgrid = std::min<unsigned>(grid_for(f->nblocks, kWavesPerBlock), f->pipe_grid);
#define LAUNCH_PIPE(B)                                                                                          \
	csr_binned_pipe_kernel<B><<<pgrid, kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), f->col.as<int32_t>(),       \
			f->val.as<double>(), x, y, f->blocks.as<int32_t>(), f->nblocks, f->split_nnz, f->lanes)
		if (f->block_nnz <= 256)
			LAUNCH_PIPE(256);
		else if (f->block_nnz <= 512)
			LAUNCH_PIPE(512);
		else
			LAUNCH_PIPE(1024);
#undef LAUNCH_PIPE
	} else {
#define LAUNCH_BINNED(B, N)                                                                                 \
	csr_binned_kernel<B, N><<<grid, kBlock, 0, s>>>(f->row_ptr.as<int32_t>(), f->col.as<int32_t>(),     \
			f->val.as<double>(), x, y, f->brec.as<BlockRec>(), f->nblocks, f->split_nnz, f->lanes, \
			remap)
		if (f->block_nnz <= 256) {
			if (narrow)
				LAUNCH_BINNED(256, true);
			else
				LAUNCH_BINNED(256, false);
		} else if (f->block_nnz <= 512) {
			if (narrow)
				LAUNCH_BINNED(512, true);
			else
				LAUNCH_BINNED(512, false);
		} else {
			if (narrow)
				LAUNCH_BINNED(1024, true);
			else
				LAUNCH_BINNED(1024, false);
		}
#undef LAUNCH_BINNED
	}
	SPMVHIP_CHECK(hipGetLastError());
	if (f->nsegs > 0) {
		csr_split_finalize<<<grid_for(f->nsplit, kBlock), kBlock, 0, s>>>(f->split_rows.as<int32_t>(),
				f->seg_ptr.as<int32_t>(), f->nsplit, f->partials.as<double>(), y);
		SPMVHIP_CHECK(hipGetLastError());
	}
	return SPMVHIP_OK;
}

} // namespace spmvhip

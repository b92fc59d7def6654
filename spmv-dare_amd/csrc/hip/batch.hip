// Batched SpMV: many independent converted matrices, each with its own x and
// y, computed by ONE persistent kernel launch that takes work from queues
// (atomic tickets over (matrix, unit range) items), optionally replayed from
// a captured hipGraph. The reference has no counterpart: its driver runs one
// matrix per process (run.sh:116-151).
//
// Units are the formats' own per-wave work: a row block of the binned CSR
// format or a 64-row slice of SELL / SELL-C-sigma (kernels.hpp). A work item
// is consecutive units of one matrix (about 16k nonzeros, or `tile` units),
// taken by one wavefront. CSR rows longer than the member's split_nnz are
// cut into the standalone kernel's segments: segment work items write
// partial sums, and batch_split_finalize (the next node of the same launch
// sequence) adds each row's partials in segment order, so a heavy-tailed row
// is spread over many waves instead of holding one.
//
// XCD-aware scheduling: the matrices are dealt to 8 queues, one per XCD,
// balanced by nnz (largest first to the least loaded queue). Workgroup b is
// dispatched to XCD b % 8 and first drains queue b % 8, so each XCD's L2
// holds the x vectors of its own matrices only; a wave whose queue is empty
// moves on to the next queues, which balances the tail. Correctness does not
// depend on the dispatch order, only locality does.

#include <algorithm>
#include <numeric>
#include <vector>

#include "format.hpp"
#include "kernels.hpp"

namespace spmvhip {

constexpr int kBatchBnnzMax = 512; // largest csr_vec block_nnz a member may have
constexpr long kItemNnz = 16384;   // default work-item size in nnz (one ticket each)

struct BatchMat {
	int kind, lanes;
	long m;
	const int32_t *rp, *col, *blocks;
	const double *val;
	const int64_t *sptr;
	const int32_t *sw, *scol, *order;
	const double *sval;
	const double *x;
	double *y;
	// CSR rows longer than split_nnz: segments (row, j0, j1) summed into
	// partials by the persistent kernel, added per row by batch_split_finalize
	const int32_t *seg_desc, *seg_ptr, *split_rows;
	double *partials;
	int split_nnz;
};

enum { kItemUnits = 0, kItemSegments = 1 };

struct WorkItem {
	int mat, u0, u1, type; // units (row blocks / slices) or split-row segments [u0, u1)
};

struct SplitRef {
	int mat, i; // split row i of member mat
};

// zero the kXcds queue counters (a kernel, not a memset: the batch is
// captured into a hipGraph and replayed)
__global__ void batch_reset(int *counter)
{
	if (threadIdx.x < kXcds)
		counter[threadIdx.x] = 0;
}

// items[qstart[q] .. qstart[q + 1]) is queue q. BNNZ: LDS staging per wave
// for CSR row blocks, at least the largest member block_nnz.
template <int BNNZ>
__global__ void __launch_bounds__(kBlock) batch_kernel(const BatchMat *__restrict__ mats,
		const WorkItem *__restrict__ items, const int *__restrict__ qstart, int *__restrict__ counter)
{
	__shared__ double prod[kWavesPerBlock][2 * BNNZ];
	const int wv = threadIdx.x / kWave;
	const int home = (int) (blockIdx.x % kXcds);
	for (int k = 0; k < kXcds; k++) {
		const int q = (home + k) % kXcds;
		const int q0 = qstart[q], q1 = qstart[q + 1];
		for (;;) {
			int t = 0;
			if (lane_id() == 0)
				t = atomicAdd(counter + q, 1);
			t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, kWave));
			if (q0 + t >= q1)
				break; // queue drained: wave-uniform exit, every wave reaches it
			const WorkItem wi = items[q0 + t];
			const BatchMat &M = mats[wi.mat];
			if (wi.type == kItemSegments) {
				for (int u = wi.u0; u < wi.u1; u++) {
					const double sum = wave_row_sum(M.col, M.val, M.x, M.seg_desc[3 * u + 1], M.seg_desc[3 * u + 2]);
					if (lane_id() == 0)
						M.partials[u] = sum;
				}
			} else if (M.kind == SPMVHIP_CSR_VEC) {
				for (int u = wi.u0; u < wi.u1; u++)
					csr_block<BNNZ>(M.rp, M.col, M.val, M.x, M.y, M.blocks[u], M.blocks[u + 1], M.split_nnz,
							M.partials != nullptr, M.lanes, prod[wv]);
			} else {
				for (int u = wi.u0; u < wi.u1; u++)
					sell_slice<4>(M.sptr, M.sw, M.scol, M.sval, M.order, M.x, M.y, M.m, u, M.lanes);
			}
		}
	}
}

// y[row] = sum of the row's segment partials in segment order (the
// standalone csr_split_finalize, over every member of the batch)
__global__ void batch_split_finalize(const BatchMat *__restrict__ mats, const SplitRef *__restrict__ refs, long n)
{
	const long k = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const BatchMat &M = mats[refs[k].mat];
	const int i = refs[k].i;
	double s = 0;
	for (int j = M.seg_ptr[i]; j < M.seg_ptr[i + 1]; j++)
		s += M.partials[j];
	M.y[M.split_rows[i]] = s;
}

} // namespace spmvhip

struct spmvhip_batch {
	int tile = 0;        // units per work item, or 0: sized by item_nnz
	long item_nnz = 16384;
	int bnnz = 256; // LDS staging of the kernel instance (256 or 512)
	int queues = 8; // 8: one queue per XCD; 1: every wave draws from one shared queue
	std::vector<spmvhip_format *> fmts;
	std::vector<spmvhip::BatchMat> host_mats;
	std::vector<spmvhip::WorkItem> host_items;
	std::vector<spmvhip::SplitRef> host_splits;
	std::vector<long> mat_nnz;
	std::vector<int> host_qstart;
	spmvhip::DevBuf mats, items, qstart, counter, splits;
	hipStream_t stream = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	int grid = 0;
	bool finalized = false;
	long total_nnz = 0;
};

using namespace spmvhip;

namespace {

int batch_launch(spmvhip_batch *b, hipStream_t s)
{
	batch_reset<<<1, kWave, 0, s>>>(b->counter.as<int>());
	SPMVHIP_CHECK(hipGetLastError());
	if (!b->host_items.empty()) {
		if (b->bnnz > 256)
			batch_kernel<kBatchBnnzMax><<<b->grid, kBlock, 0, s>>>(b->mats.as<BatchMat>(), b->items.as<WorkItem>(),
					b->qstart.as<int>(), b->counter.as<int>());
		else
			batch_kernel<256><<<b->grid, kBlock, 0, s>>>(b->mats.as<BatchMat>(), b->items.as<WorkItem>(),
					b->qstart.as<int>(), b->counter.as<int>());
		SPMVHIP_CHECK(hipGetLastError());
	}
	if (!b->host_splits.empty()) {
		batch_split_finalize<<<grid_for((long) b->host_splits.size(), kBlock), kBlock, 0, s>>>(
				b->mats.as<BatchMat>(), b->splits.as<SplitRef>(), (long) b->host_splits.size());
		SPMVHIP_CHECK(hipGetLastError());
	}
	return SPMVHIP_OK;
}

} // namespace

extern "C" {

int spmvhip_batch_create(int tile_units, spmvhip_batch **out)
{
	if (!out)
		return set_error(SPMVHIP_EINVAL, "out is NULL");
	*out = nullptr;
	if (tile_units < -(1 << 24) || tile_units > 65536)
		return set_error(SPMVHIP_EINVAL, "tile_units %d", tile_units);
	spmvhip_batch *b = new (std::nothrow) spmvhip_batch();
	if (!b)
		return set_error(SPMVHIP_ENOMEM, "batch handle");
	b->tile = tile_units > 0 ? tile_units : 0;
	b->item_nnz = tile_units < 0 ? -(long) tile_units : kItemNnz;
	hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
	if (e != hipSuccess) {
		delete b;
		return set_error(SPMVHIP_EHIP, "stream: %s", hipGetErrorString(e));
	}
	*out = b;
	return SPMVHIP_OK;
}

int spmvhip_batch_set_queues(spmvhip_batch *b, int nqueues)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	if (b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch already finalized");
	if (nqueues != 1 && nqueues != kXcds)
		return set_error(SPMVHIP_EINVAL, "batch: %d queues (1 shared, or %d per XCD)", nqueues, kXcds);
	b->queues = nqueues;
	return SPMVHIP_OK;
}

int spmvhip_batch_add(spmvhip_batch *b, const spmvhip_format *f, const double *d_x, double *d_y)
{
	if (!b || !f || !d_x || !d_y)
		return set_error(SPMVHIP_EINVAL, "NULL argument");
	if (b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch already finalized");
	BatchMat M{};
	M.kind = f->kind;
	M.m = f->m;
	M.x = d_x;
	M.y = d_y;
	long units;
	if (f->kind == SPMVHIP_CSR_VEC) {
		if (f->block_nnz > kBatchBnnzMax)
			return set_error(SPMVHIP_EINVAL, "batch: csr_vec block_nnz %d > %d", f->block_nnz, kBatchBnnzMax);
		if (f->block_nnz > b->bnnz)
			b->bnnz = kBatchBnnzMax;
		M.rp = f->row_ptr.as<int32_t>();
		M.col = f->col.as<int32_t>();
		M.val = f->val.as<double>();
		M.blocks = f->blocks.as<int32_t>();
		M.lanes = f->lanes;
		M.split_nnz = f->split_nnz;
		if (f->nsegs > 0) {
			M.seg_desc = f->seg_desc.as<int32_t>();
			M.seg_ptr = f->seg_ptr.as<int32_t>();
			M.split_rows = f->split_rows.as<int32_t>();
			M.partials = f->partials.as<double>();
		}
		units = f->nblocks;
	} else if (f->kind == SPMVHIP_SELL || f->kind == SPMVHIP_SELL_C_S) {
		M.sptr = f->slice_ptr.as<int64_t>();
		M.sw = f->slice_w.as<int32_t>();
		M.scol = f->s_cols.as<int32_t>();
		M.sval = f->s_vals.as<double>();
		M.order = f->kind == SPMVHIP_SELL_C_S ? f->row_order.as<int32_t>() : nullptr;
		M.lanes = f->lanes; // lanes per row T (>= 1) of the slice layout
		units = f->nslices;
	} else {
		return set_error(SPMVHIP_EINVAL, "batch: format %s is not batchable (csr_vec, sell, sell_c_s are)",
				f->name);
	}
	if (units > INT32_MAX)
		return set_error(SPMVHIP_EINVAL, "batch: too many units");
	int mi = (int) b->host_mats.size();
	b->host_mats.push_back(M);
	b->mat_nnz.push_back(f->nnz);
	b->fmts.push_back(const_cast<spmvhip_format *>(f));
	// units per item: fixed, or enough units of this matrix's average size
	// for about item_nnz nonzeros (a device-scope ticket costs as much as
	// streaming several KB, so items are kept coarse)
	long per = b->tile;
	if (per <= 0)
		per = units > 0 ? std::max<long>(1, b->item_nnz * units / std::max<long>(1, f->nnz)) : 1;
	// segments of split rows first (the longest work of the member), each
	// item about item_nnz nonzeros; their rows are completed by
	// batch_split_finalize after the persistent kernel
	if (M.partials) {
		const long nsegs = f->nsegs, sper = std::max<long>(1, b->item_nnz / std::max(1, f->split_nnz));
		for (long u = 0; u < nsegs; u += sper)
			b->host_items.push_back({mi, (int) u, (int) std::min(nsegs, u + sper), kItemSegments});
		for (long i = 0; i < f->nsplit; i++)
			b->host_splits.push_back({mi, (int) i});
	}
	for (long u = 0; u < units; u += per)
		b->host_items.push_back({mi, (int) u, (int) std::min(units, u + per), kItemUnits});
	b->total_nnz += f->nnz;
	return SPMVHIP_OK;
}

int spmvhip_batch_finalize(spmvhip_batch *b, int use_graph)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	if (b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch already finalized");
	if (b->host_items.size() > (size_t) INT32_MAX)
		return set_error(SPMVHIP_EINVAL, "batch: too many work items");
	// deal the matrices to the XCD queues: largest first to the least loaded
	const size_t nm = b->host_mats.size();
	std::vector<int> by_size(nm), queue_of(nm);
	std::iota(by_size.begin(), by_size.end(), 0);
	std::stable_sort(by_size.begin(), by_size.end(), [&](int i, int j) { return b->mat_nnz[i] > b->mat_nnz[j]; });
	long load[kXcds] = {0};
	for (int i : by_size) {
		const int q = b->queues == 1 ? 0 : (int) (std::min_element(load, load + kXcds) - load);
		queue_of[i] = q;
		load[q] += b->mat_nnz[i] + 1;
	}
	std::vector<WorkItem> sorted;
	sorted.reserve(b->host_items.size());
	b->host_qstart.assign(kXcds + 1, 0);
	for (int q = 0; q < kXcds; q++) {
		b->host_qstart[q] = (int) sorted.size();
		for (const WorkItem &w : b->host_items)
			if (queue_of[w.mat] == q)
				sorted.push_back(w);
	}
	b->host_qstart[kXcds] = (int) sorted.size();
	b->host_items.swap(sorted);

	SPMVHIP_TRY(b->mats.alloc(std::max<size_t>(1, nm) * sizeof(BatchMat)));
	SPMVHIP_TRY(b->items.alloc(std::max<size_t>(1, b->host_items.size()) * sizeof(WorkItem)));
	SPMVHIP_TRY(b->qstart.alloc((kXcds + 1) * sizeof(int)));
	SPMVHIP_TRY(b->counter.alloc(kXcds * sizeof(int)));
	if (nm)
		SPMVHIP_CHECK(hipMemcpy(b->mats.p, b->host_mats.data(), nm * sizeof(BatchMat), hipMemcpyHostToDevice));
	if (!b->host_items.empty())
		SPMVHIP_CHECK(hipMemcpy(b->items.p, b->host_items.data(), b->host_items.size() * sizeof(WorkItem),
				hipMemcpyHostToDevice));
	SPMVHIP_CHECK(hipMemcpy(b->qstart.p, b->host_qstart.data(), (kXcds + 1) * sizeof(int), hipMemcpyHostToDevice));
	if (!b->host_splits.empty()) {
		SPMVHIP_TRY(b->splits.alloc(b->host_splits.size() * sizeof(SplitRef)));
		SPMVHIP_CHECK(hipMemcpy(b->splits.p, b->host_splits.data(), b->host_splits.size() * sizeof(SplitRef),
				hipMemcpyHostToDevice));
	}
	// persistent grid: as many workgroups as can be resident at once, a
	// multiple of the XCD count so that every queue has the same number
	int dev = 0, per_cu = 0;
	hipDeviceProp_t prop;
	SPMVHIP_CHECK(hipGetDevice(&dev));
	SPMVHIP_CHECK(hipGetDeviceProperties(&prop, dev));
	if (b->bnnz > 256)
		SPMVHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<kBatchBnnzMax>, kBlock, 0));
	else
		SPMVHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<256>, kBlock, 0));
	long g = (long) prop.multiProcessorCount * std::max(per_cu, 1);
	g = std::max<long>(kXcds, g / kXcds * kXcds);
	b->grid = (int) g;
	b->finalized = true;
	if (use_graph) {
		SPMVHIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
		int rc = batch_launch(b, b->stream);
		hipGraph_t graph = nullptr;
		hipError_t e = hipStreamEndCapture(b->stream, &graph);
		if (rc)
			return rc;
		if (e != hipSuccess)
			return set_error(SPMVHIP_EHIP, "graph capture: %s", hipGetErrorString(e));
		b->graph = graph;
		SPMVHIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
	}
	return SPMVHIP_OK;
}

int spmvhip_batch_run(spmvhip_batch *b)
{
	if (!b || !b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch not finalized");
	if (b->exec) {
		SPMVHIP_CHECK(hipGraphLaunch(b->exec, b->stream));
		return SPMVHIP_OK;
	}
	return batch_launch(b, b->stream);
}

int spmvhip_batch_synchronize(spmvhip_batch *b)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	SPMVHIP_CHECK(hipStreamSynchronize(b->stream));
	return SPMVHIP_OK;
}

int spmvhip_batch_time(spmvhip_batch *b, int per_matrix, int warmup, int iters, float *ms)
{
	if (!b || !b->finalized || iters < 1 || warmup < 0)
		return set_error(SPMVHIP_EINVAL, "bad arguments");
	auto once = [&]() -> int {
		if (!per_matrix)
			return spmvhip_batch_run(b);
		for (size_t i = 0; i < b->fmts.size(); i++)
			SPMVHIP_TRY(spmvhip_spmv_device(b->fmts[i], b->host_mats[i].x, b->host_mats[i].y, b->stream));
		return SPMVHIP_OK;
	};
	for (int i = 0; i < warmup; i++)
		SPMVHIP_TRY(once());
	hipEvent_t e0, e1;
	SPMVHIP_CHECK(hipEventCreate(&e0));
	SPMVHIP_CHECK(hipEventCreate(&e1));
	int rc = SPMVHIP_OK;
	for (int i = 0; i < iters && rc == SPMVHIP_OK; i++) {
		hipError_t he = hipEventRecord(e0, b->stream);
		if (he == hipSuccess)
			rc = once();
		if (he == hipSuccess && rc == SPMVHIP_OK)
			he = hipEventRecord(e1, b->stream);
		if (he == hipSuccess && rc == SPMVHIP_OK)
			he = hipEventSynchronize(e1);
		float t = 0;
		if (he == hipSuccess && rc == SPMVHIP_OK)
			he = hipEventElapsedTime(&t, e0, e1);
		if (he != hipSuccess && rc == SPMVHIP_OK)
			rc = set_error(SPMVHIP_EHIP, "batch timing: %s", hipGetErrorString(he));
		if (ms)
			ms[i] = t;
	}
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	return rc;
}

int spmvhip_batch_info(const spmvhip_batch *b, long *nmats, long *nitems, int *grid, long *total_nnz)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	if (nmats)
		*nmats = (long) b->host_mats.size();
	if (nitems)
		*nitems = (long) b->host_items.size();
	if (grid)
		*grid = b->grid;
	if (total_nnz)
		*total_nnz = b->total_nnz;
	return SPMVHIP_OK;
}

void spmvhip_batch_destroy(spmvhip_batch *b)
{
	if (!b)
		return;
	if (b->stream)
		(void) hipStreamSynchronize(b->stream);
	if (b->exec)
		(void) hipGraphExecDestroy(b->exec);
	if (b->graph)
		(void) hipGraphDestroy(b->graph);
	hipStream_t s = b->stream;
	delete b;
	if (s)
		(void) hipStreamDestroy(s);
}

} // extern "C"

// Batched SpMV: many independent converted matrices, each with its own x and
// y, computed by ONE persistent kernel launch, optionally replayed from a
// captured hipGraph. The reference has no counterpart: its driver runs one
// matrix per process (run.sh:116-151).
//
// Units are the formats' own per-wave work: a row block of the binned CSR
// format, a 64-row slice of SELL / SELL-C-sigma (kernels.hpp), or a segment
// of a CSR row longer than the member's split_nnz (the standalone kernel's
// segments: a partial sum per segment, and batch_split_finalize, the next
// node of the same launch sequence, adds each row's partials in segment
// order, so a heavy-tailed row is spread over many waves).
//
// Schedule. All units of all members form one list (members in insertion
// order, each member's segments before its other units), dealt over NC
// ticket counters: unit g belongs to counter g % NC, and ticket t of counter
// c is unit t * NC + c (or the `tile` units from there). Each resident wave
// draws from one counter and requests its next ticket before it runs the
// current unit, so the atomic's latency hides behind the unit. Because
// tickets are handed out in order, the units in flight at any moment are a
// window of about W adjacent units (W = resident waves), the in-order
// locality of hardware dispatch. A static sweep (wave w: units w, w + W, ...)
// lets the waves drift apart and widens that window, which raises the
// memory latency every load sees (measured: 792 instead of 686 cycles per
// L1-to-L2 read); tickets over one shared counter serialise on one line.
// The counters are device-scope atomics, performed at the memory side, each
// on a 256-byte line of its own.
// Correctness does not depend on the dispatch order, only locality does.

#include <algorithm>
#include <cmath>
#include <vector>

#include "format.hpp"
#include "kernels.hpp"

namespace spmvhip {

constexpr int kBatchBnnzMax = 512; // largest csr_vec block_nnz a member may have
constexpr int kCounterStride = 64; // ints between ticket counters: one 256-byte line each
// Ticket counters. Atomics on one counter serialise (about 12 ns each, as
// measured: 500k units over 8 counters take 0.84 ms, over one 6.1 ms); from
// 64 counters on, one 128M-nonzero matrix runs in the batch as fast as in
// its own launch (0.36 ms).
constexpr int kCountersDefault = 256;
constexpr int kCountersMax = 1024;
// Default resident workgroups per CU: all that fit. With per-unit tickets
// the full grid is the fastest measured (9.26 ms at 8 and 7 per CU against
// 9.38 at 6 and 10.5 at 4, for 2,048 generated matrices).
constexpr int kGridPerCu = 8;

// A member's device arrays, as pointer values (common.hpp: global_ptr). The
// fields every unit reads come first, so that a wave fetches them with one
// scalar load.
struct BatchMat {
	uint64_t x, y;      // double
	uint64_t col, val;  // CSR: col int32, val double; SELL*: slice columns int32, slice values double
	uint64_t rp;        // CSR: row_ptr int32; SELL*: slice_ptr int64
	uint64_t sw;        // SELL*: slice_w int32 (0 for CSR)
	uint64_t order;     // SELL-C-sigma: row_order int32 (0 otherwise)
	uint64_t partials;  // CSR with split rows: segment partial sums, double (0 otherwise)
	long m;
	int lanes, split_nnz;
	int narrow, pad_; // CSR: one stored entry per lane and load (the member's narrow_loads)
	// read by batch_split_finalize only
	uint64_t seg_ptr, split_rows; // int32
};

enum : int { kUnitCsr = 0, kUnitSell = 1, kUnitSeg = 2 };
constexpr int kKindShift = 28; // UnitRec::mk = member index | kind << 28
constexpr int kMatMask = (1 << kKindShift) - 1;

// One unit of the global list, precomputed on the device at finalize: a CSR
// row block (rows [r0, r1), nonzeros [j0, j1)), a SELL slice u, or split-row
// segment u (nonzeros [j0, j1)). With the bounds in the record a wave
// reaches its stream after two dependent scalar loads (the record, then the
// member's pointers) instead of four (record, member, block table, row
// pointers).
struct alignas(8) UnitRec {
	int mk, u;
	int r0, r1;
	int j0, j1;
};

struct SplitRef {
	int mat, i; // split row i of member mat
};

// zero the nc ticket counters (a kernel, not a memset: the batch is
// captured into a hipGraph and replayed)
__global__ void batch_reset(int *counter, int nc)
{
	if ((int) threadIdx.x < nc)
		counter[threadIdx.x * kCounterStride] = 0;
}

__device__ __forceinline__ int batch_ticket(int *ctr)
{
	int t = 0;
	if (lane_id() == 0)
		t = atomicAdd(ctr, 1);
	return t;
}

__global__ void batch_fill_csr(const int32_t *__restrict__ blocks, const int32_t *__restrict__ rp, int nb, int mk,
		UnitRec *__restrict__ out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nb)
		return;
	const int r0 = blocks[i], r1 = blocks[i + 1];
	out[i] = UnitRec{mk, i, r0, r1, rp[r0], rp[r1]};
}

__global__ void batch_fill_seg(const int32_t *__restrict__ desc, int ns, int mk, UnitRec *__restrict__ out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < ns)
		out[i] = UnitRec{mk, i, 0, 0, desc[3 * i + 1], desc[3 * i + 2]};
}

__global__ void batch_fill_sell(int ns, int mk, UnitRec *__restrict__ out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < ns)
		out[i] = UnitRec{mk, i, 0, 0, 0, 0};
}

template <int BNNZ>
__device__ __forceinline__ void batch_unit(const BatchMat *__restrict__ mats, const UnitRec &r, double *prod)
{
	const int kind = r.mk >> kKindShift;
	const BatchMat &M = mats[r.mk & kMatMask];
	const double *x = global_ptr<const double>(M.x);
	double *y = global_ptr<double>(M.y);
	const int32_t *col = global_ptr<const int32_t>(M.col);
	const double *val = global_ptr<const double>(M.val);
	if (kind == kUnitCsr) {
		if (M.narrow)
			csr_block_at<BNNZ, true>(global_ptr<const int32_t>(M.rp), col, val, x, y, r.r0, r.r1, r.j0, r.j1,
					M.split_nnz, M.partials != 0, M.lanes, prod);
		else
			csr_block_at<BNNZ, false>(global_ptr<const int32_t>(M.rp), col, val, x, y, r.r0, r.r1, r.j0, r.j1,
					M.split_nnz, M.partials != 0, M.lanes, prod);
	} else if (kind == kUnitSeg) {
		const double sum = M.narrow ? wave_row_sum(col, val, x, r.j0, r.j1)
		                            : wave_row_sum_pairs(col, val, x, r.j0, r.j1);
		if (lane_id() == 0)
			global_ptr<double>(M.partials)[r.u] = sum;
	} else {
		sell_slice<4>(const_ptr<int64_t>(M.rp), const_ptr<int32_t>(M.sw), col, val,
				M.order ? global_ptr<const int32_t>(M.order) : nullptr, x, y, M.m, r.u, M.lanes);
	}
}

// The unit list's ranges: one (the whole list) or one per XCD.
struct UnitRanges {
	int b[kXcds + 1];
};

// nc (a power of two, at most the resident wave count) ticket counters;
// `tile` units per ticket. BNNZ: LDS staging per wave for CSR row blocks, at
// least the largest member block_nnz.
//
// Unit indices are 32-bit (finalize checks the count): loop bounds are then
// compared on the scalar unit. A 64-bit signed compare has no scalar form, and
// the VGPRs it borrowed at the loop latch made the compiler wait there for
// every outstanding memory operation of the wave, the y stores included.
// The kernel holds all unit kinds' code; it is kept to 64 VGPRs so that 8
// waves per SIMD can stay resident.
template <int BNNZ>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) batch_kernel(
		const BatchMat *__restrict__ mats, const UnitRec *__restrict__ units, UnitRanges rg, int nranges, int tile,
		int nc, int *__restrict__ counter)
{
	__shared__ double prod[kWavesPerBlock][2 * BNNZ];
	const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
	// workgroup b runs on XCD b % 8 (the grid is a multiple of 8)
	const int q = nranges == 1 ? 0 : (int) (blockIdx.x % kXcds);
	const int lw = (nranges == 1 ? (int) blockIdx.x : (int) (blockIdx.x / kXcds)) * kWavesPerBlock + wv;
	const int c = lw & (nc - 1);
	const int beg = rg.b[q], end = rg.b[q + 1];
	int *ctr = counter + (q * nc + c) * kCounterStride;
	int t = __builtin_amdgcn_readfirstlane(__shfl(batch_ticket(ctr), 0, kWave));
	for (;;) {
		const long g0 = beg + ((long) t * nc + c) * tile;
		if (g0 >= end)
			break; // wave-uniform exit: tickets only grow, every wave reaches it
		const int tn = batch_ticket(ctr); // the next ticket, in flight during this one's units
		const int g1 = (int) std::min<long>(g0 + tile, end);
		for (int g = (int) g0; g < g1; g++)
			batch_unit<BNNZ>(mats, units[g], prod[wv]);
		t = __builtin_amdgcn_readfirstlane(__shfl(tn, 0, kWave));
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
	const int32_t *seg_ptr = global_ptr<const int32_t>(M.seg_ptr);
	const double *partials = global_ptr<const double>(M.partials);
	for (int j = seg_ptr[i]; j < seg_ptr[i + 1]; j++)
		s += partials[j];
	global_ptr<double>(M.y)[global_ptr<const int32_t>(M.split_rows)[i]] = s;
}

// A member's share of the unit list: `count` records from `off`, filled at
// finalize from the member's device tables.
struct UnitSpan {
	int mk;
	long off, count;
	const int32_t *blocks, *rp, *seg_desc; // CSR blocks / segments (device)
	double weight;                         // nonzeros per unit (the member's mean)
};

} // namespace spmvhip

struct spmvhip_batch {
	int tile = 1;                        // units per ticket
	int bnnz = 256;                      // LDS staging of the kernel instance (256 or 512)
	int queues = spmvhip::kCountersDefault; // ticket counters (a power of two)
	int nc = 0;                          // counters per range in use (at most the range's resident waves)
	int xcd_ranges = -1;                 // one contiguous, nnz-balanced range of units per XCD: 1 on, 0 off, -1 auto
	long max_member_units = 0;
	spmvhip::UnitRanges ranges{};
	int nranges = 1;
	int wgs_per_cu = 0;                 // persistent grid: workgroups per CU (0 = kGridPerCu)
	std::vector<spmvhip_format *> fmts;
	std::vector<spmvhip::BatchMat> host_mats;
	std::vector<spmvhip::UnitSpan> spans;
	long nunits = 0;
	std::vector<spmvhip::SplitRef> host_splits;
	spmvhip::DevBuf mats, units, counter, splits;
	// DIV_RF members: their packet groups, drawn by a second persistent kernel
	std::vector<spmvhip::BatchDivMat> host_div;
	std::vector<spmvhip::BatchDivUnit> host_div_units;
	spmvhip::DevBuf div_mats, div_units, div_counter;
	int div_buf = 0, div_acc = 0, div_grid = 0, div_nc = 1;
	hipStream_t stream = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	int grid = 0;
	bool finalized = false;
	long total_nnz = 0;
	std::vector<int> div_fmts;           // indices into fmts of the DIV members
	std::vector<std::pair<uint64_t, uint64_t>> div_xy; // their x and y
	// members of at least solo_nnz nonzeros (0: none) run as their own
	// kernels, ahead of the persistent kernels, in the same launch (and graph)
	long solo_nnz = 0;
	struct Solo {
		spmvhip_format *f;
		const double *x;
		double *y;
	};
	std::vector<Solo> solo;
};

using namespace spmvhip;

namespace {

int batch_launch(spmvhip_batch *b, hipStream_t s)
{
	// large members first, each filling the GPU with its own kernels
	for (const auto &m : b->solo)
		SPMVHIP_TRY(format_spmv(m.f, m.x, m.y, s));
	batch_reset<<<1, kCountersMax, 0, s>>>(b->counter.as<int>(), b->nc * b->nranges);
	SPMVHIP_CHECK(hipGetLastError());
	if (b->nunits > 0) {
		if (b->bnnz > 256)
			batch_kernel<kBatchBnnzMax><<<b->grid, kBlock, 0, s>>>(b->mats.as<BatchMat>(), b->units.as<UnitRec>(),
					b->ranges, b->nranges, b->tile, b->nc, b->counter.as<int>());
		else
			batch_kernel<256><<<b->grid, kBlock, 0, s>>>(b->mats.as<BatchMat>(), b->units.as<UnitRec>(),
					b->ranges, b->nranges, b->tile, b->nc, b->counter.as<int>());
		SPMVHIP_CHECK(hipGetLastError());
	}
	if (!b->host_splits.empty()) {
		batch_split_finalize<<<grid_for((long) b->host_splits.size(), kBlock), kBlock, 0, s>>>(
				b->mats.as<BatchMat>(), b->splits.as<SplitRef>(), (long) b->host_splits.size());
		SPMVHIP_CHECK(hipGetLastError());
	}
	if (!b->host_div_units.empty())
		SPMVHIP_TRY(div_batch_launch(b->div_mats.as<BatchDivMat>(), b->div_units.as<BatchDivUnit>(),
				(int) b->host_div_units.size(), b->div_nc, kCounterStride, b->div_counter.as<int>(), b->div_grid,
				b->div_buf, b->div_acc, s));
	return SPMVHIP_OK;
}

} // namespace

extern "C" {

int spmvhip_batch_create(int tile_units, spmvhip_batch **out)
{
	if (!out)
		return set_error(SPMVHIP_EINVAL, "out is NULL");
	*out = nullptr;
	if (tile_units < 0 || tile_units > 65536)
		return set_error(SPMVHIP_EINVAL, "tile_units %d (0..65536)", tile_units);
	spmvhip_batch *b = new (std::nothrow) spmvhip_batch();
	if (!b)
		return set_error(SPMVHIP_ENOMEM, "batch handle");
	if (tile_units > 0)
		b->tile = tile_units;
	const int rc = shared_stream(&b->stream); // the device's one stream (capi.hip)
	if (rc) {
		delete b;
		return rc;
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
	if (nqueues < 1 || nqueues > kCountersMax || (nqueues & (nqueues - 1)))
		return set_error(SPMVHIP_EINVAL, "batch: %d ticket counters (a power of two, 1..%d)", nqueues, kCountersMax);
	b->queues = nqueues;
	return SPMVHIP_OK;
}

int spmvhip_batch_set_xcd_ranges(spmvhip_batch *b, int on)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	if (b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch already finalized");
	if (on < -1 || on > 1)
		return set_error(SPMVHIP_EINVAL, "batch: xcd_ranges %d (1 on, 0 off, -1 automatic)", on);
	b->xcd_ranges = on;
	return SPMVHIP_OK;
}

int spmvhip_batch_set_grid(spmvhip_batch *b, int workgroups_per_cu)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	if (b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch already finalized");
	if (workgroups_per_cu < 0 || workgroups_per_cu > 8)
		return set_error(SPMVHIP_EINVAL, "batch: %d workgroups per CU (0..8)", workgroups_per_cu);
	b->wgs_per_cu = workgroups_per_cu;
	return SPMVHIP_OK;
}

int spmvhip_batch_set_solo_nnz(spmvhip_batch *b, long nnz)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	if (b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch already finalized");
	if (!b->fmts.empty())
		return set_error(SPMVHIP_EINVAL, "batch: set the solo threshold before adding members");
	if (nnz < 0)
		return set_error(SPMVHIP_EINVAL, "batch: solo threshold %ld < 0", nnz);
	b->solo_nnz = nnz;
	return SPMVHIP_OK;
}

int spmvhip_batch_add(spmvhip_batch *b, const spmvhip_format *f, const double *d_x, double *d_y)
{
	if (!b || !f || !d_x || !d_y)
		return set_error(SPMVHIP_EINVAL, "NULL argument");
	if (b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch already finalized");
	if (b->solo_nnz > 0 && f->nnz >= b->solo_nnz && f->kind != SPMVHIP_CSR_NAIVE) {
		b->solo.push_back({const_cast<spmvhip_format *>(f), d_x, d_y});
		b->fmts.push_back(const_cast<spmvhip_format *>(f));
		b->total_nnz += f->nnz;
		return SPMVHIP_OK;
	}
	BatchMat M{};
	M.m = f->m;
	M.x = ptr_bits(d_x);
	M.y = ptr_bits(d_y);
	long units, nsegs = 0;
	int kind;
	if (f->kind == SPMVHIP_CSR_VEC) {
		if (f->block_nnz > kBatchBnnzMax)
			return set_error(SPMVHIP_EINVAL, "batch: csr_vec block_nnz %d > %d", f->block_nnz, kBatchBnnzMax);
		if (f->block_nnz > b->bnnz)
			b->bnnz = kBatchBnnzMax;
		kind = kUnitCsr;
		M.rp = ptr_bits(f->row_ptr.p);
		M.col = ptr_bits(f->col.p);
		M.val = ptr_bits(f->val.p);
		M.lanes = f->lanes;
		M.split_nnz = f->split_nnz;
		M.narrow = f->p.narrow_loads != 0;
		if (f->nsegs > 0) {
			M.seg_ptr = ptr_bits(f->seg_ptr.p);
			M.split_rows = ptr_bits(f->split_rows.p);
			M.partials = ptr_bits(f->partials.p);
			nsegs = f->nsegs;
		}
		units = f->nblocks;
	} else if (is_sell(f->kind)) {
		kind = kUnitSell;
		M.rp = ptr_bits(f->slice_ptr.p);
		M.sw = ptr_bits(f->slice_w.p);
		M.col = ptr_bits(f->s_cols.p);
		M.val = ptr_bits(f->s_vals.p);
		M.order = f->row_order.p ? ptr_bits(f->row_order.p) : 0;
		M.lanes = f->lanes; // lanes per row T (>= 1) of the slice layout
		units = f->nslices;
	} else if (f->kind == SPMVHIP_DIV_RF) {
		BatchDivMat D{};
		int buf = 0, acc = 0;
		SPMVHIP_TRY(div_batch_member(f, d_x, d_y, &D, &buf, &acc));
		b->div_xy.push_back({ptr_bits(d_x), ptr_bits(d_y)});
		if (b->host_div.size() >= (size_t) INT32_MAX || f->div_ngroups > INT32_MAX ||
				(long) b->host_div_units.size() + f->div_ngroups > INT32_MAX - (long) kCountersMax)
			return set_error(SPMVHIP_EINVAL, "batch: too many DIV packet groups");
		const int di = (int) b->host_div.size();
		b->host_div.push_back(D);
		for (long k = 0; k < f->div_ngroups; k++) // in the member's group order (packet reordering)
			b->host_div_units.push_back({di, f->div_order_host.empty() ? (int) k : f->div_order_host[k]});
		b->div_buf = std::max(b->div_buf, buf);
		b->div_acc = std::max(b->div_acc, acc);
		b->fmts.push_back(const_cast<spmvhip_format *>(f));
		b->div_fmts.push_back((int) b->fmts.size() - 1);
		b->total_nnz += f->nnz;
		return SPMVHIP_OK;
	} else {
		return set_error(SPMVHIP_EINVAL, "batch: format %s is not batchable (csr_vec, sell, sell_c_s, div_rf are)",
				f->name);
	}
	if (units > INT32_MAX || nsegs > INT32_MAX || b->host_mats.size() >= (size_t) kMatMask)
		return set_error(SPMVHIP_EINVAL, "batch: too many units or members");
	const int mi = (int) b->host_mats.size();
	b->host_mats.push_back(M);
	b->fmts.push_back(const_cast<spmvhip_format *>(f));
	// the member's segments (its longest work) first, then its units
	const double w = (double) f->nnz / (double) std::max(1L, units + nsegs);
	if (nsegs > 0) {
		b->spans.push_back({mi | kUnitSeg << kKindShift, b->nunits, nsegs, nullptr, nullptr,
				f->seg_desc.as<int32_t>(), w});
		b->nunits += nsegs;
		for (long i = 0; i < f->nsplit; i++)
			b->host_splits.push_back({mi, (int) i});
	}
	if (kind == kUnitCsr)
		b->spans.push_back({mi | kind << kKindShift, b->nunits, units, f->blocks.as<int32_t>(),
				f->row_ptr.as<int32_t>(), nullptr, w});
	else
		b->spans.push_back({mi | kind << kKindShift, b->nunits, units, nullptr, nullptr, nullptr, w});
	b->nunits += units;
	b->max_member_units = std::max(b->max_member_units, units + nsegs);
	b->total_nnz += f->nnz;
	return SPMVHIP_OK;
}

int spmvhip_batch_finalize(spmvhip_batch *b, int use_graph)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	if (b->finalized)
		return set_error(SPMVHIP_EINVAL, "batch already finalized");
	// persistent grid: wgs_per_cu resident workgroups per CU (at most what
	// fits), a multiple of the XCD count
	int dev = 0, cus = 0, per_cu = 0;
	SPMVHIP_CHECK(hipGetDevice(&dev));
	SPMVHIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
	if (b->bnnz > 256)
		SPMVHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<kBatchBnnzMax>, kBlock, 0));
	else
		SPMVHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<256>, kBlock, 0));
	const int want = b->wgs_per_cu > 0 ? b->wgs_per_cu : kGridPerCu;
	long g = (long) cus * std::max(1, std::min(per_cu, want));
	g = std::max<long>(kXcds, g / kXcds * kXcds);
	b->grid = (int) g;
	// ticket counters: the requested number, at most the resident waves (a
	// counter without a wave would leave its units undone)
	const long N = b->nunits, W = g * kWavesPerBlock;
	if (N + (long) b->tile * kCountersMax > INT32_MAX)
		return set_error(SPMVHIP_EINVAL, "batch: %ld work units (the kernel indexes them with 32 bits)", N);
	// automatic: per-XCD ranges when every member is small against the batch
	// (x locality per XCD, about 1 % on 2,048 small matrices); a member holding
	// more than 1/64 of the units keeps its units spread over all XCDs (on one
	// large random-placement matrix the ranges cost up to 7 %)
	const bool ranges_on = b->xcd_ranges == 1 || (b->xcd_ranges == -1 && b->max_member_units * 64 <= N);
	b->nranges = ranges_on ? kXcds : 1;
	b->nc = std::max(1, b->queues / b->nranges);
	while (b->nc > W / b->nranges)
		b->nc >>= 1;
	b->ranges = UnitRanges{};
	for (int q = 0; q <= kXcds; q++)
		b->ranges.b[q] = (int) N;
	b->ranges.b[0] = 0;
	if (b->nranges > 1) {
		// cut points of equal nonzero weight (each member's units weigh its
		// mean nonzeros per unit)
		double total = 0, acc = 0;
		for (const UnitSpan &sp : b->spans)
			total += sp.weight * (double) sp.count;
		int q = 1;
		for (const UnitSpan &sp : b->spans) {
			const double span_w = sp.weight * (double) sp.count;
			while (q < kXcds && acc + span_w >= total * q / kXcds) {
				const double need = total * q / kXcds - acc;
				const long k = sp.weight > 0 ? std::min(sp.count, (long) std::ceil(need / sp.weight)) : 0;
				b->ranges.b[q++] = (int) (sp.off + k);
			}
			acc += span_w;
		}
		for (int i = 1; i <= kXcds; i++)
			b->ranges.b[i] = std::max(b->ranges.b[i], b->ranges.b[i - 1]);
	}

	const size_t nm = b->host_mats.size();
	SPMVHIP_TRY(b->mats.alloc(std::max<size_t>(1, nm) * sizeof(BatchMat)));
	SPMVHIP_TRY(b->units.alloc(std::max<long>(1, N) * sizeof(UnitRec)));
	SPMVHIP_TRY(b->counter.alloc(kCountersMax * kCounterStride * sizeof(int)));
	if (nm)
		SPMVHIP_CHECK(hipMemcpy(b->mats.p, b->host_mats.data(), nm * sizeof(BatchMat), hipMemcpyHostToDevice));
	// the unit records, from each member's device tables
	for (const UnitSpan &sp : b->spans) {
		if (sp.count == 0)
			continue;
		UnitRec *out = b->units.as<UnitRec>() + sp.off;
		const unsigned grid = grid_for(sp.count, kBlock);
		const int n = (int) sp.count;
		if (sp.blocks)
			batch_fill_csr<<<grid, kBlock, 0, b->stream>>>(sp.blocks, sp.rp, n, sp.mk, out);
		else if (sp.seg_desc)
			batch_fill_seg<<<grid, kBlock, 0, b->stream>>>(sp.seg_desc, n, sp.mk, out);
		else
			batch_fill_sell<<<grid, kBlock, 0, b->stream>>>(n, sp.mk, out);
		SPMVHIP_CHECK(hipGetLastError());
	}
	SPMVHIP_CHECK(hipStreamSynchronize(b->stream));
	if (!b->host_splits.empty()) {
		SPMVHIP_TRY(b->splits.alloc(b->host_splits.size() * sizeof(SplitRef)));
		SPMVHIP_CHECK(hipMemcpy(b->splits.p, b->host_splits.data(), b->host_splits.size() * sizeof(SplitRef),
				hipMemcpyHostToDevice));
	}
	if (!b->host_div_units.empty()) {
		int per_cu = 0;
		SPMVHIP_TRY(div_batch_occupancy(b->div_buf + b->div_acc, &per_cu));
		if (per_cu < 1)
			return set_error(SPMVHIP_EINVAL, "batch: DIV members need %d bytes of LDS", b->div_buf + b->div_acc);
		b->div_grid = cus * per_cu;
		b->div_nc = 1;
		while (b->div_nc * 2 <= std::min(b->queues, b->div_grid))
			b->div_nc *= 2;
		SPMVHIP_TRY(b->div_mats.alloc(b->host_div.size() * sizeof(BatchDivMat)));
		SPMVHIP_TRY(b->div_units.alloc(b->host_div_units.size() * sizeof(BatchDivUnit)));
		SPMVHIP_TRY(b->div_counter.alloc((size_t) kCountersMax * kCounterStride * sizeof(int)));
		SPMVHIP_CHECK(hipMemcpy(b->div_mats.p, b->host_div.data(), b->host_div.size() * sizeof(BatchDivMat),
				hipMemcpyHostToDevice));
		SPMVHIP_CHECK(hipMemcpy(b->div_units.p, b->host_div_units.data(),
				b->host_div_units.size() * sizeof(BatchDivUnit), hipMemcpyHostToDevice));
	}
	b->finalized = true;
	if (use_graph) {
		// captured on a stream of its own, so that no other work enqueued
		// on the shared stream meanwhile can enter the graph; replayed on
		// the shared stream
		hipStream_t cap = nullptr;
		SPMVHIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
		hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
		int rc = SPMVHIP_OK;
		hipGraph_t graph = nullptr;
		if (e == hipSuccess) {
			rc = batch_launch(b, cap);
			e = hipStreamEndCapture(cap, &graph);
		}
		(void) hipStreamDestroy(cap);
		if (rc) {
			if (graph)
				(void) hipGraphDestroy(graph);
			return rc;
		}
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
		size_t mi = 0, di = 0;
		for (size_t i = 0; i < b->fmts.size(); i++) {
			uint64_t x, y;
			if (di < b->div_fmts.size() && b->div_fmts[di] == (int) i) {
				x = b->div_xy[di].first;
				y = b->div_xy[di].second;
				di++;
			} else {
				x = b->host_mats[mi].x;
				y = b->host_mats[mi].y;
				mi++;
			}
			SPMVHIP_TRY(spmvhip_spmv_device(b->fmts[i], (const double *) (uintptr_t) x, (double *) (uintptr_t) y,
					b->stream));
		}
		return SPMVHIP_OK;
	};
	// Iterations run back to back: an event between consecutive iterations
	// and one wait at the end. The warm-up iterations are still running when
	// the first event is recorded, so the host's submission of each
	// iteration (a graph launch costs the host far more than a kernel
	// launch) overlaps the GPU's work, and ms[i] is the GPU time of
	// iteration i, as in a loop that replays the launch.
	for (int i = 0; i < warmup; i++)
		SPMVHIP_TRY(once());
	std::vector<hipEvent_t> ev((size_t) iters + 1, nullptr);
	int rc = SPMVHIP_OK;
	hipError_t he = hipSuccess;
	for (size_t i = 0; i < ev.size() && he == hipSuccess; i++)
		he = hipEventCreate(&ev[i]);
	if (he == hipSuccess)
		he = hipEventRecord(ev[0], b->stream);
	for (int i = 0; i < iters && he == hipSuccess && rc == SPMVHIP_OK; i++) {
		rc = once();
		if (rc == SPMVHIP_OK)
			he = hipEventRecord(ev[(size_t) i + 1], b->stream);
	}
	if (he == hipSuccess && rc == SPMVHIP_OK)
		he = hipEventSynchronize(ev[(size_t) iters]);
	for (int i = 0; i < iters && he == hipSuccess && rc == SPMVHIP_OK; i++) {
		float t = 0;
		he = hipEventElapsedTime(&t, ev[(size_t) i], ev[(size_t) i + 1]);
		if (ms)
			ms[i] = t;
	}
	if (he != hipSuccess && rc == SPMVHIP_OK)
		rc = set_error(SPMVHIP_EHIP, "batch timing: %s", hipGetErrorString(he));
	if (rc != SPMVHIP_OK)
		(void) hipStreamSynchronize(b->stream); // nothing in flight references the events below
	for (hipEvent_t e : ev)
		if (e)
			(void) hipEventDestroy(e);
	return rc;
}

int spmvhip_batch_info(const spmvhip_batch *b, long *nmats, long *nitems, int *grid, long *total_nnz)
{
	if (!b)
		return set_error(SPMVHIP_EINVAL, "NULL batch");
	if (nmats)
		*nmats = (long) b->fmts.size();
	if (nitems)
		*nitems = b->nunits + (long) b->host_div_units.size();
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
	delete b; // the stream is the device's shared one: it stays
}

} // extern "C"

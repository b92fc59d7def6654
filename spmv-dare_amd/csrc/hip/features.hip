// Device-computed structure statistics of a CSR matrix: the cheap features of
// the format selector. Definitions follow the reference generator's
// descriptor statistics (artificial_matrix_generation.c:503-539,
// matrix_util.h:235-266 with window 1): row length mean / standard deviation
// / maximum, per-row bandwidth (last column - first column + 1, 0 for an
// empty row) and adjacent-column pairs. Two passes with fixed-order sums:
// per-workgroup partials, then one workgroup adds them in order.

#include <mutex>

#include "format.hpp"

namespace spmvhip {

struct StatPartial {
	double len, len2, bw, adj;
	long maxlen, nonempty;
};

__global__ void __launch_bounds__(kBlock) stats_rows(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
		long m, StatPartial *__restrict__ out)
{
	__shared__ StatPartial sh[kBlock];
	const long base = (long) blockIdx.x * kBlock * 16;
	StatPartial p{0, 0, 0, 0, 0, 0};
	for (int k = 0; k < 16; k++) {
		long r = base + (long) k * kBlock + threadIdx.x;
		if (r >= m)
			break;
		int s = rp[r], e = rp[r + 1], len = e - s;
		p.len += len;
		p.len2 += (double) len * len;
		p.maxlen = len > p.maxlen ? len : p.maxlen;
		if (len > 0) {
			p.nonempty++;
			p.bw += (double) (col[e - 1] - col[s] + 1);
			int adj = 0;
			for (int j = s + 1; j < e; j++)
				adj += (col[j] == col[j - 1] + 1);
			p.adj += adj;
		}
	}
	sh[threadIdx.x] = p;
	__syncthreads();
	for (int o = kBlock / 2; o > 0; o >>= 1) {
		if (threadIdx.x < o) {
			StatPartial &a = sh[threadIdx.x];
			const StatPartial &b = sh[threadIdx.x + o];
			a.len += b.len;
			a.len2 += b.len2;
			a.bw += b.bw;
			a.adj += b.adj;
			a.maxlen = b.maxlen > a.maxlen ? b.maxlen : a.maxlen;
			a.nonempty += b.nonempty;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0)
		out[blockIdx.x] = sh[0];
}

// The partials added in order, then the six statistics (no fused
// multiply-add: every product and sum is rounded where it is written).
__global__ void stats_final(const StatPartial *__restrict__ in, long n, long m, long ncols, double *__restrict__ out6)
{
#pragma clang fp contract(off)
	if (threadIdx.x != 0)
		return;
	StatPartial a{0, 0, 0, 0, 0, 0};
	for (long i = 0; i < n; i++) {
		a.len += in[i].len;
		a.len2 += in[i].len2;
		a.bw += in[i].bw;
		a.adj += in[i].adj;
		a.maxlen = in[i].maxlen > a.maxlen ? in[i].maxlen : a.maxlen;
		a.nonempty += in[i].nonempty;
	}
	const double avg = a.len / m;
	const double var = a.len2 / m - avg * avg;
	out6[0] = avg;                                     // mean row length
	out6[1] = var > 0 ? sqrt(var) : 0;                 // standard deviation of row lengths
	out6[2] = (double) a.maxlen;                       // longest row
	out6[3] = a.bw / m / ncols;                        // mean bandwidth / n
	out6[4] = a.len > 0 ? 2.0 * a.adj / a.len : 0;     // neighbours per nnz (window 1)
	out6[5] = (double) a.nonempty;                     // non-empty rows
}

constexpr long kStatsRowsPerBlock = (long) kBlock * 16;

// Scratch of the allocation-free entry point: the per-workgroup partials,
// at most kStatsScratchParts of them (rows up to kStatsScratchParts *
// kStatsRowsPerBlock = 67,108,864), 768 KiB, held by every GPU the library
// runs on until the process ends. Host code never uses this name as an
// address: stats_scratch_ptr() gives the device's own copy.
constexpr long kStatsScratchParts = 16384;
static __device__ __align__(256) StatPartial stats_scratch[kStatsScratchParts];

constexpr int kMaxDevices = 64;
static void *g_stats_scratch[kMaxDevices];
static std::mutex g_stats_scratch_mu;

// The address of `dev`'s copy of stats_scratch (that device current),
// looked up on the first call on each device and kept.
static int stats_scratch_ptr(int dev, StatPartial **out)
{
	if (dev < 0 || dev >= kMaxDevices)
		return set_error(SPMVHIP_EINVAL, "csr_stats: device %d beyond %d", dev, kMaxDevices);
	std::lock_guard<std::mutex> lock(g_stats_scratch_mu);
	if (!g_stats_scratch[dev]) {
		void *p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(stats_scratch));
		if (e != hipSuccess || !p)
			return set_error(SPMVHIP_EHIP, "csr_stats: scratch address on device %d: %s", dev, hipGetErrorString(e));
		g_stats_scratch[dev] = p;
	}
	*out = (StatPartial *) g_stats_scratch[dev];
	return SPMVHIP_OK;
}

} // namespace spmvhip

using namespace spmvhip;

extern "C" int spmvhip_csr_stats_device(const int32_t *d_row_ptr, const int32_t *d_col_ind, long m, long n,
		double *out6)
{
	if (!d_row_ptr || !out6 || m < 1 || n < 1)
		return set_error(SPMVHIP_EINVAL, "csr_stats: bad arguments");
	long nb = (m + kStatsRowsPerBlock - 1) / kStatsRowsPerBlock;
	DevBuf part, fin;
	SPMVHIP_TRY(part.alloc(nb * sizeof(StatPartial)));
	SPMVHIP_TRY(fin.alloc(6 * sizeof(double)));
	stats_rows<<<(unsigned) nb, kBlock>>>(d_row_ptr, d_col_ind, m, part.as<StatPartial>());
	SPMVHIP_CHECK(hipGetLastError());
	stats_final<<<1, kWave>>>(part.as<StatPartial>(), nb, m, n, fin.as<double>());
	SPMVHIP_CHECK(hipGetLastError());
	SPMVHIP_CHECK(hipMemcpy(out6, fin.p, 6 * sizeof(double), hipMemcpyDeviceToHost));
	return SPMVHIP_OK;
}

extern "C" int spmvhip_csr_stats_device_noalloc(const int32_t *d_row_ptr, const int32_t *d_col_ind, long m, long n,
		double *d_out6, void *stream)
{
	if (!d_row_ptr || !d_out6 || m < 1 || n < 1)
		return set_error(SPMVHIP_EINVAL, "csr_stats_noalloc: bad arguments");
	const long nb = (m + kStatsRowsPerBlock - 1) / kStatsRowsPerBlock;
	if (nb > kStatsScratchParts)
		return set_error(SPMVHIP_EINVAL, "csr_stats_noalloc: %ld rows exceed the scratch's %ld (use spmvhip_csr_stats_device)",
				m, kStatsScratchParts * kStatsRowsPerBlock);
	int dev = 0;
	SPMVHIP_CHECK(hipGetDevice(&dev));
	StatPartial *part = nullptr;
	SPMVHIP_TRY(stats_scratch_ptr(dev, &part));
	hipStream_t s = (hipStream_t) stream;
	// every partial stats_final reads is written by stats_rows of this call
	stats_rows<<<(unsigned) nb, kBlock, 0, s>>>(d_row_ptr, d_col_ind, m, part);
	SPMVHIP_CHECK(hipGetLastError());
	stats_final<<<1, kWave, 0, s>>>(part, nb, m, n, d_out6);
	SPMVHIP_CHECK(hipGetLastError());
	return SPMVHIP_OK;
}

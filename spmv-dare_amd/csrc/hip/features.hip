// Device-computed structure statistics of a CSR matrix: the cheap features of
// the format selector. Definitions follow the reference generator's
// descriptor statistics (artificial_matrix_generation.c:503-539,
// matrix_util.h:235-266 with window 1): row length mean / standard deviation
// / maximum, per-row bandwidth (last column - first column + 1, 0 for an
// empty row) and adjacent-column pairs. Two passes with fixed-order sums:
// per-workgroup partials, then one workgroup adds them in order.

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

__global__ void stats_final(const StatPartial *__restrict__ in, long n, StatPartial *__restrict__ out)
{
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
	*out = a;
}

} // namespace spmvhip

using namespace spmvhip;

extern "C" int spmvhip_csr_stats_device(const int32_t *d_row_ptr, const int32_t *d_col_ind, long m, long n,
		double *out6)
{
	if (!d_row_ptr || !out6 || m < 1 || n < 1)
		return set_error(SPMVHIP_EINVAL, "csr_stats: bad arguments");
	long nb = (m + kBlock * 16 - 1) / (kBlock * 16);
	DevBuf part, fin;
	SPMVHIP_TRY(part.alloc(nb * sizeof(StatPartial)));
	SPMVHIP_TRY(fin.alloc(sizeof(StatPartial)));
	stats_rows<<<(unsigned) nb, kBlock>>>(d_row_ptr, d_col_ind, m, part.as<StatPartial>());
	SPMVHIP_CHECK(hipGetLastError());
	stats_final<<<1, kWave>>>(part.as<StatPartial>(), nb, fin.as<StatPartial>());
	SPMVHIP_CHECK(hipGetLastError());
	StatPartial a;
	SPMVHIP_CHECK(hipMemcpy(&a, fin.p, sizeof(a), hipMemcpyDeviceToHost));
	double avg = a.len / m;
	double var = a.len2 / m - avg * avg;
	out6[0] = avg;                                     // mean row length
	out6[1] = var > 0 ? sqrt(var) : 0;                 // standard deviation of row lengths
	out6[2] = (double) a.maxlen;                       // longest row
	out6[3] = a.bw / m / n;                            // mean bandwidth / n
	out6[4] = a.len > 0 ? 2.0 * a.adj / a.len : 0;     // neighbours per nnz (window 1)
	out6[5] = (double) a.nonempty;                     // non-empty rows
	return SPMVHIP_OK;
}

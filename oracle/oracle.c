/*
 * CPU restatement of the reference SpMV path -- TEST INFRASTRUCTURE ONLY
 * (the checker; never linked into the product). Declarations, pinning notes
 * and reference citations: oracle.h. Compiled with -ffp-contract=off, so
 * every product and sum below is rounded exactly where it is written.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "oracle.h"

/* ------------------------------------------------------------ csr_naive */

void oracle_csr_naive(const int32_t *rp, const int32_t *col, const double *val,
		const double *x, double *y, long m)
{
	for (long i = 0; i < m; i++) {
		double sum = 0;
		for (long j = rp[i]; j < rp[i + 1]; j++) {
			double prod = val[j] * x[col[j]];
			sum = sum + prod;
		}
		y[i] = sum;
	}
}

/* ------------------------------------------------------------ double-double */

static inline void two_sum(double a, double b, double *s, double *e)
{
	double t = a + b;
	double bb = t - a;
	*e = (a - (t - bb)) + (b - bb);
	*s = t;
}

void oracle_csr_dd(const int32_t *rp, const int32_t *col, const double *val,
		const double *x, double *y_hi, double *y_lo, long m)
{
#pragma omp parallel for schedule(dynamic, 1024)
	for (long i = 0; i < m; i++) {
		double hi = 0, lo = 0;
		for (long j = rp[i]; j < rp[i + 1]; j++) {
			double a = val[j], b = x[col[j]];
			double p = a * b;
			double pe = fma(a, b, -p); /* exact product error */
			double s, e;
			two_sum(hi, p, &s, &e);
			e = e + (lo + pe);
			hi = s + e;             /* fast two-sum renormalisation */
			lo = e - (hi - s);
		}
		y_hi[i] = hi;
		y_lo[i] = lo;
	}
}

/* ------------------------------------------------------------ partitioners */

void oracle_partition_iterations(long W, long w, long start, long end, long *s, long *e)
{
	long len = end - start;
	if (1 > len) { /* |incr| >= len: worker 0 takes everything */
		*s = (w == 0) ? start : end;
		*e = end;
		return;
	}
	long per = len / W, rem = len % W;
	if (rem != 0 && w < rem) {
		per += 1;
		rem = 0;
	}
	long ls = start + per * w + rem;
	long le = ls + per;
	if (w == W - 1)
		le = end;
	*s = ls;
	*e = le;
}

/* closest index to `target` in the non-decreasing sums[lo..hi]; ties in
 * distance go to the upper index (macrolib.h:537-575, default cmp/dist). */
static long closest_index(const int64_t *A, long lo, long hi, int64_t target)
{
	long s = lo, e = hi, mid;
	if (target < A[s])
		return s;
	if (target > A[e])
		return e;
	for (;;) {
		mid = (s + e) / 2;
		if (mid == s || mid == e)
			break;
		if (target > A[mid])
			s = mid;
		else
			e = mid;
	}
	if (A[s] == target)
		return s;
	if (A[e] == target)
		return e;
	int64_t ds = target - A[s], de = A[e] - target;
	return ds < de ? s : e;
}

void oracle_partition_prefix_sums(long W, long w, const int64_t *sums, long N,
		int64_t total, long *s, long *e)
{
	int64_t target = sums[0] + (total * w) / W;
	int64_t target_next = sums[0] + (total * (w + 1)) / W;
	*s = (w == 0) ? 0 : closest_index(sums, 0, N - 1, target);
	*e = (w == W - 1) ? N : closest_index(sums, 0, N - 1, target_next);
}

/* ------------------------------------------------------------ OpenMP csr_naive */

int oracle_csr_naive_omp(const int32_t *rp, const int32_t *col, const double *val,
		const double *x, double *y, long m, int threads)
{
	int T = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(T)
	{
		long s, e;
		oracle_partition_iterations(omp_get_num_threads(), omp_get_thread_num(), 0, m, &s, &e);
		long j = rp[s];
		for (long i = s; i < e; i++) {
			long je = rp[i + 1];
			double sum = 0;
			for (; j < je; j++)
				sum = sum + val[j] * x[col[j]];
			y[i] = sum;
		}
	}
	return T;
}

/* ------------------------------------------------------------ SELL-C-sigma */

void oracle_sellcs_order(const int32_t *rp, long m, long sigma, int64_t *order)
{
	/* stable counting sort by descending length inside each window: the
	 * rows of one length keep their ascending id order */
	long win = sigma > 0 ? sigma : 1;
	for (long w0 = 0; w0 < m; w0 += win) {
		long wl = m - w0 < win ? m - w0 : win;
		int32_t mx = 0;
		for (long i = 0; i < wl; i++) {
			int32_t len = rp[w0 + i + 1] - rp[w0 + i];
			mx = len > mx ? len : mx;
		}
		long *start = calloc((size_t) mx + 2, sizeof(long));
		for (long i = 0; i < wl; i++)
			start[rp[w0 + i + 1] - rp[w0 + i]]++;
		long run = 0;
		for (long len = mx; len >= 0; len--) { /* longest first */
			long c = start[len];
			start[len] = run;
			run += c;
		}
		for (long i = 0; i < wl; i++) {
			int32_t len = rp[w0 + i + 1] - rp[w0 + i];
			order[w0 + start[len]++] = w0 + i;
		}
		free(start);
	}
}

long oracle_sellcs_layout(const int32_t *rp, long m, long C, const int64_t *order,
		int width_rule, int64_t *widths, int64_t *slice_ptr)
{
	long ns = (m + C - 1) / C;
	int64_t total = 0;
	for (long s = 0; s < ns; s++) {
		int64_t w;
		if (width_rule == 0) {
			int64_t r = order[s * C];
			w = rp[r + 1] - rp[r];
		} else {
			w = 0;
			for (long l = 0; l < C && s * C + l < m; l++) {
				int64_t r = order[s * C + l];
				int64_t k = rp[r + 1] - rp[r];
				w = k > w ? k : w;
			}
		}
		widths[s] = w;
		slice_ptr[s] = total;
		total += w * C;
	}
	slice_ptr[ns] = total;
	return total;
}

long oracle_sellcs_fill(const int32_t *rp, const int32_t *col, const double *val, long m,
		long C, const int64_t *order, const int64_t *widths, const int64_t *slice_ptr,
		int64_t *cols_out, double *vals_out)
{
	long ns = (m + C - 1) / C;
	long truncated = 0;
	memset(cols_out, 0, (size_t) slice_ptr[ns] * sizeof(int64_t));
	memset(vals_out, 0, (size_t) slice_ptr[ns] * sizeof(double));
	for (long r = 0; r < m; r++) {
		long s = r / C;
		int64_t at = slice_ptr[s] + (r % C);
		int64_t src = order[r];
		long len = rp[src + 1] - rp[src];
		long keep = len <= widths[s] ? len : widths[s];
		truncated += len - keep;
		for (long k = 0; k < keep; k++) {
			cols_out[at] = col[rp[src] + k];
			vals_out[at] = val[rp[src] + k];
			at += C;
		}
	}
	return truncated;
}

void oracle_sellcs_spmv(long m, long C, const int64_t *order, const int64_t *widths,
		const int64_t *slice_ptr, const int64_t *cols, const double *vals,
		const double *x, double *y)
{
	long ns = (m + C - 1) / C;
	for (long s = 0; s < ns; s++) {
		for (long l = 0; l < C && s * C + l < m; l++) {
			double acc = 0;
			const int64_t base = slice_ptr[s] + l;
			for (long k = 0; k < widths[s]; k++)
				acc = fma(vals[base + k * C], x[cols[base + k * C]], acc);
			y[order[s * C + l]] = acc;
		}
	}
}

/* ------------------------------------------------------------ SELL */

long oracle_sell_layout(const int32_t *rp, long m, long VL, int64_t *cluster_ptr)
{
	long nc = (m + VL - 1) / VL;
	int64_t total = 0;
	for (long c = 0; c < nc; c++) {
		int64_t w = 0;
		for (long i = c * VL; i < (c + 1) * VL && i < m; i++) {
			int64_t d = rp[i + 1] - rp[i];
			w = d > w ? d : w;
		}
		cluster_ptr[c] = total;
		total += VL * w;
	}
	cluster_ptr[nc] = total;
	return total;
}

void oracle_sell_fill(const int32_t *rp, const int32_t *col, const double *val, long m,
		long VL, long T, const int64_t *cluster_ptr, int64_t *ja, double *a)
{
	long nc = (m + VL - 1) / VL;
	int64_t nnz_ext = cluster_ptr[nc];
	int64_t *rowmajor_j = malloc((size_t) (nnz_ext > 0 ? nnz_ext : 1) * sizeof(int64_t));
	double *rowmajor_a = malloc((size_t) (nnz_ext > 0 ? nnz_ext : 1) * sizeof(double));
	for (long t = 0; t < T; t++) {
		long c0, c1;
		oracle_partition_prefix_sums(T, t, cluster_ptr, nc, nnz_ext, &c0, &c1);
		int64_t colv = 0; /* per-thread carry (sell.cpp:124,165) */
		for (long c = c0; c < c1; c++) {
			int64_t w = (cluster_ptr[c + 1] - cluster_ptr[c]) / VL;
			long i0 = VL * c, i1 = i0 + VL < m ? i0 + VL : m;
			int64_t jj = cluster_ptr[c];
			for (long i = i0; i < i1; i++) {
				long j = rp[i];
				for (; j < rp[i + 1]; j++, jj++) {
					rowmajor_a[jj] = val[j];
					colv = col[j];
					rowmajor_j[jj] = colv;
				}
				for (; j < rp[i] + w; j++, jj++) {
					rowmajor_a[jj] = 0;
					rowmajor_j[jj] = colv;
				}
			}
			for (; jj < cluster_ptr[c + 1]; jj++) {
				rowmajor_a[jj] = 0;
				rowmajor_j[jj] = colv;
			}
			/* transpose VL x w (row-major) into w x VL (sell.cpp:78-92,193-194) */
			for (int64_t k = 0; k < w; k++)
				for (long l = 0; l < VL; l++) {
					ja[cluster_ptr[c] + k * VL + l] = rowmajor_j[cluster_ptr[c] + l * w + k];
					a[cluster_ptr[c] + k * VL + l] = rowmajor_a[cluster_ptr[c] + l * w + k];
				}
		}
	}
	free(rowmajor_j);
	free(rowmajor_a);
}

void oracle_sell_spmv(long m, long VL, const int64_t *cluster_ptr, const int64_t *ja,
		const double *a, const double *x, double *y)
{
	long nc = (m + VL - 1) / VL;
	for (long c = 0; c < nc; c++) {
		for (long l = 0; l < VL && c * VL + l < m; l++) {
			double acc = 0;
			for (int64_t jj = cluster_ptr[c] + l; jj < cluster_ptr[c + 1]; jj += VL)
				acc = fma(a[jj], x[ja[jj]], acc);
			y[c * VL + l] = acc;
		}
	}
}

/* ------------------------------------------------------------ SELL_SORTED */

/* Row windows of the reference's threads (sell_sorted.cpp:160-176): T
 * nnz-balanced ranges of VL-row clusters, balanced on the cluster prefix
 * sums row_ptr[c * VL]; the last range ends at row m. */
void oracle_sellsorted_windows(const int32_t *rp, long m, long VL, long T, int64_t *row_start)
{
	long nc = (m + VL - 1) / VL;
	int64_t *sums = malloc((size_t) (nc > 0 ? nc : 1) * sizeof(int64_t));
	for (long c = 0; c < nc; c++)
		sums[c] = rp[c * VL];
	for (long t = 0; t < T; t++) {
		long c0 = 0, c1 = 0;
		if (nc > 0)
			oracle_partition_prefix_sums(T, t, sums, nc, rp[m], &c0, &c1);
		row_start[t] = c0 * VL < m ? c0 * VL : m;
	}
	row_start[T] = m;
	free(sums);
}

/* Stable ascending sort of the rows of every window [row_start[w],
 * row_start[w + 1]) by length (the degree-bucket sort of sell_sorted.cpp:
 * 177-193): perm[i] = sorted position of row i. */
void oracle_sellsorted_order(const int32_t *rp, long m, long nwin, const int64_t *row_start, int32_t *perm)
{
	(void) m;
	for (long w = 0; w < nwin; w++) {
		long i0 = row_start[w], i1 = row_start[w + 1];
		int32_t dmax = 0;
		for (long i = i0; i < i1; i++)
			dmax = rp[i + 1] - rp[i] > dmax ? rp[i + 1] - rp[i] : dmax;
		long *cnt = calloc((size_t) dmax + 2, sizeof(long));
		for (long i = i0; i < i1; i++)
			cnt[rp[i + 1] - rp[i] + 1]++;
		for (long d = 0; d <= dmax; d++)
			cnt[d + 1] += cnt[d];
		for (long i = i0; i < i1; i++)
			perm[i] = (int32_t) (i0 + cnt[rp[i + 1] - rp[i]]++);
		free(cnt);
	}
}

/* Cluster sizes VL * (longest sorted row of the cluster) and their prefix
 * sums (sell_sorted.cpp:227-251). Returns nnz_ext. */
long oracle_sellsorted_layout(const int32_t *rp, long m, long VL, const int32_t *perm, int64_t *cluster_ptr)
{
	long nc = (m + VL - 1) / VL;
	int64_t *w = calloc((size_t) (nc > 0 ? nc : 1), sizeof(int64_t));
	for (long i = 0; i < m; i++) {
		int64_t d = rp[i + 1] - rp[i];
		long c = perm[i] / VL;
		w[c] = d > w[c] ? d : w[c];
	}
	int64_t total = 0;
	for (long c = 0; c < nc; c++) {
		cluster_ptr[c] = total;
		total += VL * w[c];
	}
	cluster_ptr[nc] = total;
	free(w);
	return total;
}

/* Column-major fill of each cluster (sell_sorted.cpp:252-294): the sorted
 * rows' entries, then padding (value 0, column = the last column written in
 * this window so far, 0 at the window's start), missing rows of the last
 * cluster likewise. */
void oracle_sellsorted_fill(const int32_t *rp, const int32_t *col, const double *val, long m, long VL,
		long nwin, const int64_t *row_start, const int32_t *perm, const int64_t *cluster_ptr,
		int64_t *ja, double *a)
{
	long nc = (m + VL - 1) / VL;
	int32_t *rev = malloc((size_t) (m > 0 ? m : 1) * sizeof(int32_t));
	for (long i = 0; i < m; i++)
		rev[perm[i]] = (int32_t) i;
	int64_t nnz_ext = cluster_ptr[nc];
	int64_t *rj = malloc((size_t) (nnz_ext > 0 ? nnz_ext : 1) * sizeof(int64_t));
	double *ra = malloc((size_t) (nnz_ext > 0 ? nnz_ext : 1) * sizeof(double));
	for (long win = 0; win < nwin; win++) {
		long c0 = row_start[win] / VL, c1 = (row_start[win + 1] + VL - 1) / VL;
		if (row_start[win] >= row_start[win + 1])
			continue;
		int64_t colv = 0;
		for (long c = c0; c < c1; c++) {
			int64_t w = (cluster_ptr[c + 1] - cluster_ptr[c]) / VL;
			long i0 = VL * c, i1 = i0 + VL < m ? i0 + VL : m;
			int64_t jj = cluster_ptr[c];
			for (long i = i0; i < i1; i++) {
				long r = rev[i], j = rp[r], k = 0;
				for (; j < rp[r + 1]; j++, jj++, k++) {
					ra[jj] = val[j];
					colv = col[j];
					rj[jj] = colv;
				}
				for (; k < w; k++, jj++) {
					ra[jj] = 0;
					rj[jj] = colv;
				}
			}
			for (; jj < cluster_ptr[c + 1]; jj++) {
				ra[jj] = 0;
				rj[jj] = colv;
			}
			for (int64_t k = 0; k < w; k++)
				for (long l = 0; l < VL; l++) {
					ja[cluster_ptr[c] + k * VL + l] = rj[cluster_ptr[c] + l * w + k];
					a[cluster_ptr[c] + k * VL + l] = ra[cluster_ptr[c] + l * w + k];
				}
		}
	}
	free(rj);
	free(ra);
	free(rev);
}

/* compute_sell of sell_sorted.cpp:346-427: per cluster, lane-wise fma over
 * the width; sorted row i's result goes to y[rev_permutation[i]]. */
void oracle_sellsorted_spmv(long m, long VL, const int64_t *cluster_ptr, const int64_t *ja, const double *a,
		const int32_t *perm, const double *x, double *y)
{
	for (long r = 0; r < m; r++) {
		long i = perm[r], c = i / VL, l = i % VL;
		double acc = 0;
		for (int64_t jj = cluster_ptr[c] + l; jj < cluster_ptr[c + 1]; jj += VL)
			acc = fma(a[jj], x[ja[jj]], acc);
		y[r] = acc;
	}
}

/* ------------------------------------------------------------ DIV lossy values */

double oracle_div_reduce_precision(double v, double tolerance)
{
	union {
		double d;
		uint64_t u;
	} a, lo, hi;
	a.d = v;
	if (tolerance >= 1.0)
		return NAN;
	if (tolerance == 0 || a.u == 0)
		return v;
	if (__builtin_ctzll(a.u) >= 53)
		return v;
	uint64_t bits = (uint64_t) ceil(fabs(log2(tolerance)));
	if (bits >= 52)
		return v;
	uint64_t step = 1ULL << (52 - bits);
	lo.u = a.u & ~(step - 1);
	hi.u = (a.u + (step - 1)) & ~(step - 1);
	if (fabs(a.d - lo.d) / fabs(a.d) > tolerance || fabs(a.d - hi.d) / fabs(a.d) > tolerance)
		return NAN;
	int tz_lo = lo.u ? __builtin_ctzll(lo.u) : 64;
	int tz_hi = hi.u ? __builtin_ctzll(hi.u) : 64;
	return tz_lo > tz_hi ? lo.d : hi.d;
}

/* ------------------------------------------------------------ fixtures */

uint64_t oracle_fnv1a(const void *p, long nbytes)
{
	const unsigned char *b = (const unsigned char *) p;
	uint64_t h = 1469598103934665603ULL;
	for (long i = 0; i < nbytes; i++) {
		h ^= b[i];
		h *= 1099511628211ULL;
	}
	return h;
}

/* ------------------------------------------------------------ DIV-RF decode */

static uint64_t le_bytes(const unsigned char *p, long n)
{
	uint64_t v = 0;
	for (long b = 0; b < n; b++)
		v |= (uint64_t) p[b] << (8 * b);
	return v;
}

static uint32_t le32(const unsigned char *p)
{
	return (uint32_t) le_bytes(p, 4);
}

long oracle_div_decode(const unsigned char *bytes, long nbytes, long L, int32_t *rows, int32_t *cols,
		double *vals, long cap)
{
	long at = 0, out = 0;
	while (at < nbytes) {
		const unsigned char *pk = bytes + at;
		long nv = le32(pk), nu = le32(pk + 4);
		long row_min = le32(pk + 12), col_min = le32(pk + 16);
		long row_bits = pk[20], col_bits = pk[21], nrfs = le32(pk + 24);
		long cb = (row_bits + col_bits + 7) >> 3;
		const unsigned char *rfs = pk + 28 + 4 * L;
		const unsigned char *coords = rfs + 8 * nrfs;
		const unsigned char *lens = coords + nv * cb;
		const unsigned char *streams = lens + nu;
		long lane_start[4096], cursor[4096], total = 0;
		if (L > 4096 || out + nv > cap)
			return -1;
		for (long l = 0; l < L; l++) {
			lane_start[l] = total;
			cursor[l] = 0;
			total += le32(pk + 28 + 4 * l);
		}
		long ks = 0, cs = 0;
		for (long g = 0; g < nrfs; g++) {
			long rf = le32(rfs + 4 * g), n = le32(rfs + 4 * (nrfs + g));
			uint64_t prev[4096];
			for (long l = 0; l < L; l++)
				prev[l] = 0;
			/* position q of the group: lane q % L; its coordinates follow the
			 * element-wise interleave of lib/macros/permutation.h:47-111 */
			long dv = n / L, mul_c = dv * rf * L;
			for (long q = 0; q < n; q++) {
				long lane = q % L, pp = q / L;
				unsigned lb = lens[ks + q];
				long len = lb & 15, tz = (lb >> 4) << 2;
				if (len) {
					prev[lane] += le_bytes(streams + lane_start[lane] + cursor[lane], len) << tz;
					cursor[lane] += len;
				}
				double d;
				memcpy(&d, &prev[lane], 8);
				for (long t = 0; t < rf; t++) {
					long cpos = pp < dv ? (pp * rf + t) * L + lane : mul_c + lane * rf + t;
					uint64_t c = le_bytes(coords + (cs + cpos) * cb, cb);
					uint64_t mask = row_bits ? ((1ull << row_bits) - 1) : 0;
					rows[out + cs + cpos] = (int32_t) (row_min + (long) (c & mask));
					cols[out + cs + cpos] = (int32_t) (col_min + (long) (c >> row_bits));
					vals[out + cs + cpos] = d;
				}
			}
			ks += n;
			cs += n * rf;
		}
		out += nv;
		at += 28 + 4 * L + 8 * nrfs + nv * cb + nu + total;
	}
	return out;
}

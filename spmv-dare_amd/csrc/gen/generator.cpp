// Host-side artificial matrix generator: a from-scratch restatement of the
// reference generator's algorithm (artificial-matrix-generator/
// artificial_matrix_generation.c:163-579, random draws of random.h:127-307),
// producing bit-identical CSR structure and raw values.
//
// What has to match the reference exactly, and why it is written the way it is:
//  * the random stream: glibc random_r() with a 256-byte TYPE_4 state, reseeded
//    with srandom_r(seed + i) at every anchor row (i % period == 0); glibc is
//    called directly, so the integer stream is the system library's own;
//  * the floating-point expressions feeding floor()/comparisons (degree model,
//    Box-Muller, Marsaglia-Tsang, bandwidth correction) keep the reference's
//    operation order; the file is compiled with -ffp-contract=off so that no
//    multiply-add is fused where the reference build does not fuse;
//  * the insertion semantics of the column set (reject duplicates, return the
//    set sorted) -- the reference uses an AVL tree; any set with the same
//    insert/membership answers yields the same matrix, so a stamped
//    open-addressing hash table plus one sort per row is used here.
//
// Parallelism: rows are cut into anchor blocks [k*P, (k+1)*P); each block only
// depends on its own random stream (reseeded at its first row) and on the
// column prefix sum, so blocks are generated independently in any order and
// the result never depends on the thread count.

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

#include "spmvgen.h"

namespace {

thread_local char g_err[512];

int fail(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
	return code;
}

constexpr long kRandMax = 0x7FFFFFFFL;               // random() returns [0, 2^31)
constexpr int64_t kRandMax64 = 0x3FFFFFFFFFFFFFFFLL; // two 31-bit draws concatenated

// One glibc random_r stream (random.h:127-152: statelen 256, initstate_r then srandom_r).
struct Stream {
	struct random_data buf;
	char state[256];

	explicit Stream(unsigned int seed)
	{
		memset(&buf, 0, sizeof(buf));
		memset(state, 0, sizeof(state));
		buf.state = nullptr;
		initstate_r(seed, state, sizeof(state), &buf);
		srandom_r(seed, &buf);
	}
	void reseed(unsigned int seed) { srandom_r(seed, &buf); }
	int32_t next()
	{
		int32_t v;
		random_r(&buf, &v);
		return v;
	}
	// random.h:236-246: val / 2^31 scaled to [lo, hi).
	double uniform(double lo, double hi)
	{
		double res = ((double) next()) / ((double) (kRandMax + 1L));
		res = (hi - lo) * res + lo;
		return res;
	}
	// random.h:184-232: [lo, hi) integers with modulo-bias rejection.
	long uniform_int(long lo, long hi)
	{
		int64_t range = hi - lo;
		if (range >= kRandMax) {
			int64_t lim = kRandMax64 - kRandMax64 % range;
			int64_t r;
			do {
				r = next();
				r <<= 31;
				r |= next();
			} while (r >= lim);
			return (long) (r % range + lo);
		}
		int64_t lim = kRandMax - kRandMax % range;
		int32_t v;
		do {
			v = next();
		} while (v >= lim);
		return (long) (v % range + lo);
	}
	// Box-Muller, first output only (random.h:252-265).
	double normal(double mean, double sd)
	{
		double u1 = uniform(2 * DBL_EPSILON, 1);
		double u2 = uniform(0, 1);
		double mag = sd * sqrt(-2.0 * log(u1));
		return mag * cos(2.0 * M_PI * u2) + mean;
	}
	// Marsaglia-Tsang with the k < 1 boost (random.h:280-307).
	double gamma(double k, double theta)
	{
		if (k < 1) {
			double u = uniform(2 * DBL_EPSILON, 1);
			return gamma(1.0 + k, theta) * pow(u, 1.0 / k);
		}
		double x, v, u;
		double d = k - 1.0 / 3.0;
		double c = (1.0 / 3.0) / sqrt(d);
		for (;;) {
			do {
				x = normal(0, 1);
				v = 1.0 + c * x;
			} while (v <= 0);
			v = v * v * v;
			u = uniform(2 * DBL_EPSILON, 1);
			if (u < 1 - 0.0331 * x * x * x * x)
				break;
			if (log(u) < 0.5 * x * x + d * (1 - v + log(v)))
				break;
		}
		return theta * d * v;
	}
};

// Expected bandwidth of n uniformly placed points in an interval of B slots,
// summed in log2 space (artificial_matrix_generation.c:106-129).
double expected_span(double B, double n)
{
	double be = 0;
	for (long i = 0; i < (long) n; i++)
		be += log2(n - i) - log2(B - i);
	double b = pow(2, be);
	double E = 0;
	for (long i = (long) n; i <= (long) B; i++) {
		double x = (double) i;
		double a = (B - x + 1) * n / x * (n - 1) / (x - 1);
		E += a * b * (double) i;
		be += log2(x + 1) - log2(x + 1 - n);
		b = pow(2, be);
	}
	return E;
}

// Range correction so that the expected span of n points equals B
// (artificial_matrix_generation.c:132-150).
double corrected_range(double B, double n)
{
	n = floor(n + 0.5);
	if (n < 2)
		n = 2;
	double E = expected_span(B, n);
	if (E == 0)
		return n;
	double ratio = B / E;
	return B * ratio;
}

// Parameters of the degree model: an exponential head of height MAX decaying
// with rate C over the rows, plus a normal or gamma body
// (artificial_matrix_generation.c:211-237).
struct DegreeModel {
	double MAX, C, avg_distr, std_distr;
	bool normal;

	DegreeModel(double avg, double sd, const char *distribution, double skew)
	{
		normal = (*distribution == 'n');
		MAX = avg * (1 + skew);
		double c1 = MAX / avg;
		double c2 = (sd > 0) ? MAX * MAX / (2 * sd * sd) : 0;
		C = (c1 > c2) ? c1 : c2;
		if (C <= 2 || skew < 1) {
			MAX = 0;
			C = 0;
			avg_distr = avg;
			std_distr = sd;
		} else {
			double avg_exp = MAX / C;
			double std_exp = sqrt(avg_exp * avg_exp * (C / 2 - 1));
			avg_distr = avg - MAX / C;
			std_distr = sqrt(sd * sd - std_exp * std_exp);
			if (normal && std_distr > avg_distr / 3)
				std_distr = avg_distr / 3;
		}
	}

	// artificial_matrix_generation.c:289-311 (one row; stream already positioned).
	long degree(Stream &rs, long i, long nr_rows, long nr_cols) const
	{
		double d_exp = MAX * exp(-C / nr_rows * (double) i);
		double d_distr;
		if (std_distr > 0) {
			if (normal) {
				d_distr = rs.normal(avg_distr, std_distr);
			} else {
				double k = (avg_distr * avg_distr) / (std_distr * std_distr);
				double theta = (avg_distr > 0) ? (std_distr * std_distr) / avg_distr : 1;
				d_distr = rs.gamma(k, theta);
			}
		} else {
			d_distr = avg_distr;
		}
		double d = d_exp + d_distr;
		d = (d > 0) ? floor(d + 0.5) : 0;
		if (d > nr_cols)
			d = (double) nr_cols;
		return (long) d;
	}
};

// Anchor period: rows are reseeded every P rows (artificial_matrix_generation.c:261-268).
// The reference divides by zero for nr_rows < 100; here one anchor is used instead.
long anchor_period(long nr_rows)
{
	long anchors = nr_rows / 100;
	if (anchors < 1)
		anchors = 1;
	else if (anchors > (1L << 12))
		anchors = 1L << 12;
	return nr_rows / anchors;
}

// Column set of one row: membership + insertion list, cleared in O(1) by stamping.
struct ColumnSet {
	std::vector<int32_t> key;
	std::vector<uint32_t> stamp;
	std::vector<int32_t> items;
	uint32_t cur = 0;
	uint32_t mask = 0;

	void reserve(long max_degree)
	{
		size_t cap = 16;
		while (cap < (size_t) (2 * max_degree + 2))
			cap <<= 1;
		key.assign(cap, 0);
		stamp.assign(cap, 0);
		mask = (uint32_t) (cap - 1);
		items.reserve((size_t) max_degree);
	}
	void clear()
	{
		items.clear();
		if (++cur == 0) { // stamp wrapped: really clear once every 2^32 rows
			std::fill(stamp.begin(), stamp.end(), 0u);
			cur = 1;
		}
	}
	bool insert(int32_t v)
	{
		uint32_t h = ((uint32_t) v * 0x9E3779B1u) & mask;
		while (stamp[h] == cur) {
			if (key[h] == v)
				return false;
			h = (h + 1) & mask;
		}
		stamp[h] = cur;
		key[h] = v;
		items.push_back(v);
		return true;
	}
	size_t size() const { return items.size(); }
};

struct GenParams {
	long nr_rows, nr_cols;
	double avg, sd;
	const char *distribution;
	unsigned int seed;
	const char *placement;
	double bw_scaled, skew, neigh, crs;
};

// Phase 1: degrees of every row (artificial_matrix_generation.c:283-317).
void compute_degrees(const GenParams &g, const DegreeModel &dm, long P, int32_t *deg)
{
	long nblocks = (g.nr_rows + P - 1) / P;
#pragma omp parallel
	{
		Stream rs((unsigned int) omp_get_thread_num());
#pragma omp for schedule(dynamic, 4)
		for (long b = 0; b < nblocks; b++) {
			long i_s = b * P, i_e = std::min(g.nr_rows, i_s + P);
			for (long i = i_s; i < i_e; i++) {
				if (i % P == 0)
					rs.reseed((unsigned int) (g.seed + i));
				deg[i] = (int32_t) dm.degree(rs, i, g.nr_rows, g.nr_cols);
			}
		}
	}
}

} // namespace

extern "C" const char *spmvgen_last_error(void)
{
	return g_err;
}

namespace spmvgen_detail {
// the last-error message of the library's other entry points (mtx.cpp)
int set_error(int code, const char *msg)
{
	snprintf(g_err, sizeof(g_err), "%s", msg);
	return code;
}
} // namespace spmvgen_detail

extern "C" int spmvgen_degrees(long nr_rows, long nr_cols, double avg, double sd,
		const char *distribution, unsigned int seed, double skew, int32_t *deg, long *nnz_out)
{
	if (nr_rows <= 0 || nr_cols <= 0 || !distribution || !deg)
		return fail(SPMVGEN_EINVAL, "spmvgen_degrees: bad arguments");
	GenParams g{nr_rows, nr_cols, avg, sd, distribution, seed, "", 0, skew, 0, 0};
	DegreeModel dm(avg, sd, distribution, skew);
	compute_degrees(g, dm, anchor_period(nr_rows), deg);
	long s = 0;
	for (long i = 0; i < nr_rows; i++)
		s += deg[i];
	if (nnz_out)
		*nnz_out = s;
	return SPMVGEN_OK;
}

extern "C" int spmvgen_generate(long nr_rows, long nr_cols, double avg, double sd,
		const char *distribution, unsigned int seed, const char *placement, double bw_scaled,
		double skew, double neigh, double crs, int compute_stats, spmvgen_csr *out)
{
	if (!out)
		return fail(SPMVGEN_EINVAL, "spmvgen_generate: out is NULL");
	memset(out, 0, sizeof(*out));
	if (nr_rows <= 0 || nr_cols <= 0 || nr_cols > INT_MAX || nr_rows >= INT_MAX)
		return fail(SPMVGEN_EINVAL, "spmvgen_generate: bad shape %ld x %ld", nr_rows, nr_cols);
	if (!distribution || !placement)
		return fail(SPMVGEN_EINVAL, "spmvgen_generate: distribution/placement is NULL");
	if (bw_scaled > 1) // artificial_matrix_generation.c:202-203
		bw_scaled = 1;
	GenParams g{nr_rows, nr_cols, avg, sd, distribution, seed, placement, bw_scaled, skew, neigh, crs};
	DegreeModel dm(avg, sd, distribution, skew);
	const long P = anchor_period(nr_rows);
	const double p_right = neigh / 2;                 // :181
	const double cluster_size = 1 / (1 - p_right);    // :182

	int32_t *row_ptr = (int32_t *) malloc((size_t) (nr_rows + 1) * sizeof(int32_t));
	if (!row_ptr)
		return fail(SPMVGEN_ENOMEM, "spmvgen_generate: row_ptr allocation");
	compute_degrees(g, dm, P, row_ptr);

	// Serial prefix sum of the degrees (:322-336).
	long nnz = 0, max_degree = 0;
	for (long i = 0; i < nr_rows; i++) {
		long d = row_ptr[i];
		if (d > max_degree)
			max_degree = d;
		row_ptr[i] = (int32_t) nnz;
		nnz += d;
		if (nnz > INT_MAX) {
			free(row_ptr);
			return fail(SPMVGEN_EOVERFLOW, "spmvgen_generate: nnz exceeds int32 (%ld rows)", nr_rows);
		}
	}
	row_ptr[nr_rows] = (int32_t) nnz;

	int32_t *col_ind = (int32_t *) malloc((size_t) (nnz ? nnz : 1) * sizeof(int32_t));
	double *values = (double *) malloc((size_t) (nnz ? nnz : 1) * sizeof(double));
	if (!col_ind || !values) {
		free(row_ptr);
		free(col_ind);
		free(values);
		return fail(SPMVGEN_ENOMEM, "spmvgen_generate: %ld-nnz allocation", nnz);
	}

	// Range correction with the realised average degree (:347-349).
	double avg_real = ((double) nnz) / nr_rows;
	const double bw_corr = corrected_range(bw_scaled * nr_cols, avg_real / cluster_size);

	int err = 0;
	const char kind = *placement;
	const long nblocks = (nr_rows + P - 1) / P;

	// Phase 2: column placement (:362-523).
#pragma omp parallel
	{
		Stream rs((unsigned int) omp_get_thread_num());
		ColumnSet set;
		set.reserve(max_degree);
#pragma omp for schedule(dynamic, 4)
		for (long b = 0; b < nblocks; b++) {
			if (__atomic_load_n(&err, __ATOMIC_RELAXED))
				continue;
			long i_s = b * P, i_e = std::min(nr_rows, i_s + P);
			for (long i = i_s; i < i_e; i++) {
				if (i % P == 0)
					rs.reseed((unsigned int) (seed + i));
				long j_s = row_ptr[i], j_e = row_ptr[i + 1];
				long degree = j_e - j_s;
				if (degree == 0)
					continue;
				set.clear();

				long range = (long) floor(bw_corr + 0.5);
				if (range < degree)
					range = degree;
				long half = range / 2;

				if (kind == 's') { // all nnz at the start of the row (:387-395)
					for (long j = j_s; j < j_e; j++) {
						col_ind[j] = (int32_t) (j - j_s);
						values[j] = rs.uniform(0, 1);
					}
					continue;
				}
				long rl, rr; // relaxed bounds (may leave the matrix)
				if (kind == 'd') {
					rl = i - half;
					rr = i + half + range % 2;
				} else {
					rl = nr_cols / 2 - half;
					rr = nr_cols / 2 + half + range % 2;
				}
				long bl = rl, br = rr;
				if (bl < 0) {
					bl = 0;
					if (br < degree)
						br = degree;
				}
				if (br > nr_cols) {
					br = nr_cols;
					if (bl > nr_cols - degree)
						bl = nr_cols - degree;
				}

				long j = j_s;
				// Copy of the previous row's leading columns (:423-445).
				if ((crs > 0) && (i % P != 0) && (rs.uniform(0, 1) <= crs)) {
					long dprev = row_ptr[i] - row_ptr[i - 1];
					long dmin = degree < dprev ? degree : dprev;
					long ps = j_s - dprev, pe = ps + dmin;
					for (; ps < pe; ps++) {
						long k = col_ind[ps];
						if (k >= br)
							k = bl;
						if (!set.insert((int32_t) k)) {
							__atomic_store_n(&err, 1, __ATOMIC_RELAXED);
							break;
						}
						values[j++] = rs.uniform(0, 1);
					}
					if (__atomic_load_n(&err, __ATOMIC_RELAXED))
						break;
				}

				long d1 = 0, d2 = 0;
				while (j < j_e) {
					long k;
					long retries = 0;
					do { // random probes, then a deterministic sweep from both bounds (:451-476)
						if (retries < 20) {
							k = rs.uniform_int(rl, rr);
							if (k >= br)
								k = br - 1;
							else if (k < bl)
								k = bl;
							retries++;
						} else if (d1 < d2) {
							k = bl + d1;
							d1++;
						} else {
							k = br - d2 - 1;
							d2++;
						}
					} while (!set.insert((int32_t) k));
					values[j++] = rs.uniform(0, 1);

					if (p_right > 0) { // right-neighbour clustering (:481-496)
						while (j < j_e) {
							if (rs.uniform(0, 1) > p_right)
								break;
							k++;
							if (k >= br)
								k = bl;
							if (set.insert((int32_t) k))
								values[j++] = rs.uniform(0, 1);
						}
					}
				}
				std::sort(set.items.begin(), set.items.end());
				memcpy(&col_ind[j_s], set.items.data(), (size_t) degree * sizeof(int32_t));
			}
		}
	}
	if (err) {
		free(row_ptr);
		free(col_ind);
		free(values);
		return fail(SPMVGEN_EDUP, "spmvgen_generate: cross-row copy hit an occupied column");
	}

	out->m = nr_rows;
	out->n = nr_cols;
	out->nnz = nnz;
	out->row_ptr = row_ptr;
	out->col_ind = col_ind;
	out->values = values;
	if (compute_stats) {
		int rc = spmvgen_compute_stats(row_ptr, col_ind, nr_rows, nr_cols, &out->stats);
		if (rc)
			return rc;
	}
	out->stats.bw_corrected = bw_corr;
	return SPMVGEN_OK;
}

extern "C" void spmvgen_free(spmvgen_csr *csr)
{
	if (!csr)
		return;
	free(csr->row_ptr);
	free(csr->col_ind);
	free(csr->values);
	csr->row_ptr = nullptr;
	csr->col_ind = nullptr;
	csr->values = nullptr;
}

extern "C" void spmvgen_map_values(double *values, long nnz)
{
#pragma omp parallel for schedule(static)
	for (long i = 0; i < nnz; i++)
		values[i] = 2.0 * values[i] - 1.0;
}

extern "C" int spmvgen_fill_x(double *x, long n, unsigned int seed)
{
	if (!x || n < 0)
		return fail(SPMVGEN_EINVAL, "spmvgen_fill_x: bad arguments");
	Stream rs(seed);
	for (long i = 0; i < n; i++)
		x[i] = 2.0 * rs.uniform(0, 1) - 1.0;
	return SPMVGEN_OK;
}

// Descriptor statistics; definitions of artificial_matrix_generation.c:503-570
// and matrix_util.h:235-408 (window 1). Sums are serial in row order, so the
// values are deterministic (the reference's OpenMP sums differ from these in
// the last bits).
extern "C" int spmvgen_compute_stats(const int32_t *row_ptr, const int32_t *col_ind,
		long m, long n, spmvgen_stats *s)
{
	if (!row_ptr || !s || m <= 0)
		return fail(SPMVGEN_EINVAL, "spmvgen_compute_stats: bad arguments");
	std::vector<double> bw(m), sc(m), ng(m), cl(m), sim(m), nb(m);
	long nonempty = 0;
#pragma omp parallel for schedule(static) reduction(+ : nonempty)
	for (long i = 0; i < m; i++) {
		long js = row_ptr[i], je = row_ptr[i + 1], d = je - js;
		bw[i] = sc[i] = ng[i] = cl[i] = sim[i] = nb[i] = 0;
		if (d <= 0)
			continue;
		nonempty++;
		double b = (double) (col_ind[je - 1] - col_ind[js] + 1);
		bw[i] = b;
		sc[i] = d / b;
		long groups = 1, neigh = 0;
		for (long j = js + 1; j < je; j++) {
			if (col_ind[j] > col_ind[j - 1] + 1)
				groups++;
			else
				neigh += 2; // window 1: the next column is adjacent
		}
		ng[i] = (double) groups;
		cl[i] = (double) groups / d;
		nb[i] = (double) neigh;
		long l = i + 1; // next non-empty row
		while (l < m && row_ptr[l + 1] - row_ptr[l] <= 0)
			l++;
		if (l < m) {
			long ks = row_ptr[l], ke = row_ptr[l + 1], k = ks, hits = 0;
			for (long j = js; j < je; j++) {
				while (k < ke) {
					long diff = (long) col_ind[k] - col_ind[j];
					if (labs(diff) <= 1) {
						hits++;
						break;
					}
					if (diff <= 0)
						k++;
					else
						break;
				}
			}
			sim[i] = (double) hits / d;
		}
	}
	long nnz = row_ptr[m] - row_ptr[0];
	double sum_d = 0, sum_d2 = 0, mn = 1e300, mx = -1e300;
	double sbw = 0, ssc = 0, sng = 0, scl = 0, ssim = 0, snb = 0;
	for (long i = 0; i < m; i++) {
		double d = (double) (row_ptr[i + 1] - row_ptr[i]);
		sum_d += d;
		mn = d < mn ? d : mn;
		mx = d > mx ? d : mx;
		sbw += bw[i];
		ssc += sc[i];
		sng += ng[i];
		scl += cl[i];
		ssim += sim[i];
		snb += nb[i];
	}
	double avg = sum_d / m;
	for (long i = 0; i < m; i++) {
		double d = (double) (row_ptr[i + 1] - row_ptr[i]) - avg;
		sum_d2 += d * d;
	}
	double avg_bw = sbw / m, sbw2 = 0;
	for (long i = 0; i < m; i++)
		sbw2 += (bw[i] - avg_bw) * (bw[i] - avg_bw);
	s->avg_nnz_per_row = avg;
	s->std_nnz_per_row = sqrt(sum_d2 / m);
	s->min_nnz_per_row = mn;
	s->max_nnz_per_row = mx;
	s->skew = avg > 0 ? (mx - avg) / avg : 0;
	s->avg_bw = avg_bw;
	s->std_bw = sqrt(sbw2 / m);
	s->avg_bw_scaled = avg_bw / n;
	s->avg_sc = ssc / m;
	s->avg_num_neighbours = nnz > 0 ? snb / nnz : 0;
	s->cross_row_similarity = nonempty > 0 ? ssim / nonempty : 0;
	s->avg_clustering = scl / m;
	s->avg_ngroups = sng / m;
	return SPMVGEN_OK;
}

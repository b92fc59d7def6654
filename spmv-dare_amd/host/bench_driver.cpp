// Benchmark driver for generated matrices: the counterpart of the reference's
// src/bench.cpp (main, import_artificial_matrix :241-283) and
// src/bench_spmv.cpp (bench :515-587, compute :248-512, check_accuracy
// :109-236), linked with one format object (hip_matrix_format.cpp) per
// executable as src/Makefile_in:47-74 links the reference formats.
//
//   spmv_hip_<fmt>_d.exe nr_rows nr_cols avg std distribution placement bw skew
//                        neighbours cross_row_similarity seed [name]
//   spmv_hip_<fmt>_d.exe            (no arguments: print the CSV header)
//
// Protocol: warm-up (GPU_KERNEL=1: 1000 calls, else one call), then timed
// calls until both SPMV_MIN_LOOPS (default 64) and SPMV_MIN_RUNTIME seconds
// (default 2.0) are reached, each timed with CLOCK_MONOTONIC_RAW; min /
// median / max and GFLOPS = 2 nnz / median. The CSV row goes to stderr.
// Accuracy: y against a double-double compensated y_gold on the host. The
// reference's test (max relative difference over entries with y_gold > 1e-10,
// "Test failed!" above 1e-10) is printed as informative: with values in
// [-1, 1] single entries cancel and their relative error is unbounded. The
// exit status is 3 when the normwise relative error exceeds 1e-13.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>
#include <omp.h>

#include "spmv_format.h"
#include "spmvgen.h"

namespace {

double now()
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC_RAW, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

long env_long(const char *name, long dflt)
{
	const char *v = getenv(name);
	return (v && *v) ? atol(v) : dflt;
}

double env_double(const char *name, double dflt)
{
	const char *v = getenv(name);
	return (v && *v) ? atof(v) : dflt;
}

void print_header()
{
	char buf[4096];
	int i = snprintf(buf, sizeof(buf),
			"matrix_name,distribution,placement,seed,nr_rows,nr_cols,nr_nzeros,density,mem_footprint,"
			"avg_nnz_per_row,std_nnz_per_row,avg_bw,std_bw,avg_bw_scaled,avg_sc,skew,avg_num_neighbours,"
			"cross_row_similarity,format_name,time,gflops,gbps_alg,max_rel_diff,normwise_rel_err");
	i += statistics_print_labels(buf + i, (long) sizeof(buf) - i);
	fprintf(stderr, "%s\n", buf);
}

// y_gold = A x in double-double (TwoProd via fma, TwoSum), rounded to double.
void gold(const spmvgen_csr &a, const double *x, std::vector<double> &y)
{
	y.assign(a.m, 0.0);
#pragma omp parallel for schedule(dynamic, 1024)
	for (long i = 0; i < a.m; i++) {
		double hi = 0, lo = 0;
		for (long j = a.row_ptr[i]; j < a.row_ptr[i + 1]; j++) {
			double p = a.values[j] * x[a.col_ind[j]];
			double pe = std::fma(a.values[j], x[a.col_ind[j]], -p);
			double s = hi + p, bb = s - hi;
			double e = (hi - (s - bb)) + (p - bb) + (lo + pe);
			hi = s + e;
			lo = e - (hi - s);
		}
		y[i] = hi + lo;
	}
}

} // namespace

int main(int argc, char **argv)
{
	if (argc < 12) {
		print_header();
		return 0;
	}
	int k = 1;
	long nr_rows = atol(argv[k++]), nr_cols = atol(argv[k++]);
	double avg = atof(argv[k++]), sd = atof(argv[k++]);
	const char *distr = argv[k++], *place = argv[k++];
	double bw = atof(argv[k++]), skew = atof(argv[k++]), neigh = atof(argv[k++]), crs = atof(argv[k++]);
	unsigned seed = (unsigned) atol(argv[k++]);
	char name[512];
	if (k < argc)
		snprintf(name, sizeof(name), "%s_artificial", argv[k]);
	else
		snprintf(name, sizeof(name), "%ld_%ld_%g_%g_%s_%s", nr_rows, nr_cols, avg, sd, distr, place);

	spmvgen_csr a;
	double t0 = now();
	int rc = spmvgen_generate(nr_rows, nr_cols, avg, sd, distr, seed, place, bw, skew, neigh, crs, 1, &a);
	if (rc) {
		fprintf(stderr, "generator: %s\n", spmvgen_last_error());
		return EXIT_FAILURE;
	}
	spmvgen_map_values(a.values, a.nnz);
	printf("time generate matrix = %g\n", now() - t0);

	t0 = now();
	Matrix_Format *MF = csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, 0, 0);
	printf("time convert to format = %g\n", now() - t0);

	std::vector<double> x(a.n > 0 ? a.n : 1);
	spmvgen_fill_x(x.data(), a.n, seed + 1);
	double *y = (double *) aligned_alloc(64, ((a.m + 64) * sizeof(double) + 63) / 64 * 64);
	for (long i = 0; i < a.m; i++)
		y[i] = 1.0; // formats must write empty rows (src/bench_spmv.cpp:561-564)

	t0 = now();
	if (env_long("GPU_KERNEL", 0)) {
		for (int i = 0; i < 1000; i++)
			MF->spmv(x.data(), y);
	} else {
		MF->spmv(x.data(), y);
	}
	printf("time warm up %lf\n", now() - t0);

	MF->statistics_start();
	long min_loops = env_long("SPMV_MIN_LOOPS", 64);
	double min_runtime = env_double("SPMV_MIN_RUNTIME", 2.0);
	std::vector<double> times;
	double total = 0;
	while (total < min_runtime || (long) times.size() < min_loops) {
		double s = now();
		MF->spmv(x.data(), y);
		double t = now() - s;
		times.push_back(t);
		total += t;
	}
	std::sort(times.begin(), times.end());
	double tmed = times[times.size() / 2];
	double gflops = a.nnz / tmed * 2 * 1e-9;
	double bytes = 12.0 * a.nnz + 4.0 * (a.m + 1) + 8.0 * a.n + 8.0 * a.m;
	printf("number of loops = %zu\n", times.size());
	printf("time iter: min=%g, median=%g, max=%g\n", times.front(), tmed, times.back());
	printf("GFLOPS = %lf (%s)\n", gflops, MF->format_name);

	char stats[1024];
	int sn = MF->statistics_print_data(stats, sizeof(stats)); // also brings y back from the device
	stats[sn > 0 ? sn : 0] = '\0';

	std::vector<double> yg;
	gold(a, x.data(), yg);
	double max_rel = 0, dn = 0, rn = 0;
	for (long i = 0; i < a.m; i++) {
		double d = y[i] - yg[i];
		dn += d * d;
		rn += yg[i] * yg[i];
		if (yg[i] > 1e-10) {
			double r = fabs(d) / yg[i];
			max_rel = r > max_rel ? r : max_rel;
		}
	}
	double normwise = rn > 0 ? sqrt(dn / rn) : sqrt(dn);
	printf("accuracy: max rel diff = %g, normwise rel err = %g\n", max_rel, normwise);
	if (max_rel > 1e-10)
		printf("Test failed! (max relative difference %g > 1e-10)\n", max_rel);
	if (normwise > 1e-13)
		printf("Test failed! (normwise relative error %g > 1e-13)\n", normwise);

	const spmvgen_stats &st = a.stats;
	fprintf(stderr, "%s,%s,%s,%u,%ld,%ld,%ld,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%s,%lf,%lf,%lf,%g,%g%s\n",
			name, distr, place, seed, a.m, a.n, a.nnz, 100.0 * a.nnz / ((double) a.m * a.n),
			(12.0 * a.nnz + 4.0 * (a.m + 1)) / (1024.0 * 1024.0), st.avg_nnz_per_row, st.std_nnz_per_row,
			st.avg_bw, st.std_bw, st.avg_bw_scaled, st.avg_sc, st.skew, st.avg_num_neighbours,
			st.cross_row_similarity, MF->format_name, total, gflops, bytes / tmed * 1e-9, max_rel, normwise,
			stats);
	// Matrix_Format has no virtual destructor (as in the reference); the
	// object lives until exit, like the reference driver's MF.
	free(y);
	spmvgen_free(&a);
	return normwise > 1e-13 ? 3 : 0;
}

// Benchmark driver: the counterpart of the reference's src/bench.cpp (main,
// import_artificial_matrix :241-283, import_file :126-238) and
// src/bench_spmv.cpp (bench :515-587, compute :248-512, check_accuracy
// :109-236), linked with one format object (hip_matrix_format.cpp) per
// executable as src/Makefile_in:47-74 links the reference formats.
//
//   spmv_hip_<fmt>_d.exe nr_rows nr_cols avg std distribution placement bw skew
//                        neighbours cross_row_similarity seed [name]
//   spmv_hip_<fmt>_d.exe file.mtx   (a Matrix Market file, the reference's
//                        validation mode: check_accuracy's metrics)
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

// The reference's validation of y on a real matrix (bench_spmv.cpp:109-236):
// y_gold by a compensated (Kahan) sum in _Float128 of the double products;
// "Test failed!" when the largest relative difference over the rows with
// y_gold > 1e-10 exceeds 1e-10; and its eight error metrics
// (lib/array_metrics.c: mae, max_ae, mse, mape, smape, lnQ_error, mlare,
// gmare) of y against y_gold, both as doubles, printed in its line format.
// Returns the largest relative difference; *csv gets the metrics as
// ','-prefixed columns.
double check_accuracy(const spmvgen_csr &a, const double *x, const double *y, char *csv, long csv_n)
{
	const long m = a.m;
	std::vector<_Float128> gold(m);
#pragma omp parallel for schedule(dynamic, 1024)
	for (long i = 0; i < m; i++) {
		_Float128 sum = 0, comp = 0;
		for (long j = a.row_ptr[i]; j < a.row_ptr[i + 1]; j++) {
			const _Float128 val = (_Float128) (a.values[j] * x[a.col_ind[j]]) - comp;
			const _Float128 tmp = sum + val;
			comp = (tmp - sum) - val;
			sum = tmp;
		}
		gold[i] = sum;
	}
	const double eps = 1e-10, dbl_eps = 2.220446049250313e-16;
	double max_diff = 0, mae = 0, max_ae = 0, mse = 0, mare = 0, smare = 0, lnq = 0;
	for (long i = 0; i < m; i++) {
		const _Float128 d = gold[i] > (_Float128) y[i] ? gold[i] - (_Float128) y[i] : (_Float128) y[i] - gold[i];
		if (gold[i] > (_Float128) eps) {
			const double r = (double) (d / (gold[i] < 0 ? -gold[i] : gold[i]));
			max_diff = r > max_diff ? r : max_diff;
		}
		const double av = (double) gold[i], fv = y[i], ae = fabs(av - fv);
		mae += ae;
		max_ae = ae > max_ae ? ae : max_ae;
		mse += (av - fv) * (av - fv);
		mare += ae / std::max(fabs(av), dbl_eps);
		smare += ae / std::max(fabs(av) + fabs(fv), dbl_eps);
		lnq += log10(std::max(fabs(fv), dbl_eps)) - log10(std::max(fabs(av), dbl_eps));
	}
	if (max_diff > eps)
		printf("Test failed! (%g)\n", max_diff);
	const double N = m > 0 ? (double) m : 1.0;
	mae /= N;
	mse /= N;
	const double mape = 100.0 * mare / N, smape = 100.0 * smare / N, lnq_error = lnq / N;
	const double mlare = (double) log10l(fabsl(powl(10.0L, (long double) lnq_error) - 1.0L));
	const double gmare = pow(10.0, mlare);
	printf("errors spmv: mae=%g, max_ae=%g, mse=%g, mape=%g, smape=%g, lnQ_error=%g, mlare=%g, gmare=%g\n", mae,
			max_ae, mse, mape, smape, lnq_error, mlare, gmare);
	snprintf(csv, (size_t) csv_n, ",%g,%g,%g,%g,%g,%g,%g,%g", mae, max_ae, mse, mape, smape, lnq_error, mlare, gmare);
	return max_diff;
}

// Timed calls of the reference protocol (compute, bench_spmv.cpp:248-359):
// warm-up (GPU_KERNEL=1: 1000 calls, else one), then calls until both
// SPMV_MIN_LOOPS and SPMV_MIN_RUNTIME are reached; sorted times.
std::vector<double> time_calls(Matrix_Format *MF, const double *x, double *y, double *total)
{
	double t0 = now();
	if (env_long("GPU_KERNEL", 0)) {
		for (int i = 0; i < 1000; i++)
			MF->spmv((ValueType *) x, y);
	} else {
		MF->spmv((ValueType *) x, y);
	}
	printf("time warm up %lf\n", now() - t0);
	MF->statistics_start();
	long min_loops = env_long("SPMV_MIN_LOOPS", 64);
	double min_runtime = env_double("SPMV_MIN_RUNTIME", 2.0);
	std::vector<double> times;
	*total = 0;
	while (*total < min_runtime || (long) times.size() < min_loops) {
		double s = now();
		MF->spmv((ValueType *) x, y);
		double t = now() - s;
		times.push_back(t);
		*total += t;
	}
	std::sort(times.begin(), times.end());
	return times;
}

// A Matrix Market file (the reference's validation mode): read as import_file
// does (symmetry expanded), x from the benchmark's x definition with seed 1,
// the reference protocol, then check_accuracy. Exit status 3 when the
// reference's test fails.
int run_mtx(const char *path)
{
	spmvgen_csr a;
	int symmetric = 0;
	double t0 = now();
	if (spmvgen_read_mtx(path, 1, &a, &symmetric)) {
		fprintf(stderr, "%s\n", spmvgen_last_error());
		return EXIT_FAILURE;
	}
	printf("time read: %lf\n", now() - t0);
	t0 = now();
	Matrix_Format *MF = csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, symmetric, 1);
	printf("time convert to format = %g\n", now() - t0);
	std::vector<double> x(a.n > 0 ? a.n : 1);
	spmvgen_fill_x(x.data(), a.n, 1);
	double *y = (double *) aligned_alloc(64, ((a.m + 64) * sizeof(double) + 63) / 64 * 64);
	for (long i = 0; i < a.m; i++)
		y[i] = 1.0;
	double total = 0;
	std::vector<double> times = time_calls(MF, x.data(), y, &total);
	const double tmed = times[times.size() / 2];
	const double gflops = a.nnz / tmed * 2 * 1e-9;
	printf("number of loops = %zu\n", times.size());
	printf("time iter: min=%g, median=%g, max=%g\n", times.front(), tmed, times.back());
	printf("GFLOPS = %lf (%s)\n", gflops, MF->format_name);
	char stats[1024];
	int sn = MF->statistics_print_data(stats, sizeof(stats)); // also brings y back from the device
	stats[sn > 0 ? sn : 0] = '\0';
	char acc[512];
	const double max_diff = check_accuracy(a, x.data(), y, acc, sizeof(acc));
	fprintf(stderr, "%s,%d,%ld,%ld,%ld,%ld,%d,%lf,%lf,%lf,%lf,%lf,%lf,%s,%ld,%ld,%ld,%lf,%lf,%zu%s%s\n", path,
			omp_get_max_threads(), a.m, a.n, a.nnz, a.nnz, symmetric, total, times.front(), tmed, times.back(), gflops,
			MF->csr_mem_footprint / (1024 * 1024), MF->format_name, MF->m, MF->n, MF->nnz,
			MF->mem_footprint / (1024 * 1024), MF->mem_footprint / MF->csr_mem_footprint, times.size(), acc, stats);
	free(y);
	spmvgen_free(&a);
	return max_diff > 1e-10 ? 3 : 0;
}

} // namespace

int main(int argc, char **argv)
{
	const char *art = getenv("USE_ARTIFICIAL_MATRICES");
	if (argc == 2 || (argc >= 2 && art && atol(art) == 0))
		return run_mtx(argv[1]);
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

	double total = 0;
	std::vector<double> times = time_calls(MF, x.data(), y, &total);
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

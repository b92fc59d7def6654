// Host-side memory and undefined-behaviour check (test infrastructure, never
// shipped): the host codecs and the generator of the product
// (spmv-dare_amd/csrc/div/div_codec.cpp, csrc/gen/generator.cpp, csrc/gen/mtx.cpp)
// and the CPU oracle (oracle/oracle.c), built with -fsanitize=address,undefined
// by tests/asan/Makefile, exercised on generated matrices with edge shapes:
// DIV-RF encode at several (T, L, P) -> oracle decode == the CSR triplets,
// every device layout, packet-group order, lossy rounding; the oracle's SELL,
// SELL-C-sigma and SELL_SORTED conversions and products; the Matrix Market
// fixtures. Any sanitizer report aborts the run; a wrong result exits 1.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <string>
#include <tuple>
#include <vector>

#include "div_codec.hpp"
#include "oracle.h"
#include "spmvgen.h"

static int failures = 0;

#define CHECK(cond, ...)                                  \
	do {                                                  \
		if (!(cond)) {                                    \
			fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
			fprintf(stderr, __VA_ARGS__);                 \
			fprintf(stderr, "\n");                        \
			failures++;                                   \
		}                                                 \
	} while (0)

static void check_div(const spmvgen_csr &a, int T, int L, long P)
{
	spmvdiv::Stream st;
	const char *err = nullptr;
	if (spmvdiv::encode(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, T, L, P, &st, &err)) {
		printf("  div T=%d L=%d P=%ld: rejected (%s)\n", T, L, P, err);
		return;
	}
	std::vector<int32_t> r(a.nnz + 1), c(a.nnz + 1);
	std::vector<double> v(a.nnz + 1);
	long k = oracle_div_decode(st.bytes.data(), (long) st.bytes.size(), L, r.data(), c.data(), v.data(), a.nnz);
	CHECK(k == a.nnz, "div decode count %ld != nnz %ld (T=%d L=%d P=%ld)", k, a.nnz, T, L, P);
	if (k != a.nnz)
		return;
	std::vector<std::tuple<int32_t, int32_t, double>> got(a.nnz), want(a.nnz);
	for (long j = 0; j < a.nnz; j++)
		got[j] = std::make_tuple(r[j], c[j], v[j]);
	for (long i = 0, j = 0; i < a.m; i++)
		for (; j < a.row_ptr[i + 1]; j++)
			want[j] = std::make_tuple((int32_t) i, a.col_ind[j], a.values[j]);
	std::sort(got.begin(), got.end());
	std::sort(want.begin(), want.end());
	CHECK(got == want, "div round trip differs (T=%d L=%d P=%ld)", T, L, P);
	for (int lay = 0; lay < 4; lay++) {
		spmvdiv::DeviceLayout dl;
		if (spmvdiv::to_device_layout(st, lay & 1, lay & 2, &dl, &err) == 0)
			CHECK((long) dl.pk_off.size() == st.npackets() + 1, "device layout %d packet count", lay);
	}
	std::vector<int32_t> order = spmvdiv::group_order_by_median(st, a.col_ind);
	std::vector<int32_t> sorted(order);
	std::sort(sorted.begin(), sorted.end());
	for (long g = 0; g < (long) sorted.size(); g++)
		CHECK(sorted[g] == g, "group order is not a permutation");
}

static void check_oracle(const spmvgen_csr &a)
{
	const long m = a.m, n = a.n;
	std::vector<double> x(n > 0 ? n : 1), y(m + 1), y2(m + 1), lo(m + 1);
	for (long j = 0; j < n; j++)
		x[j] = (double) ((j % 13) + 1) / 8.0 - 1.0;
	oracle_csr_naive(a.row_ptr, a.col_ind, a.values, x.data(), y.data(), m);
	oracle_csr_dd(a.row_ptr, a.col_ind, a.values, x.data(), y2.data(), lo.data(), m);
	double num = 0, den = 0;
	for (long i = 0; i < m; i++) {
		num += (y[i] - y2[i]) * (y[i] - y2[i]);
		den += y2[i] * y2[i];
	}
	CHECK(num <= 1e-24 * den + 1e-300, "csr_naive vs dd");
	for (long sigma : {64L, 256L, 4096L}) {
		std::vector<int64_t> order(m + 1), widths(m / 64 + 2), sptr(m / 64 + 3);
		oracle_sellcs_order(a.row_ptr, m, sigma, order.data());
		long total = oracle_sellcs_layout(a.row_ptr, m, 64, order.data(), 0, widths.data(), sptr.data());
		std::vector<int64_t> cols(total + 1);
		std::vector<double> vals(total + 1);
		long trunc = oracle_sellcs_fill(a.row_ptr, a.col_ind, a.values, m, 64, order.data(), widths.data(),
				sptr.data(), cols.data(), vals.data());
		CHECK(trunc == 0, "sellcs truncated %ld", trunc);
		oracle_sellcs_spmv(m, 64, order.data(), widths.data(), sptr.data(), cols.data(), vals.data(), x.data(),
				y2.data());
	}
	for (long VL : {4L, 64L}) {
		std::vector<int64_t> cptr((m + VL - 1) / VL + 1);
		long total = oracle_sell_layout(a.row_ptr, m, VL, cptr.data());
		std::vector<int64_t> ja(total + 1);
		std::vector<double> av(total + 1);
		oracle_sell_fill(a.row_ptr, a.col_ind, a.values, m, VL, 3, cptr.data(), ja.data(), av.data());
		oracle_sell_spmv(m, VL, cptr.data(), ja.data(), av.data(), x.data(), y2.data());
		std::vector<int64_t> win(5);
		oracle_sellsorted_windows(a.row_ptr, m, VL, 4, win.data());
		std::vector<int32_t> perm(m + 1);
		oracle_sellsorted_order(a.row_ptr, m, 4, win.data(), perm.data());
		total = oracle_sellsorted_layout(a.row_ptr, m, VL, perm.data(), cptr.data());
		ja.assign(total + 1, 0);
		av.assign(total + 1, 0);
		oracle_sellsorted_fill(a.row_ptr, a.col_ind, a.values, m, VL, 4, win.data(), perm.data(), cptr.data(),
				ja.data(), av.data());
		oracle_sellsorted_spmv(m, VL, cptr.data(), ja.data(), av.data(), perm.data(), x.data(), y2.data());
	}
}

int main(int argc, char **argv)
{
	const char *mtx_dir = argc > 1 ? argv[1] : "tests/golden/mtx";
	struct Line {
		long r, c;
		double avg, std;
		const char *dist, *place;
		double bw, skew, neigh, crs;
		unsigned seed;
	};
	const Line lines[] = {
		{1000, 1000, 5, 4, "normal", "random", 0.3, 10, 1, 0.2, 14},
		{100, 100, 1, 0.5, "gamma", "diagonal", 0.05, 50, 0, 0, 3},
		{3000, 50000, 40, 80, "gamma", "random", 0.9, 300, 4, 0.6, 5},
		{2000, 2000, 200, 50, "normal", "simple", 0.5, 2, 2, 0.8, 9},
		{20000, 4000000, 16, 4, "normal", "random", 0.5, 1.5, 1, 0.2, 3},
		{5000, 5000, 2, 6, "gamma", "diagonal", 0.01, 1000, 8, 0.9, 21},
	};
	for (const Line &l : lines) {
		spmvgen_csr a;
		int rc = spmvgen_generate(l.r, l.c, l.avg, l.std, l.dist, l.seed, l.place, l.bw, l.skew, l.neigh, l.crs, 1,
				&a);
		printf("matrix %ld x %ld (%s, %s): rc=%d nnz=%ld\n", l.r, l.c, l.dist, l.place, rc, rc ? 0L : a.nnz);
		if (rc)
			continue;
		check_div(a, 1, 64, 16384);
		check_div(a, 4, 4, 512);
		check_div(a, 3, 32, 97);
		check_div(a, 8, 64, 1);
		std::vector<double> lossy(a.nnz + 1);
		const char *err = nullptr;
		CHECK(spmvdiv::reduce_precision(a.values, a.nnz, 1e-6, lossy.data(), &err) == 0, "reduce_precision");
		for (long j = 0; j < a.nnz; j++)
			CHECK(std::fabs(lossy[j] - a.values[j]) <= 1e-6 * std::fabs(a.values[j]), "lossy bound");
		check_oracle(a);
		spmvgen_free(&a);
	}
	DIR *d = opendir(mtx_dir);
	CHECK(d != nullptr, "no Matrix Market directory %s", mtx_dir);
	int files = 0;
	while (d) {
		struct dirent *e = readdir(d);
		if (!e)
			break;
		std::string name = e->d_name;
		if (name.size() < 5 || name.substr(name.size() - 4) != ".mtx")
			continue;
		spmvgen_csr a;
		int sym = 0;
		int rc = spmvgen_read_mtx((std::string(mtx_dir) + "/" + name).c_str(), 1, &a, &sym);
		printf("mtx %s: rc=%d\n", name.c_str(), rc);
		files++;
		if (rc == 0) {
			check_div(a, 2, 4, 7);
			check_oracle(a);
			spmvgen_free(&a);
		}
	}
	if (d)
		closedir(d);
	CHECK(files > 0, "no Matrix Market fixtures read");
	printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
	return failures ? 1 : 0;
}

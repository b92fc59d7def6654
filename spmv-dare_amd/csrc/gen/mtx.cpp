// Matrix Market import: the reference's real-matrix input path, import_file
// (src/bench.cpp:126-238) = mtx_read (lib/storage_formats/matrix_market/
// matrix_market.c:152-331 with matrix_market_gen.c: coordinate parsing and
// symmetry expansion), mtx_values_convert_to_real (matrix_market.c:421-456)
// and coo_to_csr (lib/storage_formats/csr/csr_gen.c:182-213, rows bucketed,
// each row's columns sorted).
//
// Rules kept from the reference:
//   * a file without a %%MatrixMarket banner is read as "coordinate real
//     general"; the banner's object must be "matrix"; the symmetry word is
//     one of general / symmetric / skew-symmetric / Hermitian (exact case);
//   * empty lines are dropped; '%' lines after the banner are comments; the
//     size line "M N NNZ" must be followed by exactly NNZ entry lines;
//   * entries are 1-based; values to real: integer as is, complex by its
//     modulus, pattern 1.0;
//   * symmetric matrices (expand_symmetry): every stored off-diagonal entry
//     (r, c, v) gets a mirror (c, r, v), negated for skew-symmetric (complex:
//     conjugate, negated for skew, so its modulus is |v|), appended after the
//     stored entries in their order;
//   * rows in ascending order, each row's columns ascending; duplicate
//     entries are kept (not summed).
// Divergences: "array" (dense) files, whose entries the reference's COO path
// never fills (it leaves R and C unset and coo_to_csr reads them), are
// rejected; duplicate (row, column) entries keep their file order, where the
// reference's unstable parallel bucket sort leaves it to its threads.

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "spmvgen.h"

namespace spmvgen_detail {
int set_error(int code, const char *msg); // generator.cpp
}
using spmvgen_detail::set_error;

namespace {

int fail(const char *path, const char *what)
{
	char buf[512];
	snprintf(buf, sizeof(buf), "Matrix Market file '%s': %s", path, what);
	return set_error(SPMVGEN_EFILE, buf);
}

// Next whitespace-separated token of [p, end) as a number; false if none.
bool next_double(const char *&p, const char *end, double *v)
{
	while (p < end && (*p == ' ' || *p == '\t' || *p == '\r'))
		p++;
	if (p >= end)
		return false;
	char *q = nullptr;
	errno = 0;
	*v = strtod(p, &q);
	if (q == p || q > end)
		return false;
	p = q;
	return true;
}

bool next_long(const char *&p, const char *end, long *v)
{
	while (p < end && (*p == ' ' || *p == '\t' || *p == '\r'))
		p++;
	if (p >= end)
		return false;
	char *q = nullptr;
	*v = strtol(p, &q, 10);
	if (q == p || q > end)
		return false;
	p = q;
	return true;
}

} // namespace

extern "C" int spmvgen_read_mtx(const char *path, int expand_symmetry, spmvgen_csr *out, int *symmetric_out)
{
	if (!path || !out)
		return set_error(SPMVGEN_EINVAL, "read_mtx: NULL argument");
	memset(out, 0, sizeof(*out));
	FILE *fp = fopen(path, "rb");
	if (!fp)
		return fail(path, strerror(errno));
	std::string text;
	{
		char buf[1 << 16];
		size_t k;
		while ((k = fread(buf, 1, sizeof(buf), fp)) > 0)
			text.append(buf, k);
		fclose(fp);
	}
	// non-empty lines
	std::vector<std::pair<const char *, const char *>> lines;
	{
		const char *p = text.data(), *end = p + text.size();
		while (p < end) {
			const char *e = (const char *) memchr(p, '\n', (size_t) (end - p));
			if (!e)
				e = end;
			if (e > p && !(e - p == 1 && *p == '\r'))
				lines.emplace_back(p, e);
			p = e + 1;
		}
	}
	if (lines.empty())
		return fail(path, "empty file");
	size_t li = 0;
	std::string field = "real";
	bool symmetric = false, skew = false;
	if (strncmp(lines[0].first, "%%MatrixMarket", 14) == 0 &&
			(lines[0].second - lines[0].first == 14 || isspace((unsigned char) lines[0].first[14]))) {
		char object[64] = "", format[64] = "", fld[64] = "", sym[64] = "";
		std::string banner(lines[0].first, lines[0].second);
		if (sscanf(banner.c_str(), "%*s %63s %63s %63s %63s", object, format, fld, sym) != 4)
			return fail(path, "incomplete %%MatrixMarket banner");
		if (strcmp(object, "matrix") || (strcmp(format, "coordinate") && strcmp(format, "array")))
			return fail(path, "only matrix coordinate or array format");
		if (!strcmp(format, "array"))
			return fail(path, "array (dense) format is not read (the reference's import leaves its indices unset)");
		field = fld;
		if (field != "real" && field != "integer" && field != "complex" && field != "pattern")
			return fail(path, "unrecognized field type");
		if (!strcmp(sym, "symmetric") || !strcmp(sym, "Hermitian"))
			symmetric = true;
		else if (!strcmp(sym, "skew-symmetric"))
			symmetric = skew = true;
		else if (strcmp(sym, "general"))
			return fail(path, "unsupported symmetry type");
		li = 1;
	}
	while (li < lines.size() && lines[li].first[0] == '%')
		li++;
	if (li >= lines.size())
		return fail(path, "missing matrix sizes");
	long M = 0, N = 0, nnz_stored = 0;
	{
		const char *p = lines[li].first, *e = lines[li].second;
		if (!next_long(p, e, &M) || !next_long(p, e, &N) || !next_long(p, e, &nnz_stored))
			return fail(path, "invalid/missing matrix sizes");
		li++;
	}
	if (M < 0 || N < 0 || nnz_stored < 0)
		return fail(path, "negative matrix sizes");
	if ((long) (lines.size() - li) != nnz_stored)
		return fail(path, "remaining number of file lines does not match the number of non-zeros");
	std::vector<int32_t> R, C;
	std::vector<double> V;
	R.reserve(symmetric ? 2 * nnz_stored : nnz_stored);
	C.reserve(R.capacity());
	V.reserve(R.capacity());
	const bool cplx = field == "complex", pattern = field == "pattern";
	for (long k = 0; k < nnz_stored; k++) {
		const char *p = lines[li + k].first, *e = lines[li + k].second;
		long r, c;
		if (!next_long(p, e, &r) || !next_long(p, e, &c))
			return fail(path, "badly formed entry line");
		if (r < 1 || r > M || c < 1 || c > N)
			return fail(path, "entry index outside the matrix");
		double v = 1.0;
		if (!pattern) {
			double re = 0, im = 0;
			if (!next_double(p, e, &re) || (cplx && !next_double(p, e, &im)))
				return fail(path, "badly formed or missing value");
			v = cplx ? std::abs(std::complex<double>(re, im)) : re;
			if (field == "integer")
				v = (double) (int) re;
		}
		R.push_back((int32_t) (r - 1));
		C.push_back((int32_t) (c - 1));
		V.push_back(v);
	}
	if (symmetric && expand_symmetry) {
		const long stored = (long) R.size();
		for (long k = 0; k < stored; k++) {
			if (R[k] == C[k])
				continue;
			R.push_back(C[k]);
			C.push_back(R[k]);
			V.push_back(skew && !cplx ? -V[k] : V[k]);
		}
	}
	const long nnz = (long) R.size();
	if (nnz > INT32_MAX || M > INT32_MAX - 1 || N > INT32_MAX)
		return set_error(SPMVGEN_EOVERFLOW, "read_mtx: matrix does not fit 32-bit CSR indices");
	// CSR: entries ordered by (row, column), file order among duplicates
	std::vector<long> ord(nnz);
	std::iota(ord.begin(), ord.end(), 0L);
	std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) {
		return R[a] != R[b] ? R[a] < R[b] : C[a] < C[b];
	});
	out->m = M;
	out->n = N;
	out->nnz = nnz;
	out->row_ptr = (int32_t *) malloc((M + 1) * sizeof(int32_t));
	out->col_ind = (int32_t *) malloc((nnz ? nnz : 1) * sizeof(int32_t));
	out->values = (double *) malloc((nnz ? nnz : 1) * sizeof(double));
	if (!out->row_ptr || !out->col_ind || !out->values) {
		spmvgen_free(out);
		return set_error(SPMVGEN_ENOMEM, "read_mtx: allocation");
	}
	std::fill(out->row_ptr, out->row_ptr + M + 1, 0);
	for (long k = 0; k < nnz; k++)
		out->row_ptr[R[k] + 1]++;
	for (long i = 0; i < M; i++)
		out->row_ptr[i + 1] += out->row_ptr[i];
	for (long k = 0; k < nnz; k++) {
		out->col_ind[k] = C[ord[k]];
		out->values[k] = V[ord[k]];
	}
	if (symmetric_out)
		*symmetric_out = symmetric ? 1 : 0;
	return SPMVGEN_OK;
}

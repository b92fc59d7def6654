/*
 * Test-infrastructure harness (checker only, never shipped): reads a Matrix
 * Market file with the reference's own import path, compiled from its
 * sources in place by oracle/Makefile, and prints what the fixtures record.
 *
 * Calls, as import_file does (src/bench.cpp:126-238):
 *   mtx_read(path, expand_symmetry = 1, pattern_dummy_vals = 1)
 *       (lib/storage_formats/matrix_market/matrix_market.c:258-331)
 *   mtx_values_convert_to_real (matrix_market.c:421-456)
 *   coo_to_csr(..., sort_columns = 1, transpose = 0)
 *       (lib/storage_formats/csr/csr_gen.c:182-213, the _csr_converter_reference instance)
 *
 * Usage: mtx_dump file.mtx
 * Prints one line: m n nnz symmetric fnv(row_ptr) fnv(col_ind) fnv(values)
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>

#include "storage_formats/matrix_market/matrix_market.h"
#include "aux/csr_converter_reference.h"

static uint64_t fnv1a(const void *p, size_t nbytes)
{
	const unsigned char *b = (const unsigned char *) p;
	uint64_t h = 1469598103934665603ULL;
	for (size_t i = 0; i < nbytes; i++) {
		h ^= b[i];
		h *= 1099511628211ULL;
	}
	return h;
}

int main(int argc, char **argv)
{
	if (argc < 2) {
		fprintf(stderr, "usage: %s file.mtx\n", argv[0]);
		return 2;
	}
	struct Matrix_Market *MTX = mtx_read(argv[1], 1, 1);
	mtx_values_convert_to_real(MTX);
	long m = MTX->m, n = MTX->n, nnz = MTX->nnz_stored;
	int32_t *ia = calloc(m + 1, sizeof(int32_t));
	int32_t *ja = calloc(nnz ? nnz : 1, sizeof(int32_t));
	double *a = calloc(nnz ? nnz : 1, sizeof(double));
	coo_to_csr(MTX->R, MTX->C, (double *) MTX->V, m, n, nnz, ia, ja, a, 1, 0);
	printf("%ld %ld %ld %d %016llx %016llx %016llx\n", m, n, nnz, MTX->symmetric,
			(unsigned long long) fnv1a(ia, (m + 1) * sizeof(int32_t)),
			(unsigned long long) fnv1a(ja, nnz * sizeof(int32_t)),
			(unsigned long long) fnv1a(a, nnz * sizeof(double)));
	return 0;
}

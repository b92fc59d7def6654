/*
 * Test-infrastructure harness (checker only, never shipped): drives the
 * reference's own matrix generator, compiled from its sources in place by
 * oracle/Makefile, and prints what the golden fixtures record.
 *
 * Calls: artificial_matrix_generation()
 *        (artificial-matrix-generator/artificial_matrix_generation.c:163-579)
 *
 * Usage: gen_dump m n avg std distr placement bw skew neigh crs seed [dump.bin]
 *   (argument order of artificial-matrix-generator/artificial_matrix.c:40-50)
 * Prints one line:
 *   m n nnz fnv(row_ptr) fnv(col_ind) fnv(values) avg std min max skew avg_bw
 *   avg_neigh crs
 * With dump.bin: writes int64 m, n, nnz, then row_ptr (int32, m+1),
 * col_ind (int32, nnz), values (float64, nnz).
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include "artificial_matrix_generation.h"

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
	if (argc < 12) {
		fprintf(stderr, "usage: %s m n avg std distr placement bw skew neigh crs seed [dump.bin]\n", argv[0]);
		return 2;
	}
	long m = atol(argv[1]), n = atol(argv[2]);
	double avg = atof(argv[3]), std = atof(argv[4]);
	char *distr = argv[5], *placement = argv[6];
	double bw = atof(argv[7]), skew = atof(argv[8]), neigh = atof(argv[9]), crs = atof(argv[10]);
	unsigned seed = (unsigned) atol(argv[11]);
	struct csr_matrix *c = artificial_matrix_generation(m, n, avg, std, distr, seed, placement, bw, skew, neigh, crs);
	long nnz = c->nr_nzeros;
	printf("%ld %ld %ld %016llx %016llx %016llx %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
	       m, n, nnz,
	       (unsigned long long) fnv1a(c->row_ptr, (size_t)(m + 1) * sizeof(int)),
	       (unsigned long long) fnv1a(c->col_ind, (size_t) nnz * sizeof(int)),
	       (unsigned long long) fnv1a(c->values, (size_t) nnz * sizeof(double)),
	       c->avg_nnz_per_row, c->std_nnz_per_row, c->min_nnz_per_row, c->max_nnz_per_row,
	       c->skew, c->avg_bw, c->avg_num_neighbours, c->cross_row_similarity);
	if (argc > 12) {
		FILE *f = fopen(argv[12], "wb");
		if (!f) { perror("fopen"); return 1; }
		int64_t hdr[3] = {m, n, nnz};
		fwrite(hdr, sizeof(hdr), 1, f);
		fwrite(c->row_ptr, sizeof(int), (size_t)(m + 1), f);
		fwrite(c->col_ind, sizeof(int), (size_t) nnz, f);
		fwrite(c->values, sizeof(double), (size_t) nnz, f);
		fclose(f);
	}
	free_csr_matrix(c);
	return 0;
}

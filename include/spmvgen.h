/*
 * spmvgen.h -- C ABI of the host-side artificial matrix generator (the input
 * producer of the SpMV path) and of the benchmark's value / x initialisation.
 *
 * Replaces:
 *   struct csr_matrix * artificial_matrix_generation(long nr_rows, long nr_cols,
 *       double avg_nnz_per_row, double std_nnz_per_row, char * distribution,
 *       unsigned int seed, char * placement, double bw_scaled, double skew,
 *       double avg_num_neighbours, double cross_row_similarity);
 *   (artificial-matrix-generator/artificial_matrix_generation.h:90, body
 *    artificial_matrix_generation.c:163-579)
 * The CSR structure (row_ptr, col_ind) and the raw values it produces are
 * bit-identical to the reference for the same 11 parameters, independent of
 * the number of OpenMP threads.
 *
 * Errors: every entry point returns 0 on success and a negative SPMVGEN_E*
 * code on failure (the reference calls exit(1)); spmvgen_last_error() holds
 * the message of the calling thread's last failure.
 */
#ifndef SPMVGEN_H
#define SPMVGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	SPMVGEN_OK = 0,
	SPMVGEN_EINVAL = -1,     /* bad parameter (negative sizes, unknown distribution...) */
	SPMVGEN_ENOMEM = -2,     /* allocation failed */
	SPMVGEN_EOVERFLOW = -3,  /* nnz does not fit the int32 row pointer */
	SPMVGEN_EDUP = -4,       /* cross-row copy hit an occupied column (reference: printf("error"); exit(1)) */
	SPMVGEN_EFILE = -5,      /* Matrix Market file missing or malformed (reference: error() and exit) */
};

/* Descriptor statistics of a generated matrix
 * (artificial_matrix_generation.c:529-570, matrix_util.h:235-408). */
typedef struct spmvgen_stats {
	double avg_nnz_per_row, std_nnz_per_row, min_nnz_per_row, max_nnz_per_row;
	double skew;                 /* (max - avg) / avg */
	double avg_bw, std_bw;       /* per-row last col - first col + 1 */
	double avg_bw_scaled;        /* avg_bw / n */
	double avg_sc;               /* mean of degree / bw */
	double avg_num_neighbours;   /* window 1, matrix_util.h:235-266 */
	double cross_row_similarity; /* window 1, matrix_util.h:332-408 */
	double avg_clustering;       /* mean of ngroups / degree */
	double avg_ngroups;          /* mean number of contiguous column runs per row */
	double bw_corrected;         /* the corrected placement range used */
} spmvgen_stats;

/* A generated CSR matrix. Arrays are allocated by the library and released
 * by spmvgen_free(). values[] holds the reference's raw values
 * (k / 2^31 in [0,1)). */
typedef struct spmvgen_csr {
	long m, n, nnz;
	int32_t *row_ptr;   /* m + 1 */
	int32_t *col_ind;   /* nnz, strictly increasing within each row */
	double *values;     /* nnz */
	spmvgen_stats stats;
} spmvgen_csr;

/* Generate one matrix. distribution: "normal" or "gamma" (only the first
 * character is read, as in the reference); placement: "simple", "diagonal"
 * or anything else for random. compute_stats != 0 fills out->stats. */
int spmvgen_generate(long nr_rows, long nr_cols, double avg_nnz_per_row,
		double std_nnz_per_row, const char *distribution, unsigned int seed,
		const char *placement, double bw_scaled, double skew,
		double avg_num_neighbours, double cross_row_similarity,
		int compute_stats, spmvgen_csr *out);

void spmvgen_free(spmvgen_csr *csr);

/* Row degrees only (phase 1 of the generator): deg[nr_rows] and the total
 * nnz. Cheap way to size a matrix before generating it. */
int spmvgen_degrees(long nr_rows, long nr_cols, double avg_nnz_per_row,
		double std_nnz_per_row, const char *distribution, unsigned int seed,
		double skew, int32_t *deg, long *nnz_out);

/* Benchmark input definitions (both exact and thread-count independent):
 *   values: v' = 2 v - 1 applied in place (v = k / 2^31, so exact);
 *   x:      x_i = 2 u_i - 1, u_i = random_r() / 2^31 from ONE glibc
 *           random_r stream (256-byte state) seeded with `seed`.
 * The reference fills x per OpenMP thread (src/bench_spmv.cpp:546-560), which
 * depends on the thread count; this definition does not. */
void spmvgen_map_values(double *values, long nnz);
int spmvgen_fill_x(double *x, long n, unsigned int seed);

/* Descriptor statistics of any CSR matrix (same definitions as above). */
int spmvgen_compute_stats(const int32_t *row_ptr, const int32_t *col_ind,
		long m, long n, spmvgen_stats *out);

/* Matrix Market import, the reference's real-matrix input path (import_file,
 * src/bench.cpp:126-238: mtx_read, mtx_values_convert_to_real and coo_to_csr
 * with sorted columns; lib/storage_formats/matrix_market/matrix_market.c:
 * 152-331,421-456). Coordinate files with real, integer, complex (read as the
 * modulus) or pattern (read as 1.0) fields and general, symmetric,
 * skew-symmetric or Hermitian symmetry; a file without a banner is read as
 * coordinate real general. With expand_symmetry each stored off-diagonal
 * entry of a symmetric matrix also gives its mirror (negated for
 * skew-symmetric). Rows ascending, columns ascending within a row,
 * duplicates kept in file order. out->stats is zero; *symmetric (may be
 * NULL) receives whether the file declares a symmetry. The arrays are
 * released by spmvgen_free(). */
int spmvgen_read_mtx(const char *path, int expand_symmetry, spmvgen_csr *out, int *symmetric);

const char *spmvgen_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SPMVGEN_H */

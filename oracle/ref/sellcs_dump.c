/*
 * Test-infrastructure harness (checker only, never shipped): runs the
 * reference's SELL-C-sigma conversion, compiled from its sources in place by
 * oracle/Makefile, on a CSR matrix read from a gen_dump binary.
 *
 * Calls: sellcs_init_params()               (src/spmv_kernels/sell-C-s/sellcs_format.c:32-51)
 *        sellcs_create_matrix_from_CSR_rd() (sellcs_format.c:202-223, dfc=0 as in
 *                                            src/spmv_kernels/sell_c_s.cpp:72)
 *   which runs get_order_by_row_size()      (radix_sort.c:103-122)
 *   and csr_to_sellcs()                     (sellcs_format.c:137-200).
 *
 * Usage: sellcs_dump csr.bin C sigma out.bin
 * out.bin: int64 nslices, int64 total, then int64 row_order[m],
 *          int64 slice_widths[nslices], int64 slice_pointers[nslices+1],
 *          int64 column_indices[total], float64 values[total].
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include "sellcs-spmv.h"

int main(int argc, char **argv)
{
	if (argc < 5) {
		fprintf(stderr, "usage: %s csr.bin C sigma out.bin\n", argv[0]);
		return 2;
	}
	FILE *f = fopen(argv[1], "rb");
	if (!f) { perror("fopen"); return 1; }
	int64_t hdr[3];
	if (fread(hdr, sizeof(hdr), 1, f) != 1) return 1;
	int64_t m = hdr[0], n = hdr[1], nnz = hdr[2];
	int32_t *rp = malloc((m + 1) * sizeof(int32_t));
	int32_t *ci = malloc((nnz ? nnz : 1) * sizeof(int32_t));
	double *va = malloc((nnz ? nnz : 1) * sizeof(double));
	if (fread(rp, sizeof(int32_t), m + 1, f) != (size_t)(m + 1)) return 1;
	if (fread(ci, sizeof(int32_t), nnz, f) != (size_t) nnz) return 1;
	if (fread(va, sizeof(double), nnz, f) != (size_t) nnz) return 1;
	fclose(f);
	/* widen to the library's index type, as src/spmv_kernels/sell_c_s.cpp:64-71 does */
	index_t *rp64 = malloc((m + 1) * sizeof(index_t));
	index_t *ci64 = malloc((nnz ? nnz : 1) * sizeof(index_t));
	for (int64_t i = 0; i <= m; i++) rp64[i] = rp[i];
	for (int64_t i = 0; i < nnz; i++) ci64[i] = ci[i];

	sellcs_matrix_t s;
	sellcs_init_params((uint64_t) atol(argv[2]), (uint64_t) atol(argv[3]), &s);
	sellcs_create_matrix_from_CSR_rd(m, n, rp64, ci64, va, 0, 0, &s);
	int64_t total = s.slice_pointers[s.nslices];
	FILE *o = fopen(argv[4], "wb");
	if (!o) { perror("fopen"); return 1; }
	int64_t oh[2] = {s.nslices, total};
	fwrite(oh, sizeof(oh), 1, o);
	fwrite(s.row_order, sizeof(index_t), m, o);
	fwrite(s.slice_widths, sizeof(index_t), s.nslices, o);
	fwrite(s.slice_pointers, sizeof(index_t), s.nslices + 1, o);
	fwrite(s.column_indices, sizeof(index_t), total, o);
	fwrite(s.values, sizeof(elem_t), total, o);
	fclose(o);
	printf("%lld %lld\n", (long long) s.nslices, (long long) total);
	return 0;
}

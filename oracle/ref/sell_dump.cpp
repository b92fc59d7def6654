// Test-infrastructure harness (checker only, never shipped): builds the
// reference's SELL format, compiled from its own source in place
// (src/spmv_kernels/sell.cpp, included below so that its members are
// reachable), on a CSR matrix read from a gen_dump binary, runs its spmv once
// on x_j = ((j % 13) + 1) / 8 - 1 (exact in binary), and writes what it built.
//
// Usage: OMP_NUM_THREADS=T sell_dump csr.bin out.bin
// out.bin: int64 VL, T, num_row_clusters, nnz_ext;
//          int32 row_cluster_ptr[num_row_clusters + 1];
//          int32 ja[nnz_ext]; float64 a[nnz_ext]; float64 y[m]

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "sell.cpp"

int main(int argc, char **argv)
{
	if (argc < 3) {
		fprintf(stderr, "usage: %s csr.bin out.bin\n", argv[0]);
		return 2;
	}
	FILE *f = fopen(argv[1], "rb");
	if (!f) {
		perror("fopen");
		return 1;
	}
	int64_t hdr[3];
	if (fread(hdr, sizeof(hdr), 1, f) != 1)
		return 1;
	long m = hdr[0], n = hdr[1], nnz = hdr[2];
	INT_T *rp = (INT_T *) malloc((m + 1) * sizeof(INT_T));
	INT_T *ci = (INT_T *) malloc((nnz ? nnz : 1) * sizeof(INT_T));
	double *va = (double *) malloc((nnz ? nnz : 1) * sizeof(double));
	if (fread(rp, sizeof(INT_T), m + 1, f) != (size_t)(m + 1) || fread(ci, sizeof(INT_T), nnz, f) != (size_t) nnz ||
			fread(va, sizeof(double), nnz, f) != (size_t) nnz)
		return 1;
	fclose(f);
	struct SELLArray *S = (struct SELLArray *) csr_to_format(rp, ci, va, m, n, nnz, 0, 0);
	double *x = (double *) malloc((n ? n : 1) * sizeof(double));
	double *y = (double *) malloc((m ? m : 1) * sizeof(double));
	for (long j = 0; j < n; j++)
		x[j] = (double) ((j % 13) + 1) / 8.0 - 1.0;
	S->spmv(x, y);
	FILE *o = fopen(argv[2], "wb");
	int64_t h[4] = {VEC_LEN, omp_get_max_threads(), S->num_row_clusters, S->nnz_ext};
	fwrite(h, 8, 4, o);
	fwrite(S->row_cluster_ptr, sizeof(INT_T), S->num_row_clusters + 1, o);
	fwrite(S->ja, sizeof(INT_T), S->nnz_ext, o);
	fwrite(S->a, sizeof(double), S->nnz_ext, o);
	fwrite(y, sizeof(double), m, o);
	fclose(o);
	return 0;
}

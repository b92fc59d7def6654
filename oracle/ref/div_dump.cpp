// Test-infrastructure harness (checker only, never shipped): builds the
// reference's DIV-RF format, compiled from its own source in place
// (src/spmv_kernels/DIV/div.cpp with -DDIV_TYPE_RF, included below so that
// its file-scope state is reachable), on a CSR matrix read from a gen_dump
// binary, and writes the packet streams it produced.
//
// Usage: OMP_NUM_THREADS=T div_dump csr.bin P out.bin [reorder [tolerance]]
// (the values of DIV_PACKET_REORDERING and DIV_VC_TOLERANCE, default 0)
// out.bin: int64 T; per partition t: int64 nbytes, bytes, int64 ngroups,
//          int64 packet_group_pos[ngroups + 1]
// The packet header's two padding bytes (offsets 22-23 of every packet) are
// left as the reference wrote them (uninitialised); consumers mask them.

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "DIV/div.cpp"

int main(int argc, char **argv)
{
	if (argc < 4) {
		fprintf(stderr, "usage: %s csr.bin P out.bin [reorder [tolerance]]\n", argv[0]);
		return 2;
	}
	setenv("CSRCV_NUM_PACKET_VALS", argv[2], 1);
	setenv("DIV_VC_TOLERANCE", argc > 5 ? argv[5] : "0", 1);
	setenv("DIV_PACKET_REORDERING", argc > 4 ? argv[4] : "0", 1);
	setenv("DIV_STEALING_ENABLED", "0", 1);
	setenv("USE_PROCESSES", "0", 1);
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
	struct Matrix_Format *MF = csr_to_format(rp, ci, va, m, n, nnz, 0, 0);
	(void) MF;
	int T = omp_get_max_threads();
	FILE *o = fopen(argv[3], "wb");
	int64_t t64 = T;
	fwrite(&t64, 8, 1, o);
	for (int t = 0; t < T; t++) {
		struct thread_data *td = tds[t];
		int64_t nb = td->compr_data_size_stealable;
		fwrite(&nb, 8, 1, o);
		fwrite(td->compr_data_stealable, 1, nb, o);
		int64_t ng = td->num_packet_groups;
		fwrite(&ng, 8, 1, o);
		for (long g = 0; g <= ng; g++) {
			int64_t v = td->packet_group_pos[g];
			fwrite(&v, 8, 1, o);
		}
	}
	fclose(o);
	return 0;
}

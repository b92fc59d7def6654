/*
 * oracle.h -- CPU restatement of the reference SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (spmv-dare_amd/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg do, as the checker. Each function cites the
 * reference lines it restates.
 *
 * Pinning (what checks each function against something other than itself):
 *   oracle_csr_naive      closed form on tiny matrices (hand-computed sums),
 *                         oracle_csr_dd (independent exact-ish arithmetic),
 *                         and numpy's dense product within rounding.
 *   oracle_csr_dd         exact rational arithmetic (Python fractions) on
 *                         small cases: |y_dd - y_exact| <= 2^-100 |y|.
 *   oracle_sellcs_*       bit-exact against the reference's own conversion
 *                         (oracle/_ref/sellcs_dump, golden fixtures in
 *                         tests/golden/), brute-force stable sort on tiny m.
 *   oracle_sell_*         bit-exact against the reference's own build
 *                         (oracle/_ref/sell_dump: cluster sizes, ja with the
 *                         per-thread padding carry, a and y; golden fixtures),
 *                         closed form of the padding rule on hand-built
 *                         matrices (empty rows, ragged last slice) at any T.
 *   oracle_sellsorted_*   bit-exact against the reference's own build
 *                         (oracle/_ref/sell_sorted_dump: thread windows,
 *                         permutation, cluster sizes, ja, a and y; golden
 *                         fixtures in tests/golden/), brute-force stable
 *                         sort on tiny m.
 *   oracle_partition_*    brute force over all targets on tiny arrays.
 *   oracle_div_*          round trip (decode(encode(A)) == A), the reference's
 *                         own self-check (src/spmv_kernels/DIV/div.cpp:515-602);
 *                         oracle_div_reduce_precision: the values decoded from
 *                         the reference's own lossy streams (golden fixtures),
 *                         and the error bound / trailing-zero invariants.
 */
#ifndef SPMV_ORACLE_H
#define SPMV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* csr_naive: y[i] = sum_j a[j]*x[ja[j]], left to right, multiply then add,
 * never fused (src/spmv_kernels/csr.cpp:166-182). */
void oracle_csr_naive(const int32_t *rp, const int32_t *col, const double *val,
		const double *x, double *y, long m);

/* Double-double compensated row sums (TwoProd by fma, TwoSum): y = hi + lo
 * carries ~106 bits. Plays the role of the reference's _Float128 Kahan
 * y_gold (src/bench_spmv.cpp:152-172). */
void oracle_csr_dd(const int32_t *rp, const int32_t *col, const double *val,
		const double *x, double *y_hi, double *y_lo, long m);

/* OpenMP csr_naive with the reference's equal-row split
 * (csr.cpp:74 loop_partitioner_balance_iterations, compute_csr :188-211):
 * the CPU baseline. threads <= 0: all. Returns the threads used. */
int oracle_csr_naive_omp(const int32_t *rp, const int32_t *col, const double *val,
		const double *x, double *y, long m, int threads);

/* loop_partitioner_balance_iterations (lib/parallel_util.h:50-96), incr 1. */
void oracle_partition_iterations(long W, long w, long start, long end, long *s, long *e);

/* loop_partitioner_balance_prefix_sums (lib/parallel_util.h:156-184) with
 * the closest-prefix binary search of lib/macros/macrolib.h:537-575. */
void oracle_partition_prefix_sums(long W, long w, const int64_t *sums, long N,
		int64_t total, long *s, long *e);

/* SELL-C-sigma conversion, reference semantics
 * (sell-C-s/radix_sort.c:36-122, sellcs_format.c:137-200).
 * order: per sigma-window stable sort by descending row length.
 * width_rule 0: slice width = length of the slice's first sorted row
 *               (sellcs_format.c:152) -- the reference;
 *            1: slice width = longest row of the slice (defined for any sigma).
 * Returns the total number of stored entries. */
void oracle_sellcs_order(const int32_t *rp, long m, long sigma, int64_t *row_order);
long oracle_sellcs_layout(const int32_t *rp, long m, long C, const int64_t *row_order,
		int width_rule, int64_t *widths, int64_t *slice_ptr);
/* Zero-initialised, column-major fill (sellcs_format.c:164-186). Rows whose
 * length exceeds the slice width (width_rule 0 with sigma % C != 0) are
 * truncated here instead of overflowing; returns the number of truncated
 * entries (0 for every well-defined case). */
long oracle_sellcs_fill(const int32_t *rp, const int32_t *col, const double *val, long m,
		long C, const int64_t *row_order, const int64_t *widths, const int64_t *slice_ptr,
		int64_t *cols_out, double *vals_out);
/* Lane-sequential fused multiply-add over each slice, y[row_order[r]] = acc
 * (sellcs_mv_kernels_epi.c:247-316). */
void oracle_sellcs_spmv(long m, long C, const int64_t *row_order, const int64_t *widths,
		const int64_t *slice_ptr, const int64_t *cols, const double *vals,
		const double *x, double *y);

/* SELL (src/spmv_kernels/sell.cpp:105-199) with VL-row clusters and the
 * reference's per-thread padding column (T threads, nnz-balanced split of
 * clusters, sell.cpp:161). */
long oracle_sell_layout(const int32_t *rp, long m, long VL, int64_t *cluster_ptr);
void oracle_sell_fill(const int32_t *rp, const int32_t *col, const double *val, long m,
		long VL, long T, const int64_t *cluster_ptr, int64_t *ja, double *a);
/* compute_sell (sell.cpp:239-280): per cluster, lane-wise fma over the
 * width; only y[0..m) is written. */
void oracle_sell_spmv(long m, long VL, const int64_t *cluster_ptr, const int64_t *ja,
		const double *a, const double *x, double *y);

/* SELL_SORTED (src/spmv_kernels/sell_sorted.cpp:119-316, 346-427): rows
 * stable-sorted by ascending length inside row windows, then VL-row
 * clusters as SELL. The reference's windows are its threads' ranges
 * (oracle_sellsorted_windows); any cluster-aligned windows may be given.
 * perm[i] = sorted position of row i. */
void oracle_sellsorted_windows(const int32_t *rp, long m, long VL, long T, int64_t *row_start);
void oracle_sellsorted_order(const int32_t *rp, long m, long nwin, const int64_t *row_start, int32_t *perm);
long oracle_sellsorted_layout(const int32_t *rp, long m, long VL, const int32_t *perm, int64_t *cluster_ptr);
void oracle_sellsorted_fill(const int32_t *rp, const int32_t *col, const double *val, long m, long VL,
		long nwin, const int64_t *row_start, const int32_t *perm, const int64_t *cluster_ptr,
		int64_t *ja, double *a);
void oracle_sellsorted_spmv(long m, long VL, const int64_t *cluster_ptr, const int64_t *ja, const double *a,
		const int32_t *perm, const double *x, double *y);

/* Decode a DIV-RF packet stream with L value lanes (the inverse of
 * compress_kernel_div, src/spmv_kernels/DIV/div_kernels_rf.h:538-919, as the
 * reference's decompress path reads it, :948-1162 and :1247-1267): writes one
 * (row, col, value) per stored coordinate, packet after packet, in the
 * packet's coordinate order. Returns the number written, -1 on overflow. */
long oracle_div_decode(const unsigned char *bytes, long nbytes, long L, int32_t *rows, int32_t *cols,
		double *vals, long cap);

/* DIV_RF's lossy value rounding, reduce_precision of
 * src/spmv_kernels/DIV/div_kernels_rf.h:467-510 for one value: returns the
 * rounded value, or NaN when the reference would stop with error()
 * (tolerance >= 1 or a candidate beyond the tolerance). */
double oracle_div_reduce_precision(double v, double tolerance);

/* FNV-1a 64 of a byte range: the hash the generator fixtures record. */
uint64_t oracle_fnv1a(const void *p, long nbytes);

#ifdef __cplusplus
}
#endif

#endif

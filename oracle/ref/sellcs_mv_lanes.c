/*
 * Portable SELL-C-sigma SpMV kernel for the CPU baseline -- TEST
 * INFRASTRUCTURE (linked only into oracle/_ref/cpu_sell_c_s_d.exe).
 *
 * The reference's only x86-less kernel for this format is written with RISC-V
 * vector intrinsics (src/spmv_kernels/sell-C-s/sellcs_mv_kernels_epi.c:178-318,
 * `sellcs_mv_d`), which no x86 compiler accepts. This file restates what that
 * kernel computes, lane by lane, so that the reference's own conversion
 * (radix_sort.c, sellcs_format.c), task analyzer (sellcs_analyzer.c), task
 * scheduler (sellcs_mv.c) and driver (bench.cpp, bench_spmv.cpp) can be
 * timed on the host:
 *
 *   for each slice s in [start_slice, end_slice):
 *     gvl = min(C, nrows - s*C)               (vsetvl, :200-208)
 *     acc[lane] = 0                            (vfmv_v_f, :216-221)
 *     for i < slice_widths[s]:                 (:236-300)
 *       acc[lane] += values[ptr + i*C + lane] * x[cols[ptr + i*C + lane]]
 *                                              (vfmacc: one rounding, fma())
 *     y[row_order[s*C + lane]] = acc[lane]     (vsuxei scatter, :303-315)
 *
 * The lane loop is a plain C loop; the compiler vectorises it with AVX2
 * gathers at -O3 -march=x86-64-v3.
 */
#include <math.h>

#include "sellcs-spmv.h"
#include "sellcs_mv.h"

void sellcs_mv_d(const sellcs_matrix_t *matrix, const elem_t *restrict x, elem_t *restrict y,
		const index_t start_slice, const index_t end_slice)
{
	const index_t C = (index_t) matrix->C;
	elem_t acc[1024];
	for (index_t s = start_slice; s < end_slice; s++) {
		const index_t row0 = s << matrix->shift;
		const index_t lanes = row0 + C > matrix->nrows ? matrix->nrows - row0 : C;
		const elem_t *v = matrix->values + matrix->slice_pointers[s];
		const index_t *c = matrix->column_indices + matrix->slice_pointers[s];
		const index_t w = matrix->slice_widths[s];
		for (index_t l = 0; l < lanes; l++)
			acc[l] = 0.0;
		for (index_t i = 0; i < w; i++) {
			for (index_t l = 0; l < lanes; l++)
				acc[l] = fma(v[l], x[c[l]], acc[l]);
			v += C;
			c += C;
		}
		for (index_t l = 0; l < lanes; l++)
			y[matrix->row_order[row0 + l]] = acc[l];
	}
}

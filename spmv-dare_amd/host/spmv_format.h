// The plugin interface of the benchmark driver, declared with the same member
// layout and signatures as the reference's src/spmv_kernels/spmv_kernel.h:8-29
// so that an object compiled against either header links with the other's
// driver (the layout is the drop-in contract, hence identical by necessity).
#ifndef SPMV_FORMAT_H
#define SPMV_FORMAT_H

#ifndef INT_T
#define INT_T int32_t
#endif
#ifndef ValueType
#define ValueType double
#endif
#ifndef ValueTypeReference
#define ValueTypeReference double
#endif

#include <stdint.h>

struct Matrix_Format
{
	char * format_name;
	long m;
	long n;
	long nnz;
	double mem_footprint;
	double csr_mem_footprint;

	virtual void spmv(ValueType * x, ValueType * y) = 0;
	virtual void statistics_start() = 0;
	virtual int statistics_print_data(char * buf, long buf_n) = 0;

	Matrix_Format(long m, long n, long nnz) : m(m), n(n), nnz(nnz)
	{
		csr_mem_footprint = nnz * (sizeof(ValueType) + sizeof(INT_T)) + (m+1) * sizeof(INT_T);
	}
};

struct Matrix_Format * csr_to_format(INT_T * row_ptr, INT_T * col_ind, ValueTypeReference * values,
		long m, long n, long nnz, long symmetric, long symmetry_expanded);
int statistics_print_labels(char * buf, long buf_n);

#endif

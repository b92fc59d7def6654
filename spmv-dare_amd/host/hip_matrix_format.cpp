// HIP-backed Matrix_Format subclass: the drop-in replacement of one of the
// reference's format translation units (src/spmv_kernels/csr.cpp,
// csr_vec.cpp, sell.cpp, sell_c_s.cpp, sell_sorted.cpp, DIV/div.cpp). One object file per
// format, selected at compile time with -DSPMVHIP_FORMAT_KIND=<enum
// spmvhip_kind>, exactly as the reference links one format per executable
// (src/Makefile_in:47-74). Link it with the reference's bench.o/bench_spmv.o
// (compiled against src/spmv_kernels/spmv_kernel.h) or with
// bench_driver.cpp, plus -lspmvhip.
//
// Environment (config.sh-style, read at csr_to_format):
//   GPU_DEVICE        device ordinal (default 0)
//   GPU_KERNEL=1      device-resident mode: x/y are bound on the first call and
//                     stay on the GPU; y is copied back by statistics_print_data()
//                     and when the object is destroyed
//   SELLCS_C, SELLCS_SIGMA          SELL-C-sigma slice height / sort window
//                                   (SELLCS_SIGMA is SELL_SORTED's row window too)
//   CSRV_BLOCK_NNZ, CSRV_LANES, CSRV_SPLIT_NNZ, CSRV_NARROW_LOADS   CSR-vector row-block tuning
//   SPMV_XCD_REMAP, SELL_UNROLL, SELL_PIPELINE, SELL_X_WINDOW, SELL_Y_MODE
//   DIV_THREADS, DIV_LANES, CSRCV_NUM_PACKET_VALS, DIV_LAYOUT   DIV-RF codec / device layout
//   DIV_ENCODER                      where the DIV-RF packets are encoded: 0 automatic
//                                    (the device when possible), 1 host, 2 device
//   DIV_PACKET_REORDERING=1          packet groups in median-column order (config.sh's variable)
//   DIV_VC_TOLERANCE=t               lossy DIV-RF values, relative error <= t (config.sh's variable)

#include <cstdio>
#include <cstdlib>
#include <cstring>

#if defined(SPMVHIP_USE_REFERENCE_HEADER)
#include "spmv_kernel.h"
#else
#include "spmv_format.h"
#endif
#include "spmvhip.h"

#ifndef SPMVHIP_FORMAT_KIND
#define SPMVHIP_FORMAT_KIND SPMVHIP_CSR_VEC
#endif

namespace {

[[noreturn]] void die(const char *what, int rc)
{
	// the reference's error(): message and exit(EXIT_FAILURE) (lib/debug.h:135)
	fprintf(stderr, "%s failed (%d): %s\n", what, rc, spmvhip_last_error());
	exit(EXIT_FAILURE);
}

long env_long(const char *name, long dflt)
{
	const char *v = getenv(name);
	return (v && *v) ? atol(v) : dflt;
}

struct HipMatrixFormat : Matrix_Format {
	spmvhip_format *h = nullptr;
	bool device_resident = false;
	const double *bound_x = nullptr;
	double *bound_y = nullptr;

	HipMatrixFormat(long m, long n, long nnz) : Matrix_Format(m, n, nnz) {}

	~HipMatrixFormat()
	{
		if (h) {
			spmvhip_sync_host_y(h);
			spmvhip_destroy(h);
		}
	}

	void spmv(ValueType *x, ValueType *y) override
	{
		int rc;
		if (device_resident) {
			if (x != bound_x || y != bound_y) {
				if (bound_y)
					spmvhip_sync_host_y(h);
				rc = spmvhip_bind_host_xy(h, x, y);
				if (rc)
					die("spmvhip_bind_host_xy", rc);
				bound_x = x;
				bound_y = y;
			}
		}
		rc = spmvhip_spmv(h, x, y);
		if (rc)
			die("spmvhip_spmv", rc);
	}

	void statistics_start() override { spmvhip_statistics_start(h); }

	int statistics_print_data(char *buf, long buf_n) override
	{
		spmvhip_sync_host_y(h);
		return spmvhip_statistics_print_data(h, buf, buf_n);
	}
};

} // namespace

struct Matrix_Format *csr_to_format(INT_T *row_ptr, INT_T *col_ind, ValueTypeReference *values, long m, long n,
		long nnz, long symmetric, long symmetry_expanded)
{
	int rc = spmvhip_init((int) env_long("GPU_DEVICE", 0));
	if (rc)
		die("spmvhip_init", rc);
	spmvhip_params p;
	memset(&p, 0, sizeof(p));
	p.kind = SPMVHIP_FORMAT_KIND;
	p.C = (int) env_long("SELLCS_C", 0);
	p.sigma = env_long("SELLCS_SIGMA", 0);
	p.block_nnz = (int) env_long("CSRV_BLOCK_NNZ", 0);
	p.lanes_per_row = (int) env_long("CSRV_LANES", 0);
	p.split_nnz = (int) env_long("CSRV_SPLIT_NNZ", 0);
	p.xcd_remap = (int) env_long("SPMV_XCD_REMAP", 0);
	p.unroll = (int) env_long("SELL_UNROLL", 0);
	p.pipeline = (int) env_long("SELL_PIPELINE", 0);
	p.x_window = (int) env_long("SELL_X_WINDOW", 0);
	p.y_mode = (int) env_long("SELL_Y_MODE", 0);
	p.div_threads = (int) env_long("DIV_THREADS", 0);
	p.div_lanes = (int) env_long("DIV_LANES", 0);
	p.div_acc_copies = (int) env_long("DIV_ACC_COPIES", 0);
	p.narrow_loads = (int) env_long("CSRV_NARROW_LOADS", 0);
	p.div_layout = (int) env_long("DIV_LAYOUT", 0);
	p.div_reorder = (int) env_long("DIV_PACKET_REORDERING", 0);
	p.div_encoder = (int) env_long("DIV_ENCODER", 0);
	if (const char *tol = getenv("DIV_VC_TOLERANCE"))
		p.div_tolerance = atof(tol);
	p.div_packet_vals = env_long("CSRCV_NUM_PACKET_VALS", 0);

	HipMatrixFormat *f = new HipMatrixFormat(m, n, nnz);
	rc = spmvhip_csr_to_format(row_ptr, col_ind, values, m, n, nnz, symmetric, symmetry_expanded, &p, &f->h);
	if (rc)
		die("csr_to_format", rc);
	spmvhip_info info;
	spmvhip_get_info(f->h, &info);
	f->format_name = (char *) info.format_name;
	f->mem_footprint = info.mem_footprint;
	f->device_resident = env_long("GPU_KERNEL", 0) != 0;
	return f;
}

int statistics_print_labels(char *buf, long buf_n)
{
	return spmvhip_statistics_print_labels(buf, buf_n);
}

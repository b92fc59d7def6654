// extern "C" boundary of the gfx950 SpMV layer (declarations and the
// reference interface each entry point replaces: include/spmvhip.h).

#include <cstdarg>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "format.hpp"

namespace spmvhip {

static thread_local char g_err[1024];

int set_error(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
	return code;
}

static const char *kind_name(int kind)
{
	switch (kind) {
	case SPMVHIP_CSR_NAIVE: return "HIP_CSR_NAIVE";
	case SPMVHIP_CSR_VEC: return "HIP_CSR_VEC";
	case SPMVHIP_SELL: return "HIP_SELL";
	case SPMVHIP_SELL_C_S: return "HIP_SELL_C_s";
	case SPMVHIP_DIV_RF: return "HIP_DIV_RF";
	case SPMVHIP_SELL_SORTED: return "HIP_SELL_SORTED";
	}
	return "?";
}

static int ensure_device()
{
	static int ready = 0;
	if (ready)
		return SPMVHIP_OK;
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n == 0)
		return set_error(SPMVHIP_ENODEV, "no HIP device available (%s)",
				e != hipSuccess ? hipGetErrorString(e) : "0 devices");
	ready = 1;
	return SPMVHIP_OK;
}

// One non-blocking stream per device, shared by every format handle and
// batched launch on that device (conversions, the spmv calls that name no
// stream, batched launches). The HIP runtime backs each stream with host
// resources of its own (a worker thread per queue), so a stream per handle
// does not scale to the thousands of handles a batched workload holds (two
// or three per matrix); and streams land on different hardware queues, on
// which the same persistent launch measured 2-3 % apart, so every launch of
// the layer goes through the one queue.
constexpr int kMaxDevices = 64;
static std::mutex g_stream_mu;
static hipStream_t g_streams[kMaxDevices];

int shared_stream(hipStream_t *out)
{
	int dev = 0;
	SPMVHIP_CHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		return set_error(SPMVHIP_EINVAL, "device %d beyond %d", dev, kMaxDevices);
	std::lock_guard<std::mutex> lock(g_stream_mu);
	if (!g_streams[dev])
		SPMVHIP_CHECK(hipStreamCreateWithFlags(&g_streams[dev], hipStreamNonBlocking));
	*out = g_streams[dev];
	return SPMVHIP_OK;
}

// Structural validation of a host CSR matrix (the reference trusts its input;
// here a bad matrix is an error status, not a wild access on the device).
static int validate_csr(const int32_t *rp, const int32_t *col, long m, long n, long nnz)
{
	if (rp[0] != 0 || rp[m] != nnz)
		return set_error(SPMVHIP_EFORMAT, "row_ptr[0] = %d, row_ptr[m] = %d, nnz = %ld", rp[0], rp[m], nnz);
	for (long i = 0; i < m; i++)
		if (rp[i + 1] < rp[i])
			return set_error(SPMVHIP_EFORMAT, "row_ptr decreases at row %ld", i);
	for (long j = 0; j < nnz; j++)
		if (col[j] < 0 || col[j] >= n)
			return set_error(SPMVHIP_EFORMAT, "col_ind[%ld] = %d outside [0, %ld)", j, col[j], n);
	return SPMVHIP_OK;
}

static int convert(spmvhip_format *f, const int32_t *h_rp, const int32_t *h_col, const double *h_val)
{
	SPMVHIP_CHECK(hipEventRecord(f->ev0, f->stream));
	int rc;
	switch (f->kind) {
	case SPMVHIP_CSR_NAIVE: rc = csr_naive_convert(f); break;
	case SPMVHIP_CSR_VEC: rc = csr_vec_convert(f); break;
	case SPMVHIP_SELL: rc = sell_convert(f, kSellUnsorted); break;
	case SPMVHIP_SELL_C_S: rc = sell_convert(f, kSellDescending); break;
	case SPMVHIP_SELL_SORTED: rc = sell_convert(f, kSellAscending); break;
	case SPMVHIP_DIV_RF: rc = div_convert(f, h_rp, h_col, h_val); break;
	default: rc = set_error(SPMVHIP_EINVAL, "unknown format kind %d", f->kind);
	}
	if (rc)
		return rc;
	SPMVHIP_CHECK(hipEventRecord(f->ev1, f->stream));
	SPMVHIP_CHECK(hipEventSynchronize(f->ev1));
	float ms = 0;
	SPMVHIP_CHECK(hipEventElapsedTime(&ms, f->ev0, f->ev1));
	f->conv_ms = ms;
	return SPMVHIP_OK;
}

static int create_handle(long m, long n, long nnz, const spmvhip_params *params, spmvhip_format **out)
{
	if (!out)
		return set_error(SPMVHIP_EINVAL, "out is NULL");
	*out = nullptr;
	if (!params)
		return set_error(SPMVHIP_EINVAL, "params is NULL");
	if (m < 0 || n < 0 || nnz < 0 || m >= INT32_MAX || n > INT32_MAX || nnz > INT32_MAX)
		return set_error(SPMVHIP_EINVAL, "shape m=%ld n=%ld nnz=%ld outside the int32 index range", m, n, nnz);
	if (params->kind < 0 || params->kind >= SPMVHIP_NUM_KINDS)
		return set_error(SPMVHIP_EINVAL, "unknown format kind %d", params->kind);
	SPMVHIP_TRY(ensure_device());
	spmvhip_format *f = new (std::nothrow) spmvhip_format();
	if (!f)
		return set_error(SPMVHIP_ENOMEM, "host allocation of the handle");
	f->p = *params;
	f->kind = params->kind;
	f->m = m;
	f->n = n;
	f->nnz = nnz;
	f->name = kind_name(f->kind);
	f->csr_mem_footprint = (double) nnz * (sizeof(double) + sizeof(int32_t)) + (double) (m + 1) * sizeof(int32_t);
	int src = shared_stream(&f->stream);
	if (src) {
		delete f;
		return src;
	}
	hipError_t e = hipEventCreate(&f->ev0);
	if (e == hipSuccess)
		e = hipEventCreate(&f->ev1);
	if (e != hipSuccess) {
		spmvhip_destroy(f);
		return set_error(SPMVHIP_EHIP, "stream/event creation: %s", hipGetErrorString(e));
	}
	*out = f;
	return SPMVHIP_OK;
}

static int run_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s);

int format_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s)
{
	return run_spmv(f, x, y, s);
}

static int run_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s)
{
	switch (f->kind) {
	case SPMVHIP_CSR_NAIVE: return csr_naive_spmv(f, x, y, s);
	case SPMVHIP_CSR_VEC: return csr_vec_spmv(f, x, y, s);
	case SPMVHIP_SELL:
	case SPMVHIP_SELL_C_S:
	case SPMVHIP_SELL_SORTED: return sell_spmv(f, x, y, s);
	case SPMVHIP_DIV_RF: return div_spmv(f, x, y, s);
	}
	return set_error(SPMVHIP_EINVAL, "unknown format kind %d", f->kind);
}

static int ensure_xy(spmvhip_format *f)
{
	if (!f->dx.p)
		SPMVHIP_TRY(f->dx.alloc((size_t) (f->n ? f->n : 1) * sizeof(double)));
	if (!f->dy.p)
		SPMVHIP_TRY(f->dy.alloc((size_t) (f->m ? f->m : 1) * sizeof(double)));
	return SPMVHIP_OK;
}

} // namespace spmvhip

using namespace spmvhip;

extern "C" {

const char *spmvhip_last_error(void)
{
	return g_err;
}

int spmvhip_init(int device)
{
	SPMVHIP_TRY(ensure_device());
	SPMVHIP_CHECK(hipSetDevice(device));
	hipDeviceProp_t prop;
	SPMVHIP_CHECK(hipGetDeviceProperties(&prop, device));
	if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
		return set_error(SPMVHIP_ENODEV, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
	return SPMVHIP_OK;
}

int spmvhip_device_name(char *buf, long buf_n)
{
	SPMVHIP_TRY(ensure_device());
	int dev = 0;
	SPMVHIP_CHECK(hipGetDevice(&dev));
	hipDeviceProp_t prop;
	SPMVHIP_CHECK(hipGetDeviceProperties(&prop, dev));
	snprintf(buf, (size_t) buf_n, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
	return SPMVHIP_OK;
}

int spmvhip_malloc(void **dptr, size_t bytes)
{
	SPMVHIP_TRY(ensure_device());
	hipError_t e = hipMalloc(dptr, bytes ? bytes : 8);
	if (e != hipSuccess)
		return set_error(e == hipErrorOutOfMemory ? SPMVHIP_ENOMEM : SPMVHIP_EHIP, "hipMalloc(%zu): %s", bytes,
				hipGetErrorString(e));
	return SPMVHIP_OK;
}

int spmvhip_free(void *dptr)
{
	if (dptr)
		SPMVHIP_CHECK(hipFree(dptr));
	return SPMVHIP_OK;
}

int spmvhip_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
	if (bytes)
		SPMVHIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
	return SPMVHIP_OK;
}

int spmvhip_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
	if (bytes)
		SPMVHIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
	return SPMVHIP_OK;
}

int spmvhip_memset(void *dptr, int value, size_t bytes)
{
	if (bytes)
		SPMVHIP_CHECK(hipMemset(dptr, value, bytes));
	return SPMVHIP_OK;
}

int spmvhip_synchronize(void)
{
	SPMVHIP_CHECK(hipDeviceSynchronize());
	return SPMVHIP_OK;
}

// Device CSR arrays of a handle. col and val carry a zeroed tail of kCsrPad
// entries: the paired CSR stream loads read entries [j, j + 1] with j even,
// so the last pair of a matrix with an odd nnz reaches one entry past it.
static int alloc_csr(spmvhip_format *f)
{
	SPMVHIP_TRY(f->row_ptr.alloc((size_t) (f->m + 1) * sizeof(int32_t)));
	SPMVHIP_TRY(f->col.alloc((size_t) (f->nnz + kCsrPad) * sizeof(int32_t)));
	SPMVHIP_TRY(f->val.alloc((size_t) (f->nnz + kCsrPad) * sizeof(double)));
	SPMVHIP_CHECK(hipMemsetAsync(f->col.as<int32_t>() + f->nnz, 0, kCsrPad * sizeof(int32_t), f->stream));
	SPMVHIP_CHECK(hipMemsetAsync(f->val.as<double>() + f->nnz, 0, kCsrPad * sizeof(double), f->stream));
	return SPMVHIP_OK;
}

int spmvhip_csr_to_format(const int32_t *row_ptr, const int32_t *col_ind, const double *values, long m, long n,
		long nnz, long symmetric, long symmetry_expanded, const spmvhip_params *params, spmvhip_format **out)
{
	if (symmetric && !symmetry_expanded)
		return set_error(SPMVHIP_ESYMMETRIC,
				"symmetric matrices have to be expanded to be supported by this format");
	if (!row_ptr || (nnz > 0 && (!col_ind || !values)))
		return set_error(SPMVHIP_EINVAL, "NULL CSR array");
	if (m < 0 || n < 0 || nnz < 0)
		return set_error(SPMVHIP_EINVAL, "negative shape");
	SPMVHIP_TRY(validate_csr(row_ptr, col_ind, m, n, nnz));
	spmvhip_format *f = nullptr;
	SPMVHIP_TRY(create_handle(m, n, nnz, params, &f));
	int rc = alloc_csr(f);
	if (!rc) {
		hipError_t e = hipMemcpyAsync(f->row_ptr.p, row_ptr, (m + 1) * sizeof(int32_t), hipMemcpyHostToDevice, f->stream);
		if (e == hipSuccess && nnz)
			e = hipMemcpyAsync(f->col.p, col_ind, nnz * sizeof(int32_t), hipMemcpyHostToDevice, f->stream);
		if (e == hipSuccess && nnz)
			e = hipMemcpyAsync(f->val.p, values, nnz * sizeof(double), hipMemcpyHostToDevice, f->stream);
		if (e == hipSuccess)
			e = hipStreamSynchronize(f->stream);
		if (e != hipSuccess)
			rc = set_error(SPMVHIP_EHIP, "CSR upload: %s", hipGetErrorString(e));
	}
	if (!rc)
		rc = convert(f, row_ptr, col_ind, values);
	if (rc) {
		spmvhip_destroy(f);
		return rc;
	}
	*out = f;
	return SPMVHIP_OK;
}

int spmvhip_csr_to_format_device(const int32_t *d_row_ptr, const int32_t *d_col_ind, const double *d_values, long m,
		long n, long nnz, const spmvhip_params *params, spmvhip_format **out)
{
	if (!d_row_ptr || (nnz > 0 && (!d_col_ind || !d_values)))
		return set_error(SPMVHIP_EINVAL, "NULL CSR array");
	spmvhip_format *f = nullptr;
	SPMVHIP_TRY(create_handle(m, n, nnz, params, &f));
	int rc = alloc_csr(f);
	if (!rc) {
		hipError_t e = hipMemcpyAsync(f->row_ptr.p, d_row_ptr, (m + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice,
				f->stream);
		if (e == hipSuccess && nnz)
			e = hipMemcpyAsync(f->col.p, d_col_ind, nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, f->stream);
		if (e == hipSuccess && nnz)
			e = hipMemcpyAsync(f->val.p, d_values, nnz * sizeof(double), hipMemcpyDeviceToDevice, f->stream);
		if (e == hipSuccess)
			e = hipStreamSynchronize(f->stream);
		if (e != hipSuccess)
			rc = set_error(SPMVHIP_EHIP, "CSR device copy: %s", hipGetErrorString(e));
	}
	if (!rc)
		rc = convert(f, nullptr, nullptr, nullptr);
	if (rc) {
		spmvhip_destroy(f);
		return rc;
	}
	*out = f;
	return SPMVHIP_OK;
}

int spmvhip_spmv_device(spmvhip_format *f, const double *d_x, double *d_y, void *stream)
{
	if (!f || !d_x || !d_y)
		return set_error(SPMVHIP_EINVAL, "NULL handle or vector");
	return run_spmv(f, d_x, d_y, stream ? (hipStream_t) stream : f->stream);
}

int spmvhip_bind_host_xy(spmvhip_format *f, const double *x, double *y)
{
	if (!f || !x || !y)
		return set_error(SPMVHIP_EINVAL, "NULL handle or vector");
	SPMVHIP_TRY(ensure_xy(f));
	SPMVHIP_CHECK(hipMemcpyAsync(f->dx.p, x, f->n * sizeof(double), hipMemcpyHostToDevice, f->stream));
	SPMVHIP_CHECK(hipStreamSynchronize(f->stream));
	f->bound_x = x;
	f->bound_y = y;
	f->y_on_device_only = false;
	return SPMVHIP_OK;
}

int spmvhip_sync_host_y(spmvhip_format *f)
{
	if (!f)
		return set_error(SPMVHIP_EINVAL, "NULL handle");
	if (f->bound_y && f->y_on_device_only) {
		SPMVHIP_CHECK(hipMemcpyAsync(f->bound_y, f->dy.p, f->m * sizeof(double), hipMemcpyDeviceToHost, f->stream));
		SPMVHIP_CHECK(hipStreamSynchronize(f->stream));
		f->y_on_device_only = false;
	}
	return SPMVHIP_OK;
}

int spmvhip_spmv(spmvhip_format *f, const double *x, double *y)
{
	if (!f || !x || !y)
		return set_error(SPMVHIP_EINVAL, "NULL handle or vector");
	SPMVHIP_TRY(ensure_xy(f));
	// per-call device time for statistics_print_data: events bracket the
	// calls since statistics_start (ev0 at the first, ev1 after the last)
	if (f->drv_calls == 0)
		SPMVHIP_CHECK(hipEventRecord(f->ev0, f->stream));
	if (x == f->bound_x && y == f->bound_y) {
		SPMVHIP_TRY(run_spmv(f, f->dx.as<double>(), f->dy.as<double>(), f->stream));
		f->y_on_device_only = true;
	} else {
		SPMVHIP_CHECK(hipMemcpyAsync(f->dx.p, x, f->n * sizeof(double), hipMemcpyHostToDevice, f->stream));
		SPMVHIP_TRY(run_spmv(f, f->dx.as<double>(), f->dy.as<double>(), f->stream));
		SPMVHIP_CHECK(hipMemcpyAsync(y, f->dy.p, f->m * sizeof(double), hipMemcpyDeviceToHost, f->stream));
		SPMVHIP_CHECK(hipStreamSynchronize(f->stream));
	}
	SPMVHIP_CHECK(hipEventRecord(f->ev1, f->stream));
	f->drv_calls++;
	return SPMVHIP_OK;
}

int spmvhip_time_spmv(spmvhip_format *f, const double *d_x, double *d_y, int warmup, int iters, float *ms,
		float *total_ms)
{
	if (!f || !d_x || !d_y || iters < 0 || warmup < 0)
		return set_error(SPMVHIP_EINVAL, "bad arguments");
	for (int i = 0; i < warmup; i++)
		SPMVHIP_TRY(run_spmv(f, d_x, d_y, f->stream));
	SPMVHIP_CHECK(hipStreamSynchronize(f->stream));
	std::vector<hipEvent_t> ev(iters + 1);
	for (auto &e : ev)
		SPMVHIP_CHECK(hipEventCreate(&e));
	int rc = SPMVHIP_OK;
	hipError_t he = hipEventRecord(ev[0], f->stream);
	for (int i = 0; i < iters && he == hipSuccess && rc == SPMVHIP_OK; i++) {
		rc = run_spmv(f, d_x, d_y, f->stream);
		if (rc == SPMVHIP_OK)
			he = hipEventRecord(ev[i + 1], f->stream);
	}
	if (he == hipSuccess && rc == SPMVHIP_OK)
		he = hipEventSynchronize(ev[iters]);
	float tot = 0;
	for (int i = 0; i < iters && he == hipSuccess && rc == SPMVHIP_OK; i++) {
		float t = 0;
		he = hipEventElapsedTime(&t, ev[i], ev[i + 1]);
		if (ms)
			ms[i] = t;
		tot += t;
	}
	for (auto &e : ev)
		(void) hipEventDestroy(e);
	if (rc)
		return rc;
	if (he != hipSuccess)
		return set_error(SPMVHIP_EHIP, "event timing: %s", hipGetErrorString(he));
	if (total_ms)
		*total_ms = tot;
	f->stat_ms_sum += tot;
	f->stat_calls += iters;
	return SPMVHIP_OK;
}

int spmvhip_statistics_start(spmvhip_format *f)
{
	if (!f)
		return set_error(SPMVHIP_EINVAL, "NULL handle");
	f->stat_ms_sum = 0;
	f->stat_calls = 0;
	f->drv_calls = 0;
	return SPMVHIP_OK;
}

int spmvhip_statistics_print_labels(char *buf, long buf_n)
{
	if (!buf || buf_n <= 0)
		return 0;
	int k = snprintf(buf, (size_t) buf_n, ",gbps_alg,pct_peak,kernel_ms,conv_ms,variant");
	return k < buf_n ? k : (int) buf_n - 1;
}

int spmvhip_statistics_print_data(spmvhip_format *f, char *buf, long buf_n)
{
	if (!f || !buf || buf_n <= 0)
		return 0;
	if (f->drv_calls > 0 && hipEventSynchronize(f->ev1) == hipSuccess) {
		float t = 0;
		if (hipEventElapsedTime(&t, f->ev0, f->ev1) == hipSuccess) {
			f->stat_ms_sum = t;
			f->stat_calls = f->drv_calls;
		}
	}
	double ms = f->stat_calls ? f->stat_ms_sum / f->stat_calls : 0;
	double bytes = 12.0 * f->nnz + 4.0 * (f->m + 1) + 8.0 * f->n + 8.0 * f->m;
	double gbps = ms > 0 ? bytes / (ms * 1e-3) / 1e9 : 0;
	char variant[96];
	switch (f->kind) {
	case SPMVHIP_CSR_VEC: snprintf(variant, sizeof(variant), "bnnz%d_L%d", f->block_nnz, f->lanes); break;
	case SPMVHIP_SELL:
	case SPMVHIP_SELL_C_S:
	case SPMVHIP_SELL_SORTED: snprintf(variant, sizeof(variant), "C%d_s%ld_u%d", f->p.C, f->p.sigma, f->p.unroll); break;
	case SPMVHIP_DIV_RF: snprintf(variant, sizeof(variant), "L%d_P%ld", f->div_lanes, f->p.div_packet_vals); break;
	default: snprintf(variant, sizeof(variant), "naive");
	}
	int k = snprintf(buf, (size_t) buf_n, ",%.3f,%.2f,%.6f,%.3f,%s", gbps, 100.0 * gbps / 8000.0, ms, f->conv_ms,
			variant);
	return k < buf_n ? k : (int) buf_n - 1;
}

int spmvhip_get_info(const spmvhip_format *f, spmvhip_info *info)
{
	if (!f || !info)
		return set_error(SPMVHIP_EINVAL, "NULL argument");
	memset(info, 0, sizeof(*info));
	info->format_name = f->name;
	info->m = f->m;
	info->n = f->n;
	info->nnz = f->nnz;
	info->mem_footprint = f->mem_footprint;
	info->csr_mem_footprint = f->csr_mem_footprint;
	info->conv_ms = f->conv_ms;
	info->padded_nnz = is_sell(f->kind) ? f->padded : f->nnz;
	info->num_work_items = f->kind == SPMVHIP_CSR_VEC ? f->nblocks
		: (f->kind == SPMVHIP_DIV_RF ? f->div_npackets : f->nslices);
	info->num_split_rows = f->nsplit;
	return SPMVHIP_OK;
}

int spmvhip_export(const spmvhip_format *f, int which, void *buf, size_t *bytes_inout)
{
	if (!f || !bytes_inout)
		return set_error(SPMVHIP_EINVAL, "NULL argument");
	const void *src = nullptr;
	size_t need = 0;
	bool sell = is_sell(f->kind);
	switch (which) {
	case 0:
		if (!f->row_order.p)
			return set_error(SPMVHIP_EINVAL, "row_order exists only for SELL_C_S and SELL_SORTED");
		src = f->row_order.p;
		need = f->m * sizeof(int32_t);
		break;
	case 1: if (!sell) return set_error(SPMVHIP_EINVAL, "not a SELL format");
		src = f->slice_w.p; need = f->nslices * sizeof(int32_t); break;
	case 2: if (!sell) return set_error(SPMVHIP_EINVAL, "not a SELL format");
		src = f->slice_ptr.p; need = (f->nslices + 1) * sizeof(int64_t); break;
	case 3: if (!sell) return set_error(SPMVHIP_EINVAL, "not a SELL format");
		src = f->s_cols.p; need = f->padded * sizeof(int32_t); break;
	case 4: if (!sell) return set_error(SPMVHIP_EINVAL, "not a SELL format");
		src = f->s_vals.p; need = f->padded * sizeof(double); break;
	case 5:
		if (f->kind != SPMVHIP_CSR_VEC)
			return set_error(SPMVHIP_EINVAL, "row blocks exist only for CSR_VEC");
		src = f->blocks.p;
		need = (f->nblocks + 1) * sizeof(int32_t);
		break;
	case 6:
	case 7:
	case 8:
		if (f->kind != SPMVHIP_DIV_RF)
			return set_error(SPMVHIP_EINVAL, "decoded streams exist only for DIV_RF");
		return div_export_decoded(f, which, buf, bytes_inout);
	case 10:
	case 11:
		if (f->kind != SPMVHIP_DIV_RF)
			return set_error(SPMVHIP_EINVAL, "packet streams exist only for DIV_RF");
		src = which == 10 ? f->div_stream.p : f->div_pk_off.p;
		need = which == 10 ? (size_t) f->div_bytes : (size_t) (f->div_npackets + 1) * sizeof(int64_t);
		break;
	case 9:
		if (!sell || f->x_cap == 0)
			return set_error(SPMVHIP_EINVAL, "x windows exist only for SELL formats built with x_window");
		src = f->x_win.p;
		need = 2 * ((f->nslices + kWavesPerBlock - 1) / kWavesPerBlock) * sizeof(int32_t);
		break;
	default: return set_error(SPMVHIP_EINVAL, "unknown export %d", which);
	}
	if (!buf) {
		*bytes_inout = need;
		return SPMVHIP_OK;
	}
	if (*bytes_inout < need)
		return set_error(SPMVHIP_EINVAL, "export %d needs %zu bytes, got %zu", which, need, *bytes_inout);
	if (need)
		SPMVHIP_CHECK(hipMemcpy(buf, src, need, hipMemcpyDeviceToHost));
	*bytes_inout = need;
	return SPMVHIP_OK;
}

void spmvhip_destroy(spmvhip_format *f)
{
	if (!f)
		return;
	if (f->stream)
		(void) hipStreamSynchronize(f->stream);
	if (f->ev0)
		(void) hipEventDestroy(f->ev0);
	if (f->ev1)
		(void) hipEventDestroy(f->ev1);
	delete f; // DevBuf members free their memory; the stream is the device's shared one
}

} // extern "C"

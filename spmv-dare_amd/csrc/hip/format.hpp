// Internal state of one converted matrix (the object behind spmvhip_format*).
#pragma once

#include "common.hpp"
#include "../div/div_codec.hpp"

#include <vector>

struct spmvhip_format {
	spmvhip_params p{};
	int kind = 0;
	long m = 0, n = 0, nnz = 0;
	const char *name = "";
	hipStream_t stream = nullptr; // the device's stream shared by all handles (capi.hip)
	hipEvent_t ev0 = nullptr, ev1 = nullptr;

	// CSR arrays (all CSR kinds; DIV keeps them only until encoded)
	spmvhip::DevBuf row_ptr, col, val;

	// CSR_VEC: row blocks (row starts, nblocks + 1 entries) and split rows
	spmvhip::DevBuf blocks;
	spmvhip::DevBuf brec;         // int32 x 4 per block: (first row, end row, first nnz, end nnz)
	long nblocks = 0;
	int block_nnz = 512, split_nnz = 8192, lanes = 0;
	unsigned pipe_grid = 0;       // resident workgroups of the persistent CSR kernel (set on first use)
	spmvhip::DevBuf seg_desc;     // int32 triplets (row, j_begin, j_end) per segment
	spmvhip::DevBuf seg_ptr;      // int32, nsplit + 1
	spmvhip::DevBuf split_rows;   // int32, nsplit
	spmvhip::DevBuf partials;     // double, nsegs
	long nsplit = 0, nsegs = 0;

	// SELL / SELL_C_S
	spmvhip::DevBuf row_order;    // int32, m (SELL_C_S only)
	spmvhip::DevBuf row_inv;      // int32, m: sorted position of each row (SELL_C_S, y_mode 3)
	spmvhip::DevBuf ysorted;      // double, nslices * 64: results in sorted order (SELL_C_S, y_mode 3)
	spmvhip::DevBuf slice_w;      // int32, nslices
	spmvhip::DevBuf slice_ptr;    // int64, nslices + 1
	spmvhip::DevBuf s_cols;       // int32, padded
	spmvhip::DevBuf s_vals;       // double, padded
	long nslices = 0, padded = 0;
	spmvhip::DevBuf x_win;        // int32 (lo, width) per workgroup of 4 slices (x_window only)
	int x_cap = 0;                // doubles of x staged in LDS per workgroup (0 = no staging)
	bool x_attr_set = false;      // dynamic-LDS limit raised for the x-window kernel

	// DIV_RF
	spmvhip::DevBuf div_stream;   // packet bytes
	spmvhip::DevBuf div_pk_off;   // int64 per packet: byte offset (16-byte aligned)
	spmvhip::DevBuf div_pk_nnz;   // int64 per packet: first nnz index
	spmvhip::DevBuf div_groups;   // int32 per packet group: first packet
	spmvhip::DevBuf div_group_rows; // int32 per packet group: [row_begin,row_end)
	spmvhip::DevBuf div_order;    // int32 per packet group: the k-th group to run (packet reordering)
	std::vector<int32_t> div_order_host;
	long div_npackets = 0, div_ngroups = 0, div_bytes = 0;
	int div_lanes = 64;
	int div_buf_bytes = 0;        // LDS packet buffer of the staged decoder: largest packet + 16, or 0 if too large
	int div_acc_bytes = 2048;     // LDS row accumulators of the staged decoder (all copies)
	int div_acc_stride = 255;     // rows per accumulator copy: the most rows of any packet, odd
	bool div_slots = false;       // packets in the x-slot layout (div_codec.hpp)
	bool div_il = false;          // lane value streams interleaved by dwords (div_codec.hpp)
	int div_xs_bytes = 0;         // LDS of the staged x slots: the most slots of any packet x 8, 16-rounded

	// host-pointer contract
	const double *bound_x = nullptr;
	double *bound_y = nullptr;
	spmvhip::DevBuf dx, dy;
	bool y_on_device_only = false;

	// statistics
	double conv_ms = 0, mem_footprint = 0, csr_mem_footprint = 0;
	double stat_ms_sum = 0;
	long stat_calls = 0;
	long drv_calls = 0;   // spmvhip_spmv calls since statistics_start
};

namespace spmvhip {

// Per-kind conversion and SpMV entry points (csr.hip, sell.hip, div.hip).
int csr_naive_convert(spmvhip_format *f);
int csr_naive_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s);
int csr_vec_convert(spmvhip_format *f);
int csr_vec_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s);
// Row order of the SELL family: none (SELL), stable descending by length in
// sigma windows (SELL_C_S), stable ascending by length in row windows
// (SELL_SORTED).
enum SellOrder { kSellUnsorted = 0, kSellDescending = 1, kSellAscending = 2 };
inline bool is_sell(int kind)
{
	return kind == SPMVHIP_SELL || kind == SPMVHIP_SELL_C_S || kind == SPMVHIP_SELL_SORTED;
}
int sell_convert(spmvhip_format *f, SellOrder order);
int sell_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s);
int div_convert(spmvhip_format *f, const int32_t *h_row_ptr, const int32_t *h_col,
		const double *h_val);
int div_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s);
int div_export_decoded(const spmvhip_format *f, int which, void *buf, size_t *bytes);
// The current device's stream shared by every handle and batch (capi.hip).
int shared_stream(hipStream_t *out);
// The format's own SpMV on stream s (the kernels spmvhip_spmv_device runs).
int format_spmv(spmvhip_format *f, const double *x, double *y, hipStream_t s);

// DIV-RF members of a batched launch (batch.hip): one persistent kernel over
// the packet groups of all of them (div.hip).
struct BatchDivMat {
	uint64_t bytes, pk_off, grp_pk, grp_row, x, y; // device arrays, as pointer values
	int L, acc_stride, acc_copies, il;
};
struct BatchDivUnit {
	int mat, group;
};
// The member record of a converted DIV_RF handle (its staged decoder's LDS
// needs in *buf_bytes, *acc_bytes); SPMVHIP_EINVAL for a handle the batched
// kernel cannot run (packets decoded from global memory, x-slot layout).
int div_batch_member(const spmvhip_format *f, const double *d_x, double *d_y, BatchDivMat *out, int *buf_bytes,
		int *acc_bytes);
// DIV-RF encoding on the device (div_enc.hip): packets of at most 2048
// values and at most 64 lanes.
bool div_device_encodable(long P, int L);
// The device encoder of one handle: prepare() uploads the packet plan (and
// rounds lossy values on the device), size() computes every packet's bytes
// in a device layout (il: interleaved lane streams, else the reference
// bytes), write() lays the packets out for the last size() into the handle's
// div_stream / div_pk_off (and div_bytes).
class DivDeviceEncoder {
public:
	~DivDeviceEncoder();
	int prepare(spmvhip_format *f, const spmvdiv::PacketPlan &plan, int L);
	int size(bool il, int64_t *max_packet, uint32_t *max_rows);
	int write();

private:
	int launch(bool write, bool il, const int64_t *off, uint8_t *out);
	spmvhip_format *f_ = nullptr;
	int L_ = 64;
	long np_ = 0;
	bool sized_il_ = false;
	const double *vals_ = nullptr;
	DevBuf lossy_, i0_, j0_, nv_, info_, size_;
};
// Resident workgroups per CU of the batched DIV kernel with that much LDS.
int div_batch_occupancy(int lds_bytes, int *per_cu);
// Enqueues the counter reset (nc counters, `stride` ints apart) and the
// batched DIV kernel.
int div_batch_launch(const BatchDivMat *d_mats, const BatchDivUnit *d_units, int nunits, int nc, int stride,
		int *d_counter, int grid, int buf_bytes, int acc_bytes, hipStream_t s);

} // namespace spmvhip

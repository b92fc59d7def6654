// DIV-RF packet codec (host side): the reference's lossless DIV encoding
// with replication factors, restated from src/spmv_kernels/DIV/div.cpp:236-513
// (partition into T nnz-balanced ranges, packetisation, packet groups) and
// src/spmv_kernels/DIV/div_kernels_rf.h:437-919 (select_num_vals_complete_last_row,
// compress_kernel_div), parameterised by the number of partitions T, the
// number of value lanes L and the packet size P.
//
// Packet layout (bytes, little endian), identical to the reference's for the
// same (T, L, P):
//   header   u32 num_vals, num_vals_unique, num_rows, row_min, col_min;
//            u8 row_bits, col_bits; 2 bytes padding (written 0 here, left
//            uninitialised by the reference); u32 num_rfs; u32 lane_size[L]
//   u32 rfs[num_rfs]; u32 unique_per_rf[num_rfs]
//   coords   num_vals x coords_bytes: row_diff | col_diff << row_bits
//   u8 val_lens[num_vals_unique]: len | (trailing_zeros / 4) << 4
//   L lane byte streams (lane 0 first), lane_size[l] bytes each
#pragma once

#include <cstdint>
#include <vector>

namespace spmvdiv {

constexpr int kHeaderFixed = 28; // header bytes before lane_size[]

inline int header_bytes(int L) { return kHeaderFixed + 4 * L; }

struct Stream {
	int T = 1, L = 64;
	long P = 16384;
	std::vector<uint8_t> bytes;      // all packets, partition 0 first
	std::vector<int64_t> pk_off;     // byte offset of each packet, + total
	std::vector<int64_t> pk_nnz;     // first nnz index of each packet, + nnz
	std::vector<int32_t> grp_pk;     // first packet of each packet group, + npackets
	std::vector<int32_t> grp_row;    // first row owned by each group, + m
	std::vector<int32_t> part_grp;   // first group of each partition, + ngroups
	long npackets() const { return (long) pk_off.size() - 1; }
	long ngroups() const { return (long) grp_pk.size() - 1; }
};

// The packets of a CSR matrix before encoding (div.cpp:236-513): for every
// packet its first row (the row holding its first nonzero, empty rows
// skipped), first nonzero and value count, in partition order; the packet
// groups (a group starts with every packet that starts a row) with their
// first packet and first owned row; the first group of each partition.
// Packet boundaries chain, so the plan is built sequentially per partition.
struct PacketPlan {
	std::vector<int32_t> i0;
	std::vector<int64_t> j0;
	std::vector<int32_t> nv;
	std::vector<int32_t> grp_pk;   // first packet of each group, + npackets
	std::vector<int32_t> grp_row;  // first row owned by each group, + m
	std::vector<int32_t> part_grp; // first group of each partition, + ngroups
};
void plan_packets(const int32_t *row_ptr, long m, long nnz, int T, long P, PacketPlan *out);

// Encode a CSR matrix. Returns 0, or -1 with *err set (row span > 255 or
// column span > 32 bits in a packet, which the reference rejects with error()).
int encode(const int32_t *row_ptr, const int32_t *col, const double *val, long m, long n, long nnz,
		int T, int L, long P, Stream *out, const char **err);

// Lossy value compression (DIV_VC_TOLERANCE, src/spmv_kernels/DIV/
// div_kernels_rf.h:467-510, applied to every value before packetisation,
// :612-616): with b = ceil(|log2 tolerance|) fraction bits kept, the low
// 52 - b fraction bits of the value are cleared (truncation) or the value is
// rounded up to the next multiple of 2^(52 - b) in its bit pattern, whichever
// result has more trailing zero bits (the rounded-up one on a tie). Zero, values
// whose fraction is already zero and b >= 52 are kept as they are. Values
// then repeat more often and their byte lengths shrink. tolerance <= 0 is
// lossless. Returns 0, or -1 with *err set when tolerance >= 1 or a rounded
// value's relative error exceeds the tolerance (the reference's error()).
int reduce_precision(const double *in, long n, double tolerance, double *out, const char **err);

// The order in which the packet groups are processed with the reference's
// packet reordering (DIV_PACKET_REORDERING=1, src/spmv_kernels/DIV/div.cpp:
// 377-395, 455-498): inside each partition, the groups in ascending order of
// a key, the key of the partition's k-th group being the median column
// (numpy's "median_unbiased" quantile, lib/array_metrics.c:940-1000,
// truncated to an integer) of the partition's k-th packet -- the reference
// records one key per packet and reads them by group index, which is the
// same packet whenever every group is one packet. Ties keep the group order
// (the reference's quicksort leaves them to its partitioning). Groups own
// disjoint rows, so the order changes which groups run side by side, not y.
std::vector<int32_t> group_order_by_median(const Stream &st, const int32_t *col);

// Device layouts of a packet (the GPU decoders' forms of a reference packet;
// values and value lengths are the reference's bytes in every form).
//
// x slots: the packet's distinct columns, in ascending order, are its x
// slots; they form runs of consecutive columns. A coordinate is row_diff |
// slot << row_bits, and the decoder stages x[column of slot s] for every slot
// into LDS with loads that follow the runs, so its x reads are contiguous
// instead of one cache line per value. Layout (little endian, in order):
//   u32 num_vals, num_vals_unique, num_rows, row_min, num_slots
//   u8 row_bits, slot_bits, coord_bytes, interleaved; u32 num_rfs, num_runs
//   u32 lane_size[L]                      (the reference header's)
//   u32 rfs[num_rfs]; u32 unique_per_rf[num_rfs]
//   u32 run_col[num_runs]                 first column of each run
//   (8-byte aligned) u64 run_start_bits[ceil(num_slots / 64)]: bit s set
//                                         where slot s starts a run
//   (4-byte aligned) coords: num_vals x coord_bytes, in the reference's order
//   u8 val_lens[num_vals_unique]; the lane value streams
// Without x slots the packet is the reference's, header byte 22 (padding in
// the reference) holding the interleaved flag.
//
// Interleaved lane streams: the L value streams as rows of L dwords, lane l's
// stream dword w at byte 4 (w L + l) of the stream section, which starts at
// the next 4-byte aligned offset after the value lengths (rows up to the
// longest stream, zero filled), so that the lanes' reads of their streams
// fall in distinct LDS banks, and a lane's stream base needs no scan of the
// lane sizes. Otherwise the streams follow each other, lane 0 first.
constexpr int kSlotHeaderFixed = 32;

inline int slot_header_bytes(int L) { return kSlotHeaderFixed + 4 * L; }

struct DeviceLayout {
	std::vector<uint8_t> bytes;   // packets, each starting at a 16-byte aligned offset
	std::vector<int64_t> pk_off;  // device byte offset of each packet, + total
	int64_t max_packet = 0;       // largest packet, 16-byte rounded
	uint32_t max_slots = 0, max_rows = 0;
};

// Transcode every packet of `st` (lanes L) into a device layout: x slots
// and/or interleaved lane streams (both false: the reference bytes, each
// packet 16-byte aligned). Returns 0, or -1 with *err set when an x-slot
// coordinate exceeds 32 bits.
int to_device_layout(const Stream &st, bool slots, bool interleave, DeviceLayout *out, const char **err);

} // namespace spmvdiv

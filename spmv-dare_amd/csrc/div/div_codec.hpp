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
	long npackets() const { return (long) pk_off.size() - 1; }
	long ngroups() const { return (long) grp_pk.size() - 1; }
};

// Encode a CSR matrix. Returns 0, or -1 with *err set (row span > 255 or
// column span > 32 bits in a packet, which the reference rejects with error()).
int encode(const int32_t *row_ptr, const int32_t *col, const double *val, long m, long n, long nnz,
		int T, int L, long P, Stream *out, const char **err);

} // namespace spmvdiv

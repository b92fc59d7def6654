/*
 * spmvhip.h -- C ABI of the gfx950 (MI355X) FP64 SpMV layer.
 *
 * The reference's plugin contract is a C++ struct per format
 * (src/spmv_kernels/spmv_kernel.h:8-29):
 *
 *   struct Matrix_Format { char *format_name; long m, n, nnz;
 *       double mem_footprint, csr_mem_footprint;
 *       virtual void spmv(ValueType *x, ValueType *y) = 0;
 *       virtual void statistics_start() = 0;
 *       virtual int statistics_print_data(char *buf, long buf_n) = 0; };
 *   struct Matrix_Format *csr_to_format(INT_T *row_ptr, INT_T *col_ind,
 *       ValueTypeReference *values, long m, long n, long nnz,
 *       long symmetric, long symmetry_expanded);
 *   int statistics_print_labels(char *buf, long buf_n);
 *
 * Each entry point below is the plain-C image of one of those members, with
 * an explicit handle instead of `this`, the format chosen by a parameter block
 * instead of by which executable was linked, and a status code instead of
 * exit(). The C++ Matrix_Format subclasses (spmv-dare_amd/host/) are thin
 * wrappers over these calls.
 *
 * Types: INT_T = int32_t, ValueType = ValueTypeReference = double
 * (src/make.sh:127,134,145).
 *
 * Errors: 0 = success, negative SPMVHIP_E* otherwise; spmvhip_last_error()
 * returns the message of the calling thread's last failure.
 *
 * Threading: handles are independent (no file-scope state per format, unlike
 * the reference's thread_i_s / tds globals); one handle must not be used from
 * two host threads at once. Each handle owns one HIP stream of default
 * priority; all work of a handle is ordered on it.
 */
#ifndef SPMVHIP_H
#define SPMVHIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	SPMVHIP_OK = 0,
	SPMVHIP_EINVAL = -1,      /* invalid argument or unsupported parameter */
	SPMVHIP_ENOMEM = -2,      /* device allocation failed */
	SPMVHIP_EHIP = -3,        /* HIP runtime error */
	SPMVHIP_ENODEV = -4,      /* no usable gfx950 device */
	SPMVHIP_EFORMAT = -5,     /* input is not a valid CSR matrix */
	SPMVHIP_ESYMMETRIC = -6,  /* symmetric && !symmetry_expanded (reference: error()) */
};

/* Drop-in formats (the executables spmv_<fmt>_d.exe of config.sh:183-196). */
enum spmvhip_kind {
	SPMVHIP_CSR_NAIVE = 0,  /* src/spmv_kernels/csr.cpp: one thread per row, sequential, no FMA */
	SPMVHIP_CSR_VEC = 1,    /* src/spmv_kernels/csr_vec.cpp: row-length-binned wave64 CSR */
	SPMVHIP_SELL = 2,       /* src/spmv_kernels/sell.cpp: slices of C rows, no sort */
	SPMVHIP_SELL_C_S = 3,   /* src/spmv_kernels/sell_c_s.cpp: SELL-C-sigma */
	SPMVHIP_DIV_RF = 4,     /* src/spmv_kernels/DIV/div.cpp: DIV-RF packet stream */
	SPMVHIP_SELL_SORTED = 5, /* src/spmv_kernels/sell_sorted.cpp: rows stable-sorted by ascending
	                            length inside row windows, then C-row slices */
	SPMVHIP_NUM_KINDS = 6
};

/* Tuning / format parameters. Zero-initialise and set `kind`; zero fields
 * take the defaults noted. */
typedef struct spmvhip_params {
	int kind;             /* enum spmvhip_kind */
	int C;                /* SELL slice height; must be 64 (one wavefront) -- default 64 */
	long sigma;           /* SELL_C_S: sort window, 1..16384 -- default 4096. SELL_SORTED: the
	                         row window, a multiple of 64 up to 16384 -- default 4096 (the
	                         reference sorts inside each OpenMP thread's row range) */
	int lanes_per_row;    /* CSR_VEC: 0 = auto, else lanes of the LDS reduction per row (1..64);
	                         SELL*: lanes per row T (power of two, 0/1 = the reference layout) */
	int block_nnz;        /* CSR_VEC: nnz per row block staged in LDS (64..1024, default 256) */
	int split_nnz;        /* CSR_VEC: rows longer than this are split across waves (default 8192) */
	int xcd_remap;        /* CSR_VEC row blocks, DIV_RF packet groups (staged decoder): deal the work
	                         to the 8 XCDs in chunks of consecutive workgroups, so each XCD's L2
	                         serves one range of rows at a time. 1 = chunks of 64, n > 1 = chunks
	                         of n, 0 = in order (default) */
	int x_window;         /* SELL*: stage each workgroup's x window [min col, max col] of its
	                         256 rows in LDS when it spans <= this many doubles (1 = 2048;
	                         at most 16384); 0 = off */
	int unroll;           /* SELL*: inner unroll (1, 2, 4, 8; default 8); DIV_RF: values per lane
	                         and decode step (4, 8 when >= 8, 16 when >= 16; default 4) */
	int div_threads;      /* DIV_RF: encoder partitions T (default 1) */
	int div_lanes;        /* DIV_RF: lanes L of the value streams (default 64) */
	long div_packet_vals; /* DIV_RF: P, max values per packet (default 16384) */
	int pipeline;         /* SELL*: 1 = request the next entries' indices and values before the
	                         current x gathers (software pipelining; default 0) */
	int y_mode;           /* SELL*: how results reach y. 0 auto; 1 direct stores (y[row_order[i]]);
	                         2 staged in LDS per workgroup of 256 rows (needs sigma | 256 or no
	                         sort); 3 stored in sorted order to a device buffer, then one
	                         coalesced permutation pass y[r] = ys[inv[r]] (SELL_C_S) */
	int div_acc_copies;   /* DIV_RF: copies of the LDS row accumulators the lanes add into (a power
	                         of two, at most 64; 0 = up to 4 within 4 KiB for the largest packet) */
	int narrow_loads;     /* CSR_VEC: 1 = one stored entry per lane and load (4-byte column, 8-byte
	                         value loads); 0 (default) = two adjacent entries per lane and load
	                         (8-byte column, 16-byte value loads: a third fewer vector-memory
	                         instructions per nonzero) */
	int div_layout;       /* DIV_RF: device packet layout. 1 = the reference's packet bytes; 2 = x
	                         slots (each packet's distinct columns, ascending, are numbered,
	                         coordinates hold the slot, and the decoder stages x for every slot into
	                         LDS with loads along runs of consecutive columns); 3 = interleaved lane
	                         streams (the L value streams as rows of L dwords, so the lanes' stream
	                         reads fall in distinct LDS banks); 4 = both. 0 (default) = automatic:
	                         the fastest measured layout whose packet fits the LDS, else 1.
	                         Explicit 2..4 are an error when a packet does not fit */
	int div_reorder;      /* DIV_RF: 1 = run the packet groups of each partition in ascending order
	                         of their median column (the reference's DIV_PACKET_REORDERING=1,
	                         src/spmv_kernels/DIV/div.cpp:377-395,455-498); y is unchanged.
	                         0 (default) = in stream order */
	double div_tolerance; /* DIV_RF: lossy value compression (the reference's DIV_VC_TOLERANCE,
	                         div_kernels_rf.h:467-510): every value rounded to keep
	                         ceil(|log2 tolerance|) fraction bits, relative error <= tolerance
	                         (< 1); 0 (default) = lossless */
	int div_encoder;      /* DIV_RF: where the packets are encoded. 0 (default) = on the device when
	                         the packets hold at most 2048 values, the lanes are at most 64, the
	                         layout is 0, 1 or 3 and there is no packet reordering, else on the
	                         host; 1 = host; 2 = device (an error when not possible). The bytes
	                         are the same either way */
	int reserved[1];
} spmvhip_params;

typedef struct spmvhip_format spmvhip_format;   /* opaque */

typedef struct spmvhip_info {
	const char *format_name; /* Matrix_Format::format_name */
	long m, n, nnz;
	double mem_footprint;     /* bytes of the device representation */
	double csr_mem_footprint; /* nnz*(8+4) + (m+1)*4, as spmv_kernel.h:21-24 */
	double conv_ms;           /* device time of csr_to_format (after H2D) */
	long padded_nnz;          /* SELL*: stored entries incl. padding; else nnz */
	long num_work_items;      /* CSR_VEC: row blocks; SELL*: slices; DIV: packets */
	long num_split_rows;      /* CSR_VEC: rows split across waves */
} spmvhip_info;

/* ---- device / memory helpers (no PyTorch needed by callers) ---- */
int spmvhip_init(int device);
int spmvhip_device_name(char *buf, long buf_n);
int spmvhip_malloc(void **dptr, size_t bytes);
int spmvhip_free(void *dptr);
int spmvhip_memcpy_h2d(void *dst, const void *src, size_t bytes);
int spmvhip_memcpy_d2h(void *dst, const void *src, size_t bytes);
int spmvhip_memset(void *dptr, int value, size_t bytes);
int spmvhip_synchronize(void);
const char *spmvhip_last_error(void);

/* ---- the plugin contract ---- */

/* csr_to_format (spmv_kernel.h:28): copies the host CSR to the device and
 * converts it there. The handle keeps no host pointer. */
int spmvhip_csr_to_format(const int32_t *row_ptr, const int32_t *col_ind,
		const double *values, long m, long n, long nnz, long symmetric,
		long symmetry_expanded, const spmvhip_params *params,
		spmvhip_format **out);

/* Same, from CSR arrays already resident on the device (not retained). */
int spmvhip_csr_to_format_device(const int32_t *d_row_ptr, const int32_t *d_col_ind,
		const double *d_values, long m, long n, long nnz,
		const spmvhip_params *params, spmvhip_format **out);

/* Matrix_Format::spmv (spmv_kernel.h:17) with HOST x and y, as the reference
 * driver passes them: x (n doubles) is uploaded, y (m doubles) written back;
 * every y[i], i < m, is written (empty rows give 0) and nothing past m.
 * If x and y are the pointers bound with spmvhip_bind_host_xy, no copy is
 * made: the kernel runs on the bound device mirrors and y stays on the
 * device until spmvhip_sync_host_y. */
int spmvhip_spmv(spmvhip_format *f, const double *x, double *y);

/* y = A x with x, y in device memory; enqueued on the handle's stream
 * (stream == NULL) or on `stream` (a hipStream_t). Asynchronous. */
int spmvhip_spmv_device(spmvhip_format *f, const double *d_x, double *d_y, void *stream);

/* Device-resident mode of the driver (GPU_KERNEL=1): bind host x/y once. */
int spmvhip_bind_host_xy(spmvhip_format *f, const double *x, double *y);
int spmvhip_sync_host_y(spmvhip_format *f);

/* Matrix_Format::statistics_start / statistics_print_data and
 * statistics_print_labels (spmv_kernel.h:18-19,29): ','-prefixed CSV
 * columns gbps_alg,pct_peak,kernel_ms,conv_ms,variant. */
int spmvhip_statistics_start(spmvhip_format *f);
int spmvhip_statistics_print_data(spmvhip_format *f, char *buf, long buf_n);
int spmvhip_statistics_print_labels(char *buf, long buf_n);

int spmvhip_get_info(const spmvhip_format *f, spmvhip_info *info);

/* Time `iters` back-to-back spmv_device calls after `warmup` untimed ones,
 * with HIP events on the handle's stream; ms[i] receives each iteration's
 * time (ms may be NULL, then only *total_ms). */
int spmvhip_time_spmv(spmvhip_format *f, const double *d_x, double *d_y,
		int warmup, int iters, float *ms, float *total_ms);

/* Converted arrays, copied to host for parity checks.
 * which: 0 row_order (int32, m)      [SELL_C_S, SELL_SORTED]
 *        1 slice_widths (int32, nslices)   [SELL, SELL_C_S]
 *        2 slice_ptr (int64, nslices+1)    [SELL, SELL_C_S]
 *        3 cols (int32, padded_nnz)        [SELL, SELL_C_S]
 *        4 vals (float64, padded_nnz)      [SELL, SELL_C_S]
 *        5 row blocks (int32, work_items+1) [CSR_VEC]
 *        6 DIV decoded rows (int32, nnz), 7 DIV decoded cols (int32, nnz),
 *        8 DIV decoded values (float64, nnz) -- device decode, packet order [DIV_RF]
 *        9 x windows (int32 pairs lo, width per workgroup of 4 slices; width 0 =
 *          not staged)                     [SELL, SELL_C_S with x_window]
 *        10 packet bytes in the device layout (uint8, the packets without the
 *          trailing padding), 11 their offsets (int64, npackets + 1) [DIV_RF] */
/* *bytes_inout: capacity in, bytes written out (or needed, with buf NULL). */
int spmvhip_export(const spmvhip_format *f, int which, void *buf, size_t *bytes_inout);

void spmvhip_destroy(spmvhip_format *f);

/* ---- batched SpMV (no counterpart in the reference, whose driver runs one
 * matrix per process, run.sh:116-151) ----
 * Many converted matrices, each with its own device x and y, computed by ONE
 * persistent kernel launch: the resident wavefronts draw the units of all
 * members (CSR row blocks and long-row segments, SELL slices) in list order
 * by atomic ticket, unit g from counter g % nqueues, each wave requesting its
 * next ticket before it runs the current unit; long-row segments are
 * completed by a second kernel of the same sequence. With use_graph the
 * sequence (counter reset + kernel + segment finalize) is captured once into
 * a hipGraph and replayed. DIV_RF members (LDS-staged decoder, div_layout 1 or
 * 3) are decoded by a second persistent kernel of the same sequence that draws
 * all DIV members' packet groups by ticket in the same way. Members must be
 * CSR_VEC (block_nnz <= 512), SELL, SELL_C_S or DIV_RF handles, must outlive
 * the batch and must not run elsewhere meanwhile. tile_units: consecutive
 * units per ticket (0 = 1, at most 65536). */
typedef struct spmvhip_batch spmvhip_batch;
int spmvhip_batch_create(int tile_units, spmvhip_batch **out);
/* Ticket counters (before finalize): a power of two, 1..1024 (default 256),
 * each on its own 256-byte line; reduced at finalize to at most the number
 * of resident wavefronts. */
int spmvhip_batch_set_queues(spmvhip_batch *b, int nqueues);
/* Persistent grid (before finalize): workgroups of 4 wavefronts per CU, 1..8,
 * at most what the kernel's resources allow; 0 (default) = the measured best. */
int spmvhip_batch_set_grid(spmvhip_batch *b, int workgroups_per_cu);
/* XCD ranges (before finalize): 1 = the unit list is cut into 8 contiguous
 * ranges of equal nonzero weight, XCD q's workgroups draw only from range q
 * (queues / 8 counters each), so a member's x is fetched into one or two
 * XCDs' L2; 0 = every wave draws from the whole list; -1 (default) =
 * automatic: ranges when no member holds more than 1/64 of the units. */
int spmvhip_batch_set_xcd_ranges(spmvhip_batch *b, int on);
/* Large members (before any add): a member of at least nnz nonzeros (0, the
 * default: none) runs as its own kernels -- the ones spmvhip_spmv_device
 * runs -- at the start of the same launch sequence (and hipGraph), ahead of
 * the persistent kernels, instead of as units in them: such a matrix fills
 * the GPU alone, and its tuned standalone kernels beat the shared unit loop. */
int spmvhip_batch_set_solo_nnz(spmvhip_batch *b, long nnz);
int spmvhip_batch_add(spmvhip_batch *b, const spmvhip_format *f, const double *d_x, double *d_y);
int spmvhip_batch_finalize(spmvhip_batch *b, int use_graph);
int spmvhip_batch_run(spmvhip_batch *b);          /* asynchronous on the device's shared stream (the
                                                     one the format handles use) */
int spmvhip_batch_synchronize(spmvhip_batch *b);
/* ms[i]: GPU time of iteration i of the batched launch (per_matrix = 0) or
 * of one standalone spmv launch per member, back to back (per_matrix = 1).
 * After `warmup` untimed iterations, the `iters` timed ones are submitted
 * back to back with an event between consecutive iterations and one wait at
 * the end, so the host's submission overlaps the GPU's work. */
int spmvhip_batch_time(spmvhip_batch *b, int per_matrix, int warmup, int iters, float *ms);
/* nitems: units of the batch (row blocks, segments, slices); grid: workgroups of the launch. */
int spmvhip_batch_info(const spmvhip_batch *b, long *nmats, long *nitems, int *grid, long *total_nnz);
void spmvhip_batch_destroy(spmvhip_batch *b);

/* Bandwidth reference (no counterpart in the reference): mode 0 streams n
 * doubles of src once (16-byte loads), mode 1 copies src to dst, mode 2
 * streams src through direct global-to-LDS loads; ms[i] is
 * iteration i's time. The measured ceiling for the SpMV kernels and the
 * known-size traffic that calibrates the profiler's byte counters. */
int spmvhip_bandwidth_probe(int mode, const double *src, double *dst, long n, int iters, float *ms);
/* The vector-memory cost of one 64-lane load, in cycles per CU at clock_ghz,
 * by pattern (mode 0 contiguous 4-byte, 1 contiguous 8-byte, 2 contiguous
 * 16-byte, 3 8-byte gather over 64 cache lines, 4 the same with half the lanes
 * active, 5 16-byte gather over 64 lines, 6 8-byte gather over 16 lines, 7
 * 8-byte gather over 4 lines, 8..13 mode 3 as buffer loads with the cache
 * policy none, sc0, nt, sc1, sc0 sc1, sc0 sc1 nt; n <= 2^27 for those) on the
 * first n doubles of buf (n a power of two, >= 1024: small n stays in the L1,
 * large in the L2). sink: >= 32 x CUs
 * doubles. No reference counterpart (a measurement aid). */
int spmvhip_gather_probe(int mode, const double *buf, long n, int iters, double clock_ghz, double *sink,
		double *cycles_per_load);

/* ---- format selector support (no counterpart in the reference) ---- */

/* Structure statistics of a device CSR matrix, the selector's cheap
 * features, with the definitions of the reference generator's descriptor
 * statistics (artificial_matrix_generation.c:503-539, matrix_util.h:235-266):
 * out6 = {mean row length, std of row lengths, longest row, mean bandwidth / n,
 * neighbours per nnz (window 1), non-empty rows}. */
int spmvhip_csr_stats_device(const int32_t *d_row_ptr, const int32_t *d_col_ind, long m, long n,
		double *out6);

/* The same statistics with no allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation). No counterpart in the reference. Enqueued on
 * `stream` (a hipStream_t, NULL = the default stream) and asynchronous: the
 * six doubles land in DEVICE memory d_out6. The per-workgroup partials live
 * in a file-scope device array of 768 KiB that every GPU the library runs on
 * holds until the process ends; it limits m to 67,108,864 rows (SPMVHIP_EINVAL
 * beyond). Calls that overlap on one GPU share that array and race. */
int spmvhip_csr_stats_device_noalloc(const int32_t *d_row_ptr, const int32_t *d_col_ind, long m, long n,
		double *d_out6, void *stream);

/* FP32 MLP n_in -> h1 -> h2 -> n_out (ReLU hidden, linear out) with
 * hand-written forward, backward (masked mean squared error) and Adam
 * kernels. Parameters: one flat array, per layer W[n_in][n_out] then b. */
typedef struct spmvhip_mlp spmvhip_mlp;
int spmvhip_mlp_create(int n_in, int h1, int h2, int n_out, unsigned int seed, spmvhip_mlp **out);
long spmvhip_mlp_num_params(const spmvhip_mlp *net);
int spmvhip_mlp_set_params(spmvhip_mlp *net, const float *host, long count);
int spmvhip_mlp_get_params(const spmvhip_mlp *net, float *host, long count);
int spmvhip_mlp_predict(spmvhip_mlp *net, const float *X, int batch, float *Y);
int spmvhip_mlp_gradients(spmvhip_mlp *net, const float *X, const float *T, const float *M, int batch,
		float *grads, float *loss);
int spmvhip_mlp_train(spmvhip_mlp *net, const float *X, const float *T, const float *M, long rows,
		int batch, int epochs, float lr, float *epoch_losses);
int spmvhip_mlp_adam_step(spmvhip_mlp *net, const float *grads, float lr);
/* Adam state for checkpoint / resume: the first and second moment estimates
 * (count = spmvhip_mlp_num_params floats each) and the step counter.
 * set_params resets it; set_adam_state after set_params restores it. */
int spmvhip_mlp_get_adam_state(const spmvhip_mlp *net, float *m1, float *m2, long count, long *step);
int spmvhip_mlp_set_adam_state(spmvhip_mlp *net, const float *m1, const float *m2, long count, long step);
void spmvhip_mlp_destroy(spmvhip_mlp *net);

/* The host half of csr_to_format for DIV_RF, exposed so that the encoder can
 * be checked without a GPU: encodes the CSR matrix into the reference's
 * packet layout with T nnz-balanced partitions, L value lanes and packets
 * of at most P values (src/spmv_kernels/DIV/div.cpp:236-513,
 * div_kernels_rf.h:538-919). Two calls: with bytes == NULL only the sizes
 * are returned in *nbytes, *npackets and *ngroups; then the caller passes
 * buffers of those sizes: bytes[*nbytes], pk_off[*npackets + 1] (byte offset
 * of each packet), grp_pk[*ngroups + 1] (first packet of each packet group).
 * Needs no device. */
int spmvhip_div_encode(const int32_t *row_ptr, const int32_t *col_ind, const double *values,
		long m, long n, long nnz, int T, int L, long P, uint8_t *bytes, size_t *nbytes,
		int64_t *pk_off, long *npackets, int32_t *grp_pk, long *ngroups);

/* The packet-group order of the reference's packet reordering
 * (DIV_PACKET_REORDERING=1, src/spmv_kernels/DIV/div.cpp:377-395,455-498)
 * for the stream spmvhip_div_encode writes: order[k] = the k-th group to
 * run; inside each partition, ascending median column of the partition's
 * k-th packet, ties in stream order. With order == NULL only *ngroups is
 * returned. Needs no device. */
int spmvhip_div_group_order(const int32_t *row_ptr, const int32_t *col_ind, const double *values,
		long m, long n, long nnz, int T, int L, long P, int32_t *order, long *ngroups);

/* The value rounding of DIV_RF's lossy compression (DIV_VC_TOLERANCE,
 * src/spmv_kernels/DIV/div_kernels_rf.h:467-510, applied to each value at
 * :612-616): out[k] = in[k] with ceil(|log2 tolerance|) fraction bits kept,
 * truncated or rounded up in its bit pattern, whichever leaves more trailing
 * zero bits (rounded up on a tie). tolerance <= 0 copies; tolerance >= 1 or a
 * rounding beyond it is SPMVHIP_EINVAL (the reference's error()). in == out
 * is allowed. Needs no device. */
int spmvhip_div_reduce_precision(const double *in, long n, double tolerance, double *out);

/* The device form of the same stream in device layout `layout` (the
 * div_layout codes 1..4 of spmvhip_params: 1 the reference bytes, 2 x slots,
 * 3 interleaved lane streams, 4 both), as spmv-dare_amd/csrc/div/div_codec.hpp
 * describes, each packet at a 16-byte aligned offset. Two calls as
 * spmvhip_div_encode: bytes[*nbytes], pk_off[*npackets + 1]. No reference
 * counterpart (the reference's DIV decoder reads its own packets); needs no
 * device. */
int spmvhip_div_device_encode(const int32_t *row_ptr, const int32_t *col_ind, const double *values,
		long m, long n, long nnz, int T, int L, long P, int layout, uint8_t *bytes, size_t *nbytes,
		int64_t *pk_off, long *npackets);

#ifdef __cplusplus
}
#endif

#endif /* SPMVHIP_H */

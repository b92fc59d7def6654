// DIV-RF packet encoder (layout and reference citations: div_codec.hpp).
// Packet boundaries are found sequentially per partition (they chain), then
// the packets are encoded in parallel and concatenated in partition order.

#include "div_codec.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <omp.h>

namespace spmvdiv {
namespace {

// Number of bits of the range [0, v] (0 for v == 0), lib/bit_ops.h:225-243.
inline uint64_t req_bits(uint64_t v)
{
	return v == 0 ? 0 : 64 - __builtin_clzll(v);
}

// Closest prefix-sum index, ties to the later index (lib/macros/macrolib.h:537-575).
long closest(const int32_t *A, long lo, long hi, int64_t target)
{
	long s = lo, e = hi;
	if (target < A[s])
		return s;
	if (target > A[e])
		return e;
	for (;;) {
		long mid = (s + e) / 2;
		if (mid == s || mid == e)
			break;
		if (target > A[mid])
			s = mid;
		else
			e = mid;
	}
	if (A[s] == target)
		return s;
	if (A[e] == target)
		return e;
	return (target - A[s]) < (A[e] - target) ? s : e;
}

// nnz-balanced row range of partition t of T (lib/parallel_util.h:156-184 on row_ptr).
void partition(long T, long t, const int32_t *rp, long m, long nnz, long *s, long *e)
{
	int64_t target = rp[0] + (nnz * t) / T;
	int64_t target_next = rp[0] + (nnz * (t + 1)) / T;
	*s = (t == 0) ? 0 : closest(rp, 0, m - 1, target);
	*e = (t == T - 1) ? m : closest(rp, 0, m - 1, target_next);
}

// Packet size with a complete last row where possible, at most 255 rows
// (div_kernels_rf.h:437-464 with force_row_index_lte_1_byte).
long select_num_vals(const int32_t *rp, long i_s, long j_s, long num_vals)
{
	long j_e = j_s + num_vals, i_e = i_s;
	while (rp[i_e] < j_e)
		i_e++;
	if (j_e != rp[i_e] && i_e - 1 > i_s) {
		i_e--;
		j_e = rp[i_e];
	}
	if (i_e - i_s > 255)
		i_e = i_s + 255;
	long r = rp[i_e] - j_s;
	return r > num_vals ? num_vals : r;
}

// Interleave N elements (units of U) into S segments: element-wise for the
// full rounds, unit-wise for the remainder (lib/macros/permutation.h:47-111
// with solid_units_main = 0, solid_units_rem = 1).
template <class T>
void interleave(const T *in, T *out, long N, long S, long U)
{
	long units = N / U, rem = N - units * U;
	long div = units / S, mod = units % S;
	long mul = div * U * S;
	std::vector<long> seg(S);
	for (long s = 0; s < S; s++) {
		seg[s] = U * (s * div + (s < mod ? s : mod));
		if (s > mod)
			seg[s] += rem;
	}
	for (long j = 0; j < mul; j++)
		out[j] = in[seg[j % S] + j / S];
	for (long j = mul; j < N; j++)
		out[j] = in[seg[(j / U) % S] + U * (j / (U * S)) + j % U];
}

struct Key {
	uint64_t v;
	uint32_t row, col;
	bool operator<(const Key &o) const
	{
		if (v != o.v)
			return v < o.v;
		if (row != o.row)
			return row < o.row;
		return col < o.col;
	}
};

void put32(std::vector<uint8_t> &b, size_t at, uint32_t v)
{
	memcpy(&b[at], &v, 4);
}

// One packet (div_kernels_rf.h:538-919, lossless: DIV_VC_TOLERANCE = 0).
int compress(const int32_t *rp, const int32_t *col, const double *val, long i0, long j0, long nv, int L,
		std::vector<uint8_t> &out, const char **err)
{
	std::vector<Key> key(nv);
	long i = i0;
	uint32_t cmin = (uint32_t) col[j0], cmax = cmin;
	for (long k = 0; k < nv; k++) {
		long j = j0 + k;
		while (j >= rp[i + 1])
			i++;
		uint64_t u;
		memcpy(&u, &val[j], 8);
		uint32_t c = (uint32_t) col[j];
		cmin = c < cmin ? c : cmin;
		cmax = c > cmax ? c : cmax;
		key[k] = {u, (uint32_t) i, c};
	}
	uint64_t row_bits = req_bits((uint64_t) (i - i0));
	if (row_bits > 8) {
		*err = "row_bits > 8";
		return -1;
	}
	uint64_t col_bits = req_bits(cmax - cmin);
	uint64_t cb = (row_bits + col_bits + 7) >> 3;
	if (row_bits % 8) {
		uint64_t cmpl = 8 - (row_bits % 8);
		if ((((row_bits + cmpl) + col_bits + 7) >> 3) == cb)
			row_bits += cmpl;
	}
	col_bits = (cb << 3) - row_bits;
	if (col_bits > 32) {
		*err = "col_bits > 32";
		return -1;
	}
	const uint64_t num_rows = (uint64_t) (i - i0) + 1;

	std::sort(key.begin(), key.end());

	// replication factor of every element: length of its run of equal value bits
	std::vector<int32_t> rf(nv);
	long max_rf = 1, nuniq = 0;
	for (long a = 0; a < nv;) {
		long b = a + 1;
		while (b < nv && key[b].v == key[a].v)
			b++;
		for (long t = a; t < b; t++)
			rf[t] = (int32_t) (b - a);
		max_rf = std::max(max_rf, b - a);
		nuniq++;
		a = b;
	}
	// stable counting sort by replication factor (bucketsort_gen.c:164-193)
	std::vector<long> start(max_rf + 2, 0);
	for (long k = 0; k < nv; k++)
		start[rf[k] + 1]++;
	for (long b = 1; b < max_rf + 2; b++)
		start[b] += start[b - 1];
	std::vector<long> fill(start.begin(), start.end() - 1);
	std::vector<Key> sk(nv);
	for (long k = 0; k < nv; k++)
		sk[fill[rf[k]]++] = key[k];

	// unique values in that order, then per replication factor: interleave
	// the values and (in units of rf) their coordinates over the L lanes
	std::vector<uint64_t> uniq(nuniq), uniq_il(nuniq);
	std::vector<Key> sk_il(nv);
	std::vector<uint32_t> rfs, per_rf;
	long u0 = 0;
	for (long r = 1; r <= max_rf; r++) {
		long ls = start[r], le = start[r + 1];
		if (le == ls)
			continue;
		long n_unique = (le - ls) / r;
		for (long t = 0; t < n_unique; t++) // bucket r holds runs of r equal values
			uniq[u0 + t] = sk[ls + t * r].v;
		rfs.push_back((uint32_t) r);
		per_rf.push_back((uint32_t) n_unique);
		interleave(&uniq[u0], &uniq_il[u0], n_unique, L, 1);
		interleave(&sk[ls], &sk_il[ls], le - ls, L, r);
		u0 += n_unique;
	}
	const long num_rfs = (long) rfs.size();

	// lane streams: per rf group the lane's previous value restarts at 0
	std::vector<std::vector<uint8_t>> lanes(L);
	std::vector<uint8_t> lens(nuniq);
	u0 = 0;
	for (long g = 0; g < num_rfs; g++) {
		std::vector<uint64_t> prev(L, 0);
		for (long q = 0; q < (long) per_rf[g]; q++) {
			long k = u0 + q;
			int lane = (int) (q % L);
			uint64_t diff = uniq_il[k] - prev[lane];
			prev[lane] += diff;
			if (diff == 0) {
				lens[k] = 0;
				continue;
			}
			uint64_t lz = __builtin_clzll(diff), tz = __builtin_ctzll(diff);
			uint64_t rem_tz = tz & 3;
			uint64_t len_bits = 64 - lz - tz;
			uint64_t len = (len_bits + 7) >> 3;
			uint64_t rem = (len << 3) - len_bits;
			if (rem >= rem_tz) {
				rem -= rem_tz;
				tz -= rem_tz;
			}
			if (rem <= lz) {
				lz -= rem;
				rem = 0;
			} else {
				rem -= lz;
				lz = 0;
			}
			tz -= rem;
			uint64_t tz4 = tz >> 2;
			tz = tz4 << 2;
			len_bits = 64 - lz - tz;
			len = (len_bits + 7) >> 3;
			uint64_t payload = diff >> tz;
			for (uint64_t b = 0; b < len; b++)
				lanes[lane].push_back((uint8_t) (payload >> (8 * b)));
			lens[k] = (uint8_t) (len | (tz4 << 4));
		}
		u0 += per_rf[g];
	}

	// assemble
	size_t hb = (size_t) header_bytes(L);
	size_t total = hb + 8 * (size_t) num_rfs + (size_t) nv * cb + (size_t) nuniq;
	for (int l = 0; l < L; l++)
		total += lanes[l].size();
	out.assign(total, 0);
	put32(out, 0, (uint32_t) nv);
	put32(out, 4, (uint32_t) nuniq);
	put32(out, 8, (uint32_t) num_rows);
	put32(out, 12, (uint32_t) i0);
	put32(out, 16, cmin);
	out[20] = (uint8_t) row_bits;
	out[21] = (uint8_t) col_bits;
	put32(out, 24, (uint32_t) num_rfs);
	for (int l = 0; l < L; l++)
		put32(out, kHeaderFixed + 4 * l, (uint32_t) lanes[l].size());
	size_t at = hb;
	for (long g = 0; g < num_rfs; g++)
		put32(out, at + 4 * g, rfs[g]);
	at += 4 * num_rfs;
	for (long g = 0; g < num_rfs; g++)
		put32(out, at + 4 * g, per_rf[g]);
	at += 4 * num_rfs;
	for (long k = 0; k < nv; k++) {
		uint64_t c = (uint64_t) (sk_il[k].row - (uint32_t) i0) | ((uint64_t) (sk_il[k].col - cmin) << row_bits);
		for (uint64_t b = 0; b < cb; b++)
			out[at + k * cb + b] = (uint8_t) (c >> (8 * b));
	}
	at += nv * cb;
	memcpy(&out[at], lens.data(), nuniq);
	at += nuniq;
	for (int l = 0; l < L; l++) {
		if (!lanes[l].empty())
			memcpy(&out[at], lanes[l].data(), lanes[l].size());
		at += lanes[l].size();
	}
	return 0;
}

} // namespace

int reduce_precision(const double *in, long n, double tolerance, double *out, const char **err)
{
	*err = nullptr;
	if (tolerance >= 1.0) {
		*err = "DIV_VC_TOLERANCE must be < 1";
		return -1;
	}
	if (!(tolerance > 0)) {
		if (out != in && n > 0)
			memcpy(out, in, (size_t) n * sizeof(double));
		return 0;
	}
	const long keep = (long) ceil(fabs(log2(tolerance)));
	int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
	for (long k = 0; k < n; k++) {
		uint64_t u;
		memcpy(&u, &in[k], 8);
		double r = in[k];
		if (u != 0 && __builtin_ctzll(u) < 53 && keep < 52) {
			const uint64_t low = (1ull << (52 - keep)) - 1;
			const uint64_t dn = u & ~low, up = (u + low) & ~low;
			double vd, vu;
			memcpy(&vd, &dn, 8);
			memcpy(&vu, &up, 8);
			if (fabs(in[k] - vd) / fabs(in[k]) > tolerance || fabs(in[k] - vu) / fabs(in[k]) > tolerance)
				bad = 1;
			const int tz_dn = dn ? __builtin_ctzll(dn) : 64, tz_up = up ? __builtin_ctzll(up) : 64;
			r = tz_dn > tz_up ? vd : vu;
		}
		out[k] = r;
	}
	if (bad) {
		*err = "a rounded value exceeds DIV_VC_TOLERANCE";
		return -1;
	}
	return 0;
}

void plan_packets(const int32_t *rp, long m, long nnz, int T, long P, PacketPlan *out)
{
	out->i0.clear();
	out->j0.clear();
	out->nv.clear();
	out->grp_pk.clear();
	out->grp_row.clear();
	out->part_grp.clear();
	std::vector<long> grp_first_row;
	for (long t = 0; t < T; t++) {
		out->part_grp.push_back((int32_t) grp_first_row.size());
		if (m == 0)
			continue;
		long i_s, i_e;
		partition(T, t, rp, m, nnz, &i_s, &i_e);
		long j = rp[i_s], j_e = rp[i_e];
		long i = i_s;
		bool first = true;
		while (j < j_e) {
			while (rp[i] < j)
				i++;
			if (rp[i] != j)
				i--;
			while (rp[i] == rp[i + 1]) // empty rows are skipped
				i++;
			long nmax = (j + P <= j_e) ? P : j_e - j;
			long nv = select_num_vals(rp, i, j, nmax);
			if (j == rp[i]) { // a packet that starts a row starts a packet group
				out->grp_pk.push_back((int32_t) out->i0.size());
				grp_first_row.push_back(first ? i_s : i);
			}
			out->i0.push_back((int32_t) i);
			out->j0.push_back(j);
			out->nv.push_back((int32_t) nv);
			first = false;
			j += nv;
		}
	}
	out->grp_pk.push_back((int32_t) out->i0.size());
	for (size_t g = 0; g < grp_first_row.size(); g++)
		out->grp_row.push_back(g == 0 ? 0 : (int32_t) grp_first_row[g]);
	out->grp_row.push_back((int32_t) m);
	out->part_grp.push_back((int32_t) grp_first_row.size());
}

int encode(const int32_t *rp, const int32_t *col, const double *val, long m, long n, long nnz, int T, int L,
		long P, Stream *out, const char **err)
{
	(void) n;
	*err = nullptr;
	out->T = T;
	out->L = L;
	out->P = P;
	PacketPlan plan;
	plan_packets(rp, m, nnz, T, P, &plan);
	const long np = (long) plan.i0.size();
	std::vector<std::vector<uint8_t>> pk(np);
	int fail = 0;
	const char *ferr = nullptr;
#pragma omp parallel for schedule(dynamic, 8)
	for (long p = 0; p < np; p++) {
		const char *e = nullptr;
		if (compress(rp, col, val, plan.i0[p], plan.j0[p], plan.nv[p], L, pk[p], &e)) {
#pragma omp critical
			{
				fail = 1;
				ferr = e;
			}
		}
	}
	if (fail) {
		*err = ferr;
		return -1;
	}
	size_t total = 0;
	for (long p = 0; p < np; p++)
		total += pk[p].size();
	out->bytes.resize(total);
	out->pk_off.assign(1, 0);
	out->pk_nnz.assign(plan.j0.begin(), plan.j0.end());
	out->pk_nnz.push_back(nnz);
	size_t at = 0;
	for (long p = 0; p < np; p++) {
		if (!pk[p].empty())
			memcpy(&out->bytes[at], pk[p].data(), pk[p].size());
		at += pk[p].size();
		out->pk_off.push_back((int64_t) at);
	}
	out->grp_pk.swap(plan.grp_pk);
	out->grp_row.swap(plan.grp_row);
	out->part_grp.swap(plan.part_grp);
	return 0;
}

std::vector<int32_t> group_order_by_median(const Stream &st, const int32_t *col)
{
	const long ng = st.ngroups();
	std::vector<int32_t> order(ng);
	std::vector<double> v;
	for (size_t t = 0; t + 1 < st.part_grp.size(); t++) {
		const int32_t g0 = st.part_grp[t], g1 = st.part_grp[t + 1];
		if (g0 == g1)
			continue;
		const int32_t p0 = st.grp_pk[g0];
		std::vector<long> key(g1 - g0);
		for (int32_t k = 0; k < g1 - g0; k++) {
			const int32_t p = p0 + k; // the partition's k-th packet
			const int64_t j0 = st.pk_nnz[p], N = st.pk_nnz[p + 1] - j0;
			// median_unbiased: virtual index (N - 1) q + q / 3 + 1 / 3 = N / 2 at q = 1/2
			const double virt = (double) (N - 1) * 0.5 + (1.0 / 3 * 0.5 + 1.0 / 3);
			const long idx = (long) virt;
			const double frac = virt - (double) idx;
			v.assign(col + j0, col + j0 + N);
			double med;
			if (idx >= N - 1) {
				med = *std::max_element(v.begin(), v.end());
			} else {
				std::nth_element(v.begin(), v.begin() + idx, v.end());
				const double lo = v[idx];
				const double hi = *std::min_element(v.begin() + idx + 1, v.end());
				med = (1 - frac) * lo + frac * hi;
			}
			key[k] = (long) med;
		}
		std::vector<int32_t> ids(g1 - g0);
		for (int32_t k = 0; k < g1 - g0; k++)
			ids[k] = k;
		std::stable_sort(ids.begin(), ids.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
		for (int32_t k = 0; k < g1 - g0; k++)
			order[g0 + k] = g0 + ids[k];
	}
	return order;
}

namespace {

uint32_t get32(const uint8_t *p)
{
	uint32_t v;
	memcpy(&v, p, 4);
	return v;
}

uint64_t get_le(const uint8_t *p, int nbytes)
{
	uint64_t v = 0;
	for (int b = 0; b < nbytes; b++)
		v |= (uint64_t) p[b] << (8 * b);
	return v;
}

struct SlotSizes {
	size_t runs_at, bits_at, coords_at, lens_at, streams_at;
};

SlotSizes slot_sizes(int L, uint32_t nv, uint32_t nu, uint32_t nrfs, uint32_t nruns, uint32_t nslots, int cb)
{
	SlotSizes z;
	z.runs_at = (size_t) slot_header_bytes(L) + 8 * (size_t) nrfs;
	z.bits_at = (z.runs_at + 4 * (size_t) nruns + 7) / 8 * 8;
	z.coords_at = z.bits_at + 8 * (((size_t) nslots + 63) / 64);
	z.lens_at = z.coords_at + (size_t) nv * cb;
	z.streams_at = z.lens_at + nu;
	return z;
}

// The lane streams of a packet (lane sizes from its header), contiguous or
// interleaved by dwords, appended to `o` from byte `at`; returns the end.
size_t put_streams(std::vector<uint8_t> &o, size_t at, const uint8_t *lane_sizes, const uint8_t *streams, int L,
		bool interleave)
{
	size_t total = 0, maxw = 0;
	for (int l = 0; l < L; l++) {
		const size_t sz = get32(lane_sizes + 4 * l);
		total += sz;
		maxw = std::max(maxw, (sz + 3) / 4);
	}
	if (interleave)
		at = (at + 3) & ~(size_t) 3; // dword-aligned rows
	const size_t end = at + (interleave ? maxw * 4 * (size_t) L : total);
	o.resize((end + 15) / 16 * 16, 0);
	if (!interleave) {
		memcpy(&o[at], streams, total);
		return end;
	}
	size_t base = 0;
	for (int l = 0; l < L; l++) {
		const size_t sz = get32(lane_sizes + 4 * l);
		for (size_t b = 0; b < sz; b++)
			o[at + (b >> 2) * 4 * (size_t) L + 4 * (size_t) l + (b & 3)] = streams[base + b];
		base += sz;
	}
	return end;
}

} // namespace

int to_device_layout(const Stream &st, bool slots, bool interleave, DeviceLayout *out, const char **err)
{
	const long np = st.npackets();
	const int L = st.L;
	std::vector<std::vector<uint8_t>> pk(np);
	int fail = 0;
	uint32_t max_slots = 0, max_rows = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : max_slots, max_rows)
	for (long p = 0; p < np; p++) {
		const uint8_t *src = &st.bytes[st.pk_off[p]];
		const uint32_t nv = get32(src), nu = get32(src + 4), nrows = get32(src + 8), row_min = get32(src + 12);
		const uint32_t col_min = get32(src + 16);
		const int row_bits = src[20], col_bits = src[21];
		const uint32_t nrfs = get32(src + 24);
		const int cb = (row_bits + col_bits + 7) >> 3;
		const size_t hb = (size_t) header_bytes(L);
		const uint8_t *coords = src + hb + 8 * (size_t) nrfs;
		const uint8_t *lens = coords + (size_t) nv * cb;
		std::vector<uint8_t> &o = pk[p];
		max_rows = std::max(max_rows, nrows);
		if (!slots) {
			const size_t streams_at = (size_t) (lens - src) + nu;
			o.assign(streams_at, 0);
			memcpy(o.data(), src, streams_at);
			o[22] = interleave ? 1 : 0;
			put_streams(o, streams_at, src + kHeaderFixed, src + streams_at, L, interleave);
			continue;
		}
		const uint64_t rmask = row_bits ? (1ull << row_bits) - 1 : 0;
		std::vector<uint32_t> rows(nv), cols(nv);
		for (uint32_t k = 0; k < nv; k++) {
			const uint64_t c = get_le(coords + (size_t) k * cb, cb);
			rows[k] = (uint32_t) (c & rmask);
			cols[k] = col_min + (uint32_t) (c >> row_bits);
		}
		std::vector<uint32_t> ucol(cols);
		std::sort(ucol.begin(), ucol.end());
		ucol.erase(std::unique(ucol.begin(), ucol.end()), ucol.end());
		const uint32_t nslots = (uint32_t) ucol.size();
		std::vector<uint32_t> run_col;
		std::vector<uint64_t> bits((nslots + 63) / 64, 0);
		for (uint32_t s = 0; s < nslots; s++) {
			if (s == 0 || ucol[s] != ucol[s - 1] + 1) {
				run_col.push_back(ucol[s]);
				bits[s >> 6] |= 1ull << (s & 63);
			}
		}
		const int rb = (int) req_bits(nrows > 0 ? nrows - 1 : 0);
		const int sb = (int) req_bits(nslots > 0 ? nslots - 1 : 0);
		if (rb + sb > 32) {
#pragma omp critical
			{
				fail = 1;
				*err = "x-slot coordinate above 32 bits";
			}
			continue;
		}
		const int ncb = std::max(1, (rb + sb + 7) >> 3);
		const SlotSizes z = slot_sizes(L, nv, nu, nrfs, (uint32_t) run_col.size(), nslots, ncb);
		o.assign(z.streams_at, 0);
		put32(o, 0, nv);
		put32(o, 4, nu);
		put32(o, 8, nrows);
		put32(o, 12, row_min);
		put32(o, 16, nslots);
		o[20] = (uint8_t) rb;
		o[21] = (uint8_t) sb;
		o[22] = (uint8_t) ncb;
		o[23] = interleave ? 1 : 0;
		put32(o, 24, nrfs);
		put32(o, 28, (uint32_t) run_col.size());
		memcpy(&o[kSlotHeaderFixed], src + kHeaderFixed, 4 * (size_t) L);            // lane sizes
		memcpy(&o[slot_header_bytes(L)], src + hb, 8 * (size_t) nrfs);                // rfs, unique_per_rf
		for (size_t r = 0; r < run_col.size(); r++)
			put32(o, z.runs_at + 4 * r, run_col[r]);
		if (!bits.empty())
			memcpy(&o[z.bits_at], bits.data(), 8 * bits.size());
		for (uint32_t k = 0; k < nv; k++) {
			const uint32_t slot = (uint32_t) (std::lower_bound(ucol.begin(), ucol.end(), cols[k]) - ucol.begin());
			const uint64_t c = (uint64_t) rows[k] | ((uint64_t) slot << rb);
			for (int b = 0; b < ncb; b++)
				o[z.coords_at + (size_t) k * ncb + b] = (uint8_t) (c >> (8 * b));
		}
		memcpy(&o[z.lens_at], lens, nu); // val_lens, unchanged
		put_streams(o, z.streams_at, src + kHeaderFixed, lens + nu, L, interleave);
		max_slots = std::max(max_slots, nslots);
	}
	if (fail)
		return -1;
	out->pk_off.assign(np + 1, 0);
	int64_t at = 0, mx = 0;
	for (long p = 0; p < np; p++) {
		out->pk_off[p] = at;
		at += (int64_t) pk[p].size();
		mx = std::max<int64_t>(mx, (int64_t) pk[p].size());
	}
	out->pk_off[np] = at;
	out->bytes.assign((size_t) at + 16, 0); // + 16: the decoder reads whole aligned words
#pragma omp parallel for schedule(static)
	for (long p = 0; p < np; p++)
		if (!pk[p].empty())
			memcpy(&out->bytes[out->pk_off[p]], pk[p].data(), pk[p].size());
	out->max_packet = mx;
	out->max_slots = max_slots;
	out->max_rows = max_rows;
	return 0;
}

} // namespace spmvdiv

/*
 * Cleaning kernels for gfx950 (MI355X), one wave64 = 64 lanes.
 *
 * Pass structure of one call (all on the caller's stream, no host round trip):
 *   zero counters and flags -> first pass (prologue + dense detection) ->
 *   replace(fill holes) -> maxiter x [block pass + detect -> replace] ->
 *   epilogue
 * Every per-frame decision of the reference loop (hole fill, early exit, the
 * "no finite replacement donors" error) is taken on the device from counters;
 * the host reads them once after the last pass.
 */
#include "common.hpp"
#include "networks.inc"

namespace cusmic {

/* ------------------------------------------------------------------------ */
/* Small helpers                                                             */

__device__ __forceinline__ int
block_sum(int v, int *scratch)
{
	/* wave reduction then one LDS slot per wave; BLOCK = 4 waves */
	for (int o = 32; o; o >>= 1)
		v += __shfl_xor(v, o, 64);
	int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	__syncthreads();
	if (!lane)
		scratch[wave] = v;
	__syncthreads();
	int s = scratch[0] + scratch[1] + scratch[2] + scratch[3];
	__syncthreads();
	return s;
}

/* Append this lane's item to a block-shared list when pred holds; returns
 * nothing, the list length lives in *count (LDS). Order is not deterministic,
 * and no consumer depends on it. */
template <class T>
__device__ __forceinline__ void
lds_append(bool pred, T item, T *list, int *count)
{
	unsigned long long m = __ballot(pred);
	if (!m)
		return;
	int lane = threadIdx.x & 63;
	int leader = __ffsll((long long)m) - 1;
	int base = 0;
	if (lane == leader)
		base = atomicAdd(count, __popcll(m));
	base = __shfl(base, leader, 64);
	if (pred)
		list[base + __popcll(m & ((1ull << lane) - 1))] = item;
}

template <class T>
__device__ __forceinline__ void
list_push(bool pred, T item, T *list, unsigned int *count)
{
	unsigned long long m = __ballot(pred);
	if (!m)
		return;
	int lane = threadIdx.x & 63, leader = __ffsll((long long)m) - 1;
	unsigned base = 0;
	if (lane == leader)
		base = atomicAdd(count, (unsigned)__popcll(m));
	base = __shfl(base, leader, 64);
	if (pred)
		list[base + __popcll(m & ((1ull << lane) - 1))] = item;
}

/* list_push into a list that fills from its top: entry k at end[-1 - k]. */
template <class T>
__device__ __forceinline__ void
list_push_top(bool pred, T item, T *end, unsigned int *count)
{
	unsigned long long m = __ballot(pred);
	if (!m)
		return;
	int lane = threadIdx.x & 63, leader = __ffsll((long long)m) - 1;
	unsigned base = 0;
	if (lane == leader)
		base = atomicAdd(count, (unsigned)__popcll(m));
	base = __shfl(base, leader, 64);
	if (pred)
		end[-1 - long(base + __popcll(m & ((1ull << lane) - 1)))] = item;
}

/* Per-workgroup output buffer of list entries: lanes append in LDS, then one
 * global atomic per flush reserves room in the shared list (a single counter
 * hammered by every wave serialises in L2). */
template <int CAP>
struct BlockList {
	unsigned long long item[CAP];
	int n;
	unsigned base;
};

template <int CAP>
__device__ __forceinline__ void
block_list_flush(BlockList<CAP> &b, unsigned long long *list, unsigned int *count)
{
	__syncthreads();
	const int n = b.n;
	if (!threadIdx.x && n)
		b.base = atomicAdd(count, (unsigned)n);
	__syncthreads();
	for (int k = threadIdx.x; k < n; k += BLOCK)
		list[b.base + k] = b.item[k];
	__syncthreads();
	if (!threadIdx.x)
		b.n = 0;
	__syncthreads();
}

/* Add each lane's value to the counter of its frame (base + f * stride), one
 * atomic per distinct frame in the wave. */
__device__ __forceinline__ void
wave_frame_sum(int value, int f, int *base, int stride)
{
	unsigned long long m = __ballot(value != 0);
	int lane = threadIdx.x & 63;
	while (m) {
		int leader = __ffsll((long long)m) - 1;
		int lf = __shfl(f, leader, 64);
		bool same = value != 0 && f == lf;
		unsigned long long sm = __ballot(same);
		int v = same ? value : 0;
		for (int o = 32; o; o >>= 1)
			v += __shfl_xor(v, o, 64);
		if (lane == leader)
			atomicAdd(base + (size_t)lf * stride, v);
		m &= ~sm;
	}
}

__device__ __forceinline__ double
noise_model(double med5, double g, double rn)
{
	/* src/filters.cuh:107-123 */
	double m = med5 < SIGNAL_FLOOR ? SIGNAL_FLOOR : med5;
	return sqrt(rn * rn + g * m) / g;
}

/* ------------------------------------------------------------------------ */
/* Call set-up                                                               */

__global__ void
zero_words(unsigned int *words, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
	     i += (size_t)gridDim.x * blockDim.x)
		words[i] = 0;
}

/* Zero n 16-byte groups (the flags plane) and, when mask is given, the
 * mask_bytes of the caller's mask plane, whose start may have any alignment:
 * its bytes before the first 16-byte boundary and after the last one go one
 * by one. */
__global__ void
zero_groups(uint4 *groups, size_t n, uint8_t *mask, size_t mask_bytes)
{
	const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
	for (size_t i = t; i < n; i += stride)
		groups[i] = make_uint4(0, 0, 0, 0);
	if (!mask)
		return;
	const size_t head = std::min<size_t>((16 - (reinterpret_cast<uintptr_t>(mask) & 15)) & 15, mask_bytes);
	const size_t body = (mask_bytes - head) / 16, tail = head + body * 16;
	uint4 *aligned = reinterpret_cast<uint4 *>(mask + head);
	for (size_t i = t; i < body; i += stride)
		aligned[i] = make_uint4(0, 0, 0, 0);
	if (t < head)
		mask[t] = 0;
	if (t < mask_bytes - tail)
		mask[tail + t] = 0;
}

/* Fold the pending per-frame count shards into the totals (and, after a
 * detection pass, the pass's new-pixel count) and clear them: SHARDS lanes
 * per frame, one shard each, summed across the lanes. */
/* The fold of global thread t (t / SHARDS is the frame): every lane of the
 * wave calls it, with the waves' t aligned to SHARDS. */
__device__ __forceinline__ void
fold_shards(const Params &p, int iter, int first, unsigned t)
{
	static_assert(SHARDS == 32, "a frame's shards fill half a wave");
	const unsigned f = t / SHARDS;
	int d[SHARD_FIELDS] = {};
	if (f < unsigned(p.nf)) {
		int *sh = p.shard + (size_t)t * SHARD_FIELDS;
#pragma unroll
		for (int k = 0; k < SHARD_FIELDS; ++k) {
			d[k] = sh[k];
			sh[k] = 0;
		}
	}
#pragma unroll
	for (int k = 0; k < SHARD_FIELDS; ++k)
		for (int o = SHARDS / 2; o; o >>= 1)
			d[k] += __shfl_xor(d[k], o, 64);
	if (f < unsigned(p.nf) && !(t % SHARDS)) {
		p.counts[f].holes += d[SHARD_HOLES];
		p.counts[f].allowed += d[SHARD_ALLOWED];
		p.counts[f].donors += d[SHARD_DONORS];
		if (!first) {
			const int fresh = d[SHARD_FRESH];
			p.todo[todo_index(p, iter, f)] = fresh;
			if (fresh > 0) {
				p.counts[f].fresh_passes += 1;
				/* every writer stores the same value */
				__atomic_store_n(&p.nlist[nl_active(p, iter)], 1u, __ATOMIC_RELAXED);
			}
		}
	}
}

__global__ void
fold_counts(Params p, int iter, int first)
{
	fold_shards(p, iter, first, blockIdx.x * blockDim.x + threadIdx.x);
}

/* Copy the n counter words the host reads into its pinned block (after the
 * block's first 64 words), then write seq to the block's first word, which
 * the host polls. Every lane's stores are made visible at system scope before
 * the barrier, and the sequence word is released after it, so a host that
 * sees seq sees the words. One workgroup of 256 lanes. */
__global__ __launch_bounds__(256) void
post_counters(const unsigned int *words, unsigned int *box, unsigned n, unsigned seq)
{
	for (unsigned i = threadIdx.x; i < n; i += 256)
		box[64 + i] = words[i];
	__threadfence_system();
	__syncthreads();
	if (!threadIdx.x)
		__hip_atomic_store(box, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

/* Clear the ring slots of passes [lo, hi] before a chunk reuses them (at most
 * PASS_RING passes; a fill kernel, not a memset, so that the call can be
 * captured in a graph). */
__global__ void
clear_pass_slots(Params p, int lo, int hi)
{
	const int n = (hi - lo + 1) * PASS_FIELDS;
	for (int t = threadIdx.x; t < n; t += blockDim.x)
		p.nlist[nl_pass(lo + t / PASS_FIELDS, t % PASS_FIELDS)] = 0;
}

/* Donor records start as (-1, -1): no donor. */
__global__ void
reset_donors(int32_t *donors, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
	     i += (size_t)gridDim.x * blockDim.x)
		donors[i] = -1;
}

/* Iterate the pixels of one tile: lane -> column, 4 rows per pass. */
struct TilePixels {
	int x0, y0, w, h;
};

__device__ __forceinline__ TilePixels
tile_of(const Frame &fr, int tile)
{
	int lt = tile - fr.tile0;
	TilePixels t;
	t.x0 = (lt % fr.tiles_x) * TILE_W;
	t.y0 = (lt / fr.tiles_x) * TILE_H;
	t.w = fr.w;
	t.h = fr.h;
	return t;
}

/* chunk -> frame and run -> frame tables (binary search once per entry) */
__global__ void
set_tables(Params p)
{
	for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < p.nchunks + p.nruns;
	     t += gridDim.x * blockDim.x) {
		const bool chunk = t < p.nchunks;
		int c = chunk ? t : t - p.nchunks, lo = 0, hi = p.nf - 1;
		while (lo < hi) {
			int mid = (lo + hi + 1) >> 1;
			if ((chunk ? p.frames[mid].chunk0 : p.frames[mid].run0) <= c)
				lo = mid;
			else
				hi = mid - 1;
		}
		(chunk ? p.chunk_frame : p.run_frame)[c] = lo;
	}
}

/* maxiter = 0 (src/image.cuh:72-92): background arithmetic only, one
 * workgroup per chunk of CHUNK consecutive pixels of a frame. */
__global__ __launch_bounds__(BLOCK) void
copy_without_detection(Params p)
{
	const int f = p.chunk_frame[blockIdx.x];
	const Frame fr = p.frames[f];
	const int n = fr.w * fr.h, start = (blockIdx.x - fr.chunk0) * CHUNK + threadIdx.x;
	const bool bg_map = fr.background != nullptr, bg = bg_map || !isnan(p.bg);
	for (int i = start; i < start + CHUNK && i < n; i += BLOCK) {
		double d = fr.data[i];
		double v = isfinite(d) ? d : 0.0;
		if (bg) {
			double b = bg_map ? fr.background[i] : p.bg;
			v += 1 * b;
			v += -1 * b;
		}
		p.clean[fr.off + i] = isfinite(d) ? v : d;
		p.mask[fr.off + i] = 0;
	}
}

/* Frame that holds batch pixel g, searching forward from frame f. */
__device__ __forceinline__ int
frame_from(const Frame *frames, int nf, int f, long long g)
{
	while (f + 1 < nf && frames[f + 1].off <= g)
		++f;
	return f;
}

/* Epilogue without a background, where out = C except at holes: the mask
 * plane was zeroed with the flags, and only the flagged pixels (the flagged
 * list, from the bottom of list) and the holes (the hole list, from its top)
 * are touched: mask 1 at the former, out = D (payload bits kept) at the
 * latter. Holes are never flagged. The same output as epilogue. */
__global__ __launch_bounds__(BLOCK) void
epilogue_sparse(Params p)
{
	const unsigned flagged = p.nlist[NL_LIST], holes = p.nlist[NL_HOLES];
	for (unsigned e = blockIdx.x * BLOCK + threadIdx.x; e < flagged + holes; e += gridDim.x * BLOCK) {
		const bool hole = e >= flagged;
		const unsigned long long entry = hole ? p.list[p.list_cap - 1 - (e - flagged)] : p.list[e];
		const Frame &fr = p.frames[entry >> 32];
		const long long i = (long long)(entry & 0xffffffffu);
		if (hole)
			p.clean[fr.off + i] = fr.data[i];
		else
			p.mask[fr.off + i] = 1;
	}
}

/* Epilogue (src/image.cuh:50-70): out = finite D ? C - bg : D (payload bits
 * kept), mask = flagged bit. A workgroup step covers 8 BLOCK batch pixels.
 * The mask takes eight pixels per lane, so that the byte planes move in 64-bit
 * accesses (bytewise where the caller's mask is not 8-byte aligned). The clean
 * image is then walked with consecutive pixels in consecutive lanes, and only
 * in steps that hold a hole or lose a background. */
__global__ __launch_bounds__(BLOCK) void
epilogue(Params p, long long total, int bg_maps)
{
	constexpr unsigned long long LOW = 0x0101010101010101ull;
	constexpr int SPAN = 8 * BLOCK;
	static_assert(HOLE == 1 && CRAY == 4, "bit positions used below");
	const bool vec_mask = !(reinterpret_cast<uintptr_t>(p.mask) & 7);
	const bool bg_scalar = !isnan(p.bg);
	for (long long b0 = (long long)blockIdx.x * SPAN; b0 < total; b0 += (long long)gridDim.x * SPAN) {
		const long long g0 = b0 + 8 * threadIdx.x;
		unsigned long long fl = 0;
		if (g0 + 8 <= total) {
			fl = *reinterpret_cast<const unsigned long long *>(p.flags + g0);
		} else {
			for (int k = 0; g0 + k < total; ++k)
				fl |= (unsigned long long)p.flags[g0 + k] << (8 * k);
		}
		const unsigned long long m = (fl >> 2) & LOW;
		if (g0 + 8 <= total && vec_mask) {
			*reinterpret_cast<unsigned long long *>(p.mask + g0) = m;
		} else {
			for (int k = 0; k < 8 && g0 + k < total; ++k)
				p.mask[g0 + k] = uint8_t(m >> (8 * k));
		}
		const bool holes = __syncthreads_or((fl & LOW) != 0);
		if (!holes && !bg_maps) {
			/* no frame lookup: every pixel of the step loses the scalar */
			if (bg_scalar) {
				for (int k = 0; k < 8; ++k) {
					const long long g = b0 + k * BLOCK + threadIdx.x;
					if (g < total)
						p.clean[g] = p.clean[g] + -1 * p.bg;
				}
			}
			continue;
		}
		int f = 0;
		{
			const long long g = min(b0 + threadIdx.x, total - 1);
			int lo = 0, hi = p.nf - 1;
			while (lo < hi) {
				const int mid = (lo + hi + 1) >> 1;
				if (p.frames[mid].off <= g)
					lo = mid;
				else
					hi = mid - 1;
			}
			f = lo;
		}
		for (int k = 0; k < 8; ++k) {
			const long long g = b0 + k * BLOCK + threadIdx.x;
			if (g >= total)
				break;
			f = frame_from(p.frames, p.nf, f, g);
			const Frame &fr = p.frames[f];
			const long long i = g - fr.off;
			if (p.flags[g] & HOLE)
				p.clean[g] = fr.data[i];
			else if (fr.background || bg_scalar)
				p.clean[g] = p.clean[g] + -1 * (fr.background ? fr.background[i] : p.bg);
		}
	}
}

/* ------------------------------------------------------------------------ */
/* Replacement (src/replace.cuh:112-153)                                     */

__device__ __forceinline__ bool
is_donor(const uint8_t *flags, const double *clean, int q, double &v)
{
	if (flags[q] & (EXCL | CRAY))
		return false;
	v = clean[q];
	return isfinite(v);
}

/* The rank-th smallest donor key in the clipped window by MSB-first binary
 * search over the 64 key bits: each step counts donors below the candidate
 * prefix. Used only when the 5x5 window holds no donor. */
__device__ unsigned long long
window_select(const uint8_t *flags, const double *clean, int w, int x0, int x1, int y0, int y1,
	int rank)
{
	unsigned long long prefix = 0;
	for (int bit = 63; bit >= 0; --bit) {
		unsigned long long hi_mask = bit == 63 ? 0 : ~0ull << (bit + 1);
		unsigned long long probe = prefix;  /* candidate with this bit clear */
		int below = 0;
		for (int y = y0; y <= y1; ++y)
			for (int x = x0; x <= x1; ++x) {
				double v;
				if (!is_donor(flags, clean, y * w + x, v))
					continue;
				unsigned long long k = order_key(v);
				below += (k & hi_mask) == (probe & hi_mask) && !((k >> bit) & 1);
			}
		if (rank >= below) {
			prefix |= 1ull << bit;
			rank -= below;
		}
	}
	return prefix;
}

/* Index of the element of rank r in the (key, pixel index) order, given its
 * key: skip r - (#keys below) equal keys in raster order. */
__device__ int
rank_index(const uint8_t *flags, const double *clean, int w, int x0, int x1, int y0, int y1,
	unsigned long long key, int r)
{
	int less = 0;
	for (int y = y0; y <= y1; ++y)
		for (int x = x0; x <= x1; ++x) {
			double v;
			if (is_donor(flags, clean, y * w + x, v))
				less += order_key(v) < key;
		}
	int skip = r - less;
	for (int y = y0; y <= y1; ++y)
		for (int x = x0; x <= x1; ++x) {
			double v;
			if (is_donor(flags, clean, y * w + x, v) && order_key(v) == key) {
				if (!skip)
					return y * w + x;
				--skip;
			}
		}
	return -1;
}

/* Key of the rank-r element when the value network returned v: equal values
 * differ only for zeros, where -0 ranks below +0. cnt_neg0 counts donors whose
 * key is at most key(-0). */
__device__ __forceinline__ unsigned long long
rank_key(double v, int r, int cnt_neg0)
{
	if (v == 0.0)
		return order_key(r < cnt_neg0 ? -0.0 : 0.0);
	return order_key(v);
}

/* Record that pixel i of frame f changed (changed = this lane's pixel did)
 * in the replacement before pass next: the first such pixel of a block puts
 * the block on that pass's changed list (one atomic per wave). */
__device__ __forceinline__ bool
changed_block(const Params &p, bool changed, int f, int i, int next, unsigned long long &entry)
{
	const unsigned tag = tag_stamp(next);
	unsigned b = 0;
	bool fresh = false;
	if (changed) {
		const Frame fr = p.frames[f];
		const int x = i % fr.w, y = i / fr.w, ox = x % BLK, oy = y % BLK;
		b = unsigned(fr.block0 + (y / BLK) * fr.blocks_x + x / BLK);
		auto sides_of = [&](int r) {
			return (ox < r ? SIDE_L : 0u) | (ox >= BLK - r ? SIDE_R : 0u) | (oy < r ? SIDE_U : 0u) |
			       (oy >= BLK - r ? SIDE_D : 0u);
		};
		const unsigned sides = sides_of(p.reach) | sides_of(p.reach_s) << SIDES_S;
		if ((__atomic_load_n(&p.changed_sides[b], __ATOMIC_RELAXED) & sides) != sides)
			atomicOr(&p.changed_sides[b], sides);
		fresh = __atomic_load_n(&p.changed_tag[b], __ATOMIC_RELAXED) != tag &&
			atomicExch(&p.changed_tag[b], tag) != tag;
	}
	entry = list_entry(f, (int)b);
	return fresh;
}

/* Record that pixel i of frame f changed (changed: this lane's pixel did):
 * its block joins the changed list of pass next once (whole wave). */
__device__ __forceinline__ void
mark_dirty(const Params &p, bool changed, int f, int i, int next)
{
	unsigned long long entry;
	const bool fresh = changed_block(p, changed, f, i, next, entry);
	list_push(fresh, entry, p.changed_list, &p.nlist[nl_changed(p, next)]);
}

/* Blocks of pass iter: the neighbours of a changed block on the sides its
 * changes reach (the reach is below BLK, so only adjacent blocks), diagonal
 * ones when both of their sides are reached. Within the S' reach a block
 * recomputes S'; only within the detection reach, it re-emits its seeds from
 * the S' it keeps (RESEED entries). Under WRAP every block within two
 * recomputes, because a partial block at the seam is narrower than the reach.
 * Each block joins the list once per kind (tagged with the pass); the sides
 * are reset for the next pass. */
__global__ __launch_bounds__(BLOCK) void
expand_dirty(Params p, int iter)
{
	static_assert(REACH < BLK, "a change reaches the adjacent blocks only");
	__shared__ BlockList<BLOCK * 25> out;
	__shared__ int wave_total[BLOCK / 64];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned n = p.nlist[nl_changed(p, iter)], tag = tag_stamp(iter);
	const bool wrap = p.mode == CUSMIC_WRAP;
	const int r = wrap ? 2 : 1, side = 2 * r + 1;
	unsigned int *count = &p.nlist[nl_dirty(p, iter)];
	if (!threadIdx.x)
		out.n = 0;
	/* the previous pass's new pixels stop being new (NEWP served its
	 * replacement) */
	{
		const unsigned first = p.nlist[nl_start(p, iter - 1)], last = p.nlist[0];
		for (unsigned e = first + blockIdx.x * BLOCK + threadIdx.x; e < last; e += gridDim.x * BLOCK) {
			const unsigned long long entry = p.list[e];
			const Frame &fr = p.frames[entry >> 32];
			flags_clear(&p.flags[fr.off + (entry & 0xffffffffu)], NEWP);
		}
	}
	for (unsigned base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
		const unsigned e = base + threadIdx.x;
		int f = 0, bx = 0, by = 0, nbx = 1, nby = 1, block0 = 0;
		unsigned sides = 0;
		if (e < n) {
			unsigned long long entry = p.changed_list[e];
			f = (int)(entry >> 32);
			const Frame fr = p.frames[f];
			block0 = fr.block0;
			nbx = fr.blocks_x;
			nby = (fr.h + BLK - 1) / BLK;
			const int b = (int)(entry & 0xffffffffu), lb = b - block0;
			bx = lb % nbx;
			by = lb / nbx;
			sides = p.changed_sides[b];
			p.changed_sides[b] = 0;
		}
		/* Every reached neighbour's tag first; then the fresh ones go to the
		 * buffer in lane order, each lane's neighbours together, so that the
		 * blocks around one change sit side by side in the list and the block
		 * pass takes them together (under one L2). */
		auto neighbour = [&](int k, int &q, bool &near_s) {
			const int dx = k % side - r, dy = k / side - r;
			int qx = bx + dx, qy = by + dy;
			auto reaches = [&](unsigned sd) {
				return (dx == 0 || (sd & (dx < 0 ? SIDE_L : SIDE_R))) &&
				       (dy == 0 || (sd & (dy < 0 ? SIDE_U : SIDE_D)));
			};
			near_s = wrap || reaches(sides >> SIDES_S);
			const bool reached = near_s || reaches(sides & 15u);
			if (wrap) {
				qx = (qx % nbx + nbx) % nbx;
				qy = (qy % nby + nby) % nby;
			}
			q = block0 + qy * nbx + qx;
			return e < n && reached && qx >= 0 && qx < nbx && qy >= 0 && qy < nby;
		};
		unsigned fresh = 0;
		for (int k = 0; k < side * side; ++k) {
			int q;
			bool near_s;
			if (neighbour(k, q, near_s)) {
				unsigned int *tags = near_s ? p.dirty_tag : p.reseed_tag;
				if (__atomic_load_n(&tags[q], __ATOMIC_RELAXED) != tag && atomicExch(&tags[q], tag) != tag)
					fresh |= 1u << k;
			}
		}
		const int cnt = __popc(fresh);
		int incl = cnt;
		for (int d = 1; d < 64; d <<= 1) {
			const int t = __shfl_up(incl, d, 64);
			if (lane >= d)
				incl += t;
		}
		if (lane == 63)
			wave_total[wave] = incl;
		__syncthreads();
		int at = incl - cnt, total = 0;
		for (int k = 0; k < BLOCK / 64; ++k) {
			at += k < wave ? wave_total[k] : 0;
			total += wave_total[k];
		}
		for (unsigned m = fresh; m; m &= m - 1) {
			int q;
			bool near_s;
			neighbour(__ffs(m) - 1, q, near_s);
			out.item[at++] = list_entry(f, int(unsigned(q) | (near_s ? 0u : RESEED)));
		}
		if (!threadIdx.x)
			out.n = total;
		block_list_flush(out, p.dirty_list, count);
	}
}

/* Value of a replacement from the two middle donors (src/replace.cuh:141-150):
 * odd count -> 0.0 + lo; even -> overflow-safe midpoint. When both middles are
 * zeros, the key order (-0 below +0) that the reference uses for windows above
 * 32 pixels decides the sign: -0 iff more than n/2 donors have a key at most
 * key(-0) (neg0 counts them); smaller windows take network_zero. */
__device__ __forceinline__ double
replacement(int n, double lo, double hi, int neg0)
{
	if (n & 1)
		return 0.0 + lo;
	if (lo == 0.0 && hi == 0.0)
		return neg0 > (n >> 1) ? -0.0 : 0.0;
	return midpoint(lo, hi);
}

/* The sign of an even-count replacement whose two middle donors are zeros, in
 * a window of at most 32 pixels. The reference sorts such windows by value on
 * a 32-lane bitonic network (src/replace.cuh:68-83,132-145): lane pos holds the
 * pixel at raster position pos of the window if it is a donor, else +inf, and
 * a compare-exchange moves values only when they differ, so -0 and +0 keep
 * their lanes against each other. Whether an exchange that involves a zero
 * happens depends only on the classes (negative, zero, positive) of the two
 * lanes, and exchanges within a class leave the class layout as it was, so the
 * network runs on three lane masks: negatives, zeros, and the zeros that are
 * -0. The midpoint of two zeros is -0 iff both are -0. */
__device__ double
network_zero(const uint8_t *flags, const double *clean, int w, int x0, int x1, int y0, int y1, int n)
{
	const int ww = x1 - x0 + 1, area = ww * (y1 - y0 + 1);
	uint32_t neg = 0, zero = 0, nzero = 0;
	for (int k = 0; k < area; ++k) {
		double v;
		if (is_donor(flags, clean, (y0 + k / ww) * w + x0 + k % ww, v)) {
			neg |= uint32_t(v < 0) << k;
			zero |= uint32_t(v == 0.0) << k;
			nzero |= uint32_t(v == 0.0 && signbit(v)) << k;
		}
	}
	/* LANES[t]: lanes whose bit t is clear */
	constexpr uint32_t LANES[5] = {0x55555555u, 0x33333333u, 0x0f0f0f0fu, 0x00ff00ffu, 0x0000ffffu};
	for (int s = 1; s <= 5; ++s) {     /* span 2^s */
		const uint32_t keeps_min = s < 5 ? LANES[s] : ~0u;
		for (int t = s - 1; t >= 0; --t) { /* step 2^t: lane l pairs with l + step */
			const int step = 1 << t;
			const uint32_t pos = ~(neg | zero);
			const uint32_t nu = neg >> step, zu = zero >> step, pu = pos >> step;
			const uint32_t l_above = (zero & nu) | (pos & (nu | zu));
			const uint32_t u_above = (zu & neg) | (pu & (neg | zero));
			const uint32_t swap = LANES[t] & ((keeps_min & l_above) | (~keeps_min & u_above));
			uint32_t d = (neg ^ nu) & swap;
			neg ^= d | (d << step);
			d = (zero ^ zu) & swap;
			zero ^= d | (d << step);
			d = (nzero ^ (nzero >> step)) & swap;
			nzero ^= d | (d << step);
		}
	}
	const int a = (n - 1) >> 1, b = n >> 1;
	return (nzero >> a) & (nzero >> b) & 1u ? -0.0 : 0.0;
}

/* Store the replacement of pixel i; record the donors when asked (ties in
 * raster order within the final window), unless the caller found them
 * (known: ia, ib). Returns true if the bits changed. */
__device__ bool
store_replacement(const Params &p, const Frame &fr, int i, int32_t *donors, int n, double lo,
	double hi, int neg0, int x0, int x1, int y0, int y1, bool known = false, int ia = -1, int ib = -1)
{
	const double *clean = p.clean + fr.off;
	const uint8_t *flags = p.flags + fr.off;
	double out = replacement(n, lo, hi, neg0);
	if (!(n & 1) && lo == 0.0 && hi == 0.0 && (x1 - x0 + 1) * (y1 - y0 + 1) <= 32)
		out = network_zero(flags, clean, fr.w, x0, x1, y0, y1, n);
	if (donors) {
		int a = (n - 1) >> 1, b = n >> 1;
		if (!known) {
			ia = rank_index(flags, clean, fr.w, x0, x1, y0, y1, rank_key(lo, a, neg0), a);
			ib = a == b ? ia : rank_index(flags, clean, fr.w, x0, x1, y0, y1, rank_key(hi, b, neg0), b);
		}
		donors[2 * (fr.off + i)] = ia;
		donors[2 * (fr.off + i) + 1] = ib;
	}
	double *dst = p.clean + fr.off + i;
	double old = *dst;
	*dst = out;
	return __double_as_longlong(old) != __double_as_longlong(out);
}

/* Gather the K neighbours of the (2R+1)^2 window around (x, y) (centre
 * excluded): donor values, +inf elsewhere. Flags and values are loaded together
 * (one memory round trip); a value read at a flagged neighbour, which another
 * lane may be rewriting, is discarded. */
template <int R, int K>
__device__ __forceinline__ int
gather_window(const Params &p, const Frame &fr, int x, int y, double (&v)[K], int &neg0,
	bool *any_new = nullptr)
{
	constexpr int S = 2 * R + 1, C = R * S + R;
	const double *clean = p.clean + fr.off;
	const uint8_t *flags = p.flags + fr.off;
	const int w = fr.w, h = fr.h, i = y * w + x;
	uint32_t fl[K];
#pragma unroll
	for (int k = 0; k < K; ++k) {
		int o = k < C ? k : k + 1, qx = x + o % S - R, qy = y + o / S - R;
		bool in = qx >= 0 && qx < w && qy >= 0 && qy < h;
		int q = in ? qy * w + qx : i;
		fl[k] = in ? flags[q] : EXCL;
		v[k] = clean[q];
	}
	int n = 0;
	neg0 = 0;
	if (any_new) {
		bool a = false;
#pragma unroll
		for (int k = 0; k < K; ++k)
			a |= (fl[k] & NEWP) != 0;
		*any_new = a;
	}
#pragma unroll
	for (int k = 0; k < K; ++k) {
		bool ok = !(fl[k] & (EXCL | CRAY)) && isfinite(v[k]);
		v[k] = ok ? v[k] : INFINITY;
		n += ok;
		neg0 += ok && (v[k] < 0 || (v[k] == 0.0 && signbit(v[k])));
	}
	return n;
}

/* rank_index over the 5x5 window around (x, y) as gather_window holds it:
 * its order is the window's raster order without the centre (never a donor),
 * and a non-donor holds +inf. Donor values do not change during a pass, so
 * this equals the scan of the window in memory. */
__device__ __forceinline__ int
rank_index_near(const double (&v)[24], int x, int y, int w, unsigned long long key, int r)
{
	int less = 0;
#pragma unroll
	for (int k = 0; k < 24; ++k)
		less += v[k] != INFINITY && order_key(v[k]) < key;
	int skip = r - less, idx = -1;
#pragma unroll
	for (int k = 0; k < 24; ++k) {
		const int o = k < 12 ? k : k + 1;
		const bool hit = v[k] != INFINITY && order_key(v[k]) == key;
		if (hit && skip == 0)
			idx = (y + o / 5 - 2) * w + x + o % 5 - 2;
		skip -= hit;
	}
	return idx;
}

/* Replacement from the 5x5 window (src/replace.cuh:118-150 with r = 2).
 * Returns 1 if the pixel changed, 0 if not, -1 if the window has no donor. */
__device__ int
replace_near(const Params &p, const Frame &fr, int i, int32_t *donors)
{
	const int w = fr.w, h = fr.h, x = i % w, y = i / w;
	double v[24], s[13];
	int neg0;
	int n = gather_window<2, 24>(p, fr, x, y, v, neg0);
	if (!n)
		return -1;
	sort_low13_of24(v, s);
	int a = (n - 1) >> 1, b = n >> 1;
	double lo = s[0], hi = s[0];
#pragma unroll
	for (int j = 1; j < 13; ++j) {
		lo = j == a ? s[j] : lo;
		hi = j == b ? s[j] : hi;
	}
	/* the donors from the window in registers */
	int ia = -1, ib = -1;
	if (donors) {
		ia = rank_index_near(v, x, y, w, rank_key(lo, a, neg0), a);
		ib = a == b ? ia : rank_index_near(v, x, y, w, rank_key(hi, b, neg0), b);
	}
	return store_replacement(p, fr, i, donors, n, lo, hi, neg0, max(0, x - 2), min(w - 1, x + 2),
		max(0, y - 2), min(h - 1, y + 2), true, ia, ib);
}


/* Replacement from a widened window, r = 3, 4, ... (a 5x5 window without
 * donors). Returns true if the pixel changed. With only_if_new, the pixel is
 * left alone unless its final window holds a pixel flagged in this pass
 * (NEWP): donors only disappear, so its previous window and donors are then
 * unchanged. */
__device__ bool
replace_wide(const Params &p, const Frame &fr, int i, int32_t *donors, bool only_if_new)
{
	const double *clean = p.clean + fr.off;
	const uint8_t *flags = p.flags + fr.off;
	const int w = fr.w, h = fr.h, x = i % w, y = i / w;
	{
		double v[48], s[25];
		int neg0;
		bool any_new = false;
		int n = gather_window<3, 48>(p, fr, x, y, v, neg0, &any_new);
		if (n) {
			if (only_if_new && !any_new)
				return false;
			sort_low25_of48(v, s);
			int a = (n - 1) >> 1, b = n >> 1;
			double lo = s[0], hi = s[0];
#pragma unroll
			for (int j = 1; j < 25; ++j) {
				lo = j == a ? s[j] : lo;
				hi = j == b ? s[j] : hi;
			}
			return store_replacement(p, fr, i, donors, n, lo, hi, neg0, max(0, x - 3),
				min(w - 1, x + 3), max(0, y - 3), min(h - 1, y + 3));
		}
	}
	int rmax = max(w, h), n = 0, x0 = 0, x1 = 0, y0 = 0, y1 = 0;
	for (int r = 4; r <= rmax; ++r) {
		x0 = max(0, x - r);
		x1 = min(w - 1, x + r);
		y0 = max(0, y - r);
		y1 = min(h - 1, y + r);
		n = 0;
		for (int yy = y0; yy <= y1; ++yy)
			for (int xx = x0; xx <= x1; ++xx) {
				double t;
				n += is_donor(flags, clean, yy * w + xx, t);
			}
		if (n)
			break;
	}
	if (!n)
		return false; /* unreachable: callers check the frame has donors */
	if (only_if_new) {
		bool any_new = false;
		for (int yy = y0; yy <= y1 && !any_new; ++yy)
			for (int xx = x0; xx <= x1; ++xx)
				any_new |= (flags[yy * w + xx] & NEWP) != 0;
		if (!any_new)
			return false;
	}
	/* windows this wide are rare: the select reads the window from memory */
	double lo = key_value(window_select(flags, clean, w, x0, x1, y0, y1, (n - 1) >> 1));
	double hi = n & 1 ? lo : key_value(window_select(flags, clean, w, x0, x1, y0, y1, n >> 1));
	int neg0 = 0;
	for (int yy = y0; yy <= y1; ++yy)
		for (int xx = x0; xx <= x1; ++xx) {
			double t;
			if (is_donor(flags, clean, yy * w + xx, t))
				neg0 += t < 0 || (t == 0.0 && signbit(t));
		}
	return store_replacement(p, fr, i, donors, n, lo, hi, neg0, x0, x1, y0, y1);
}

/* A 64-bit value that is the same in every lane, as a wave-uniform value. */
__device__ __forceinline__ unsigned long long
uniform_u64(unsigned long long v)
{
	unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
	unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
	return (unsigned long long)hi << 32 | lo;
}

/* Window radius a wave handles in registers: the (2 RW + 1)^2 neighbourhood
 * in chunks of 64 positions. */
constexpr int RW = 10;
constexpr int RW_CHUNKS = ((2 * RW + 1) * (2 * RW + 1) + 63) / 64;

__device__ __forceinline__ int
wave_sum(int v)
{
	for (int o = 32; o; o >>= 1)
		v += __shfl_xor(v, o, 64);
	return v;
}

/* Donors of a window whose donors all lie on its rim (at most 8 RW keys):
 * compacted into RING_CHUNKS chunks of 64, unused entries at the maximum key. */
constexpr int RING_CHUNKS = (8 * RW + 63) / 64;

/* rank-th smallest of the keys held in the lanes' chunks, by MSB-first bit
 * search: each step counts the keys under the candidate prefix whose next bit
 * is 0 (ballot + popcount per chunk). */
__device__ __forceinline__ unsigned long long
wave_select(const unsigned long long (&key)[RING_CHUNKS], int rank)
{
	unsigned long long prefix = 0;
	for (int bit = 63; bit >= 0; --bit) {
		const unsigned long long hi_mask = bit == 63 ? 0 : ~0ull << (bit + 1);
		int below = 0;
#pragma unroll
		for (int c = 0; c < RING_CHUNKS; ++c)
			below += __popcll(__ballot((key[c] & hi_mask) == (prefix & hi_mask) && !((key[c] >> bit) & 1)));
		if (rank >= below) {
			prefix |= 1ull << bit;
			rank -= below;
		}
	}
	return prefix;
}

/* Replacement of pixel i from its smallest window wider than 5 x 5 that holds
 * donors (src/replace.cuh:118-150, r = 3, 4, ...), by one whole wave: the
 * neighbourhood up to RW is loaded at once, r is the smallest Chebyshev
 * distance of a donor (the window's donors are exactly those at distance r),
 * and the middle keys come from a wave radix select. With only_if_new the
 * pixel is left alone unless its window holds a pixel flagged in this pass
 * (NEWP): donors only disappear, so its window and donors are then unchanged.
 * Larger windows fall back to replace_wide on one lane. Returns whether the
 * pixel changed (the same in every lane). */
__device__ bool
replace_wide_wave(const Params &p, const Frame &fr, int i, int32_t *donors, bool only_if_new,
	unsigned long long (*ring_keys)[64 * RING_CHUNKS])
{
	const int lane = threadIdx.x & 63;
	const double *clean = p.clean + fr.off;
	const uint8_t *flags = p.flags + fr.off;
	const int w = fr.w, h = fr.h, x = i % w, y = i / w;
	const int x0 = max(0, x - RW), y0 = max(0, y - RW);
	const int side = min(w - 1, x + RW) - x0 + 1, npos = side * (min(h - 1, y + RW) - y0 + 1);
	unsigned long long key[RW_CHUNKS];
	int dist[RW_CHUNKS];
	bool don[RW_CHUNKS], fresh[RW_CHUNKS];
	int rmin = RW + 1;
#pragma unroll
	for (int c = 0; c < RW_CHUNKS; ++c) {
		const int k = 64 * c + lane;
		don[c] = fresh[c] = false;
		dist[c] = RW + 1;
		key[c] = 0;
		if (k < npos) {
			const int xx = x0 + k % side, yy = y0 + k / side, q = yy * w + xx;
			const uint8_t fl = flags[q];
			const double v = clean[q];
			dist[c] = max(abs(xx - x), abs(yy - y));
			don[c] = !(fl & (EXCL | CRAY)) && isfinite(v);
			fresh[c] = (fl & NEWP) != 0;
			key[c] = order_key(v);
			if (don[c])
				rmin = min(rmin, dist[c]);
		}
	}
	for (int o = 32; o; o >>= 1)
		rmin = min(rmin, __shfl_xor(rmin, o, 64));
	if (rmin > RW) {
		/* no donor within RW: the lane search over growing windows */
		bool changed = false;
		if (!lane)
			changed = replace_wide(p, fr, i, donors, only_if_new);
		return __shfl(changed ? 1 : 0, 0, 64) != 0;
	}
	const int r = rmin;
	bool any_new = false;
	int n = 0, neg0 = 0;
#pragma unroll
	for (int c = 0; c < RW_CHUNKS; ++c) {
		any_new |= __ballot(fresh[c] && dist[c] <= r) != 0;
		/* every donor within r is at distance exactly r */
		n += don[c] && dist[c] == r;
		const double v = key_value(key[c]);
		neg0 += don[c] && dist[c] == r && (v < 0 || (v == 0.0 && signbit(v)));
	}
	if (only_if_new && !any_new)
		return false;
	n = wave_sum(n);
	neg0 = wave_sum(neg0);
	/* compact the rim donors (n <= 8r) into the wave's LDS slots */
	unsigned long long *slot = ring_keys[threadIdx.x >> 6];
	int base = 0;
#pragma unroll
	for (int c = 0; c < RW_CHUNKS; ++c) {
		const bool sel = don[c] && dist[c] == r;
		const unsigned long long m = __ballot(sel);
		if (sel)
			slot[base + __popcll(m & ((1ull << lane) - 1))] = key[c];
		base += __popcll(m);
	}
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	unsigned long long rk[RING_CHUNKS];
#pragma unroll
	for (int c = 0; c < RING_CHUNKS; ++c) {
		const int k = 64 * c + lane;
		rk[c] = k < n ? slot[k] : ~0ull;
	}
	__builtin_amdgcn_wave_barrier();
	const int a = (n - 1) >> 1;
	const unsigned long long klo = wave_select(rk, a);
	/* the upper middle: the same key if more than a + 1 keys are <= it, else
	 * the smallest key above it */
	unsigned long long khi = klo;
	if (!(n & 1)) {
		int le = 0;
		unsigned long long above = ~0ull;
#pragma unroll
		for (int c = 0; c < RING_CHUNKS; ++c) {
			le += __popcll(__ballot(64 * c + lane < n && rk[c] <= klo));
			if (64 * c + lane < n && rk[c] > klo)
				above = min(above, rk[c]);
		}
		for (int o = 32; o; o >>= 1)
			above = min(above, (unsigned long long)__shfl_xor((long long)above, o, 64));
		khi = le > a + 1 ? klo : above;
	}
	const double lo = key_value(klo), hi = key_value(khi);
	bool changed = false;
	if (!lane)
		changed = store_replacement(p, fr, i, donors, n, lo, hi, neg0, max(0, x - r),
			min(w - 1, x + r), max(0, y - r), min(h - 1, y + r));
	return __shfl(changed ? 1 : 0, 0, 64) != 0;
}

/* May the pixels of frame f be replaced in this pass? Sets the no-donor error. */
__device__ __forceinline__ bool
replace_allowed(const Params &p, int f, int iter, int fill)
{
	Counts c = p.counts[f];
	if (fill) {
		if (!c.allowed)
			return false;
		if (!c.donors) {
			atomicOr(p.status, ERR_NO_DONORS);
			return false;
		}
		return true;
	}
	if (!p.todo[todo_index(p, iter, f)])
		return false;
	if (c.donors <= 0) {
		atomicOr(p.status, ERR_NO_DONORS);
		return false;
	}
	return true;
}

/* Hole fill before the loop (fill = 1, src/cleaner.cuh:78-100) or the
 * replacement pass of iteration iter (src/cleaner.cuh:175-181). Donor values
 * never change (replacement writes flagged pixels only), so a flagged pixel's
 * value changes only when its window loses a donor: the pass replaces the
 * pixels it flagged, then the older ones within 2 pixels of them (redo list);
 * pixels that needed a window wider than 5 x 5 are redone by replace_widened.
 * Pixels whose 5x5 window has no donor go to the widened list. */
__global__ __launch_bounds__(BLOCK) void
replace_pass(Params p, int iter, int fill)
{
	/* changed blocks, flushed when the next round might not fit */
	__shared__ BlockList<2 * BLOCK> changed_out;
	if (!threadIdx.x)
		changed_out.n = 0;
	__syncthreads();
	const bool marking = !fill && iter + 1 < p.maxiter;
	const unsigned start = fill ? 0 : p.nlist[nl_start(p, iter)];
	const unsigned n_new = fill ? p.nlist[1] : p.nlist[0] - start;
	const unsigned n = n_new + (fill ? 0 : p.nlist[nl_redo(p, iter)]);
	unsigned int *wide = &p.nlist[fill ? NL_FILL_WIDE : nl_wide(p, iter)];
	if (!fill && !blockIdx.x && !threadIdx.x)
		p.nlist[nl_wide_base(p, iter)] = p.nlist[nl_wide_all(p)];
	for (unsigned base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
		unsigned e = base + threadIdx.x;
		bool widen = false, changed = false;
		unsigned long long entry = 0;
		int f = 0, i = 0;
		if (e < n) {
			entry = fill ? p.list[p.list_cap - 1 - e] : e < n_new ? p.list[start + e] : p.list_b[e - n_new];
			f = (int)(entry >> 32);
			i = (int)(entry & 0xffffffffu);
			if (replace_allowed(p, f, iter, fill)) {
				const Frame fr = p.frames[f];
				int r = replace_near(p, fr, i, fill ? nullptr : p.donors);
				widen = r < 0;
				changed = r > 0;
				if (!fill)
					flags_clear(&p.flags[fr.off + i], REDO);
			}
		}
		list_push(widen, entry, p.list_c, wide);
		if (marking) {
			unsigned long long block_entry;
			const bool fresh = changed_block(p, changed, f, i, iter + 1, block_entry);
			lds_append(fresh, block_entry, changed_out.item, &changed_out.n);
			__syncthreads();
			if (changed_out.n > BLOCK)
				block_list_flush(changed_out, p.changed_list, &p.nlist[nl_changed(p, iter + 1)]);
		}
	}
	if (marking)
		block_list_flush(changed_out, p.changed_list, &p.nlist[nl_changed(p, iter + 1)]);
}

/* Windows wider than 5 x 5, one wave per pixel: this pass's widened pixels,
 * then (after the hole fill) every pixel widened in an earlier pass, whose
 * window may have lost donors. A pixel widened for the first time joins
 * wide_all. */
__global__ __launch_bounds__(BLOCK) void
replace_widened(Params p, int iter, int fill)
{
	__shared__ unsigned long long ring_keys[BLOCK / 64][64 * RING_CHUNKS];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned n_pass = p.nlist[fill ? NL_FILL_WIDE : nl_wide(p, iter)];
	const unsigned n = n_pass + (fill ? 0 : p.nlist[nl_wide_base(p, iter)]);
	for (unsigned e = blockIdx.x * (BLOCK / 64) + wave; e < n; e += gridDim.x * (BLOCK / 64)) {
		const bool in_pass = e < n_pass;
		const unsigned long long entry = uniform_u64(in_pass ? p.list_c[e] : p.wide_all[e - n_pass]);
		const int f = (int)(entry >> 32), i = (int)(entry & 0xffffffffu);
		bool changed = false, join = false;
		if (in_pass || replace_allowed(p, f, iter, 0)) {
			const Frame fr = p.frames[f];
			changed = replace_wide_wave(p, fr, i, fill ? nullptr : p.donors, !in_pass, ring_keys);
			if (!fill && in_pass && !lane)
				join = !(flags_or(&p.flags[fr.off + i], WIDE) & WIDE);
		}
		list_push(join, entry, p.wide_all, &p.nlist[nl_wide_all(p)]);
		if (!fill && iter + 1 < p.maxiter)
			mark_dirty(p, changed && !lane, f, i, iter + 1);
	}
}

/* Older flagged pixels within 2 pixels of a pixel that pass iter flagged have
 * lost a donor from their 5x5 window: each joins the pass's redo list once
 * (REDO bit). Pixels flagged in this pass (NEWP) are replaced anyway and
 * pixels on wide_all (WIDE) are redone anyway. The launch also folds the
 * pass's count shards, which saves a launch of fold_counts. */
__global__ __launch_bounds__(BLOCK) void
redo_scan(Params p, int iter)
{
	__shared__ BlockList<BLOCK * 24> out;
	const unsigned start = p.nlist[nl_start(p, iter)], n = p.nlist[0] - start;
	unsigned int *count = &p.nlist[nl_redo(p, iter)];
	if (!threadIdx.x)
		out.n = 0;
	/* the pass's counts first (fold_counts for passes after the first: nothing
	 * here reads them, the replacement after this launch does) */
	for (unsigned t0 = blockIdx.x * BLOCK; t0 < unsigned(p.nf) * SHARDS; t0 += gridDim.x * BLOCK)
		fold_shards(p, iter, 0, t0 + threadIdx.x);
	for (unsigned base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
		const unsigned e = base + threadIdx.x;
		int f = 0, x = 0, y = 0, w = 1, h = 1;
		long long off = 0;
		if (e < n) {
			unsigned long long entry = p.list[start + e];
			f = (int)(entry >> 32);
			int i = (int)(entry & 0xffffffffu);
			const Frame fr = p.frames[f];
			w = fr.w;
			h = fr.h;
			off = fr.off;
			x = i % w;
			y = i / w;
		}
		/* the 24 neighbours' flags in one round trip, then the atomics of the
		 * few to redo, then one buffer flush */
		uint32_t fl[24];
#pragma unroll
		for (int k = 0; k < 24; ++k) {
			const int o = k < 12 ? k : k + 1; /* skip the centre */
			const int qx = x + o % 5 - 2, qy = y + o / 5 - 2;
			fl[k] = e < n && qx >= 0 && qx < w && qy >= 0 && qy < h ? uint32_t(p.flags[off + qy * w + qx]) : 0u;
		}
		bool redo[24];
#pragma unroll
		for (int k = 0; k < 24; ++k) {
			const int o = k < 12 ? k : k + 1;
			const int q = (y + o / 5 - 2) * w + x + o % 5 - 2;
			redo[k] = (fl[k] & CRAY) && !(fl[k] & (NEWP | WIDE | REDO)) &&
				!(flags_or(&p.flags[off + q], REDO) & REDO);
		}
		__syncthreads();
#pragma unroll
		for (int k = 0; k < 24; ++k) {
			const int o = k < 12 ? k : k + 1;
			lds_append(redo[k], list_entry(f, (y + o / 5 - 2) * w + x + o % 5 - 2), out.item, &out.n);
		}
		block_list_flush(out, p.list_b, count);
	}
}

/* ------------------------------------------------------------------------ */
/* Detection (src/filters.cuh:86-193, src/cleaner.cuh:22-37,102-126)         */
/*
 * One detection pass is four stages:
 *   significance  S = L / (2 N) over a region and a 2-pixel halo in LDS;
 *                 S' = S - med5(S) only where S > min(t, t_n) (S' <= S because
 *                 S >= 0), written to a sparse plane; seeds (S' > t) go to a
 *                 list. The first pass is dense: it walks runs of run_tiles
 *                 tiles down a tile column, so rows shared by two tiles are
 *                 computed once, and it makes C from the input as it goes (the
 *                 prologue). The block pass processes the BLK x BLK blocks
 *                 marked dirty, one wave per block: in pass 0 those of the
 *                 tiles whose S waits for the hole fill, later those near a
 *                 pixel the last replacement changed.
 *   fine          per seed: F = max((m3 - med7(m3)) / N, 0.01), m3 = med3(C);
 *                 S' / F > contrast puts the seed on the candidate list (two
 *                 launches: fine, one lane per seed whose 9 x 9 samples lie
 *                 inside the frame, then fine_edge, one wave per seed near an
 *                 edge)
 *   grow_first    per candidate: neighbours with S' > t join the growth list
 *   grow_second   per grown pixel: neighbours with S' > t_n are flagged; a
 *                 pixel flagged for the first time is new (accumulate)
 * A block that a pass skips had no change of the clean image within REACH of
 * its pixels since the pass that last processed it, so its S' values and the
 * growth of its candidates are still those of that pass: the sparse passes
 * give the reference's dense result exactly.
 */

constexpr int HS = 2;                      /* halo of S: the 5x5 median */
constexpr int SW = TILE_W + 2 * HS + 2;    /* 64: one wave of C columns */
constexpr int STEP_ROWS = TILE_H + 2 * HS; /* rows of S the medians of one tile read */
constexpr int RING = 40;                   /* rows of S kept in LDS */
static_assert(SW == 64, "the significance region is one wave wide");
static_assert(STEP_ROWS % 4 == 0 && TILE_H % 4 == 0, "four waves share the rows of a step");
static_assert(RING >= STEP_ROWS && RING <= 64, "the ring holds the rows of one tile");

/* The 5x5 median of S, v the window in row-major order. Without a NaN every
 * correct selection gives the same value, so the min/max network does. A NaN
 * (S made from infinite samples) stops the reference's compare-exchange
 * (src/sort.cuh:18-36: odd-even transposition passes that swap only when
 * b < a), so its result depends on where the NaN lies; such a window takes
 * that sort. */
__device__ __forceinline__ double
s_median25(const double *v)
{
	bool nan = false;
#pragma unroll
	for (int k = 0; k < 25; ++k)
		nan |= isnan(v[k]);
	double m;
	if (!nan) {
		median25(v, &m);
		return m;
	}
	double s[25];
#pragma unroll
	for (int k = 0; k < 25; ++k)
		s[k] = v[k];
#pragma unroll
	for (int pass = 0; pass < 13; ++pass)
#pragma unroll
		for (int start = 0; start < 2; ++start)
#pragma unroll
			for (int i = start; i + 1 < 25; i += 2) {
				const double a = s[i], b = s[i + 1];
				const bool swap = b < a;
				s[i] = swap ? b : a;
				s[i + 1] = swap ? a : b;
			}
	return s[12];
}

/* max(v, 0) as the reference writes it (v < 0 ? 0 : v), which keeps a NaN.
 * Without infinite samples (TAME) no NaN can arise and one max instruction
 * does; the two differ at most in the sign of a zero, which no threshold test
 * sees. */
template <bool TAME>
__device__ __forceinline__ double
clip(double v)
{
	return TAME ? fmax(v, 0.0) : (v < 0 ? 0.0 : v);
}

/* Laplacian of the 2x-replicated frame (src/filters.cuh:86-105) from the
 * quarter samples (x / 4) of the pixel and of its up, down, left and right
 * neighbours. EXACT is the L.A.Cosmic (scipy) accumulation order, otherwise
 * the reference CUDA order; narrow marks a one-pixel-wide frame, whose 2x2
 * block sum numpy runs sequentially. A negated first operand stands for the
 * leading 0 - x of the scipy sum: that changes only the sign of a zero. */
template <bool EXACT, bool TAME>
__device__ __forceinline__ double
laplacian_q(double c, double u, double d, double l, double r, bool narrow)
{
	if (!EXACT) {
		double up = c - u, down = c - d, left = c - l, right = c - r;
		return (clip<TAME>(up + left) + clip<TAME>(up + right)) +
		       (clip<TAME>(down + left) + clip<TAME>(down + right));
	}
	double c4 = 4 * c;
	double ul = clip<TAME>((((-u - l) + c4) - c) - c);
	double ur = clip<TAME>((((-u - c) + c4) - r) - c);
	double dl = clip<TAME>((((-c - l) + c4) - c) - d);
	/* (-c - c) + 4c is 2c exactly for finite c (sign of a zero aside); an
	 * infinite c makes it NaN, which only a background can bring */
	double dr = clip<TAME>(((TAME ? c + c : (-c - c) + c4) - r) - d);
	if (narrow)
		return ((ul + ur) + dl) + dr;
	return (ul + ur) + (dl + dr);
}

/* A noise value that breaks S >= 0 (zero or negative; NaN only makes S NaN). */
__device__ __forceinline__ bool
invalid_noise(double n)
{
	return !(n > 0) && !isnan(n);
}

/* Laplacian sample of the clean image at (x, y), at most one step outside the
 * frame for in-frame centres: mirror/reflect/nearest repeat the edge pixel,
 * wrap wraps, constant reads 0 (src/filters.cuh:72-105). Far-outside
 * coordinates (halo slots that are later overwritten) stay in bounds. */
__device__ __forceinline__ double
lap_sample(const double *c, int x, int y, int w, int h, int mode)
{
	if (x >= 0 && x < w && y >= 0 && y < h)
		return c[y * w + x];
	if (mode == CUSMIC_CONSTANT)
		return 0.0;
	if (mode == CUSMIC_WRAP) {
		x = border_index(x, w, mode);
		y = border_index(y, h, mode);
	} else {
		x = min(max(x, 0), w - 1);
		y = min(max(y, 0), h - 1);
	}
	return c[y * w + x];
}

/* Median-filter sample under the border mode (scipy semantics). */
__device__ __forceinline__ double
med_sample(const double *c, int x, int y, int w, int h, int mode)
{
	int bx = border_index(x, w, mode), by = border_index(y, h, mode);
	return bx < 0 || by < 0 ? 0.0 : c[by * w + bx];
}

/* In-bounds index for a map read at a halo slot. */
__device__ __forceinline__ int
map_index(int x, int y, int w, int h, int mode)
{
	if (mode == CUSMIC_WRAP) {
		x = border_index(x, w, mode);
		y = border_index(y, h, mode);
	} else {
		x = min(max(x, 0), w - 1);
		y = min(max(y, 0), h - 1);
	}
	return y * w + x;
}

/* C at index i of a frame, made from the input as the prologue of
 * src/image.cuh:24-48 makes it: C = (hole ? 0 : D) + bg, where a nonfinite
 * data value is a hole (its C changes again in the hole fill). */
template <bool BG>
__device__ __forceinline__ double
input_c(double d, double b, bool has_bg, bool &hole)
{
	const bool hl = !isfinite(d);
	hole |= hl;
	double c = hl ? 0.0 : d;
	if (BG && has_bg)
		c += 1 * b;
	return c;
}

template <bool BG>
__device__ __forceinline__ double
input_at(const Frame &fr, double bg, bool has_bg, long long i, bool &hole)
{
	return input_c<BG>(fr.data[i], BG && fr.background ? fr.background[i] : bg, has_bg, hole);
}

/* input_at under the Laplacian's border rule (lap_sample). */
template <bool BG>
__device__ __forceinline__ double
lap_input(const Frame &fr, double bg, bool has_bg, int x, int y, int mode, bool &hole)
{
	const int w = fr.w, h = fr.h;
	if (!(x >= 0 && x < w && y >= 0 && y < h)) {
		if (mode == CUSMIC_CONSTANT)
			return 0.0;
		if (mode == CUSMIC_WRAP) {
			x = border_index(x, w, mode);
			y = border_index(y, h, mode);
		} else {
			x = min(max(x, 0), w - 1);
			y = min(max(y, 0), h - 1);
		}
	}
	return input_at<BG>(fr, bg, has_bg, (long long)y * w + x, hole);
}

/* input_at under the median's border rule (med_sample). */
template <bool BG>
__device__ __forceinline__ double
med_input(const Frame &fr, double bg, bool has_bg, int x, int y, int mode, bool &hole)
{
	const int bx = border_index(x, fr.w, mode), by = border_index(y, fr.h, mode);
	return bx < 0 || by < 0 ? 0.0 : input_at<BG>(fr, bg, has_bg, (long long)by * fr.w + bx, hole);
}

/* frames[f] through scalar loads, for a uniform f: the table is written
 * before the kernel that reads it starts, so the scalar cache holds no stale
 * copy. */
__device__ __forceinline__ Frame
frame_scalar(const Frame *frames, int f)
{
	static_assert(sizeof(Frame) % 4 == 0, "the table is read in 32-bit words");
	using cword = const __attribute__((address_space(4))) unsigned;
	Frame fr;
	unsigned *dst = reinterpret_cast<unsigned *>(&fr);
	cword *src = (cword *)(frames + f);
#pragma unroll
	for (unsigned k = 0; k < sizeof(Frame) / 4; ++k)
		dst[k] = src[k];
	return fr;
}

/* Raw buffer access: a resource whose base is a uniform address, a lane's byte
 * offset in a vector register and a row's byte offset in a scalar one, so that
 * a row's loads and stores need no vector address arithmetic. The record
 * count is the largest offset the hardware accepts; callers keep offsets far
 * below it. Word 3 is the gfx9 raw-buffer format (32-bit data format, no
 * swizzle, no index). */
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t
buffer_at(const void *base)
{
	return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, 0x7fffffff, 0x00020000);
}

__device__ __forceinline__ double
buffer_load_f64(__amdgpu_buffer_rsrc_t r, unsigned lane_off, unsigned row_off)
{
	return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, row_off, 0));
}

/* Non-temporal store (cache policy bit 1: nt). */
__device__ __forceinline__ void
buffer_store_f64_nt(double v, __amdgpu_buffer_rsrc_t r, unsigned lane_off, unsigned row_off)
{
	__builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, lane_off, row_off, 2);
}

/* Two neighbouring doubles in one 16-byte access. */
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void
buffer_load_f64x2(__amdgpu_buffer_rsrc_t r, unsigned lane_off, unsigned row_off, double &x, double &y)
{
	const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, lane_off, row_off, 0);
	x = __hiloint2double(int(v.y), int(v.x));
	y = __hiloint2double(int(v.w), int(v.z));
}

__device__ __forceinline__ void
buffer_store_f64x2_nt(double x, double y, __amdgpu_buffer_rsrc_t r, unsigned lane_off, unsigned row_off)
{
	const u32x4 v = {unsigned(__double2loint(x)), unsigned(__double2hiint(x)), unsigned(__double2loint(y)),
		unsigned(__double2hiint(y))};
	__builtin_amdgcn_raw_buffer_store_b128(v, r, lane_off, row_off, 2);
}


/* Lane i takes lane i - 1's value (wave_shr:1); lane 0 gets 0 (bound_ctrl,
 * so no register has to hold an old value). */
__device__ __forceinline__ double
wave_shift_right(double v)
{
	const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
	const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
	return __hiloint2double(hi, lo);
}

/* Lane i takes lane i + 1's value (wave_shl:1); lane 63 gets 0. */
__device__ __forceinline__ double
wave_shift_left(double v)
{
	const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
	const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
	return __hiloint2double(hi, lo);
}

constexpr int SEED_BUF = 512; /* seeds buffered per workgroup between flushes */

/* Tile geometry of a first pass. The noise-model pass takes TILE_W x TILE_H
 * tiles and a 64-column window, one column per lane (the window's columns
 * are frame columns X0 - OFF ..). The error-map pass takes TILE_W2 x TILE_H
 * tiles and a 128-column window, two neighbouring columns per lane: its loads
 * and stores move 16 bytes per lane, which that pass's memory access pattern
 * is bound by. Window columns 0 and WIN - 1 carry C only; S is read at
 * columns 1 .. WIN - 2. */
template <bool WIDE>
struct TileGeom {
	static constexpr int TW = WIDE ? TILE_W2 : TILE_W; /* owned columns */
	static constexpr int WIN = WIDE ? 128 : SW;        /* window columns */
	static constexpr int OFF = WIDE ? 4 : HS + 1;      /* first owned window column */
	static constexpr int CBITS = WIDE ? 7 : 6;         /* bits of a window column */
	static_assert(OFF >= HS + 1 && OFF + TW + HS + 1 <= WIN, "the window holds the tile, S halo and ring");
};
static_assert(TILE_W2 % BLK == 0 && TILE_W2 % 2 == 0, "wide tiles cover whole blocks and 16-byte pairs");

template <bool WIDE>
struct SigShared {
	using G = TileGeom<WIDE>;
	/* S, row r of a run in slot r % RING (16-byte aligned: the wide row loop
	 * writes two columns at once) */
	__attribute__((aligned(16))) double s[RING][G::WIN];
	int step_bits[2][BLOCK / 64];    /* per step parity and wave: bit 0 an invalid
					  * noise value among its rows, bit 1 a hole
					  * among the samples of C they read */
	uint16_t alist[G::TW * TILE_H];  /* screened pixels of the tile: row << CBITS | column */
	unsigned long long seeds[SEED_BUF];
	int na, nseed, item;
	unsigned flush_base;
	int own[3];                      /* the run's holes, allowed pixels, donors */
};

/* This lane's counts of the pixels of a run it owns (the run's pixels less
 * its holes and masked pixels are the allowed ones). */
struct OwnCounts {
	int holes;            /* holes (skipped tiles) */
	int masked;           /* masked pixels that are no hole */
	int nonfinite;        /* pixels that are no hole with a nonfinite C */
	int masked_nonfinite; /* masked ones of those */
	bool infinite;        /* an infinite C */
};

/* A workgroup barrier that orders LDS only. The first pass's global stores
 * are read by later kernels alone, and no two of its waves store to the same
 * address, so its barriers need not wait for them (a full barrier would wait
 * for every store in flight). */
__device__ __forceinline__ void
lds_barrier()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

/* Move the buffered seeds of this workgroup to the global list. */
template <class SH>
__device__ void
flush_seeds(SH &sh, unsigned long long *list, unsigned int *count)
{
	lds_barrier();
	const int n = sh.nseed;
	if (!threadIdx.x && n)
		sh.flush_base = atomicAdd(count, (unsigned)n);
	lds_barrier();
	for (int k = threadIdx.x; k < n; k += BLOCK)
		list[sh.flush_base + k] = sh.seeds[k];
	lds_barrier();
	if (!threadIdx.x)
		sh.nseed = 0;
}

/* Store one row of S (this lane's slot). The screen of a row is taken from
 * the ring after the step's barrier. */
template <class SH>
__device__ __forceinline__ void
emit_row(SH &sh, int &slot, int lane, double v)
{
	sh.s[slot][lane] = v;
	slot = slot + 1 == RING ? 0 : slot + 1;
}

/* The noise and hole flags of one wave's rows of a step: one word, written
 * once after its rows (flags: this lane saw an invalid noise value / a hole). */
__device__ __forceinline__ void
emit_step_bits(int *word, int lane, bool bad, bool hole)
{
	const int bits = int(__ballot(bad) != 0) | int(__ballot(hole) != 0) << 1;
	if (!lane)
		*word = bits;
}

/* The prologue's work at a row of owned pixels (own: this lane's pixel i
 * belongs to the run): C (unless the caller stores it later) and, with a
 * background, the counts of nonfinite C behind the donors (a hole's C is the
 * finite background). Branch-free. Masked pixels are flagged and counted per
 * tile, holes where their tile skips. */
template <bool BG>
__device__ __forceinline__ void
own_row(const Params &p, const Frame &fr, bool own, long long i, double c, bool store_c, OwnCounts &acc)
{
	if (store_c && own)
		p.clean[fr.off + i] = c;
	if (BG) {
		acc.nonfinite += own && !isfinite(c);
		acc.infinite |= own && isinf(c);
	}
}

/* The samples of C that one step of the noise-model first pass reads, in LDS:
 * region row k is run row rS0 - HS + k (rS0: the step's first S row), region
 * column c is frame column X0 - 2 HS - 1 + c, both under the median's border
 * mode (CONSTANT: 0). The first step has STEP_ROWS S rows and reads the most:
 * its last wave's last four-row strip ends three rows below them. */
constexpr int CREG_COLS = SW + 2 * HS;                /* 68 */
constexpr int CREG_ROWS = STEP_ROWS + 2 * HS + 4;     /* 44 */
static_assert(CREG_ROWS <= 64, "one bit per region row");

struct ModelRegion {
	double c[CREG_ROWS][CREG_COLS];
	unsigned long long hole_lo[CREG_ROWS]; /* columns 0..63 that are holes */
	unsigned long long hole_hi[CREG_ROWS]; /* columns 64.. that are holes */
};

/* Load region rows [0, nrows) for the frame rows from yc0 (whole workgroup;
 * wave w loads rows w, w + 4, ..., lane l columns l and 64 + l). */
template <bool BG>
__device__ __forceinline__ void
load_region(ModelRegion &R, const Params &p, const Frame &fr, int X0, int yc0, int nrows)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int mode = p.mode, w = fr.w, h = fr.h;
	const bool has_bg = BG && (fr.background || !isnan(p.bg));
	const int bx0 = border_index(X0 - 2 * HS - 1 + lane, w, mode);
	const bool second = lane < CREG_COLS - 64;
	const int bx1 = second ? border_index(X0 - 2 * HS - 1 + 64 + lane, w, mode) : -1;
	/* every load of the wave's rows first, then C, the LDS stores and the
	 * hole ballots: one memory round trip per step instead of one per row */
	constexpr int MAXK = (CREG_ROWS + BLOCK / 64 - 1) / (BLOCK / 64);
	double d0[MAXK], d1[MAXK], g0[MAXK], g1[MAXK];
	bool in0[MAXK], in1[MAXK];
	long long i0[MAXK], i1[MAXK];
#pragma unroll
	for (int j = 0; j < MAXK; ++j) {
		const int k = wave + j * (BLOCK / 64);
		const int by = k < nrows ? border_index(yc0 + k, h, mode) : -1;
		in0[j] = by >= 0 && bx0 >= 0;
		in1[j] = by >= 0 && bx1 >= 0;
		i0[j] = (long long)by * w + bx0;
		i1[j] = (long long)by * w + bx1;
		d0[j] = in0[j] ? fr.data[i0[j]] : 0.0;
		d1[j] = in1[j] ? fr.data[i1[j]] : 0.0;
		g0[j] = BG && fr.background && in0[j] ? fr.background[i0[j]] : p.bg;
		g1[j] = BG && fr.background && in1[j] ? fr.background[i1[j]] : p.bg;
	}
#pragma unroll
	for (int j = 0; j < MAXK; ++j) {
		const int k = wave + j * (BLOCK / 64);
		if (k >= nrows)
			break;
		bool h0 = false, h1 = false;
		const double v0 = in0[j] ? input_c<BG>(d0[j], g0[j], has_bg, h0) : 0.0;
		const double v1 = in1[j] ? input_c<BG>(d1[j], g1[j], has_bg, h1) : 0.0;
		R.c[k][lane] = v0;
		if (second)
			R.c[k][64 + lane] = v1;
		const unsigned long long m0 = __ballot(h0), m1 = __ballot(h1);
		if (!lane) {
			R.hole_lo[k] = m0;
			R.hole_hi[k] = m1;
		}
	}
}

/* One row of S under the noise model from the region: rows holds the region's
 * rows around it (row k + 2 of the strip is this row), med its 5x5 median of
 * C, hm whether a sample of those rows is a hole, hc whether its own pixel
 * is. The Laplacian's neighbours are the strip's; at a frame edge the
 * mirror, reflect and nearest modes repeat the edge pixel for it (the
 * median's mirror mode reflects instead, as in the region). */
template <bool EXACT, bool BG>
__device__ __forceinline__ void
model_row_region(SigShared<false> &sh, int &slot, int lane, double t_min, unsigned long long cols,
	const Params &p, const Frame &fr, int gx, int gy, const double *rows, int k, double med, bool hm,
	bool hc, bool own, OwnCounts &acc, bool &any_bad, bool &any_hole)
{
	const int mode = p.mode, w = fr.w, h = fr.h;
	const double cc = rows[(k + 2) * 5 + 2];
	double u = rows[(k + 1) * 5 + 2], d = rows[(k + 3) * 5 + 2];
	double l = rows[(k + 2) * 5 + 1], r = rows[(k + 2) * 5 + 3];
	const bool in = gx >= 0 && gx < w && gy >= 0 && gy < h;
	if (in && mode != CUSMIC_WRAP && mode != CUSMIC_CONSTANT) {
		u = gy - 1 < 0 ? cc : u;
		d = gy + 1 >= h ? cc : d;
		l = gx - 1 < 0 ? cc : l;
		r = gx + 1 >= w ? cc : r;
	}
	const int mi = map_index(gx, gy, w, h, mode);
	const double gg = fr.gain ? fr.gain[mi] : fr.g;
	const double rn = fr.readnoise ? fr.readnoise[mi] : fr.rn;
	const double nz = noise_model(med, gg, rn);
	own_row<BG>(p, fr, own, (long long)gy * w + gx, cc, true, acc);
	/* lanes 0 and 63 carry C only: their S is never read */
	const bool s_lane = lane >= 1 && lane <= SW - 2;
	emit_row(sh, slot, lane,
		laplacian_q<EXACT, !BG>(cc * 0.25, u * 0.25, d * 0.25, l * 0.25, r * 0.25, w == 1) / (2 * nz));
	any_bad |= s_lane && (in || mode == CUSMIC_WRAP) && invalid_noise(nz);
	any_hole |= hm;
}

/* Whether a step of NR rows from frame row ya takes the in-frame row loop:
 * its columns, rows and their one-pixel ring lie inside the frame, and the
 * rows' byte offsets fit the buffer offsets. */
template <int NR>
__device__ __forceinline__ bool
inner_rows(int X0, int ya, int w, int h)
{
	return X0 - HS - 1 >= 0 && X0 - HS - 1 + SW <= w && ya - 1 >= 0 && ya + NR < h &&
		(long long)(NR + 2) * w * 8 < (1ll << 31);
}

/* One wave: S of run rows [r0, r0 + NR) into the ring (run row r is frame row
 * yb + r, lane l is column X0 - HS - 1 + l), with each row's screened columns
 * (among cols), whether it holds an invalid noise value and whether a sample
 * of C it read is a hole. The S that the medians read are those of lanes
 * 1..62 (columns X0 - HS .. X0 + TILE_W + HS - 1); lanes 0 and 63 hold the
 * C samples left and right of them, so every Laplacian neighbour is a
 * neighbouring lane's centre, and their own S is never read. C is made from
 * the input here and stored at once for the run's own pixels: the rows
 * [own0, own1) and the columns of cols (own_row). Storing inside the row loop
 * keeps no C in registers, which leaves them to loads five rows ahead. Without
 * a background no sample of C is infinite, so the Laplacian clips with one max
 * instruction; with one the reference's NaN-keeping clip is used. Slots
 * outside the frame are computed from in-bounds samples; the caller overwrites
 * the ones the medians read, unless the mode is WRAP, under which every
 * stencil commutes with the periodic extension. NR is a constant so that the
 * row loop unrolls around its ballots and the loads of several rows are in
 * flight. */
template <bool MODEL, bool EXACT, bool BG, bool BGMAP, int NR>
__device__ __forceinline__ void
significance_rows(SigShared<false> &sh, const Params &p, const Frame &fr, int f, int X0, int yb, int r0,
	double t_min, unsigned long long cols, int own0, int own1, OwnCounts &acc,
	const ModelRegion *R, int rS0, int *step_word)
{
	constexpr bool TAME = !BG;
	/* an opaque lane keeps the compiler from hoisting this step's per-lane
	 * addresses out of the step loop, where they would stay live across the
	 * medians */
	int lane = threadIdx.x & 63;
	asm volatile("" : "+v"(lane));
	const int mode = p.mode, w = fr.w, h = fr.h, nr = NR;
	const bool wrap = mode == CUSMIC_WRAP, narrow = w == 1;
	const bool has_bg = BG && (fr.background || !isnan(p.bg));
	const bool own_lane = (cols >> lane) & 1;
	/* lanes whose S the medians may read */
	const bool s_lane = lane >= 1 && lane <= SW - 2;
	const int gx = X0 - HS - 1 + lane, ya = yb + r0;
	int slot = r0 % RING;
	bool any_bad = false, any_hole = false;
	const bool inner = inner_rows<NR>(X0, ya, w, h);
	if ((!MODEL || fr.error) && inner) {
		/* buffer resources at the wave's first column in row -1 (data,
		 * background) and row 0 (noise, clean): row k of a plane is a scalar
		 * offset, the lane a fixed vector one */
		const long long rb = (long long)ya * w + (X0 - HS - 1);
		const unsigned W8 = unsigned(w) * 8u, lane8 = unsigned(lane) * 8u;
		const double *bp = BGMAP && fr.background ? fr.background : nullptr;
		const __amdgpu_buffer_rsrc_t rd = buffer_at(fr.data + (rb - w));
		const __amdgpu_buffer_rsrc_t re = buffer_at(fr.error + rb);
		const __amdgpu_buffer_rsrc_t rc = buffer_at(p.clean + fr.off + rb);
		const __amdgpu_buffer_rsrc_t rg = buffer_at(bp ? bp + (rb - w) : fr.data);
		/* the samples of row r: the pixel below and the noise (and the
		 * background below), fetched PF rows ahead */
		constexpr int PF = BGMAP ? 4 : 5;
		static_assert(PF <= NR, "the first rows are loaded before the loop");
		double dn[NR], nz[NR], bn[NR];
#define FETCH_ROW(k)                                                                       \
	do {                                                                               \
		dn[k] = buffer_load_f64(rd, lane8, unsigned((k) + 2) * W8);                \
		nz[k] = buffer_load_f64(re, lane8, unsigned(k) * W8);                      \
		bn[k] = p.bg;                                                              \
		if (BG && bp)                                                              \
			bn[k] = buffer_load_f64(rg, lane8, unsigned((k) + 2) * W8);        \
	} while (0)
		/* Without a background C is the input itself, also at a hole: a
		 * hole's C is written again by the hole fill before anything reads
		 * it, and the output takes the input there. So every sample of a
		 * row is used as loaded, and a hole among them shows as a NaN in
		 * the sum d * 0 over every sample this lane loaded (lanes 0 and 63:
		 * also the rows above and below theirs, which only makes the flag
		 * cautious). */
		double hsum = 0.0;
		bool hu = false, hc0 = false;
		const double du = buffer_load_f64(rd, lane8, 0), d0 = buffer_load_f64(rd, lane8, W8);
#pragma unroll
		for (int r = 0; r < PF; ++r)
			FETCH_ROW(r);
		double qu, cc;
		if (TAME) {
			hsum = __builtin_fma(du, 0.0, __builtin_fma(d0, 0.0, hsum));
			qu = du * 0.25;
			cc = d0;
		} else {
			qu = input_c<BG>(du, bp ? buffer_load_f64(rg, lane8, 0) : p.bg, has_bg, hu) * 0.25;
			cc = input_c<BG>(d0, bp ? buffer_load_f64(rg, lane8, W8) : p.bg, has_bg, hc0);
		}
		double qc = cc * 0.25;
		/* fully unrolled and branch-free (constant NR, convergent ballots):
		 * the loads run PF rows ahead */
#pragma unroll
		for (int r = 0; r < NR; ++r) {
			if (r + PF < NR)
				FETCH_ROW(r + PF);
			bool hd = false;
			double cd;
			if (TAME) {
				cd = dn[r];
				hsum = __builtin_fma(cd, 0.0, hsum);
			} else {
				cd = input_c<BG>(dn[r], bn[r], has_bg, hd);
			}
			/* lanes 0 and 63 keep their own centre: their S is not read */
			const double ql = wave_shift_right(qc), qr = wave_shift_left(qc);
			const double qd = cd * 0.25;
			/* C at the run's own pixels. Non-temporal: C is read again only by
			 * later kernels, so it should not displace the input lines that
			 * neighbouring tiles read from L2. */
			if (own_lane && ya + r >= own0 && ya + r < own1)
				buffer_store_f64_nt(cc, rc, lane8, unsigned(r) * W8);
			if (BG) {
				const int gy = ya + r;
				own_row<BG>(p, fr, own_lane && gy >= own0 && gy < own1, rb + r * w + lane, cc, false, acc);
			}
			emit_row(sh, slot, lane, laplacian_q<EXACT, TAME>(qc, qu, qd, ql, qr, false) / (2 * nz[r]));
			any_bad |= s_lane && nz[r] <= 0;
			any_hole |= hd;
			qu = qc;
			qc = qd;
			cc = cd;
		}
		any_hole |= hu || hc0 || hsum != 0.0;
#undef FETCH_ROW
	} else if (!MODEL || fr.error) {
		bool hu = false, hc = false;
		double qu = lap_input<BG>(fr, p.bg, has_bg, gx, ya - 1, mode, hu) * 0.25;
		double cc = lap_input<BG>(fr, p.bg, has_bg, gx, ya, mode, hc);
#pragma unroll 1
		for (int r = 0; r < nr; ++r) {
			const int gy = ya + r;
			bool hd = false, hs = false;
			const double cd = lap_input<BG>(fr, p.bg, has_bg, gx, gy + 1, mode, hd);
			const double ql = lap_input<BG>(fr, p.bg, has_bg, gx - 1, gy, mode, hs) * 0.25;
			const double qr = lap_input<BG>(fr, p.bg, has_bg, gx + 1, gy, mode, hs) * 0.25;
			const double nz = fr.error[map_index(gx, gy, w, h, mode)];
			const bool in = gx >= 0 && gx < w && gy >= 0 && gy < h;
			own_row<BG>(p, fr, own_lane && gy >= own0 && gy < own1, (long long)gy * w + gx, cc, true, acc);
			emit_row(sh, slot, lane, laplacian_q<EXACT, TAME>(cc * 0.25, qu, cd * 0.25, ql, qr, narrow) / (2 * nz));
			any_bad |= s_lane && (in || wrap) && invalid_noise(nz);
			any_hole |= hc || (s_lane && (hu || hd || hs));
			qu = cc * 0.25;
			cc = cd;
			hu = hc;
			hc = hd;
		}
	} else {
		/* noise model: C from the step's region (loaded by the caller), N
		 * from strip-shared 5x5 medians, four rows at a time; the region's
		 * rows with a hole, one bit each */
		const unsigned long long hole_rows =
			__ballot(lane < CREG_ROWS && (R->hole_lo[lane] | R->hole_hi[lane]) != 0);
		const int kb = r0 - rS0; /* region row of the strip's first row */
#pragma unroll 1
		for (int g = 0; g < nr; g += 4) {
			double rows[8 * 5], med[4];
#pragma unroll
			for (int j = 0; j < 8; ++j)
#pragma unroll
				for (int c = 0; c < 5; ++c)
					rows[j * 5 + c] = R->c[kb + g + j][lane + c];
			strip_median5x4(rows, med);
			const int n_out = min(4, nr - g);
#pragma unroll
			for (int k = 0; k < 4; ++k) {
				if (k >= n_out)
					break;
				const int kc = kb + g + k + HS, gy = ya + g + k;
				const bool hm = (hole_rows >> (kb + g + k)) & 0x1full;
				const bool hc = lane + HS < 64 ? (R->hole_lo[kc] >> (lane + HS)) & 1
							      : (R->hole_hi[kc] >> (lane + HS - 64)) & 1;
				model_row_region<EXACT, BG>(sh, slot, lane, t_min, cols, p, fr, gx, gy, rows, k, med[k],
					hm, hc, own_lane && gy >= own0 && gy < own1, acc, any_bad, any_hole);
			}
		}
	}
	emit_step_bits(step_word, lane, any_bad, any_hole);
}

/* One wave: S of run rows [r0, r0 + NR) into the ring of a wide tile
 * (TileGeom<true>). Lane l holds window columns 2 l and 2 l + 1 (frame columns
 * X0 - 4 + 2 l and the next), so inside the frame its loads and stores move
 * 16 bytes: the pixel below and the noise value of both columns, PF rows
 * ahead, and C at the run's own pixels. The rows above and below a lane's
 * samples are its own earlier and later rows; the left neighbour of column
 * 2 l is lane l - 1's second column and the right one of 2 l + 1 lane
 * l + 1's first (wave shifts). Window columns 0 and 127 carry C only.
 *
 * Near a frame edge the same loop runs on mapped samples: a window column or
 * run row outside the frame reads the sample the Laplacian's border rule
 * gives (src/filters.cuh:72-105: the edge pixel for mirror, reflect and
 * nearest, the wrapped one for wrap, 0 for constant), with 8-byte loads where
 * a lane's two columns are not neighbours in memory. So every in-frame S is
 * exact; an out-of-frame slot is exact under wrap (the periodic extension)
 * and otherwise rewritten by the caller, as are its noise values, which can
 * only make the noise flag cautious. Frames whose byte size passes the
 * buffer offsets take the per-sample loop of significance_rows instead.
 * Otherwise as significance_rows: without a background C is the input itself
 * and a hole shows as a NaN in the sum d * 0; the step's noise and hole flags
 * are emitted at the end. */
template <bool EXACT, bool BG, bool BGMAP, int NR>
__device__ __forceinline__ void
significance_rows_wide(SigShared<true> &sh, const Params &p, const Frame &fr, int X0, int yb, int r0,
	int ncols, int own0, int own1, OwnCounts &acc, int *step_word)
{
	using G = TileGeom<true>;
	constexpr bool TAME = !BG;
	int lane = threadIdx.x & 63;
	asm volatile("" : "+v"(lane));
	const int mode = p.mode, w = fr.w, h = fr.h;
	const bool wrap = mode == CUSMIC_WRAP, constant = mode == CUSMIC_CONSTANT;
	const bool has_bg = BG && (fr.background || !isnan(p.bg));
	const int c0 = 2 * lane, c1 = c0 + 1;  /* window columns */
	const int x0 = X0 - G::OFF;           /* frame column of window column 0 */
	const int gx0 = x0 + c0, gx1 = gx0 + 1;
	const bool own_x = c0 >= G::OFF && c0 < G::OFF + ncols;
	const bool own_y = c1 >= G::OFF && c1 < G::OFF + ncols;
	/* columns whose S the medians may read */
	const bool s_x = c0 >= 1, s_y = c1 <= G::WIN - 2;
	const int ya = yb + r0;
	bool any_bad = false, any_hole = false;
	if ((long long)w * h * 8 < (1ll << 31)) {
		/* the source column of each sample, and whether it is a zero of the
		 * constant mode; a lane's two columns are neighbours in memory when
		 * the window lies inside the frame */
		auto src = [&](int x) { return wrap ? border_index(x, w, mode) : min(max(x, 0), w - 1); };
		const bool zx = constant && (gx0 < 0 || gx0 >= w), zy = constant && (gx1 < 0 || gx1 >= w);
		const bool pairs = x0 >= 0 && x0 + G::WIN <= w;
		const unsigned vx = unsigned(pairs ? gx0 : src(gx0)) * 8u, vy = unsigned(pairs ? gx1 : src(gx1)) * 8u;
		const unsigned W8 = unsigned(w) * 8u;
		/* the byte offset of a run row's source row; zero: the row is outside
		 * the frame under the constant mode */
		auto row_off = [&](int y, bool &zero) -> unsigned {
			zero = constant && (y < 0 || y >= h);
			return unsigned(wrap ? border_index(y, h, mode) : min(max(y, 0), h - 1)) * W8;
		};
		const double *bp = BGMAP && fr.background ? fr.background : nullptr;
		const __amdgpu_buffer_rsrc_t rd = buffer_at(fr.data);
		const __amdgpu_buffer_rsrc_t re = buffer_at(fr.error);
		const __amdgpu_buffer_rsrc_t rc = buffer_at(p.clean + fr.off);
		const __amdgpu_buffer_rsrc_t rg = buffer_at(bp ? bp : fr.data);
		auto load2 = [&](__amdgpu_buffer_rsrc_t r, unsigned roff, double &a, double &b) {
			if (pairs) {
				buffer_load_f64x2(r, vx, roff, a, b);
			} else {
				a = buffer_load_f64(r, vx, roff);
				b = buffer_load_f64(r, vy, roff);
			}
		};
		/* C from a loaded sample pair of a row (zero: constant-mode row
		 * outside the frame) */
		bool hh = false;
		double hsum = 0.0;
		auto make_c = [&](double d0, double d1, double b0, double b1, bool zero, double &o0, double &o1) {
			if (TAME) {
				hsum = __builtin_fma(d0, 0.0, __builtin_fma(d1, 0.0, hsum));
				o0 = d0;
				o1 = d1;
			} else {
				o0 = input_c<BG>(d0, b0, has_bg, hh);
				o1 = input_c<BG>(d1, b1, has_bg, hh);
			}
			if (constant) {
				o0 = zero || zx ? 0.0 : o0;
				o1 = zero || zy ? 0.0 : o1;
			}
		};
		constexpr int PF = BGMAP ? 3 : 4;
		static_assert(PF <= NR, "the first rows are loaded before the loop");
		double dx[NR], dy[NR], nx[NR], ny[NR], bx[NR], by[NR];
		bool zr[NR];
#define FETCH_ROW(k)                                                                       \
	do {                                                                               \
		bool zn;                                                                   \
		const unsigned ro = row_off(ya + (k) + 1, zr[k]);                          \
		load2(rd, ro, dx[k], dy[k]);                                               \
		load2(re, row_off(ya + (k), zn), nx[k], ny[k]);                            \
		bx[k] = by[k] = p.bg;                                                      \
		if (BG && bp)                                                              \
			load2(rg, ro, bx[k], by[k]);                                       \
	} while (0)
		bool zu, zc;
		const unsigned ru = row_off(ya - 1, zu), rcur = row_off(ya, zc);
		double ux, uy, cx, cy, gux = p.bg, guy = p.bg, gcx = p.bg, gcy = p.bg;
		load2(rd, ru, ux, uy);
		load2(rd, rcur, cx, cy);
		if (BG && bp) {
			load2(rg, ru, gux, guy);
			load2(rg, rcur, gcx, gcy);
		}
#pragma unroll
		for (int r = 0; r < PF; ++r)
			FETCH_ROW(r);
		make_c(ux, uy, gux, guy, zu, ux, uy);
		make_c(cx, cy, gcx, gcy, zc, cx, cy);
		double qux = ux * 0.25, quy = uy * 0.25, qcx = cx * 0.25, qcy = cy * 0.25;
		int slot = r0 % RING;
#pragma unroll
		for (int r = 0; r < NR; ++r) {
			if (r + PF < NR)
				FETCH_ROW(r + PF);
			double ex, ey;
			make_c(dx[r], dy[r], bx[r], by[r], zr[r], ex, ey);
			const double qdx = ex * 0.25, qdy = ey * 0.25;
			const double qlx = wave_shift_right(qcy), qry = wave_shift_left(qcx);
			const int gy = ya + r;
			const bool own_r = gy >= own0 && gy < own1;
			/* C at the run's own pixels (inside the frame), non-temporal as in
			 * significance_rows */
			const unsigned roff = unsigned(gy) * W8, lo = unsigned(gx0) * 8u;
			if (own_r && own_x && own_y && pairs)
				buffer_store_f64x2_nt(cx, cy, rc, lo, roff);
			else {
				if (own_r && own_x)
					buffer_store_f64_nt(cx, rc, lo, roff);
				if (own_r && own_y)
					buffer_store_f64_nt(cy, rc, lo + 8u, roff);
			}
			if (BG) {
				const long long i = (long long)gy * w + gx0;
				own_row<BG>(p, fr, own_r && own_x, i, cx, false, acc);
				own_row<BG>(p, fr, own_r && own_y, i + 1, cy, false, acc);
			}
			const bool narrow = w == 1;
			const double sx = laplacian_q<EXACT, TAME>(qcx, qux, qdx, qlx, qcy, narrow) / (2 * nx[r]);
			const double sy = laplacian_q<EXACT, TAME>(qcy, quy, qdy, qcx, qry, narrow) / (2 * ny[r]);
			*reinterpret_cast<f64x2 *>(&sh.s[slot][c0]) = f64x2{sx, sy};
			slot = slot + 1 == RING ? 0 : slot + 1;
			any_bad |= (s_x && nx[r] <= 0) || (s_y && ny[r] <= 0);
			qux = qcx;
			quy = qcy;
			qcx = qdx;
			qcy = qdy;
			cx = ex;
			cy = ey;
		}
		any_hole |= hh || hsum != 0.0;
#undef FETCH_ROW
	} else {
#pragma unroll
		for (int e = 0; e < 2; ++e) {
			const int gx = gx0 + e, col = c0 + e;
			const bool own_c = e ? own_y : own_x, s_c = e ? s_y : s_x;
			int slot = r0 % RING;
			bool hu = false, hc = false;
			double qu = lap_input<BG>(fr, p.bg, has_bg, gx, ya - 1, mode, hu) * 0.25;
			double cc = lap_input<BG>(fr, p.bg, has_bg, gx, ya, mode, hc);
#pragma unroll 1
			for (int r = 0; r < NR; ++r) {
				const int gy = ya + r;
				bool hd = false, hs = false;
				const double cd = lap_input<BG>(fr, p.bg, has_bg, gx, gy + 1, mode, hd);
				const double ql = lap_input<BG>(fr, p.bg, has_bg, gx - 1, gy, mode, hs) * 0.25;
				const double qr = lap_input<BG>(fr, p.bg, has_bg, gx + 1, gy, mode, hs) * 0.25;
				const double nz = fr.error[map_index(gx, gy, w, h, mode)];
				const bool in = gx >= 0 && gx < w && gy >= 0 && gy < h;
				own_row<BG>(p, fr, own_c && gy >= own0 && gy < own1, (long long)gy * w + gx, cc, true, acc);
				sh.s[slot][col] = laplacian_q<EXACT, TAME>(cc * 0.25, qu, cd * 0.25, ql, qr, w == 1) / (2 * nz);
				slot = slot + 1 == RING ? 0 : slot + 1;
				any_bad |= s_c && (in || wrap) && invalid_noise(nz);
				any_hole |= hc || (s_c && (hu || hd || hs));
				qu = cc * 0.25;
				cc = cd;
				hu = hc;
				hc = hd;
			}
		}
	}
	emit_step_bits(step_word, lane, any_bad, any_hole);
}

/* The first detection pass, which also does the prologue's work (C, flags,
 * counts, hole list) for the pixels of each run. A tile whose samples of C
 * hold a hole gets no S' here: its S depends on the hole fill, which runs
 * after this pass, so its blocks go to the dirty list of the block pass that
 * follows the fill (tile_skip marks it). Elsewhere S' is exact, and the
 * screened pixels carry SCR. */
template <bool MODEL, bool EXACT, bool BG, bool BGMAP>
__device__ __forceinline__ void
significance_pass(SigShared<!MODEL> &sh, ModelRegion *region, const Params &p, int iter)
{
	/* the error-map pass takes wide tiles (TileGeom) */
	constexpr bool WIDE = !MODEL;
	using G = TileGeom<WIDE>;
	/* uniform, so that a wave's rows and row addresses live in scalar
	 * registers */
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int mode = p.mode;
	const bool wrap = mode == CUSMIC_WRAP;
	const double t_cr = p.threshold, t_min = fmin(p.threshold, p.neighbor);
	const unsigned dtag = tag_stamp(iter);
	unsigned int *seeds = &p.nlist[nl_seed(p, iter)];
	/* The runs are split into XCDS contiguous ranges, one queue each; a
	 * workgroup claims from the queue of its XCD (workgroups are dispatched to
	 * the XCDs in turn) and, once that is empty, from the others in turn.
	 * Consecutive runs are neighbouring tile columns, so the runs that share
	 * a halo column start together under one L2. Every queue ends empty, so
	 * every workgroup exits. */
	int tries = 0;
	auto claim = [&]() -> int {
		while (tries < XCDS) {
			const unsigned x = (blockIdx.x + tries) % XCDS;
			const unsigned lo = unsigned((long long)p.nruns * x / XCDS);
			const unsigned hi = unsigned((long long)p.nruns * (x + 1) / XCDS);
			const unsigned k = atomicAdd(&p.queue[x], 1u);
			if (k < hi - lo)
				return int(lo + k);
			++tries;
		}
		return -1;
	};
	int next = -1;
	unsigned long long work = 0; /* in-frame pixels processed, thread 0 */
	if (!threadIdx.x) {
		next = claim();
		sh.nseed = 0;
		sh.own[0] = sh.own[1] = sh.own[2] = 0;
	}

	for (;;) {
		lds_barrier();
		if (!threadIdx.x) {
			sh.item = next;
			if (next >= 0)
				next = claim();
		}
		lds_barrier();
		const int item = __builtin_amdgcn_readfirstlane(sh.item);
		if (item < 0)
			break;
		/* the run: tile column tx, tile rows [ty0, ty0 + K) of frame f */
		/* scalar loads (the tables are written before this kernel starts), so
		 * the frame's pointers and sizes are uniform registers that buffer
		 * resources can take */
		const int f = *(const __attribute__((address_space(4))) int *)(p.run_frame + item);
		const Frame fr = frame_scalar(p.frames, f);
		const int lr = item - fr.run0, tx = lr % fr.tiles_x, ty0 = lr / fr.tiles_x * p.run_tiles;
		const int K = min(p.run_tiles, fr.tiles_y - ty0);
		const int w = fr.w, h = fr.h, X0 = tx * G::TW, Y0 = ty0 * TILE_H, yb = Y0 - HS;
		const int ncols = min(G::TW, w - X0), own1 = min(Y0 + TILE_H * K, h);
		/* the tile's columns in the narrow window (one lane each) */
		const unsigned long long cols = WIDE ? 0 : ((1ull << ncols) - 1) << (HS + 1);
		const bool has_bg = BG && (fr.background || !isnan(p.bg));
		OwnCounts acc{0, 0, 0, 0, false};

		for (int s = 0; s < K; ++s) {
			/* per-step opaque thread index, as in significance_rows */
			int tid = threadIdx.x;
			asm volatile("" : "+v"(tid));
			const int lane = tid & 63;
			/* 1. S rows: the first tile needs rows [0, STEP_ROWS) of the run,
			 * each later one TILE_H more */
			/* the step's first S row; under the noise model its C samples
			 * go to the region first */
			const int rS0 = s == 0 ? 0 : TILE_H * s + 2 * HS;
			if (s > 0)
				lds_barrier(); /* the last tile's medians are done with the ring */
			if (MODEL && !fr.error) {
				load_region<BG>(*region, p, fr, X0, yb + rS0 - HS, s == 0 ? CREG_ROWS : TILE_H + 2 * HS);
				lds_barrier();
			}
			if constexpr (WIDE) {
				if (s == 0)
					significance_rows_wide<EXACT, BG, BGMAP, STEP_ROWS / 4>(sh, p, fr, X0, yb,
						wave * (STEP_ROWS / 4), ncols, Y0, own1, acc, &sh.step_bits[s & 1][wave]);
				else
					significance_rows_wide<EXACT, BG, BGMAP, TILE_H / 4>(sh, p, fr, X0, yb,
						TILE_H * s + 2 * HS + wave * (TILE_H / 4), ncols, Y0, own1, acc,
						&sh.step_bits[s & 1][wave]);
			} else if (s == 0) {
				const int r0 = wave * (STEP_ROWS / 4);
				significance_rows<MODEL, EXACT, BG, BGMAP, STEP_ROWS / 4>(sh, p, fr, f, X0, yb, r0, t_min,
					cols, Y0, own1, acc, region, rS0, &sh.step_bits[s & 1][wave]);
			} else {
				const int r0 = TILE_H * s + 2 * HS + wave * (TILE_H / 4);
				significance_rows<MODEL, EXACT, BG, BGMAP, TILE_H / 4>(sh, p, fr, f, X0, yb, r0, t_min,
					cols, Y0, own1, acc, region, rS0, &sh.step_bits[s & 1][wave]);
			}
			const int ty = Y0 + TILE_H * s; /* first frame row of the tile */
			const int nrows = min(TILE_H, h - ty);
			if (!tid) {
				work += (unsigned long long)(ncols * nrows);
				sh.na = 0;
			}
			lds_barrier();

			/* 2. out-of-frame slots the medians read: the mapped in-frame slot
			 * (median border mode), or 0 for CONSTANT */
			const int r_lo = TILE_H * s, y_lo = yb + r_lo;
			const bool edge = X0 - HS < 0 || X0 + G::TW + HS > w || y_lo < 0 || y_lo + STEP_ROWS > h;
			if (edge && !wrap) {
				for (int k = tid; k < G::WIN * STEP_ROWS; k += BLOCK) {
					int sx = k & (G::WIN - 1), sr = k / G::WIN, x = X0 - G::OFF + sx, y = y_lo + sr;
					if ((x >= 0 && x < w && y >= 0 && y < h) || x < -2 || x > w + 1 || y < -2 || y > h + 1)
						continue;
					double v = 0.0;
					if (mode != CUSMIC_CONSTANT)
						v = sh.s[(border_index(y, h, mode) - yb) % RING][border_index(x, w, mode) - (X0 - G::OFF)];
					sh.s[(r_lo + sr) % RING][sx] = v;
				}
				lds_barrier();
			}

			/* 3. screened pixels of the tile: S' <= S, so only S > min(t, t_n)
			 * can pass a threshold; with an invalid noise value among the rows
			 * the medians read, every pixel is kept. With a hole among the
			 * samples, the tile's blocks are left to the block pass after the
			 * fill. */
			/* the tile's medians read rows [r_lo, r_lo + STEP_ROWS): this
			 * step's rows and the last four of the previous step, which its
			 * last wave computed */
			int rb = 0;
#pragma unroll
			for (int k = 0; k < BLOCK / 64; ++k)
				rb |= sh.step_bits[s & 1][k];
			if (s > 0)
				rb |= sh.step_bits[(s - 1) & 1][BLOCK / 64 - 1];
			const bool unsafe = (rb & 1) != 0, skip = (rb & 2) != 0;
			if (skip) {
				const int bx0 = X0 / BLK, by0 = ty / BLK;
				const int nbx = (X0 + ncols - 1) / BLK - bx0 + 1;
				const int nb = nbx * ((ty + nrows - 1) / BLK - by0 + 1); /* at most 9 x 4 */
				bool fresh = false;
				int q = 0;
				if (tid < nb) {
					q = fr.block0 + (by0 + tid / nbx) * fr.blocks_x + bx0 + tid % nbx;
					fresh = __atomic_load_n(&p.dirty_tag[q], __ATOMIC_RELAXED) != dtag &&
						atomicExch(&p.dirty_tag[q], dtag) != dtag;
				}
				list_push(fresh, list_entry(f, q), p.dirty_list, &p.nlist[nl_dirty(p, iter)]);
				if (!tid)
					p.tile_skip[fr.tile0 + (ty0 + s) * fr.tiles_x + tx] = 1;
				/* the tile's holes (only a skipped tile has any): the fill's
				 * list and count. The trip count is uniform per wave. */
				for (int k = tid; k < G::TW * TILE_H; k += BLOCK) {
					const int row = k / G::TW, col = k % G::TW, i = (ty + row) * w + X0 + col;
					const bool hole = row < nrows && col < ncols && !isfinite(fr.data[i]);
					if (hole)
						p.flags[fr.off + i] = uint8_t(HOLE | EXCL);
					acc.holes += hole;
					list_push_top(hole, list_entry(f, i), p.list + p.list_cap, &p.nlist[1]);
				}
			} else {
				for (int j = 0; j < TILE_H / 4; ++j) {
					int row = wave * (TILE_H / 4) + j;
					if (row >= nrows)
						break;
					/* the row's screen: S > min(t, t_n) at the tile's columns */
					const double *srow = sh.s[(r_lo + HS + row) % RING];
					if constexpr (WIDE) {
						const int c0 = 2 * lane, c1 = c0 + 1;
						const bool p0 = c0 >= G::OFF && c0 < G::OFF + ncols && (unsafe || srow[c0] > t_min);
						const bool p1 = c1 >= G::OFF && c1 < G::OFF + ncols && (unsafe || srow[c1] > t_min);
						lds_append(p0, (uint16_t)((row << G::CBITS) | c0), sh.alist, &sh.na);
						lds_append(p1, (uint16_t)((row << G::CBITS) | c1), sh.alist, &sh.na);
					} else {
						unsigned long long m = unsafe ? cols : __ballot(srow[lane] > t_min) & cols;
						lds_append(((m >> lane) & 1) != 0, (uint16_t)((row << G::CBITS) | lane), sh.alist,
							&sh.na);
					}
				}
			}
			/* EXCL at the tile's masked pixels (the flags plane starts zeroed;
			 * holes are flagged above). A screened pixel's flags are written
			 * once, with SCR, by the medians below. */
			if (fr.mask) {
				/* every mask byte of this thread is loaded at once; the data
				 * only at masked pixels */
				constexpr int PER = (G::TW * TILE_H + BLOCK - 1) / BLOCK;
				uint8_t mv[PER];
#pragma unroll
				for (int j = 0; j < PER; ++j) {
					const int k = tid + j * BLOCK, row = k / G::TW, col = k % G::TW;
					const bool in = k < G::TW * TILE_H && row < nrows && col < ncols;
					mv[j] = in ? fr.mask[(ty + row) * w + X0 + col] : 0;
				}
#pragma unroll
				for (int j = 0; j < PER; ++j) {
					if (!mv[j])
						continue;
					const int k = tid + j * BLOCK, row = k / G::TW, col = k % G::TW;
					const int i = (ty + row) * w + X0 + col;
					const double d = fr.data[i];
					if (!isfinite(d))
						continue;
					const bool screened = !skip && (unsafe ||
						sh.s[(r_lo + HS + row) % RING][col + G::OFF] > t_min);
					if (!screened)
						p.flags[fr.off + i] = uint8_t(EXCL);
					++acc.masked;
					if (BG) {
						bool hh = false;
						const double c = input_c<BG>(d, fr.background ? fr.background[i] : p.bg,
							has_bg, hh);
						acc.masked_nonfinite += !isfinite(c);
					}
				}
			}
			lds_barrier();

			/* 4. S' at the screened pixels, which get SCR (no hole lies in the
			 * tile); those above t are buffered as seeds (a batch adds at most
			 * BLOCK, and the buffer keeps room for that many). Excluded pixels
			 * are dropped by the consumers, which read the flags anyway. */
			const int na = sh.na;
			for (int base = 0; base < na; base += BLOCK) {
				int idx = base + tid;
				bool seed = false;
				int i = 0;
				if (idx < na) {
					int pos = sh.alist[idx], row = pos >> G::CBITS, c = pos & ((1 << G::CBITS) - 1);
					i = (ty + row) * w + X0 - G::OFF + c;
					int rc = r_lo + HS + row;
					double v[25];
#pragma unroll
					for (int dy = 0; dy < 5; ++dy) {
						int sl = (rc - 2 + dy) % RING;
#pragma unroll
						for (int dx = 0; dx < 5; ++dx)
							v[dy * 5 + dx] = sh.s[sl][c - 2 + dx];
					}
					double sp = v[12] - s_median25(v);
					p.sprime[fr.off + i] = sp;
					p.flags[fr.off + i] = uint8_t(SCR | (fr.mask && fr.mask[i] ? EXCL : 0u));
					seed = sp > t_cr;
				}
				lds_append(seed, list_entry(f, i), sh.seeds, &sh.nseed);
				lds_barrier();
				if (sh.nseed + BLOCK > SEED_BUF)
					flush_seeds(sh, p.list_b, seeds);
			}
		}
		/* the run's counts: one LDS atomic per wave, one global atomic per
		 * count and run */
		{
			/* the run's own pixels, counted once, less this lane's holes and
			 * masked pixels */
			const int own_total = !threadIdx.x ? ncols * (own1 - Y0) : 0;
			const int allowed = own_total - acc.holes - acc.masked;
			int c3[3] = {acc.holes, allowed, allowed - (BG ? acc.nonfinite - acc.masked_nonfinite : 0)};
#pragma unroll
			for (int k = 0; k < 3; ++k) {
				for (int o = 32; o; o >>= 1)
					c3[k] += __shfl_xor(c3[k], o, 64);
			}
			const bool inf = BG && __ballot(acc.infinite) != 0;
			if (!(threadIdx.x & 63)) {
#pragma unroll
				for (int k = 0; k < 3; ++k)
					if (c3[k])
						atomicAdd(&sh.own[k], c3[k]);
				if (inf)
					atomicOr(&p.status[1], 1);
			}
		}
		lds_barrier();
		if (!threadIdx.x) {
			int *shard = p.shard + ((size_t)f * SHARDS + blockIdx.x % SHARDS) * SHARD_FIELDS;
			if (sh.own[0])
				atomicAdd(shard + SHARD_HOLES, sh.own[0]);
			if (sh.own[1])
				atomicAdd(shard + SHARD_ALLOWED, sh.own[1]);
			if (sh.own[2])
				atomicAdd(shard + SHARD_DONORS, sh.own[2]);
			sh.own[0] = sh.own[1] = sh.own[2] = 0;
		}
	}
	flush_seeds(sh, p.list_b, seeds);
	/* one atomic per workgroup: a counter that every tile hit would
	 * serialise the whole pass in L2 */
	if (!threadIdx.x && work)
		atomicAdd(p.work, work);
}

/* S at an in-frame pixel (x, y); bad = its noise value is invalid. */
template <bool MODEL, bool EXACT, bool TAME>
__device__ __forceinline__ double
significance_at(const Params &p, const Frame &fr, int x, int y, bool &bad)
{
	const int mode = p.mode, w = fr.w, h = fr.h;
	const double *clean = p.clean + fr.off;
	double qc = clean[y * w + x] * 0.25;
	double qu = lap_sample(clean, x, y - 1, w, h, mode) * 0.25;
	double qd = lap_sample(clean, x, y + 1, w, h, mode) * 0.25;
	double ql = lap_sample(clean, x - 1, y, w, h, mode) * 0.25;
	double qr = lap_sample(clean, x + 1, y, w, h, mode) * 0.25;
	double nz;
	if (!MODEL || fr.error) {
		nz = fr.error[y * w + x];
	} else {
		double v[25], m;
#pragma unroll
		for (int k = 0; k < 25; ++k)
			v[k] = med_sample(clean, x - 2 + k % 5, y - 2 + k / 5, w, h, mode);
		median25(v, &m);
		nz = noise_model(m, fr.gain ? fr.gain[y * w + x] : fr.g,
			fr.readnoise ? fr.readnoise[y * w + x] : fr.rn);
	}
	bad = invalid_noise(nz);
	return laplacian_q<EXACT, TAME>(qc, qu, qd, ql, qr, w == 1) / (2 * nz);
}

/* S' at pixel (x, y) of a frame if the last pass that processed the pixel
 * computed it there (0 at holes, src/filters.cuh:137-146), else -inf: neither
 * can pass a threshold t >= 0. A block that no block pass has processed is the
 * first pass's (SCR); otherwise the pixel's stamp must be the block's. */
__device__ __forceinline__ double
sprime_at(const Params &p, long long off, int block0, int blocks_x, int x, int y, int w)
{
	long long g = off + (long long)y * w + x;
	uint32_t fw = p.flags[g];
	unsigned last = p.block_pass[block0 + (y / BLK) * blocks_x + x / BLK];
	bool current = last ? p.stamp[g] == last : (fw & SCR) != 0;
	return current && !(fw & HOLE) ? p.sprime[g] : -INFINITY;
}

/* S' (as sprime_at) at the 3 x 3 neighbours of (x, y) in row-major order,
 * -inf outside the frame or for an invalid lane: every neighbour's flags,
 * block stamp and pixel stamp are loaded before any is used, then S' where it
 * is current, so a lane waits for two round trips instead of nine. */
__device__ __forceinline__ void
neighbour_sprime(const Params &p, long long off, int block0, int blocks_x, int x, int y, int w, int h,
	bool valid, double sp[9])
{
	uint32_t fw[9];
	unsigned last[9], st[9];
	bool in[9];
#pragma unroll
	for (int k = 0; k < 9; ++k) {
		const int qx = x + k % 3 - 1, qy = y + k / 3 - 1;
		in[k] = valid && qx >= 0 && qx < w && qy >= 0 && qy < h;
		fw[k] = 0;
		last[k] = 0;
		st[k] = 0;
		if (in[k]) {
			const long long g = off + (long long)qy * w + qx;
			fw[k] = p.flags[g];
			last[k] = p.block_pass[block0 + (qy / BLK) * blocks_x + qx / BLK];
			st[k] = p.stamp[g];
		}
	}
#pragma unroll
	for (int k = 0; k < 9; ++k) {
		const bool current = last[k] ? st[k] == last[k] : (fw[k] & SCR) != 0;
		sp[k] = -INFINITY;
		if (in[k] && current && !(fw[k] & HOLE))
			sp[k] = p.sprime[off + (long long)(y + k / 3 - 1) * w + x + k % 3 - 1];
	}
}

constexpr int BR = BLK + 2 * HS; /* edge of the S region of a block */
constexpr int WAVE_SEEDS = 256;  /* seeds buffered per wave between flushes */
static_assert(BLK * BLK == 64, "one lane per pixel of a block");

constexpr int CR = BR + 2; /* edge of the C region of a block: S and its ring */

struct BlockShared {
	double s[BLOCK / 64][BR * BR];
	double c[BLOCK / 64][CR * CR];
	unsigned long long seeds[BLOCK / 64][WAVE_SEEDS];
	int nseed[BLOCK / 64];
	unsigned base;
	unsigned long long work;
};


/* Move this wave's buffered seeds to the global list (whole wave). */
__device__ __forceinline__ void
flush_wave_seeds(unsigned long long *buf, int *n, unsigned long long *list, unsigned int *count)
{
	const int lane = threadIdx.x & 63, k = *n;
	unsigned base = 0;
	if (!lane && k)
		base = atomicAdd(count, (unsigned)k);
	base = __shfl(base, 0, 64);
	for (int j = lane; j < k; j += 64)
		list[base + j] = buf[j];
	__builtin_amdgcn_wave_barrier();
	if (!lane)
		*n = 0;
	__builtin_amdgcn_wave_barrier();
}

/* Block pass: one wave per dirty block, independent of the other waves. S
 * over the block and a 2-pixel halo goes to the wave's LDS region; a slot
 * outside the frame holds S at its mapped in-frame pixel (median border mode;
 * 0 for CONSTANT), which WRAP's periodic stencils also give. Screened pixels
 * of the block get S' as in the first pass, and every pixel of the block the
 * pass's stamp (screened) or 0. In pass 0 the blocks are those of the tiles
 * that the first pass left to this pass after the hole fill: only pixels of
 * those tiles become seeds (the first pass has the others). */
template <bool MODEL, bool EXACT, bool TAME>
__device__ __forceinline__ void
significance_blocks_pass(BlockShared &sh, const Params &p, int iter)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int mode = p.mode;
	const double t_cr = p.threshold, t_min = fmin(p.threshold, p.neighbor);
	const unsigned stamp = pass_stamp(iter), tag = tag_stamp(iter);
	const unsigned n = p.nlist[nl_dirty(p, iter)];
	unsigned int *seeds = &p.nlist[nl_seed(p, iter)];
	double *S = sh.s[wave];
	unsigned long long *buf = sh.seeds[wave];
	int *nseed = &sh.nseed[wave];
	unsigned long long work = 0;
	if (!lane)
		*nseed = 0;
	if (!threadIdx.x)
		sh.work = 0;
	__syncthreads();
	/* Each XCD takes a contiguous part of the list (workgroups go to the
	 * XCDs in turn; the grid is a multiple of XCDS), so that neighbouring
	 * blocks, which sit together in the list, share one L2. */
	const unsigned xcd = blockIdx.x % XCDS, waves = gridDim.x / XCDS * (BLOCK / 64);
	const unsigned hi = unsigned((unsigned long long)n * (xcd + 1) / XCDS);
	unsigned e = unsigned((unsigned long long)n * xcd / XCDS) + blockIdx.x / XCDS * (BLOCK / 64) + wave;
	/* the next block's list entry is loaded while this one is processed */
	unsigned long long next = e < hi ? p.dirty_list[e] : 0;
	for (; e < hi; e += waves) {
		const unsigned long long entry = uniform_u64(next);
		if (e + waves < hi)
			next = p.dirty_list[e + waves];
		const int f = (int)(entry >> 32), b = (int)((unsigned)entry & ~RESEED);
		const Frame fr = p.frames[f];
		const int lb = b - fr.block0, w = fr.w, h = fr.h;
		const int X0 = lb % fr.blocks_x * BLK, Y0 = lb / fr.blocks_x * BLK;
		if ((unsigned)entry & RESEED) {
			/* S' is still current here: re-emit the seeds, unless the block
			 * also recomputes in this pass */
			if (__atomic_load_n(&p.dirty_tag[b], __ATOMIC_RELAXED) == tag)
				continue;
			const int x = X0 + lane % BLK, y = Y0 + lane / BLK;
			const bool seed = x < w && y < h &&
				sprime_at(p, fr.off, fr.block0, fr.blocks_x, x, y, w) > t_cr;
			lds_append(seed, list_entry(f, y * w + x), buf, nseed);
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			if (*nseed > WAVE_SEEDS - 64)
				flush_wave_seeds(buf, nseed, p.list_b, seeds);
			continue;
		}
		if (!lane) {
			p.block_pass[b] = stamp;
			if (iter)
				work += (unsigned long long)(min(BLK, w - X0) * min(BLK, h - Y0));
		}
		bool bad = false;
		/* the region and its one-pixel ring lie inside the frame */
		const bool inner = X0 - HS - 1 >= 0 && X0 + BLK + HS < w && Y0 - HS - 1 >= 0 && Y0 + BLK + HS < h;
		if (MODEL && fr.error && inner) {
			/* (the model kernel keeps its registers: five loads per S) */
			const double *cb = p.clean + fr.off + (long long)(Y0 - HS) * w + (X0 - HS);
			const double *eb = fr.error + (long long)(Y0 - HS) * w + (X0 - HS);
#pragma unroll
			for (int j = 0; j < 3; ++j) {
				const int k = lane + 64 * j;
				if (j < 2 || k < BR * BR) {
					const int o = (k / BR) * w + k % BR;
					double nz = eb[o];
					double q = laplacian_q<EXACT, TAME>(cb[o] * 0.25, cb[o - w] * 0.25,
						cb[o + w] * 0.25, cb[o - 1] * 0.25, cb[o + 1] * 0.25, false);
					bad |= invalid_noise(nz);
					S[k] = q / (2 * nz);
				}
			}
		} else if (!MODEL && inner) {
			/* the C region (S and its ring) once into the wave's LDS, then
			 * the stencil from there */
			const double *cb = p.clean + fr.off + (long long)(Y0 - HS - 1) * w + (X0 - HS - 1);
			const double *eb = fr.error + (long long)(Y0 - HS) * w + (X0 - HS);
			double *Ct = sh.c[wave];
			double nz[3];
#pragma unroll
			for (int j = 0; j < 3; ++j) {
				const int k = lane + 64 * j;
				nz[j] = j < 2 || k < BR * BR ? eb[(k / BR) * w + k % BR] : 1.0;
			}
#pragma unroll
			for (int j = 0; j < 4; ++j) {
				const int k = lane + 64 * j;
				if (j < 3 || k < CR * CR)
					Ct[k] = cb[(k / CR) * w + k % CR];
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
			for (int j = 0; j < 3; ++j) {
				const int k = lane + 64 * j;
				if (j < 2 || k < BR * BR) {
					const int c = (k / BR + 1) * CR + k % BR + 1;
					double q = laplacian_q<EXACT, TAME>(Ct[c] * 0.25, Ct[c - CR] * 0.25,
						Ct[c + CR] * 0.25, Ct[c - 1] * 0.25, Ct[c + 1] * 0.25, false);
					bad |= invalid_noise(nz[j]);
					S[k] = q / (2 * nz[j]);
				}
			}
		} else {
			for (int k = lane; k < BR * BR; k += 64) {
				int mx = border_index(X0 - HS + k % BR, w, mode);
				int my = border_index(Y0 - HS + k / BR, h, mode);
				double v = 0.0;
				if (mx >= 0 && my >= 0) {
					bool bk;
					v = significance_at<MODEL, EXACT, TAME>(p, fr, mx, my, bk);
					bad |= bk;
				}
				S[k] = v;
			}
		}
		const bool unsafe = __ballot(bad) != 0;
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

		const int ox = lane % BLK, oy = lane / BLK, x = X0 + ox, y = Y0 + oy;
		const int c = (oy + HS) * BR + ox + HS;
		bool seed = false;
		if (x < w && y < h) {
			const long long g = fr.off + (long long)y * w + x;
			const bool screened = unsafe || S[c] > t_min;
			if (screened) {
				double v[25];
#pragma unroll
				for (int k = 0; k < 25; ++k)
					v[k] = S[c + (k / 5 - 2) * BR + k % 5 - 2];
				double sp = v[12] - s_median25(v);
				p.sprime[g] = sp;
				seed = sp > t_cr;
			}
			p.stamp[g] = uint16_t(screened ? stamp : 0u);
			if (!iter && seed)
				seed = p.tile_skip[fr.tile0 + (y / TILE_H) * fr.tiles_x + x / p.tile_w] != 0;
		}
		lds_append(seed, list_entry(f, y * w + x), buf, nseed);
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		if (*nseed > WAVE_SEEDS - 64)
			flush_wave_seeds(buf, nseed, p.list_b, seeds);
	}
	/* the workgroup's remaining seeds and its pixel count: one global atomic
	 * each for all four waves */
	if (!lane && work)
		atomicAdd(&sh.work, work);
	__syncthreads();
	if (!threadIdx.x) {
		unsigned total = 0;
		for (int k = 0; k < BLOCK / 64; ++k)
			total += sh.nseed[k];
		sh.base = total ? atomicAdd(seeds, total) : 0;
		if (sh.work)
			atomicAdd(p.work, sh.work);
	}
	__syncthreads();
	unsigned at = sh.base;
	for (int k = 0; k < wave; ++k)
		at += sh.nseed[k];
	for (int j = lane; j < *nseed; j += 64)
		p.list_b[at + j] = buf[j];
}

template <bool MODEL, bool EXACT>
__global__ __launch_bounds__(BLOCK) void
significance_blocks(Params p, int iter)
{
	__shared__ BlockShared sh;
	if (p.status[1])
		significance_blocks_pass<MODEL, EXACT, false>(sh, p, iter);
	else
		significance_blocks_pass<MODEL, EXACT, true>(sh, p, iter);
}

/* The first pass when every frame has an error map, on wide tiles. Its LDS
 * (the 128-column ring) allows three workgroups per CU, and the register
 * budget is set to match. BG: some frame has a background; BGMAP: some frame
 * has a background map. */
template <bool EXACT, bool BG, bool BGMAP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void
significance_maps(Params p, int iter)
{
	static_assert(BG || !BGMAP, "a background map is a background");
	__shared__ SigShared<true> sh;
	significance_pass<false, EXACT, BG, BGMAP>(sh, nullptr, p, iter);
}

/* Some frame uses the noise model (5x5 medians of C in the dense part):
 * two waves per SIMD. */
template <bool EXACT, bool BG>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2))) void
significance_model(Params p, int iter)
{
	__shared__ SigShared<false> sh;
	__shared__ ModelRegion region;
	significance_pass<true, EXACT, BG, BG>(sh, &region, p, iter);
}

__device__ __forceinline__ double
med3(double a, double b, double c)
{
	return fmax(fmin(a, b), fmin(fmax(a, b), c));
}

/* Feed row j of the 7 x 7 med3 block to the incremental median of 49. */
__device__ __forceinline__ void
median49_feed(int j, const double *acc, const double *row, double *out)
{
	switch (j) {
	case 0: median49_step0(acc, row, out); break;
	case 1: median49_step1(acc, row, out); break;
	case 2: median49_step2(acc, row, out); break;
	case 3: median49_step3(acc, row, out); break;
	case 4: median49_step4(acc, row, out); break;
	case 5: median49_step5(acc, row, out); break;
	default: median49_step6(acc, row, out); break;
	}
}

/* The 9 x 9 block of C samples around (x, y) lies inside the frame. */
__device__ __forceinline__ bool
fine_inner(int x, int y, int w, int h)
{
	return x >= 4 && x + 4 < w && y >= 4 && y + 4 < h;
}

/* F at a seed whose 9 x 9 samples lie inside the frame (fine_inner;
 * src/filters.cuh:148-162): m3 = med3 of the clean image over the 7 x 7
 * positions around it, med7 = their median, fed row by row so that at most 22
 * ranks stay live, and N at the seed. Rows are loaded one at a time through
 * one row pointer (about 220 VGPRs, two waves per SIMD). fine_edge takes the
 * seeds near an edge. */
__device__ double
fine_structure(const Params &p, const Frame &fr, int x, int y)
{
	const double *clean = p.clean + fr.off;
	const int w = fr.w;
	auto load_row = [&](int yy, double *dst) {
		const double *r = clean + (long long)yy * w + x - 4;
#pragma unroll
		for (int i = 0; i < 9; ++i)
			dst[i] = r[i];
	};
	double acc[2][22], a[9], b[9], c[9], centre = 0.0;
	load_row(y - 4, a);
	load_row(y - 3, b);
#pragma unroll
	for (int j = 0; j < 7; ++j) {
		double row[7];
		load_row(y - 2 + j, c);
		/* 3x3 medians from sorted columns: med9 = med3(max of minima, med3
		 * of middles, min of maxima) */
		double lo[9], md[9], hi[9];
#pragma unroll
		for (int i = 0; i < 9; ++i) {
			double l1 = fmin(a[i], b[i]), h1 = fmax(a[i], b[i]);
			double t = fmax(l1, c[i]);
			lo[i] = fmin(l1, c[i]);
			md[i] = fmin(h1, t);
			hi[i] = fmax(h1, t);
		}
#pragma unroll
		for (int i = 0; i < 7; ++i) {
			double L = fmax(fmax(lo[i], lo[i + 1]), lo[i + 2]);
			double H = fmin(fmin(hi[i], hi[i + 1]), hi[i + 2]);
			double M = med3(md[i], md[i + 1], md[i + 2]);
			row[i] = med3(L, M, H);
		}
#pragma unroll
		for (int i = 0; i < 9; ++i) {
			a[i] = b[i];
			b[i] = c[i];
		}
		if (j == 3)
			centre = row[3];
		median49_feed(j, acc[(j + 1) & 1], row, acc[j & 1]);
		/* keep the loads of later rows from being hoisted here */
		__builtin_amdgcn_sched_barrier(0);
	}
	const double med7 = acc[0][0];
	double nz;
	if (fr.error) {
		nz = fr.error[y * w + x];
	} else {
		double v25[25], m5;
		for (int j = 0; j < 5; ++j) {
			double r9[9];
			load_row(y - 2 + j, r9);
			for (int i = 0; i < 5; ++i)
				v25[j * 5 + i] = r9[i + 2];
		}
		median25(v25, &m5);
		double g = fr.gain ? fr.gain[y * w + x] : fr.g;
		double rn = fr.readnoise ? fr.readnoise[y * w + x] : fr.rn;
		nz = noise_model(m5, g, rn);
	}
	double fine = (centre - med7) / nz;
	return fine < FINE_FLOOR ? FINE_FLOOR : fine;
}

/* Candidates among the seeds of pass iter: S' / F > contrast. This launch
 * takes the fine_inner seeds and lists the others not excluded (edge_list),
 * which fine_edge takes. */
__global__ __launch_bounds__(BLOCK) void
fine(Params p, int iter)
{
	__shared__ BlockList<2 * BLOCK> out; /* flushed when the next round might not fit */
	const unsigned n = p.nlist[nl_seed(p, iter)];
	const unsigned long long *seeds = p.list_b;
	unsigned int *count = &p.nlist[nl_cand(p, iter)];
	if (!threadIdx.x)
		out.n = 0;
	for (unsigned base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
		unsigned e = base + threadIdx.x;
		bool cand = false, edge = false;
		unsigned long long entry = 0;
		if (e < n) {
			entry = seeds[e];
			int f = (int)(entry >> 32), i = (int)(entry & 0xffffffffu);
			const Frame fr = p.frames[f];
			const int x = i % fr.w, y = i / fr.w;
			if (!(p.flags[fr.off + i] & EXCL)) {
				if (!fine_inner(x, y, fr.w, fr.h)) {
					edge = true;
				} else {
					double sp = p.sprime[fr.off + i];
					cand = sp / fine_structure(p, fr, x, y) > p.contrast;
				}
			}
		}
		list_push(edge, entry, p.edge_list, &p.nlist[nl_edge(p, iter)]);
		__syncthreads();
		lds_append(cand, entry, out.item, &out.n);
		__syncthreads();
		if (out.n > BLOCK)
			block_list_flush(out, p.list_c, count);
	}
	block_list_flush(out, p.list_c, count);
}

/* Rank of lane k's value among the first n of vals (ties by lane): the
 * values' ranks are 0 .. n - 1, each once, so the lane of rank n / 2 holds
 * the value that a median network over them returns. */
__device__ __forceinline__ int
lane_rank(const double *vals, int n, int k)
{
	const double v = vals[k];
	int r = 0;
	for (int j = 0; j < n; ++j)
		r += vals[j] < v || (vals[j] == v && j < k);
	return r;
}

/* The seeds that fine left on the edge list (their 9 x 9 samples cross
 * the frame edge): F as in fine_structure, one wave per seed. Lane k < 49
 * takes m3 at position k of the 7 x 7 window under the same border rules
 * (med3 of mapped samples; CONSTANT m3 outside the frame 0; NEAREST off the
 * inner ring at the clamped position), and the wave selects med7 by rank,
 * as does the noise model's 5 x 5 median. C holds no nonfinite value by now,
 * and a signed zero among the medians does not change F (centre - med7 and
 * the noise floor), so F equals fine_structure's. */
__global__ __launch_bounds__(BLOCK) void
fine_edge(Params p, int iter)
{
	__shared__ BlockList<2 * BLOCK> out; /* flushed when the next round might not fit */
	__shared__ double vals[BLOCK / 64][64];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const unsigned n = p.nlist[nl_edge(p, iter)];
	unsigned int *count = &p.nlist[nl_cand(p, iter)];
	const int mode = p.mode;
	if (!threadIdx.x)
		out.n = 0;
	double *v = vals[wave];
	for (unsigned base = blockIdx.x * (BLOCK / 64); base < n; base += gridDim.x * (BLOCK / 64)) {
		const unsigned e = base + wave;
		bool cand = false;
		unsigned long long entry = 0;
		if (e < n) {
			entry = uniform_u64(p.edge_list[e]);
			const int f = (int)(entry >> 32), i = (int)(entry & 0xffffffffu);
			const Frame fr = p.frames[f];
			const int w = fr.w, h = fr.h, x = i % w, y = i / w;
			const double *clean = p.clean + fr.off;
			const bool inside = x >= 3 && x + 3 < w && y >= 3 && y + 3 < h;
			double m = 0.0;
			if (lane < 49) {
				const int qx = x - 3 + lane % 7, qy = y - 3 + lane / 7;
				double v9[9];
				if (inside || mode != CUSMIC_NEAREST) {
#pragma unroll
					for (int k = 0; k < 9; ++k)
						v9[k] = med_sample(clean, qx - 1 + k % 3, qy - 1 + k / 3, w, h, mode);
					median9(v9, &m);
					const bool outside = qx < 0 || qx >= w || qy < 0 || qy >= h;
					if (mode == CUSMIC_CONSTANT && outside)
						m = 0.0;
				} else {
					const int cx = min(max(qx, 0), w - 1), cy = min(max(qy, 0), h - 1);
#pragma unroll
					for (int k = 0; k < 9; ++k)
						v9[k] = med_sample(clean, cx - 1 + k % 3, cy - 1 + k / 3, w, h, mode);
					median9(v9, &m);
				}
				v[lane] = m;
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			const double centre = v[24];
			const unsigned long long mid = __ballot(lane < 49 && lane_rank(v, 49, lane) == 24);
			const double med7 = __shfl(m, __ffsll((long long)mid) - 1, 64);
			double nz;
			if (fr.error) {
				nz = fr.error[i];
			} else {
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
				double c = 0.0;
				if (lane < 25) {
					c = med_sample(clean, x - 2 + lane % 5, y - 2 + lane / 5, w, h, mode);
					v[lane] = c;
				}
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
				const unsigned long long m5l = __ballot(lane < 25 && lane_rank(v, 25, lane) == 12);
				const double m5 = __shfl(c, __ffsll((long long)m5l) - 1, 64);
				const double g = fr.gain ? fr.gain[i] : fr.g;
				const double rn = fr.readnoise ? fr.readnoise[i] : fr.rn;
				nz = noise_model(m5, g, rn);
			}
			double fine = (centre - med7) / nz;
			fine = fine < FINE_FLOOR ? FINE_FLOOR : fine;
			cand = !(p.flags[fr.off + i] & EXCL) && p.sprime[fr.off + i] / fine > p.contrast;
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		}
		__syncthreads();
		lds_append(cand && !lane, entry, out.item, &out.n);
		__syncthreads();
		if (out.n > BLOCK)
			block_list_flush(out, p.list_c, count);
	}
	block_list_flush(out, p.list_c, count);
}

/* First growth (src/filters.cuh:177-193 at t): every pixel next to a
 * candidate with S' > t joins the growth list once. */
__global__ __launch_bounds__(BLOCK) void
grow_first(Params p, int iter)
{
	__shared__ BlockList<BLOCK * 9> out;
	const unsigned n = p.nlist[nl_cand(p, iter)];
	unsigned int *count = &p.nlist[nl_grow(p, iter)];
	if (!threadIdx.x)
		out.n = 0;
	/* the flagged list before this pass: its tail will be the new pixels */
	if (!blockIdx.x && !threadIdx.x)
		p.nlist[nl_start(p, iter)] = p.nlist[0];
	for (unsigned base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
		unsigned e = base + threadIdx.x;
		int f = 0, x = 0, y = 0, w = 1, h = 1, block0 = 0, blocks_x = 1;
		long long off = 0;
		if (e < n) {
			unsigned long long entry = p.list_c[e];
			f = (int)(entry >> 32);
			int i = (int)(entry & 0xffffffffu);
			const Frame fr = p.frames[f];
			w = fr.w;
			h = fr.h;
			off = fr.off;
			block0 = fr.block0;
			blocks_x = fr.blocks_x;
			x = i % w;
			y = i / w;
		}
		__syncthreads();
		/* all neighbours' S' in one round trip, then their atomics */
		double sp[9];
		neighbour_sprime(p, off, block0, blocks_x, x, y, w, h, e < n, sp);
		uint32_t old[9];
#pragma unroll
		for (int k = 0; k < 9; ++k) {
			old[k] = GROW;
			if (sp[k] > p.threshold)
				old[k] = flags_or(&p.flags[off + (y + k / 3 - 1) * w + x + k % 3 - 1], GROW);
		}
#pragma unroll
		for (int k = 0; k < 9; ++k)
			lds_append(!(old[k] & GROW), list_entry(f, (y + k / 3 - 1) * w + x + k % 3 - 1), out.item,
				&out.n);
		block_list_flush(out, p.list_b, count);
	}
}

/* Second growth at t_n and accumulation (src/cleaner.cuh:22-37): a pixel next
 * to a grown pixel with S' > t_n is flagged; if it was not flagged before it
 * is new, joins the flagged list and leaves the donor count. */
__global__ __launch_bounds__(BLOCK) void
grow_second(Params p, int iter)
{
	__shared__ BlockList<BLOCK * 9> out;
	const unsigned n = p.nlist[nl_grow(p, iter)];
	if (!threadIdx.x)
		out.n = 0;
	for (unsigned base = blockIdx.x * BLOCK; base < n; base += gridDim.x * BLOCK) {
		unsigned e = base + threadIdx.x;
		int f = 0, x = 0, y = 0, w = 1, h = 1, block0 = 0, blocks_x = 1, fresh_n = 0, lost_n = 0;
		long long off = 0;
		if (e < n) {
			unsigned long long entry = p.list_b[e];
			f = (int)(entry >> 32);
			int i = (int)(entry & 0xffffffffu);
			const Frame fr = p.frames[f];
			w = fr.w;
			h = fr.h;
			off = fr.off;
			block0 = fr.block0;
			blocks_x = fr.blocks_x;
			x = i % w;
			y = i / w;
			flags_clear(&p.flags[off + i], GROW);
		}
		__syncthreads();
		/* all neighbours' S' in one round trip, then their atomics */
		double sp[9];
		neighbour_sprime(p, off, block0, blocks_x, x, y, w, h, e < n, sp);
		uint32_t old[9];
#pragma unroll
		for (int k = 0; k < 9; ++k) {
			old[k] = CRAY;
			if (sp[k] > p.neighbor)
				old[k] = flags_or(&p.flags[off + (y + k / 3 - 1) * w + x + k % 3 - 1], CRAY);
		}
		bool fresh[9];
#pragma unroll
		for (int k = 0; k < 9; ++k) {
			const long long q = off + (y + k / 3 - 1) * w + x + k % 3 - 1;
			fresh[k] = !(old[k] & CRAY);
			if (fresh[k]) {
				flags_or(&p.flags[q], NEWP);
				fresh_n += 1;
				lost_n += !(old[k] & EXCL) && isfinite(p.clean[q]);
			}
		}
#pragma unroll
		for (int k = 0; k < 9; ++k)
			lds_append(fresh[k], list_entry(f, (y + k / 3 - 1) * w + x + k % 3 - 1), out.item, &out.n);
		int *shard = p.shard + (blockIdx.x % SHARDS) * SHARD_FIELDS;
		wave_frame_sum(fresh_n, f, shard + SHARD_FRESH, SHARDS * SHARD_FIELDS);
		wave_frame_sum(-lost_n, f, shard + SHARD_DONORS, SHARDS * SHARD_FIELDS);
		block_list_flush(out, p.list, &p.nlist[0]);
	}
}

/* ------------------------------------------------------------------------ */
/* Adjoint with the mask held fixed                                          */

/* The gradient of a donor q is the identity (unflagged q) plus, for every
 * flagged pixel p whose recorded median donors include q, p's upstream
 * gradient times its weight (1 for one middle donor, 1/2 for two), added in
 * raster order of p. The donors of a 5 x 5 window lie within 2 of p; those of
 * a widened window of radius r all lie at distance exactly r > 2 (the window
 * widened because no donor was nearer). So adjoint_gather sums over the 5 x 5
 * neighbourhood of every pixel, and the few targets of widened windows are
 * listed and recomputed by adjoint_wide over the frame's widest radius. */

constexpr int ADJ_NEAR = 2; /* donor distance of a 5 x 5 window */

/* g[p] times p's weight on donor q (0 if q is not one of p's donors). */
__device__ __forceinline__ double
donor_weight_term(const int32_t *donors, const double *grad_clean, long long pg, int q, bool &hit)
{
	const int lo = donors[2 * pg], hi = donors[2 * pg + 1];
	hit = lo == q || hi == q;
	return lo == hi ? grad_clean[pg] : 0.5 * grad_clean[pg];
}

/* One workgroup per tile. GATHER: grad_data over the 5 x 5 neighbourhoods.
 * Both forms append the donor targets of the tile's widened windows
 * (frame << 32 | q, possibly repeated) to list while fewer than cap are
 * listed, count them all in *count, and raise reach[f] to the widest
 * window's radius.
 *
 * Lane l is column x0 - 2 + l, so a wave's ballot over a row is the row's
 * flagged pixels around the tile's TILE_W columns (lanes 2..TILE_W + 1). Each wave walks
 * its 8 rows with the masks of the five rows around the current one: a pixel
 * whose 5 x 5 window holds no flagged pixel (nearly all) gets the identity,
 * the others add the terms of the flagged pixels of the window, whose bits
 * are taken in raster order. */
template <bool GATHER>
__global__ __launch_bounds__(BLOCK) void
adjoint_gather(const Frame *frames, int nf, const uint8_t *mask, const int32_t *donors,
	const double *grad_clean, double *grad_data, int *reach, unsigned int *count,
	unsigned long long *list, unsigned cap)
{
	static_assert(TILE_W + 2 * ADJ_NEAR <= 64, "a tile row and its window fit one wave");
	const int tile = blockIdx.x;
	const int f = find_frame(frames, nf, tile);
	const Frame fr = frames[f];
	const TilePixels t = tile_of(fr, tile);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int x = t.x0 - ADJ_NEAR + lane;
	const bool col_in = x >= 0 && x < t.w;
	const bool out_col = lane >= ADJ_NEAR && lane < ADJ_NEAR + TILE_W && x < t.w;
	const long long base = fr.off;
	const int ya = t.y0 + wave * (TILE_H / 4);
	auto row_mask = [&](int yy) {
		return __ballot(col_in && yy >= 0 && yy < t.h && mask[base + (long long)yy * t.w + x] != 0);
	};
	unsigned long long m[5];
#pragma unroll
	for (int k = 0; k < 4; ++k)
		m[k] = row_mask(ya - ADJ_NEAR + k);
	int widest = 0;
	for (int r = 0; r < TILE_H / 4; ++r) {
		const int y = ya + r;
		m[4] = row_mask(y + ADJ_NEAR);
		const bool out = out_col && y < t.h;
		const int q = y * t.w + x;
		const bool flagged = out && ((m[2] >> lane) & 1);
		int targets[2] = {-1, -1};
		if (flagged) {
			const int lo = donors[2 * (base + q)], hi = donors[2 * (base + q) + 1];
			const int d = lo < 0 ? 0 : max(abs(lo % t.w - x), abs(lo / t.w - y));
			if (d > ADJ_NEAR) {
				widest = max(widest, d);
				targets[0] = lo;
				targets[1] = hi != lo ? hi : -1;
			}
		}
		for (int k = 0; k < 2; ++k) {
			const bool emit = targets[k] >= 0;
			const unsigned long long b = __ballot(emit);
			if (!b)
				continue;
			const int leader = __ffsll((long long)b) - 1;
			unsigned at = 0;
			if (lane == leader)
				at = atomicAdd(count, (unsigned)__popcll(b));
			at = __shfl(at, leader, 64) + __popcll(b & ((1ull << lane) - 1));
			if (emit && at < cap)
				list[at] = list_entry(f, targets[k]);
		}
		if (GATHER && out) {
			double s = flagged ? 0.0 : grad_clean[base + q];
#pragma unroll
			for (int k = 0; k < 5; ++k) {
				unsigned bits = unsigned(m[k] >> (lane - ADJ_NEAR)) & 0x1fu;
				while (bits) {
					const int c = __builtin_ctz(bits);
					bits &= bits - 1;
					const long long pg = base + (long long)(y - ADJ_NEAR + k) * t.w + x - ADJ_NEAR + c;
					bool hit;
					const double term = donor_weight_term(donors, grad_clean, pg, q, hit);
					if (hit)
						s += term;
				}
			}
			grad_data[base + q] = s;
		}
#pragma unroll
		for (int k = 0; k < 4; ++k)
			m[k] = m[k + 1];
	}
	for (int o = 32; o; o >>= 1)
		widest = max(widest, __shfl_xor(widest, o, 64));
	if (!lane && widest)
		atomicMax(&reach[f], widest);
}

/* One wave per listed target q (entries [0, min(*count, limit))): grad_data[q]
 * from all flagged pixels within the frame's widest radius, in raster order.
 * The window is scanned in chunks of 64 positions (lane order is raster
 * order); the terms found in a chunk are added one after another, so the sum
 * is the same as a sequential scan. A target listed twice gets the same
 * value twice. */
__global__ __launch_bounds__(BLOCK) void
adjoint_wide(const Frame *frames, const uint8_t *mask, const int32_t *donors,
	const double *grad_clean, double *grad_data, const int *reach, const unsigned int *count,
	const unsigned long long *list, unsigned limit)
{
	const int lane = threadIdx.x & 63;
	const unsigned n = min(*count, limit);
	for (unsigned e = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6); e < n;
	     e += gridDim.x * (BLOCK / 64)) {
		const unsigned long long entry = uniform_u64(list[e]);
		const int f = (int)(entry >> 32), q = (int)(entry & 0xffffffffu);
		const Frame &fr = frames[f];
		const int w = fr.w, h = fr.h, x = q % w, y = q / w, R = reach[f];
		const int x0 = max(0, x - R), x1 = min(w - 1, x + R), y0 = max(0, y - R), y1 = min(h - 1, y + R);
		const int ww = x1 - x0 + 1, npos = ww * (y1 - y0 + 1);
		const long long base = fr.off;
		double s = mask[base + q] ? 0.0 : grad_clean[base + q];
		for (int k0 = 0; k0 < npos; k0 += 64) {
			const int k = k0 + lane;
			bool hit = false;
			double term = 0.0;
			if (k < npos) {
				const long long pg = base + (long long)(y0 + k / ww) * w + x0 + k % ww;
				if (mask[pg])
					term = donor_weight_term(donors, grad_clean, pg, q, hit);
			}
			unsigned long long m = __ballot(hit);
			while (m) {
				const int j = __ffsll((long long)m) - 1;
				s += __shfl(term, j, 64);
				m &= m - 1;
			}
		}
		if (!lane)
			grad_data[base + q] = s;
	}
}

__global__ void
selftest_arith(const double *a, const double *b, double *quotient, double *root, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
	     i += (size_t)gridDim.x * blockDim.x) {
		quotient[i] = a[i] / b[i];
		root[i] = sqrt(a[i]);
	}
}

} // namespace cusmic

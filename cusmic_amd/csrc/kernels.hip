/*
 * Cleaning kernels for gfx950 (MI355X), one wave64 = 64 lanes.
 *
 * Pass structure of one call (all on the caller's stream, no host round trip):
 *   reset_counters -> prologue -> replace(fill holes) ->
 *   maxiter x [detect -> replace] -> epilogue
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

__device__ __forceinline__ double
positive(double v)
{
	return v < 0 ? 0.0 : v;
}

/* Laplacian of the 2x-replicated frame (src/filters.cuh:86-105).
 * c, u, d, l, r are the pixel and its up/down/left/right samples; exact is the
 * L.A.Cosmic (scipy) accumulation order, otherwise the reference CUDA order;
 * narrow marks a one-pixel-wide frame, whose 2x2 block sum numpy runs
 * sequentially. */
__device__ __forceinline__ double
laplacian(double c, double u, double d, double l, double r, bool exact, bool narrow)
{
	c = c / 4;
	u = u / 4;
	d = d / 4;
	l = l / 4;
	r = r / 4;
	if (!exact) {
		double up = c - u, down = c - d, left = c - l, right = c - r;
		return (positive(up + left) + positive(up + right)) +
		       (positive(down + left) + positive(down + right));
	}
	double ul = positive(((((0.0 - u) - l) + 4 * c) - c) - c);
	double ur = positive(((((0.0 - u) - c) + 4 * c) - r) - c);
	double dl = positive(((((0.0 - c) - l) + 4 * c) - c) - d);
	double dr = positive(((((0.0 - c) - c) + 4 * c) - r) - d);
	if (narrow)
		return ((ul + ur) + dl) + dr;
	return (ul + ur) + (dl + dr);
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

/* Prologue (src/image.cuh:24-48): holes, exclusions, C = (hole ? 0 : D) + bg,
 * per-frame counts of holes, allowed pixels and finite donors, and the list of
 * holes for the fill. One workgroup per tile. */
__global__ __launch_bounds__(BLOCK) void
prologue(Params p)
{
	__shared__ int red[4];
	int tile = blockIdx.x;
	int f = find_frame(p.frames, p.nf, tile);
	const Frame fr = p.frames[f];
	TilePixels t = tile_of(fr, tile);
	bool bg_map = fr.background != nullptr, bg = bg_map || !isnan(p.bg);
	int holes = 0, allowed = 0, donors = 0;
	if (!threadIdx.x)
		p.tile_frame[tile] = f;
	int x = t.x0 + (threadIdx.x & 63);
	for (int r = threadIdx.x >> 6; r < TILE_H; r += 4) {
		int y = t.y0 + r;
		bool in = x < t.w && y < t.h;
		bool hole = false;
		int i = y * t.w + x;
		if (in) {
			double d = fr.data[i];
			hole = !isfinite(d);
			bool allow = !hole && !(fr.mask && fr.mask[i]);
			double c = hole ? 0.0 : d;
			if (bg)
				c += 1 * (bg_map ? fr.background[i] : p.bg);
			long long g = fr.off + i;
			p.clean[g] = c;
			p.flags[g] = (hole ? HOLE : 0) | (allow ? 0 : EXCL);
			if (p.donors) {
				p.donors[2 * g] = -1;
				p.donors[2 * g + 1] = -1;
			}
			holes += hole;
			allowed += allow;
			donors += allow && isfinite(c);
		}
		/* hole list (only consulted when maxiter > 0) */
		unsigned long long m = __ballot(hole);
		if (m) {
			int lane = threadIdx.x & 63, leader = __ffsll((long long)m) - 1;
			unsigned base = 0;
			if (lane == leader)
				base = atomicAdd(&p.nlist[1], (unsigned)__popcll(m));
			base = __shfl(base, leader, 64);
			if (hole)
				p.list[base + __popcll(m & ((1ull << lane) - 1))] = list_entry(f, i);
		}
	}
	holes = block_sum(holes, red);
	allowed = block_sum(allowed, red);
	donors = block_sum(donors, red);
	if (!threadIdx.x) {
		if (holes)
			atomicAdd(&p.counts[f].holes, holes);
		if (allowed)
			atomicAdd(&p.counts[f].allowed, allowed);
		if (donors)
			atomicAdd(&p.counts[f].donors, donors);
	}
}

/* maxiter = 0 (src/image.cuh:72-92): background arithmetic only. */
__global__ __launch_bounds__(BLOCK) void
copy_without_detection(Params p)
{
	int tile = blockIdx.x;
	int f = find_frame(p.frames, p.nf, tile);
	const Frame fr = p.frames[f];
	TilePixels t = tile_of(fr, tile);
	bool bg_map = fr.background != nullptr, bg = bg_map || !isnan(p.bg);
	int x = t.x0 + (threadIdx.x & 63);
	for (int r = threadIdx.x >> 6; r < TILE_H; r += 4) {
		int y = t.y0 + r;
		if (x >= t.w || y >= t.h)
			continue;
		int i = y * t.w + x;
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

/* Epilogue (src/image.cuh:50-70): out = finite D ? C - bg : D (payload bits
 * kept), mask = flagged bit. */
__global__ __launch_bounds__(BLOCK) void
epilogue(Params p)
{
	int tile = blockIdx.x;
	int f = find_frame(p.frames, p.nf, tile);
	const Frame fr = p.frames[f];
	TilePixels t = tile_of(fr, tile);
	bool bg_map = fr.background != nullptr, bg = bg_map || !isnan(p.bg);
	int x = t.x0 + (threadIdx.x & 63);
	for (int r = threadIdx.x >> 6; r < TILE_H; r += 4) {
		int y = t.y0 + r;
		if (x >= t.w || y >= t.h)
			continue;
		int i = y * t.w + x;
		long long g = fr.off + i;
		uint8_t fl = p.flags[g];
		if (fl & HOLE)
			p.clean[g] = fr.data[i];
		else if (bg)
			p.clean[g] = p.clean[g] + -1 * (bg_map ? fr.background[i] : p.bg);
		p.mask[g] = (fl & CRAY) ? 1 : 0;
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

/* Mark the tiles whose detection may change after pixel (x, y) changed. */
__device__ void
mark_dirty(const Params &p, const Frame &fr, int x, int y, int next)
{
	int tx[3], ty[3], nx = 0, ny = 0;
	if (p.mode == CUSMIC_WRAP) {
		if (fr.w <= 2 * REACH + 1 || fr.tiles_x <= 3) {
			nx = -1;
		} else {
			int xs[3] = {x - REACH, x, x + REACH};
			for (int k = 0; k < 3; ++k) {
				int v = xs[k] < 0 ? xs[k] + fr.w : xs[k] >= fr.w ? xs[k] - fr.w : xs[k];
				tx[nx++] = v / TILE_W;
			}
		}
		if (fr.h <= 2 * REACH + 1 || fr.tiles_y <= 3) {
			ny = -1;
		} else {
			int ys[3] = {y - REACH, y, y + REACH};
			for (int k = 0; k < 3; ++k) {
				int v = ys[k] < 0 ? ys[k] + fr.h : ys[k] >= fr.h ? ys[k] - fr.h : ys[k];
				ty[ny++] = v / TILE_H;
			}
		}
	} else {
		int a = max(0, x - REACH) / TILE_W, b = min(fr.w - 1, x + REACH) / TILE_W;
		for (int v = a; v <= b; ++v)
			tx[nx++] = v;
		a = max(0, y - REACH) / TILE_H;
		b = min(fr.h - 1, y + REACH) / TILE_H;
		for (int v = a; v <= b; ++v)
			ty[ny++] = v;
	}
	int cx = nx < 0 ? fr.tiles_x : nx, cy = ny < 0 ? fr.tiles_y : ny;
	unsigned int *flag = p.dirty + (size_t)(next & 1) * p.ntiles;
	unsigned int *list = p.dirty_list + (size_t)(next & 1) * p.ntiles;
	for (int j = 0; j < cy; ++j)
		for (int i = 0; i < cx; ++i) {
			int tile = fr.tile0 + (ny < 0 ? j : ty[j]) * fr.tiles_x + (nx < 0 ? i : tx[i]);
			if (!atomicExch(&flag[tile], 1u)) {
				unsigned slot = atomicAdd(&p.nlist[2 + next], 1u);
				list[slot] = tile;
			}
		}
}

/* Replace pixel i of frame fr from the donors of its smallest non-empty
 * clipped window. Returns true if the stored bits changed. */
__device__ bool
replace_pixel(const Params &p, const Frame &fr, int i, int32_t *donors)
{
	const double *clean = p.clean + fr.off;
	const uint8_t *flags = p.flags + fr.off;
	int w = fr.w, h = fr.h, x = i % w, y = i / w;
	/* flags and values of the 24 neighbours are loaded together (one memory
	 * round trip); a value read at a flagged neighbour, which another lane may
	 * be rewriting, is discarded */
	double v[24];
	uint8_t fl[24];
	int n = 0, neg0 = 0;
#pragma unroll
	for (int k = 0; k < 24; ++k) {
		int o = k < 12 ? k : k + 1, qx = x + o % 5 - 2, qy = y + o / 5 - 2;
		bool in = qx >= 0 && qx < w && qy >= 0 && qy < h;
		int q = in ? qy * w + qx : i;
		fl[k] = in ? flags[q] : EXCL;
		v[k] = clean[q];
	}
#pragma unroll
	for (int k = 0; k < 24; ++k) {
		bool ok = !(fl[k] & (EXCL | CRAY)) && isfinite(v[k]);
		v[k] = ok ? v[k] : INFINITY;
		n += ok;
		neg0 += ok && (v[k] < 0 || (v[k] == 0.0 && signbit(v[k])));
	}
	double lo, hi;
	int r = 2;
	int x0 = max(0, x - 2), x1 = min(w - 1, x + 2), y0 = max(0, y - 2), y1 = min(h - 1, y + 2);
	if (n) {
		double s[13];
		sort_low13_of24(v, s);
		int a = (n - 1) >> 1, b = n >> 1;
		lo = s[0];
		hi = s[0];
#pragma unroll
		for (int j = 1; j < 13; ++j) {
			lo = j == a ? s[j] : lo;
			hi = j == b ? s[j] : hi;
		}
	} else {
		/* rare: widen the window until it holds a donor */
		int rmax = max(w, h);
		for (r = 3; r <= rmax; ++r) {
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
		lo = key_value(window_select(flags, clean, w, x0, x1, y0, y1, (n - 1) >> 1));
		hi = n & 1 ? lo : key_value(window_select(flags, clean, w, x0, x1, y0, y1, n >> 1));
		neg0 = 0;
		for (int yy = y0; yy <= y1; ++yy)
			for (int xx = x0; xx <= x1; ++xx) {
				double t;
				if (is_donor(flags, clean, yy * w + xx, t))
					neg0 += t < 0 || (t == 0.0 && signbit(t));
			}
	}
	double out;
	if (n & 1)
		out = 0.0 + lo;
	else if (lo == 0.0 && hi == 0.0)
		out = neg0 > (n >> 1) ? -0.0 : 0.0;
	else
		out = midpoint(lo, hi);
	if (donors) {
		int a = (n - 1) >> 1, b = n >> 1;
		int ia = rank_index(flags, clean, w, x0, x1, y0, y1, rank_key(lo, a, neg0), a);
		int ib = a == b ? ia : rank_index(flags, clean, w, x0, x1, y0, y1, rank_key(hi, b, neg0), b);
		donors[2 * (fr.off + i)] = ia;
		donors[2 * (fr.off + i) + 1] = ib;
	}
	double *dst = p.clean + fr.off + i;
	double old = *dst;
	*dst = out;
	return __double_as_longlong(old) != __double_as_longlong(out);
}

/* Hole fill before the loop (fill = 1, src/cleaner.cuh:78-100) or the
 * replacement pass of iteration iter over every flagged pixel. */
__global__ __launch_bounds__(BLOCK) void
replace_pass(Params p, int iter, int fill)
{
	unsigned n = fill ? p.nlist[1] : p.nlist[0];
	for (unsigned e = blockIdx.x * BLOCK + threadIdx.x; e < n; e += gridDim.x * BLOCK) {
		unsigned long long entry = p.list[e];
		int f = (int)(entry >> 32), i = (int)(entry & 0xffffffffu);
		const Frame fr = p.frames[f];
		Counts c = p.counts[f];
		if (fill) {
			if (!c.allowed)
				continue;
			if (!c.donors) {
				atomicOr(p.status, ERR_NO_DONORS);
				continue;
			}
			replace_pixel(p, fr, i, nullptr);
		} else {
			if (!p.todo[iter * p.nf + f])
				continue;
			if (c.donors <= 0) {
				atomicOr(p.status, ERR_NO_DONORS);
				continue;
			}
			bool changed = replace_pixel(p, fr, i, p.donors);
			if (changed && iter + 1 < p.maxiter)
				mark_dirty(p, fr, i % fr.w, i / fr.w, iter + 1);
		}
	}
}

/* ------------------------------------------------------------------------ */
/* Detection (src/filters.cuh:86-193, src/cleaner.cuh:22-37,102-126)         */

/* LDS regions, in slots relative to the tile origin: the clean image with a
 * halo of 6, the significance with a halo of 4, per-pixel bits and growth
 * with halos of 2 and 1. */
constexpr int HC = 6, HS = 4, HP = 2;
constexpr int CW = TILE_W + 2 * HC, CH = TILE_H + 2 * HC;  /* 76 x 44 */
constexpr int SW = TILE_W + 2 * HS, SH = TILE_H + 2 * HS;  /* 72 x 40 */
constexpr int PW = TILE_W + 2 * HP, PH = TILE_H + 2 * HP;  /* 68 x 36 */
constexpr int GW = TILE_W + 2, GH = TILE_H + 2;            /* 66 x 34 */
constexpr int PER_THREAD = (PW * PH + BLOCK - 1) / BLOCK;  /* 10 */
constexpr int C_PER = (CW * CH + BLOCK - 1) / BLOCK;       /* 14 */
constexpr int S_PER = (SW * SH + BLOCK - 1) / BLOCK;       /* 12 */
constexpr int P_PER = PER_THREAD;

/* bits of the per-slot byte over the halo-2 region */
constexpr uint8_t S_HOLE = 1, S_EXCL = 2, S_IN = 8, S_A = 16, S_CAND = 32;

struct DetectShared {
	double c[CH][CW];
	double s[SH][SW];
	uint8_t bits[PH][PW];
	uint8_t g1[GH][GW];
	uint16_t alist[PW * PH];
	uint16_t blist[PW * PH];
	int na, nb, tile, unsafe;
	int red[4];
};

__global__ __launch_bounds__(BLOCK, 2) void
detect(Params p, int iter)
{
	__shared__ DetectShared sh;
	const int tid = threadIdx.x;
	const int mode = p.mode;
	const bool wrap = mode == CUSMIC_WRAP;
	const bool exact = p.laplacian == 0;
	const double t_cr = p.threshold, t_nb = p.neighbor, t_min = fmin(t_cr, t_nb);
	unsigned limit = iter == 0 ? (unsigned)p.ntiles : p.nlist[2 + iter];
	const unsigned int *dlist = p.dirty_list + (size_t)(iter & 1) * p.ntiles;
	/* thread 0 keeps the claim on the following tile in a register, so the
	 * atomic's latency overlaps the current tile's loads */
	unsigned next = 0;
	if (!tid)
		next = atomicAdd(&p.queue[iter], 1u);

	for (;;) {
		__syncthreads(); /* the previous tile's shared state is no longer read */
		if (!tid) {
			unsigned q = next;
			sh.tile = q < limit ? (iter == 0 ? (int)q : (int)dlist[q]) : -1;
			sh.na = sh.nb = sh.unsafe = 0;
			if (q < limit)
				next = atomicAdd(&p.queue[iter], 1u);
		}
		__syncthreads();
		const int tile = sh.tile;
		if (tile < 0)
			break;
		if (iter > 0 && !tid)
			p.dirty[(size_t)(iter & 1) * p.ntiles + tile] = 0;
		const int f = p.tile_frame[tile];
		const Frame fr = p.frames[f];
		const TilePixels t = tile_of(fr, tile);
		const int w = t.w, h = t.h, X0 = t.x0, Y0 = t.y0;
		const double *clean = p.clean + fr.off;
		const bool narrow = w == 1;
		if (!tid)
			atomicAdd(&p.work[iter], (unsigned long long)(min(TILE_W, w - X0) * min(TILE_H, h - Y0)));
		/* the halo-6 window lies inside the frame: no border mapping anywhere */
		const bool inner = X0 >= HC && Y0 >= HC && X0 + TILE_W + HC <= w && Y0 + TILE_H + HC <= h;

		/* 1. issue every global load of the tile before the first use: the clean
		 * image with halo 6 (median border mode), the flags with halo 2 and the
		 * error map with halo 4. The slot arithmetic is recomputed per tile
		 * (lane is opaque to the optimiser) rather than held in registers
		 * across the persistent loop. */
		int lane = tid;
		asm volatile("" : "+v"(lane));
		double cv[C_PER];
#pragma unroll
		for (int j = 0; j < C_PER; ++j) {
			int k = lane + j * BLOCK;
			cv[j] = 0.0;
			if (k < CW * CH) {
				int sx = k % CW, sy = k / CW;
				int bx = X0 - HC + sx, by = Y0 - HC + sy;
				if (!inner) {
					bx = border_index(bx, w, mode);
					by = border_index(by, h, mode);
				}
				if (bx >= 0 && by >= 0)
					cv[j] = clean[by * w + bx];
			}
		}
		uint8_t fv[P_PER];
#pragma unroll
		for (int j = 0; j < P_PER; ++j) {
			int k = lane + j * BLOCK;
			fv[j] = 0;
			if (k < PW * PH) {
				int gx = X0 - HP + k % PW, gy = Y0 - HP + k / PW;
				if (gx >= 0 && gx < w && gy >= 0 && gy < h)
					fv[j] = (uint8_t)((p.flags[fr.off + gy * w + gx] & (HOLE | EXCL)) | S_IN);
			}
		}
		double ev[S_PER];
		if (fr.error) {
#pragma unroll
			for (int j = 0; j < S_PER; ++j) {
				int k = lane + j * BLOCK;
				ev[j] = 1.0;
				if (k < SW * SH) {
					int gx = X0 - HS + k % SW, gy = Y0 - HS + k / SW;
					bool in = gx >= 0 && gx < w && gy >= 0 && gy < h;
					if (wrap && !in) {
						gx = border_index(gx, w, mode);
						gy = border_index(gy, h, mode);
						in = true;
					}
					if (in)
						ev[j] = fr.error[gy * w + gx];
				}
			}
			/* staged in the significance array; each slot is later read and
			 * overwritten by the thread that stored it */
#pragma unroll
			for (int j = 0; j < S_PER; ++j) {
				int k = lane + j * BLOCK;
				if (k < SW * SH)
					sh.s[k / SW][k % SW] = ev[j];
			}
		}
#pragma unroll
		for (int j = 0; j < C_PER; ++j) {
			int k = lane + j * BLOCK;
			if (k < CW * CH)
				sh.c[k / CW][k % CW] = cv[j];
		}
#pragma unroll
		for (int j = 0; j < P_PER; ++j) {
			int k = lane + j * BLOCK;
			if (k < PW * PH)
				sh.bits[k / PW][k % PW] = fv[j];
		}
		__syncthreads();

		/* 2. significance S = L / (2 N) over the halo-4 region: in-frame slots,
		 * or every slot under WRAP, whose stencils commute with the periodic
		 * extension. */
		if (fr.error) {
			for (int k = tid; k < SW * SH; k += BLOCK) {
				int sx = k % SW, sy = k / SW, gx = X0 - HS + sx, gy = Y0 - HS + sy;
				bool in = gx >= 0 && gx < w && gy >= 0 && gy < h;
				if (!in && !wrap)
					continue;
				int cx = sx + (HC - HS), cy = sy + (HC - HS);
				double c = sh.c[cy][cx];
				double u, d, l, r;
				if (wrap || mode == CUSMIC_CONSTANT) {
					u = sh.c[cy - 1][cx];
					d = sh.c[cy + 1][cx];
					l = sh.c[cy][cx - 1];
					r = sh.c[cy][cx + 1];
				} else {
					/* mirror/reflect/nearest: a replicated edge samples itself */
					u = gy > 0 ? sh.c[cy - 1][cx] : c;
					d = gy < h - 1 ? sh.c[cy + 1][cx] : c;
					l = gx > 0 ? sh.c[cy][cx - 1] : c;
					r = gx < w - 1 ? sh.c[cy][cx + 1] : c;
				}
				double nz = sh.s[sy][sx];
				if (!(nz > 0) && !isnan(nz))
					sh.unsafe = 1;
				sh.s[sy][sx] = laplacian(c, u, d, l, r, exact, narrow) / (2 * nz);
			}
		} else {
			/* noise model: N from the 5x5 median of C, 4 rows per item */
			constexpr int STRIPS = SH / 4;
			for (int k = tid; k < SW * STRIPS; k += BLOCK) {
				int sx = k % SW, sy0 = (k / SW) * 4;
				double rows[8 * 5], med[4];
#pragma unroll
				for (int j = 0; j < 8; ++j)
#pragma unroll
					for (int c = 0; c < 5; ++c)
						rows[j * 5 + c] = sh.c[sy0 + j][sx + c];
				strip_median5x4(rows, med);
#pragma unroll
				for (int o = 0; o < 4; ++o) {
					int sy = sy0 + o, gx = X0 - HS + sx, gy = Y0 - HS + sy;
					bool in = gx >= 0 && gx < w && gy >= 0 && gy < h;
					if (!in && !wrap)
						continue;
					int cx = sx + (HC - HS), cy = sy + (HC - HS);
					double c = sh.c[cy][cx];
					double u, d, l, r;
					if (wrap || mode == CUSMIC_CONSTANT) {
						u = sh.c[cy - 1][cx];
						d = sh.c[cy + 1][cx];
						l = sh.c[cy][cx - 1];
						r = sh.c[cy][cx + 1];
					} else {
						u = gy > 0 ? sh.c[cy - 1][cx] : c;
						d = gy < h - 1 ? sh.c[cy + 1][cx] : c;
						l = gx > 0 ? sh.c[cy][cx - 1] : c;
						r = gx < w - 1 ? sh.c[cy][cx + 1] : c;
					}
					int bx = border_index(gx, w, mode), by = border_index(gy, h, mode);
					double g = fr.gain ? fr.gain[by * w + bx] : fr.g;
					double rn = fr.readnoise ? fr.readnoise[by * w + bx] : fr.rn;
					double nz = noise_model(med[o], g, rn);
					if (!(nz > 0) && !isnan(nz))
						sh.unsafe = 1;
					sh.s[sy][sx] = laplacian(c, u, d, l, r, exact, narrow) / (2 * nz);
				}
			}
		}
		__syncthreads();

		/* 3. out-of-frame significance slots read by the 5x5 median */
		const bool edge = X0 - HS < 0 || Y0 - HS < 0 || X0 + TILE_W + HS > w || Y0 + TILE_H + HS > h;
		if (edge && !wrap) {
			for (int k = tid; k < SW * SH; k += BLOCK) {
				int sx = k % SW, sy = k / SW, gx = X0 - HS + sx, gy = Y0 - HS + sy;
				if (gx >= 0 && gx < w && gy >= 0 && gy < h)
					continue;
				if (gx < -2 || gx > w + 1 || gy < -2 || gy > h + 1)
					continue;
				double v = 0.0;
				if (mode != CUSMIC_CONSTANT) {
					int bx = border_index(gx, w, mode), by = border_index(gy, h, mode);
					v = sh.s[by - (Y0 - HS)][bx - (X0 - HS)];
				}
				sh.s[sy][sx] = v;
			}
			__syncthreads();
		}

		/* 4. screen: S' = S - med5(S) <= S because S >= 0 and so med5(S) >= 0;
		 * where S <= min(t, t_n) no threshold test can pass. An invalid noise
		 * value (<= 0) disables the screen for the tile. */
		const bool unsafe = sh.unsafe != 0;
		for (int k = tid; k < PW * PH; k += BLOCK) {
			int sx = k % PW, sy = k / PW;
			uint8_t b = sh.bits[sy][sx];
			bool a = (b & S_IN) && !(b & S_HOLE) &&
				 (unsafe || sh.s[sy + (HS - HP)][sx + (HS - HP)] > t_min);
			lds_append(a, (uint16_t)k, sh.alist, &sh.na);
		}
		__syncthreads();

		/* 5. S' at screened pixels; written back after every read is done */
		const int na = sh.na;
		double sp[PER_THREAD];
#pragma unroll
		for (int j = 0; j < PER_THREAD; ++j) {
			int idx = tid + j * BLOCK;
			sp[j] = 0.0;
			if (idx < na) {
				int pos = sh.alist[idx], sx = pos % PW + (HS - HP), sy = pos / PW + (HS - HP);
				double v[25], m;
#pragma unroll
				for (int dy = 0; dy < 5; ++dy)
#pragma unroll
					for (int dx = 0; dx < 5; ++dx)
						v[dy * 5 + dx] = sh.s[sy - 2 + dy][sx - 2 + dx];
				median25(v, &m);
				sp[j] = sh.s[sy][sx] - m;
			}
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < PER_THREAD; ++j) {
			int idx = tid + j * BLOCK;
			bool b = false;
			uint16_t pos = 0;
			if (idx < na) {
				pos = sh.alist[idx];
				int sx = pos % PW, sy = pos / PW;
				sh.s[sy + (HS - HP)][sx + (HS - HP)] = sp[j];
				sh.bits[sy][sx] |= S_A;
				b = !(sh.bits[sy][sx] & S_EXCL) && sp[j] > t_cr;
			}
			if (j * BLOCK < na)
				lds_append(b, pos, sh.blist, &sh.nb);
		}
		__syncthreads();
		const int nb = sh.nb;
		if (!nb)
			continue; /* no seed candidates near this tile: nothing new */

		/* 6. fine structure at seed candidates: F = max((m3 - med7(m3)) / N,
		 * 0.01) with m3 = med3(C); cand = S' / F > contrast */
		for (int idx = tid; idx < nb; idx += BLOCK) {
			int pos = sh.blist[idx], px = pos % PW, py = pos / PW;
			int gx = X0 - HP + px, gy = Y0 - HP + py;
			double m3[49];
#pragma unroll
			for (int dy = -3; dy <= 3; ++dy)
#pragma unroll
				for (int dx = -3; dx <= 3; ++dx) {
					int qx = gx + dx, qy = gy + dy;
					bool zero = false;
					if (!wrap && (qx < 0 || qx >= w || qy < 0 || qy >= h)) {
						if (mode == CUSMIC_CONSTANT)
							zero = true;
						else {
							qx = border_index(qx, w, mode);
							qy = border_index(qy, h, mode);
						}
					}
					double val = 0.0;
					if (!zero) {
						int cx = qx - X0 + HC, cy = qy - Y0 + HC;
						double v9[9];
#pragma unroll
						for (int j = 0; j < 3; ++j)
#pragma unroll
							for (int i = 0; i < 3; ++i)
								v9[j * 3 + i] = sh.c[cy - 1 + j][cx - 1 + i];
						median9(v9, &val);
					}
					m3[(dy + 3) * 7 + dx + 3] = val;
				}
			double med7;
			median49(m3, &med7);
			double nz;
			if (fr.error) {
				nz = fr.error[gy * w + gx];
			} else {
				int cx = gx - X0 + HC, cy = gy - Y0 + HC;
				double v25[25], m5;
#pragma unroll
				for (int j = 0; j < 5; ++j)
#pragma unroll
					for (int i = 0; i < 5; ++i)
						v25[j * 5 + i] = sh.c[cy - 2 + j][cx - 2 + i];
				median25(v25, &m5);
				double g = fr.gain ? fr.gain[gy * w + gx] : fr.g;
				double rn = fr.readnoise ? fr.readnoise[gy * w + gx] : fr.rn;
				nz = noise_model(m5, g, rn);
			}
			double fine = (m3[24] - med7) / nz;
			fine = fine < FINE_FLOOR ? FINE_FLOOR : fine;
			double sig = sh.s[py + (HS - HP)][px + (HS - HP)];
			if (sig / fine > p.contrast)
				sh.bits[py][px] |= S_CAND;
		}
		__syncthreads();

		/* 7. first growth at the cosmic-ray threshold over the halo-1 region */
		for (int k = tid; k < GW * GH; k += BLOCK) {
			int gxs = k % GW, gys = k / GW, px = gxs + 1, py = gys + 1;
			uint8_t b = sh.bits[py][px];
			bool g = false;
			if ((b & S_A) && sh.s[py + (HS - HP)][px + (HS - HP)] > t_cr) {
#pragma unroll
				for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
					for (int dx = -1; dx <= 1; ++dx)
						g |= (sh.bits[py + dy][px + dx] & S_CAND) != 0;
			}
			sh.g1[gys][gxs] = g;
		}
		__syncthreads();

		/* 8. second growth at the neighbour threshold over the tile, then
		 * accumulate: new = grown & !flagged (src/cleaner.cuh:22-37) */
		int fresh = 0, lost = 0;
		for (int k = tid; k < TILE_W * TILE_H; k += BLOCK) {
			int tx = k % TILE_W, ty = k / TILE_W, px = tx + HP, py = ty + HP;
			uint8_t b = sh.bits[py][px];
			bool g = false;
			if ((b & S_A) && sh.s[py + (HS - HP)][px + (HS - HP)] > t_nb) {
#pragma unroll
				for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
					for (int dx = -1; dx <= 1; ++dx)
						g |= sh.g1[ty + 1 + dy][tx + 1 + dx] != 0;
			}
			bool isnew = false;
			int i = (Y0 + ty) * w + X0 + tx;
			if (g) {
				uint8_t fl = p.flags[fr.off + i];
				if (!(fl & CRAY)) {
					isnew = true;
					p.flags[fr.off + i] = fl | CRAY;
					fresh += 1;
					lost += !(fl & EXCL) && isfinite(sh.c[ty + HC][tx + HC]);
				}
			}
			unsigned long long m = __ballot(isnew);
			if (m) {
				int lane = tid & 63, leader = __ffsll((long long)m) - 1;
				unsigned base = 0;
				if (lane == leader)
					base = atomicAdd(&p.nlist[0], (unsigned)__popcll(m));
				base = __shfl(base, leader, 64);
				if (isnew)
					p.list[base + __popcll(m & ((1ull << lane) - 1))] = list_entry(f, i);
			}
		}
		fresh = block_sum(fresh, sh.red);
		lost = block_sum(lost, sh.red);
		if (!tid) {
			if (fresh)
				atomicAdd(&p.todo[iter * p.nf + f], fresh);
			if (lost)
				atomicSub(&p.counts[f].donors, lost);
		}
	}
}

/* ------------------------------------------------------------------------ */
/* Adjoint with the mask held fixed                                          */

/* Chebyshev reach of each frame's recorded donors (>= 2), for the gather. */
__global__ __launch_bounds__(BLOCK) void
donor_reach(const Frame *frames, int nf, const uint8_t *mask, const int32_t *donors, int *reach)
{
	int tile = blockIdx.x;
	int f = find_frame(frames, nf, tile);
	const Frame fr = frames[f];
	TilePixels t = tile_of(fr, tile);
	int x = t.x0 + (threadIdx.x & 63), r = 0;
	for (int row = threadIdx.x >> 6; row < TILE_H; row += 4) {
		int y = t.y0 + row;
		if (x >= t.w || y >= t.h)
			continue;
		long long g = fr.off + (long long)y * t.w + x;
		if (!mask[g])
			continue;
		for (int k = 0; k < 2; ++k) {
			int d = donors[2 * g + k];
			if (d < 0)
				continue;
			int dx = abs(d % t.w - x), dy = abs(d / t.w - y);
			r = max(r, max(dx, dy));
		}
	}
	for (int o = 32; o; o >>= 1)
		r = max(r, __shfl_xor(r, o, 64));
	if (!(threadIdx.x & 63) && r > 2)
		atomicMax(&reach[f], r);
}

/* grad_data[q] = (q unflagged) * g[q] + sum over flagged p, in raster order of
 * p, of g[p] times p's weight on donor q. */
__global__ __launch_bounds__(BLOCK) void
adjoint_gather(const Frame *frames, int nf, const uint8_t *mask, const int32_t *donors,
	const int *reach, const double *grad_clean, double *grad_data)
{
	int tile = blockIdx.x;
	int f = find_frame(frames, nf, tile);
	const Frame fr = frames[f];
	TilePixels t = tile_of(fr, tile);
	const int R = max(2, reach[f]);
	int x = t.x0 + (threadIdx.x & 63);
	for (int row = threadIdx.x >> 6; row < TILE_H; row += 4) {
		int y = t.y0 + row;
		if (x >= t.w || y >= t.h)
			continue;
		int q = y * t.w + x;
		const long long base = fr.off;
		double s = mask[base + q] ? 0.0 : grad_clean[base + q];
		for (int py = max(0, y - R); py <= min(t.h - 1, y + R); ++py)
			for (int px = max(0, x - R); px <= min(t.w - 1, x + R); ++px) {
				long long pg = base + (long long)py * t.w + px;
				if (!mask[pg])
					continue;
				int lo = donors[2 * pg], hi = donors[2 * pg + 1];
				if (lo == hi) {
					if (lo == q)
						s += grad_clean[pg];
				} else {
					if (lo == q)
						s += 0.5 * grad_clean[pg];
					if (hi == q)
						s += 0.5 * grad_clean[pg];
				}
			}
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

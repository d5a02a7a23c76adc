/*
 * Device-side data layout shared by the cleaning kernels (gfx950).
 *
 * A call cleans a batch of frames of individual shapes. Every per-pixel plane
 * (clean, flags, mask, donors) is the frames' pixels concatenated in
 * descriptor order; frame f starts at Frame::off. Frames are cut into
 * tiles of Params::tile_w x TILE_H numbered frame-major, row-major; Frame::tile0 is the
 * first tile of frame f.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/cusmic.h"

namespace cusmic {

/* Detection tile: 58 x 32 outputs, so that the significance region with the
 * 2-pixel halo of the 5x5 median and the 1-pixel ring of the Laplacian is
 * exactly one wave (64 lanes) wide. The first pass over error-map frames
 * takes wide tiles of TILE_W2 x TILE_H, two columns per lane (120 + 2 x 4 =
 * 128); Params::tile_w says which width a call's tile table uses. */
constexpr int TILE_W = 58;
constexpr int TILE_W2 = 120;
constexpr int TILE_H = 32;
constexpr int BLOCK = 256;

/* The first detection pass walks each frame in runs of up to RUN_TILES
 * vertically adjacent tiles, so that the significance rows shared by two
 * tiles of a run are computed once. A call with few tiles uses shorter runs
 * (Params::run_tiles, a power of two), so that its workgroups get several
 * runs each and the last ones do not finish long after the others. */
constexpr int RUN_TILES = 8;

/* Passes after the first recompute detection on BLK x BLK blocks marked
 * dirty by the previous replacement. */
constexpr int BLK = 8;

/* Per-pixel flags (u8). The first pass marks the pixels where it computed S'
 * with SCR. A block that a block pass has processed (Params::block_pass
 * nonzero) keeps in a u16 plane, for each of its pixels, the stamp of the pass
 * that computed S' there, or 0: S' is current iff that stamp equals the
 * block's. */
constexpr uint32_t HOLE = 1;    /* input pixel is nonfinite */
constexpr uint32_t EXCL = 2;    /* hole or masked: never a seed or donor */
constexpr uint32_t CRAY = 4;    /* flagged as cosmic ray (the output mask) */
constexpr uint32_t GROW = 8;    /* on this pass's first-growth list */
constexpr uint32_t NEWP = 16;   /* flagged in the latest detection pass */
constexpr uint32_t REDO = 32;   /* older flagged pixel on this pass's redo list */
constexpr uint32_t WIDE = 64;   /* on the list of pixels replaced from a
				 * window wider than 5 x 5 */
constexpr uint32_t SCR = 128;   /* the first pass computed S' here */

constexpr double FINE_FLOOR = 0.01;   /* src/filters.cuh:21 */
constexpr double SIGNAL_FLOOR = 1e-5; /* src/filters.cuh:22 */

/* Detection reach: a change of the clean image at one pixel can change the
 * contribution of a candidate (its own test and its two growth steps) only
 * within this Chebyshev radius: S' reaches 3 with an error map (Laplacian 1 +
 * median of S 2) and 4 under the noise model (its N is a 5x5 median of C),
 * growth adds 2; the fine structure reaches 4 (median 1 + median 3), below
 * either. */
constexpr int REACH_MAPS = 5;
constexpr int REACH = 6;
/* Reach of a change on S' itself: blocks within it recompute S'; blocks only
 * within the detection reach keep their S' and re-emit their seeds. */
constexpr int REACH_S_MAPS = 3;
constexpr int REACH_S = 4;

/* Sides of a block that a change at in-block offset (ox, oy) reaches: bits
 * 0-3 for the detection reach, bits 4-7 for the S' reach. */
constexpr unsigned SIDE_L = 1, SIDE_R = 2, SIDE_U = 4, SIDE_D = 8;
constexpr int SIDES_S = 4; /* shift of the S' reach's sides */

/* Dirty-list entries of blocks that only re-emit seeds carry this bit in their
 * block index. */
constexpr unsigned RESEED = 1u << 31;

struct Frame {
	const double *data, *error, *gain, *readnoise, *background;
	const uint8_t *mask;
	double g, rn;          /* scalar gain and read noise of this frame */
	long long off;         /* first pixel in the batch planes */
	int w, h;
	int tile0, tiles_x, tiles_y;
	int chunk0;            /* first CHUNK-pixel chunk of the frame */
	int run0;              /* first first-pass run of the frame */
	int block0, blocks_x;  /* first dirty block, blocks per block row */
};

/* Pointwise passes stream frames in chunks of CHUNK consecutive pixels. */
constexpr int CHUNK = 8192;

/* Per-frame counters are updated in SHARDS copies, one per group of
 * workgroups (blockIdx.x % SHARDS), and folded into the totals by a small
 * kernel: many workgroups adding to one address serialise at the memory side
 * (hundreds of nanoseconds each on this device). */
constexpr int SHARDS = 32;

enum { SHARD_HOLES, SHARD_ALLOWED, SHARD_DONORS, SHARD_FRESH, SHARD_FIELDS };

/* Accelerator dies (XCDs) of the device, each with its own L2; workgroups
 * are dispatched to them in turn. */
constexpr int XCDS = 8;

/* Per-frame counters of one call, zeroed at its start. fresh_passes counts
 * the detection passes that flagged a new pixel: the frame ran fresh_passes
 * replacement passes and min(fresh_passes + 1, maxiter) detection passes. */
struct Counts {
	int holes, allowed, donors, fresh_passes;
};

/* The host enqueues the detection loop in chunks of CHUNK_PASSES passes and
 * stops after a chunk in which no frame flagged a new pixel (the reference
 * loop's early exit, src/cleaner.cuh:175-181). The per-pass counters live in
 * a ring of PASS_RING slots: a pass writes its successor's changed-block
 * count and reads its predecessor's list start, so a slot is cleared before
 * the chunk that reuses it and stays intact for the pass after it. */
constexpr int CHUNK_PASSES = 4;
constexpr int PASS_RING = 2 * CHUNK_PASSES;

/* Params::nlist: global counters, then PASS_FIELDS counters per ring slot. */
enum {
	NL_LIST,        /* length of the flagged-pixel list (from the bottom of list) */
	NL_HOLES,       /* length of the hole list (from the top of list) */
	NL_WIDE_ALL,    /* length of wide_all */
	NL_FILL_WIDE,   /* widened replacements of the hole fill */
	NL_GLOBALS
};
enum {
	PF_DIRTY,       /* dirty blocks of the pass */
	PF_SEED,        /* seeds */
	PF_CAND,        /* candidates */
	PF_GROW,        /* first-growth pixels */
	PF_WIDE,        /* widened replacements */
	PF_CHANGED,     /* blocks changed by the previous pass's replacement */
	PF_START,       /* flagged-list length before the pass's new pixels */
	PF_REDO,        /* redo list */
	PF_WIDE_BASE,   /* wide_all length before the pass */
	PF_EDGE,        /* seeds near a frame edge */
	PF_ACTIVE,      /* nonzero if some frame flagged a new pixel */
	PASS_FIELDS
};

struct Params {
	const Frame *frames;
	int nf, ntiles, maxiter, mode, laplacian;
	int run_tiles;                   /* tiles per first-pass run (<= RUN_TILES) */
	int tile_w;                      /* tile width of the tile table: TILE_W or TILE_W2 */
	int reach, reach_s;              /* REACH_MAPS or REACH, REACH_S_MAPS or REACH_S */
	double contrast, threshold, neighbor, bg;  /* bg: scalar background or NaN */
	double *clean;
	uint8_t *flags;
	uint16_t *stamp;                 /* S' stamps of processed blocks */
	uint8_t *mask;
	double *sprime;                  /* S' where the stamp is current */
	unsigned long long *list_b;      /* seed candidates, then first-growth pixels */
	unsigned long long *list_c;      /* candidates that passed the contrast test,
					  * then this pass's widened replacements */
	unsigned long long *wide_all;    /* every flagged pixel replaced from a
					  * window wider than 5 x 5 so far */
	int32_t *donors;                 /* 2 per pixel or null */
	unsigned long long *list;        /* hole list, then the flagged-pixel list */
	Counts *counts;                  /* per frame */
	int *shard;                      /* [nf][SHARDS][SHARD_FIELDS] pending counts */
	int *todo;                       /* [PASS_RING][nf] new pixels of the pass
					  * in each ring slot */
	unsigned int *queue;             /* [XCDS] run queue heads of the first pass */
	unsigned int *nlist;             /* NL_* globals, then [PASS_RING][PASS_FIELDS] */
	unsigned int *changed_tag;       /* [nblocks] stamp of the last pass whose
					  * preceding replacement changed a pixel of
					  * the block */
	unsigned int *changed_sides;     /* [nblocks] SIDE_* bits the block's
					  * changes reach; reset when expanded */
	unsigned int *reseed_tag;        /* [nblocks] stamp of the last pass that
					  * listed the block to re-emit its seeds */
	unsigned int *dirty_tag;         /* [nblocks] stamp of the last pass that
					  * listed the block as dirty */
	unsigned long long *changed_list; /* frame << 32 | block, per pass */
	unsigned long long *edge_list;   /* frame << 32 | pixel of the seeds of a
					  * pass near a frame edge (not fine_inner) */
	unsigned long long *dirty_list;  /* frame << 32 | block of the blocks the
					  * block pass processes */
	int *status;                     /* [0] error bits, [1] nonzero if a clean
					  * sample is infinite (D + bg overflowed) */
	unsigned long long *work;        /* in-frame pixels that detection processed */
	int *chunk_frame;                /* [nchunks] frame of each chunk */
	int nchunks;
	int *run_frame;                  /* [nruns] frame of each first-pass run */
	int nruns;
	int nblocks;
	unsigned list_cap;               /* entries that list holds: the flagged
					  * pixels from the bottom, the holes (hole
					  * list) from the top */
	unsigned int *block_pass;        /* [nblocks] stamp of the last block pass
					  * that processed the block, or 0 */
	unsigned int *tile_skip;         /* [ntiles] nonzero if the first pass left
					  * the tile to the block pass after the hole
					  * fill */
};

/* Runs per tile column of a frame; runs are numbered frame-major, then
 * row-major over blocks of run_tiles tile rows. */
__device__ __host__ __forceinline__ int
runs_y(int tiles_y, int run_tiles)
{
	return (tiles_y + run_tiles - 1) / run_tiles;
}

constexpr int ERR_NO_DONORS = 1;

/* S' stamp of pass iter, as kept in the u16 stamp plane and in block_pass (0
 * is never a stamp). A block pass writes a block's stamp and the stamps of
 * all its pixels together, so the values may repeat every 65535 passes. */
__device__ __forceinline__ unsigned
pass_stamp(int iter)
{
	return unsigned(iter % 65535) + 1;
}

/* Tag of pass iter in the per-block list tags (changed, dirty, reseed): never
 * 0 and never repeated within a call (iter < maxiter <= INT_MAX). */
__device__ __forceinline__ unsigned
tag_stamp(int iter)
{
	return unsigned(iter) + 1;
}

__device__ __host__ __forceinline__ int
pass_slot(int iter)
{
	return iter % PASS_RING;
}

/* Offsets of the per-pass list counters in Params::nlist. */
__device__ __host__ __forceinline__ int
nl_pass(int iter, int field)
{
	return NL_GLOBALS + pass_slot(iter) * PASS_FIELDS + field;
}
__device__ __forceinline__ int nl_dirty(const Params &, int it) { return nl_pass(it, PF_DIRTY); }
__device__ __forceinline__ int nl_seed(const Params &, int it) { return nl_pass(it, PF_SEED); }
__device__ __forceinline__ int nl_cand(const Params &, int it) { return nl_pass(it, PF_CAND); }
__device__ __forceinline__ int nl_grow(const Params &, int it) { return nl_pass(it, PF_GROW); }
__device__ __forceinline__ int nl_wide(const Params &, int it) { return nl_pass(it, PF_WIDE); }
__device__ __forceinline__ int nl_changed(const Params &, int it) { return nl_pass(it, PF_CHANGED); }
__device__ __forceinline__ int nl_start(const Params &, int it) { return nl_pass(it, PF_START); }
__device__ __forceinline__ int nl_redo(const Params &, int it) { return nl_pass(it, PF_REDO); }
__device__ __forceinline__ int nl_wide_base(const Params &, int it) { return nl_pass(it, PF_WIDE_BASE); }
__device__ __forceinline__ int nl_edge(const Params &, int it) { return nl_pass(it, PF_EDGE); }
__device__ __forceinline__ int nl_active(const Params &, int it) { return nl_pass(it, PF_ACTIVE); }
__device__ __forceinline__ int nl_wide_all(const Params &) { return NL_WIDE_ALL; }

/* Row of pass iter in Params::todo. */
__device__ __forceinline__ int
todo_index(const Params &p, int iter, int f)
{
	return pass_slot(iter) * p.nf + f;
}

/* scipy.ndimage border mapping of coordinate x on an axis of n samples;
 * -1 means a CONSTANT-mode zero (src/filters.cuh:24-44). */
__device__ __forceinline__ int
border_index(int x, int n, int mode)
{
	if (x >= 0 && x < n)
		return x;
	if (mode == CUSMIC_CONSTANT)
		return -1;
	if (mode == CUSMIC_NEAREST)
		return x < 0 ? 0 : n - 1;
	if (n == 1)
		return 0;
	if (mode == CUSMIC_WRAP) {
		x %= n;
		return x < 0 ? x + n : x;
	}
	int p = mode == CUSMIC_REFLECT ? 2 * n : 2 * (n - 1);
	x %= p;
	if (x < 0)
		x += p;
	if (x < n)
		return x;
	return mode == CUSMIC_REFLECT ? p - 1 - x : p - x;
}

/* Order-preserving key: -0 below +0 (src/replace.cuh:37-48). */
__device__ __forceinline__ unsigned long long
order_key(double v)
{
	unsigned long long b = (unsigned long long)__double_as_longlong(v);
	return (b >> 63) ? ~b : b ^ (1ull << 63);
}

__device__ __forceinline__ double
key_value(unsigned long long k)
{
	return __longlong_as_double((long long)((k >> 63) ? k ^ (1ull << 63) : ~k));
}

/* Overflow-safe midpoint (src/replace.cuh:50-56) */
__device__ __forceinline__ double
midpoint(double lo, double hi)
{
	if (fabs(lo) < 1 && fabs(hi) < 1)
		return (lo + hi) / 2;
	return lo / 2 + hi / 2;
}

__device__ __forceinline__ int
find_frame(const Frame *frames, int nf, int tile)
{
	int lo = 0, hi = nf - 1;
	while (lo < hi) {
		int mid = (lo + hi + 1) >> 1;
		if (frames[mid].tile0 <= tile)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

__device__ __forceinline__ int
find_frame_of_block(const Frame *frames, int nf, int block)
{
	int lo = 0, hi = nf - 1;
	while (lo < hi) {
		int mid = (lo + hi + 1) >> 1;
		if (frames[mid].block0 <= block)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

/* Atomic OR / AND-NOT of flag bits on one byte of the flags plane, through
 * the aligned 32-bit word that holds it; return the byte's old value. */
__device__ __forceinline__ uint32_t
flags_or(uint8_t *flag, uint32_t bits)
{
	uintptr_t a = reinterpret_cast<uintptr_t>(flag);
	int sh = int(a & 3) * 8;
	return (atomicOr(reinterpret_cast<uint32_t *>(a & ~uintptr_t(3)), bits << sh) >> sh) & 0xffu;
}

__device__ __forceinline__ uint32_t
flags_clear(uint8_t *flag, uint32_t bits)
{
	uintptr_t a = reinterpret_cast<uintptr_t>(flag);
	int sh = int(a & 3) * 8;
	return (atomicAnd(reinterpret_cast<uint32_t *>(a & ~uintptr_t(3)), ~(bits << sh)) >> sh) & 0xffu;
}

__device__ __forceinline__ unsigned long long
list_entry(int frame, int index)
{
	return ((unsigned long long)(unsigned)frame << 32) | (unsigned)index;
}

} // namespace cusmic

/*
 * CPU oracle for the L.A.Cosmic cleaning loop -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm, used as the parity checker
 * of the HIP kernels. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library never does. Build with
 * -ffp-contract=off (see oracle/Makefile) so every product is rounded before
 * it is added, as in the reference build (Makefile:92,102-104).
 *
 * Each function names the reference lines it follows. Parity pins (tests):
 *   whole loop, error-map path   test/data fixture: 6,518-pixel mask and all
 *                                pixels bit-exact (tests/test_oracle.py)
 *   medians, all border modes    scipy.ndimage.median_filter vectors
 *   Laplacian, exact order       scipy convolve on the 2x-replicated frame,
 *                                clip, 2x2 block sum vectors; KATs of
 *                                test/test_cleaner.py:33-47
 *   dilation                     scipy.ndimage.binary_dilation vectors
 *   replacement                  brute-force numpy median + KATs of
 *                                test/test_cleaner.py:135-151
 *   noise model                  closed-form flag thresholds: raised pixels on
 *                                flat frames flip at a = t N for scalar and
 *                                map calibration, the 1e-5 floor, and plateau
 *                                scenes that separate the 5x5 median from the
 *                                3x3 and 7x7 ones (tests/test_oracle_noise.py)
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FINE_FLOOR 0.01   /* src/filters.cuh:21 */
#define SIGNAL_FLOOR 1e-5 /* src/filters.cuh:22 */

/* scipy.ndimage border index mapping; src/filters.cuh:24-44. Returns -1 for a
 * CONSTANT-mode sample outside the frame (read as 0). */
long
oracle_border_index(long x, long n, int mode)
{
	long p;
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
	p = mode == CUSMIC_REFLECT ? 2 * n : 2 * (n - 1);
	x %= p;
	if (x < 0)
		x += p;
	if (x < n)
		return x;
	return mode == CUSMIC_REFLECT ? p - 1 - x : p - x;
}

/* Median of the size x size window with scipy border semantics, the window
 * in row-major order (src/filters.cuh:46-70). The reference sorts it with
 * odd-even transposition passes that swap only when b < a (src/sort.cuh:18-36);
 * the insertion sort below gives the same array: neither moves a NaN, and each
 * run of values between NaNs ends sorted in place, so a window with a NaN (S
 * made from infinite samples) has a median that depends on the positions. */
void
oracle_median(const double *in, double *out, long w, long h, int size, int mode)
{
	long r = size / 2;
#pragma omp parallel
	{
		double v[49];
		long xs[7], ys[7];
#pragma omp for schedule(static)
		for (long y = 0; y < h; ++y) {
			for (long k = 0; k < size; ++k)
				ys[k] = oracle_border_index(y + k - r, h, mode);
			for (long x = 0; x < w; ++x) {
				int n = 0;
				for (long k = 0; k < size; ++k)
					xs[k] = oracle_border_index(x + k - r, w, mode);
				for (long j = 0; j < size; ++j)
					for (long i = 0; i < size; ++i)
						v[n++] = xs[i] < 0 || ys[j] < 0 ? 0.0 : in[ys[j] * w + xs[i]];
				/* insertion sort: stable, ties keep their order */
				for (int a = 1; a < n; ++a) {
					double t = v[a];
					int b = a - 1;
					while (b >= 0 && t < v[b]) {
						v[b + 1] = v[b];
						--b;
					}
					v[b + 1] = t;
				}
				out[y * w + x] = v[n / 2];
			}
		}
	}
}

static double
positive(double v)
{
	return v < 0 ? 0.0 : v;
}

static double
lap_sample(const double *in, long x, long y, long w, long h, int mode)
{
	x = oracle_border_index(x, w, mode);
	y = oracle_border_index(y, h, mode);
	return x < 0 || y < 0 ? 0.0 : in[y * w + x];
}

/* Laplacian of the 2x-replicated frame, clipped at zero and rebinned 2x2.
 * Neighbours of a replicated edge pixel under MIRROR/REFLECT are the edge
 * pixel itself, i.e. NEAREST (src/filters.cuh:86-105; mod/cusmic/filters.py:24-31).
 *   close order: the reference CUDA/CuPy expression (src/filters.cuh:103-104).
 *   exact order: the accumulation order of scipy.ndimage.convolve over the
 *     kernel [[0,-1,0],[-1,4,-1],[0,-1,0]] on the replicated frame, whose
 *     row-major footprint visits up, left, centre, right, down, followed by a
 *     2x2 block sum (UL+UR)+(DL+DR), or ((UL+UR)+DL)+DR for one-pixel-wide
 *     frames, as numpy reduces astropy's block view (L.A.Cosmic 1.4.0
 *     remove_cosmics, restated by the reference test test/test_cleaner.py:33-47). */
void
oracle_laplacian(const double *in, double *out, long w, long h, int mode, int laplacian)
{
	int m = mode == CUSMIC_MIRROR || mode == CUSMIC_REFLECT ? CUSMIC_NEAREST : mode;
#pragma omp parallel for schedule(static)
	for (long y = 0; y < h; ++y)
		for (long x = 0; x < w; ++x) {
			double c = in[y * w + x] / 4;
			double u = lap_sample(in, x, y - 1, w, h, m) / 4;
			double d = lap_sample(in, x, y + 1, w, h, m) / 4;
			double l = lap_sample(in, x - 1, y, w, h, m) / 4;
			double r = lap_sample(in, x + 1, y, w, h, m) / 4;
			double v;
			if (laplacian == ORACLE_LAP_CLOSE) {
				double up = c - u, down = c - d, left = c - l, right = c - r;
				v = (positive(up + left) + positive(up + right)) +
				    (positive(down + left) + positive(down + right));
			} else {
				double ul = ((((0.0 - u) - l) + 4 * c) - c) - c;
				double ur = ((((0.0 - u) - c) + 4 * c) - r) - c;
				double dl = ((((0.0 - c) - l) + 4 * c) - c) - d;
				double dr = ((((0.0 - c) - c) + 4 * c) - r) - d;
				/* np.sum over the 2x2 block axes coalesces them into one
				 * sequential run of four when the frame is one pixel wide */
				if (w == 1)
					v = ((positive(ul) + positive(ur)) + positive(dl)) + positive(dr);
				else
					v = (positive(ul) + positive(ur)) + (positive(dl) + positive(dr));
			}
			out[y * w + x] = v;
		}
}

/* 3x3 binary dilation clipped at the frame edge (zero border), as
 * scipy.ndimage.binary_dilation with a full 3x3 structure
 * (src/filters.cuh:177-193; mod/cusmic/filters.py:63-73). */
void
oracle_dilate(const uint8_t *in, uint8_t *out, long w, long h)
{
	for (long y = 0; y < h; ++y)
		for (long x = 0; x < w; ++x) {
			uint8_t v = 0;
			for (long yy = y > 0 ? y - 1 : 0; yy <= (y + 1 < h ? y + 1 : h - 1); ++yy)
				for (long xx = x > 0 ? x - 1 : 0; xx <= (x + 1 < w ? x + 1 : w - 1); ++xx)
					v |= in[yy * w + xx] != 0;
			out[y * w + x] = v;
		}
}

/* Order-preserving unsigned key of a double: -0 sorts below +0
 * (src/replace.cuh:37-48). */
static uint64_t
order_key(double v)
{
	uint64_t b;
	memcpy(&b, &v, sizeof(b));
	return b >> 63 ? ~b : b ^ (UINT64_C(1) << 63);
}

struct donor {
	uint64_t key;
	double value;
	long index;
};

static int
cmp_donor(const void *a, const void *b)
{
	const struct donor *x = a, *y = b;
	if (x->key != y->key)
		return x->key < y->key ? -1 : 1;
	return (x->index > y->index) - (x->index < y->index);
}

/* Overflow-safe midpoint (src/replace.cuh:50-56). */
static double
midpoint(double lo, double hi)
{
	if (fabs(lo) < 1 && fabs(hi) < 1)
		return (lo + hi) / 2;
	return lo / 2 + hi / 2;
}

/* The reference's replacement sort for windows of at most 32 pixels
 * (src/replace.cuh:68-83,132-145): a 32-lane bitonic network in which lane pos
 * holds the pixel at raster position pos of the clipped window (+inf for a
 * non-donor and past the window). In each compare-exchange the lane that keeps
 * the minimum takes its partner's value only when that is strictly smaller,
 * so +0 and -0 stay where they are against each other. */
void
oracle_bitonic32(double v[32])
{
	for (int span = 2; span <= 32; span *= 2)
		for (int step = span / 2; step; step /= 2)
			for (int l = 0; l < 32; ++l) {
				int u = l | step;
				if (l & step)
					continue;
				double a = v[l], b = v[u];
				/* l keeps the minimum when its span bit is clear */
				int swap = !(l & span) ? b < a : a < b;
				if (swap) {
					v[l] = b;
					v[u] = a;
				}
			}
}

/* Replace every target with the median of the donors in the smallest clipped
 * window [x-r, x+r] x [y-r, y+r], r = 2, 3, ..., that holds a donor. Donor
 * values come from the pre-pass snapshot (targets are never donors), odd
 * counts give 0.0 + middle and even counts the midpoint of the two middles.
 * Values come from the reference's value network for windows of at most 32
 * pixels and from key order (-0 below +0) for larger ones; the two differ only
 * in the sign of an even-count median whose two middle donors are zeros. The
 * recorded donors order ties by value key, then by pixel index
 * (src/replace.cuh:112-153; mod/cusmic/replace.py:28-70). */
int
oracle_replace(double *clean, const uint8_t *target, const uint8_t *donor, long w, long h,
	int32_t *donors)
{
	long n = w * h, rmax = w > h ? w : h;
	double *next = malloc(sizeof(double) * (size_t)n);
	struct donor *list = NULL;
	size_t cap = 0;
	int status = 0;
	if (!next)
		return -2;
	memcpy(next, clean, sizeof(double) * (size_t)n);
	for (long i = 0; i < n && !status; ++i) {
		if (!target[i])
			continue;
		long x = i % w, y = i / w;
		int found = 0;
		for (long r = 2; r <= (rmax > 2 ? rmax : 2) && !found; ++r) {
			long x0 = x - r < 0 ? 0 : x - r, x1 = x + r >= w ? w - 1 : x + r;
			long y0 = y - r < 0 ? 0 : y - r, y1 = y + r >= h ? h - 1 : y + r;
			size_t need = (size_t)((x1 - x0 + 1) * (y1 - y0 + 1)), cnt = 0;
			if (need > cap) {
				struct donor *grown = realloc(list, need * sizeof(*list));
				if (!grown) {
					status = -2;
					break;
				}
				list = grown;
				cap = need;
			}
			for (long yy = y0; yy <= y1; ++yy)
				for (long xx = x0; xx <= x1; ++xx) {
					long j = yy * w + xx;
					if (donor[j] && isfinite(clean[j])) {
						list[cnt].key = order_key(clean[j]);
						list[cnt].value = clean[j];
						list[cnt].index = j;
						++cnt;
					}
				}
			if (!cnt)
				continue;
			qsort(list, cnt, sizeof(*list), cmp_donor);
			struct donor lo = list[(cnt - 1) / 2], hi = list[cnt / 2];
			double vlo = lo.value, vhi = hi.value;
			if (need <= 32) {
				double v[32];
				long ww = x1 - x0 + 1;
				for (int k = 0; k < 32; ++k) {
					long j = (y0 + k / ww) * w + x0 + k % ww;
					v[k] = (size_t)k < need && donor[j] && isfinite(clean[j]) ? clean[j] : INFINITY;
				}
				oracle_bitonic32(v);
				vlo = v[(cnt - 1) / 2];
				vhi = v[cnt / 2];
			}
			next[i] = cnt % 2 ? 0.0 + vlo : midpoint(vlo, vhi);
			if (donors) {
				donors[2 * i] = (int32_t)lo.index;
				donors[2 * i + 1] = (int32_t)hi.index;
			}
			found = 1;
		}
		if (!found && !status)
			status = -1;
	}
	memcpy(clean, next, sizeof(double) * (size_t)n);
	free(next);
	free(list);
	return status;
}

static void
set_msg(char *msg, size_t len, const char *text)
{
	if (msg && len)
		snprintf(msg, len, "%s", text);
}

/* Per-frame working state of the loop (src/cleaner.cuh:39-61). */
struct frame_state {
	long w, h, n;
	const struct cusmic_image *im;
	double g, rn;
	double *clean;
	uint8_t *hole, *excl, *crmask, *cand, *grown, *donor;
	double *lap, *noise, *sig, *tmp, *med3, *med7, *fine;
	int converged;
};

/* Argument checks of the reference ABI, extended to per-frame shapes
 * (src/api.cu:47-81). */
static const char *
check_arguments(const struct cusmic_image *ims, size_t nf, const struct cusmic_options *o,
	const double *fg, const double *frn)
{
	if (!ims || !o || !nf || nf > 65535)
		return "invalid images, batch size or output buffers";
	if (o->maxiter < 0 || o->border < 0 || o->border > CUSMIC_WRAP || !isfinite(o->contrast) ||
		o->contrast < 0 || !isfinite(o->cr_threshold) || o->cr_threshold < 0 ||
		!isfinite(o->neighbor_threshold) || o->neighbor_threshold < 0)
		return "invalid detection settings";
	for (size_t f = 0; f < nf; ++f) {
		const struct cusmic_image *im = &ims[f];
		double g = fg ? fg[f] : o->gain, rn = frn ? frn[f] : o->readnoise;
		if (!im->width || !im->height || im->width > (size_t)INT_MAX / 4 / im->height)
			return "invalid image dimensions";
		if (!im->data)
			return "frames must have data and the same shape";
		if (!im->error) {
			int needs_gain = o->maxiter || !isnan(g);
			if (!im->gain && needs_gain && (!isfinite(g) || g <= 0))
				return "positive gain or error is required";
			if (!im->readnoise && (!isfinite(rn) || rn < 0))
				return "read noise must be finite and nonnegative";
		}
		if (!im->background && !isnan(o->background) && !isfinite(o->background))
			return "background must be finite";
	}
	return NULL;
}

static double
background_at(const struct frame_state *s, const struct cusmic_options *o, long i)
{
	return s->im->background ? s->im->background[i] : o->background;
}

static int
has_background(const struct frame_state *s, const struct cusmic_options *o)
{
	return s->im->background || !isnan(o->background);
}

/* Detection of one pass: Laplacian, noise, significance, fine structure,
 * detect and two growth steps (src/cleaner.cuh:102-126, A.2-A.6). */
static void
find_cosmics(struct frame_state *s, const struct cusmic_options *o, int laplacian)
{
	long w = s->w, h = s->h, n = s->n;
	const struct cusmic_image *im = s->im;
	const double *noise;
	oracle_laplacian(s->clean, s->lap, w, h, o->border, laplacian);
	if (im->error) {
		noise = im->error;
	} else {
		/* src/filters.cuh:107-123 */
		oracle_median(s->clean, s->tmp, w, h, 5, o->border);
		for (long i = 0; i < n; ++i) {
			double g = im->gain ? im->gain[i] : s->g;
			double rn = im->readnoise ? im->readnoise[i] : s->rn;
			double m = s->tmp[i] < SIGNAL_FLOOR ? SIGNAL_FLOOR : s->tmp[i];
			s->noise[i] = sqrt(rn * rn + g * m) / g;
		}
		noise = s->noise;
	}
	/* src/filters.cuh:125-146 */
	for (long i = 0; i < n; ++i)
		s->sig[i] = s->lap[i] / (2 * noise[i]);
	oracle_median(s->sig, s->tmp, w, h, 5, o->border);
	for (long i = 0; i < n; ++i)
		s->sig[i] = s->hole[i] ? 0.0 : s->sig[i] - s->tmp[i];
	/* src/filters.cuh:148-162 */
	oracle_median(s->clean, s->med3, w, h, 3, o->border);
	oracle_median(s->med3, s->med7, w, h, 7, o->border);
	for (long i = 0; i < n; ++i) {
		double v = (s->med3[i] - s->med7[i]) / noise[i];
		s->fine[i] = v < FINE_FLOOR ? FINE_FLOOR : v;
	}
	/* src/filters.cuh:164-193 */
	for (long i = 0; i < n; ++i)
		s->cand[i] = !s->excl[i] && s->sig[i] > o->cr_threshold &&
			     s->sig[i] / s->fine[i] > o->contrast;
	oracle_dilate(s->cand, s->grown, w, h);
	for (long i = 0; i < n; ++i)
		s->grown[i] = s->grown[i] && s->sig[i] > o->cr_threshold;
	oracle_dilate(s->grown, s->cand, w, h);
	for (long i = 0; i < n; ++i)
		s->cand[i] = s->cand[i] && s->sig[i] > o->neighbor_threshold;
}

static void
free_state(struct frame_state *s)
{
	free(s->hole);
	free(s->lap);
}

static int
alloc_state(struct frame_state *s)
{
	size_t n = (size_t)s->n;
	s->hole = calloc(6 * n, 1);
	s->lap = malloc(7 * n * sizeof(double));
	if (!s->hole || !s->lap)
		return -1;
	s->excl = s->hole + n;
	s->crmask = s->excl + n;
	s->cand = s->crmask + n;
	s->grown = s->cand + n;
	s->donor = s->grown + n;
	s->noise = s->lap + n;
	s->sig = s->noise + n;
	s->tmp = s->sig + n;
	s->med3 = s->tmp + n;
	s->med7 = s->med3 + n;
	s->fine = s->med7 + n;
	return 0;
}

/* The whole call (src/cleaner.cuh:144-187 with the prologue/epilogue of
 * src/image.cuh:24-92 and the hole fill of src/cleaner.cuh:78-100). Frames are
 * independent; a frame that converges is idempotent under further passes, so
 * stopping it early equals the reference's batch-wide loop. */
int
oracle_clean(const struct cusmic_image *ims, size_t nf, const struct cusmic_options *o,
	const double *frame_gain, const double *frame_readnoise, int laplacian, int threads,
	double *clean, uint8_t *mask, int32_t *donors, int32_t *passes, char *msg, size_t len)
{
	const char *err;
	struct frame_state *st;
	size_t off = 0;
	int status = CUSMIC_OK;
	if (msg && len)
		msg[0] = 0;
	if (!clean || !mask) {
		set_msg(msg, len, "invalid images, batch size or output buffers");
		return CUSMIC_INVALID;
	}
	err = check_arguments(ims, nf, o, frame_gain, frame_readnoise);
	if (err) {
		set_msg(msg, len, err);
		return CUSMIC_INVALID;
	}
#ifdef _OPENMP
	omp_set_num_threads(threads > 0 ? threads : 1);
#else
	(void)threads;
#endif
	st = calloc(nf, sizeof(*st));
	if (!st) {
		set_msg(msg, len, "std::bad_alloc");
		return CUSMIC_NO_MEMORY;
	}
	for (size_t f = 0; f < nf; ++f) {
		struct frame_state *s = &st[f];
		s->w = (long)ims[f].width;
		s->h = (long)ims[f].height;
		s->n = s->w * s->h;
		s->im = &ims[f];
		s->g = frame_gain ? frame_gain[f] : o->gain;
		s->rn = frame_readnoise ? frame_readnoise[f] : o->readnoise;
		s->clean = clean + off;
		if (donors)
			for (long i = 0; i < 2 * s->n; ++i)
				donors[2 * off + i] = -1;
		if (passes)
			passes[f] = 0;
		off += (size_t)s->n;
	}

	if (!o->maxiter) {
		/* src/image.cuh:72-92; src/api.cu:120-130 */
		for (size_t f = 0; f < nf; ++f) {
			struct frame_state *s = &st[f];
			for (long i = 0; i < s->n; ++i) {
				double d = s->im->data[i];
				double v = isfinite(d) ? d : 0.0;
				if (has_background(s, o)) {
					double bg = background_at(s, o, i);
					v += 1 * bg;
					v += -1 * bg;
				}
				s->clean[i] = isfinite(d) ? v : d;
			}
		}
		memset(mask, 0, off);
		free(st);
		return CUSMIC_OK;
	}

	for (size_t f = 0; f < nf && !status; ++f)
		if (alloc_state(&st[f])) {
			status = CUSMIC_NO_MEMORY;
			set_msg(msg, len, "std::bad_alloc");
		}

	/* Prologue and hole fill (src/image.cuh:24-48; src/cleaner.cuh:78-100). */
	for (size_t f = 0; f < nf && !status; ++f) {
		struct frame_state *s = &st[f];
		long holes = 0, allowed = 0, finite_donors = 0;
		for (long i = 0; i < s->n; ++i) {
			double d = s->im->data[i];
			int hole = !isfinite(d);
			int allow = !hole && !(s->im->mask && s->im->mask[i]);
			s->hole[i] = (uint8_t)hole;
			s->excl[i] = (uint8_t)!allow;
			s->clean[i] = hole ? 0.0 : d;
			if (has_background(s, o))
				s->clean[i] += 1 * background_at(s, o, i);
			holes += hole;
			allowed += allow;
		}
		if (!holes || !allowed)
			continue;
		for (long i = 0; i < s->n; ++i) {
			s->donor[i] = !s->excl[i] && isfinite(s->clean[i]);
			finite_donors += s->donor[i];
		}
		if (!finite_donors) {
			status = CUSMIC_INVALID;
			set_msg(msg, len, "no finite replacement donors");
			break;
		}
		oracle_replace(s->clean, s->hole, s->donor, s->w, s->h, NULL);
	}

	/* The detection loop (src/cleaner.cuh:175-181; src/cleaner.cuh:22-37,128-142). */
	for (int it = 0; it < o->maxiter && !status; ++it) {
		int any = 0;
		for (size_t f = 0; f < nf; ++f) {
			struct frame_state *s = &st[f];
			long todo = 0, ndonor = 0;
			if (s->converged)
				continue;
			find_cosmics(s, o, laplacian);
			if (passes)
				passes[f] = it + 1;
			for (long i = 0; i < s->n; ++i) {
				todo += s->cand[i] && !s->crmask[i];
				s->crmask[i] |= s->cand[i];
				s->donor[i] = !s->crmask[i] && !s->excl[i] && isfinite(s->clean[i]);
				ndonor += s->donor[i];
			}
			if (todo && !ndonor) {
				status = CUSMIC_INVALID;
				set_msg(msg, len, "no finite replacement donors");
				break;
			}
			if (!todo) {
				s->converged = 1;
				continue;
			}
			any = 1;
		}
		if (status || !any)
			break;
		for (size_t f = 0; f < nf; ++f) {
			struct frame_state *s = &st[f];
			if (s->converged)
				continue;
			size_t base = (size_t)(s->clean - clean);
			oracle_replace(s->clean, s->crmask, s->donor, s->w, s->h,
				donors ? donors + 2 * base : NULL);
		}
	}

	/* Epilogue (src/image.cuh:50-70). */
	for (size_t f = 0; f < nf && !status; ++f) {
		struct frame_state *s = &st[f];
		size_t base = (size_t)(s->clean - clean);
		for (long i = 0; i < s->n; ++i) {
			double v = s->clean[i];
			double d = s->im->data[i];
			if (has_background(s, o))
				v += -1 * background_at(s, o, i);
			s->clean[i] = isfinite(d) ? v : d;
			mask[base + (size_t)i] = s->crmask[i];
			if (donors && !s->crmask[i])
				donors[2 * (base + (size_t)i)] = donors[2 * (base + (size_t)i) + 1] = -1;
		}
	}
	for (size_t f = 0; f < nf; ++f)
		free_state(&st[f]);
	free(st);
	return status;
}

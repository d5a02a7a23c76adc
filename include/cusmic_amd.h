/*
 * Extensions of the cusmic C ABI that have no counterpart in the reference
 * (src/cusmic.h). The six reference entry points are declared in cusmic.h and
 * behave as there; everything here is additive.
 */
#ifndef CUSMIC_AMD_H
#define CUSMIC_AMD_H

#include "cusmic.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Laplacian accumulation order. EXACT follows L.A.Cosmic 1.4.0 (scipy's
 * convolution order on the 2x-replicated frame); CLOSE is the expression of the
 * reference CUDA kernel (src/filters.cuh:103-104). Both gave the saved
 * reference's mask; they can differ by rounding near a threshold. */
enum cusmic_laplacian { CUSMIC_LAPLACIAN_EXACT = 0, CUSMIC_LAPLACIAN_CLOSE = 1 };

/* Per-call extras of remove_cosmics_ragged_device; zero-initialise and set what
 * is needed. All pointers may be NULL. */
struct cusmic_ragged_options {
	const double *frame_gain;      /* host, nf values: per-frame scalar gain */
	const double *frame_readnoise; /* host, nf values: per-frame scalar read noise */
	int32_t *donors;    /* device, 2 per pixel: final median donors (in-frame
	                       indices) of flagged pixels, -1 elsewhere */
	int32_t *passes;    /* host, nf values: detection passes each frame ran */
	int laplacian;      /* enum cusmic_laplacian */
	double *profile;    /* host, CUSMIC_PROFILE_SIZE values, or NULL: when set,
	                       the call times its kernels with hipEvents on the
	                       caller's stream (see enum cusmic_profile) */
	int64_t *new_pixels; /* host, new_pixels_size values, or NULL: pixels
	                       flagged for the first time in each detection pass,
	                       over all frames (0 for passes that did not run) */
	size_t new_pixels_size;
};

/* Slots of cusmic_ragged_options.profile. Times are summed over launches in
 * milliseconds. PROLOGUE_MS covers the set-up launches; FIRST_PASS_MS the
 * dense first detection pass, which also makes the working image and the
 * flags from the input; SIGNIFICANCE_MS the block passes (DETECT_LAUNCHES of
 * them: one per pass, the first after the hole fill); DETECT_MS whole
 * detection passes, both of these included. DETECT_PIXELS counts the in-frame
 * pixels that the first pass and the later block passes processed. SEEDS ..
 * WIDENED are list lengths summed over passes. DETECT_PASS_PIXELS and
 * REPLACE_PASS_PIXELS sum frame pixels times the detection and replacement
 * passes the frame ran. */
enum cusmic_profile {
	CUSMIC_PROFILE_DETECT_MS, CUSMIC_PROFILE_REPLACE_MS, CUSMIC_PROFILE_PROLOGUE_MS,
	CUSMIC_PROFILE_EPILOGUE_MS, CUSMIC_PROFILE_DETECT_LAUNCHES, CUSMIC_PROFILE_DETECT_PIXELS,
	CUSMIC_PROFILE_FLAGGED, CUSMIC_PROFILE_SEEDS, CUSMIC_PROFILE_CANDIDATES,
	CUSMIC_PROFILE_GROWN, CUSMIC_PROFILE_WIDENED, CUSMIC_PROFILE_SIGNIFICANCE_MS,
	CUSMIC_PROFILE_DETECT_PASS_PIXELS, CUSMIC_PROFILE_REPLACE_PASS_PIXELS,
	CUSMIC_PROFILE_FIRST_PASS_MS, CUSMIC_PROFILE_SIZE
};

/* Ragged batch: 1..65535 contiguous device frames of individual shapes,
 * cleaned in one call on the caller's stream (persistent tile-queue kernels;
 * the host enqueues the passes in chunks of four and, between chunks, reads
 * one word to stop once no frame flags a new pixel, so maxiter <= 4 makes no
 * round trip inside the loop). Outputs are the frames concatenated in
 * descriptor order. Per-frame scalars in ext replace options->gain/readnoise.
 * Same validation, status codes and completion semantics as
 * remove_cosmics_batch_device (src/cusmic.h:80-85), which is the special
 * case of equal shapes. No counterpart in the reference. */
int remove_cosmics_ragged_device(const struct cusmic_image *ims, size_t nf,
	const struct cusmic_options *options, const struct cusmic_ragged_options *ext,
	double *clean, uint8_t *mask, void *stream, char *msg, size_t message_size);

/* Size of the scratch buffer that remove_cosmics_batch_device_noalloc uses, in
 * MiB. A call needs about 44 bytes per pixel plus a small table per frame:
 * at most about 3 million pixels, e.g. one 1024 x 1024 frame (46 MB) or eleven
 * 512 x 512 frames. */
#define CUSMIC_NOALLOC_SCRATCH_MIB 128

/* remove_cosmics_batch_device (src/cusmic.h:80-86) without any device memory
 * allocation: its scratch is one static buffer of CUSMIC_NOALLOC_SCRATCH_MIB
 * per device, which every device the library runs on holds until the process
 * ends. No counterpart in the reference; for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). Same arguments, validation, results, status codes and
 * completion semantics as remove_cosmics_batch_device; a batch that needs more
 * scratch than the buffer holds is rejected with CUSMIC_INVALID before any
 * work. Every call on one device shares that device's buffer: calls that
 * overlap on one device (several host threads or streams) race. */
int remove_cosmics_batch_device_noalloc(const struct cusmic_image *ims, size_t nf,
	const struct cusmic_options *options, double *clean, uint8_t *mask, void *stream, char *msg,
	size_t message_size);

/* Adjoint of the cleaning map with the detection mask held fixed:
 * grad_data = grad_clean on unflagged pixels, plus, for every flagged pixel,
 * grad_clean scattered to its recorded donors (weight 1 for an odd donor count,
 * 1/2 each for the two middle donors of an even count), summed in raster order
 * of the flagged pixels: a deterministic gather, no atomics on the gradient.
 * The cost is that of a 5 x 5 gather plus, for the donors of windows wider than
 * 5 x 5, a rescan of the frame's widest window per donor. The gradient with
 * respect to a background map is grad_data - grad_clean (a scalar
 * background's is exactly zero); the caller forms it. Frames, mask and donors
 * are laid out as returned by remove_cosmics_ragged_device; only width/height
 * of ims are read. No counterpart in the reference. */
int cusmic_adjoint_device(const struct cusmic_image *ims, size_t nf, const uint8_t *mask,
	const int32_t *donors, const double *grad_clean, double *grad_data, void *stream, char *msg,
	size_t message_size);

/* Diagnostic: quotient[i] = a[i] / b[i] and root[i] = sqrt(a[i]) computed on
 * the GPU with the kernels' own compile flags, for bitwise checks of IEEE
 * rounding against the host. Device arrays. No counterpart in the reference. */
int cusmic_selftest_arith(const double *a, const double *b, double *quotient, double *root,
	size_t n, void *stream);

#ifdef __cplusplus
}
#endif
#endif

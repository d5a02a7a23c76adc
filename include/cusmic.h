/*
 * cusmic C ABI, MI355X (gfx950) implementation.
 *
 * Drop-in for the reference header src/cusmic.h: the structs, enum values and
 * the six exported functions keep their names, layouts, argument meanings and
 * status codes, so existing callers relink without source changes.
 * Extensions that have no counterpart in the reference live in cusmic_amd.h.
 */
#ifndef CUSMIC_H
#define CUSMIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Border handling of the median and Laplacian stencils (scipy.ndimage modes).
 * Values follow src/cusmic.h:26. */
enum cusmic_border { CUSMIC_REFLECT, CUSMIC_CONSTANT, CUSMIC_NEAREST, CUSMIC_MIRROR, CUSMIC_WRAP };

/* Status codes, src/cusmic.h:27. CUSMIC_CUDA_ERROR keeps its name for ABI
 * compatibility and reports HIP runtime failures here. */
enum cusmic_status { CUSMIC_OK, CUSMIC_INVALID, CUSMIC_CUDA_ERROR, CUSMIC_NO_MEMORY };

/* One frame, src/cusmic.h:37-43. Row-major float64 pixels, width is the fast
 * axis. Optional per-pixel maps (NULL when absent) have the same shape:
 *   error      positive 1-sigma noise in image units; overrides gain/readnoise
 *   gain       positive electrons per image unit (L.A.Cosmic effective_gain)
 *   readnoise  nonnegative electrons
 *   background added before cleaning and subtracted afterwards
 *   mask       nonzero bytes exclude seeds and replacement donors; detections
 *              may still grow into masked pixels
 * Nonfinite data pixels are preserved bit for bit and never flagged. */
struct cusmic_image {
	const double *data, *error;
	const double *gain;
	const double *readnoise, *background;
	const uint8_t *mask;
	size_t width, height;
};

/* Detection settings, src/cusmic.h:45-51. maxiter = 0 disables detection but
 * still applies the background arithmetic. gain/background NaN mean unset. */
struct cusmic_options {
	double contrast, cr_threshold, neighbor_threshold;
	double gain;
	double readnoise, background;
	int maxiter;
	int border;
};

/* Library version string (src/cusmic.h:53-54). */
const char *cusmic_version(void);

/* Runtime version (src/cusmic.h:56-57). The reference returns CUDART_VERSION;
 * this build returns HIP_VERSION of the headers it was compiled with. */
int cusmic_cuda_version(void);

/* Defaults {3, 5, 3, NaN, 0, NaN, 4, MIRROR}; NULL-safe (src/cusmic.h:59-63,
 * src/api.cu:40-45). */
void cusmic_default_options(struct cusmic_options *options);

/* Host arrays, borrowed; outputs must not overlap inputs. Validates map
 * contents. Host outputs are untouched on failure. maxiter = 0 runs on the
 * CPU and needs no GPU. Replaces src/cusmic.h:65-70 (src/api.cu:155-175). */
int remove_cosmics(const struct cusmic_image *image, const struct cusmic_options *options,
	double *clean, uint8_t *mask, char *msg, size_t message_size);

/* Device arrays on the caller's current GPU; stream is a hipStream_t cast to
 * void * (NULL = default stream). Work is ordered on that stream and the call
 * returns after completion. No content validation. Replaces
 * src/cusmic.h:72-78 (src/api.cu:177-182). */
int remove_cosmics_device(const struct cusmic_image *image, const struct cusmic_options *options,
	double *clean, uint8_t *mask, void *stream, char *msg, size_t message_size);

/* 1..65535 same-shaped contiguous device frames; calibration pointers may be
 * shared between descriptors; outputs are contiguous in descriptor order,
 * clean[f * width * height + i]. Replaces src/cusmic.h:80-85
 * (src/api.cu:184-192). */
int remove_cosmics_batch_device(const struct cusmic_image *ims, size_t nf,
	const struct cusmic_options *options, double *clean, uint8_t *mask, void *stream, char *msg,
	size_t message_size);

#ifdef __cplusplus
}
#endif
#endif

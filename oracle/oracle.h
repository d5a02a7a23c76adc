/*
 * CPU oracle for the L.A.Cosmic cleaning loop -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference algorithm used as the parity
 * checker for the HIP implementation. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it. The product library never links or
 * calls it.
 */
#ifndef CUSMIC_ORACLE_H
#define CUSMIC_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/cusmic.h"

#ifdef __cplusplus
extern "C" {
#endif

enum oracle_laplacian { ORACLE_LAP_EXACT = 0, ORACLE_LAP_CLOSE = 1 };

/* Clean nf frames of individual shapes (a ragged batch). Outputs are the
 * frames concatenated in descriptor order. frame_gain/frame_readnoise, when
 * non-NULL, give per-frame scalars that replace the option scalars.
 * donors (2 int32 per pixel, may be NULL) receives the in-frame indices of the
 * final replacement's lower/upper median donors for flagged pixels, -1
 * elsewhere. passes (nf ints, may be NULL) receives the number of detection
 * passes each frame ran. threads <= 0 means one thread. Returns a
 * cusmic_status value and writes msg on failure. */
int oracle_clean(const struct cusmic_image *ims, size_t nf, const struct cusmic_options *o,
	const double *frame_gain, const double *frame_readnoise, int laplacian, int threads,
	double *clean, uint8_t *mask, int32_t *donors, int32_t *passes, char *msg, size_t len);

/* Stage functions exposed for golden-vector checks. */
void oracle_median(const double *in, double *out, long w, long h, int size, int mode);
void oracle_laplacian(const double *in, double *out, long w, long h, int mode, int laplacian);
void oracle_dilate(const uint8_t *in, uint8_t *out, long w, long h);
long oracle_border_index(long x, long n, int mode);
/* Replace masked targets from donors in expanding windows; returns 0 on
 * success or -1 when a target finds no donor anywhere. */
int oracle_replace(double *clean, const uint8_t *target, const uint8_t *donor, long w, long h,
	int32_t *donors);
/* The reference's 32-lane value network for small replacement windows. */
void oracle_bitonic32(double v[32]);

#ifdef __cplusplus
}
#endif
#endif

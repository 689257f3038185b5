/* mplbws.h -- C ABI of the learned warm start kernels (libmplbws.so, gfx950).
 *
 * No reference counterpart: the reference starts every run from the App's
 * initial macros (SetInitialMacrosVars, lbm2d_cavity.cpp:44-66).  These
 * entry points let a caller train and run a small U-Net that predicts the
 * steady cavity flow, whose prediction then initialises the lattice
 * (mplbhip_init_lift).
 *
 * All pointers are device pointers to FP32 arrays, channels last
 * (x[n][y][x][c]).  Convolutions are 3x3, zero padded, stride 1, computed
 * with BF16 operands (rounded to nearest even on load) and FP32
 * accumulation on the matrix cores.  Weights: w[tap = 3*ky + kx][cin][cout].
 * Work is queued on the default stream; mplbws_sync waits for it, and the
 * copies of mplbws_upload/mplbws_download are synchronous.
 * Return 0 on success, a negative MPLBWS_ERR_* code otherwise.
 */
#ifndef MPLBWS_H
#define MPLBWS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPLBWS_ERR_ARG (-1)
#define MPLBWS_ERR_HIP (-2)
/* pixels per weight-gradient partial (whole rows of one image: at least
 * one row, at most one image) */
#define MPLBWS_WGRAD_CHUNK 4096

/* out = [relu](conv3x3(in, w) + bias); cin a multiple of 8; bias may be 0 */
int mplbws_conv3x3_fwd(const float* in, const float* w, const float* bias, float* out, int n,
                       int h, int wd, int cin, int cout, int relu);
/* wt[tap][cout][cin] = w[8 - tap][cin][cout]: with it, conv3x3_fwd of the
 * output gradient (cout channels, a multiple of 8) gives the input gradient */
int mplbws_flip_weights(const float* w, float* wt, int cin, int cout);
/* floats of workspace mplbws_conv3x3_wgrad needs for this shape */
int64_t mplbws_wgrad_workspace(int n, int h, int wd, int cin, int cout);
/* dw[tap][cin][cout] = sum over pixels of in(shifted) * dout, dbias[cout] =
 * sum of dout (dbias may be 0); sums are taken in a fixed order */
int mplbws_conv3x3_wgrad(const float* in, const float* dout, float* dw, float* dbias,
                         float* workspace, int n, int h, int wd, int cin, int cout);
/* din = dout where out > 0, else 0 */
int mplbws_relu_bwd(const float* dout, const float* out, float* din, int64_t count);
/* 2x2 average pooling of an h x wd image (both even) and its adjoint */
int mplbws_pool2_fwd(const float* in, float* out, int n, int h, int wd, int c);
int mplbws_pool2_bwd(const float* dout, float* din, int n, int h, int wd, int c);
/* nearest 2x upsampling of an h x wd image into channels [c0, c0 + c) of a
 * ctot-channel tensor, and its adjoint */
int mplbws_up2_fwd(const float* in, float* out, int n, int h, int wd, int c, int ctot, int c0);
int mplbws_up2_bwd(const float* dout, float* din, int n, int h, int wd, int c, int ctot, int c0);
/* direction 0: dst[:, c0:c0+c] = src (c channels); 1: dst = src[:, c0:c0+c];
 * 2: dst += src[:, c0:c0+c] */
int mplbws_channels(const float* src, float* dst, int64_t pixels, int c, int ctot, int c0,
                    int direction);
/* dst[p][ch] = src[p][ch] where mask[p] == 0 and ch < cu, else 0 (c
 * channels; src and dst may be the same array) */
int mplbws_mask_channels(const float* src, const float* mask, float* dst, int64_t pixels, int c,
                         int cu);
/* dst[i][0 .. elems) = src[idx[i]][0 .. elems) for i < count (idx on the
 * device; the caller keeps every idx[i] inside src) */
int mplbws_gather(const float* src, float* dst, const int32_t* idx, int count, int64_t elems);
/* masked squared error over channels [0, cu) of pred (ctot channels) and
 * target (cu channels), pixels with mask == 0: grad = 2 (pred - target) *
 * inv_count there, 0 elsewhere; partial[npartial] receives block sums of the
 * squared errors (their sum times inv_count is the loss) */
int mplbws_mse(const float* pred, const float* target, const float* mask, float* grad,
               float* partial, int npartial, int64_t pixels, int cu, int ctot, float inv_count);
/* one Adam step (Kingma & Ba 2015) on count parameters; step counts from 1 */
int mplbws_adam(float* p, const float* g, float* m, float* v, int64_t count, float lr, float b1,
                float b2, float eps, int step);
int mplbws_sync(void);
/* device memory for callers without their own allocator (the Python layer):
 * hipMalloc/hipFree, synchronous copies, and a zero fill (a kernel) */
int mplbws_alloc(void** p, int64_t bytes);
int mplbws_free(void* p);
int mplbws_upload(void* dst, const void* src, int64_t bytes);
int mplbws_download(void* dst, const void* src, int64_t bytes);
int mplbws_zero(void* p, int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif

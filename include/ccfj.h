/*
 * ccfj.h -- C ABI of the MI355X F-J transform library (libccfj.so).
 *
 * Two layers:
 *   1. Reference-compatible entry points with the argument order and meaning
 *      of the reference's native F-J functions.  Host pointers, synchronous,
 *      out[nc][nf] float32, U_f[nr][nf] float32, r ascending in metres.
 *   2. A context-based API: device selection, cached device workspace, one
 *      stream per context, the fused multi-spectrum modes that fj_noise and
 *      fj_earthquake need, device-pointer (stream-ordered) calls, on-device
 *      per-column normalisation and kernel timing.
 *
 * Every function returns 0 on success and a negative CCFJ_E* code on
 * failure; ccfj_last_error() then describes the failure (per thread).  No
 * function exits the process (the reference's GPU helper calls exit(),
 * include/cuda_helper.h:74-84).
 */
#ifndef CCFJ_H
#define CCFJ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCFJ_OK 0
#define CCFJ_EINVAL (-1)   /* bad argument */
#define CCFJ_EDEVICE (-2)  /* no usable GPU / HIP runtime error */
#define CCFJ_ENOMEM (-3)   /* device allocation failed */

/* Transform modes (how the spectra and bases are combined per (f,c)). */
#define CCFJ_MODE_J 0          /* FJ[U0]                       J0 basis   */
#define CCFJ_MODE_Y 1          /* FY[U0]                       Y0 basis   */
#define CCFJ_MODE_J_MINUS_Y 2  /* FJ[U0] - FY[U1]              fj_noise func=1 (U1 = Hilbert(U0)) */
#define CCFJ_MODE_ABS_J 3      /* |FJ[U0] + i FJ[U1]|          fj_earthquake func=0 (U0 = Re, U1 = Im) */
#define CCFJ_MODE_ABS_H 4      /* |FJ[U0]-FY[U1] + i(FJ[U1]+FY[U0])|  fj_earthquake func=1 */

/* ---- 1. reference-compatible entry points --------------------------------
 * type: 0 = trapezoid, 1 = linear segments.  A type outside {0,1} returns
 * CCFJ_EINVAL and leaves out untouched (the reference silently does nothing,
 * src/FJcpu.cpp:403-407).  The current HIP device is used. */

/* Replaces FJ, include/FJgpu.hh:13 (GPU build; J0 basis). */
int FJ(float *u_f, float *r, float *f, float *out, float *c, int nc, int nr,
       int nf, int type);
/* Replaces FHr, include/FJgpu.hh:14 (J0 basis, the GPU build's FP64 variant;
 * here identical to FJ). */
int FHr(float *u_f, float *r, float *f, float *out, float *c, int nc, int nr,
        int nf, int type);
/* Replaces FHi, include/FJgpu.hh:15 (Y0 basis). */
int FHi(float *u_f, float *r, float *f, float *out, float *c, int nc, int nr,
        int nf, int type);
/* Replaces FH, include/FJcpu.h:20 (CPU build's Y0 basis; nThread ignored). */
int FH(float *u_f, float *r, float *f, float *out, float *c, int nc, int nr,
       int nf, int type, int nThread);

/* ---- 2. context API --------------------------------------------------------
 * No counterpart in the reference (its GPU code allocates, copies and frees
 * on every call, src/FJgpu.cu:398-417).  A context owns a HIP stream and a
 * grow-only device workspace on one device; calls on one context are
 * serialised by the caller (one context per host thread). */
typedef struct ccfj_ctx ccfj_ctx;

/* device < 0 selects the current HIP device. */
int ccfj_create(int device, ccfj_ctx **ctx);
int ccfj_destroy(ccfj_ctx *ctx);
const char *ccfj_last_error(void);
/* Number of visible GPUs (0 when none); never fails. */
int ccfj_device_count(void);

/* Host pointers.  U0, U1: [nr][nf] float32 (U1 only for the two-spectrum
 * modes, else NULL); r[nr] ascending (metres); f[nf]; c[nc]; out[nc][nf].
 * normalize != 0 divides every column by its max |value| (src/ccfj.pyx:225).
 * Synchronous. */
int ccfj_transform(ccfj_ctx *ctx, int mode, int itype, int normalize,
                   const float *U0, const float *U1, const float *r,
                   const float *f, const float *c, int nr, int nf, int nc,
                   float *out);

/* Device pointers, enqueued on `stream` (a hipStream_t; NULL = the
 * context's stream).  Returns once enqueued. */
int ccfj_transform_dev(ccfj_ctx *ctx, int mode, int itype, int normalize,
                       const float *U0, const float *U1, const float *r,
                       const float *f, const float *c, int nr, int nf, int nc,
                       float *out, void *stream);

/* The same transform with no device allocation: no counterpart in the
 * reference (its GPU code allocates on every call, src/FJgpu.cu:398-417).
 * For callers that cannot allocate -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Its work arrays (station
 * arrays, records, [nf][nc] results, column scales) live in one 64 MiB
 * device array per GPU that the library holds for the life of the process;
 * a size needing more returns CCFJ_EINVAL with the bytes it would need (the
 * workload of 300 pairs x 512 f x 500 c needs about 3.6 MB).  Calls on one
 * GPU share that array, so two calls whose work overlaps in time on one GPU
 * (other streams, other contexts, other threads) race: serialise them.  Every
 * work array is written before it is read; the timing of ccfj_set_timing is
 * not recorded.  Same arguments and results as ccfj_transform_dev. */
int ccfj_transform_dev_noalloc(ccfj_ctx *ctx, int mode, int itype, int normalize,
                               const float *U0, const float *U1, const float *r,
                               const float *f, const float *c, int nr, int nf, int nc,
                               float *out, void *stream);

int ccfj_synchronize(ccfj_ctx *ctx);

/* Device memory on the context's device (for callers without their own
 * allocator).  ccfj_memcpy is synchronous on the context's stream; kind:
 * 0 host->device, 1 device->host, 2 device->device. */
int ccfj_malloc(ccfj_ctx *ctx, size_t bytes, void **ptr);
int ccfj_free(ccfj_ctx *ctx, void *ptr);
int ccfj_memcpy(ccfj_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);

/* Kernel timing: when enabled, every main F-J kernel launch is bracketed by
 * HIP events on its stream; ccfj_get_timing returns the summed kernel time
 * (ms) and launch count since the last reset. */
int ccfj_set_timing(ccfj_ctx *ctx, int enable);
int ccfj_get_timing(ccfj_ctx *ctx, double *kernel_ms, int *launches);

/* Plain single-instance FP64 GPU F-J: one thread per (f,c), serial station
 * loop, double-precision Bessel (ROCm device library) and Struve
 * (Zhang & Jin) functions, device buffers allocated, copied and freed per
 * call -- the structure of the reference's GPU engine
 * (src/FJgpu.cu:160-313, 390-483) in FP64.  Used as the speed-up baseline.
 * ybasis: 0 = J0, 1 = Y0.  kernel_ms (may be NULL) receives the kernel time. */
int ccfj_plain_fp64(int device, int ybasis, int itype, const float *U,
                    const float *r, const float *f, const float *c, int nr,
                    int nf, int nc, float *out, double *kernel_ms);

/* ---- 3. batched sub-arrays -------------------------------------------------
 * No counterpart in the reference (one array per call there).  nb
 * independent sub-arrays that share f[nf] and c[nc]; sub-array b owns rows
 * soff[b] .. soff[b+1]-1 of the concatenated U0/U1 ([rows][nf] float32) and
 * of r (each sub-array's rows sorted by r).  out: [nb][nc][nf] float32, each
 * sub-array normalised per column when normalize != 0.  One persistent
 * kernel processes every (sub-array, frequency, velocity tile) item from a
 * work queue, largest sub-arrays first; its velocity tile (64/128/256) is
 * chosen by timing on the first run of a plan. */
typedef struct ccfj_batch ccfj_batch;

/* Plan (host soff[nb+1]); allocates the plan's workspace on ctx's device. */
int ccfj_batch_create(ccfj_ctx *ctx, int mode, int itype, int normalize, int nb,
                      const long long *soff, int nf, int nc, ccfj_batch **plan);
/* Device pointers, enqueued on `stream` (NULL = the context's stream).
 * use_graph != 0: the launch sequence is captured into a HIP graph on first
 * use and replayed while the pointers are unchanged (needs a non-null
 * stream, i.e. not the legacy default stream). */
int ccfj_batch_run(ccfj_batch *plan, const float *U0, const float *U1, const float *r,
                   const float *f, const float *c, float *out, void *stream, int use_graph);
/* Velocity tile chosen by the autotuner (0 before the first run). */
int ccfj_batch_tile(ccfj_batch *plan);
int ccfj_batch_destroy(ccfj_batch *plan);
/* Host pointers: plan, copy in, run, copy out, free.  Synchronous. */
int ccfj_transform_batched(ccfj_ctx *ctx, int mode, int itype, int normalize, int nb,
                           const long long *soff, const float *U0, const float *U1,
                           const float *r, const float *f, const float *c, int nf, int nc,
                           float *out);

/* ---- 4. multi-window F-J on the device ------------------------------------
 * Replaces the per-window loop of MWFJ (src/ccfj.pyx:252-268): tapering,
 * the rFFT (double precision) and the F-J of every window run on the GPU;
 * for itype 0 the windows are right-hand sides of one MFMA contraction per
 * frequency.  Host pointers; synchronous.
 *   u[nr][npts]            traces, rows already sorted by r (double)
 *   r[nr]                  sorted distances (float32, as the reference casts)
 *   f[nf], c[nc]           float32; bins 0, fstride, ... of the rFFT are used
 *   n1, n2, delta [nwin][nr]  each window as win() describes it
 *                          (src/ccfj.pyx:148-158): zero outside [n1, n2),
 *                          sine ramps of delta samples; sorted station order
 *   xidx[nwin][nr]         -1, or the row of xrows[nx][npts] holding an
 *                          explicit window (used for degenerate windows)
 *   out[nwin][nc][nf]      float64, |..| normalised per (window, column). */
int ccfj_mwfj(ccfj_ctx *ctx, int itype, int func, int nr, int npts, int nwin, int nf,
              int nc, int fstride, const double *u, const float *r, const float *f,
              const float *c, const int *n1, const int *n2, const int *delta,
              const int *xidx, const double *xrows, int nx, double *out);
/* The same on device pointers (every array above, out included, in device
 * memory), enqueued on `stream` (NULL = the context's stream); returns once
 * enqueued.  Descriptions are not checked on the host: a (window, trace)
 * whose xidx lies outside [-1, nx) gives a zero trace.  Work arrays live in
 * the context and are reused, so a repeated call of the same or smaller
 * size allocates nothing; calls on one context must not overlap. */
int ccfj_mwfj_dev(ccfj_ctx *ctx, int itype, int func, int nr, int npts, int nwin, int nf,
                  int nc, int fstride, const double *u, const float *r, const float *f,
                  const float *c, const int *n1, const int *n2, const int *delta,
                  const int *xidx, const double *xrows, int nx, double *out, void *stream);

/* ---- 5. noise F-J end to end on the device -------------------------------
 * Row-wise Hilbert transform along frequency with scipy.fftpack's
 * convention (y = ifft(i sign(k) fft(x)), zero and even-length Nyquist bins
 * dropped; double precision inside), replacing the reference's Python loop
 * (src/ccfj.pyx:217-219).  x, y: device [rows][n] float32. */
int ccfj_hilbert_dev(ccfj_ctx *ctx, const float *x, float *y, int rows, int n, void *stream);
/* fj_noise after its host-side sort and float32 cast (src/ccfj.pyx:202-227):
 * uf[nr][ncol] real part, rows sorted by r; columns 0, fstride, ... (nf of
 * them) are used; func 1 takes the Hilbert transform of the full rows first.
 * out[nc][nf] normalised.  Host pointers; synchronous. */
int ccfj_fj_noise(ccfj_ctx *ctx, int itype, int func, int fstride, const float *uf,
                  const float *r, const float *f, const float *c, int nr, int ncol,
                  int nf, int nc, float *out);

/* ---- 5b. cross-correlation (the producer of the noise spectra) ----------
 * Reference: src/CrossCorr.cpp:33-159, include/CrossCorr.h:14-35.  Windows
 * of fftlen samples every steplen samples (fftlen >= npts: one window of
 * npts); station i uses window k when k*steplen >= startend[2i] and
 * k*steplen + fftlen < startend[2i+1]; optional one-bit normalisation and
 * spectral whitening (X / max(|X|, 1e-8)); for pair p = (Pairs[2p],
 * Pairs[2p+1]) the mean over shared windows of conj(X_a) X_b at bins
 * j*fstride, j < nf, into ncfsr/ncfsi[npairs][nf] (zero without a shared
 * window; the arrays are overwritten).  Host pointers; synchronous.
 * Single-precision FFT on the GPU. */
int ccfj_cross_correlate(ccfj_ctx *ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                         int steplen, int onebit, int whiten, int npairs, const int *pairs,
                         const int *startend, const float *data, float *ncfsr, float *ncfsi);
/* The same with the result as interleaved complex ncfs[npairs][nf][2]
 * (numpy's complex64 layout, what the reference's Python CC returns after
 * combining its two arrays, src/ccfj.pyx:145).  No counterpart in the
 * reference's C layer. */
int ccfj_cross_correlate_c64(ccfj_ctx *ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                             int steplen, int onebit, int whiten, int npairs, const int *pairs,
                             const int *startend, const float *data, float *ncfs);
/* The same on device pointers (pairs, startend, data, ncfsr, ncfsi all in
 * device memory), enqueued on `stream` (NULL = the context's stream);
 * returns once enqueued.  data is replaced by its signs when onebit is set
 * (the reference does this to its caller's array).  Pairs are not checked
 * on the host: a pair naming a station outside [0, nsta) gives zeros.  Work
 * arrays live in the context and are reused; calls on one context must not
 * overlap. */
int ccfj_cross_correlate_dev(ccfj_ctx *ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                             int steplen, int onebit, int whiten, int npairs, const int *pairs,
                             const int *startend, float *data, float *ncfsr, float *ncfsi,
                             void *stream);
/* The reference's entry point and argument block (include/CrossCorr.h:
 * 14-35), on the current device's default context.  As the reference, the
 * sums start from the values already in ncfsr/ncfsi (its wrapper passes
 * zeros), and data is replaced by its signs when ifonebit is set; nThreads
 * is ignored. */
typedef struct CC_data {
  int nsta;
  int npts;
  int nf;
  int npairs;
  int fstride;
  int fftlen;
  int steplen;
  int *Pairs;
  int *startend;
  float *data;
  float *ncfsr;
  float *ncfsi;
  int ifonebit;
  int ifspecwhitenning;
  int nThreads;
} CC_data;
int CrossCorrelation(CC_data d);

/* ---- 6. learned coarse-to-fine ----------------------------------------------
 * No counterpart in the reference.  A small CNN proposes, from a coarse F-J
 * spectrum, the velocity bands that hold modal energy per frequency; the
 * dense velocity grid is then evaluated only inside them.  All pointers are
 * device pointers unless noted; calls are enqueued on the context's stream
 * (or `stream` where given). */

/* Dense F-J on a list of items (frequency index, first velocity index,
 * count <= 256), items[n][3] int32 on the device.  out[nc][nf] is written
 * only at listed entries (not normalised).  Modes J and J_MINUS_Y.  The
 * kernel skips an item whose frequency or first velocity lies outside the
 * grid and clips its count to the grid (the list is device data, which
 * the host does not inspect). */
int ccfj_transform_items(ccfj_ctx *ctx, int mode, int itype, const float *U0, const float *U1,
                         const float *r, const float *f, const float *c, int nr, int nf, int nc,
                         const int *items, int nitems, float *out, void *stream);

/* Synthetic multimode fields: U[B*nr][nf] = sum_m A_m J0(2 pi f r / c_m(f))
 * + sigma_b * N(0,1), c_m(f) = cinf + d exp(-f/phi); modes[B][4][4] =
 * (cinf, d, phi, A) with A = 0 for unused modes; r[B][nr] sorted; the noise
 * is a counter-based hash of (seed, b, s, i). */
int ccfj_synth_fields(ccfj_ctx *ctx, int B, int nr, int nf, const float *r, const float *f,
                      const float *modes, const float *sigma, unsigned long long seed, float *U,
                      void *stream);
/* Ridge labels on the coarse grid: Y[B][H][nf] = 1 where |cgrid[h] - c_m(f)|
 * <= halfwidth for a mode with A >= amin. */
int ccfj_synth_labels(ccfj_ctx *ctx, int B, int H, int nf, const float *cgrid, const float *f,
                      const float *modes, float halfwidth, float amin, float *Y, void *stream);

/* out[B][H][W] = max over s < S of |in[B][H*S + s][W]|: the coarse spectrum
 * as the envelope of S sub-velocities per cell. */
int ccfj_pool_abs_max(ccfj_ctx *ctx, const float *in, int B, int H, int S, int W, float *out,
                      void *stream);
/* out[B][H][W] = max over s in [ranges[2h], ranges[2h+1]) of |in[B][s][W]|
 * (in[B][N][W]; ranges: device int32 [H][2], clipped to [0, N), an empty
 * range giving 0): the coarse spectrum as the envelope of the sub-velocities
 * inside each cell, any number per cell. */
int ccfj_pool_abs_max_ranges(ccfj_ctx *ctx, const float *in, int B, int N, int H, int W,
                             const int *ranges, float *out, void *stream);
/* Y[b][h*][w] = 1 where h* is the cell of the column maximum of X[b][:][w]
 * (X >= 0; first on ties); other entries of Y unchanged. */
int ccfj_label_argmax(ccfj_ctx *ctx, const float *X, int B, int H, int W, float *Y, void *stream);

/* Band-proposal CNN: conv3x3 1->32->32->32->1 (ReLU between), bf16 weights,
 * fp32 accumulation, input and output [B][1][H][W]. */
typedef struct ccfj_cnn ccfj_cnn;
int ccfj_cnn_param_count(void);
int ccfj_cnn_create(ccfj_ctx *ctx, int H, int W, int max_batch, unsigned long long seed,
                    ccfj_cnn **net);
int ccfj_cnn_destroy(ccfj_cnn *net);
/* fp32 master parameters, host arrays of ccfj_cnn_param_count() floats;
 * set also resets the optimizer state. */
int ccfj_cnn_set_params(ccfj_cnn *net, const float *params);
int ccfj_cnn_get_params(ccfj_cnn *net, float *params);
/* logits[B][H][W] (device) from x[B][H][W] (device). */
int ccfj_cnn_forward(ccfj_cnn *net, const float *x, int B, float *logits, void *stream);
/* Gradients of the mean binary cross-entropy w.r.t. every parameter (host
 * grads[param_count]) and the loss (host). */
int ccfj_cnn_gradients(ccfj_cnn *net, const float *x, const float *target, int B, float *grads,
                       float *loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8) at learning rate lr; loss (host,
 * may be NULL) receives the batch loss before the step. */
int ccfj_cnn_train_step(ccfj_cnn *net, const float *x, const float *target, int B, float lr,
                        float *loss);

#ifdef __cplusplus
}
#endif
#endif /* CCFJ_H */

# cython: language_level=3, boundscheck=False, wraparound=False
"""Cython shim between the ccfj Python API and libccfj.so (include/ccfj.h).

Mirrors the reference's typed-memoryview shims (src/ccfj-gpu.pyx:75-135
fj_full / fHr_full / fHi_full, src/ccfj.pyx:95-113 fH_full): float32
buffers are passed as raw pointers to the C ABI.  Unlike the reference the
GIL is released during the native call and a non-zero status raises
RuntimeError carrying ccfj_last_error().
"""
import numpy as np
cimport numpy as cnp
from cython.parallel cimport prange

cdef extern from "ccfj.h" nogil:
    ctypedef struct ccfj_ctx:
        pass
    int CCFJ_MODE_J
    int CCFJ_MODE_Y
    int CCFJ_MODE_J_MINUS_Y
    int CCFJ_MODE_ABS_J
    int CCFJ_MODE_ABS_H
    int FJ(float *u_f, float *r, float *f, float *out, float *c, int nc, int nr, int nf, int type)
    int FHr(float *u_f, float *r, float *f, float *out, float *c, int nc, int nr, int nf, int type)
    int FHi(float *u_f, float *r, float *f, float *out, float *c, int nc, int nr, int nf, int type)
    int FH(float *u_f, float *r, float *f, float *out, float *c, int nc, int nr, int nf, int type, int nThread)
    ctypedef struct CC_data:
        int nsta
        int npts
        int nf
        int npairs
        int fstride
        int fftlen
        int steplen
        int *Pairs
        int *startend
        float *data
        float *ncfsr
        float *ncfsi
        int ifonebit
        int ifspecwhitenning
        int nThreads
    int CrossCorrelation(CC_data d)
    int ccfj_create(int device, ccfj_ctx **ctx)
    int ccfj_destroy(ccfj_ctx *ctx)
    const char *ccfj_last_error()
    int ccfj_device_count()
    int ccfj_transform(ccfj_ctx *ctx, int mode, int itype, int normalize,
                       const float *U0, const float *U1, const float *r,
                       const float *f, const float *c, int nr, int nf, int nc,
                       float *out)
    int ccfj_transform_dev(ccfj_ctx *ctx, int mode, int itype, int normalize,
                           const float *U0, const float *U1, const float *r,
                           const float *f, const float *c, int nr, int nf, int nc,
                           float *out, void *stream)
    int ccfj_synchronize(ccfj_ctx *ctx)
    int ccfj_set_timing(ccfj_ctx *ctx, int enable)
    int ccfj_get_timing(ccfj_ctx *ctx, double *kernel_ms, int *launches)
    int ccfj_plain_fp64(int device, int ybasis, int itype, const float *U,
                        const float *r, const float *f, const float *c, int nr,
                        int nf, int nc, float *out, double *kernel_ms)
    ctypedef struct ccfj_batch:
        pass
    int ccfj_batch_create(ccfj_ctx *ctx, int mode, int itype, int normalize, int nb,
                          const long long *soff, int nf, int nc, ccfj_batch **plan)
    int ccfj_batch_run(ccfj_batch *plan, const float *U0, const float *U1, const float *r,
                       const float *f, const float *c, float *out, void *stream, int use_graph)
    int ccfj_batch_tile(ccfj_batch *plan)
    int ccfj_batch_destroy(ccfj_batch *plan)
    int ccfj_transform_batched(ccfj_ctx *ctx, int mode, int itype, int normalize, int nb,
                               const long long *soff, const float *U0, const float *U1,
                               const float *r, const float *f, const float *c, int nf, int nc,
                               float *out)
    int ccfj_hilbert_dev(ccfj_ctx *ctx, const float *x, float *y, int rows, int n, void *stream)
    int ccfj_fj_noise(ccfj_ctx *ctx, int itype, int func, int fstride, const float *uf,
                      const float *r, const float *f, const float *c, int nr, int ncol,
                      int nf, int nc, float *out)
    int ccfj_malloc(ccfj_ctx *ctx, size_t bytes, void **ptr)
    int ccfj_free(ccfj_ctx *ctx, void *ptr)
    int ccfj_memcpy(ccfj_ctx *ctx, void *dst, const void *src, size_t bytes, int kind)
    int ccfj_transform_items(ccfj_ctx *ctx, int mode, int itype, const float *U0, const float *U1,
                             const float *r, const float *f, const float *c, int nr, int nf, int nc,
                             const int *items, int nitems, float *out, void *stream)
    int ccfj_synth_fields(ccfj_ctx *ctx, int B, int nr, int nf, const float *r, const float *f,
                          const float *modes, const float *sigma, unsigned long long seed, float *U,
                          void *stream)
    int ccfj_synth_labels(ccfj_ctx *ctx, int B, int H, int nf, const float *cgrid, const float *f,
                          const float *modes, float halfwidth, float amin, float *Y, void *stream)
    int ccfj_pool_abs_max(ccfj_ctx *ctx, const float *inp, int B, int H, int S, int W, float *out,
                          void *stream)
    int ccfj_pool_abs_max_ranges(ccfj_ctx *ctx, const float *inp, int B, int N, int H, int W,
                                 const int *ranges, float *out, void *stream)
    int ccfj_label_argmax(ccfj_ctx *ctx, const float *X, int B, int H, int W, float *Y,
                          void *stream)
    int ccfj_cross_correlate(ccfj_ctx *ctx, int nsta, int npts, int nf, int fstride, int fftlen,
                             int steplen, int onebit, int whiten, int npairs, const int *pairs,
                             const int *startend, const float *data, float *ncfsr,
                             float *ncfsi) nogil
    int ccfj_cross_correlate_c64(ccfj_ctx *ctx, int nsta, int npts, int nf, int fstride,
                                 int fftlen, int steplen, int onebit, int whiten, int npairs,
                                 const int *pairs, const int *startend, const float *data,
                                 float *ncfs)
    int ccfj_cross_correlate_dev(ccfj_ctx *ctx, int nsta, int npts, int nf, int fstride,
                                 int fftlen, int steplen, int onebit, int whiten, int npairs,
                                 const int *pairs, const int *startend, float *data, float *ncfsr,
                                 float *ncfsi, void *stream)
    ctypedef struct ccfj_cnn:
        pass
    int ccfj_cnn_param_count()
    int ccfj_cnn_create(ccfj_ctx *ctx, int H, int W, int max_batch, unsigned long long seed,
                        ccfj_cnn **net)
    int ccfj_cnn_destroy(ccfj_cnn *net)
    int ccfj_cnn_set_params(ccfj_cnn *net, const float *params)
    int ccfj_cnn_get_params(ccfj_cnn *net, float *params)
    int ccfj_cnn_forward(ccfj_cnn *net, const float *x, int B, float *logits, void *stream)
    int ccfj_cnn_gradients(ccfj_cnn *net, const float *x, const float *target, int B, float *grads,
                           float *loss)
    int ccfj_cnn_train_step(ccfj_cnn *net, const float *x, const float *target, int B, float lr,
                            float *loss)
    int ccfj_mwfj(ccfj_ctx *ctx, int itype, int func, int nr, int npts, int nwin, int nf,
                  int nc, int fstride, const double *u, const float *r, const float *f,
                  const float *c, const int *n1, const int *n2, const int *delta,
                  const int *xidx, const double *xrows, int nx, double *out)
    int ccfj_mwfj_dev(ccfj_ctx *ctx, int itype, int func, int nr, int npts, int nwin, int nf,
                      int nc, int fstride, const double *u, const float *r, const float *f,
                      const float *c, const int *n1, const int *n2, const int *delta,
                      const int *xidx, const double *xrows, int nx, double *out,
                      void *stream)

MODE_J = CCFJ_MODE_J
MODE_Y = CCFJ_MODE_Y
MODE_J_MINUS_Y = CCFJ_MODE_J_MINUS_Y
MODE_ABS_J = CCFJ_MODE_ABS_J
MODE_ABS_H = CCFJ_MODE_ABS_H


cdef _check(int rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, ccfj_last_error().decode()))


def device_count():
    return ccfj_device_count()


def fj_full(float[:] U_f, float[:] r, float[:] f, float[:] out, float[:] c,
            int nc, int nr, int nf, int type):
    """J0 basis, the reference's fj_full (src/ccfj-gpu.pyx:75-92)."""
    cdef int rc
    with nogil:
        rc = FJ(&U_f[0], &r[0], &f[0], &out[0], &c[0], nc, nr, nf, type)
    _check(rc, "FJ")


def fHr_full(float[:] U_f, float[:] r, float[:] f, float[:] out, float[:] c,
             int nc, int nr, int nf, int type):
    """J0 basis, the reference's fHr_full (src/ccfj-gpu.pyx:97-114)."""
    cdef int rc
    with nogil:
        rc = FHr(&U_f[0], &r[0], &f[0], &out[0], &c[0], nc, nr, nf, type)
    _check(rc, "FHr")


def fHi_full(float[:] U_f, float[:] r, float[:] f, float[:] out, float[:] c,
             int nc, int nr, int nf, int type):
    """Y0 basis, the reference's fHi_full (src/ccfj-gpu.pyx:118-135)."""
    cdef int rc
    with nogil:
        rc = FHi(&U_f[0], &r[0], &f[0], &out[0], &c[0], nc, nr, nf, type)
    _check(rc, "FHi")


def fH_full(float[:] U_f, float[:] r, float[:] f, float[:] out, float[:] c,
            int nc, int nr, int nf, int type, int nThread=20):
    """Y0 basis under the CPU build's name (src/ccfj.pyx:95-113)."""
    cdef int rc
    with nogil:
        rc = FH(&U_f[0], &r[0], &f[0], &out[0], &c[0], nc, nr, nf, type, nThread)
    _check(rc, "FH")


def CC_full(int npts, int nsta, int nf, int fstride, int fftlen, int steplen, int ifonebit,
            int ifspecwhitenning, int nThreads, int npairs, int[:] Pairs, int[:] startend,
            float[:] data, float[:] ncfsr, float[:] ncfsi):
    """The reference's CC_full (src/ccfj.pyx:44-68, src/ccfj-gpu.pyx:46-70):
    fills the CC_data block and calls CrossCorrelation, which adds each
    pair's mean cross-spectrum to ncfsr/ncfsi [npairs * nf] and, with
    ifonebit, replaces data by its signs in place."""
    cdef CC_data d
    d.npts = npts
    d.nsta = nsta
    d.nf = nf
    d.fstride = fstride
    d.fftlen = fftlen
    d.steplen = steplen
    d.ifonebit = ifonebit
    d.ifspecwhitenning = ifspecwhitenning
    d.nThreads = nThreads
    d.npairs = npairs
    d.Pairs = &Pairs[0]
    d.startend = &startend[0]
    d.data = &data[0]
    d.ncfsr = &ncfsr[0]
    d.ncfsi = &ncfsi[0]
    cdef int rc
    with nogil:
        rc = CrossCorrelation(d)
    _check(rc, "CrossCorrelation")


def plain_fp64(int device, int ybasis, int itype, float[:, ::1] U, float[:] r,
               float[:] f, float[:] c):
    """Plain FP64 GPU F-J (speed-up baseline).  Returns (out[nc,nf], kernel_ms)."""
    cdef int nr = U.shape[0], nf = U.shape[1], nc = c.shape[0]
    out = np.empty((nc, nf), np.float32)
    cdef float[:, ::1] o = out
    cdef double ms = 0
    cdef int rc
    with nogil:
        rc = ccfj_plain_fp64(device, ybasis, itype, &U[0, 0], &r[0], &f[0], &c[0],
                             nr, nf, nc, &o[0, 0], &ms)
    _check(rc, "ccfj_plain_fp64")
    return out, ms


cdef class Context:
    """One device, one HIP stream, a grow-only device workspace."""
    cdef ccfj_ctx *ctx
    cdef readonly int device

    def __cinit__(self, int device=-1):
        self.ctx = NULL
        _check(ccfj_create(device, &self.ctx), "ccfj_create")
        self.device = device

    def __dealloc__(self):
        if self.ctx != NULL:
            ccfj_destroy(self.ctx)
            self.ctx = NULL

    def transform(self, int mode, int itype, bint normalize, U0, U1, r, f, c):
        """Host arrays in, float32 [nc, nf] out.  U0/U1: C-contiguous float32
        [nr, nf]; r, f, c: float32 vectors."""
        cdef float[:, ::1] u0 = U0
        cdef float[:, ::1] u1 = U1 if U1 is not None else U0
        cdef float[::1] rv = r
        cdef float[::1] fv = f
        cdef float[::1] cv = c
        cdef int nr = u0.shape[0], nf = u0.shape[1], nc = cv.shape[0]
        if rv.shape[0] != nr or fv.shape[0] != nf:
            raise ValueError("shape mismatch: U %s, r %d, f %d" % ((nr, nf), rv.shape[0], fv.shape[0]))
        if u1.shape[0] != nr or u1.shape[1] != nf:
            raise ValueError("U1 shape differs from U0")
        out = np.zeros((nc, nf), np.float32)
        cdef float[:, ::1] o = out
        cdef const float *p0 = &u0[0, 0] if nr > 0 and nf > 0 else NULL
        cdef const float *p1 = &u1[0, 0] if nr > 0 and nf > 0 else NULL
        cdef const float *pr = &rv[0] if nr > 0 else NULL
        cdef const float *pf = &fv[0] if nf > 0 else NULL
        cdef const float *pc = &cv[0] if nc > 0 else NULL
        cdef float *po = &o[0, 0] if nc > 0 and nf > 0 else NULL
        cdef int rc
        with nogil:
            rc = ccfj_transform(self.ctx, mode, itype, normalize, p0, p1, pr, pf, pc,
                                nr, nf, nc, po)
        _check(rc, "ccfj_transform")
        return out

    def transform_dev(self, int mode, int itype, bint normalize, size_t U0, size_t U1,
                      size_t r, size_t f, size_t c, int nr, int nf, int nc, size_t out,
                      size_t stream=0):
        """Device pointers (integers), enqueued on `stream` (0 = own stream)."""
        cdef int rc
        with nogil:
            rc = ccfj_transform_dev(self.ctx, mode, itype, normalize, <const float *>U0,
                                    <const float *>U1, <const float *>r, <const float *>f,
                                    <const float *>c, nr, nf, nc, <float *>out, <void *>stream)
        _check(rc, "ccfj_transform_dev")

    def transform_batched(self, int mode, int itype, bint normalize, soff, U0, U1, r, f, c):
        """Host arrays: soff int64 [nb+1], U0/U1 float32 [rows, nf], r float32
        [rows], f [nf], c [nc] -> float32 [nb, nc, nf]."""
        cdef long long[::1] so = soff
        cdef float[:, ::1] u0 = U0
        cdef float[:, ::1] u1 = U1 if U1 is not None else U0
        cdef float[::1] rv = r
        cdef float[::1] fv = f
        cdef float[::1] cv = c
        cdef int nb = so.shape[0] - 1, nf = fv.shape[0], nc = cv.shape[0]
        if u0.shape[0] != so[nb] or u0.shape[1] != nf or rv.shape[0] != so[nb]:
            raise ValueError("batched shapes do not match the offsets")
        out = np.zeros((nb, nc, nf), np.float32)
        cdef float[:, :, ::1] o = out
        cdef int rc
        with nogil:
            rc = ccfj_transform_batched(self.ctx, mode, itype, normalize, nb, &so[0],
                                        &u0[0, 0], &u1[0, 0], &rv[0], &fv[0], &cv[0],
                                        nf, nc, &o[0, 0, 0])
        _check(rc, "ccfj_transform_batched")
        return out

    def fj_noise(self, int itype, int func, int fstride, uf, r, f, c):
        """uf float32 [nr, ncol] (rows sorted by r), r/f/c float32 ->
        normalised float32 [nc, nf]."""
        cdef float[:, ::1] u = uf
        cdef float[::1] rv = r
        cdef float[::1] fv = f
        cdef float[::1] cv = c
        cdef int nr = u.shape[0], ncol = u.shape[1], nf = fv.shape[0], nc = cv.shape[0]
        out = np.zeros((nc, nf), np.float32)
        cdef float[:, ::1] o = out
        cdef const float *pu = &u[0, 0] if nr > 0 and ncol > 0 else NULL
        cdef const float *pr = &rv[0] if nr > 0 else NULL
        cdef const float *pf = &fv[0] if nf > 0 else NULL
        cdef const float *pc = &cv[0] if nc > 0 else NULL
        cdef float *po = &o[0, 0] if nc > 0 and nf > 0 else NULL
        cdef int rc
        with nogil:
            rc = ccfj_fj_noise(self.ctx, itype, func, fstride, pu, pr, pf, pc, nr, ncol, nf, nc, po)
        _check(rc, "ccfj_fj_noise")
        return out

    def hilbert_dev(self, size_t x, size_t y, int rows, int n, size_t stream=0):
        cdef int rc
        with nogil:
            rc = ccfj_hilbert_dev(self.ctx, <const float *>x, <float *>y, rows, n, <void *>stream)
        _check(rc, "ccfj_hilbert_dev")

    def mwfj(self, int itype, int func, int fstride, u, r, f, c, n1, n2, delta, xidx, xrows):
        """Device multi-window F-J.  u float64 [nr, npts] (rows sorted by r),
        r/f/c float32, n1/n2/delta/xidx int32 [nwin, nr], xrows float64
        [nx, npts].  Returns float64 [nwin, nc, nf]."""
        cdef double[:, ::1] uu = u
        cdef float[::1] rv = r
        cdef float[::1] fv = f
        cdef float[::1] cv = c
        cdef int[:, ::1] a1 = n1
        cdef int[:, ::1] a2 = n2
        cdef int[:, ::1] ad = delta
        cdef int[:, ::1] ax = xidx
        cdef double[:, ::1] xr = xrows
        cdef int nr = uu.shape[0], npts = uu.shape[1], nwin = a1.shape[0]
        cdef int nf = fv.shape[0], nc = cv.shape[0], nx = xr.shape[0]
        out = np.empty((nwin, nc, nf), np.float64)
        cdef double[:, :, ::1] o = out
        cdef const double *px = &xr[0, 0] if nx > 0 else NULL
        cdef int rc
        with nogil:
            rc = ccfj_mwfj(self.ctx, itype, func, nr, npts, nwin, nf, nc, fstride, &uu[0, 0],
                           &rv[0], &fv[0], &cv[0], &a1[0, 0], &a2[0, 0], &ad[0, 0], &ax[0, 0],
                           px, nx, &o[0, 0, 0])
        _check(rc, "ccfj_mwfj")
        return out

    def mwfj_dev(self, int itype, int func, int fstride, int nr, int npts, int nwin, int nf,
                 int nc, size_t u, size_t r, size_t f, size_t c, size_t n1, size_t n2,
                 size_t delta, size_t xidx, size_t xrows, int nx, size_t out, size_t stream=0):
        """Device multi-window F-J on device pointers (layouts as mwfj),
        enqueued on `stream` (0 = the context's stream)."""
        cdef int rc
        with nogil:
            rc = ccfj_mwfj_dev(self.ctx, itype, func, nr, npts, nwin, nf, nc, fstride,
                               <const double *>u, <const float *>r, <const float *>f,
                               <const float *>c, <const int *>n1, <const int *>n2,
                               <const int *>delta, <const int *>xidx, <const double *>xrows, nx,
                               <double *>out, <void *>stream)
        _check(rc, "ccfj_mwfj_dev")

    def cross_correlate(self, int nsta, int npts, int nf, int fstride, int fftlen, int steplen,
                        bint onebit, bint whiten, pairs, startend, data):
        """Device cross-correlation (host arrays): pairs int32 [2 npairs],
        startend int32 [2 nsta], data float32 [nsta * npts] ->
        (ncfsr, ncfsi) float32 [npairs, nf]."""
        cdef int[::1] pv = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        cdef int[::1] se = np.ascontiguousarray(startend, dtype=np.int32).reshape(-1)
        cdef float[::1] dv = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
        cdef int npairs = pv.shape[0] // 2
        if se.shape[0] < 2 * nsta or dv.shape[0] < <size_t>nsta * npts:
            raise ValueError("startend or data shorter than nsta / npts require")
        re = np.zeros((npairs, nf), np.float32)
        im = np.zeros((npairs, nf), np.float32)
        cdef float[:, ::1] rv = re
        cdef float[:, ::1] iv = im
        cdef const int *pp = &pv[0] if npairs > 0 else NULL
        cdef float *rp = &rv[0, 0] if npairs > 0 else NULL
        cdef float *ip = &iv[0, 0] if npairs > 0 else NULL
        cdef int rc
        with nogil:
            rc = ccfj_cross_correlate(self.ctx, nsta, npts, nf, fstride, fftlen, steplen, onebit,
                                      whiten, npairs, pp, &se[0], &dv[0], rp, ip)
        _check(rc, "ccfj_cross_correlate")
        return re, im

    def cross_correlate_c64(self, int nsta, int npts, int nf, int fstride, int fftlen,
                            int steplen, bint onebit, bint whiten, pairs, startend, data):
        """As cross_correlate, returning complex64 [npairs, nf] directly."""
        cdef int[::1] pv = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        cdef int[::1] se = np.ascontiguousarray(startend, dtype=np.int32).reshape(-1)
        cdef float[::1] dv = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
        cdef int npairs = pv.shape[0] // 2
        if se.shape[0] < 2 * nsta or dv.shape[0] < <size_t>nsta * npts:
            raise ValueError("startend or data shorter than nsta / npts require")
        out = np.empty((npairs, nf), np.complex64)
        cdef float[:, ::1] ov = out.view(np.float32)
        cdef const int *pp = &pv[0] if npairs > 0 else NULL
        cdef float *op = &ov[0, 0] if npairs > 0 else NULL
        cdef int rc
        with nogil:
            rc = ccfj_cross_correlate_c64(self.ctx, nsta, npts, nf, fstride, fftlen, steplen,
                                          onebit, whiten, npairs, pp, &se[0], &dv[0], op)
        _check(rc, "ccfj_cross_correlate_c64")
        return out

    def cross_correlate_dev(self, int nsta, int npts, int nf, int fstride, int fftlen,
                            int steplen, bint onebit, bint whiten, int npairs, size_t pairs,
                            size_t startend, size_t data, size_t ncfsr, size_t ncfsi,
                            size_t stream=0):
        """Device cross-correlation on device pointers (layouts as
        cross_correlate), enqueued on `stream` (0 = the context's stream)."""
        cdef int rc
        with nogil:
            rc = ccfj_cross_correlate_dev(self.ctx, nsta, npts, nf, fstride, fftlen, steplen,
                                          onebit, whiten, npairs, <const int *>pairs,
                                          <const int *>startend, <float *>data, <float *>ncfsr,
                                          <float *>ncfsi, <void *>stream)
        _check(rc, "ccfj_cross_correlate_dev")

    def batch_plan(self, int mode, int itype, bint normalize, soff, int nf, int nc):
        return BatchPlan(self, mode, itype, normalize, soff, nf, nc)

    def synchronize(self):
        cdef int rc
        with nogil:
            rc = ccfj_synchronize(self.ctx)
        _check(rc, "ccfj_synchronize")

    def set_timing(self, bint enable):
        _check(ccfj_set_timing(self.ctx, enable), "ccfj_set_timing")

    def get_timing(self):
        """(summed main-kernel ms, launches) since the last call."""
        cdef double ms = 0
        cdef int n = 0
        _check(ccfj_get_timing(self.ctx, &ms, &n), "ccfj_get_timing")
        return ms, n


cdef class BatchPlan:
    """A batched-transform plan on a Context (device pointers, optional
    HIP-graph replay)."""
    cdef ccfj_batch *plan
    cdef Context owner

    def __cinit__(self, Context ctx, int mode, int itype, bint normalize, soff, int nf, int nc):
        cdef long long[::1] so = np.ascontiguousarray(soff, dtype=np.int64)
        self.plan = NULL
        self.owner = ctx
        _check(ccfj_batch_create(ctx.ctx, mode, itype, normalize, so.shape[0] - 1, &so[0],
                                 nf, nc, &self.plan), "ccfj_batch_create")

    def __dealloc__(self):
        if self.plan != NULL:
            ccfj_batch_destroy(self.plan)
            self.plan = NULL

    def run(self, size_t U0, size_t U1, size_t r, size_t f, size_t c, size_t out,
            size_t stream=0, bint use_graph=False):
        cdef int rc
        with nogil:
            rc = ccfj_batch_run(self.plan, <const float *>U0, <const float *>U1,
                                <const float *>r, <const float *>f, <const float *>c,
                                <float *>out, <void *>stream, use_graph)
        _check(rc, "ccfj_batch_run")

    @property
    def tile(self):
        return ccfj_batch_tile(self.plan)


cdef class DeviceArray:
    """A float32 or int32 array in device memory of a Context."""
    cdef void *ptr
    cdef Context owner
    cdef readonly tuple shape
    cdef readonly object dtype
    cdef readonly size_t nbytes

    def __cinit__(self, Context ctx, shape, dtype=np.float32):
        self.ptr = NULL
        self.owner = ctx
        self.shape = tuple(int(n) for n in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.float32), np.dtype(np.int32)):
            raise TypeError("DeviceArray holds float32 or int32")
        self.nbytes = int(np.prod(self.shape)) * 4
        _check(ccfj_malloc(ctx.ctx, self.nbytes if self.nbytes else 4, &self.ptr), "ccfj_malloc")

    def __dealloc__(self):
        if self.ptr != NULL:
            ccfj_free(self.owner.ctx, self.ptr)
            self.ptr = NULL

    @property
    def addr(self):
        return <size_t>self.ptr

    def offset(self, size_t elements):
        """Address of element `elements` (flat index)."""
        return <size_t>self.ptr + 4 * elements

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        cdef const unsigned char[::1] v = a.reshape(-1).view(np.uint8)
        if self.nbytes:
            _check(ccfj_memcpy(self.owner.ctx, self.ptr, &v[0], self.nbytes, 0), "ccfj_memcpy")
        return self

    def download(self):
        out = np.empty(self.shape, self.dtype)
        cdef unsigned char[::1] v = out.reshape(-1).view(np.uint8)
        if self.nbytes:
            _check(ccfj_memcpy(self.owner.ctx, &v[0], self.ptr, self.nbytes, 1), "ccfj_memcpy")
        return out


def device_array(Context ctx, a, dtype=np.float32):
    a = np.ascontiguousarray(a, dtype=dtype)
    return DeviceArray(ctx, a.shape, dtype).upload(a)


def transform_items(Context ctx, int mode, int itype, size_t U0, size_t U1, size_t r, size_t f,
                    size_t c, int nr, int nf, int nc, size_t items, int nitems, size_t out,
                    size_t stream=0):
    cdef int rc
    with nogil:
        rc = ccfj_transform_items(ctx.ctx, mode, itype, <const float *>U0, <const float *>U1,
                                  <const float *>r, <const float *>f, <const float *>c, nr, nf, nc,
                                  <const int *>items, nitems, <float *>out, <void *>stream)
    _check(rc, "ccfj_transform_items")


def synth_fields(Context ctx, int B, int nr, int nf, size_t r, size_t f, size_t modes,
                 size_t sigma, unsigned long long seed, size_t U):
    _check(ccfj_synth_fields(ctx.ctx, B, nr, nf, <const float *>r, <const float *>f,
                             <const float *>modes, <const float *>sigma, seed, <float *>U, NULL),
           "ccfj_synth_fields")


def synth_labels(Context ctx, int B, int H, int nf, size_t cgrid, size_t f, size_t modes,
                 float halfwidth, float amin, size_t Y):
    _check(ccfj_synth_labels(ctx.ctx, B, H, nf, <const float *>cgrid, <const float *>f,
                             <const float *>modes, halfwidth, amin, <float *>Y, NULL),
           "ccfj_synth_labels")


def pool_abs_max(Context ctx, size_t x, int B, int H, int S, int W, size_t out):
    _check(ccfj_pool_abs_max(ctx.ctx, <const float *>x, B, H, S, W, <float *>out, NULL),
           "ccfj_pool_abs_max")


def pool_abs_max_ranges(Context ctx, size_t x, int B, int N, int H, int W, size_t ranges,
                        size_t out):
    _check(ccfj_pool_abs_max_ranges(ctx.ctx, <const float *>x, B, N, H, W, <const int *>ranges,
                                    <float *>out, NULL), "ccfj_pool_abs_max_ranges")


def label_argmax(Context ctx, size_t x, int B, int H, int W, size_t y):
    _check(ccfj_label_argmax(ctx.ctx, <const float *>x, B, H, W, <float *>y, NULL),
           "ccfj_label_argmax")


def cnn_param_count():
    return ccfj_cnn_param_count()


cdef class CNN:
    """Band-proposal CNN on a Context (HIP kernels, bf16 weights)."""
    cdef ccfj_cnn *net
    cdef Context owner
    cdef readonly int H, W, max_batch

    def __cinit__(self, Context ctx, int H, int W, int max_batch, unsigned long long seed=0):
        self.net = NULL
        self.owner = ctx
        self.H, self.W, self.max_batch = H, W, max_batch
        _check(ccfj_cnn_create(ctx.ctx, H, W, max_batch, seed, &self.net), "ccfj_cnn_create")

    def __dealloc__(self):
        if self.net != NULL:
            ccfj_cnn_destroy(self.net)
            self.net = NULL

    def get_params(self):
        p = np.empty(ccfj_cnn_param_count(), np.float32)
        cdef float[::1] v = p
        _check(ccfj_cnn_get_params(self.net, &v[0]), "ccfj_cnn_get_params")
        return p

    def set_params(self, params):
        cdef float[::1] v = np.ascontiguousarray(params, dtype=np.float32)
        if v.shape[0] != ccfj_cnn_param_count():
            raise ValueError("parameter count mismatch")
        _check(ccfj_cnn_set_params(self.net, &v[0]), "ccfj_cnn_set_params")

    def forward(self, size_t x, int B, size_t logits):
        _check(ccfj_cnn_forward(self.net, <const float *>x, B, <float *>logits, NULL),
               "ccfj_cnn_forward")

    def gradients(self, size_t x, size_t target, int B):
        g = np.empty(ccfj_cnn_param_count(), np.float32)
        cdef float[::1] v = g
        cdef float loss = 0
        _check(ccfj_cnn_gradients(self.net, <const float *>x, <const float *>target, B, &v[0],
                                  &loss), "ccfj_cnn_gradients")
        return g, loss

    def train_step(self, size_t x, size_t target, int B, float lr):
        cdef float loss = 0
        _check(ccfj_cnn_train_step(self.net, <const float *>x, <const float *>target, B, lr,
                                   &loss), "ccfj_cnn_train_step")
        return loss


def gather_real_rows(src, order):
    """float32 [nr, ncol] = real part of src[order] (rows permuted), for a
    C-contiguous complex64 or float32 src [nr, ncol]; OpenMP threads over
    rows.  The same values as np.array(src[order], dtype=np.float32)."""
    if src.ndim != 2 or not src.flags.c_contiguous:
        raise ValueError("src must be a C-contiguous 2-D array")
    cdef Py_ssize_t nr = src.shape[0], ncol = src.shape[1]
    cdef cnp.int64_t[::1] idx = np.ascontiguousarray(order, dtype=np.int64)
    if idx.shape[0] != nr:
        raise ValueError("order has %d entries for %d rows" % (idx.shape[0], nr))
    if nr and (np.min(order) < 0 or np.max(order) >= nr):
        raise ValueError("order out of range")
    cdef int step
    if src.dtype == np.complex64:
        step = 2
    elif src.dtype == np.float32:
        step = 1
    else:
        raise TypeError("src dtype %s" % src.dtype)
    out = np.empty((nr, ncol), np.float32)
    if nr == 0 or ncol == 0:
        return out
    cdef float[:, ::1] s = src.view(np.float32)
    cdef float[:, ::1] o = out
    cdef Py_ssize_t i, j, k
    for i in prange(nr, nogil=True, schedule="static"):
        k = idx[i]
        for j in range(ncol):
            o[i, j] = s[k, step * j]
    return out


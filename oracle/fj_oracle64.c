/*
 * fj_oracle64.c -- FP64 oracle for the F-J transform.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ (and smoke()) as the accuracy
 * judge of the HIP kernels.  The product path never links or calls it.
 *
 * It evaluates the integral that the reference discretises, in double
 * precision, on the float32 inputs the reference's Python layer hands to the
 * native code (src/ccfj.pyx:161-179):
 *
 *   itype 0 (src/FJcpu.cpp:348-357, trapezoid over sorted stations):
 *     O[j,i] = sum_{s=1}^{nr-1} d_s/2 * (U_{s-1} Z0(k r_{s-1}) r_{s-1}
 *                                       + U_s Z0(k r_s) r_s)
 *   itype 1 (src/FJcpu.cpp:253-268, exact integral of the linear segment):
 *     O[j,i] = sum_s int_{r_{s-1}}^{r_s} (g1 + (rho - r_{s-1}) b) Z0(k rho) rho drho
 *     with b = (U_s - U_{s-1}) / d_s,
 *   where k = 2 pi f_i / c_j, d_s = max(r_s - r_{s-1}, 0.1) and Z0 is J0
 *   (FJ) or Y0 (FH).  The reference forms k in float with pi = 3.14159265358
 *   and the argument k*r in float; the oracle uses the exact double k and
 *   k*rho instead (that is the integral the reference approximates).
 *
 * itype 1 is evaluated by composite Gauss-Legendre quadrature: each segment
 * is cut into panels no wider than 0.5 in x = k*rho, each integrated with
 * GL_N nodes; the integrand is entire in rho (J basis) so the rule is exact
 * to far below 1e-12 relative.  For the Y basis rho*Y0(k rho) has a
 * rho*log(rho) singularity at rho = 0, so below x = 0.5 the panels are
 * graded geometrically away from it (each panel [a, 2a]).
 *
 * Special functions: the C library's double j0, y0 (glibc); their accuracy
 * is checked against mpmath in the tests.
 *
 * f = 0 with itype 1: the reference divides by k = 0 and yields a NaN column
 * (src/FJcpu.cpp:262-267).  The oracle returns NaN there too, so that the
 * comparison covers the same entries.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define GL_N 8
static double gl_x[GL_N], gl_w[GL_N];
static int gl_ready = 0;

/* Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_n. */
static void gl_init(void)
{
    if (gl_ready) return;
    const int n = GL_N;
    for (int i = 0; i < n; i++) {
        double x = cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
        }
        gl_x[i] = x;
        gl_w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    gl_ready = 1;
}

static inline double z0f(double x, int ybasis) { return ybasis ? y0(x) : j0(x); }

/* int_{a}^{b} (g1 + (rho - r1) * slope) * Z0(k rho) * rho drho by GL panels. */
static double panel_integral(double a, double b, double r1, double g1,
                             double slope, double k, int ybasis)
{
    double s = 0.0;
    double hm = 0.5 * (b - a), cm = 0.5 * (b + a);
    for (int q = 0; q < GL_N; q++) {
        double rho = cm + hm * gl_x[q];
        double L = g1 + (rho - r1) * slope;
        s += gl_w[q] * L * z0f(k * rho, ybasis) * rho;
    }
    return s * hm;
}

/* int over [lo0, r2] (lo0 <= r2) of the line anchored at (r1, g1). */
static double span_integral(double lo0, double r2, double r1, double g1, double slope,
                            double k, int ybasis)
{
    double h = r2 - lo0;
    if (h <= 0.0) return 0.0;
    double s = 0.0;
    double lo = lo0;
    const double wmax = 0.5 / k;  /* panel width limit: 0.5 in x */
    if (ybasis) {
        /* rho*Y0(k rho) has a rho*log(rho) singularity at rho = 0: panels
         * are graded geometrically away from it while x < 0.5. */
        if (lo <= 0.0) {
            double edge = fmin(h, wmax);
            double e = edge;
            for (int g = 0; g < 60; g++) {
                s += panel_integral(e * 0.5, e, r1, g1, slope, k, ybasis);
                e *= 0.5;
            }
            lo = edge;
        }
        while (lo < r2 && k * lo < 0.5) {
            double b = fmin(fmin(2.0 * lo, r2), lo + wmax);
            s += panel_integral(lo, b, r1, g1, slope, k, ybasis);
            lo = b;
        }
        if (lo >= r2) return s;
    }
    int np = (int)ceil((r2 - lo) / wmax);
    if (np < 1) np = 1;
    double w = (r2 - lo) / np;
    for (int p = 0; p < np; p++) {
        double a = lo + p * w;
        double b = (p == np - 1) ? r2 : a + w;
        s += panel_integral(a, b, r1, g1, slope, k, ybasis);
    }
    return s;
}

/* Signed segment integral: a reversed segment (r2 < r1, unsorted input at
 * the C level) integrates from r1 down to r2, as the reference's closed
 * form does. */
static double segment_integral(double r1, double r2, double g1, double slope,
                               double k, int ybasis)
{
    if (r2 >= r1) return span_integral(r1, r2, r1, g1, slope, k, ybasis);
    return -span_integral(r2, r1, r1, g1, slope, k, ybasis);
}

static double column_entry(const float *U, const float *r, int nr, int nf,
                           int i, double k, int itype, int ybasis)
{
    double acc = 0.0;
    if (itype == 0) {
        for (int s = 1; s < nr; s++) {
            double r1 = r[s - 1], r2 = r[s];
            double d = fmax(r2 - r1, 0.1);
            double g1 = U[i + (int64_t)(s - 1) * nf], g2 = U[i + (int64_t)s * nf];
            acc += 0.5 * d * (g1 * z0f(k * r1, ybasis) * r1 + g2 * z0f(k * r2, ybasis) * r2);
        }
        return acc;
    }
    if (k == 0.0) return NAN;
    for (int s = 1; s < nr; s++) {
        double r1 = r[s - 1], r2 = r[s];
        double d = fmax(r2 - r1, 0.1);
        double g1 = U[i + (int64_t)(s - 1) * nf], g2 = U[i + (int64_t)s * nf];
        acc += segment_integral(r1, r2, g1, (g2 - g1) / d, k, ybasis);
    }
    return acc;
}

/* out[j][q] (double, shape [nc][nfi]) for frequency columns fidx[q]
 * (fidx == NULL: all nf columns, nfi == nf).  U is [nr][nf] float32 with r
 * ascending.  ybasis: 0 = J0 (FJ), 1 = Y0 (FH). */
int oracle64_fj(const float *U, const float *r, const float *f, const float *c,
                int nc, int nr, int nf, int itype, int ybasis, double *out,
                int nthreads, const int *fidx, int nfi)
{
    if (itype != 0 && itype != 1) return -1;
    gl_init();
    if (nthreads < 1) nthreads = 1;
    int64_t total = (int64_t)nfi * nc;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int64_t t = 0; t < total; t++) {
        int q = (int)(t / nc);
        int j = (int)(t % nc);
        int i = fidx ? fidx[q] : q;
        double k = 2.0 * M_PI * (double)f[i] / (double)c[j];
        out[q + (int64_t)j * nfi] = column_entry(U, r, nr, nf, i, k, itype, ybasis);
    }
    return 0;
}

/* Double-precision special functions, exported so the tests can pin them
 * against mpmath. */
double oracle64_j0(double x) { return j0(x); }
double oracle64_y0(double x) { return y0(x); }

/* One segment integral, exported for the closed-form checks in the tests. */
double oracle64_segment(double r1, double r2, double g1, double slope,
                        double k, int ybasis)
{
    gl_init();
    return segment_integral(r1, r2, g1, slope, k, ybasis);
}

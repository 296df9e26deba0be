/*
 * fjcpu_port.c -- CPU restatement of the reference F-J engine (FJcpu).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the "port" oracle: it restates the
 * float/double arithmetic of src/FJcpu.cpp so that its float32 outputs match
 * the reference bit for bit (same libm calls, same operation order, same
 * float roundings).  It is used by tests/ as a checker and by bench.py's
 * cpu_baseline leg as the timed CPU baseline.  The product path never links
 * or calls it.
 *
 * Reference map (file:line under the reference tree):
 *   bessel_j0_nr   <- src/FJcpu.cpp:21-49   (Numerical Recipes 6.5 J0)
 *   bessel_j1_nr   <- src/FJcpu.cpp:51-80   (Numerical Recipes 6.5 J1)
 *   bessel_y0_nr   <- src/FJcpu.cpp:82-110  (Numerical Recipes 6.5 Y0; its
 *                     asymptotic Q coefficient differs from J0's in the 5th
 *                     digit and is kept as the reference has it)
 *   bessel_y1_nr   <- src/FJcpu.cpp:114-144 (Numerical Recipes 6.5 Y1)
 *   struve_h0_zj   <- src/FJcpu.cpp:146-186 (Zhang & Jin 1996, STVH0)
 *   struve_h1_zj   <- src/FJcpu.cpp:188-228 (Zhang & Jin 1996, STVH1)
 *   linseg_sum     <- src/FJcpu.cpp:230-273 (integral_J) and :276-319
 *                     (integral_Y): closed-form linear-segment sum
 *   trapz_sum      <- src/FJcpu.cpp:323-362 (trap_J, glibc j0) and :364-399
 *                     (trap_Y, NR Y0)
 *   fjport_FJ/FH   <- src/FJcpu.cpp:401-419 and include/FJcpu.h:19-20
 *
 * Differences from the reference, on purpose: trap_J's debug print of f[0..9]
 * (src/FJcpu.cpp:329-331, reads past f when nf<10) is not reproduced, and
 * the thread count is a per-call OpenMP clause, not a global
 * omp_set_num_threads side effect.
 *
 * Build: plain C, -O3 -fopenmp, no -march (x86-64 baseline: no FMA
 * contraction), exactly the flag set of the reference's CPU build.
 */
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/FJcpu.cpp:16 redefines M_PI with this truncated literal. */
#define REF_PI 3.14159265358
/* Zhang & Jin use the full double pi inside the Struve routines. */
#define ZJ_PI 3.141592653589793

/* ---- Numerical Recipes 6.5 rational / asymptotic Bessel forms ---------- */

static double bessel_j0_nr(double x)
{
    double ax = fabs(x);
    if (ax < 8.0) {
        double y = x * x;
        double num = 57568490574.0 + y * (-13362590354.0 + y * (651619640.7
                   + y * (-11214424.18 + y * (77392.33017 + y * (-184.9052456)))));
        double den = 57568490411.0 + y * (1029532985.0 + y * (9494680.718
                   + y * (59272.64853 + y * (267.8532712 + y * 1.0))));
        return num / den;
    }
    double z = 8.0 / ax;
    double y = z * z;
    double ph = ax - 0.785398164;
    double p = 1.0 + y * (-0.1098628627e-2 + y * (0.2734510407e-4
             + y * (-0.2073370639e-5 + y * 0.2093887211e-6)));
    double q = -0.1562499995e-1 + y * (0.1430488765e-3
             + y * (-0.6911147651e-5 + y * (0.7621095161e-6
             - y * 0.934935152e-7)));
    return sqrt(0.636619772 / ax) * (cos(ph) * p - z * sin(ph) * q);
}

static double bessel_j1_nr(double x)
{
    double ax = fabs(x);
    if (ax < 8.0) {
        double y = x * x;
        double num = x * (72362614232.0 + y * (-7895059235.0 + y * (242396853.1
                   + y * (-2972611.439 + y * (15704.48260 + y * (-30.16036606))))));
        double den = 144725228442.0 + y * (2300535178.0 + y * (18583304.74
                   + y * (99447.43394 + y * (376.9991397 + y * 1.0))));
        return num / den;
    }
    double z = 8.0 / ax;
    double y = z * z;
    double ph = ax - 2.356194491;
    double p = 1.0 + y * (0.183105e-2 + y * (-0.3516396496e-4
             + y * (0.2457520174e-5 + y * (-0.240337019e-6))));
    double q = 0.04687499995 + y * (-0.2002690873e-3
             + y * (0.8449199096e-5 + y * (-0.88228987e-6
             + y * 0.105787412e-6)));
    double v = sqrt(0.636619772 / ax) * (cos(ph) * p - z * sin(ph) * q);
    if (x < 0.0) v = -v;
    return v;
}

static double bessel_y0_nr(double x)
{
    if (x < 8.0) {
        double y = x * x;
        double num = -2957821389.0 + y * (7062834065.0 + y * (-512359803.6
                   + y * (10879881.29 + y * (-86327.92757 + y * 228.4622733))));
        double den = 40076544269.0 + y * (745249964.8 + y * (7189466.438
                   + y * (47447.26470 + y * (226.1030244 + y * 1.0))));
        return (num / den) + 0.636619772 * bessel_j0_nr(x) * log(x);
    }
    double z = 8.0 / x;
    double y = z * z;
    double ph = x - 0.785398164;
    double p = 1.0 + y * (-0.1098628627e-2 + y * (0.2734510407e-4
             + y * (-0.2073370639e-5 + y * 0.2093887211e-6)));
    double q = -0.1562499995e-1 + y * (0.1430488765e-3
             + y * (-0.6911147651e-5 + y * (0.7621095161e-6
             + y * (-0.934945152e-7))));
    return sqrt(0.636619772 / x) * (sin(ph) * p + z * cos(ph) * q);
}

static double bessel_y1_nr(double x)
{
    if (x < 8.0) {
        double y = x * x;
        double num = x * (-0.4900604943e13 + y * (0.1275274390e13
                   + y * (-0.5153438139e11 + y * (0.7349264551e9
                   + y * (-0.4237922726e7 + y * 0.8511937935e4)))));
        double den = 0.2499580570e14 + y * (0.4244419664e12
                   + y * (0.3733650367e10 + y * (0.2245904002e8
                   + y * (0.1020426050e6 + y * (0.3549632885e3 + y)))));
        return (num / den) + 0.636619772 * (bessel_j1_nr(x) * log(x) - 1.0 / x);
    }
    double z = 8.0 / x;
    double y = z * z;
    double ph = x - 2.356194491;
    double p = 1.0 + y * (0.183105e-2 + y * (-0.3516396496e-4
             + y * (0.2457520174e-5 + y * (-0.240337019e-6))));
    double q = 0.04687499995 + y * (-0.2002690873e-3
             + y * (0.8449199096e-5 + y * (-0.88228987e-6
             + y * 0.105787412e-6)));
    return sqrt(0.636619772 / x) * (sin(ph) * p + z * cos(ph) * q);
}

/* ---- Zhang & Jin (1996) Struve H0, H1 ----------------------------------- */

static double struve_h0_zj(double x)
{
    double s = 1.0, r = 1.0;
    if (x <= 20.0) {
        double a0 = 2.0 * x / ZJ_PI;
        for (int k = 1; k < 61; k++) {
            r = -r * x / (2.0 * k + 1.0) * x / (2.0 * k + 1.0);
            s = s + r;
            if (fabs(r) < fabs(s) * 1.0e-12) break;
        }
        return a0 * s;
    }
    int km = (int)(0.5 * (x + 1.0));
    if (x >= 50.0) km = 25;
    for (int k = 1; k <= km; k++) {
        r = -r * (2.0 * k - 1.0) * (2.0 * k - 1.0) / x / x;
        s = s + r;
        if (fabs(r) < fabs(s) * 1.0e-12) break;
    }
    double t = 4.0 / x;
    double t2 = t * t;
    double p0 = ((((-.37043e-5 * t2 + .173565e-4) * t2 - .487613e-4) * t2
                + .17343e-3) * t2 - 0.1753062e-2) * t2 + .3989422793;
    double q0 = t * (((((.32312e-5 * t2 - 0.142078e-4) * t2 + 0.342468e-4) * t2
                - 0.869791e-4) * t2 + 0.4564324e-3) * t2 - 0.0124669441);
    double ta0 = x - 0.25 * ZJ_PI;
    double by0 = 2.0 / sqrt(x) * (p0 * sin(ta0) + q0 * cos(ta0));
    return 2.0 / (ZJ_PI * x) * s + by0;
}

static double struve_h1_zj(double x)
{
    double r = 1.0, s;
    if (x <= 20.0) {
        s = 0.0;
        double a0 = -2.0 / ZJ_PI;
        for (int k = 1; k <= 60; k++) {
            r = -r * x * x / (4.0 * k * k - 1.0);
            s = s + r;
            if (fabs(r) < fabs(s) * 1.0e-12) break;
        }
        return a0 * s;
    }
    s = 1.0;
    int km = (int)(0.5 * x);
    if (x > 50.0) km = 25;
    for (int k = 1; k <= km; k++) {
        r = -r * (4.0 * k * k - 1.0) / (x * x);
        s = s + r;
        if (fabs(r) < fabs(s) * 1.0e-12) break;
    }
    double t = 4.0 / x;
    double t2 = t * t;
    double p1 = ((((0.42414e-5 * t2 - 0.20092e-4) * t2 + 0.580759e-4) * t2
                - 0.223203e-3) * t2 + 0.29218256e-2) * t2 + 0.3989422819;
    double q1 = t * (((((-0.36594e-5 * t2 + 0.1622e-4) * t2 - 0.398708e-4) * t2
                + 0.1064741e-3) * t2 - 0.63904e-3) * t2 + 0.0374008364);
    double ta1 = x - 0.75 * ZJ_PI;
    double by1 = 2.0 / sqrt(x) * (p1 * sin(ta1) + q1 * cos(ta1));
    return 2.0 / ZJ_PI * (1.0 + s / (x * x)) + by1;
}

/* ---- per-(f,c) sums ------------------------------------------------------ */

typedef double (*bessel_fn)(double);

/* Linear-segment closed form (src/FJcpu.cpp:253-268).  z0/z1 are the order-0
 * and order-1 functions of the basis.  Every `acc +=` rounds to float, and
 * k, a, b, dr0, B01, B02 are floats, as in the reference. */
static inline __attribute__((always_inline))
float linseg_sum(const float *U, const float *r, int nr, int nf, int i,
                 float k, bessel_fn z0, bessel_fn z1)
{
    float acc = 0.0f;
    for (int ir = 1; ir < nr; ir++) {
        float g1 = U[i + (ir - 1) * nf];
        float g2 = U[i + ir * nf];
        float r1 = r[ir - 1];
        float r2 = r[ir];
        float dr0 = fmaxf((r2 - r1), 0.1);
        float a = g1 - r1 * (g2 - g1) / dr0;
        float b = (g2 - g1) / dr0;
        acc += a * (r2 * z1(k * r2) - r1 * z1(k * r1)) / k;
        acc += b * (r2 * r2 * z1(k * r2) - r1 * r1 * z1(k * r1)) / k;
        acc += b * (r2 * z0(k * r2) - r1 * z0(k * r1)) / k / k;
        float bz2 = k * r2 * z0(k * r2)
                  + REF_PI * k * r2 * (z1(k * r2) * struve_h0_zj(k * r2)
                                      - z0(k * r2) * struve_h1_zj(k * r2)) / 2;
        float bz1 = k * r1 * z0(k * r1)
                  + REF_PI * k * r1 * (z1(k * r1) * struve_h0_zj(k * r1)
                                      - z0(k * r1) * struve_h1_zj(k * r1)) / 2;
        acc += -b * (bz2 - bz1) / k / k / k;
    }
    return acc;
}

/* Trapezoid sum (src/FJcpu.cpp:348-357 and :386-394).  The J basis calls the
 * C library's j0; the Y basis calls the NR Y0 form.  The J variant scales by
 * 0.5 and the Y variant divides by 2, which are the same in double. */
static inline __attribute__((always_inline))
float trapz_sum(const float *U, const float *r, int nr, int nf, int i,
                float k, bessel_fn z0)
{
    float acc = 0.0f;
    for (int ir = 1; ir < nr; ir++) {
        float g1 = U[i + (ir - 1) * nf];
        float g2 = U[i + ir * nf];
        float r1 = r[ir - 1];
        float r2 = r[ir];
        float dr0 = fmaxf((r2 - r1), 0.1);
        acc += (g1 * z0(k * r1) * r1 + g2 * z0(k * r2) * r2) * dr0 * 0.5;
    }
    return acc;
}

static double glibc_j0(double x) { return j0(x); }

/* One (f,c) output.  The four basis/itype combinations are spelled out so
 * that, after inlining, every special-function call is a direct call, as in
 * the reference's four separate loops. */
static float one_output(const float *U, const float *r, int nr, int nf, int i,
                        float k, int itype, int ybasis)
{
    if (itype == 1)
        return ybasis ? linseg_sum(U, r, nr, nf, i, k, bessel_y0_nr, bessel_y1_nr)
                      : linseg_sum(U, r, nr, nf, i, k, bessel_j0_nr, bessel_j1_nr);
    return ybasis ? trapz_sum(U, r, nr, nf, i, k, bessel_y0_nr)
                  : trapz_sum(U, r, nr, nf, i, k, glibc_j0);
}

/* Frequency columns are independent; OpenMP splits them across threads as
 * the reference does (src/FJcpu.cpp:236-243).  `fidx` (may be NULL) selects
 * a subset of frequency columns; out then has shape [nc][nfi] and column q
 * holds frequency fidx[q].  With fidx == NULL, nfi must equal nf. */
static int run(const float *U, const float *r, const float *f, float *out,
               const float *c, int nc, int nr, int nf, int itype, int ybasis,
               int nthreads, const int *fidx, int nfi)
{
    if (itype != 0 && itype != 1) return 0;  /* reference: silent no-op */
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int q = 0; q < nfi; q++) {
        int i = fidx ? fidx[q] : q;
        for (int j = 0; j < nc; j++) {
            float k = 2 * REF_PI * f[i] / c[j];
            out[q + (int64_t)j * nfi] = one_output(U, r, nr, nf, i, k, itype, ybasis);
        }
    }
    return 0;
}

/* Same argument order and meaning as FJ/FH in include/FJcpu.h:19-20. */
int fjport_FJ(const float *u_f, const float *r, const float *f, float *out,
              const float *c, int nc, int nr, int nf, int type, int nThread)
{
    return run(u_f, r, f, out, c, nc, nr, nf, type, 0, nThread, 0, nf);
}

int fjport_FH(const float *u_f, const float *r, const float *f, float *out,
              const float *c, int nc, int nr, int nf, int type, int nThread)
{
    return run(u_f, r, f, out, c, nc, nr, nf, type, 1, nThread, 0, nf);
}

/* Column-subset variant used to time a bounded sample of a large workload. */
int fjport_columns(const float *u_f, const float *r, const float *f, float *out,
                   const float *c, int nc, int nr, int nf, int type, int ybasis,
                   int nThread, const int *fidx, int nfi)
{
    return run(u_f, r, f, out, c, nc, nr, nf, type, ybasis, nThread, fidx, nfi);
}

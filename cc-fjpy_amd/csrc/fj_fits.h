/* Generated by gen_fits.py -- do not edit by hand. */
#pragma once
#ifndef FJ_HD
#define FJ_HD
#endif
/* double tables: __constant__ on the device, plain on the host */
enum { FIT_LITERAL = 0, FIT_TABLE = 1, FIT_SCALAR = 2 };
/* FJ_SCONST(v, lo, hi): the double v (bits hi:lo); on the device two
   s_mov_b32 into a scalar register pair where the value is used (a
   volatile asm, so it is neither hoisted nor kept live across a loop) */
#if defined(__HIP_DEVICE_COMPILE__)
template <unsigned LO, unsigned HI>
__device__ inline double fj_sconst() {
  unsigned lo, hi;
  asm volatile("s_mov_b32 %0, %2\n\ts_mov_b32 %1, %3" : "=s"(lo), "=s"(hi) : "n"(LO), "n"(HI));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
#define FJ_SCONST(v, lo, hi) fj_sconst<lo, hi>()
#else
#define FJ_SCONST(v, lo, hi) (v)
#endif
#if defined(__HIPCC__)
#define FJ_TABLE(name, n, ...) static __constant__ double name##_d[n] = __VA_ARGS__; static const double name##_h[n] = __VA_ARGS__;
#if defined(__HIP_DEVICE_COMPILE__)
#define FJ_TABLE_PTR(name) name##_d
#else
#define FJ_TABLE_PTR(name) name##_h
#endif
#else
#define FJ_TABLE(name, n, ...) static const double name##_h[n] = __VA_ARGS__;
#define FJ_TABLE_PTR(name) name##_h
#endif

/* large argument x >= 8, t = (8/x)^2, float */
/* max abs error of the fit on (0,1]: 8.30e-11 */
FJ_HD inline float fit_P0(float t) {
  float p = 2.0872274757632412e-07f;
  p = __builtin_fmaf(p, t, -2.072035746898442e-06f);
  p = __builtin_fmaf(p, t, 2.7344267112284302e-05f);
  p = __builtin_fmaf(p, t, -0.0010986284596301117f);
  p = __builtin_fmaf(p, t, 0.9999999999157112f);
  return p;
}
/* max abs error of the fit on (0,1]: 4.50e-11 */
FJ_HD inline float fit_q0(float t) {
  float p = -9.308052403636231e-08f;
  p = __builtin_fmaf(p, t, 7.612796799996811e-07f);
  p = __builtin_fmaf(p, t, -6.910627331787113e-06f);
  p = __builtin_fmaf(p, t, 0.00014304877194255163f);
  p = __builtin_fmaf(p, t, -0.015624999954253075f);
  return p;
}
/* max abs error of the fit on (0,1]: 9.30e-11 */
FJ_HD inline float fit_P1(float t) {
  float p = -2.396020151809229e-07f;
  p = __builtin_fmaf(p, t, 2.4560468976503757e-06f);
  p = __builtin_fmaf(p, t, -3.516304122374271e-05f);
  p = __builtin_fmaf(p, t, 0.001831049815192328f);
  p = __builtin_fmaf(p, t, 1.000000000094397f);
  return p;
}
/* max abs error of the fit on (0,1]: 4.99e-11 */
FJ_HD inline float fit_q1(float t) {
  float p = 1.0533367480691048e-07f;
  p = __builtin_fmaf(p, t, -8.813803538404306e-07f);
  p = __builtin_fmaf(p, t, 8.448628806224969e-06f);
  p = __builtin_fmaf(p, t, -0.00020026897267309463f);
  p = __builtin_fmaf(p, t, 0.04687499994925341f);
  return p;
}
/* max abs error of the fit on (0,1]: 9.01e-10 */
FJ_HD inline float fit_R0(float t) {
  float p = -2.820982558380003e-05f;
  p = __builtin_fmaf(p, t, 0.0001375252158385671f);
  p = __builtin_fmaf(p, t, -0.0003148555691661709f);
  p = __builtin_fmaf(p, t, 0.0005059846072811284f);
  p = __builtin_fmaf(p, t, -0.0008291186708186722f);
  p = __builtin_fmaf(p, t, 0.0021942019537319547f);
  p = __builtin_fmaf(p, t, -0.015624870399185526f);
  p = __builtin_fmaf(p, t, 0.9999999990698373f);
  return p;
}
/* max abs error of the fit on (0,1]: 5.10e-10 */
FJ_HD inline float fit_R1(float t) {
  float p = -5.35064546265832e-06f;
  p = __builtin_fmaf(p, t, 2.603089300202931e-05f);
  p = __builtin_fmaf(p, t, -6.712250189155052e-05f);
  p = __builtin_fmaf(p, t, 0.00016461168687600765f);
  p = __builtin_fmaf(p, t, -0.0007314297629187565f);
  p = __builtin_fmaf(p, t, 0.015624944443527137f);
  p = __builtin_fmaf(p, t, 1.0000000005252478f);
  return p;
}

/* far argument x >= 20, t = (8/x)^2 <= 0.16, float */
constexpr float kFarX = 20.0f;
/* max abs error of the fit on (0,0.16]: 2.66e-10 */
FJ_HD inline float fitf_P0(float t) {
  float p = 2.6872226653787023e-05f;
  p = __builtin_fmaf(p, t, -0.00109860259568224f);
  p = __builtin_fmaf(p, t, 0.999999999732733f);
  return p;
}
/* max abs error of the fit on (0,0.16]: 9.89e-11 */
FJ_HD inline float fitf_q0(float t) {
  float p = -6.740746258951341e-06f;
  p = __builtin_fmaf(p, t, 0.000143039894382146f);
  p = __builtin_fmaf(p, t, -0.015624999900625058f);
  return p;
}
/* max abs error of the fit on (0,0.16]: 3.15e-10 */
FJ_HD inline float fitf_P1(float t) {
  float p = -3.4602122207834955e-05f;
  p = __builtin_fmaf(p, t, 0.0018310189187639315f);
  p = __builtin_fmaf(p, t, 1.0000000003164382f);
  return p;
}
/* max abs error of the fit on (0,0.16]: 1.14e-10 */
FJ_HD inline float fitf_q1(float t) {
  float p = 8.251415608533157e-06f;
  p = __builtin_fmaf(p, t, -0.00020025860089406176f);
  p = __builtin_fmaf(p, t, 0.04687499988512577f);
  return p;
}
/* max abs error of the fit on (0,0.16]: 2.27e-09 */
FJ_HD inline float fitf_R0(float t) {
  float p = -0.0006925691349526906f;
  p = __builtin_fmaf(p, t, 0.0021819717373648033f);
  p = __builtin_fmaf(p, t, -0.015624529368443808f);
  p = __builtin_fmaf(p, t, 0.999999997700073f);
  return p;
}
/* max abs error of the fit on (0,0.16]: 3.52e-10 */
FJ_HD inline float fitf_R1(float t) {
  float p = 0.0001468112808747115f;
  p = __builtin_fmaf(p, t, -0.0007300897303956774f);
  p = __builtin_fmaf(p, t, 0.015624927646223167f);
  p = __builtin_fmaf(p, t, 1.0000000003552583f);
  return p;
}

/* far argument, modulus and phase of the order-0 and order-1 functions:
   Mn(t), and the phase shift in revolutions over 8/x */
/* max abs error of the fit on (0,0.16]: 2.52e-10 */
FJ_HD inline float fitf_M0(float t) {
  float p = 2.47902815240584e-05f;
  p = __builtin_fmaf(p, t, -0.0009765338659266368f);
  p = __builtin_fmaf(p, t, 0.9999999997467515f);
  return p;
}
/* max abs error of the fit on (0,0.16]: 3.17e-09 */
FJ_HD inline float fitf_ph0(float t) {
  float p = 2.0077624210811506e-05f;
  p = __builtin_fmaf(p, t, -0.002486792805295915f);
  return p;
}
/* max abs error of the fit on (0,0.16]: 3.76e-10 */
FJ_HD inline float fitf_M1(float t) {
  float p = -4.648938447752519e-05f;
  p = __builtin_fmaf(p, t, 0.002929644837346829f);
  p = __builtin_fmaf(p, t, 1.0000000003775265f);
  return p;
}
/* max abs error of the fit on (0,0.16]: 5.63e-09 */
FJ_HD inline float fitf_ph1(float t) {
  float p = -5.071470336918935e-05f;
  p = __builtin_fmaf(p, t, 0.007460382304125815f);
  return p;
}

/* x >= 8 (t in (0, 1]), the F-J lane's far forms: modulus and phase, and P, Q */
/* max abs error of the fit on (0,1]: 4.96e-08 */
FJ_HD inline float fitm_M0(float t) {
  float p = 2.2643268583106974e-05f;
  p = __builtin_fmaf(p, t, -0.0009756294194968865f);
  p = __builtin_fmaf(p, t, 0.9999999495300606f);
  return p;
}
/* max abs error of the fit on (0,1]: 2.75e-09 */
FJ_HD inline float fitm_ph0(float t) {
  float p = -8.67582837049008e-07f;
  p = __builtin_fmaf(p, t, 2.018513326570912e-05f);
  p = __builtin_fmaf(p, t, -0.0024867931705683244f);
  return p;
}
/* max abs error of the fit on (0,1]: 7.57e-08 */
FJ_HD inline float fitm_M1(float t) {
  float p = -4.323915890770389e-05f;
  p = __builtin_fmaf(p, t, 0.0029282710195231427f);
  p = __builtin_fmaf(p, t, 1.0000000768364674f);
  return p;
}
/* max abs error of the fit on (0,1]: 4.11e-09 */
FJ_HD inline float fitm_ph1(float t) {
  float p = 1.5796516822635373e-06f;
  p = __builtin_fmaf(p, t, -5.09207762010708e-05f);
  p = __builtin_fmaf(p, t, 0.00746038375925628f);
  return p;
}
/* max abs error of the fit on (0,1]: 5.25e-08 */
FJ_HD inline float fitm_P0(float t) {
  float p = 2.4601478300604937e-05f;
  p = __builtin_fmaf(p, t, -0.00109764557192691f);
  p = __builtin_fmaf(p, t, 0.9999999465791198f);
  return p;
}
/* max abs error of the fit on (0,1]: 1.83e-08 */
FJ_HD inline float fitm_q0(float t) {
  float p = -5.931598728855359e-06f;
  p = __builtin_fmaf(p, t, 0.0001427019975810866f);
  p = __builtin_fmaf(p, t, -0.015624981254604237f);
  return p;
}
/* max abs error of the fit on (0,1]: 6.27e-08 */
FJ_HD inline float fitm_P1(float t) {
  float p = -3.1898274403362295e-05f;
  p = __builtin_fmaf(p, t, 0.0018298779405753972f);
  p = __builtin_fmaf(p, t, 1.0000000637426278f);
  return p;
}
/* max abs error of the fit on (0,1]: 2.14e-08 */
FJ_HD inline float fitm_q1(float t) {
  float p = 7.310892206382535e-06f;
  p = __builtin_fmaf(p, t, -0.00019986536318951833f);
  p = __builtin_fmaf(p, t, 0.04687497816655271f);
  return p;
}

/* small argument x < 8, y = (x/8)^2, double */
/* max abs error of the fit on (0,1]: 8.04e-11 */
FJ_TABLE(k_fit_J0s, 12, {-0.005618801482900782, 0.07079209695405728, -0.5037809434408522, 2.625220559644334, -10.557357884793742, 32.35918804582165, -72.81660855664654, 113.77757421102329, -113.77775682787487, 63.99999888465921, -15.999999976722172, 0.9999999999195268})
template <int TAB = FIT_LITERAL>
FJ_HD inline double fit_J0s(double y) {
  if (TAB == FIT_TABLE) {
    const double* c = FJ_TABLE_PTR(k_fit_J0s);
    double p = c[0];
    p = __builtin_fma(p, y, c[1]);
    p = __builtin_fma(p, y, c[2]);
    p = __builtin_fma(p, y, c[3]);
    p = __builtin_fma(p, y, c[4]);
    p = __builtin_fma(p, y, c[5]);
    p = __builtin_fma(p, y, c[6]);
    p = __builtin_fma(p, y, c[7]);
    p = __builtin_fma(p, y, c[8]);
    p = __builtin_fma(p, y, c[9]);
    p = __builtin_fma(p, y, c[10]);
    p = __builtin_fma(p, y, c[11]);
    return p;
  }
  if (TAB == FIT_SCALAR) {
    double p = FJ_SCONST(-0.005618801482900782, 0x89c9d4c1u, 0xbf7703bdu);
    p = __builtin_fma(p, y, FJ_SCONST(0.07079209695405728, 0x4d3ba1c6u, 0x3fb21f6eu));
    p = __builtin_fma(p, y, FJ_SCONST(-0.5037809434408522, 0x368da5c6u, 0xbfe01ef9u));
    p = __builtin_fma(p, y, FJ_SCONST(2.625220559644334, 0xa303ac82u, 0x40050073u));
    p = __builtin_fma(p, y, FJ_SCONST(-10.557357884793742, 0x033eb6b6u, 0xc0251d5eu));
    p = __builtin_fma(p, y, FJ_SCONST(32.35918804582165, 0xdfb6f585u, 0x40402df9u));
    p = __builtin_fma(p, y, FJ_SCONST(-72.81660855664654, 0x50891b90u, 0xc0523443u));
    p = __builtin_fma(p, y, FJ_SCONST(113.77757421102329, 0xc69fa3b7u, 0x405c71c3u));
    p = __builtin_fma(p, y, FJ_SCONST(-113.77775682787487, 0xc492fda6u, 0xc05c71c6u));
    p = __builtin_fma(p, y, FJ_SCONST(63.99999888465921, 0xf6a4d2ebu, 0x404fffffu));
    p = __builtin_fma(p, y, FJ_SCONST(-15.999999976722172, 0xff380b84u, 0xc02fffffu));
    p = __builtin_fma(p, y, FJ_SCONST(0.9999999999195268, 0xfff4f09au, 0x3fefffffu));
    return p;
  }
  double p = -0.005618801482900782;
  p = __builtin_fma(p, y, 0.07079209695405728);
  p = __builtin_fma(p, y, -0.5037809434408522);
  p = __builtin_fma(p, y, 2.625220559644334);
  p = __builtin_fma(p, y, -10.557357884793742);
  p = __builtin_fma(p, y, 32.35918804582165);
  p = __builtin_fma(p, y, -72.81660855664654);
  p = __builtin_fma(p, y, 113.77757421102329);
  p = __builtin_fma(p, y, -113.77775682787487);
  p = __builtin_fma(p, y, 63.99999888465921);
  p = __builtin_fma(p, y, -15.999999976722172);
  p = __builtin_fma(p, y, 0.9999999999195268);
  return p;
}
/* max abs error of the fit on (0,1]: 1.21e-10 */
FJ_TABLE(k_fit_J1s, 11, {0.001933119992191311, -0.022146255813530545, 0.1417755820614683, -0.6564575115820559, 2.309572547363542, -6.06743715646006, 11.377627299279547, -14.222203592242003, 10.66666548155003, -3.9999999705566402, 0.4999999998789634})
template <int TAB = FIT_LITERAL>
FJ_HD inline double fit_J1s(double y) {
  if (TAB == FIT_TABLE) {
    const double* c = FJ_TABLE_PTR(k_fit_J1s);
    double p = c[0];
    p = __builtin_fma(p, y, c[1]);
    p = __builtin_fma(p, y, c[2]);
    p = __builtin_fma(p, y, c[3]);
    p = __builtin_fma(p, y, c[4]);
    p = __builtin_fma(p, y, c[5]);
    p = __builtin_fma(p, y, c[6]);
    p = __builtin_fma(p, y, c[7]);
    p = __builtin_fma(p, y, c[8]);
    p = __builtin_fma(p, y, c[9]);
    p = __builtin_fma(p, y, c[10]);
    return p;
  }
  if (TAB == FIT_SCALAR) {
    double p = FJ_SCONST(0.001933119992191311, 0xc9533cfau, 0x3f5fac17u);
    p = __builtin_fma(p, y, FJ_SCONST(-0.022146255813530545, 0x11cab6f6u, 0xbf96ad82u));
    p = __builtin_fma(p, y, FJ_SCONST(0.1417755820614683, 0xc829a5bfu, 0x3fc225b3u));
    p = __builtin_fma(p, y, FJ_SCONST(-0.6564575115820559, 0x2eeeabccu, 0xbfe501b3u));
    p = __builtin_fma(p, y, FJ_SCONST(2.309572547363542, 0x2bf5539bu, 0x40027a01u));
    p = __builtin_fma(p, y, FJ_SCONST(-6.06743715646006, 0x3ef61ff1u, 0xc018450eu));
    p = __builtin_fma(p, y, FJ_SCONST(11.377627299279547, 0x5d88f705u, 0x4026c158u));
    p = __builtin_fma(p, y, FJ_SCONST(-14.222203592242003, 0xab53775cu, 0xc02c71c4u));
    p = __builtin_fma(p, y, FJ_SCONST(10.66666548155003, 0x2d91424au, 0x40255555u));
    p = __builtin_fma(p, y, FJ_SCONST(-3.9999999705566402, 0xfc0c5578u, 0xc00fffffu));
    p = __builtin_fma(p, y, FJ_SCONST(0.4999999998789634, 0xffdebaceu, 0x3fdfffffu));
    return p;
  }
  double p = 0.001933119992191311;
  p = __builtin_fma(p, y, -0.022146255813530545);
  p = __builtin_fma(p, y, 0.1417755820614683);
  p = __builtin_fma(p, y, -0.6564575115820559);
  p = __builtin_fma(p, y, 2.309572547363542);
  p = __builtin_fma(p, y, -6.06743715646006);
  p = __builtin_fma(p, y, 11.377627299279547);
  p = __builtin_fma(p, y, -14.222203592242003);
  p = __builtin_fma(p, y, 10.66666548155003);
  p = __builtin_fma(p, y, -3.9999999705566402);
  p = __builtin_fma(p, y, 0.4999999998789634);
  return p;
}
/* max abs error of the fit on (0,1]: 1.62e-10 */
FJ_TABLE(k_fit_Y0s, 12, {0.011007520614961696, -0.13607675855273377, 0.9417493466708342, -4.732348709939299, 18.20275922795894, -52.85781250766169, 111.22092650701482, -159.2993419475141, 141.1914151817568, -65.8389707815695, 11.366785032612231, -0.07380429494619678})
template <int TAB = FIT_LITERAL>
FJ_HD inline double fit_Y0s(double y) {
  if (TAB == FIT_TABLE) {
    const double* c = FJ_TABLE_PTR(k_fit_Y0s);
    double p = c[0];
    p = __builtin_fma(p, y, c[1]);
    p = __builtin_fma(p, y, c[2]);
    p = __builtin_fma(p, y, c[3]);
    p = __builtin_fma(p, y, c[4]);
    p = __builtin_fma(p, y, c[5]);
    p = __builtin_fma(p, y, c[6]);
    p = __builtin_fma(p, y, c[7]);
    p = __builtin_fma(p, y, c[8]);
    p = __builtin_fma(p, y, c[9]);
    p = __builtin_fma(p, y, c[10]);
    p = __builtin_fma(p, y, c[11]);
    return p;
  }
  if (TAB == FIT_SCALAR) {
    double p = FJ_SCONST(0.011007520614961696, 0x68691346u, 0x3f868b1cu);
    p = __builtin_fma(p, y, FJ_SCONST(-0.13607675855273377, 0x95dd6626u, 0xbfc16af6u));
    p = __builtin_fma(p, y, FJ_SCONST(0.9417493466708342, 0x869f6112u, 0x3fee22cfu));
    p = __builtin_fma(p, y, FJ_SCONST(-4.732348709939299, 0xd1f9d40du, 0xc012edecu));
    p = __builtin_fma(p, y, FJ_SCONST(18.20275922795894, 0x075d0bbdu, 0x403233e8u));
    p = __builtin_fma(p, y, FJ_SCONST(-52.85781250766169, 0xccdd40dcu, 0xc04a6dccu));
    p = __builtin_fma(p, y, FJ_SCONST(111.22092650701482, 0xa8ee9cafu, 0x405bce23u));
    p = __builtin_fma(p, y, FJ_SCONST(-159.2993419475141, 0x35905c9cu, 0xc063e994u));
    p = __builtin_fma(p, y, FJ_SCONST(141.1914151817568, 0x12bb334fu, 0x4061a620u));
    p = __builtin_fma(p, y, FJ_SCONST(-65.8389707815695, 0xb28148ffu, 0xc05075b1u));
    p = __builtin_fma(p, y, FJ_SCONST(11.366785032612231, 0x3f6f76b2u, 0x4026bbcbu));
    p = __builtin_fma(p, y, FJ_SCONST(-0.07380429494619678, 0x99192720u, 0xbfb2e4d6u));
    return p;
  }
  double p = 0.011007520614961696;
  p = __builtin_fma(p, y, -0.13607675855273377);
  p = __builtin_fma(p, y, 0.9417493466708342);
  p = __builtin_fma(p, y, -4.732348709939299);
  p = __builtin_fma(p, y, 18.20275922795894);
  p = __builtin_fma(p, y, -52.85781250766169);
  p = __builtin_fma(p, y, 111.22092650701482);
  p = __builtin_fma(p, y, -159.2993419475141);
  p = __builtin_fma(p, y, 141.1914151817568);
  p = __builtin_fma(p, y, -65.8389707815695);
  p = __builtin_fma(p, y, 11.366785032612231);
  p = __builtin_fma(p, y, -0.07380429494619678);
  return p;
}
/* max abs error of the fit on (0,1]: 2.41e-10 */
FJ_TABLE(k_fit_Y1s, 11, {-0.0037279760680778515, 0.04184985392300427, -0.2599904328974209, 1.1572271941470782, -3.8770978479644107, 9.589115060130059, -16.654007220047983, 18.780663548716326, -12.104928295564273, 3.478315983625302, -0.19605709040518532})
template <int TAB = FIT_LITERAL>
FJ_HD inline double fit_Y1s(double y) {
  if (TAB == FIT_TABLE) {
    const double* c = FJ_TABLE_PTR(k_fit_Y1s);
    double p = c[0];
    p = __builtin_fma(p, y, c[1]);
    p = __builtin_fma(p, y, c[2]);
    p = __builtin_fma(p, y, c[3]);
    p = __builtin_fma(p, y, c[4]);
    p = __builtin_fma(p, y, c[5]);
    p = __builtin_fma(p, y, c[6]);
    p = __builtin_fma(p, y, c[7]);
    p = __builtin_fma(p, y, c[8]);
    p = __builtin_fma(p, y, c[9]);
    p = __builtin_fma(p, y, c[10]);
    return p;
  }
  if (TAB == FIT_SCALAR) {
    double p = FJ_SCONST(-0.0037279760680778515, 0xe95d8314u, 0xbf6e8a21u);
    p = __builtin_fma(p, y, FJ_SCONST(0.04184985392300427, 0x13e223f7u, 0x3fa56d58u));
    p = __builtin_fma(p, y, FJ_SCONST(-0.2599904328974209, 0xe9a44ebbu, 0xbfd0a3aeu));
    p = __builtin_fma(p, y, FJ_SCONST(1.1572271941470782, 0xa98e74eau, 0x3ff28400u));
    p = __builtin_fma(p, y, FJ_SCONST(-3.8770978479644107, 0xe0633168u, 0xc00f044bu));
    p = __builtin_fma(p, y, FJ_SCONST(9.589115060130059, 0x7d39adeau, 0x40232da0u));
    p = __builtin_fma(p, y, FJ_SCONST(-16.654007220047983, 0x04657437u, 0xc030a76du));
    p = __builtin_fma(p, y, FJ_SCONST(18.780663548716326, 0x90faea7au, 0x4032c7d9u));
    p = __builtin_fma(p, y, FJ_SCONST(-12.104928295564273, 0x295bbf45u, 0xc02835b9u));
    p = __builtin_fma(p, y, FJ_SCONST(3.478315983625302, 0x54969913u, 0x400bd397u));
    p = __builtin_fma(p, y, FJ_SCONST(-0.19605709040518532, 0x13b83742u, 0xbfc91866u));
    return p;
  }
  double p = -0.0037279760680778515;
  p = __builtin_fma(p, y, 0.04184985392300427);
  p = __builtin_fma(p, y, -0.2599904328974209);
  p = __builtin_fma(p, y, 1.1572271941470782);
  p = __builtin_fma(p, y, -3.8770978479644107);
  p = __builtin_fma(p, y, 9.589115060130059);
  p = __builtin_fma(p, y, -16.654007220047983);
  p = __builtin_fma(p, y, 18.780663548716326);
  p = __builtin_fma(p, y, -12.104928295564273);
  p = __builtin_fma(p, y, 3.478315983625302);
  p = __builtin_fma(p, y, -0.19605709040518532);
  return p;
}
/* max abs error of the fit on (0,1]: 5.06e-12 */
FJ_TABLE(k_fit_GJs, 11, {8.887397533409992e-05, -0.0010796865524819253, 0.007517728519888622, -0.03868451112092575, 0.1540241037174691, -0.4667509709541983, 1.034337151432889, -1.5802461351138604, 1.5238094742576516, -0.7999999987682918, 0.16666666666160115})
template <int TAB = FIT_LITERAL>
FJ_HD inline double fit_GJs(double y) {
  if (TAB == FIT_TABLE) {
    const double* c = FJ_TABLE_PTR(k_fit_GJs);
    double p = c[0];
    p = __builtin_fma(p, y, c[1]);
    p = __builtin_fma(p, y, c[2]);
    p = __builtin_fma(p, y, c[3]);
    p = __builtin_fma(p, y, c[4]);
    p = __builtin_fma(p, y, c[5]);
    p = __builtin_fma(p, y, c[6]);
    p = __builtin_fma(p, y, c[7]);
    p = __builtin_fma(p, y, c[8]);
    p = __builtin_fma(p, y, c[9]);
    p = __builtin_fma(p, y, c[10]);
    return p;
  }
  if (TAB == FIT_SCALAR) {
    double p = FJ_SCONST(8.887397533409992e-05, 0x45256592u, 0x3f174c3bu);
    p = __builtin_fma(p, y, FJ_SCONST(-0.0010796865524819253, 0x9bb3aa23u, 0xbf51b088u));
    p = __builtin_fma(p, y, FJ_SCONST(0.007517728519888622, 0xe2214af3u, 0x3f7ecae8u));
    p = __builtin_fma(p, y, FJ_SCONST(-0.03868451112092575, 0xcc409361u, 0xbfa3ce74u));
    p = __builtin_fma(p, y, FJ_SCONST(0.1540241037174691, 0xd4219124u, 0x3fc3b70fu));
    p = __builtin_fma(p, y, FJ_SCONST(-0.4667509709541983, 0x76e7f843u, 0xbfdddf3fu));
    p = __builtin_fma(p, y, FJ_SCONST(1.034337151432889, 0x1ce712abu, 0x3ff08ca5u));
    p = __builtin_fma(p, y, FJ_SCONST(-1.5802461351138604, 0x2bdf1c60u, 0xbff948b0u));
    p = __builtin_fma(p, y, FJ_SCONST(1.5238094742576516, 0x0b145857u, 0x3ff86186u));
    p = __builtin_fma(p, y, FJ_SCONST(-0.7999999987682918, 0x98f050b8u, 0xbfe99999u));
    p = __builtin_fma(p, y, FJ_SCONST(0.16666666666160115, 0x55528c6du, 0x3fc55555u));
    return p;
  }
  double p = 8.887397533409992e-05;
  p = __builtin_fma(p, y, -0.0010796865524819253);
  p = __builtin_fma(p, y, 0.007517728519888622);
  p = __builtin_fma(p, y, -0.03868451112092575);
  p = __builtin_fma(p, y, 0.1540241037174691);
  p = __builtin_fma(p, y, -0.4667509709541983);
  p = __builtin_fma(p, y, 1.034337151432889);
  p = __builtin_fma(p, y, -1.5802461351138604);
  p = __builtin_fma(p, y, 1.5238094742576516);
  p = __builtin_fma(p, y, -0.7999999987682918);
  p = __builtin_fma(p, y, 0.16666666666160115);
  return p;
}
/* max abs error of the fit on (0,1]: 1.55e-10 */
FJ_TABLE(k_fit_GYs, 12, {-0.010493016794276316, 0.12932686955425834, -0.8908749111050289, 4.447592307093985, -16.958832589184595, 48.66961558750545, -100.72670076329833, 140.7051499670384, -119.54298053751586, 51.0414296530347, -6.446088190525072, -0.6366197725231723})
template <int TAB = FIT_LITERAL>
FJ_HD inline double fit_GYs(double y) {
  if (TAB == FIT_TABLE) {
    const double* c = FJ_TABLE_PTR(k_fit_GYs);
    double p = c[0];
    p = __builtin_fma(p, y, c[1]);
    p = __builtin_fma(p, y, c[2]);
    p = __builtin_fma(p, y, c[3]);
    p = __builtin_fma(p, y, c[4]);
    p = __builtin_fma(p, y, c[5]);
    p = __builtin_fma(p, y, c[6]);
    p = __builtin_fma(p, y, c[7]);
    p = __builtin_fma(p, y, c[8]);
    p = __builtin_fma(p, y, c[9]);
    p = __builtin_fma(p, y, c[10]);
    p = __builtin_fma(p, y, c[11]);
    return p;
  }
  if (TAB == FIT_SCALAR) {
    double p = FJ_SCONST(-0.010493016794276316, 0xdfbe0b96u, 0xbf857d5cu);
    p = __builtin_fma(p, y, FJ_SCONST(0.12932686955425834, 0x699d6375u, 0x3fc08dc8u));
    p = __builtin_fma(p, y, FJ_SCONST(-0.8908749111050289, 0x1a00bc78u, 0xbfec820cu));
    p = __builtin_fma(p, y, FJ_SCONST(4.447592307093985, 0xa343a3b1u, 0x4011ca55u));
    p = __builtin_fma(p, y, FJ_SCONST(-16.958832589184595, 0x0d74e308u, 0xc030f576u));
    p = __builtin_fma(p, y, FJ_SCONST(48.66961558750545, 0xf6ac9d26u, 0x404855b5u));
    p = __builtin_fma(p, y, FJ_SCONST(-100.72670076329833, 0x43eb160du, 0xc0592e82u));
    p = __builtin_fma(p, y, FJ_SCONST(140.7051499670384, 0x96a9e693u, 0x40619690u));
    p = __builtin_fma(p, y, FJ_SCONST(-119.54298053751586, 0x3170bfb0u, 0xc05de2c0u));
    p = __builtin_fma(p, y, FJ_SCONST(51.0414296530347, 0x911e6f30u, 0x4049854du));
    p = __builtin_fma(p, y, FJ_SCONST(-6.446088190525072, 0x57b5bf7du, 0xc019c8cbu));
    p = __builtin_fma(p, y, FJ_SCONST(-0.6366197725231723, 0x6ddf2ae1u, 0xbfe45f30u));
    return p;
  }
  double p = -0.010493016794276316;
  p = __builtin_fma(p, y, 0.12932686955425834);
  p = __builtin_fma(p, y, -0.8908749111050289);
  p = __builtin_fma(p, y, 4.447592307093985);
  p = __builtin_fma(p, y, -16.958832589184595);
  p = __builtin_fma(p, y, 48.66961558750545);
  p = __builtin_fma(p, y, -100.72670076329833);
  p = __builtin_fma(p, y, 140.7051499670384);
  p = __builtin_fma(p, y, -119.54298053751586);
  p = __builtin_fma(p, y, 51.0414296530347);
  p = __builtin_fma(p, y, -6.446088190525072);
  p = __builtin_fma(p, y, -0.6366197725231723);
  return p;
}

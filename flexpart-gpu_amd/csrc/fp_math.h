/*
 * fp_math.h - deterministic single-precision transcendentals for the
 * per-particle timestep.
 *
 * The particle step calls exp, log, sin/cos and pow of f32 arguments
 * (Hanna profiles, Box-Muller, Ornstein-Uhlenbeck decay, deposition).
 * Platform libraries (glibc on the host, the ROCm device library on the
 * GPU) round these differently in the last bit, which makes two runs of the
 * same trajectory drift apart through the Hanna regime switches and the PBL
 * reflection.  Every function here is written only with IEEE-754 operations
 * that are correctly rounded on both x86-64 and gfx950 (+, -, *, /, sqrt,
 * fused multiply-add, rint, floor, integer bit manipulation), so the same
 * source gives bit-identical results on the CPU and on the GPU.  The file
 * compiles as C99 (gcc) and as HIP device/host code (hipcc).
 *
 * Both sides must be compiled with floating-point contraction disabled
 * (-ffp-contract=off); every fused multiply-add below is written out
 * explicitly with fmaf/fma.
 *
 * Accuracy (measured against glibc over dense sweeps in the CPU tests):
 *   fpm_expf      <= 1 ulp         (Cody-Waite reduction, degree-7 Taylor)
 *   fpm_logf      <= 1 ulp         (atanh series in s = f/(2+f), 4 terms)
 *   fpm_sincosf   <= 1 ulp on |x| <= 2*pi, the Box-Muller angle range
 *                 (3-part Cody-Waite reduction by pi/2, Taylor kernels)
 *   fpm_powf      correctly rounded except within ~1e-8 ulp of a midpoint
 *                 (exp(y*log(x)) evaluated in binary64 with table-driven
 *                 reductions, rounded once)
 */
#ifndef FP_MATH_H
#define FP_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define FPM_FN __host__ __device__ static inline
#else
#define FPM_FN static inline
#endif

#define FPM_INF __builtin_inff()
#define FPM_NAN __builtin_nanf("")

FPM_FN uint32_t fpm_f2u(float x) { uint32_t u; __builtin_memcpy(&u, &x, 4); return u; }
FPM_FN float fpm_u2f(uint32_t u) { float x; __builtin_memcpy(&x, &u, 4); return x; }
FPM_FN uint64_t fpm_d2u(double x) { uint64_t u; __builtin_memcpy(&u, &x, 8); return u; }
FPM_FN double fpm_u2d(uint64_t u) { double x; __builtin_memcpy(&x, &u, 8); return x; }

/* n / d, correctly rounded, for operands that need no range scaling: n zero
 * or 2^-100 <= |n| <= 2^100, 2^-100 <= |d| <= 2^100, and |n / d| normal.  On
 * the GPU this is the IEEE division sequence the compiler emits
 * (v_div_scale, v_rcp, Newton refinement of the reciprocal, two corrections of
 * the quotient, v_div_fmas, v_div_fixup) without the steps that only act
 * outside that range: v_div_scale returns its operand unchanged with the
 * scale flag clear, so v_div_fmas is a plain fused multiply-add, and
 * v_div_fixup passes a finite normal quotient through.  Same result, three
 * instructions fewer.  On the CPU it is the division itself. */
FPM_FN float fpm_div_unscaled(float n, float d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float nd = -d;
    float r = __builtin_amdgcn_rcpf(d);
    const float e0 = __builtin_fmaf(nd, r, 1.0f);
    r = __builtin_fmaf(e0, r, r);
    float q = n * r;
    const float e1 = __builtin_fmaf(nd, q, n);
    q = __builtin_fmaf(e1, r, q);
    const float e2 = __builtin_fmaf(nd, q, n);
    return __builtin_fmaf(e2, r, q);
#else
    return n / d;
#endif
}

/* min(max(x, lo), hi) for lo <= hi.  On the GPU one median-of-three
 * (v_med3_f32), which gives the same value: for a quiet NaN x it returns lo,
 * as min(max(x, lo), hi) does.  (Callers here discard the clamped value of a
 * NaN argument, so a signaling NaN, for which the two differ, changes no
 * result either.) */
FPM_FN float fpm_clampf(float x, float lo, float hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(x, lo, hi);
#else
    return __builtin_fminf(__builtin_fmaxf(x, lo), hi);
#endif
}

/* fpm_div_unscaled that also returns the Newton-refined reciprocal of d it
 * multiplies by (about one unit in the last place of 1/d).  The quotient is
 * the same. */
FPM_FN float fpm_div_unscaled_rcp(float n, float d, float *rcp)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float nd = -d;
    float r = __builtin_amdgcn_rcpf(d);
    const float e0 = __builtin_fmaf(nd, r, 1.0f);
    r = __builtin_fmaf(e0, r, r);
    float q = n * r;
    const float e1 = __builtin_fmaf(nd, q, n);
    q = __builtin_fmaf(e1, r, q);
    const float e2 = __builtin_fmaf(nd, q, n);
    *rcp = r;
    return __builtin_fmaf(e2, r, q);
#else
    *rcp = 1.0f / d;
    return n / d;
#endif
}

/* e^x, single precision.  Special cases are selected at the end rather than
 * branched on, so a wavefront never splits inside the function. */
FPM_FN float fpm_expf_core(float x)
{
    const float xc = fpm_clampf(x, -104.0f, 89.0f);
    const float k = __builtin_rintf(xc * 1.44269502162933349609375f);
    /* ln2 = 0.693145751953125 (exact, 12 trailing zero bits) + 1.4286068e-6 */
    float r = __builtin_fmaf(k, -0.693145751953125f, xc);
    r = __builtin_fmaf(k, -1.428606765330187e-06f, r);
    /* e^r = 1 + r + r^2 * q(r), q from the Taylor series 1/2!, ..., 1/7! */
    float q = 1.98412701e-04f;
    q = __builtin_fmaf(q, r, 1.38888892e-03f);
    q = __builtin_fmaf(q, r, 8.33333377e-03f);
    q = __builtin_fmaf(q, r, 4.16666679e-02f);
    q = __builtin_fmaf(q, r, 1.66666672e-01f);
    q = __builtin_fmaf(q, r, 0.5f);
    const float p = __builtin_fmaf(q, r * r, r);
    /* ldexp rounds y * 2^k once, exactly like fpm_scale2; it overflows to
     * +inf for x in (88.72, 89]. */
    return __builtin_ldexpf(1.0f + p, (int)k);
}
FPM_FN float fpm_expf(float x)
{
    /* every x below -104 (and -inf) is evaluated at -104, whose result is 0
     * (e^-104 is below half the smallest subnormal); only NaN is selected */
    const float res = fpm_expf_core(x);
    return x != x ? x : res;
}
/* fpm_expf of an x known not to be NaN (the select leaves it unchanged). */
FPM_FN float fpm_expf_not_nan(float x) { return fpm_expf_core(x); }

/* Natural logarithm, single precision (special cases selected, not branched). */
FPM_FN float fpm_logf(float x)
{
    const uint32_t ix0 = fpm_f2u(x);
    const int sub = ix0 < 0x00800000u; /* +0 or subnormal: scale into the normal range */
    const uint32_t ix = sub ? fpm_f2u(x * 8388608.0f) : ix0;
    /* Split x = 2^e * m with m in [sqrt(1/2), sqrt(2)). */
    const uint32_t t = ix - 0x3f3504f3u;
    const int e = (((int32_t)t) >> 23) - (sub ? 23 : 0);
    const float m = fpm_u2f(ix - (t & 0xff800000u));
    const float f = m - 1.0f; /* exact (Sterbenz) */
    /* m in [sqrt(1/2), sqrt(2)) for every ix: f is 0 or 2^-24 <= |f| < 0.42,
     * 2 + f in (1.7, 2.5), so the quotient needs no scaling */
    const float s = fpm_div_unscaled(f, 2.0f + f);
    const float z = s * s;
    /* log(1+f) = f - hfsq + s*(hfsq + R), R = 2z/3 + 2z^2/5 + 2z^3/7 + 2z^4/9 */
    float R = 2.22222224e-01f;
    R = __builtin_fmaf(R, z, 2.85714298e-01f);
    R = __builtin_fmaf(R, z, 4.00000006e-01f);
    R = __builtin_fmaf(R, z, 6.66666687e-01f);
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float dk = (float)e;
    float res = dk * 0.693145751953125f -
                ((hfsq - (s * (hfsq + R) + dk * 1.428606765330187e-06f)) - f);
    res = x == FPM_INF ? x : res;
    res = x == 0.0f ? -FPM_INF : res;
    res = x < 0.0f ? FPM_NAN : res;
    return x != x ? x : res;
}

/* fpm_logf restricted to normal x in (0, 1]: the same arithmetic without the
 * subnormal rescaling and special-value selects, hence the same values on
 * that domain (the Box-Muller radius argument lies in [2^-126, 1)). */
FPM_FN float fpm_logf_unit(float x)
{
    const uint32_t ix = fpm_f2u(x);
    const uint32_t t = ix - 0x3f3504f3u;
    const int e = ((int32_t)t) >> 23;
    const float m = fpm_u2f(ix - (t & 0xff800000u));
    const float f = m - 1.0f;
    const float s = fpm_div_unscaled(f, 2.0f + f);
    const float z = s * s;
    float R = 2.22222224e-01f;
    R = __builtin_fmaf(R, z, 2.85714298e-01f);
    R = __builtin_fmaf(R, z, 4.00000006e-01f);
    R = __builtin_fmaf(R, z, 6.66666687e-01f);
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float dk = (float)e;
    return dk * 0.693145751953125f -
           ((hfsq - (s * (hfsq + R) + dk * 1.428606765330187e-06f)) - f);
}

/* sin(x) and cos(x), single precision; accurate on |x| <= 2*pi (the first
 * reduction step is exact there), usable with growing error up to |x|~1e4. */
FPM_FN void fpm_sincosf(float x, float *sin_out, float *cos_out)
{
    const float q = __builtin_rintf(x * 0.636619747f);
    /* pi/2 = 1.57079637 - 4.37113883e-08 - 1.71512451e-15 */
    float r = __builtin_fmaf(q, -1.57079637050628662109375f, x);
    r = __builtin_fmaf(q, 4.371138828673793e-08f, r);
    r = __builtin_fmaf(q, 1.7151245100058819e-15f, r);
    const float z = r * r;
    /* sin r = r + r^3 * (-1/3! + z/5! - z^2/7! + z^3/9!) */
    float ps = 2.75573188e-06f;
    ps = __builtin_fmaf(ps, z, -1.98412701e-04f);
    ps = __builtin_fmaf(ps, z, 8.33333377e-03f);
    ps = __builtin_fmaf(ps, z, -1.66666672e-01f);
    const float sr = __builtin_fmaf(r * z, ps, r);
    /* cos r = 1 + z * (-1/2 + z/4! - z^2/6! + z^3/8! - z^4/10!) */
    float pc = -2.75573200e-07f;
    pc = __builtin_fmaf(pc, z, 2.48015876e-05f);
    pc = __builtin_fmaf(pc, z, -1.38888892e-03f);
    pc = __builtin_fmaf(pc, z, 4.16666679e-02f);
    pc = __builtin_fmaf(pc, z, -0.5f);
    const float cr = __builtin_fmaf(z, pc, 1.0f);
    const int quadrant = ((int)q) & 3;
    const float s = (quadrant & 1) ? cr : sr; /* quadrants 1, 3 swap sin and cos */
    const float c = (quadrant & 1) ? sr : cr;
    *sin_out = (quadrant & 2) ? -s : s;       /* quadrants 2, 3 negate sin */
    *cos_out = ((quadrant + 1) & 2) ? -c : c; /* quadrants 1, 2 negate cos */
}

/* ---- binary64 kernels used by fpm_powf ------------------------------- */

/* Tables of the binary64 kernels (constants of the device code on the GPU).
 * FPM_LOG_TAB[2i] = 1/c_i rounded to binary64, c_i = 1 + (i + 1/2)/128, the
 * centre of the i-th 1/128 mantissa interval; FPM_LOG_TAB[2i+1] = -log of that
 * rounded reciprocal, to binary64 (so that log m = T_i + log(m * invc_i)
 * exactly); FPM_EXP2_J[j] = 2^(j/32) rounded to binary64.  Generated with
 * 60-digit decimal arithmetic.  The _tab variants below read the tables
 * through a pointer, so that a GPU kernel can pass a copy held in LDS; the
 * values, and so the results, are the same. */
#if defined(__cplusplus)
#define FPM_TABLE static constexpr
#else
#define FPM_TABLE static const
#endif
FPM_TABLE double FPM_LOG_TAB[256] = {
    0.9961089494163424, 0.003898640415657309, 0.9884169884169884, 0.01165061721997525,
    0.9808429118773946, 0.019342962843130987, 0.973384030418251, 0.026976587698202083,
    0.9660377358490566, 0.03455238150665973, 0.9588014981273408, 0.042071213920687044,
    0.9516728624535316, 0.049533935122276676, 0.9446494464944649, 0.05694137640013845,
    0.9377289377289377, 0.06429435070539725, 0.9309090909090909, 0.07159365318700882,
    0.924187725631769, 0.078840061707776, 0.9175627240143369, 0.08603433734180316,
    0.9110320284697508, 0.09317722485418334, 0.9045936395759717, 0.10026945316367517,
    0.8982456140350877, 0.10731173578908804, 0.89198606271777, 0.11430477128005863,
    0.8858131487889274, 0.12124924363286965, 0.8797250859106529, 0.12814582269193006,
    0.8737201365187713, 0.13499516453750482, 0.8677966101694915, 0.1417979118602574,
    0.8619528619528619, 0.1485546943231372, 0.8561872909698997, 0.15526612891112396,
    0.8504983388704319, 0.16193282026931324, 0.8448844884488449, 0.16855536102980664,
    0.839344262295082, 0.17513433212784915, 0.8338762214983714, 0.18167030310763463,
    0.8284789644012945, 0.18816383241818294, 0.8231511254019293, 0.19461546769967167,
    0.8178913738019169, 0.2010257460605908, 0.8126984126984127, 0.2073951943460706,
    0.807570977917981, 0.21372432939771818, 0.8025078369905956, 0.22001365830528213,
    0.7975077881619937, 0.2262636786504534, 0.7925696594427245, 0.232474878743094,
    0.7876923076923077, 0.238647737850175, 0.7828746177370031, 0.24478272641769092,
    0.7781155015197568, 0.25088030628580943, 0.7734138972809668, 0.2569409308975004,
    0.7687687687687688, 0.26296504550088134, 0.764179104477612, 0.26895308734550394,
    0.7596439169139466, 0.2749054858727992, 0.7551622418879056, 0.2808226629008878,
    0.750733137829912, 0.2867050328039543, 0.7463556851311953, 0.29255300268637746,
    0.7420289855072464, 0.2983669725517973, 0.7377521613832853, 0.3041473354672968,
    0.7335243553008596, 0.3098944777228647, 0.7293447293447294, 0.3156087789863033,
    0.7252124645892352, 0.32129061245373425, 0.7211267605633803, 0.3269403449958533,
    0.7170868347338936, 0.3325583373000766, 0.713091922005571, 0.3381449440087164,
    0.7091412742382271, 0.34370051385331846, 0.7052341597796143, 0.3492253897852883,
    0.7013698630136986, 0.354719909102929, 0.6975476839237057, 0.3601844035750078,
    0.6937669376693767, 0.3656191995609647, 0.6900269541778976, 0.37102461812787263,
    0.6863270777479893, 0.376400975164253, 0.6826666666666666, 0.3817485814908484,
    0.6790450928381963, 0.3870677429684483, 0.6754617414248021, 0.3923587606028639,
    0.6719160104986877, 0.3976219306471385, 0.6684073107049608, 0.4028575447010835,
    0.6649350649350649, 0.4080658898082217, 0.661498708010336, 0.41324724855021927,
    0.6580976863753213, 0.41840189913888387, 0.6547314578005116, 0.4235301155058032,
    0.6513994910941476, 0.42863216738969867, 0.6481012658227848, 0.4337083204215594,
    0.6448362720403022, 0.43875883620762796, 0.6416040100250626, 0.44378397241030104,
    0.6384039900249376, 0.4487839828270067, 0.6352357320099256, 0.4537591174671205,
    0.6320987654320988, 0.4587096226269767, 0.628992628992629, 0.46363574096303256,
    0.6259168704156479, 0.46853771156323926, 0.6228710462287105, 0.4734157700166721,
    0.6198547215496368, 0.47827014848147026, 0.6168674698795181, 0.48310107575113576,
    0.6139088729016786, 0.48790877731923904, 0.6109785202863962, 0.4926934754425752,
    0.6080760095011877, 0.4974553892028189, 0.6052009456264775, 0.5021947345667155,
    0.6023529411764705, 0.5069117244448544, 0.5995316159250585, 0.5116065687490621,
    0.5967365967365967, 0.5162794744484545, 0.5939675174013921, 0.5209306456241853,
    0.5912240184757506, 0.5255602835229274, 0.5885057471264368, 0.5301685866091216,
    0.585812356979405, 0.5347557506160276, 0.5831435079726651, 0.5393219685956089,
    0.5804988662131519, 0.5438674309672835, 0.5778781038374717, 0.5483923255655733,
    0.5752808988764045, 0.5528968376866776, 0.5727069351230425, 0.5573811501340064,
    0.5701559020044543, 0.5618454432626918, 0.5676274944567627, 0.5662898950231159,
    0.565121412803532, 0.5707146810034716, 0.5626373626373626, 0.575119974471388,
    0.5601750547045952, 0.5795059464146423, 0.5577342047930284, 0.5838727655809826,
    0.5553145336225597, 0.588220598517086, 0.5529157667386609, 0.5925496096066716,
    0.5505376344086022, 0.5968599611077938, 0.5481798715203426, 0.6011518131893347,
    0.5458422174840085, 0.6054253239667169, 0.5435244161358811, 0.6096806495368553,
    0.5412262156448203, 0.6139179440123704, 0.5389473684210526, 0.6181373595550788,
    0.5366876310272537, 0.6223390464087787, 0.534446764091858, 0.6265231529313529,
    0.5322245322245323, 0.6306898256261987, 0.5300207039337475, 0.6348392091730102,
    0.5278350515463918, 0.6389714464579207, 0.5256673511293635, 0.6430866786030273,
    0.523517382413088, 0.6471850449953095, 0.5213849287169042, 0.6512666833149582,
    0.5192697768762677, 0.6553317295631277, 0.5171717171717172, 0.6593803180891278,
    0.5150905432595574, 0.6634125816170662, 0.5130260521042084, 0.6674286512719563,
    0.5109780439121756, 0.6714286566053024, 0.5089463220675944, 0.6754127256201768,
    0.5069306930693069, 0.6793809847957973, 0.504930966469428, 0.6833335591116206,
    0.5029469548133595, 0.6872705720709603, 0.5009784735812133, 0.691192145724142,
};
FPM_TABLE double FPM_EXP2_J[32] = {
    1.0, 1.0218971486541166, 1.0442737824274138, 1.0671404006768237,
    1.0905077326652577, 1.1143867425958924, 1.1387886347566916, 1.1637248587775775,
    1.189207115002721, 1.215247359980469, 1.241857812073484, 1.2690509571917332,
    1.2968395546510096, 1.3252366431597413, 1.3542555469368927, 1.383909881963832,
    1.4142135623730951, 1.4451808069770467, 1.4768261459394993, 1.5091644275934228,
    1.5422108254079407, 1.5759808451078865, 1.6104903319492543, 1.645755478153965,
    1.681792830507429, 1.718619298122478, 1.7562521603732995, 1.7947090750031072,
    1.8340080864093424, 1.8741676341103, 1.9152065613971474, 1.9571441241754002,
};

/* log(x) for a positive, finite, normal binary64 x; ~1e-16 absolute.
 * x = 2^e m, m in [1, 2); the top 7 bits of m select c_i, r = m/c_i - 1
 * (|r| <= 2^-8, one rounding by the fused multiply-add) and
 * log m = T_i + log1p(r), log1p by its Taylor series to degree 7. */
FPM_FN double fpm_log_d_tab(double x, const double *tab)
{
    const uint64_t ix = fpm_d2u(x);
    const int e = (int)((ix >> 52) & 0x7ff) - 1023;
    const double m = fpm_u2d((ix & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    const int i = (int)((ix >> 45) & 127);
    const double r = __builtin_fma(m, tab[2 * i], -1.0);
    double q = 1.0 / 7.0;
    q = __builtin_fma(q, r, -1.0 / 6.0);
    q = __builtin_fma(q, r, 1.0 / 5.0);
    q = __builtin_fma(q, r, -0.25);
    q = __builtin_fma(q, r, 1.0 / 3.0);
    q = __builtin_fma(q, r, -0.5);
    const double l1p = __builtin_fma(q, r * r, r);
    const double de = (double)e;
    /* ln2 = 0.6931471805598903 (exact, 11 trailing zero bits) + 5.4979e-14 */
    return __builtin_fma(de, 0.69314718055989033, tab[2 * i + 1]) +
           __builtin_fma(de, 5.497923018708371e-14, l1p);
}
FPM_FN double fpm_log_d(double x) { return fpm_log_d_tab(x, FPM_LOG_TAB); }

/* e^t rounded to single precision, for t in [-104, 89]; ~1e-16 relative
 * before the final rounding (overflow and underflow selected at the end).
 * t = (32 k' + j) ln2/32 + r, |r| <= ln2/64: e^t = 2^k' 2^(j/32) e^r with
 * e^r - 1 by its Taylor series to degree 6. */
FPM_FN float fpm_exp_d_to_f_core(double tc, const double *exp2j)
{
    const double k = __builtin_rint(tc * 46.16624130844683);
    /* ln2/32 = 0.021660849392446835 (exact, 16 trailing zero bits) + 5.1456e-14 */
    double r = __builtin_fma(k, -0.021660849392446835, tc);
    r = __builtin_fma(k, -5.145609244655338e-14, r);
    double p = 1.0 / 720.0;
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r * r, r);
    const int ki = (int)k;
    const int j = ki & 31, ke = ki >> 5;
    const double y = __builtin_fma(exp2j[j], p, exp2j[j]);
    /* binary64 scaling is exact for ke in [-152, 130] */
    return (float)(y * fpm_u2d((uint64_t)(ke + 1023) << 52));
}
FPM_FN float fpm_exp_d_to_f_tab(double t, const double *exp2j)
{
    const double tc = __builtin_fmin(__builtin_fmax(t, -104.0), 89.0);
    float res = fpm_exp_d_to_f_core(tc, exp2j);
    res = t > 89.0 ? FPM_INF : res;
    return t < -104.0 ? 0.0f : res;
}
/* fpm_exp_d_to_f_tab of a t known to lie in [-104, 89], where the clamp and
 * the overflow and underflow selects leave every value unchanged. */
FPM_FN float fpm_exp_d_to_f_in_range(double t, const double *exp2j) { return fpm_exp_d_to_f_core(t, exp2j); }
FPM_FN float fpm_exp_d_to_f(double t) { return fpm_exp_d_to_f_tab(t, FPM_EXP2_J); }

/* x^y, single precision, for x >= 0 (negative x gives NaN). */
FPM_FN float fpm_powf_tab(float x, float y, const double *logtab, const double *exp2j)
{
    const float xs = (x > 0.0f && x != FPM_INF) ? x : 1.0f;
    float res = fpm_exp_d_to_f_tab((double)y * fpm_log_d_tab((double)xs, logtab), exp2j);
    /* special cases, lowest priority first */
    res = x == FPM_INF ? (y > 0.0f ? FPM_INF : 0.0f) : res;
    res = x == 0.0f ? (y > 0.0f ? 0.0f : FPM_INF) : res;
    res = x < 0.0f ? FPM_NAN : res;
    res = x == 1.0f ? 1.0f : res;
    res = (x != x || y != y) ? x + y : res;
    return y == 0.0f ? 1.0f : res;
}
FPM_FN float fpm_powf(float x, float y) { return fpm_powf_tab(x, y, FPM_LOG_TAB, FPM_EXP2_J); }

FPM_FN float fpm_sqrtf(float x) { return __builtin_sqrtf(x); }

#endif /* FP_MATH_H */

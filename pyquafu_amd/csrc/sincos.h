// sin and cos of one angle, for the phase factors of diagonal runs.
//
// Method: Cody-Waite argument reduction by pi/2 with a three-part constant (the first two parts
// carry 32 significant bits in double, 12 in float, so k * part is exact while |k| < 2^20 / 2^12),
// then the Taylor series of sin and cos on [-pi/4, pi/4] evaluated with Horner's rule (to r^15 /
// r^16 in double, whose truncation error is below 5e-17, and to r^11 / r^12 in float), and the
// quadrant k mod 4 picks and signs the two results. The split of pi/2 comes from an exact
// rational value of pi (Machin's formula). Accurate to a few ulp for |x| < 2^20 (double) and
// |x| < 6000 (float); the angles of a diagonal run are sums of gate phases, far below that.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define QFVM_HD __host__ __device__
#else
#define QFVM_HD
#endif

namespace qfvm {

QFVM_HD inline void sin_cos(double x, double* s, double* c) {
    const double k = rint(x * 0x1.45f306dc9c883p-1);  // x * 2/pi
    double r = fma(-k, 0x1.921fb54400000p+0, x);
    r = fma(-k, 0x1.0b4611a600000p-34, r);
    r = fma(-k, 0x1.3198a2e037073p-69, r);
    const double r2 = r * r;
    double ps = -1.0 / 1307674368000.0;  // -1/15!
    ps = fma(ps, r2, 1.0 / 6227020800.0);
    ps = fma(ps, r2, -1.0 / 39916800.0);
    ps = fma(ps, r2, 1.0 / 362880.0);
    ps = fma(ps, r2, -1.0 / 5040.0);
    ps = fma(ps, r2, 1.0 / 120.0);
    ps = fma(ps, r2, -1.0 / 6.0);
    const double sr = fma(r * r2, ps, r);
    double pc = 1.0 / 20922789888000.0;  // 1/16!
    pc = fma(pc, r2, -1.0 / 87178291200.0);
    pc = fma(pc, r2, 1.0 / 479001600.0);
    pc = fma(pc, r2, -1.0 / 3628800.0);
    pc = fma(pc, r2, 1.0 / 40320.0);
    pc = fma(pc, r2, -1.0 / 720.0);
    pc = fma(pc, r2, 1.0 / 24.0);
    const double cr = fma(r2 * r2, pc, fma(-0.5, r2, 1.0));
    const int q = (int)((long long)k & 3);
    const double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

QFVM_HD inline void sin_cos(float x, float* s, float* c) {
    const float k = rintf(x * 0x1.45f306p-1f);
    float r = fmaf(-k, 0x1.92p+0f, x);
    r = fmaf(-k, 0x1.fb4p-12f, r);
    r = fmaf(-k, 0x1.4442d2p-24f, r);
    const float r2 = r * r;
    float ps = -1.0f / 39916800.0f;
    ps = fmaf(ps, r2, 1.0f / 362880.0f);
    ps = fmaf(ps, r2, -1.0f / 5040.0f);
    ps = fmaf(ps, r2, 1.0f / 120.0f);
    ps = fmaf(ps, r2, -1.0f / 6.0f);
    const float sr = fmaf(r * r2, ps, r);
    float pc = 1.0f / 479001600.0f;
    pc = fmaf(pc, r2, -1.0f / 3628800.0f);
    pc = fmaf(pc, r2, 1.0f / 40320.0f);
    pc = fmaf(pc, r2, -1.0f / 720.0f);
    pc = fmaf(pc, r2, 1.0f / 24.0f);
    const float cr = fmaf(r2 * r2, pc, fmaf(-0.5f, r2, 1.0f));
    const int q = (int)((long long)k & 3);
    const float sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

}  // namespace qfvm

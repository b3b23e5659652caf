// dexp2.hpp -- FP64 2^x for the one-pass FP64 kernels.
//
// Table-driven method (P. T. P. Tang, "Table-driven implementation of the
// exponential function in IEEE floating-point arithmetic", ACM TOMS 15(2),
// 1989), base 2: with k = rint(64 x), e = k >> 6, j = k & 63 and
// f = x - k / 64 in [-1/128, 1/128] (exact),
//   2^x = 2^e * 2^(j/64) * 2^f.
// 2^(j/64) comes from a 64-entry table held as a leading double t and a
// trailing correction t_lo (Tang's lead/trail split);
// q = 2^f - 1 = sum_{n=1..6} (ln 2)^n f^n / n! (Taylor; the first omitted
// term is below 3e-20); the result t + fma(t, q, t_lo) is scaled by ldexp,
// which also rounds subnormal results.  On 2e7 random arguments 0.2% of the
// results are one ulp from the correctly rounded value, none further.
// x < -1100 gives 0 and x > 1100 gives +inf (the argument is clamped there),
// a NaN argument gives a quiet NaN.  The tables and the
// coefficients were computed with 60-digit decimal arithmetic and rounded to
// nearest.  Callers stage the 128 table doubles in LDS (exp2_tab_load) and
// pass that copy.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace regot {
namespace dev {

static __constant__ double kExp2Tab[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};

// 2^(j/64) - kExp2Tab[j], rounded to nearest.
static __constant__ double kExp2TabLo[64] = {
    0x0.0p+0, -0x1.19083535b085dp-56, 0x1.d73e2a475b465p-55, 0x1.186be4bb284ffp-57,
    0x1.8a62e4adc610bp-54, 0x1.03a1727c57b53p-59, -0x1.6c51039449b3ap-54, -0x1.32fbf9af1369ep-54,
    -0x1.19041b9d78a76p-55, 0x1.e5b4c7b4968e4p-55, 0x1.e016e00a2643cp-54, 0x1.dc775814a8495p-55,
    0x1.9b07eb6c70573p-54, 0x1.2bd339940e9d9p-55, 0x1.612e8afad1255p-55, 0x1.0024754db41d5p-54,
    0x1.6f46ad23182e4p-55, 0x1.32721843659a6p-54, -0x1.63aeabf42eae2p-54, -0x1.5e436d661f5e3p-56,
    0x1.ada0911f09ebcp-55, -0x1.ef3691c309278p-58, 0x1.89b7a04ef80d0p-59, 0x1.3c1a3b69062f0p-56,
    0x1.d4397afec42e2p-56, -0x1.4b309d25957e3p-54, -0x1.07abe1db13cadp-55, 0x1.9bb2c011d93adp-54,
    0x1.6324c054647adp-54, 0x1.ba6f93080e65ep-54, -0x1.383c17e40b497p-54, -0x1.bb60987591c34p-54,
    -0x1.bdd3413b26456p-54, -0x1.bbe3a683c88abp-57, -0x1.16e4786887a99p-55, -0x1.0245957316dd3p-54,
    -0x1.41577ee04992fp-55, 0x1.05d02ba15797ep-56, -0x1.d4c1dd41532d8p-54, -0x1.fc6f89bd4f6bap-54,
    0x1.6e9f156864b27p-54, 0x1.5cc13a2e3976cp-55, -0x1.75fc781b57ebcp-57, -0x1.d185b7c1b85d1p-54,
    0x1.c7c46b071f2bep-56, -0x1.359495d1cd533p-54, -0x1.d2f6edb8d41e1p-54, 0x1.0fac90ef7fd31p-54,
    0x1.7a1cd345dcc81p-54, -0x1.2805e3084d708p-57, -0x1.5584f7e54ac3bp-56, 0x1.23dd07a2d9e84p-55,
    0x1.11065895048ddp-55, 0x1.2884dff483cadp-54, 0x1.503cbd1e949dbp-56, -0x1.cbc3743797a9cp-54,
    0x1.2ed02d75b3707p-55, 0x1.c2300696db532p-54, -0x1.1a5cd4f184b5cp-54, 0x1.39e8980a9cc8fp-55,
    -0x1.e9c23179c2893p-54, 0x1.dc7f486a4b6b0p-54, 0x1.9d3e12dd8a18bp-54, 0x1.74853f3a5931ep-55,
};

// Copy both tables into a 16-byte aligned 128-double LDS array, interleaved
// (lead, trail) so one 16-byte LDS read fetches both (all threads of the
// workgroup call it; the caller synchronises before use).
__device__ __forceinline__ void exp2_tab_load(double* lds)
{
    for (int i = threadIdx.x; i < 128; i += blockDim.x)
        lds[i] = (i & 1) ? kExp2TabLo[i >> 1] : kExp2Tab[i >> 1];
}

// 2^(y / 64) for y = 64 x.  Callers that form x as a product with k pass
// 64 k instead: scaling by a power of two commutes with every rounding (no
// overflow or underflow at these magnitudes), so y is exactly 64 x and the
// result is bit for bit dexp2(x).  Then k = rint(y) comes from the
// 1.5 * 2^52 shifter (one add; |y| < 2^51 after the clamp, ties to even as
// rint; the low word of the sum is k), and the polynomial runs on
// d = y - k = 64 f (exact) with the coefficients scaled by 64^-n, which again
// changes no rounding.
__device__ __forceinline__ double dexp2_64(double y, const double* tab)
{
    constexpr double c1 = 0x1.62e42fefa39efp-1 * 0x1p-6, c2 = 0x1.ebfbdff82c58fp-3 * 0x1p-12;
    constexpr double c3 = 0x1.c6b08d704a0c0p-5 * 0x1p-18, c4 = 0x1.3b2ab6fba4e77p-7 * 0x1p-24;
    constexpr double c5 = 0x1.5d87fe78a6731p-10 * 0x1p-30, c6 = 0x1.430912f86c787p-13 * 0x1p-36;
    // clamp: beyond +-1100 * 64 the result is already 0 or +inf (NaN last)
    const double yc = fmin(fmax(y, -70400.0), 70400.0);
    const double sh = yc + 0x1.8p52;
    const double kx = sh - 0x1.8p52;
    const int i = __double2loint(sh);
    const double d = yc - kx;
    const double q = d * (c1 + d * (c2 + d * (c3 + d * (c4 + d * (c5 + d * c6)))));
    const double2 t = *reinterpret_cast<const double2*>(tab + 2 * (i & 63));
    const double r = ldexp(t.x + fma(t.x, q, t.y), i >> 6);
    // a NaN argument gives a quiet NaN (exponent all ones, top mantissa bit set)
    return __hiloint2double((y != y) ? 0x7ff80000 : __double2hiint(r), __double2loint(r));
}

__device__ __forceinline__ double dexp2(double x, const double* tab)
{
    return dexp2_64(x * 64.0, tab);
}

}  // namespace dev
}  // namespace regot

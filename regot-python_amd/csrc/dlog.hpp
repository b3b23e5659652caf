// dlog.hpp -- FP64 natural logarithm for the APDAGD plan pass.
//
// Table-driven method (P. T. P. Tang, "Table-driven implementation of the
// logarithm function in IEEE floating-point arithmetic", ACM TOMS 16(4),
// 1990): x = 2^k m with m in [1, 2), and the top seven mantissa bits j pick a
// table entry.  For j >= 64 the reduced argument is m / 2 with k + 1, so it
// lies in [0.75, 1.5).  c_j is about 1 / (midpoint of the reduced interval),
// r = fma(m', c_j, -1) is |r| < 2^-7, and
//   log x = k ln2 - ln(c_j) + log1p(r),
// with log1p(r) = r - r^2/2 + ... - r^8/8 (Taylor; the first omitted term is
// below 2^-59 |r|).  c_0 = c_127 = 1, so next to x = 1 the result keeps its
// relative accuracy.  ln2 is held as a 42-bit lead (k * lead is exact) plus a
// trail, -ln(c_j) as a lead plus a trail, and k * lead - ln(c_j) is summed
// exactly (Fast2Sum: |k ln2| >= 0.69 > |ln c_j|, or k = 0).  Subnormal x is
// scaled by 2^54 first.  log(0) = -inf, log(+inf) = +inf, and a negative or
// NaN argument gives a quiet NaN.  The table comes from gen_dlog_table.py
// (60-digit decimal arithmetic, each double rounded to nearest).  Callers
// stage the 384 table doubles in LDS (log_tab_load) and pass that copy.
#pragma once

#include <hip/hip_runtime.h>

namespace regot {
namespace dev {

static __constant__ double kLogTab[384] = {
#include "dlog_table.inc"
};

constexpr double kLn2Lead = 0x1.62e42fefa38p-1;
constexpr double kLn2Trail = 0x1.ef35793c7673p-45;

__device__ __forceinline__ void log_tab_load(double* lds)
{
    for (int i = threadIdx.x; i < 384; i += blockDim.x) lds[i] = kLogTab[i];
}

__device__ __forceinline__ double dlog(double x, const double* tab)
{
    const bool sub = x < 0x1p-1022;
    const double xs = sub ? x * 0x1p54 : x;
    const int hx = __double2hiint(xs);
    const int j = (hx >> 13) & 127;
    const int up = j >> 6;
    const int k = ((hx >> 20) & 0x7ff) - 1023 + up - (sub ? 54 : 0);
    const double mr = __hiloint2double((hx & 0x000fffff) | (0x3ff00000 - (up << 20)),
                                       __double2loint(xs));
    const double c = tab[3 * j], lhi = tab[3 * j + 1], llo = tab[3 * j + 2];
    const double r = fma(mr, c, -1.0);
    double p = -0.125;
    p = fma(p, r, 0x1.2492492492492p-3);    // 1/7
    p = fma(p, r, -0x1.5555555555555p-3);   // -1/6
    p = fma(p, r, 0.2);
    p = fma(p, r, -0.25);
    p = fma(p, r, 0x1.5555555555555p-2);    // 1/3
    p = fma(p, r, -0.5);
    const double kd = static_cast<double>(k);
    const double big = kd * kLn2Lead;
    const double s = big + lhi;
    const double e = (big - s) + lhi;
    const double lo = fma(kd, kLn2Trail, llo) + e;
    double y = s + (r + fma(r * r, p, lo));
    y = (x == 0.0) ? -__builtin_inf() : y;
    y = (x == __builtin_inf()) ? x : y;
    y = (x < 0.0 || x != x) ? __builtin_nan("") : y;
    return y;
}

}  // namespace dev
}  // namespace regot

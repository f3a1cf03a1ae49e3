// r3d_dtlim.hpp - per-cell time-step limits and the screen of the last
// stage (host and device: the CPU tests run them directly).
//
// Included by r3d_stage.hpp inside its namespace choice (R3D_KNS), so that
// the bit-exact and the FMA-contracted builds each get their own copy.
#pragma once

#include <cmath>

#include "r3d_dev.hpp"

#ifndef R3D_KNS
#define R3D_KNS r3d
#endif

namespace R3D_KNS {
using namespace r3d;

__host__ __device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }
__host__ __device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }

// Per-cell time-step limits of a final state value (reference
// src/step.cpp:125-209): pm[0] = max(ogamma*UU/T), pm[1..3] = min of
// dd/sqrt(fast speed), the thermal and the viscous limit, pm[4] = min of the
// resistive limit (MHD).  With MAG the fast speed includes the Alfven term
// b2 = (bx^2+by^2+bz^2)*obeta/ro (src/step.cpp:144-156).
template <bool MAG>
__host__ __device__ __forceinline__ void dt_terms(const Consts& C, double ru, double rv, double rw,
                                         double ro, double tt, double bx, double by, double bz,
                                         double dd, double rk, double pm[NPART]) {
    const double uu = (ru * ru + rv * rv + rw * rw) * (1.0 / ro) * (1.0 / ro);
    pm[0] = dmax(pm[0], C.og * uu / tt);
    double fast;
    if constexpr (MAG) {
        const double b2 = (bx * bx + by * by + bz * bz) * C.obeta / ro;
        fast = uu + C.gamma * tt + b2 + 2.0 * sqrt(uu * (C.gamma * tt + b2));
    } else {
        fast = uu + C.gamma * tt + 2.0 * sqrt(uu * C.gamma * tt);
    }
    pm[1] = dmin(pm[1], dd / sqrt(fast));
    pm[2] = dmin(pm[2], 0.5 * dd * dd * C.repr * C.cv * ro / rk);
    pm[3] = dmin(pm[3], 0.375 * dd * dd * C.re * ro);
    if constexpr (MAG) pm[4] = dmin(pm[4], 0.5 * dd * dd * C.rm);
}

// Constants of the time-step screen, formed once per block from the
// previous extrema: the thresholds thr_um (ogamma*UU/T) and thr_lim (every
// limit) with a relative margin of 1e-12 folded in, and the coefficients of
// the limits (reference src/step.cpp:125-209).
enum { SC_UM, SC_T, SC_T2, SC_TH, SC_VI, SC_RS, NSC };
__host__ __device__ __forceinline__ void screen_consts(const Consts& C, double lim_prev, double um_prev,
                                              double* k) {
    const double up = 1.0 + 1e-12, dn = 1.0 - 1e-12;
    const double thr_lim = lim_prev * C.scr_up, thr_um = um_prev * C.scr_dn;
    k[SC_UM] = thr_um * dn;
    k[SC_T] = thr_lim * up;
    k[SC_T2] = thr_lim * thr_lim * up;
    k[SC_TH] = 0.5 * C.repr * C.cv;
    k[SC_VI] = 0.375 * C.re;
    k[SC_RS] = 0.5 * C.rm;
}

// Whether a final state value may set an extremum of the next reduction:
// one of its limits at most thr_lim, or ogamma*UU/T above thr_um (a cell
// with ogamma*UU/T = thr_um = 0 cannot raise the maximum above the identity
// 0).  With UU = m2/rho^2 the tests are multiplied out, so no division or
// square root is needed: ogamma m2 > thr_um T rho^2, thermal dd^2 rho (0.5 repr cv) <=
// thr_lim rk, viscous, resistive, and for dd/sqrt(fast) the bound fast =
// (sqrt(UU) + sqrt(G))^2 <= 5 UU + 1.25 G (G = gamma T + b2, from
// 2 sqrt(UU G) <= 4 UU + G/4).  Each side is a product of a few rounded
// factors, so its relative error, like that of the exact limits, is far
// below the margin 1e-12 in k: a cell that fails every test has every exact
// limit above thr_lim and ogamma*UU/T at most thr_um.  The limit tests are
// negated strict comparisons, m2 is NaN in the umach test only if a limit
// test sees it too, and a non-positive rho, T or conductivity always hits,
// so a NaN or an unphysical value takes the exact evaluation.
template <bool MAG>
__host__ __device__ __forceinline__ bool dt_screen(const Consts& C, double ru, double rv, double rw,
                                          double ro, double tt, double bx, double by, double bz,
                                          double dd, double rk, const double* k) {
    const double m2 = ru * ru + rv * rv + rw * rw;
    const double r2 = ro * ro;
    const double d2 = dd * dd;
    const double tr2 = tt * r2;
    const double x = d2 * ro;
    double g2 = C.gamma * tr2;  // G rho^2
    if constexpr (MAG) g2 = g2 + (bx * bx + by * by + bz * bz) * C.obeta * ro;
    bool hit = !(ro > 0.0 && tt > 0.0 && rk > 0.0);
    hit |= C.og * m2 > k[SC_UM] * tr2;  // strict: a state at rest passes no cell
    hit |= !(x * k[SC_TH] > k[SC_T] * rk);
    hit |= !(x * k[SC_VI] > k[SC_T]);
    hit |= !(d2 * r2 > k[SC_T2] * (5.0 * m2 + 1.25 * g2));
    if constexpr (MAG) hit |= !(d2 * k[SC_RS] > k[SC_T]);
    return hit;
}

// The screened reduction's extrema are the exact ones if the least limit
// found is at most the widened threshold lim_prev * scr_up and the largest
// ogamma*UU/T at least um_prev * scr_dn (the thresholds exactly as
// screen_consts forms them): a cell with a smaller limit or a larger
// ogamma*UU/T would have passed the screen and been included.
__host__ __device__ __forceinline__ bool screen_proven(const Consts& C, double lim_min,
                                                      double um_max, double lim_prev,
                                                      double um_prev) {
    return lim_min <= lim_prev * C.scr_up && um_max >= um_prev * C.scr_dn;
}

}  // namespace R3D_KNS

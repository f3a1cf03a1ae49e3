// r3d_adj.hpp - per-cell hydrodynamic flux and its hand-written adjoint.
//
// The discrete adjoint of the time step needs, per RK stage, the product of
// the transposed flux Jacobian with an adjoint vector: lamX = (dF/dX)^T lamF.
// F at a cell depends on the primaries at 19 points (the cell, its 6 face and
// 12 edge neighbours) plus, on the walls, the plane two cells in.  cell_flux
// evaluates F from such a stencil with exactly the operations of the stage
// kernel (so it also reproduces the reference's FLUXES bit for bit), and
// cell_flux_vjp runs that evaluation in reverse: from lamF it returns the
// adjoint of every stencil primary and the derivatives of lamF . F with respect
// to the Reynolds-number coefficients ore = 1/Re and tk = ocv/(Re Pr).
//
// Both are __host__ __device__ so that the CPU tests can check them against
// the oracle's FLUXES and against finite differences.
#pragma once

#include "r3d_dev.hpp"

namespace r3d {

// Stencil points: the cell, faces, edges in the xy, xz and yz planes.
enum {
    P_C = 0, P_E, P_W, P_N, P_S, P_U, P_D,
    P_NE, P_NW, P_SE, P_SW,
    P_EU, P_ED, P_WU, P_WD,
    P_NU, P_ND, P_SU, P_SD,
    NPTS
};

// (di, dj, dk) of each point; U/D are k+1/k-1.
__host__ __device__ inline void pt_offset(int p, int& di, int& dj, int& dk) {
    const signed char o[NPTS][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                                    {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, 1, 0},  {1, -1, 0},
                                    {-1, -1, 0}, {1, 0, 1},  {1, 0, -1}, {-1, 0, 1},  {-1, 0, -1},
                                    {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
    di = o[p][0];
    dj = o[p][1];
    dk = o[p][2];
}

// A stencil point: the primaries and, in a state stencil, 1/RO and the
// velocities derived from them (derive_point; the operations of the stage
// kernel's derive, so the same doubles).  An adjoint stencil uses the first
// five members only.
struct Pt {
    double ru, rv, rw, ro, tt;
    double ri, u, v, w;
};
__host__ __device__ inline void derive_point(Pt& q) {
    q.ri = 1.0 / q.ro;
    q.u = q.ru * q.ri;
    q.v = q.rv * q.ri;
    q.w = q.rw * q.ri;
}

struct Stencil {
    Pt p[NPTS];
    double rw2;  // RW two planes in: k+2 on the top wall, k-2 on the bottom wall
};

// Metrics of the cell's column and plane, and which wall it is on.
struct CellMet {
    double dxx, d2xx, dyy, d2yy, dzz, d2zz, rk, dk;
    int wall;     // 0 interior, 1 top (k = 1), 2 bottom (k = npz)
    int mag_bot;  // LMAG one-sided stencils on this bottom-wall cell (npez > 1)
};

// Magnetic field at the stencil points, and the wall extras of the LMAG
// one-sided stencils: RO two planes in (for WW there), BX, BY two and three
// planes in (k+2, k+3 on the top wall, k-2, k-3 on the bottom wall).
struct BPt {
    double bx, by, bz;
};
struct MagStencil {
    BPt p[NPTS];
    double ro2, bx2, by2, bx3, by3;
};

// Stencil entries of the per-cell transpose that can be nonzero: bit f of
// VJP_MASK[p] for primary f at point p.  The in-plane and vertical edge
// points only carry velocity adjoints (u, v at NE..SW; u, w at EU..WD; v, w
// at NU..SD), which reach the momenta and the density; every other entry of
// cell_flux_vjp / cell_flux_mag_vjp there is exactly zero.  mag: the whole
// MHD transpose; lmag_only: cell_flux_mag_extra_vjp alone, which reaches the
// face points only and never the temperature.
__host__ __device__ constexpr unsigned vjp_mask(int p, bool mag, bool lmag_only = false) {
    if (lmag_only) return p <= P_D ? 0xffu & ~(1u << QTT) : 0u;
    const unsigned all = mag ? 0xffu : 0x1fu;
    const unsigned UV = (1u << QRU) | (1u << QRV) | (1u << QRO);
    const unsigned UW = (1u << QRU) | (1u << QRW) | (1u << QRO);
    const unsigned VW = (1u << QRV) | (1u << QRW) | (1u << QRO);
    return p <= P_D ? all : (p <= P_SW ? UV : (p <= P_WD ? UW : VW));
}

// F = (FU, FV, FW, FR, FT) in the order of the primaries QRU..QTT.
__host__ __device__ inline void cell_flux(const Stencil& s, const CellMet& m, const Consts& C,
                                          double F[5]) {
    double ri[NPTS], u[NPTS], v[NPTS], w[NPTS], pr[NPTS];
    for (int q = 0; q < NPTS; ++q) {
        ri[q] = s.p[q].ri;
        u[q] = s.p[q].u;
        v[q] = s.p[q].v;
        w[q] = s.p[q].w;
        pr[q] = s.p[q].ro * s.p[q].tt;
    }
    const double hxd = C.hx, dxx = m.dxx, d2xx = m.d2xx, dyy = m.dyy, d2yy = m.d2yy;
    const double dzz = m.dzz, d2zz = m.d2zz;
    const double tmpy = C.hy * dyy, tmpy1 = C.hy * d2yy, tmpy2 = C.h2y * dyy * dyy;
    const double tmpz = C.hz * dzz, tmpz1 = C.hz * d2zz, tmpz2 = C.h2z * dzz * dzz;
    const Pt* P = s.p;
#define LAPX(e, ww, c2) ((e - ww) * hxd * d2xx + (e - c2 + ww) * C.h2x * dxx * dxx)
    // FR
    {
        const double w1top = (4.0 * P[P_U].rw - s.rw2) * tmpz;
        const double w1bot = (s.rw2 - 4.0 * P[P_D].rw) * tmpz;
        const double w1mid = (P[P_U].rw - P[P_D].rw) * tmpz;
        const double w1 = m.wall == 1 ? w1top : (m.wall == 2 ? w1bot : w1mid);
        double fr = (P[P_W].ru - P[P_E].ru) * hxd * dxx;
        fr = fr - (P[P_N].rv - P[P_S].rv) * tmpy;
        fr = fr - w1;
        F[QRO] = fr;
    }
    // FU
    {
        const double uE = u[P_E], uW = u[P_W], uN = u[P_N], uS = u[P_S], u2 = 2.0 * u[P_C];
        double fu = (pr[P_W] - pr[P_E]) * hxd * dxx;
        fu = fu - (P[P_E].ru * uE - P[P_W].ru * uW) * hxd * dxx;
        fu = fu - (P[P_N].ru * v[P_N] - P[P_S].ru * v[P_S]) * tmpy;
        fu = fu - (P[P_U].ru * w[P_U] - P[P_D].ru * w[P_D]) * tmpz;
        fu = fu + C.ore * (C.c43 * LAPX(uE, uW, u2) + (uN - uS) * tmpy1 + (uN - u2 + uS) * tmpy2 +
                           (u[P_U] - u[P_D]) * tmpz1 + (u[P_U] - u2 + u[P_D]) * tmpz2);
        const double vxN = (v[P_NE] - v[P_NW]) * hxd * dxx;
        const double vxS = (v[P_SE] - v[P_SW]) * hxd * dxx;
        fu = fu + C.t2 * (vxN - vxS) * C.hy * dyy;
        const double wzE = (w[P_EU] - w[P_ED]) * C.hz * dzz, wzW = (w[P_WU] - w[P_WD]) * C.hz * dzz;
        fu = fu + C.t2 * (wzE - wzW) * hxd * dxx;
        if (C.lrot) fu = fu + C.omz * P[P_C].rv;
        F[QRU] = fu;
    }
    // FV
    {
        const double vE = v[P_E], vW = v[P_W], vN = v[P_N], vS = v[P_S], v2 = 2.0 * v[P_C];
        double fv = (pr[P_S] - pr[P_N]) * tmpy;
        fv = fv - (P[P_E].rv * u[P_E] - P[P_W].rv * u[P_W]) * hxd * dxx;
        fv = fv - (P[P_N].rv * vN - P[P_S].rv * vS) * tmpy;
        fv = fv - (P[P_U].rv * w[P_U] - P[P_D].rv * w[P_D]) * tmpz;
        fv = fv + C.ore * (LAPX(vE, vW, v2) + C.c43 * ((vN - vS) * tmpy1 + (vN - v2 + vS) * tmpy2) +
                           (v[P_U] - v[P_D]) * tmpz1 + (v[P_U] - v2 + v[P_D]) * tmpz2);
        const double uxN = (u[P_NE] - u[P_NW]) * hxd * dxx;
        const double uxS = (u[P_SE] - u[P_SW]) * hxd * dxx;
        fv = fv + C.t2 * (uxN - uxS) * C.hy * dyy;
        const double wzN = (w[P_NU] - w[P_ND]) * C.hz * dzz, wzS = (w[P_SU] - w[P_SD]) * C.hz * dzz;
        fv = fv + C.t2 * (wzN - wzS) * C.hy * dyy;
        if (C.lrot) fv = fv + (-C.omz * P[P_C].ru + C.omx * P[P_C].rw);
        F[QRV] = fv;
    }
    // FW
    {
        const double wE = w[P_E], wW = w[P_W], wN = w[P_N], wS = w[P_S], w2 = 2.0 * w[P_C];
        double fw = C.grav * P[P_C].ro;
        fw = fw - (pr[P_U] - pr[P_D]) * tmpz;
        fw = fw - (P[P_E].rw * u[P_E] - P[P_W].rw * u[P_W]) * hxd * dxx;
        fw = fw - (P[P_N].rw * v[P_N] - P[P_S].rw * v[P_S]) * tmpy;
        fw = fw - (P[P_U].rw * w[P_U] - P[P_D].rw * w[P_D]) * tmpz;
        fw = fw + C.ore * (LAPX(wE, wW, w2) + (wN - wS) * tmpy1 + (wN - w2 + wS) * tmpy2 +
                           C.c43 * ((w[P_U] - w[P_D]) * tmpz1 + (w[P_U] - w2 + w[P_D]) * tmpz2));
        const double uzE = (u[P_EU] - u[P_ED]) * C.hz * dzz, uzW = (u[P_WU] - u[P_WD]) * C.hz * dzz;
        const double vzN = (v[P_NU] - v[P_ND]) * C.hz * dzz, vzS = (v[P_SU] - v[P_SD]) * C.hz * dzz;
        fw = fw + C.t2 * ((uzE - uzW) * hxd * dxx + (vzN - vzS) * C.hy * dyy);
        if (C.lrot) fw = fw + -C.omx * P[P_C].rv;
        F[QRW] = fw;
    }
    // FT
    {
        const double R = ri[P_C], rk = m.rk, dk = m.dk;
        const double tt0 = P[P_C].tt, ttU = P[P_U].tt, ttD = P[P_D].tt;
        const double ttE = P[P_E].tt, ttW = P[P_W].tt, ttN = P[P_N].tt, ttS = P[P_S].tt;
        const double t2c = 2.0 * tt0;
        const double wz = (w[P_U] - w[P_D]) * C.hz * dzz;
        const double uz = (u[P_U] - u[P_D]) * C.hz * dzz;
        const double vz = (v[P_U] - v[P_D]) * C.hz * dzz;
        double ft = -1.0 * u[P_C] * (ttE - ttW) * hxd * dxx;
        ft = ft - v[P_C] * (ttN - ttS) * tmpy;
        ft = ft - w[P_C] * (ttU - ttD) * tmpz;
        ft = ft + ((ttE - ttW) * hxd * d2xx + (ttE - t2c + ttW) * C.h2x * dxx * dxx) * R * C.tk * rk;
        ft = ft + ((ttN - ttS) * tmpy1 + (ttN - t2c + ttS) * tmpy2) * R * C.tk * rk;
        ft = ft + ((ttU - ttD) * tmpz1 + (ttU - t2c + ttD) * tmpz2) * R * C.tk * rk;
        ft = ft + (ttU - ttD) * C.hz * dzz * R * C.tk * dk;
        const double ux = (u[P_E] - u[P_W]) * hxd * dxx;
        const double vx = (v[P_E] - v[P_W]) * hxd * dxx;
        const double wx = (w[P_E] - w[P_W]) * hxd * dxx;
        ft = ft + (C.c43 * ux * ux + vx * vx + wx * wx) * R * C.t1;
        ft = ft - C.ocv * ux * tt0;
        const double uy = (u[P_N] - u[P_S]) * C.hy * dyy;
        ft = ft + (uy * uy + 2.0 * uy * vx) * R * C.t1;
        const double vy = (v[P_N] - v[P_S]) * C.hy * dyy;
        ft = ft + vy * vy * R * C.t3;
        ft = ft - C.ocv * vy * tt0;
        ft = ft + (wz * wz - ux * wz - ux * vy - vy * wz) * R * C.t3;
        ft = ft - C.ocv * wz * tt0;
        const double wy = (w[P_N] - w[P_S]) * C.hy * dyy;
        ft = ft + (uz * uz + vz * vz + wy * wy + 2.0 * vz * wy) * R * C.t1;
        ft = ft + 2.0 * uz * wx * R * C.t1;
        F[QTT] = ft;
    }
#undef LAPX
}

// Reverse of cell_flux.  lam[f] weights F[f]; A receives d(lam . F)/d(primary)
// at every stencil point (A.rw2 for the wall plane two cells in), and
// g_ore, g_tk the derivatives with respect to the coefficients ore and tk
// (every other coefficient held fixed).  The arithmetic is the plain chain
// rule; it is not ordered to reproduce any reference rounding.
__host__ __device__ inline void cell_flux_vjp(const Stencil& s, const CellMet& m, const Consts& C,
                                              const double lam[5], Stencil& A, double& g_ore,
                                              double& g_tk) {
    double ri[NPTS], u[NPTS], v[NPTS], w[NPTS];
    for (int q = 0; q < NPTS; ++q) {
        ri[q] = s.p[q].ri;
        u[q] = s.p[q].u;
        v[q] = s.p[q].v;
        w[q] = s.p[q].w;
    }
    // adjoints of the derived point quantities
    double au[NPTS], av[NPTS], aw[NPTS], ap[NPTS];
    for (int q = 0; q < NPTS; ++q) {
        au[q] = av[q] = aw[q] = ap[q] = 0.0;
        A.p[q].ru = A.p[q].rv = A.p[q].rw = A.p[q].ro = A.p[q].tt = 0.0;
    }
    A.rw2 = 0.0;
    double ari = 0.0;  // adjoint of 1/RO at the cell (thermal diffusion, heating)
    g_ore = 0.0;
    g_tk = 0.0;
    const Pt* P = s.p;
    Pt* AP = A.p;
    const double hxd = C.hx * m.dxx;  // d/dx of a centred difference
    const double hyd = C.hy * m.dyy;
    const double hzd = C.hz * m.dzz;
    const double tmpy1 = C.hy * m.d2yy, tmpy2 = C.h2y * m.dyy * m.dyy;
    const double tmpz1 = C.hz * m.d2zz, tmpz2 = C.h2z * m.dzz * m.dzz;
    const double lx1 = C.hx * m.d2xx, lx2 = C.h2x * m.dxx * m.dxx;
    const double LU = lam[QRU], LV = lam[QRV], LW = lam[QRW], LR = lam[QRO], LT = lam[QTT];

    // ---- FR
    AP[P_W].ru += LR * hxd;
    AP[P_E].ru -= LR * hxd;
    AP[P_N].rv -= LR * hyd;
    AP[P_S].rv += LR * hyd;
    if (m.wall == 1) {
        AP[P_U].rw -= 4.0 * LR * hzd;
        A.rw2 += LR * hzd;
    } else if (m.wall == 2) {
        A.rw2 -= LR * hzd;
        AP[P_D].rw += 4.0 * LR * hzd;
    } else {
        AP[P_U].rw -= LR * hzd;
        AP[P_D].rw += LR * hzd;
    }
    // ---- advection of the three momenta and the pressure gradient
    // F_c -= d/dx(m_c u) + d/dy(m_c v) + d/dz(m_c w); m_c = ru, rv, rw
    {
        const double Lm[3] = {LU, LV, LW};
        for (int c = 0; c < 3; ++c) {
            const double L = Lm[c];
            auto mom = [&](int q) { return c == 0 ? P[q].ru : (c == 1 ? P[q].rv : P[q].rw); };
            auto amom = [&](int q) -> double& { return c == 0 ? AP[q].ru : (c == 1 ? AP[q].rv : AP[q].rw); };
            amom(P_E) -= L * hxd * u[P_E];
            au[P_E] -= L * hxd * mom(P_E);
            amom(P_W) += L * hxd * u[P_W];
            au[P_W] += L * hxd * mom(P_W);
            amom(P_N) -= L * hyd * v[P_N];
            av[P_N] -= L * hyd * mom(P_N);
            amom(P_S) += L * hyd * v[P_S];
            av[P_S] += L * hyd * mom(P_S);
            amom(P_U) -= L * hzd * w[P_U];
            aw[P_U] -= L * hzd * mom(P_U);
            amom(P_D) += L * hzd * w[P_D];
            aw[P_D] += L * hzd * mom(P_D);
        }
        ap[P_W] += LU * hxd;
        ap[P_E] -= LU * hxd;
        ap[P_S] += LV * hyd;
        ap[P_N] -= LV * hyd;
        ap[P_U] -= LW * hzd;
        ap[P_D] += LW * hzd;
        AP[P_C].ro += LW * C.grav;
    }
    // ---- viscous Laplacians: the c43 factor sits on the component's own axis
    {
        double* aq[3] = {au, av, aw};
        const double* q3[3] = {u, v, w};
        const double Lm[3] = {LU, LV, LW};
        for (int c = 0; c < 3; ++c) {
            const double* q = q3[c];
            const double L = Lm[c] * C.ore;
            const double fx = c == 0 ? C.c43 : 1.0, fy = c == 1 ? C.c43 : 1.0,
                         fz = c == 2 ? C.c43 : 1.0;
            aq[c][P_E] += L * fx * (lx1 + lx2);
            aq[c][P_W] += L * fx * (-lx1 + lx2);
            aq[c][P_N] += L * fy * (tmpy1 + tmpy2);
            aq[c][P_S] += L * fy * (-tmpy1 + tmpy2);
            aq[c][P_U] += L * fz * (tmpz1 + tmpz2);
            aq[c][P_D] += L * fz * (-tmpz1 + tmpz2);
            aq[c][P_C] += -2.0 * L * (fx * lx2 + fy * tmpy2 + fz * tmpz2);
            // d/d(ore) of the whole bracket
            const double lap = fx * ((q[P_E] - q[P_W]) * lx1 + (q[P_E] - 2.0 * q[P_C] + q[P_W]) * lx2) +
                               fy * ((q[P_N] - q[P_S]) * tmpy1 + (q[P_N] - 2.0 * q[P_C] + q[P_S]) * tmpy2) +
                               fz * ((q[P_U] - q[P_D]) * tmpz1 + (q[P_U] - 2.0 * q[P_C] + q[P_D]) * tmpz2);
            g_ore += Lm[c] * lap;
        }
    }
    // ---- cross terms t2 = c13*ore (src/fluxes.cpp:362-514)
    {
        const double t2 = C.t2, c13 = C.c13;
        // FU: t2 (vx(N) - vx(S)) hy dyy + t2 (wz(E) - wz(W)) hx dxx
        const double cu1 = LU * t2 * hyd * hxd, cu2 = LU * t2 * hxd * hzd;
        av[P_NE] += cu1; av[P_NW] -= cu1; av[P_SE] -= cu1; av[P_SW] += cu1;
        aw[P_EU] += cu2; aw[P_ED] -= cu2; aw[P_WU] -= cu2; aw[P_WD] += cu2;
        // FV: t2 (ux(N) - ux(S)) hy dyy + t2 (wz(N) - wz(S)) hy dyy
        const double cv1 = LV * t2 * hyd * hxd, cv2 = LV * t2 * hyd * hzd;
        au[P_NE] += cv1; au[P_NW] -= cv1; au[P_SE] -= cv1; au[P_SW] += cv1;
        aw[P_NU] += cv2; aw[P_ND] -= cv2; aw[P_SU] -= cv2; aw[P_SD] += cv2;
        // FW: t2 ((uz(E) - uz(W)) hx dxx + (vz(N) - vz(S)) hy dyy)
        const double cw1 = LW * t2 * hxd * hzd, cw2 = LW * t2 * hyd * hzd;
        au[P_EU] += cw1; au[P_ED] -= cw1; au[P_WU] -= cw1; au[P_WD] += cw1;
        av[P_NU] += cw2; av[P_ND] -= cw2; av[P_SU] -= cw2; av[P_SD] += cw2;
        // d/d(ore) through t2 = c13 * ore
        const double xu = ((v[P_NE] - v[P_NW]) - (v[P_SE] - v[P_SW])) * hxd * hyd +
                          ((w[P_EU] - w[P_ED]) - (w[P_WU] - w[P_WD])) * hzd * hxd;
        const double xv = ((u[P_NE] - u[P_NW]) - (u[P_SE] - u[P_SW])) * hxd * hyd +
                          ((w[P_NU] - w[P_ND]) - (w[P_SU] - w[P_SD])) * hzd * hyd;
        const double xw = ((u[P_EU] - u[P_ED]) - (u[P_WU] - u[P_WD])) * hzd * hxd +
                          ((v[P_NU] - v[P_ND]) - (v[P_SU] - v[P_SD])) * hzd * hyd;
        g_ore += c13 * (LU * xu + LV * xv + LW * xw);
    }
    // ---- rotation
    if (C.lrot) {
        AP[P_C].rv += LU * C.omz;
        AP[P_C].ru -= LV * C.omz;
        AP[P_C].rw += LV * C.omx;
        AP[P_C].rv -= LW * C.omx;
    }
    // ---- energy equation
    {
        const double R = ri[P_C];
        const double tt0 = P[P_C].tt;
        const double ttE = P[P_E].tt, ttW = P[P_W].tt, ttN = P[P_N].tt, ttS = P[P_S].tt;
        const double ttU = P[P_U].tt, ttD = P[P_D].tt;
        // advection -u dT/dx - v dT/dy - w dT/dz
        au[P_C] -= LT * (ttE - ttW) * hxd;
        av[P_C] -= LT * (ttN - ttS) * hyd;
        aw[P_C] -= LT * (ttU - ttD) * hzd;
        AP[P_E].tt -= LT * u[P_C] * hxd;
        AP[P_W].tt += LT * u[P_C] * hxd;
        AP[P_N].tt -= LT * v[P_C] * hyd;
        AP[P_S].tt += LT * v[P_C] * hyd;
        AP[P_U].tt -= LT * w[P_C] * hzd;
        AP[P_D].tt += LT * w[P_C] * hzd;
        // conduction (D_x + D_y + D_z) R tk rk + dT/dz R tk dk
        const double sk = R * C.tk * m.rk;
        AP[P_E].tt += LT * sk * (lx1 + lx2);
        AP[P_W].tt += LT * sk * (-lx1 + lx2);
        AP[P_N].tt += LT * sk * (tmpy1 + tmpy2);
        AP[P_S].tt += LT * sk * (-tmpy1 + tmpy2);
        AP[P_U].tt += LT * sk * (tmpz1 + tmpz2);
        AP[P_D].tt += LT * sk * (-tmpz1 + tmpz2);
        AP[P_C].tt += -2.0 * LT * sk * (lx2 + tmpy2 + tmpz2);
        const double sd = R * C.tk * m.dk * hzd;
        AP[P_U].tt += LT * sd;
        AP[P_D].tt -= LT * sd;
        const double dsum = (ttE - ttW) * lx1 + (ttE - 2.0 * tt0 + ttW) * lx2 +
                            (ttN - ttS) * tmpy1 + (ttN - 2.0 * tt0 + ttS) * tmpy2 +
                            (ttU - ttD) * tmpz1 + (ttU - 2.0 * tt0 + ttD) * tmpz2;
        const double cond = dsum * m.rk + (ttU - ttD) * hzd * m.dk;
        ari += LT * C.tk * cond;
        g_tk += LT * R * cond;
        // viscous heating and compression (velocity gradients at the cell)
        const double ux = (u[P_E] - u[P_W]) * hxd, vx = (v[P_E] - v[P_W]) * hxd;
        const double wx = (w[P_E] - w[P_W]) * hxd;
        const double uy = (u[P_N] - u[P_S]) * hyd, vy = (v[P_N] - v[P_S]) * hyd;
        const double wy = (w[P_N] - w[P_S]) * hyd;
        const double uz = (u[P_U] - u[P_D]) * hzd, vz = (v[P_U] - v[P_D]) * hzd;
        const double wz = (w[P_U] - w[P_D]) * hzd;
        const double t1 = C.t1, t3 = C.t3, ocv = C.ocv, c43 = C.c43;
        const double q1 = c43 * ux * ux + vx * vx + wx * wx + uy * uy + 2.0 * uy * vx + uz * uz +
                          vz * vz + wy * wy + 2.0 * vz * wy + 2.0 * uz * wx;
        const double q3 = vy * vy + wz * wz - ux * wz - ux * vy - vy * wz;
        double aux = LT * (2.0 * c43 * ux * R * t1 - ocv * tt0 + (-wz - vy) * R * t3);
        double avx = LT * (2.0 * vx + 2.0 * uy) * R * t1;
        double awx = LT * (2.0 * wx + 2.0 * uz) * R * t1;
        double auy = LT * (2.0 * uy + 2.0 * vx) * R * t1;
        double avy = LT * ((2.0 * vy - ux - wz) * R * t3 - ocv * tt0);
        double awy = LT * (2.0 * wy + 2.0 * vz) * R * t1;
        double auz = LT * (2.0 * uz + 2.0 * wx) * R * t1;
        double avz = LT * (2.0 * vz + 2.0 * wy) * R * t1;
        double awz = LT * ((2.0 * wz - ux - vy) * R * t3 - ocv * tt0);
        AP[P_C].tt -= LT * ocv * (ux + vy + wz);
        ari += LT * (q1 * t1 + q3 * t3);
        // t1 = ocv ore, t3 = c43 ocv ore
        g_ore += LT * R * ocv * (q1 + c43 * q3);
        au[P_E] += aux * hxd; au[P_W] -= aux * hxd;
        av[P_E] += avx * hxd; av[P_W] -= avx * hxd;
        aw[P_E] += awx * hxd; aw[P_W] -= awx * hxd;
        au[P_N] += auy * hyd; au[P_S] -= auy * hyd;
        av[P_N] += avy * hyd; av[P_S] -= avy * hyd;
        aw[P_N] += awy * hyd; aw[P_S] -= awy * hyd;
        au[P_U] += auz * hzd; au[P_D] -= auz * hzd;
        av[P_U] += avz * hzd; av[P_D] -= avz * hzd;
        aw[P_U] += awz * hzd; aw[P_D] -= awz * hzd;
    }
    // ---- derived quantities back to the primaries
    for (int q = 0; q < NPTS; ++q) {
        AP[q].ru += ri[q] * au[q];
        AP[q].rv += ri[q] * av[q];
        AP[q].rw += ri[q] * aw[q];
        AP[q].ro += P[q].tt * ap[q] - ri[q] * (u[q] * au[q] + v[q] * av[q] + w[q] * aw[q]);
        AP[q].tt += P[q].ro * ap[q];
    }
    AP[P_C].ro -= ri[P_C] * ri[P_C] * ari;
}

// F of an MHD cell: the hydrodynamic F followed by the LMAG terms of the
// Fortran FLUXES (golden/fortran/3dmhdsub.f:2146-2651) in the stage kernel's
// operation order; F = (FU, FV, FW, FR, FT, WW1, WW2, WW3).
__host__ __device__ inline void cell_flux_mag(const Stencil& s, const MagStencil& b,
                                              const CellMet& m, const Consts& C, double F[8]) {
    cell_flux(s, m, C, F);
    double u[NPTS], v[NPTS], w[NPTS], bx[NPTS], by[NPTS], bz[NPTS];
    for (int q = 0; q < NPTS; ++q) {
        u[q] = s.p[q].u;
        v[q] = s.p[q].v;
        w[q] = s.p[q].w;
        bx[q] = b.p[q].bx;
        by[q] = b.p[q].by;
        bz[q] = b.p[q].bz;
    }
    const double R = s.p[P_C].ri;
    const double hxd = C.hx, dxx = m.dxx, d2xx = m.d2xx, dyy = m.dyy, d2yy = m.d2yy;
    const double dzz = m.dzz, d2zz = m.d2zz;
#define DXS(q) ((q[P_E] - q[P_W]) * hxd * dxx)
#define DYS(q) ((q[P_N] - q[P_S]) * C.hy * dyy)
#define DZC(q) ((q[P_U] - q[P_D]) * C.hz * dzz)
#define LAPX(q, c2) ((q[P_E] - q[P_W]) * hxd * d2xx + (q[P_E] - c2 + q[P_W]) * C.h2x * dxx * dxx)
#define LAPY(q, c2) ((q[P_N] - q[P_S]) * C.hy * d2yy + (q[P_N] - c2 + q[P_S]) * C.h2y * dyy * dyy)
#define LAPZ(q, c2) ((q[P_U] - q[P_D]) * C.hz * d2zz + (q[P_U] - c2 + q[P_D]) * C.h2z * dzz * dzz)
    {
        const double bxy = DYS(bx), byx = DXS(by), bzx = DXS(bz), bxz = DZC(bx);
        double fu = F[QRU];
        fu = fu + C.obeta * by[P_C] * (bxy - byx) - C.obeta * bz[P_C] * bzx;
        fu = fu + C.obeta * bz[P_C] * bxz;
        F[QRU] = fu;
    }
    {
        const double byx = DXS(by), bxy = DYS(bx), byz = DZC(by), bzy = DYS(bz);
        double fv = F[QRV];
        fv = fv + C.obeta * bx[P_C] * (byx - bxy);
        fv = fv + C.obeta * bz[P_C] * (byz - bzy);
        F[QRV] = fv;
    }
    {
        const double bzx = DXS(bz), bxz = DZC(bx), byz = DZC(by), bzy = DYS(bz);
        double fw = F[QRW];
        fw = fw + C.obeta * bx[P_C] * bzx;
        fw = fw - C.obeta * bx[P_C] * bxz;
        fw = fw - C.obeta * by[P_C] * (byz - bzy);
        F[QRW] = fw;
    }
    {
        const double bxy = DYS(bx), byx = DXS(by), bzx = DXS(bz), bxz = DZC(bx);
        const double byz = DZC(by), bzy = DYS(bz);
        double ft = F[QTT];
        ft = ft + (bxy * bxy + byx * byx + bzx * bzx - 2.0 * byx * bxy) * R * C.ohm;
        ft = ft + (bxz * bxz - 2.0 * bxz * bzx) * R * C.ohm;
        ft = ft + (byz * byz + bzy * bzy - 2.0 * byz * bzy) * R * C.ohm;
        F[QTT] = ft;
    }
    const bool top = m.wall == 1, bot = m.wall == 2 && m.mag_bot;
    const double w2 = (top || bot) ? s.rw2 * (1.0 / b.ro2) : 0.0;
    double wzm = DZC(w);
    if (top) wzm = (4.0 * w[P_U] - w2) * C.hz * dzz;
    if (bot) wzm = (w2 - 4.0 * w[P_D]) * C.hz * dzz;
    {
        const double b2c = 2.0 * bx[P_C];
        double lbz = LAPZ(bx, b2c);
        if (top)
            lbz = (-3.0 * bx[P_C] + 4.0 * bx[P_U] - b.bx2) * C.hz * d2zz +
                  (2.0 * bx[P_C] - 5.0 * bx[P_U] + 4.0 * b.bx2 - b.bx3) * C.h2z * dzz * dzz;
        if (bot)
            lbz = (3.0 * bx[P_C] - 4.0 * bx[P_D] + b.bx2) * C.hz * d2zz +
                  (2.0 * bx[P_C] - 5.0 * bx[P_D] + 4.0 * b.bx2 - b.bx3) * C.h2z * dzz * dzz;
        double w1 = by[P_C] * DYS(u) - bx[P_C] * DYS(v);
        w1 = w1 + bz[P_C] * DZC(u) - bx[P_C] * wzm;
        w1 = w1 - u[P_C] * DXS(bx) - v[P_C] * DYS(bx);
        w1 = w1 + C.orm * (LAPX(bx, b2c) + LAPY(bx, b2c));
        w1 = w1 - w[P_C] * DZC(bx);
        w1 = w1 + C.orm * lbz;
        F[QBX] = w1;
    }
    {
        const double b2c = 2.0 * by[P_C];
        double lbz = LAPZ(by, b2c);
        if (top)
            lbz = (-3.0 * by[P_C] + 4.0 * by[P_U] - b.by2) * C.hz * d2zz +
                  (2.0 * by[P_C] - 5.0 * by[P_U] + 4.0 * b.by2 - b.by3) * C.h2z * dzz * dzz;
        if (bot)
            lbz = (3.0 * by[P_C] - 4.0 * by[P_D] + b.by2) * C.hz * d2zz +
                  (2.0 * by[P_C] - 5.0 * by[P_D] + 4.0 * b.by2 - b.by3) * C.h2z * dzz * dzz;
        double w2f = bx[P_C] * DXS(v) - by[P_C] * DXS(u);
        w2f = w2f + bz[P_C] * DZC(v) - by[P_C] * wzm;
        w2f = w2f - u[P_C] * DXS(by);
        w2f = w2f + C.orm * LAPX(by, b2c);
        w2f = w2f - v[P_C] * DYS(by);
        w2f = w2f + C.orm * LAPY(by, b2c);
        w2f = w2f - w[P_C] * DZC(by);
        w2f = w2f + C.orm * lbz;
        F[QBY] = w2f;
    }
    {
        const double b2c = 2.0 * bz[P_C];
        double w3 = bx[P_C] * DXS(w) - bz[P_C] * DXS(u);
        w3 = w3 + by[P_C] * DYS(w) - bz[P_C] * DYS(v);
        w3 = w3 - u[P_C] * DXS(bz);
        w3 = w3 + C.orm * LAPX(bz, b2c);
        w3 = w3 - v[P_C] * DYS(bz);
        w3 = w3 + C.orm * LAPY(bz, b2c);
        w3 = w3 - w[P_C] * DZC(bz);
        w3 = w3 + C.orm * LAPZ(bz, b2c);
        F[QBZ] = w3;
    }
#undef DXS
#undef DYS
#undef DZC
#undef LAPX
#undef LAPY
#undef LAPZ
}

// Reverse of cell_flux_mag: lam[0..7] weights F; A (hydrodynamic primaries,
// A.rw2 included) and AB (B at the points and the wall extras) receive the
// derivatives; g_ore, g_tk as for cell_flux_vjp (the LMAG terms do not
// depend on Re).
// The LMAG terms alone (what cell_flux_mag adds to the hydrodynamic F, and
// the induction fluxes): E = (dFU, dFV, dFW, 0, dFT, WW1, WW2, WW3).  They use
// the cell and its six face neighbours only (points P_C..P_D) and the wall
// extras.  Used for <F, lamF>, so the sums are not ordered like the reference.
constexpr int NFACE = P_D + 1;
__host__ __device__ inline void cell_flux_mag_extra(const Stencil& s, const MagStencil& b,
                                                    const CellMet& m, const Consts& C,
                                                    double E[8]) {
    double u[NFACE], v[NFACE], w[NFACE], bx[NFACE], by[NFACE], bz[NFACE];
    for (int q = 0; q < NFACE; ++q) {
        u[q] = s.p[q].u;
        v[q] = s.p[q].v;
        w[q] = s.p[q].w;
        bx[q] = b.p[q].bx;
        by[q] = b.p[q].by;
        bz[q] = b.p[q].bz;
    }
    const double R = s.p[P_C].ri;
    const double hxd = C.hx * m.dxx, hyd = C.hy * m.dyy, hzd = C.hz * m.dzz;
    const double lx1 = C.hx * m.d2xx, lx2 = C.h2x * m.dxx * m.dxx;
    const double ly1 = C.hy * m.d2yy, ly2 = C.h2y * m.dyy * m.dyy;
    const double lz1 = C.hz * m.d2zz, lz2 = C.h2z * m.dzz * m.dzz;
    auto dx = [&](const double* q) { return (q[P_E] - q[P_W]) * hxd; };
    auto dy = [&](const double* q) { return (q[P_N] - q[P_S]) * hyd; };
    auto dz = [&](const double* q) { return (q[P_U] - q[P_D]) * hzd; };
    auto lxy = [&](const double* q) {
        return (q[P_E] - q[P_W]) * lx1 + (q[P_E] - 2.0 * q[P_C] + q[P_W]) * lx2 +
               (q[P_N] - q[P_S]) * ly1 + (q[P_N] - 2.0 * q[P_C] + q[P_S]) * ly2;
    };
    const double bxy = dy(bx), byx = dx(by), bzx = dx(bz), bxz = dz(bx), byz = dz(by);
    const double bzy = dy(bz);
    const double ob = C.obeta;
    E[QRU] = ob * by[P_C] * (bxy - byx) - ob * bz[P_C] * bzx + ob * bz[P_C] * bxz;
    E[QRV] = ob * bx[P_C] * (byx - bxy) + ob * bz[P_C] * (byz - bzy);
    E[QRW] = ob * bx[P_C] * bzx - ob * bx[P_C] * bxz - ob * by[P_C] * (byz - bzy);
    E[QRO] = 0.0;
    E[QTT] = (bxy * bxy + byx * byx + bzx * bzx - 2.0 * byx * bxy + bxz * bxz - 2.0 * bxz * bzx +
              byz * byz + bzy * bzy - 2.0 * byz * bzy) * R * C.ohm;
    const bool top = m.wall == 1, bot = m.wall == 2 && m.mag_bot;
    const double w2 = (top || bot) ? s.rw2 * (1.0 / b.ro2) : 0.0;
    double wzm = dz(w);
    if (top) wzm = (4.0 * w[P_U] - w2) * hzd;
    if (bot) wzm = (w2 - 4.0 * w[P_D]) * hzd;
    auto lz = [&](const double* q, double q2, double q3) {
        if (top)
            return (-3.0 * q[P_C] + 4.0 * q[P_U] - q2) * lz1 +
                   (2.0 * q[P_C] - 5.0 * q[P_U] + 4.0 * q2 - q3) * lz2;
        if (bot)
            return (3.0 * q[P_C] - 4.0 * q[P_D] + q2) * lz1 +
                   (2.0 * q[P_C] - 5.0 * q[P_D] + 4.0 * q2 - q3) * lz2;
        return (q[P_U] - q[P_D]) * lz1 + (q[P_U] - 2.0 * q[P_C] + q[P_D]) * lz2;
    };
    const double u0 = u[P_C], v0 = v[P_C], w0 = w[P_C];
    E[QBX] = by[P_C] * dy(u) - bx[P_C] * dy(v) + bz[P_C] * dz(u) - bx[P_C] * wzm - u0 * dx(bx) -
             v0 * dy(bx) - w0 * dz(bx) + C.orm * (lxy(bx) + lz(bx, b.bx2, b.bx3));
    E[QBY] = bx[P_C] * dx(v) - by[P_C] * dx(u) + bz[P_C] * dz(v) - by[P_C] * wzm - u0 * dx(by) -
             v0 * dy(by) - w0 * dz(by) + C.orm * (lxy(by) + lz(by, b.by2, b.by3));
    E[QBZ] = bx[P_C] * dx(w) - bz[P_C] * dx(u) + by[P_C] * dy(w) - bz[P_C] * dy(v) - u0 * dx(bz) -
             v0 * dy(bz) - w0 * dz(bz) +
             C.orm * (lxy(bz) + (bz[P_U] - bz[P_D]) * lz1 + (bz[P_U] - 2.0 * bz[P_C] + bz[P_D]) * lz2);
}

// Transpose of the LMAG terms: adds their adjoint into A at points
// P_C..P_D (and A.rw2) and sets AB (zero first).  With cell_flux_vjp before
// it this is the transpose of the whole MHD F.
__host__ __device__ inline void cell_flux_mag_extra_vjp(const Stencil& s, const MagStencil& b,
                                                        const CellMet& m, const Consts& C,
                                                        const double lam[8], Stencil& A,
                                                        MagStencil& AB) {
    // the LMAG terms reach the face points only
    double ri[NFACE], u[NFACE], v[NFACE], w[NFACE], bx[NFACE], by[NFACE], bz[NFACE];
    double au[NFACE], av[NFACE], aw[NFACE];
    for (int q = 0; q < NPTS; ++q) AB.p[q].bx = AB.p[q].by = AB.p[q].bz = 0.0;
    for (int q = 0; q < NFACE; ++q) {
        ri[q] = s.p[q].ri;
        u[q] = s.p[q].u;
        v[q] = s.p[q].v;
        w[q] = s.p[q].w;
        bx[q] = b.p[q].bx;
        by[q] = b.p[q].by;
        bz[q] = b.p[q].bz;
        au[q] = av[q] = aw[q] = 0.0;
    }
    AB.ro2 = AB.bx2 = AB.by2 = AB.bx3 = AB.by3 = 0.0;
    const double hxd = C.hx * m.dxx, hyd = C.hy * m.dyy, hzd = C.hz * m.dzz;
    const double lx1 = C.hx * m.d2xx, lx2 = C.h2x * m.dxx * m.dxx;
    const double ly1 = C.hy * m.d2yy, ly2 = C.h2y * m.dyy * m.dyy;
    const double lz1 = C.hz * m.d2zz, lz2 = C.h2z * m.dzz * m.dzz;
    const double LU = lam[QRU], LV = lam[QRV], LW = lam[QRW], LT = lam[QTT];
    const double L1 = lam[QBX], L2 = lam[QBY], L3 = lam[QBZ];
    const double ob = C.obeta, oh = C.ohm, R = ri[P_C];
    const double bx0 = bx[P_C], by0 = by[P_C], bz0 = bz[P_C];
    const double u0 = u[P_C], v0 = v[P_C], w0 = w[P_C];
    auto dx = [&](const double* q) { return (q[P_E] - q[P_W]) * hxd; };
    auto dy = [&](const double* q) { return (q[P_N] - q[P_S]) * hyd; };
    auto dz = [&](const double* q) { return (q[P_U] - q[P_D]) * hzd; };
    const double bxy = dy(bx), byx = dx(by), bzx = dx(bz), bxz = dz(bx), byz = dz(by);
    const double bzy = dy(bz), bxx = dx(bx), byy = dy(by), bzz = dz(bz);
    const double ux = dx(u), vx = dx(v), wx = dx(w), uy = dy(u), vy = dy(v), wy = dy(w);
    const double uz = dz(u), vz = dz(v);
    const bool top = m.wall == 1, bot = m.wall == 2 && m.mag_bot;
    const double ri2 = (top || bot) ? 1.0 / b.ro2 : 0.0;
    const double w2 = s.rw2 * ri2;
    double wzm = dz(w);
    if (top) wzm = (4.0 * w[P_U] - w2) * hzd;
    if (bot) wzm = (w2 - 4.0 * w[P_D]) * hzd;

    // adjoints of the first derivatives of B
    const double a_bxy = LU * ob * by0 - LV * ob * bx0 + LT * R * oh * (2.0 * bxy - 2.0 * byx) - L1 * v0;
    const double a_byx = -LU * ob * by0 + LV * ob * bx0 + LT * R * oh * (2.0 * byx - 2.0 * bxy) - L2 * u0;
    const double a_bzx = -LU * ob * bz0 + LW * ob * bx0 + LT * R * oh * (2.0 * bzx - 2.0 * bxz) - L3 * u0;
    const double a_bxz = LU * ob * bz0 - LW * ob * bx0 + LT * R * oh * (2.0 * bxz - 2.0 * bzx) - L1 * w0;
    const double a_byz = LV * ob * bz0 - LW * ob * by0 + LT * R * oh * (2.0 * byz - 2.0 * bzy) - L2 * w0;
    const double a_bzy = -LV * ob * bz0 + LW * ob * by0 + LT * R * oh * (2.0 * bzy - 2.0 * byz) - L3 * v0;
    const double a_bxx = -L1 * u0, a_byy = -L2 * v0, a_bzz = -L3 * w0;
    auto add_d = [&](double a, int pp, int pm, double h, double BPt::*f) {
        AB.p[pp].*f += a * h;
        AB.p[pm].*f -= a * h;
    };
    add_d(a_bxy, P_N, P_S, hyd, &BPt::bx);
    add_d(a_byx, P_E, P_W, hxd, &BPt::by);
    add_d(a_bzx, P_E, P_W, hxd, &BPt::bz);
    add_d(a_bxz, P_U, P_D, hzd, &BPt::bx);
    add_d(a_byz, P_U, P_D, hzd, &BPt::by);
    add_d(a_bzy, P_N, P_S, hyd, &BPt::bz);
    add_d(a_bxx, P_E, P_W, hxd, &BPt::bx);
    add_d(a_byy, P_N, P_S, hyd, &BPt::by);
    add_d(a_bzz, P_U, P_D, hzd, &BPt::bz);
    // resistive Laplacians (orm times each component's full Laplacian)
    const double Lb[3] = {L1 * C.orm, L2 * C.orm, L3 * C.orm};
    double BPt::*fld[3] = {&BPt::bx, &BPt::by, &BPt::bz};
    for (int c = 0; c < 3; ++c) {
        double BPt::*f = fld[c];
        const double L = Lb[c];
        AB.p[P_E].*f += L * (lx1 + lx2);
        AB.p[P_W].*f += L * (-lx1 + lx2);
        AB.p[P_N].*f += L * (ly1 + ly2);
        AB.p[P_S].*f += L * (-ly1 + ly2);
        AB.p[P_C].*f += L * (-2.0 * lx2 - 2.0 * ly2);
        if (c < 2 && top) {
            AB.p[P_C].*f += L * (-3.0 * lz1 + 2.0 * lz2);
            AB.p[P_U].*f += L * (4.0 * lz1 - 5.0 * lz2);
            (c == 0 ? AB.bx2 : AB.by2) += L * (-lz1 + 4.0 * lz2);
            (c == 0 ? AB.bx3 : AB.by3) += L * (-lz2);
        } else if (c < 2 && bot) {
            AB.p[P_C].*f += L * (3.0 * lz1 + 2.0 * lz2);
            AB.p[P_D].*f += L * (-4.0 * lz1 - 5.0 * lz2);
            (c == 0 ? AB.bx2 : AB.by2) += L * (lz1 + 4.0 * lz2);
            (c == 0 ? AB.bx3 : AB.by3) += L * (-lz2);
        } else {
            AB.p[P_U].*f += L * (lz1 + lz2);
            AB.p[P_D].*f += L * (-lz1 + lz2);
            AB.p[P_C].*f += L * (-2.0 * lz2);
        }
    }
    // B at the cell
    AB.p[P_C].bx += LV * ob * (byx - bxy) + LW * ob * (bzx - bxz) - L1 * vy - L1 * wzm + L2 * vx +
                    L3 * wx;
    AB.p[P_C].by += LU * ob * (bxy - byx) - LW * ob * (byz - bzy) + L1 * uy - L2 * ux - L2 * wzm +
                    L3 * wy;
    AB.p[P_C].bz += -LU * ob * bzx + LU * ob * bxz + LV * ob * (byz - bzy) + L1 * uz + L2 * vz -
                    L3 * ux - L3 * vy;
    // velocities: at the cell and through their derivatives
    au[P_C] += -L1 * bxx - L2 * byx - L3 * bzx;
    av[P_C] += -L1 * bxy - L2 * byy - L3 * bzy;
    aw[P_C] += -L1 * bxz - L2 * byz - L3 * bzz;
    const double a_uy = L1 * by0, a_vy = -L1 * bx0 - L3 * bz0, a_uz = L1 * bz0;
    const double a_wzm = -L1 * bx0 - L2 * by0, a_vx = L2 * bx0, a_ux = -L2 * by0 - L3 * bz0;
    const double a_vz = L2 * bz0, a_wx = L3 * bx0, a_wy = L3 * by0;
    au[P_N] += a_uy * hyd; au[P_S] -= a_uy * hyd;
    av[P_N] += a_vy * hyd; av[P_S] -= a_vy * hyd;
    au[P_U] += a_uz * hzd; au[P_D] -= a_uz * hzd;
    av[P_U] += a_vz * hzd; av[P_D] -= a_vz * hzd;
    av[P_E] += a_vx * hxd; av[P_W] -= a_vx * hxd;
    au[P_E] += a_ux * hxd; au[P_W] -= a_ux * hxd;
    aw[P_E] += a_wx * hxd; aw[P_W] -= a_wx * hxd;
    aw[P_N] += a_wy * hyd; aw[P_S] -= a_wy * hyd;
    double aw2 = 0.0;
    if (top) {
        aw[P_U] += 4.0 * a_wzm * hzd;
        aw2 -= a_wzm * hzd;
    } else if (bot) {
        aw2 += a_wzm * hzd;
        aw[P_D] -= 4.0 * a_wzm * hzd;
    } else {
        aw[P_U] += a_wzm * hzd;
        aw[P_D] -= a_wzm * hzd;
    }
    // Ohmic heating's 1/RO at the cell
    const double q1 = bxy * bxy + byx * byx + bzx * bzx - 2.0 * byx * bxy;
    const double q2 = bxz * bxz - 2.0 * bxz * bzx;
    const double q3 = byz * byz + bzy * bzy - 2.0 * byz * bzy;
    const double aR = LT * oh * (q1 + q2 + q3);
    // derived quantities back to the primaries
    for (int q = 0; q < NFACE; ++q) {
        A.p[q].ru += ri[q] * au[q];
        A.p[q].rv += ri[q] * av[q];
        A.p[q].rw += ri[q] * aw[q];
        A.p[q].ro -= ri[q] * (u[q] * au[q] + v[q] * av[q] + w[q] * aw[q]);
    }
    A.p[P_C].ro -= R * R * aR;
    A.rw2 += ri2 * aw2;
    AB.ro2 -= w2 * ri2 * aw2;
}

// The wall extras of the LMAG transpose: RW and RO two planes in, BX and BY
// two and three planes in (the one-sided stencils of a wall cell).
struct MagWallAdj {
    double rw2, ro2, bx2, by2, bx3, by3;
};

// cell_flux_mag_extra_vjp in point-major form, for kernels that add each
// adjoint entry into an accumulator as soon as it is known: the cell's
// coefficients (adjoints of the derivatives of B and of the velocities,
// the resistive Laplacian weights) first, then each face point's seven
// entries in closed form, so that no per-point adjoint array is held.
// pt(q, ri, u, v, w) returns 1/RO and the velocities at face point q (a
// kernel re-reads them where they are needed); out(q, f, value) receives the
// entry of primary f (QRU..QRO, QBX..QBZ; the temperature's is zero) at face
// point q, each (q, f) once.  The same entries as cell_flux_mag_extra_vjp,
// which sums some of them in another order.
template <class PT, class OUT>
__host__ __device__ inline void cell_flux_mag_extra_vjp_points(const Stencil& s,
                                                               const MagStencil& b,
                                                               const CellMet& m, const Consts& C,
                                                               const double lam[8], PT&& pt,
                                                               OUT&& out, MagWallAdj& W) {
    double u[NFACE], v[NFACE], w[NFACE], bx[NFACE], by[NFACE], bz[NFACE];
    for (int q = 0; q < NFACE; ++q) {
        u[q] = s.p[q].u;
        v[q] = s.p[q].v;
        w[q] = s.p[q].w;
        bx[q] = b.p[q].bx;
        by[q] = b.p[q].by;
        bz[q] = b.p[q].bz;
    }
    const double hxd = C.hx * m.dxx, hyd = C.hy * m.dyy, hzd = C.hz * m.dzz;
    const double lx1 = C.hx * m.d2xx, lx2 = C.h2x * m.dxx * m.dxx;
    const double ly1 = C.hy * m.d2yy, ly2 = C.h2y * m.dyy * m.dyy;
    const double lz1 = C.hz * m.d2zz, lz2 = C.h2z * m.dzz * m.dzz;
    const double LU = lam[QRU], LV = lam[QRV], LW = lam[QRW], LT = lam[QTT];
    const double L1 = lam[QBX], L2 = lam[QBY], L3 = lam[QBZ];
    const double ob = C.obeta, oh = C.ohm, R = s.p[P_C].ri;
    const double bx0 = bx[P_C], by0 = by[P_C], bz0 = bz[P_C];
    const double u0 = u[P_C], v0 = v[P_C], w0 = w[P_C];
    auto dx = [&](const double* q) { return (q[P_E] - q[P_W]) * hxd; };
    auto dy = [&](const double* q) { return (q[P_N] - q[P_S]) * hyd; };
    auto dz = [&](const double* q) { return (q[P_U] - q[P_D]) * hzd; };
    const double bxy = dy(bx), byx = dx(by), bzx = dx(bz), bxz = dz(bx), byz = dz(by);
    const double bzy = dy(bz), bxx = dx(bx), byy = dy(by), bzz = dz(bz);
    const double ux = dx(u), vx = dx(v), wx = dx(w), uy = dy(u), vy = dy(v), wy = dy(w);
    const double uz = dz(u), vz = dz(v);
    const bool top = m.wall == 1, bot = m.wall == 2 && m.mag_bot;
    const double ri2 = (top || bot) ? 1.0 / b.ro2 : 0.0;
    const double w2 = s.rw2 * ri2;
    double wzm = dz(w);
    if (top) wzm = (4.0 * w[P_U] - w2) * hzd;
    if (bot) wzm = (w2 - 4.0 * w[P_D]) * hzd;

    // adjoints of the first derivatives of B
    const double a_bxy = LU * ob * by0 - LV * ob * bx0 + LT * R * oh * (2.0 * bxy - 2.0 * byx) - L1 * v0;
    const double a_byx = -LU * ob * by0 + LV * ob * bx0 + LT * R * oh * (2.0 * byx - 2.0 * bxy) - L2 * u0;
    const double a_bzx = -LU * ob * bz0 + LW * ob * bx0 + LT * R * oh * (2.0 * bzx - 2.0 * bxz) - L3 * u0;
    const double a_bxz = LU * ob * bz0 - LW * ob * bx0 + LT * R * oh * (2.0 * bxz - 2.0 * bzx) - L1 * w0;
    const double a_byz = LV * ob * bz0 - LW * ob * by0 + LT * R * oh * (2.0 * byz - 2.0 * bzy) - L2 * w0;
    const double a_bzy = -LV * ob * bz0 + LW * ob * by0 + LT * R * oh * (2.0 * bzy - 2.0 * byz) - L3 * v0;
    const double a_bxx = -L1 * u0, a_byy = -L2 * v0, a_bzz = -L3 * w0;
    // resistive Laplacian weights (orm times each component's adjoint)
    const double Lb0 = L1 * C.orm, Lb1 = L2 * C.orm, Lb2 = L3 * C.orm;
    // adjoints of the velocity derivatives
    const double a_uy = L1 * by0, a_vy = -L1 * bx0 - L3 * bz0, a_uz = L1 * bz0;
    const double a_wzm = -L1 * bx0 - L2 * by0, a_vx = L2 * bx0, a_ux = -L2 * by0 - L3 * bz0;
    const double a_vz = L2 * bz0, a_wx = L3 * bx0, a_wy = L3 * by0;
    // Ohmic heating's 1/RO at the cell
    const double q1 = bxy * bxy + byx * byx + bzx * bzx - 2.0 * byx * bxy;
    const double q2 = bxz * bxz - 2.0 * bxz * bzx;
    const double q3 = byz * byz + bzy * bzy - 2.0 * byz * bzy;
    const double aR = LT * oh * (q1 + q2 + q3);

    // one face point: velocity adjoints to the momenta and the density with
    // the point's own 1/RO and velocities, and the adjoint of B
    auto point = [&](int q, double au, double av, double aw, double abx, double aby,
                     double abz, double ro_extra) {
        double ri, uq, vq, wq;
        pt(q, ri, uq, vq, wq);
        out(q, QRU, ri * au);
        out(q, QRV, ri * av);
        out(q, QRW, ri * aw);
        out(q, QRO, -(ri * (uq * au + vq * av + wq * aw)) + ro_extra);
        out(q, QBX, abx);
        out(q, QBY, aby);
        out(q, QBZ, abz);
    };
    // the cell itself: B of the induction and Lorentz terms, the velocities
    // of the induction terms, every Laplacian's centre weight
    {
        const double cxy = -2.0 * lx2 - 2.0 * ly2;
        const double czb = top ? -3.0 * lz1 + 2.0 * lz2 : (bot ? 3.0 * lz1 + 2.0 * lz2 : -2.0 * lz2);
        const double abx = Lb0 * cxy + Lb0 * czb + (LV * ob * (byx - bxy) + LW * ob * (bzx - bxz) -
                                                    L1 * vy - L1 * wzm + L2 * vx + L3 * wx);
        const double aby = Lb1 * cxy + Lb1 * czb + (LU * ob * (bxy - byx) - LW * ob * (byz - bzy) +
                                                    L1 * uy - L2 * ux - L2 * wzm + L3 * wy);
        const double abz = Lb2 * cxy + Lb2 * (-2.0 * lz2) +
                           (-LU * ob * bzx + LU * ob * bxz + LV * ob * (byz - bzy) + L1 * uz +
                            L2 * vz - L3 * ux - L3 * vy);
        point(P_C, -L1 * bxx - L2 * byx - L3 * bzx, -L1 * bxy - L2 * byy - L3 * bzy,
              -L1 * bxz - L2 * byz - L3 * bzz, abx, aby, abz, -(R * R * aR));
    }
    const double le = lx1 + lx2, lw = -lx1 + lx2, ln = ly1 + ly2, ls = -ly1 + ly2;
    point(P_E, a_ux * hxd, a_vx * hxd, a_wx * hxd, a_bxx * hxd + Lb0 * le, a_byx * hxd + Lb1 * le,
          a_bzx * hxd + Lb2 * le, 0.0);
    point(P_W, -(a_ux * hxd), -(a_vx * hxd), -(a_wx * hxd), -(a_bxx * hxd) + Lb0 * lw,
          -(a_byx * hxd) + Lb1 * lw, -(a_bzx * hxd) + Lb2 * lw, 0.0);
    point(P_N, a_uy * hyd, a_vy * hyd, a_wy * hyd, a_bxy * hyd + Lb0 * ln, a_byy * hyd + Lb1 * ln,
          a_bzy * hyd + Lb2 * ln, 0.0);
    point(P_S, -(a_uy * hyd), -(a_vy * hyd), -(a_wy * hyd), -(a_bxy * hyd) + Lb0 * ls,
          -(a_byy * hyd) + Lb1 * ls, -(a_bzy * hyd) + Lb2 * ls, 0.0);
    // vertical points: the one-sided wall stencils of W (for WZ) and of the
    // Laplacians of BX and BY; BZ's Laplacian is centred everywhere
    {
        const double awu = top ? 4.0 * a_wzm * hzd : (bot ? 0.0 : a_wzm * hzd);
        const double lu = top ? 4.0 * lz1 - 5.0 * lz2 : lz1 + lz2;
        point(P_U, a_uz * hzd, a_vz * hzd, awu, a_bxz * hzd + (bot ? 0.0 : Lb0 * lu),
              a_byz * hzd + (bot ? 0.0 : Lb1 * lu), a_bzz * hzd + Lb2 * (lz1 + lz2), 0.0);
        const double awd = top ? 0.0 : (bot ? -(4.0 * a_wzm * hzd) : -(a_wzm * hzd));
        const double ld = bot ? -4.0 * lz1 - 5.0 * lz2 : -lz1 + lz2;
        point(P_D, -(a_uz * hzd), -(a_vz * hzd), awd, -(a_bxz * hzd) + (top ? 0.0 : Lb0 * ld),
              -(a_byz * hzd) + (top ? 0.0 : Lb1 * ld), -(a_bzz * hzd) + Lb2 * (-lz1 + lz2), 0.0);
    }
    const double aw2 = top ? -(a_wzm * hzd) : (bot ? a_wzm * hzd : 0.0);
    W.rw2 = ri2 * aw2;
    W.ro2 = -(w2 * ri2 * aw2);
    const double l2 = top ? -lz1 + 4.0 * lz2 : lz1 + 4.0 * lz2;
    W.bx2 = (top || bot) ? Lb0 * l2 : 0.0;
    W.by2 = (top || bot) ? Lb1 * l2 : 0.0;
    W.bx3 = (top || bot) ? Lb0 * (-lz2) : 0.0;
    W.by3 = (top || bot) ? Lb1 * (-lz2) : 0.0;
}

// Transpose of the whole MHD F (cell_flux_mag): the hydrodynamic part, then
// the LMAG terms.
__host__ __device__ inline void cell_flux_mag_vjp(const Stencil& s, const MagStencil& b,
                                                  const CellMet& m, const Consts& C,
                                                  const double lam[8], Stencil& A,
                                                  MagStencil& AB, double& g_ore, double& g_tk) {
    cell_flux_vjp(s, m, C, lam, A, g_ore, g_tk);
    cell_flux_mag_extra_vjp(s, b, m, C, lam, A, AB);
}

}  // namespace r3d

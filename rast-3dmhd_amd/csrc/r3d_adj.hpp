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

struct Pt {
    double ru, rv, rw, ro, tt;
};

struct Stencil {
    Pt p[NPTS];
    double rw2;  // RW two planes in: k+2 on the top wall, k-2 on the bottom wall
};

// Metrics of the cell's column and plane, and which wall it is on.
struct CellMet {
    double dxx, d2xx, dyy, d2yy, dzz, d2zz, rk, dk;
    int wall;  // 0 interior, 1 top (k = 1), 2 bottom (k = npz)
};

// F = (FU, FV, FW, FR, FT) in the order of the primaries QRU..QTT.
__host__ __device__ inline void cell_flux(const Stencil& s, const CellMet& m, const Consts& C,
                                          double F[5]) {
    double ri[NPTS], u[NPTS], v[NPTS], w[NPTS], pr[NPTS];
    for (int q = 0; q < NPTS; ++q) {
        ri[q] = 1.0 / s.p[q].ro;
        u[q] = s.p[q].ru * ri[q];
        v[q] = s.p[q].rv * ri[q];
        w[q] = s.p[q].rw * ri[q];
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
        ri[q] = 1.0 / s.p[q].ro;
        u[q] = s.p[q].ru * ri[q];
        v[q] = s.p[q].rv * ri[q];
        w[q] = s.p[q].rw * ri[q];
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

}  // namespace r3d

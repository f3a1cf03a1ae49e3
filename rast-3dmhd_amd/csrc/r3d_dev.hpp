// r3d_dev.hpp - device-side data structures shared by the gfx950 kernels.
//
// Global layout of every 3-D field: one padded array per field covering the
// whole domain with one ghost layer, x fastest.  Element (i, j, k) with
// i in [0, npx+1], j in [0, npy+1], k in [0, npz+1] lives at
//     XOFF + i + px*j + pp*k,   px = row pitch (multiple of 16 doubles),
//                               pp = px*(npy+2)
// XOFF = 15 puts the first interior column (i = 1) on a 128-byte boundary, so
// a wavefront's 64 consecutive interior doubles are exactly four cache lines.
#pragma once

#include <hip/hip_runtime.h>

namespace r3d {

constexpr int XOFF = 15;
constexpr int NXF = 8;  // primaries per state: ru rv rw ro tt bx by bz

// Indices of the primaries inside a state.
enum { QRU = 0, QRV, QRW, QRO, QTT, QBX, QBY, QBZ };

// Plume / boundary branch codes for the top plane (reference src/bcon.cpp:43-94).
enum { TOP_GAUSS_T = 0, TOP_GAUSS_FLUX = 1, TOP_CONST_T = 2, TOP_CONST_FLUX = 3, TOP_NONE = 4 };

struct Geo {
    int npx, npy, npz;
    long px, pp;
    __host__ __device__ long at(int i, int j, int k) const {
        return XOFF + i + px * j + pp * (long)k;
    }
};

// Scalars of the run that live on the device, so that a step needs no host
// round trip (dt of step n+1 is reduced at the end of step n).
struct DevScal {
    double dt, timc, timt, timi, umach;
    double dt_next, umach_next;
    int nit, pad;
    // the last reduction's extrema: the least time-step limit over all cells
    // and limits, and the largest ogamma*UU/T; they set the next screen (below)
    double lim_prev, um_prev;
    int dt_redo;  // the screened reduction could not prove its extrema: redo exactly
    int pad2;
};

// Time-step screen of stage 3.  A cell's limits can only set the next dt if
// one of them is at most the previous step's least limit times (1 + margin),
// and its ogamma*UU/T can only set umach if it is at least the previous
// largest times (1 - margin).  Cells that pass neither bound (tested with
// products only, no division or square root) skip the exact evaluation; the
// reduction then proves that its extrema are the exact ones, or has the step
// redo them over every cell.  The margin is Consts::scr_up - 1.

// Constants of the step: the reference's Derived/Params values that the
// kernels read, combined exactly as the reference combines them
// (e.g. t1 = ocv*ore as in src/fluxes.cpp:344).
struct Consts {
    double c13, c23, c43, ore, ocv, grav;
    double hx, h2x, hy, h2y, hz, h2z;
    double omx, omz;
    double tk;   // ocv/repr          (src/fluxes.cpp:247)
    double t1;   // ocv*ore           (src/fluxes.cpp:344)
    double t2;   // c13*ore           (src/fluxes.cpp:362)
    double t3;   // c43*ocv*ore       (src/fluxes.cpp:399)
    double og;   // 1/gamma           (src/step.cpp:119)
    double gamma, re, repr, cv, sff, rm, obeta, orm;
    double ohm;  // obeta*orm*ocv     (golden/fortran/3dmhdsub.f:2307)
    double theta, tb, hh, tpr, tp;
    double gam[3], zeta[2];
    double dtfix;  // > 0: every step takes this dt instead of the CFL one (adjoint runs)
    // ITCON=2 embedded heat loss amplitude TP*OCV*ramp(TIMT) of this step
    // (golden/fortran/3dmhdsub.f:1799-1822); read by the HEAT stage kernels only
    double hloss;
    // density diffusion (ID != 0): the horizontal and vertical terms are on
    // (DH != 0, DV != 0); read by the DEN stage kernels only
    int den_h, den_v;
    int rlx;  // flux relaxation on: FT gains the per-plane RLAX*TFAC*CORRECT (DEN kernels)
    // LREM (DEN kernels): FT diffuses TT - TTM(K) with unit conductivity and
    // gains the mean profile's radiative heating (golden/fortran/3dmhdsub.f:1551-1596);
    // BCON takes TU, DZTU, DZTB (:2852-2952)
    int lrem;
    double tu, dztu, dztb;
    double scr_up, scr_dn;  // time-step screen: 1 + margin, 1 - margin
    int lrot, lmag, top_mode, half_t, izcon, bz_zero;
    int mag_bot;  // LMAG one-sided z stencils on the bottom wall (npez > 1,
                  // golden/fortran/3dmhdsub.f:1126, 1392, 1513 sit in an ELSE IF)
};

// Metric arrays on the device (global indexing, see r3d_hip.h).
struct Metrics {
    const double *dxx, *d2xx, *ddx;
    const double *dyy, *d2yy, *ddy;
    const double *dzz, *d2zz, *ddz, *rkapa, *dkapa;
    const double *e1, *e2;  // plume Gaussians exp(...) per x / per y (host libm)
    // ITCON=2 heat-loss Gaussians exp(CLN*(EXX-XP)**2), ... per x / y / z (host libm)
    const double *hgx, *hgy, *hgz;
    // density diffusion per global plane (0..npz+1), rewritten before every
    // stage: ROM (horizontal mean of RO), DH*ORE*WWZ and DV*ORE*WWZ
    const double *rom, *dth, *dtv;
    // flux relaxation per global plane: RLAX*TFAC*CORRECT(K), rewritten
    // before every stage; and the plane coordinates
    const double *ftc, *zee;
    // LREM per global plane, rewritten before every stage: TTM (horizontal
    // mean of TT) and the radiative heating HRAD(K) of the mean profile
    const double *ttm, *hrad;
    // conductivity in the thermal time-step limit: rkapa, or 1 + rkapa with
    // LREM (golden/fortran/3dmhdsub.f:845-851)
    const double* rkdt;
    double dz_top, dz_bot;  // f4(1/(npz-1)) / dzzdz at the boundary planes
};

struct State {
    double* q[NXF];
};

// Arguments of one RK stage.
struct StageArgs {
    Geo g;
    State xin, xout, z;
    double* partial;   // per-block dt partial extrema (stage 3)
    DevScal* scal;
    int stage;         // 0, 1, 2
    int zmode;         // 0: Z' = mix (X' + zeta dt F); 1: Z' = X (save_state); 2: no Z write
    int want_dt;       // stage 3 of a full step: emit dt partials of X'
    int screen;        // stage 3: screen cells against the previous extrema (DevScal)
    int ntx, nty, nkc, kc, nblocks;
};

// Number of dt partial slots per block: max of (ogamma uu/tt), min of the
// advective, thermal, viscous and (MHD) resistive limits.
constexpr int NPART = 5;

}  // namespace r3d

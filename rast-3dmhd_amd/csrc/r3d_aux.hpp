// r3d_aux.hpp - boundary-plane, dt and bookkeeping kernels (gfx950).
//
// k_bound   : per stage, after k_stage.  BCON on the global top (k = 1) and
//             bottom (k = npz) planes (reference src/bcon.cpp:12-141, with
//             BZ = 0 under LMAG), the
//             deferred Z' mix on those planes, their dt limits in the last
//             stage, and the
//             1/RO flip of the two outer ghost planes that the reference's
//             FLUXES performs on its full array (src/fluxes.cpp:133-138) but
//             its halo exchange never refreshes.
// halo_thread: periodic x/y ghost images of a state, written before a download
//             (the step itself reads wrapped interior cells instead).
// k_dt      : dt limits of the current state when no fused value is
//             available (first step, after an upload or a truncated step).
// k_finish  : reduces the dt partials and advances the step counters.
#pragma once

#include "r3d_stage.hpp"

namespace R3D_KNS {

constexpr int BOUND_THREADS = 256;
constexpr int DT_THREADS = 256;

// Block bx of nbb (per side) of the boundary pass, NT threads a block; side
// 0 = top, 1 = bottom.  Its dt partial goes to slot partial_base + side * nbb
// + bx.
template <int NT = BOUND_THREADS>
__device__ __forceinline__ void bound_block(const StageArgs& a, const Consts& C, const Metrics& M,
                                            int partial_base, int bx, int nbb, int side) {
    __shared__ double red[NT / 64][NPART];
    const Geo g = a.g;
    const int W = g.npx + 2, H = g.npy + 2;
    const int idx = bx * NT + threadIdx.x;
    double pm[NPART] = {0.0, 1e300, 1e300, 1e300, 1e300};
    const double dt = a.scal->dt;
    const double cz = (a.stage < 2) ? C.zeta[a.stage] * dt : 0.0;

    if (idx < W * H) {
        const int i = idx % W, j = idx / W;
        // outer ghost plane: RO <- 1/RO over the whole plane
        const long no = g.at(i, j, side ? g.npz + 1 : 0);
        a.xout.q[QRO][no] = 1.0 / a.xin.q[QRO][no];

        if (i >= 1 && i <= g.npx && j >= 1 && j <= g.npy) {
            const int kb = side ? g.npz : 1;
            const long n = g.at(i, j, kb);
            const long d1 = side ? -g.pp : g.pp;
            double* const* X = a.xout.q;
            const double ro = X[QRO][n], ro1 = X[QRO][n + d1], ro2 = X[QRO][n + 2 * d1];
            const double ru = ro * (C.c43 * X[QRU][n + d1] / ro1 - C.c13 * X[QRU][n + 2 * d1] / ro2);
            const double rv = ro * (C.c43 * X[QRV][n + d1] / ro1 - C.c13 * X[QRV][n + 2 * d1] / ro2);
            const double rw = 0.0;
            double tt = X[QTT][n];
            const double tt1 = X[QTT][n + d1], tt2 = X[QTT][n + 2 * d1];
            if (side == 0) {
                switch (C.top_mode) {
                    case TOP_GAUSS_T:
                        tt = 1.0 - (1.0 - C.tpr) * M.e1[i] / C.hh * M.e2[j] / C.hh;
                        break;
                    case TOP_GAUSS_FLUX:
                        tt = C.c43 * tt1 - C.c13 * tt2 -
                             C.c23 * M.dz_top *
                                 (C.theta - (C.theta - C.tpr) * M.e1[i] / C.hh * M.e2[j] / C.hh);
                        break;
                    case TOP_CONST_T:  // LREM: TU (golden/fortran/3dmhdsub.f:2852-2853)
                        tt = C.lrem ? C.tu : 1.0;
                        if (!C.lrem && C.half_t && i <= (g.npx + 2) / 2) tt = 1.0 - C.tp;
                        break;
                    case TOP_CONST_FLUX:  // LREM: + DZTU (:2876-2879)
                        tt = C.lrem ? C.c43 * tt1 - C.c13 * tt2 + C.dztu
                                    : C.c43 * tt1 - C.c13 * tt2 - C.c23 * M.dz_top * C.theta;
                        break;
                    default:
                        break;
                }
            } else {
                if (C.izcon == 0)
                    tt = C.tb;
                else if (C.lrem)  // + DZTB (golden/fortran/3dmhdsub.f:2949-2952)
                    tt = C.c43 * tt1 - C.c13 * tt2 + C.dztb;
                else
                    tt = C.c43 * tt1 - C.c13 * tt2 + C.c23 * M.dz_bot * C.theta / M.rkapa[kb];
            }
            X[QRU][n] = ru;
            X[QRV][n] = rv;
            X[QRW][n] = rw;
            X[QTT][n] = tt;
            // LMAG: BZ = 0 on both walls (src/bcon.cpp:95-98, 136-139); BX, BY
            // keep their updated values
            double bx = 0.0, by = 0.0, bz = 0.0;
            if (C.lmag) {
                bx = X[QBX][n];
                by = X[QBY][n];
                bz = C.bz_zero ? 0.0 : X[QBZ][n];
                X[QBZ][n] = bz;
            }

            if (a.zmode == 0 && a.stage < 2) {
                // mix for the next stage with the post-BCON state; the stage
                // kernel left F in Z on this plane
                double* const* Z = a.z.q;
                Z[QRU][n] = ru + cz * Z[QRU][n];
                Z[QRV][n] = rv + cz * Z[QRV][n];
                Z[QRW][n] = rw + cz * Z[QRW][n];
                Z[QRO][n] = ro + cz * Z[QRO][n];
                Z[QTT][n] = tt + cz * Z[QTT][n];
                if (C.lmag) {
                    Z[QBX][n] = bx + cz * Z[QBX][n];
                    Z[QBY][n] = by + cz * Z[QBY][n];
                    Z[QBZ][n] = bz + cz * Z[QBZ][n];
                }
            }
            if (a.stage == 2 && a.want_dt) {
                const double dd = dmin(dmin(M.ddx[i], M.ddy[j]), M.ddz[kb]);
                if (C.lmag)
                    dt_terms<true>(C, ru, rv, rw, ro, tt, bx, by, bz, dd, M.rkdt[kb], pm);
                else
                    dt_terms<false>(C, ru, rv, rw, ro, tt, 0.0, 0.0, 0.0, dd, M.rkdt[kb], pm);
            }
        }
    }
    if (a.stage == 2 && a.want_dt)
        block_partials<NT / 64>(pm, red, a.partial, partial_base + side * nbb + bx);
}

// dt limits of the state `s` (reference src/step.cpp:121-209): block bx of
// nbx (NT threads each), one partial per block.
template <int NT = DT_THREADS>
__device__ __forceinline__ void dt_block(const Geo& g, const State& s, const Consts& C,
                                         const Metrics& M, double* partial, int bx, int nbx) {
    __shared__ double red[NT / 64][NPART];
    double pm[NPART] = {0.0, 1e300, 1e300, 1e300, 1e300};
    const long total = (long)g.npx * g.npy * g.npz;
    for (long q = (long)bx * NT + threadIdx.x; q < total; q += (long)nbx * NT) {
        const int i = 1 + (int)(q % g.npx);
        const long rest = q / g.npx;
        const int j = 1 + (int)(rest % g.npy);
        const int k = 1 + (int)(rest / g.npy);
        const long n = g.at(i, j, k);
        const double dd = dmin(dmin(M.ddx[i], M.ddy[j]), M.ddz[k]);
        if (C.lmag)
            dt_terms<true>(C, s.q[QRU][n], s.q[QRV][n], s.q[QRW][n], s.q[QRO][n], s.q[QTT][n],
                           s.q[QBX][n], s.q[QBY][n], s.q[QBZ][n], dd, M.rkdt[k], pm);
        else
            dt_terms<false>(C, s.q[QRU][n], s.q[QRV][n], s.q[QRW][n], s.q[QRO][n], s.q[QTT][n],
                            0.0, 0.0, 0.0, dd, M.rkdt[k], pm);
    }
    block_partials<NT / 64>(pm, red, partial, bx);
}


// Periodic x/y ghost images on planes 1..npz (what the reference's halo
// exchange, src/communicate.cpp:70-169, leaves in the ghost cells of one
// global domain): every ghost cell takes the interior cell it wraps to.
// grid = (perimeter blocks, npz).
constexpr int HALO_THREADS = 256;
__device__ __forceinline__ void halo_thread(const Geo& g, const State& x, int nf, int q, int k) {
    const int W = g.npx + 2;
    int i, j;
    if (q < 2 * W) {  // ghost rows j = 0 and j = npy+1, all i
        i = q % W;
        j = q < W ? 0 : g.npy + 1;
    } else if (q < 2 * W + 2 * g.npy) {  // ghost columns i = 0 and i = npx+1, interior j
        const int m = q - 2 * W;
        j = 1 + m % g.npy;
        i = m < g.npy ? 0 : g.npx + 1;
    } else {
        return;
    }
    const int si = i == 0 ? g.npx : i == g.npx + 1 ? 1 : i;
    const int sj = j == 0 ? g.npy : j == g.npy + 1 ? 1 : j;
    const long d = g.at(i, j, k), s = g.at(si, sj, k);
    for (int f = 0; f < nf; ++f) x.q[f][d] = x.q[f][s];
}

// One block of NT threads.  reduce: fold partial[0..npart) into dt_next /
// umach_next (dt = sff * min of the limits, src/step.cpp:224-229; umach =
// sqrt(max), src/step.cpp:141).  advance: the end-of-step counters
// (src/step.cpp:254-256).
// screened: the partials of stage 3 were screened against sc->lim_prev /
// um_prev; the extrema found are exact only if they reach the screen's
// thresholds (the limit at most lim_prev*scr_up, ogamma*UU/T at least
// um_prev*scr_dn): a cell with a smaller limit or a larger ogamma*UU/T would
// have passed the screen.  Otherwise dt_redo is set and dt_next is left to
// the exact pass.
template <int NT = 1024>
__device__ __forceinline__ void finish_block(DevScal* sc, const double* partial, int npart,
                                             const Consts& C, int reduce, int advance,
                                             int screened = 0) {
    __shared__ double red[NT / 64][NPART];
    double pm[NPART] = {0.0, 1e300, 1e300, 1e300, 1e300};
    if (reduce)
        for (int q = threadIdx.x; q < npart; q += NT) {
            pm[0] = dmax(pm[0], partial[(long)q * NPART]);
            for (int v = 1; v < NPART; ++v) pm[v] = dmin(pm[v], partial[(long)q * NPART + v]);
        }
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    double v[NPART];
    v[0] = wave_max(pm[0]);
    for (int q = 1; q < NPART; ++q) v[q] = wave_min(pm[q]);
    if (ln == 0)
        for (int q = 0; q < NPART; ++q) red[wv][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        if (advance) {
            sc->nit += 1;
            sc->timc += sc->dt;
            sc->timt = sc->timi + sc->timc;
        }
        if (reduce) {
            double o[NPART];
            for (int q = 0; q < NPART; ++q) o[q] = red[0][q];
            for (int w = 1; w < NT / 64; ++w) {
                o[0] = dmax(o[0], red[w][0]);
                for (int q = 1; q < NPART; ++q) o[q] = dmin(o[q], red[w][q]);
            }
            double m = dmin(dmin(o[1], o[2]), o[3]);
            if (C.lmag) m = dmin(m, o[4]);
            const bool proven =
                !screened || screen_proven(C, m, o[0], sc->lim_prev, sc->um_prev);
            sc->dt_redo = proven ? 0 : 1;
            if (proven) {
                sc->umach_next = sqrt(o[0]);
                sc->dt_next = C.dtfix > 0.0 ? C.dtfix : C.sff * m;
                sc->lim_prev = m;
                sc->um_prev = o[0];
            }
        }
    }
}


}  // namespace R3D_KNS

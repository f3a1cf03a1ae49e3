// r3d_adjoint.hip - discrete adjoint of the hydrodynamic time step
// (include/r3d_adj.h).
//
// Forward step n from X_n with step size dt_n = sff * min(limits(X_n)):
//     X1 = B(X_n + g1 dt F(X_n)),   Z1 = X1 + z1 dt F(X_n)
//     X2 = B(Z1  + g2 dt F(X1)),    Z2 = X2 + z2 dt F(X1)
//     X_{n+1} = B(Z2 + g3 dt F(X2))
// where B is the boundary pass (reference src/bcon.cpp) applied to the wall
// planes.  The backward pass runs this in reverse for each step, with
//     lamX <- B^T lamX,   lamF = g dt lamX (+ z dt lamZ),   lamX += (dF/dX)^T lamF
// (the flux transpose is cell_flux_vjp of r3d_adj.hpp, scattered to the 19
// stencil points), accumulates lam_dt = sum over the stages of <F, lamF>/dt,
// and adds lam_dt * d(dt)/dX_n at the cell and limit that attain the
// minimum.  Re enters through ore = 1/Re and tk = ocv/(Re Pr) in the fluxes
// and through the thermal and viscous dt limits; theta through the boundary
// heat-flux conditions.
//
// This file is compiled with -ffp-contract=fast (the transposes need no
// reference rounding); the objective and the dt argmin, which must round as
// the forward does, switch contraction off.
//
// The states X_n, X1, X2 of every step of one segment are recomputed from
// the segment's checkpoint by one forward pass that keeps the stage states.
// The last segment's are kept by the forward itself: its stages write into
// the segment buffers, so the backward starts without a recompute.
// Three adjoint buffers rotate through the stages, so that every flux
// transpose reads its inputs from buffers it does not write.
#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "r3d_adj.h"
#include "r3d_adj.hpp"
#include "r3d_ctx.hpp"

using namespace r3d;

namespace {

constexpr int AT = 256;          // threads per block of the adjoint kernels
constexpr int ADJ_BLOCKS = 2048; // grid-stride blocks of the cell kernels
constexpr int NFH = 5;           // hydrodynamic primaries

__device__ __forceinline__ void cell_of(const Geo& g, long q, int& i, int& j, int& k) {
    i = 1 + (int)(q % g.npx);
    const long r = q / g.npx;
    j = 1 + (int)(r % g.npy);
    k = 1 + (int)(r / g.npy);
}
__device__ __forceinline__ int wrapx(const Geo& g, int i) {
    return i == 0 ? g.npx : (i == g.npx + 1 ? 1 : i);
}
__device__ __forceinline__ int wrapy(const Geo& g, int j) {
    return j == 0 ? g.npy : (j == g.npy + 1 ? 1 : j);
}

// The flux stencil of cell (i, j, k) from the global array: periodic x/y
// neighbours are read from the cells they wrap to, as the stage kernel does.
template <int NP = NPTS>
__device__ void gather_stencil(const Geo& g, const State& X, const Metrics& M, int i, int j, int k,
                               Stencil& s, CellMet& m) {
    for (int p = 0; p < NP; ++p) {
        int di, dj, dk;
        pt_offset(p, di, dj, dk);
        const long n = g.at(wrapx(g, i + di), wrapy(g, j + dj), k + dk);
        s.p[p] = {X.q[QRU][n], X.q[QRV][n], X.q[QRW][n], X.q[QRO][n], X.q[QTT][n]};
        derive_point(s.p[p]);
    }
    m.wall = k == 1 ? 1 : (k == g.npz ? 2 : 0);
    s.rw2 = m.wall == 1 ? X.q[QRW][g.at(i, j, k + 2)]
                        : (m.wall == 2 ? X.q[QRW][g.at(i, j, k - 2)] : 0.0);
    m.dxx = M.dxx[i];
    m.d2xx = M.d2xx[i];
    m.dyy = M.dyy[j];
    m.d2yy = M.d2yy[j];
    m.dzz = M.dzz[k];
    m.d2zz = M.d2zz[k];
    m.rk = M.rkapa[k];
    m.dk = M.dkapa[k];
}

// B at the stencil points and the LMAG wall extras.
template <int NP = NPTS>
__device__ void gather_mag(const Geo& g, const State& X, int i, int j, int k, const CellMet& m,
                           MagStencil& b) {
    for (int p = 0; p < NP; ++p) {
        int di, dj, dk;
        pt_offset(p, di, dj, dk);
        const long n = g.at(wrapx(g, i + di), wrapy(g, j + dj), k + dk);
        b.p[p] = {X.q[QBX][n], X.q[QBY][n], X.q[QBZ][n]};
    }
    b.ro2 = b.bx2 = b.by2 = b.bx3 = b.by3 = 0.0;
    if (m.wall == 1 || m.mag_bot) {
        const int dir = m.wall == 1 ? 1 : -1;
        const long n2 = g.at(i, j, k + 2 * dir), n3 = g.at(i, j, k + 3 * dir);
        b.ro2 = X.q[QRO][n2];
        b.bx2 = X.q[QBX][n2];
        b.by2 = X.q[QBY][n2];
        b.bx3 = X.q[QBX][n3];
        b.by3 = X.q[QBY][n3];
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum v over the block into out[slot] (or atomically into *acc).
__device__ void block_sum(double v, double* red, double* acc) {
    v = wave_sum(v);
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
        atomicAdd(acc, s);
    }
    __syncthreads();
}

// One reverse stage: lamF = ca * LA + cb * LB at every interior cell, then
// LO += LB (the stage input's own adjoint, passed through) + (dF/dX)^T lamF.
// acc[0] += d(lamF.F)/d(ore), acc[1] += d(lamF.F)/d(tk), acc[2] += <F, lamF>.
// LO must not alias LA or LB.
template <bool MAG>
__global__ __launch_bounds__(AT) void k_flux_vjp(const Geo g, const State X, const Consts C,
                                                 const Metrics M, const State LA, double ca,
                                                 const State LB, double cb, const State LO,
                                                 double* acc) {
    constexpr int NF = MAG ? 8 : 5;
    __shared__ double red[AT / 64];
    double go = 0.0, gt = 0.0, dot = 0.0;
    const long total = (long)g.npx * g.npy * g.npz;
    for (long q = (long)blockIdx.x * AT + threadIdx.x; q < total; q += (long)gridDim.x * AT) {
        int i, j, k;
        cell_of(g, q, i, j, k);
        const long n = g.at(i, j, k);
        double lam[NF], own[NF];
        for (int f = 0; f < NF; ++f) {
            own[f] = LB.q[f][n];
            lam[f] = cb * own[f] + (ca != 0.0 ? ca * LA.q[f][n] : 0.0);
        }
        Stencil s, A;
        CellMet m;
        gather_stencil(g, X, M, i, j, k, s, m);
        m.mag_bot = m.wall == 2 && C.mag_bot;
        MagStencil b, AB;
        double F[NF];
        double a, bb;
        if constexpr (MAG) {
            gather_mag(g, X, i, j, k, m, b);
            cell_flux_mag(s, b, m, C, F);
            cell_flux_mag_vjp(s, b, m, C, lam, A, AB, a, bb);
            AB.p[P_C].bx += own[QBX];
            AB.p[P_C].by += own[QBY];
            AB.p[P_C].bz += own[QBZ];
        } else {
            cell_flux(s, m, C, F);
            cell_flux_vjp(s, m, C, lam, A, a, bb);
        }
        for (int f = 0; f < NF; ++f) dot += F[f] * lam[f];
        go += a;
        gt += bb;
        A.p[P_C].ru += own[QRU];
        A.p[P_C].rv += own[QRV];
        A.p[P_C].rw += own[QRW];
        A.p[P_C].ro += own[QRO];
        A.p[P_C].tt += own[QTT];
        for (int p = 0; p < NPTS; ++p) {
            int di, dj, dk;
            pt_offset(p, di, dj, dk);
            const int kk = k + dk;
            if (kk < 1 || kk > g.npz) continue;  // outer planes are not state
            const long t = g.at(wrapx(g, i + di), wrapy(g, j + dj), kk);
            atomicAdd(&LO.q[QRU][t], A.p[p].ru);
            atomicAdd(&LO.q[QRV][t], A.p[p].rv);
            atomicAdd(&LO.q[QRW][t], A.p[p].rw);
            atomicAdd(&LO.q[QRO][t], A.p[p].ro);
            atomicAdd(&LO.q[QTT][t], A.p[p].tt);
            if constexpr (MAG) {
                atomicAdd(&LO.q[QBX][t], AB.p[p].bx);
                atomicAdd(&LO.q[QBY][t], AB.p[p].by);
                atomicAdd(&LO.q[QBZ][t], AB.p[p].bz);
            }
        }
        if (m.wall) atomicAdd(&LO.q[QRW][g.at(i, j, m.wall == 1 ? k + 2 : k - 2)], A.rw2);
        if (MAG && (m.wall == 1 || m.mag_bot)) {
            const int dir = m.wall == 1 ? 1 : -1;
            const long n2 = g.at(i, j, k + 2 * dir), n3 = g.at(i, j, k + 3 * dir);
            atomicAdd(&LO.q[QRO][n2], AB.ro2);
            atomicAdd(&LO.q[QBX][n2], AB.bx2);
            atomicAdd(&LO.q[QBY][n2], AB.by2);
            atomicAdd(&LO.q[QBX][n3], AB.bx3);
            atomicAdd(&LO.q[QBY][n3], AB.by3);
        }
    }
    block_sum(go, red, &acc[0]);
    block_sum(gt, red, &acc[1]);
    block_sum(dot, red, &acc[2]);
}

// f[k * pp] plane base (uniform over the workgroup) plus a per-lane byte offset
__device__ __forceinline__ double ld_at(const double* f, long pp, int k, uint32_t boff) {
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(f + pp * (long)k) +
                                            boff);
}

// Tiled transpose: one workgroup of 256 threads owns a VTW x VTH column tile
// (thread t at x = t % VTW, y = t / VTW) and a chunk of kc planes, and marches
// up the chunk.  The tile is square so that its frame and halo, whose entries
// other workgroups also reach and which go out as global atomics, are the
// smallest share of its cells.
//
// State.  The workgroup stages planes k-1, k, k+1 of the tile plus a one-cell
// x/y halo (periodic neighbours loaded from the cells they wrap to) in an LDS
// ring, each staged cell with 1/RO and the velocities already derived (once
// per cell instead of once per stencil point), and every cell reads its
// stencil from there; plane k+2 is fetched into registers while plane k is
// computed.
//
// Adjoint.  Each cell's stencil transpose is added into a second LDS ring of
// three planes; once plane k is done, plane k-1 has received every
// contribution this workgroup makes to it and is written out.  A cell that
// nothing else reaches (off the tile's frame, in a chunk-interior plane, not on
// a periodic x/y edge, which halos wrap onto, and not in a plane the wall
// cells' extra entries reach) is stored; every other entry is added with a
// global atomic, LO being zero on entry.  Same results as k_flux_vjp up to
// rounding (sum order; this file's transpose kernels contract products and
// sums into fused multiply-adds).
//
// MAG: per cell, the hydrodynamic transpose (cell_flux_vjp, 19 points) is
// emitted first and the transpose of the LMAG terms (cell_flux_mag_extra_vjp,
// face points) after it, so that the two never hold their stencils in
// registers at the same time.
constexpr int VTW = 16, VTH = 256 / VTW;  // tile width (x) and height (y)
constexpr int VRW = VTW + 2, VRH = VTH + 2, VPL = VRH * VRW;
enum { SS_RU = 0, SS_RV, SS_RW, SS_RO, SS_TT, SS_RI, SS_U, SS_V, SS_W, SS_BX, SS_BY, SS_BZ };

template <bool MAG>
__global__ __launch_bounds__(256) void k_vjp_tile(const Geo g, const State X, const Consts C,
                                                  const Metrics M, const State LA, double ca,
                                                  const State LB, double cb, const State LO,
                                                  double* acc, int kc) {
    constexpr int NF = MAG ? 8 : 5;           // primaries: lamF components, adjoint ring fields
    constexpr int NS = MAG ? 12 : 9;          // staged state fields (SS_*)
    constexpr int SLOT = NF * VPL, SSLOT = NS * VPL;
    constexpr int NPC = (VPL + 255) / 256;  // staged cells per thread and plane
    __shared__ double ring[3 * SLOT];
    __shared__ double st[3 * SSLOT];
    __shared__ double red[4];
    const int ntx = (g.npx + VTW - 1) / VTW, nty = (g.npy + VTH - 1) / VTH;
    const int tx = blockIdx.x % ntx, ty = (blockIdx.x / ntx) % nty, tz = blockIdx.x / (ntx * nty);
    const int lx = threadIdx.x % VTW, ly = threadIdx.x / VTW;
    const int i = tx * VTW + lx + 1, j = ty * VTH + ly + 1;
    const bool act = i <= g.npx && j <= g.npy;
    const int k0 = 1 + tz * kc, k1 = min(g.npz, k0 + kc - 1);
    for (int e = threadIdx.x; e < 3 * SLOT; e += 256) ring[e] = 0.0;
    auto slot = [&](int kk) { return (kk - k0 + 1) % 3; };

    // ring cell (rr, c) <-> global (i, j) = (tx*VTW + c, ty*VTH + rr), wrapped
    auto wrap_i = [&](int gi) { return gi == 0 ? g.npx : (gi == g.npx + 1 ? 1 : gi); };
    auto wrap_j = [&](int gj) { return gj == 0 ? g.npy : (gj == g.npy + 1 ? 1 : gj); };
    // this thread's staged cells: byte offsets within a plane (~0u: none)
    uint32_t poff[NPC];
    for (int t = 0; t < NPC; ++t) {
        const int e = threadIdx.x + 256 * t;
        const int c = e % VRW, rr = e / VRW;
        const int gi = tx * VTW + c, gj = ty * VTH + rr;
        poff[t] = (e < VPL && gi <= g.npx + 1 && gj <= g.npy + 1)
                      ? (uint32_t)(8 * (XOFF + wrap_i(gi) + g.px * (long)wrap_j(gj)))
                      : 0xffffffffu;
    }
    auto fetch = [&](int kk, double (*pf)[NF]) {
        for (int t = 0; t < NPC; ++t) {
            uint32_t off = poff[t];
            asm volatile("" : "+v"(off));
            for (int f = 0; f < NF; ++f)
                pf[t][f] = off != 0xffffffffu ? ld_at(X.q[f], g.pp, kk, off) : 1.0;
        }
    };
    auto put = [&](int kk, double (*pf)[NF]) {
        double* d = st + slot(kk) * SSLOT;
        for (int t = 0; t < NPC; ++t) {
            const int e = threadIdx.x + 256 * t;
            if (e >= VPL) continue;
            Pt q{pf[t][QRU], pf[t][QRV], pf[t][QRW], pf[t][QRO], pf[t][QTT]};
            derive_point(q);
            d[SS_RU * VPL + e] = q.ru;
            d[SS_RV * VPL + e] = q.rv;
            d[SS_RW * VPL + e] = q.rw;
            d[SS_RO * VPL + e] = q.ro;
            d[SS_TT * VPL + e] = q.tt;
            d[SS_RI * VPL + e] = q.ri;
            d[SS_U * VPL + e] = q.u;
            d[SS_V * VPL + e] = q.v;
            d[SS_W * VPL + e] = q.w;
            if constexpr (MAG) {
                d[SS_BX * VPL + e] = pf[t][QBX];
                d[SS_BY * VPL + e] = pf[t][QBY];
                d[SS_BZ * VPL + e] = pf[t][QBZ];
            }
        }
    };
    // write adjoint ring plane kk out and clear it
    auto flush = [&](int kk) {
        double* r = ring + slot(kk) * SLOT;
        // planes 3, 4 and npz-2, npz-3 also receive the wall cells' extra
        // entries (two and three planes in), added by atomics as they occur
        const bool inner_plane = kk > k0 && kk < k1 && kk != 3 && kk != 4 && kk != g.npz - 2 &&
                                 kk != g.npz - 3;
        for (int e = threadIdx.x; e < SLOT; e += 256) {
            const double v = r[e];
            if (v == 0.0) continue;
            r[e] = 0.0;
            const int f = e / VPL, rc = e % VPL;
            const int c = rc % VRW, rr = rc / VRW;
            const int gi = tx * VTW + c, gj = ty * VTH + rr;  // 0 .. npx+1 / npy+1
            if (gi > g.npx + 1 || gj > g.npy + 1) continue;  // past a partial tile: never set
            const bool owned = inner_plane && c >= 2 && c <= VTW - 1 && rr >= 2 && rr <= VTH - 1 &&
                               gi >= 2 && gi <= g.npx - 1 && gj >= 2 && gj <= g.npy - 1;
            double* out = &LO.q[f][g.at(wrap_i(gi), wrap_j(gj), kk)];
            if (owned)
                *out = v;
            else
                atomicAdd(out, v);
        }
    };
    // add a cell's hydrodynamic stencil entries (those vjp_mask allows) into
    // the ring
    auto emit = [&](int k, const Stencil& A) {
#pragma unroll
        for (int p = 0; p < NPTS; ++p) {
            int di, dj, dk;
            pt_offset(p, di, dj, dk);
            const int kk = k + dk;
            if (kk < 1 || kk > g.npz) continue;  // outer planes are not state
            double* rp = ring + slot(kk) * SLOT + (ly + 1 + dj) * VRW + (lx + 1 + di);
            const unsigned msk = vjp_mask(p, false);
            const double vals[5] = {A.p[p].ru, A.p[p].rv, A.p[p].rw, A.p[p].ro, A.p[p].tt};
#pragma unroll
            for (int f = 0; f < 5; ++f)
                if (msk & (1u << f)) atomicAdd(rp + f * VPL, vals[f]);
        }
    };

    // Global accesses are ordered so that no wait drains a later prefetch: the
    // vector-memory counter completes in issue order, so each plane first
    // issues the next plane's lamF inputs, then the state prefetch two planes
    // ahead, and its flush atomics last; the plane's compute then waits only
    // for lamF loads issued one plane earlier.
    double go = 0.0, gt = 0.0, dot = 0.0;
    double pf[NPC][NF];
    double lbn[NF], lan[NF];  // LB, LA at (i, j) of the next plane
    const uint32_t coff = (uint32_t)(8 * (XOFF + min(i, g.npx) + g.px * (long)min(j, g.npy)));
    auto fetch_lam = [&](int kk) {
        uint32_t off = coff;
        asm volatile("" : "+v"(off));
        for (int f = 0; f < NF; ++f) {
            lbn[f] = ld_at(LB.q[f], g.pp, kk, off);
            lan[f] = ca != 0.0 ? ld_at(LA.q[f], g.pp, kk, off) : 0.0;
        }
    };
    // per-column metrics, loaded once
    const int ic = min(i, g.npx), jc = min(j, g.npy);
    const double mdxx = M.dxx[ic], md2xx = M.d2xx[ic], mdyy = M.dyy[jc], md2yy = M.d2yy[jc];
    fetch(k0 - 1, pf);
    put(k0 - 1, pf);
    fetch(k0, pf);
    put(k0, pf);
    fetch_lam(k0);
    fetch(k0 + 1, pf);  // k0 + 1 <= npz + 1
    for (int k = k0; k <= k1; ++k) {
        put(k + 1, pf);
        __syncthreads();  // planes k-1..k+1 staged; adjoint plane k+1 cleared
        double own[8], lam[8];
        for (int f = 0; f < NF; ++f) {
            own[f] = lbn[f];
            lam[f] = cb * lbn[f] + ca * lan[f];
        }
        if (k + 1 <= k1) fetch_lam(k + 1);
        if (k + 2 <= g.npz + 1 && k + 1 <= k1) fetch(k + 2, pf);
        if (act) {
            const double* sl[3] = {st + slot(k - 1) * SSLOT, st + slot(k) * SSLOT,
                                   st + slot(k + 1) * SSLOT};
            CellMet m;
            m.wall = k == 1 ? 1 : (k == g.npz ? 2 : 0);
            m.mag_bot = m.wall == 2 && C.mag_bot;
            m.dxx = mdxx;
            m.d2xx = md2xx;
            m.dyy = mdyy;
            m.d2yy = md2yy;
            m.dzz = M.dzz[k];
            m.d2zz = M.d2zz[k];
            m.rk = M.rkapa[k];
            m.dk = M.dkapa[k];
            const double rw2 = m.wall ? X.q[QRW][g.at(i, j, m.wall == 1 ? k + 2 : k - 2)] : 0.0;
            // hydrodynamic part
            {
                Stencil s, A;
#pragma unroll
                for (int p = 0; p < NPTS; ++p) {
                    int di, dj, dk;
                    pt_offset(p, di, dj, dk);
                    const double* pl = sl[dk + 1] + (ly + 1 + dj) * VRW + (lx + 1 + di);
                    s.p[p] = {pl[SS_RU * VPL], pl[SS_RV * VPL], pl[SS_RW * VPL], pl[SS_RO * VPL],
                              pl[SS_TT * VPL], pl[SS_RI * VPL], pl[SS_U * VPL], pl[SS_V * VPL],
                              pl[SS_W * VPL]};
                }
                s.rw2 = rw2;
                double F[5], a, bb;
                cell_flux(s, m, C, F);
                cell_flux_vjp(s, m, C, lam, A, a, bb);
                go += a;
                gt += bb;
                for (int f = 0; f < 5; ++f) dot += F[f] * lam[f];
                A.p[P_C].ru += own[QRU];
                A.p[P_C].rv += own[QRV];
                A.p[P_C].rw += own[QRW];
                A.p[P_C].ro += own[QRO];
                A.p[P_C].tt += own[QTT];
                emit(k, A);
                if (m.wall) atomicAdd(&LO.q[QRW][g.at(i, j, m.wall == 1 ? k + 2 : k - 2)], A.rw2);
            }
            // LMAG terms; the empty asm keeps the compiler from hoisting their
            // stencil reads above the hydrodynamic part (the two stencils
            // would then be live together and spill)
            if constexpr (MAG) {
                asm volatile("" ::: "memory");
                Stencil s;
                MagStencil b;
                auto at_pt = [&](int p) {
                    int di, dj, dk;
                    pt_offset(p, di, dj, dk);
                    return sl[dk + 1] + (ly + 1 + dj) * VRW + (lx + 1 + di);
                };
#pragma unroll
                for (int p = 0; p < NFACE; ++p) {
                    const double* pl = at_pt(p);
                    s.p[p].ri = pl[SS_RI * VPL];
                    s.p[p].u = pl[SS_U * VPL];
                    s.p[p].v = pl[SS_V * VPL];
                    s.p[p].w = pl[SS_W * VPL];
                    b.p[p] = {pl[SS_BX * VPL], pl[SS_BY * VPL], pl[SS_BZ * VPL]};
                }
                s.rw2 = rw2;
                b.ro2 = b.bx2 = b.by2 = b.bx3 = b.by3 = 0.0;
                if (m.wall == 1 || m.mag_bot) {
                    const int dir = m.wall == 1 ? 1 : -1;
                    const long n2 = g.at(i, j, k + 2 * dir), n3 = g.at(i, j, k + 3 * dir);
                    b.ro2 = X.q[QRO][n2];
                    b.bx2 = X.q[QBX][n2];
                    b.by2 = X.q[QBY][n2];
                    b.bx3 = X.q[QBX][n3];
                    b.by3 = X.q[QBY][n3];
                }
                {
                    double E[8];
                    cell_flux_mag_extra(s, b, m, C, E);
                    for (int f = 0; f < 8; ++f) dot += E[f] * lam[f];
                }
                // point-major transpose: each face point's entries go into the
                // ring as soon as they are known (the point's 1/RO and
                // velocities re-read from the staged planes), so that no
                // per-point adjoint array is held
                MagWallAdj W;
                cell_flux_mag_extra_vjp_points(
                    s, b, m, C, lam,
                    [&](int q, double& ri, double& u, double& v, double& w) {
                        const double* pl = at_pt(q);
                        ri = pl[SS_RI * VPL];
                        u = pl[SS_U * VPL];
                        v = pl[SS_V * VPL];
                        w = pl[SS_W * VPL];
                    },
                    [&](int q, int f, double val) {
                        int di, dj, dk;
                        pt_offset(q, di, dj, dk);
                        const int kk = k + dk;
                        if (kk < 1 || kk > g.npz) return;  // outer planes are not state
                        if (q == P_C && f >= QBX) val += own[f];
                        atomicAdd(ring + slot(kk) * SLOT + (ly + 1 + dj) * VRW + (lx + 1 + di) +
                                      f * VPL,
                                  val);
                    },
                    W);
                if (m.wall == 1 || m.mag_bot) {
                    const int dir = m.wall == 1 ? 1 : -1;
                    const long n2 = g.at(i, j, k + 2 * dir), n3 = g.at(i, j, k + 3 * dir);
                    atomicAdd(&LO.q[QRW][n2], W.rw2);
                    atomicAdd(&LO.q[QRO][n2], W.ro2);
                    atomicAdd(&LO.q[QBX][n2], W.bx2);
                    atomicAdd(&LO.q[QBY][n2], W.by2);
                    atomicAdd(&LO.q[QBX][n3], W.bx3);
                    atomicAdd(&LO.q[QBY][n3], W.by3);
                }
            }
        }
        __syncthreads();  // plane k done: k-1 complete; state plane k-1 free
        if (k - 1 >= 1) flush(k - 1);
    }
    __syncthreads();
    flush(k1);
    if (k1 + 1 <= g.npz) flush(k1 + 1);
    // block sums
    auto bsum = [&](double v, double* out) {
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
        __syncthreads();
    };
    bsum(go, &acc[0]);
    bsum(gt, &acc[1]);
    bsum(dot, &acc[2]);
}

// LX <- B^T LX on the two wall planes, with X the post-boundary state;
// acc[3] += d/d(theta), acc[5] += d/d(tb) (the fixed bottom temperature).
// grid = (ceil(npx*npy / AT), 2 sides).
__global__ __launch_bounds__(AT) void k_bcon_adj(const Geo g, const State X, const Consts C,
                                                 const Metrics M, const State LX, double* acc) {
    __shared__ double red[AT / 64];
    double gth = 0.0, gtb = 0.0;
    const int side = blockIdx.y;
    const int q = blockIdx.x * AT + threadIdx.x;
    if (q < g.npx * g.npy) {
        const int i = 1 + q % g.npx, j = 1 + q / g.npx;
        const int kw = side ? g.npz : 1;
        const long n = g.at(i, j, kw);
        const long d1 = side ? -g.pp : g.pp;
        const long n1 = n + d1, n2 = n + 2 * d1;
        const double ro = X.q[QRO][n], ro1 = X.q[QRO][n1], ro2 = X.q[QRO][n2];
        for (int f = QRU; f <= QRV; ++f) {
            // m = ro * (c43 m1 / ro1 - c13 m2 / ro2)  (src/bcon.cpp:22-37)
            const double a = LX.q[f][n];
            const double m1 = X.q[f][n1], m2 = X.q[f][n2];
            LX.q[QRO][n] += a * (C.c43 * m1 / ro1 - C.c13 * m2 / ro2);
            LX.q[f][n1] += a * ro * C.c43 / ro1;
            LX.q[QRO][n1] -= a * ro * C.c43 * m1 / (ro1 * ro1);
            LX.q[f][n2] -= a * ro * C.c13 / ro2;
            LX.q[QRO][n2] += a * ro * C.c13 * m2 / (ro2 * ro2);
            LX.q[f][n] = 0.0;
        }
        LX.q[QRW][n] = 0.0;
        if (C.lmag && C.bz_zero) LX.q[QBZ][n] = 0.0;  // BZ = 0 on the walls
        const double aT = LX.q[QTT][n];
        bool fixed = false, extrap = false;
        double dth = 0.0;
        if (side == 0) {
            switch (C.top_mode) {
                case TOP_GAUSS_FLUX:
                    extrap = true;
                    dth = -C.c23 * M.dz_top * (1.0 - M.e1[i] / C.hh * M.e2[j] / C.hh);
                    break;
                case TOP_CONST_FLUX:
                    extrap = true;
                    dth = -C.c23 * M.dz_top;
                    break;
                case TOP_GAUSS_T:
                case TOP_CONST_T:
                    fixed = true;
                    break;
                default:
                    break;
            }
        } else if (C.izcon == 0) {
            fixed = true;  // T = tb
            gtb = aT;
        } else {
            extrap = true;
            dth = C.c23 * M.dz_bot / M.rkapa[kw];
        }
        if (extrap) {
            LX.q[QTT][n1] += C.c43 * aT;
            LX.q[QTT][n2] -= C.c13 * aT;
            gth = aT * dth;
        }
        if (fixed || extrap) LX.q[QTT][n] = 0.0;
    }
    block_sum(gth, red, &acc[3]);
    block_sum(gtb, red, &acc[5]);
}

// Objective per block (deterministic host sum) and, if LX is given, its
// gradient: J = sum vol (ru^2+rv^2+rw^2)/(2 ro), vol = 1/(dxx dyy dzz).
// X0's outer z planes (k = 0 and npz + 1) of every field into nbuf state
// buffers: each buffer a stage writes into must hold them (the stage kernels
// write interior cells only).
__global__ __launch_bounds__(AT) void k_outer_planes(double* const* dst, int nbuf,
                                                     const double* x0, long fe, int nf, long pp,
                                                     int npz) {
    const long per = 2 * pp;  // the two planes of one field
    const long total = (long)nbuf * nf * per;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long)gridDim.x * blockDim.x) {
        const long e = q % per, bf = q / per;
        const long off = (bf % nf) * fe + (e < pp ? e : (long)(npz + 1) * pp + (e - pp));
        dst[bf / nf][off] = x0[off];
    }
}

__global__ __launch_bounds__(AT) void k_ke(const Geo g, const State X, const Metrics M,
                                           const State LX, int want_grad, double* part) {
#pragma clang fp contract(off)
    __shared__ double red[AT / 64];
    double sum = 0.0;
    const long total = (long)g.npx * g.npy * g.npz;
    for (long q = (long)blockIdx.x * AT + threadIdx.x; q < total; q += (long)gridDim.x * AT) {
        int i, j, k;
        cell_of(g, q, i, j, k);
        const long n = g.at(i, j, k);
        const double vol = 1.0 / (M.dxx[i] * M.dyy[j] * M.dzz[k]);
        const double ru = X.q[QRU][n], rv = X.q[QRV][n], rw = X.q[QRW][n], ro = X.q[QRO][n];
        const double m2 = ru * ru + rv * rv + rw * rw;
        sum += vol * 0.5 * m2 / ro;
        if (want_grad) {
            LX.q[QRU][n] = vol * ru / ro;
            LX.q[QRV][n] = vol * rv / ro;
            LX.q[QRW][n] = vol * rw / ro;
            LX.q[QRO][n] = -vol * 0.5 * m2 / (ro * ro);
            LX.q[QTT][n] = 0.0;
        }
    }
    sum = wave_sum(sum);
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) red[wv] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < AT / 64; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// Per-block minimum of the dt limits (the formulas of dt_terms) with the key
// cell * 4 + limit, ties to the smaller key.
__global__ __launch_bounds__(AT) void k_dt_argmin(const Geo g, const State X, const Consts C,
                                                  const Metrics M, double* bval, long* bkey) {
    // the limits must round as the forward's dt reduction does, so that the
    // minimum falls on the same cell and limit
#pragma clang fp contract(off)
    __shared__ double rv[AT / 64];
    __shared__ long rk[AT / 64];
    double best = 1e300;
    long key = LONG_MAX;
    const long total = (long)g.npx * g.npy * g.npz;
    for (long q = (long)blockIdx.x * AT + threadIdx.x; q < total; q += (long)gridDim.x * AT) {
        int i, j, k;
        cell_of(g, q, i, j, k);
        const long n = g.at(i, j, k);
        const double ru = X.q[QRU][n], rv_ = X.q[QRV][n], rw = X.q[QRW][n];
        const double ro = X.q[QRO][n], tt = X.q[QTT][n];
        const double dd = dmin(dmin(M.ddx[i], M.ddy[j]), M.ddz[k]);
        const double uu = (ru * ru + rv_ * rv_ + rw * rw) * (1.0 / ro) * (1.0 / ro);
        double fast;
        if (C.lmag) {
            const double bx = X.q[QBX][n], by = X.q[QBY][n], bz = X.q[QBZ][n];
            const double b2 = (bx * bx + by * by + bz * bz) * C.obeta / ro;
            fast = uu + C.gamma * tt + b2 + 2.0 * sqrt(uu * (C.gamma * tt + b2));
        } else {
            fast = uu + C.gamma * tt + 2.0 * sqrt(uu * C.gamma * tt);
        }
        const double L[4] = {dd / sqrt(fast), 0.5 * dd * dd * C.repr * C.cv * ro / M.rkapa[k],
                             0.375 * dd * dd * C.re * ro, 0.5 * dd * dd * C.rm};
        for (int t = 0; t < (C.lmag ? 4 : 3); ++t) {
            const long kk = q * 4 + t + 1;
            if (L[t] < best || (L[t] == best && kk < key)) {
                best = L[t];
                key = kk;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const long ok = __shfl_xor(key, o, 64);
        if (ob < best || (ob == best && ok < key)) {
            best = ob;
            key = ok;
        }
    }
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) {
        rv[wv] = best;
        rk[wv] = key;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < AT / 64; ++w)
            if (rv[w] < best || (rv[w] == best && rk[w] < key)) {
                best = rv[w];
                key = rk[w];
            }
        bval[blockIdx.x] = best;
        bkey[blockIdx.x] = key;
    }
}

// Single block: the cell and limit attaining the minimum over the block
// results, then LX[cell] += sff * lam_dt * dL/dX there, lam_dt = acc[2] / dt,
// and acc[4] += sff * lam_dt * dL/dRe.  At a cell at rest the advective
// limit's velocity derivative is taken as 0.
__global__ __launch_bounds__(64) void k_dt_apply(const Geo g, const State X, const Consts C,
                                                 const double* bval, const long* bkey, int nb,
                                                 double dt, const State LX, double* acc) {
    double best = 1e300;
    long key = LONG_MAX;
    for (int b = threadIdx.x; b < nb; b += 64)
        if (bval[b] < best || (bval[b] == best && bkey[b] < key)) {
            best = bval[b];
            key = bkey[b];
        }
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const long ok = __shfl_xor(key, o, 64);
        if (ob < best || (ob == best && ok < key)) {
            best = ob;
            key = ok;
        }
    }
    if (threadIdx.x != 0) return;
    const long q = (key - 1) / 4;
    const int t = (int)((key - 1) % 4);  // 0 advective, 1 thermal, 2 viscous
    int i, j, k;
    cell_of(g, q, i, j, k);
    const long n = g.at(i, j, k);
    const double ru = X.q[QRU][n], rv = X.q[QRV][n], rw = X.q[QRW][n];
    const double ro = X.q[QRO][n], tt = X.q[QTT][n];
    const double L = best, s = C.sff * acc[2] / dt;
    double dv[NXF] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, dre = 0.0;
    if (t == 0) {
        // L = dd / sqrt(fast), fast = uu + a + 2 sqrt(uu a), uu = |m|^2 / ro^2,
        // a = gamma T (+ b2 = |B|^2 obeta / ro with LMAG)
        const double m2 = ru * ru + rv * rv + rw * rw;
        const double uu = m2 / (ro * ro);
        double bx = 0.0, by = 0.0, bz = 0.0, b2 = 0.0;
        if (C.lmag) {
            bx = X.q[QBX][n];
            by = X.q[QBY][n];
            bz = X.q[QBZ][n];
            b2 = (bx * bx + by * by + bz * bz) * C.obeta / ro;
        }
        const double aa = C.gamma * tt + b2;
        const double sq = sqrt(uu * aa);
        const double fast = uu + aa + 2.0 * sq;
        const double dLdf = -0.5 * L / fast;
        const double dfdu = uu > 0.0 ? 1.0 + aa / sq : 0.0;
        const double dfda = 1.0 + (uu > 0.0 ? uu / sq : 0.0);
        dv[QRU] = dLdf * dfdu * 2.0 * ru / (ro * ro);
        dv[QRV] = dLdf * dfdu * 2.0 * rv / (ro * ro);
        dv[QRW] = dLdf * dfdu * 2.0 * rw / (ro * ro);
        dv[QRO] = dLdf * (dfdu * (-2.0 * uu / ro) + dfda * (-b2 / ro));
        dv[QTT] = dLdf * dfda * C.gamma;
        dv[QBX] = dLdf * dfda * 2.0 * bx * C.obeta / ro;
        dv[QBY] = dLdf * dfda * 2.0 * by * C.obeta / ro;
        dv[QBZ] = dLdf * dfda * 2.0 * bz * C.obeta / ro;
    } else if (t < 3) {
        // thermal 0.5 dd^2 Re Pr cv ro / rk and viscous 0.375 dd^2 Re ro: L ~ Re ro
        dv[QRO] = L / ro;
        dre = L / C.re;
    }  // t = 3: the resistive limit 0.5 dd^2 Rm depends on neither
    for (int f = 0; f < (C.lmag ? 8 : 5); ++f) LX.q[f][n] += s * dv[f];
    acc[4] += s * dre;
}

}  // namespace

struct r3d_adj_ctx {
    r3d_hip_ctx* work = nullptr;   // forward steps (plain path)
    int K = 20;                     // checkpoint cadence
    int nsteps = 0;
    int nf = 5;                     // primaries: 5, or 8 with LMAG
    size_t fe = 0;                  // doubles per field array
    std::vector<double*> ckpt;      // X at steps 0, K, 2K, ...
    std::vector<double*> segx, seg1, seg2;  // X_n, X1, X2 of the steps of one segment
    std::vector<double> segdt;
    double* dt_pinned = nullptr;    // dt of the recomputed steps, read back once per segment
    double* scratch = nullptr;      // a state buffer for the unneeded last output of a segment
    double** seg_bufs = nullptr;    // device list: the segment buffers and scratch
    int nseg_bufs = 0;
    bool seg_outer = false;         // segment buffers hold X0's outer z planes, xfinal all of X0
    bool kept_last = false;         // segment buffers and segdt hold the last forward's last segment
    double *x0 = nullptr, *xfinal = nullptr;
    double* lam[3] = {nullptr, nullptr, nullptr};  // rotating adjoint buffers
    double* acc = nullptr;          // g_ore, g_tk, <F, lamF>, g_theta, dt's d/dRe, g_tb
    double dJ_dtb = 0.0;            // of the last backward
    double* part = nullptr;         // k_ke block sums
    double* bval = nullptr;
    long* bkey = nullptr;
    double J = 0.0;
    double fwd_ms = 0.0, bwd_ms = 0.0;
    unsigned long long bytes = 0;
    bool tiled = true;  // k_vjp_tile; R3D_ADJ_VJP=atomic selects k_flux_vjp
};

namespace {

State state_of(const r3d_adj_ctx* a, double* base) {
    State s{};
    for (int f = 0; f < a->nf; ++f) s.q[f] = base + f * a->fe;
    return s;
}

int alloc_state(r3d_adj_ctx* a, double** p) {
    if (hipMalloc((void**)p, a->fe * a->nf * sizeof(double)) != hipSuccess) {
        *p = nullptr;
        return set_err(R3D_ENOMEM, "adjoint: state buffer of %zu bytes",
                       a->fe * a->nf * sizeof(double));
    }
    a->bytes += a->fe * a->nf * sizeof(double);
    return R3D_OK;
}

// work context's current X <-> a state buffer.  put_state fills both X
// buffers, as an upload does: the stage kernels write interior cells only, so
// the outer z planes of every stage state (read by the wall cells' fluxes)
// are those of the state put here.
int put_state(r3d_adj_ctx* a, const double* src) {
    r3d_hip_ctx* w = a->work;
    for (int s = 0; s < 2; ++s)
        for (int f = 0; f < a->nf; ++f)
            HIPCHK(hipMemcpyAsync(w->xs[s][f], src + f * a->fe, a->fe * sizeof(double),
                                  hipMemcpyDeviceToDevice, w->stream));
    w->next_valid = false;
    return R3D_OK;
}
int get_state(r3d_adj_ctx* a, double* dst) {
    r3d_hip_ctx* w = a->work;
    for (int f = 0; f < a->nf; ++f)
        HIPCHK(hipMemcpyAsync(dst + f * a->fe, w->xs[w->cur][f], a->fe * sizeof(double),
                              hipMemcpyDeviceToDevice, w->stream));
    return R3D_OK;
}

// The stage kernels write interior cells only, and the outer z planes keep
// X0's values (RO's is rewritten by the boundary kernel every stage), so a
// buffer that a stage writes into must hold them: copy X0's planes 0 and
// npz+1 into every segment buffer, once per initial state.
// The final state of the forward is the last stage's output buffer: it gets
// the whole of X0 (ghost entries included), as the work context's buffers do
// from put_state.
int seg_outer_planes(r3d_adj_ctx* a) {
    if (a->seg_outer) return R3D_OK;
    r3d_hip_ctx* w = a->work;
    const long total = (long)a->nseg_bufs * a->nf * 2 * w->g.pp;
    hipLaunchKernelGGL(k_outer_planes, dim3((unsigned)std::min<long>(4096, (total + AT - 1) / AT)),
                       dim3(AT), 0, w->stream, a->seg_bufs, a->nseg_bufs, a->x0, (long)a->fe,
                       a->nf, (long)w->g.pp, w->g.npz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(a->xfinal, a->x0, a->fe * a->nf * sizeof(double),
                          hipMemcpyDeviceToDevice, w->stream));
    a->seg_outer = true;
    return R3D_OK;
}

int blocks_for(const Geo& g) {
    const long total = (long)g.npx * g.npy * g.npz;
    return (int)std::min<long>(ADJ_BLOCKS, (total + AT - 1) / AT);
}

int objective(r3d_adj_ctx* a, const double* x, double* lx, double* J) {
    r3d_hip_ctx* w = a->work;
    const int nb = blocks_for(w->g);
    hipLaunchKernelGGL(k_ke, dim3(nb), dim3(AT), 0, w->stream, w->g, state_of(a, (double*)x),
                       w->M, state_of(a, lx ? lx : a->lam[0]), lx ? 1 : 0, a->part);
    HIPCHK(hipGetLastError());
    std::vector<double> h(nb);
    HIPCHK(hipMemcpyAsync(h.data(), a->part, nb * sizeof(double), hipMemcpyDeviceToHost,
                          w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    double s = 0.0;
    for (double v : h) s += v;
    *J = s;
    return R3D_OK;
}

// LO = LB + (dF/dX)^T (ca LA + cb LB) at the state x (LO zeroed first).
int vjp(r3d_adj_ctx* a, const double* x, const double* la, double ca, const double* lb, double cb,
        double* lo) {
    r3d_hip_ctx* w = a->work;
    HIPCHK(hipMemsetAsync(lo, 0, a->fe * a->nf * sizeof(double), w->stream));
    const State sx = state_of(a, (double*)x), sa = state_of(a, (double*)(la ? la : lb));
    const State sb = state_of(a, (double*)lb), so = state_of(a, lo);
    const Geo& g = w->g;
    if (a->tiled) {
        // chunks of >= 8 planes, at least about 512 workgroups in all
        auto grid_of = [&](int& kc) {
            const long tiles = (long)((g.npx + VTW - 1) / VTW) * ((g.npy + VTH - 1) / VTH);
            long nkz = std::max<long>(1, std::min<long>(g.npz / 8, (512 + tiles - 1) / tiles));
            kc = (int)((g.npz + nkz - 1) / nkz);
            nkz = (g.npz + kc - 1) / kc;
            return dim3((unsigned)(tiles * nkz));
        };
        int kc;
        const dim3 grid = grid_of(kc);
        if (a->nf == 8)
            hipLaunchKernelGGL(k_vjp_tile<true>, grid, dim3(256), 0, w->stream, g, sx, w->C, w->M,
                               sa, la ? ca : 0.0, sb, cb, so, a->acc, kc);
        else
            hipLaunchKernelGGL(k_vjp_tile<false>, grid, dim3(256), 0, w->stream, g, sx, w->C,
                               w->M, sa, la ? ca : 0.0, sb, cb, so, a->acc, kc);
        HIPCHK(hipGetLastError());
        return R3D_OK;
    }
    const dim3 grid(blocks_for(w->g)), blk(AT);
    if (a->nf == 8)
        hipLaunchKernelGGL(k_flux_vjp<true>, grid, blk, 0, w->stream, w->g, sx, w->C, w->M, sa,
                           la ? ca : 0.0, sb, cb, so, a->acc);
    else
        hipLaunchKernelGGL(k_flux_vjp<false>, grid, blk, 0, w->stream, w->g, sx, w->C, w->M, sa,
                           la ? ca : 0.0, sb, cb, so, a->acc);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

int bcon_adj(r3d_adj_ctx* a, const double* x, double* lx) {
    r3d_hip_ctx* w = a->work;
    const int nb = (w->g.npx * w->g.npy + AT - 1) / AT;
    hipLaunchKernelGGL(k_bcon_adj, dim3(nb, 2), dim3(AT), 0, w->stream, w->g,
                       state_of(a, (double*)x), w->C, w->M, state_of(a, lx), a->acc);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

int dt_adjoint(r3d_adj_ctx* a, const double* x, double dt, double* lx) {
    r3d_hip_ctx* w = a->work;
    const int nb = blocks_for(w->g);
    hipLaunchKernelGGL(k_dt_argmin, dim3(nb), dim3(AT), 0, w->stream, w->g,
                       state_of(a, (double*)x), w->C, w->M, a->bval, a->bkey);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_dt_apply, dim3(1), dim3(64), 0, w->stream, w->g, state_of(a, (double*)x),
                       w->C, a->bval, a->bkey, nb, dt, state_of(a, lx), a->acc);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

}  // namespace

extern "C" {

int r3d_adj_create(const r3d_hip_cfg* cfg, int checkpoint_every, r3d_adj_ctx** out) {
    if (!cfg || !out || checkpoint_every < 1) return set_err(R3D_EINVAL, "bad arguments");
    *out = nullptr;
    if (cfg->tc != 0.0) return set_err(R3D_EUNSUP, "adjoint: tc != 0 is not supported");
    if (cfg->itcon == 2) return set_err(R3D_EUNSUP, "adjoint: ITCON=2 heat loss is not supported");
    if (cfg->id != 0) return set_err(R3D_EUNSUP, "adjoint: density diffusion is not supported");
    if (cfg->rlax > 0.0) return set_err(R3D_EUNSUP, "adjoint: flux relaxation is not supported");
    if (cfg->lrem) return set_err(R3D_EUNSUP, "adjoint: LREM is not supported");
    auto* a = new r3d_adj_ctx();
    a->K = checkpoint_every;
    a->nf = cfg->lmag ? 8 : 5;
    if (const char* e = std::getenv("R3D_ADJ_VJP")) a->tiled = std::strcmp(e, "atomic") != 0;
    int rc = r3d_hip_create(cfg, &a->work);
    if (rc) {
        delete a;
        return rc;
    }
    std::fill(a->work->uploaded.begin(), a->work->uploaded.end(), 1);
    a->fe = field_elems(a->work->g);
    auto fail = [&](int code) {
        r3d_adj_destroy(a);
        return code;
    };
    double** bufs[] = {&a->x0, &a->xfinal, &a->lam[0], &a->lam[1], &a->lam[2]};
    for (double** b : bufs)
        if ((rc = alloc_state(a, b))) return fail(rc);
    if (hipMalloc((void**)&a->acc, 6 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&a->part, ADJ_BLOCKS * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&a->bval, ADJ_BLOCKS * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&a->bkey, ADJ_BLOCKS * sizeof(long)) != hipSuccess)
        return fail(set_err(R3D_ENOMEM, "adjoint: small buffers"));
    for (auto* v : {&a->segx, &a->seg1, &a->seg2}) {
        v->assign(a->K, nullptr);
        for (auto& p : *v)
            if ((rc = alloc_state(a, &p))) return fail(rc);
    }
    a->segdt.assign(a->K, 0.0);
    if ((rc = alloc_state(a, &a->scratch))) return fail(rc);
    {
        std::vector<double*> bufs(a->segx);
        bufs.insert(bufs.end(), a->seg1.begin(), a->seg1.end());
        bufs.insert(bufs.end(), a->seg2.begin(), a->seg2.end());
        bufs.push_back(a->scratch);
        a->nseg_bufs = (int)bufs.size();
        if (hipMalloc((void**)&a->seg_bufs, bufs.size() * sizeof(double*)) != hipSuccess)
            return fail(set_err(R3D_ENOMEM, "adjoint: buffer list"));
        if (hipMemcpy(a->seg_bufs, bufs.data(), bufs.size() * sizeof(double*),
                      hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(R3D_EHIP, "adjoint: buffer list upload"));
    }
    if (hipHostMalloc((void**)&a->dt_pinned, a->K * sizeof(double)) != hipSuccess)
        return fail(set_err(R3D_ENOMEM, "adjoint: pinned dt buffer"));
    // ghost entries of the initial state and the adjoint fields stay zero
    for (double* p : {a->x0, a->xfinal, a->lam[0], a->lam[1], a->lam[2]})
        if (hipMemset(p, 0, a->fe * a->nf * sizeof(double)) != hipSuccess)
            return fail(set_err(R3D_EHIP, "adjoint: memset"));
    *out = a;
    return R3D_OK;
}

int r3d_adj_destroy(r3d_adj_ctx* a) {
    if (!a) return R3D_OK;
    if (a->work) (void)hipStreamSynchronize(a->work->stream);
    for (double* p : {a->x0, a->xfinal, a->lam[0], a->lam[1], a->lam[2], a->acc, a->part,
                      a->bval, a->scratch})
        if (p) (void)hipFree(p);
    if (a->dt_pinned) (void)hipHostFree(a->dt_pinned);
    if (a->bkey) (void)hipFree(a->bkey);
    if (a->seg_bufs) (void)hipFree(a->seg_bufs);
    for (auto* v : {&a->ckpt, &a->segx, &a->seg1, &a->seg2})
        for (double* p : *v)
            if (p) (void)hipFree(p);
    if (a->work) r3d_hip_destroy(a->work);
    delete a;
    return R3D_OK;
}

int r3d_adj_set_metrics(r3d_adj_ctx* a, const r3d_hip_metrics* m) {
    if (!a) return set_err(R3D_EINVAL, "null adjoint context");
    return r3d_hip_set_metrics(a->work, m);
}

namespace {
int set_initial(r3d_adj_ctx* a, const double* const* fields, bool device) {
    if (!a || !fields) return set_err(R3D_EINVAL, "null argument");
    for (int f = 0; f < a->nf; ++f) {
        if (!fields[f]) return set_err(R3D_EINVAL, "field %d missing", f);
        double* src = const_cast<double*>(fields[f]);
        int rc = device ? global_copy_device(a->work, a->x0 + f * a->fe, src, false)
                        : global_copy(a->work, a->x0 + f * a->fe, src, false);
        if (rc) return rc;
    }
    a->nsteps = 0;
    a->seg_outer = false;
    a->kept_last = false;
    return R3D_OK;
}
}  // namespace

int r3d_adj_set_initial(r3d_adj_ctx* a, const double* const* fields) {
    return set_initial(a, fields, false);
}

int r3d_adj_set_initial_device(r3d_adj_ctx* a, const double* const* fields) {
    return set_initial(a, fields, true);
}

int r3d_adj_forward(r3d_adj_ctx* a, int nsteps, double* J) {
    if (!a || nsteps < 1) return set_err(R3D_EINVAL, "bad arguments");
    r3d_hip_ctx* w = a->work;
    int rc;
    if (!w->metrics_set) return set_err(R3D_ESTATE, "metrics not uploaded");
    const auto t0 = std::chrono::steady_clock::now();
    const int nck = (nsteps + a->K - 1) / a->K;
    while ((int)a->ckpt.size() < nck) {
        double* p = nullptr;
        if ((rc = alloc_state(a, &p))) return rc;
        a->ckpt.push_back(p);
    }
    a->kept_last = false;
    a->nsteps = 0;
    if ((rc = seg_outer_planes(a))) return rc;
    w->cur = 0;
    if ((rc = put_state(a, a->x0))) return rc;
    r3d_hip_scalars s0{};
    if ((rc = r3d_hip_set_scalars(w, &s0))) return rc;
    const int n0 = (nck - 1) * a->K;  // first step of the last segment
    for (int n = 0; n < n0; ++n) {
        if (n % a->K == 0 && (rc = get_state(a, a->ckpt[n / a->K]))) return rc;
        if ((rc = r3d_hip_step(w, 1, 3))) return rc;
    }
    // The last segment: its checkpoint, then its steps with the stage outputs
    // redirected into the segment buffers (as the backward's recompute writes
    // them) and the final state into xfinal.  The dt of step n0 is already
    // reduced (next_valid is kept), so the steps are those of the plain path.
    if ((rc = get_state(a, a->ckpt[nck - 1]))) return rc;
    {
        double* own[2][NXF];
        std::memcpy(own, w->xs, sizeof own);
        w->cur = 0;
        for (int f = 0; f < a->nf; ++f) w->xs[0][f] = a->ckpt[nck - 1] + f * a->fe;
        for (int n = n0; n < nsteps; ++n) {
            double* outs[3] = {a->seg1[n - n0], a->seg2[n - n0],
                               n + 1 < nsteps ? a->segx[n + 1 - n0] : a->xfinal};
            rc = step_one(w, 3, nullptr, outs);
            if (!rc && hipMemcpyAsync(a->dt_pinned + (n - n0), &w->scal->dt, sizeof(double),
                                      hipMemcpyDeviceToHost, w->stream) != hipSuccess)
                rc = set_err(R3D_EHIP, "dt readback failed");
            if (rc) break;
        }
        std::memcpy(w->xs, own, sizeof own);
        w->cur = 0;
        w->next_valid = false;
        if (rc) return rc;
    }
    if ((rc = objective(a, a->xfinal, nullptr, &a->J))) return rc;
    HIPCHK(hipStreamSynchronize(w->stream));
    for (int n = n0; n < nsteps; ++n) a->segdt[n - n0] = a->dt_pinned[n - n0];
    a->kept_last = true;
    a->nsteps = nsteps;
    if (J) *J = a->J;
    a->fwd_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
    return R3D_OK;
}

namespace {
int backward(r3d_adj_ctx* a, double* const* grad, double* dJ_dre, double* dJ_dth, bool device) {
    if (!a) return set_err(R3D_EINVAL, "null adjoint context");
    if (a->nsteps < 1) return set_err(R3D_ESTATE, "no forward run to differentiate");
    r3d_hip_ctx* w = a->work;
    const Consts& C = w->C;
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    double* L = a->lam[0];
    double* Mb = a->lam[1];
    double* Nb = a->lam[2];
    double Jdummy;
    HIPCHK(hipMemsetAsync(L, 0, a->fe * a->nf * sizeof(double), w->stream));
    if ((rc = objective(a, a->xfinal, L, &Jdummy))) return rc;  // L = dJ/dX_N (B: 0)
    HIPCHK(hipMemsetAsync(a->acc, 0, 6 * sizeof(double), w->stream));
    const double g1 = C.gam[0], g2 = C.gam[1], g3 = C.gam[2];
    const double z1 = C.zeta[0], z2 = C.zeta[1];
    const int N = a->nsteps, nseg = (N + a->K - 1) / a->K;
    if ((rc = seg_outer_planes(a))) return rc;
    for (int sgm = nseg - 1; sgm >= 0; --sgm) {
        const int n0 = sgm * a->K, n1 = std::min(N, n0 + a->K);
        // X_n, X1, X2 and dt of every step of this segment, from its
        // checkpoint.  The stages read and write the checkpoint and segment
        // buffers directly (the work context's state pointers are redirected,
        // then restored), and dt comes back by one copy per segment.  The
        // forward left the last segment's in place.
        if (!(sgm == nseg - 1 && a->kept_last)) {
            a->kept_last = false;  // the recompute overwrites the kept segment
            double* own[2][NXF];
            std::memcpy(own, w->xs, sizeof own);
            w->cur = 0;
            for (int f = 0; f < a->nf; ++f) w->xs[0][f] = a->ckpt[sgm] + f * a->fe;
            w->next_valid = false;
            for (int n = n0; n < n1; ++n) {
                // the last step's X_{n+1} (a checkpoint or the final state) is
                // stored already; it goes to a scratch buffer
                double* outs[3] = {a->seg1[n - n0], a->seg2[n - n0],
                                   n + 1 < n1 ? a->segx[n + 1 - n0] : a->scratch};
                rc = step_one(w, 3, nullptr, outs);
                if (!rc && hipMemcpyAsync(a->dt_pinned + (n - n0), &w->scal->dt, sizeof(double),
                                          hipMemcpyDeviceToHost, w->stream) != hipSuccess)
                    rc = set_err(R3D_EHIP, "dt readback failed");
                if (rc) break;
            }
            std::memcpy(w->xs, own, sizeof own);
            w->cur = 0;
            w->next_valid = false;
            if (rc) return rc;
            HIPCHK(hipStreamSynchronize(w->stream));
            for (int n = n0; n < n1; ++n) a->segdt[n - n0] = a->dt_pinned[n - n0];
        }
        for (int n = n1 - 1; n >= n0; --n) {
            const double* xn = n == n0 ? a->ckpt[sgm] : a->segx[n - n0];
            const double* x1b = a->seg1[n - n0];
            const double* x2b = a->seg2[n - n0];
            const double* xn1 =
                n + 1 < n1 ? a->segx[n + 1 - n0] : (n1 == N ? a->xfinal : a->ckpt[sgm + 1]);
            const double dt = a->segdt[n - n0];
            HIPCHK(hipMemsetAsync(a->acc + 2, 0, sizeof(double), w->stream));
            // stage 3: L = lam X_{n+1} -> B^T; lamF = g3 dt L; lamZ2 = L
            if ((rc = bcon_adj(a, xn1, L))) return rc;
            if ((rc = vjp(a, x2b, nullptr, 0.0, L, g3 * dt, Mb))) return rc;
            // stage 2: lamF = z2 dt lamZ2 + g2 dt lamX2; lamZ1 = lamX2
            if ((rc = bcon_adj(a, x2b, Mb))) return rc;
            if ((rc = vjp(a, x1b, L, z2 * dt, Mb, g2 * dt, Nb))) return rc;
            // stage 1: lamF = z1 dt lamZ1 + g1 dt lamX1; Z0 = X_n
            if ((rc = bcon_adj(a, x1b, Nb))) return rc;
            if ((rc = vjp(a, xn, Mb, z1 * dt, Nb, g1 * dt, L))) return rc;
            if (C.dtfix <= 0.0 && (rc = dt_adjoint(a, xn, dt, L))) return rc;
        }
    }
    double acc[6];
    HIPCHK(hipMemcpyAsync(acc, a->acc, sizeof acc, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    a->bwd_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
    // ore = 1/Re, tk = ocv/(Re Pr); acc[4] holds the dt limits' d/dRe
    const double g_re = acc[0] * (-C.ore * C.ore) + acc[1] * (-C.tk / C.re) + acc[4];
    if (grad)
        for (int f = 0; f < a->nf; ++f)
            if (grad[f] && (rc = device ? global_copy_device(w, L + f * a->fe, grad[f], true)
                                        : global_copy(w, L + f * a->fe, grad[f], true)))
                return rc;
    if (dJ_dre) *dJ_dre = g_re;
    if (dJ_dth) *dJ_dth = acc[3];
    a->dJ_dtb = acc[5];
    return R3D_OK;
}
}  // namespace

int r3d_adj_backward(r3d_adj_ctx* a, double* const* grad, double* dJ_dre, double* dJ_dth) {
    return backward(a, grad, dJ_dre, dJ_dth, false);
}

int r3d_adj_backward_device(r3d_adj_ctx* a, double* const* grad, double* dJ_dre,
                            double* dJ_dth) {
    return backward(a, grad, dJ_dre, dJ_dth, true);
}

int r3d_adj_set_fixed_dt(r3d_adj_ctx* a, double dt) {
    if (!a) return set_err(R3D_EINVAL, "null adjoint context");
    if (!(dt >= 0.0) || !std::isfinite(dt)) return set_err(R3D_EINVAL, "fixed dt %g", dt);
    a->work->C.dtfix = dt;
    a->nsteps = 0;  // a forward run under the new rule is needed before backward
    a->kept_last = false;
    return R3D_OK;
}

int r3d_adj_dJ_dtb(r3d_adj_ctx* a, double* dJ_dtb) {
    if (!a || !dJ_dtb) return set_err(R3D_EINVAL, "null argument");
    if (a->nsteps < 1) return set_err(R3D_ESTATE, "no forward run to differentiate");
    *dJ_dtb = a->dJ_dtb;
    return R3D_OK;
}

int r3d_adj_timing(r3d_adj_ctx* a, double* fwd, double* bwd) {
    if (!a) return set_err(R3D_EINVAL, "null adjoint context");
    if (fwd) *fwd = a->fwd_ms;
    if (bwd) *bwd = a->bwd_ms;
    return R3D_OK;
}

int r3d_adj_download_final(r3d_adj_ctx* a, double* const* fields) {
    if (!a || !fields) return set_err(R3D_EINVAL, "null argument");
    for (int f = 0; f < a->nf; ++f)
        if (fields[f]) {
            int rc = global_copy(a->work, a->xfinal + f * a->fe, fields[f], true);
            if (rc) return rc;
        }
    return R3D_OK;
}

}  // extern "C"

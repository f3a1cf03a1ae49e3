// r3d_stage.hpp - the fused RK3 stage kernel for gfx950 (hydrodynamic and MHD).
//
// One kernel launch = one RK stage over the whole domain:
//     F  = FLUXES(X)                      reference src/fluxes.cpp:44-533
//     X' = Z + gam_s dt F                 reference src/step.cpp:53-80
//     Z' = X' + zeta_s dt F               reference src/step.cpp:83-110 (next stage's mix)
// and, in the last stage, the per-cell dt limits of X' for the next step.
//
// Halo.  The reference refreshes ghost cells by a halo exchange after every
// stage (src/communicate.cpp:70-169).  Here a virtual rank's ghost cells are
// its neighbours' interior cells in the one global array, and the periodic
// x/y ghosts are never read during a step: a column or row past the global
// x/y edge is loaded from the interior cell it wraps to, which holds exactly
// the value the halo exchange would have copied.  The ghost images in HBM are
// written only when a state is downloaded (k_halo).
//
// Work decomposition.  A workgroup owns a 64 x TY column tile and a chunk of
// kc planes.  Waves 0..TY+1 each own one row of 64 columns (rows j0-1..j0+TY,
// the outer two are the y halo); wave TY+2 owns the 2*(TY+2) x-halo columns.
// Every thread marches up its column in z, holding planes k-1, k, k+1 of its
// own column in registers (the z stencil) and publishing the plane-k values
// its x/y neighbours need to a double-buffered LDS plane (one barrier per
// plane).  Only waves 1..TY compute fluxes; the halo waves only publish.
//
// MHD (MAG = true) adds BX, BY, BZ to the state and the LMAG terms of the
// Fortran FLUXES (golden/fortran/3dmhdsub.f:2146-2651): the Lorentz force on
// FU, FV, FW, the Ohmic heating on FT and the induction + resistive diffusion
// fluxes of B.  The reference C++ has no LMAG branch; these follow the
// Fortran's left-to-right evaluation of each statement.
//
// Bit-exactness.  This file must be compiled with -ffp-contract=off: every
// product and sum below is written in the operand order of the reference
// CPU sweep it cites, so each intermediate rounds identically.  Quantities the
// reference computes in a separate sweep over a wider index range (the
// velocities, P = rho T, the derivative scratch WW1..WW3) are recomputed at
// the neighbour from the same operands, which yields the same doubles.
#pragma once

#include "r3d_dev.hpp"

// The kernels live in namespace R3D_KNS: r3d for the bit-exact build; the
// FMA-contracted variant compiles this same code a second time, in its own
// namespace and translation unit, with -ffp-contract=fast (r3d_fma.hip).
#ifndef R3D_KNS
#define R3D_KNS r3d
#endif

#include "r3d_dtlim.hpp"

namespace R3D_KNS {
using namespace r3d;

// Quantities a column publishes to its in-plane neighbours.
enum { SH_RU = 0, SH_RV, SH_RW, SH_TT, SH_P, SH_U, SH_V, SH_W, SH_WZ, SH_UZ, SH_VZ, NSH,
       SH_BX = NSH, SH_BY, SH_BZ, NSH_MAG };

struct Raw {
    double ru, rv, rw, ro, tt, bx, by, bz;
};

template <bool B>
struct BoolConst {
    static constexpr bool value = B;
};

// Per-plane values of one column (reference src/fluxes.cpp:131-166: P = RO*TT
// before the inversion, 1/RO, velocities = momentum * (1/RO)); ruw..rww are
// the vertical advective fluxes m*w of src/fluxes.cpp:183-185.
struct Col {
    double ru, rv, rw, ro, tt, p, ri, u, v, w, ruw, rvw, rww, bx, by, bz;
};

// Global accesses of the stage kernel use a plane base that is uniform over
// the workgroup (field pointer + k * plane bytes, held in scalar registers)
// plus one 32-bit per-lane byte offset within the plane, so that no per-lane
// 64-bit address has to stay live across the z loop.
//
// They go through address-space-1 (global) pointers.  A generic pointer the
// compiler cannot trace to a kernel argument becomes a FLAT access, which
// also counts on the LDS counter: every later LDS read of the flux blocks
// would then wait (lgkmcnt(0)) until the preceding HBM store has completed.
//
// Stores and the Z loads are non-temporal (streaming): X' and Z' are not read
// again in this launch and Z is read by its own lane only, so they should not
// displace the X planes that neighbouring tiles still read through L2.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ const char* plane_base(const double* f, long pp, int k) {
    return reinterpret_cast<const char*>(f) + (size_t)pp * 8u * (size_t)k;
}
__device__ __forceinline__ double ldp(const double* f, long pp, int k, uint32_t voff) {
    return *(const gdouble*)(plane_base(f, pp, k) + voff);
}
__device__ __forceinline__ double ldp_stream(const double* f, long pp, int k, uint32_t voff) {
    return __builtin_nontemporal_load((const gdouble*)(plane_base(f, pp, k) + voff));
}
__device__ __forceinline__ void stp(double* f, long pp, int k, uint32_t voff, double v) {
    __builtin_nontemporal_store(v, (gdouble*)(const_cast<char*>(plane_base(f, pp, k)) + voff));
}

template <bool MAG>
__device__ __forceinline__ Raw load_raw(const State& s, long pp, int k, uint32_t voff) {
    Raw r;
    r.ru = ldp(s.q[QRU], pp, k, voff);
    r.rv = ldp(s.q[QRV], pp, k, voff);
    r.rw = ldp(s.q[QRW], pp, k, voff);
    r.ro = ldp(s.q[QRO], pp, k, voff);
    r.tt = ldp(s.q[QTT], pp, k, voff);
    if constexpr (MAG) {
        r.bx = ldp(s.q[QBX], pp, k, voff);
        r.by = ldp(s.q[QBY], pp, k, voff);
        r.bz = ldp(s.q[QBZ], pp, k, voff);
    } else {
        r.bx = r.by = r.bz = 0.0;
    }
    return r;
}

__device__ __forceinline__ Col derive(const Raw& r) {
    Col c;
    c.ru = r.ru;
    c.rv = r.rv;
    c.rw = r.rw;
    c.ro = r.ro;
    c.tt = r.tt;
    c.p = r.ro * r.tt;
    c.ri = 1.0 / r.ro;
    c.u = r.ru * c.ri;
    c.v = r.rv * c.ri;
    c.w = r.rw * c.ri;
    c.ruw = c.ru * c.w;
    c.rvw = c.rv * c.w;
    c.rww = c.rw * c.w;
    c.bx = r.bx;
    c.by = r.by;
    c.bz = r.bz;
    return c;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Give each XCD a contiguous range of logical tiles, y fastest, so
// that tiles resident at the same time on one XCD are y-neighbours and share
// their halo rows through that XCD's L2.  gridDim.x is a multiple of 8.
__device__ __forceinline__ int logical_block() {
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o, 64));
    return v;
}

// Write the block's dt extrema to partial[slot*NPART ...].  red must hold
// NPART doubles per wave.
template <int NW>
__device__ __forceinline__ void block_partials(double pm[NPART], double (*red)[NPART],
                                               double* partial, int slot) {
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    double v[NPART];
    v[0] = wave_max(pm[0]);
    for (int q = 1; q < NPART; ++q) v[q] = wave_min(pm[q]);
    if (ln == 0)
        for (int q = 0; q < NPART; ++q) red[wv][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[NPART];
        for (int q = 0; q < NPART; ++q) o[q] = red[0][q];
        for (int w = 1; w < NW; ++w) {
            o[0] = dmax(o[0], red[w][0]);
            for (int q = 1; q < NPART; ++q) o[q] = dmin(o[q], red[w][q]);
        }
        for (int q = 0; q < NPART; ++q) partial[(long)slot * NPART + q] = o[q];
    }
}

// Waves of a tile TW columns wide and TY rows high.  TW = 64: one row per
// wave (waves 1..TY the flux rows, 0 and TY+1 the y-halo rows, TY+2 the
// x-halo columns).  TW = 32 or 16: RPW = 64 / TW rows per wave (waves
// 0..TY/RPW-1 the flux rows, TY/RPW both y-halo rows, TY/RPW+1 the x-halo
// columns).
template <int TY, int TW>
struct TileShape {
    static constexpr int RPW = 64 / TW;  // rows per wave
    static_assert((TW == 64 || TW == 32 || TW == 16) && TY % RPW == 0 && TY + 2 <= 32,
                  "tile width 64, 32 or 16 with whole waves of flux rows");
    static constexpr int ROWS = TY + 2;
    static constexpr int NFW = TY / RPW;  // flux waves
    static constexpr int NW = TW == 64 ? TY + 3 : NFW + 2;
};

// LDS of one stage tile.  The kernel that runs the tile declares it, so that
// a persistent kernel can run the three stages in one allocation (a union).
template <int TY, int STAGE, bool MAG, bool HEAT, bool DEN, int TW = 64>
struct StageLds {
    static constexpr int ROWS = TY + 2, NW = TileShape<TY, TW>::NW, NFW = TileShape<TY, TW>::NFW;
    static constexpr int KCMAX = 128;
    static constexpr int NZX = (DEN ? 6 : 0) + (HEAT ? 1 : 0);
    static constexpr int NPM = MAG ? NPART : NPART - 1;  // extrema in use (no resistive limit without B)
    double sh[2][(MAG ? NSH_MAG : NSH) + (DEN ? 1 : 0)][ROWS][TW + 2];
    double red[NW][NPART];
    // Per-plane z metrics of this block's chunk, staged once: they are
    // uniform per plane, and fetching them inside the z loop would put a
    // global-memory round trip on every plane.
    // dzz, d2zz, ddz, the conductivity (rkapa; 1 + rkapa with LREM, whose
    // diffusion does not read it), dkapa
    double zm[5][KCMAX];
    // DEN / HEAT per-plane values, staged with the metrics for the same
    // reason: ROM at k-1..k+1 (row 0, offset by one plane), DH*ORE*WWZ,
    // DV*ORE*WWZ, RLAX*TFAC*CORRECT, then the heat loss z Gaussian; and the heat loss x / y
    // Gaussians of the tile's columns and rows, so that they hold no
    // registers across the z loop
    // and with LREM TTM at k-1..k+1 (row ZX_DEN + 4, offset by one plane) and HRAD
    double zx[NZX > 0 ? NZX : 1][KCMAX + 2];
    double hgx_s[HEAT ? TW + 2 : 1], hgy_s[HEAT ? ROWS : 1];
    // Stage 3: the column's min(ddx, ddy) and each flux row's running dt
    // extrema live in LDS rather than in registers across the z loop.  Held
    // in registers they are spilled, and a scratch reload after a plane's
    // stores waits for all of them (vector memory operations complete in
    // issue order); an LDS access does not.  The extrema are kept per flux
    // wave, merged by a wave reduction on the rare planes where some cell
    // of the wave passes the time-step screen.
    double ddxy_s[STAGE == 2 ? ROWS : 1][TW + 2];
    double pmw_s[STAGE == 2 ? NPM : 1][STAGE == 2 ? NFW : 1];
    // stage 3 screen thresholds from the previous reduction (DevScal); read
    // from LDS where they are used, so they hold no register across the loop
    double scr_s[NSC];
};

// One tile t (0 <= t < a.nblocks) of one RK stage; the caller maps blocks to
// tiles and provides the LDS.  ZMODE: 0 = Z' is the next stage's mix, 1 = Z
// keeps the saved state (a step truncated after stage 1), 2 = Z is not
// written (last stage).
// HEAT: the ITCON=2 embedded heat loss term is added to FT; DEN: density
// diffusion is added to FR.  Separate instantiations, so that the default
// kernels carry none of their registers or LDS.
template <int TY, int STAGE, int ZMODE, bool MAG, bool HEAT = false, bool DEN = false, int TW = 64>
__device__ __forceinline__ void stage_tile(const StageArgs& a, const Consts& C, const Metrics& M,
                                           const int t,
                                           StageLds<TY, STAGE, MAG, HEAT, DEN, TW>& L) {
    using LT = StageLds<TY, STAGE, MAG, HEAT, DEN, TW>;
    constexpr int ROWS = LT::ROWS;
    constexpr int NW = LT::NW;
    constexpr int NF = MAG ? 8 : 5;  // primaries per state
    constexpr int SH_RO = MAG ? NSH_MAG : NSH;  // DEN only: raw RO of the plane
    constexpr int KCMAX = LT::KCMAX;
    constexpr int ZX_DEN = 0, ZX_HEAT = DEN ? 6 : 0;
    constexpr int NPM = LT::NPM;
    auto& sh = L.sh;
    auto& red = L.red;
    auto& zm = L.zm;
    auto& zx = L.zx;
    auto& hgx_s = L.hgx_s;
    auto& hgy_s = L.hgy_s;
    auto& ddxy_s = L.ddxy_s;
    auto& pmw_s = L.pmw_s;
    auto& scr_s = L.scr_s;
    // Stage 1 takes the dt reduced at the end of the previous step (or by the
    // standalone dt pass) and makes it the step's dt.
    const double dt = (STAGE == 0) ? a.scal->dt_next : a.scal->dt;
    if (STAGE == 0 && t == 0 && threadIdx.x == 0) {
        a.scal->dt = a.scal->dt_next;
        a.scal->umach = a.scal->umach_next;
    }
    // x fastest: the tiles an XCD runs at the same time are x-neighbours
    // (whole tile rows), whose x-halo columns lie in each other's cache lines
    const int tx = t % a.ntx;
    const int rem = t / a.ntx;
    const int ty = rem % a.nty;
    const int kz = rem / a.nty;
    const Geo g = a.g;
    const int i0 = 1 + TW * tx, j0 = 1 + TY * ty;
    const int k0 = 1 + a.kc * kz;
    const int k1 = min(g.npz, k0 + a.kc - 1);

    // the wave index is uniform over the wave; from a scalar register, the
    // role branch below is a scalar branch
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = threadIdx.x & 63;
    constexpr int NFW = LT::NFW;
    // LDS row and column of this lane (column 1..TW inside the tile)
    int r, c, gi, gj;
    bool act;
    if (TW == 64 ? wv < ROWS : wv <= NFW) {
        if constexpr (TW == 64) {
            r = wv;
            c = ln + 1;
        } else {  // RPW rows per wave; wave NFW holds the two y-halo rows
            constexpr int RPW = TileShape<TY, TW>::RPW;
            const int q = ln / TW;
            r = wv < NFW ? 1 + RPW * wv + q : (q == 1 ? TY + 1 : 0);
            c = ln % TW + 1;
            act = wv < NFW || q < 2;  // y-halo wave: lanes past the two rows idle
        }
        gi = i0 + c - 1;
        gj = j0 - 1 + r;
        if constexpr (TW == 64) act = true;
    } else {
        act = ln < 2 * ROWS;
        r = act ? (ln >> 1) : 0;
        c = (ln & 1) ? TW + 1 : 0;
        gi = (ln & 1) ? i0 + TW : i0 - 1;
        gj = j0 - 1 + r;
    }
    const bool heavy = TW == 64 ? (wv >= 1) && (wv <= TY) : wv < NFW;
    // the flux wave's slot for its running dt extrema
    const int fw = TW == 64 ? wv - 1 : wv;
    const bool outp = heavy && gi <= g.npx && gj <= g.npy;
    // the column this lane reads: periodic wrap at the x/y edges; lanes past
    // the wrap (partial edge tiles) read column/row 1 and their stores go to
    // the row padding
    const int li = gi == 0 ? g.npx : (gi > g.npx ? 1 : gi);
    const int lj = gj == 0 ? g.npy : (gj > g.npy ? 1 : gj);
    // byte offset of this column within a plane (a plane is < 4 GiB)
    uint32_t voff = (uint32_t)((XOFF + li + (long)g.px * lj) * 8);
    // Stores of a heavy lane that is not an output cell go to element 0 of
    // its row, inside the unused row padding (interior i = 1 starts at
    // element XOFF + 1), so that the store stream stays unconditional.
    uint32_t soff = outp ? voff : (uint32_t)((long)g.px * lj * 8);

    for (int q = threadIdx.x; q < 5 * KCMAX; q += blockDim.x) {
        const int f = q / KCMAX, kk = k0 + q % KCMAX;
        if (kk <= k1) {
            const double* src = f == 0 ? M.dzz : f == 1 ? M.d2zz : f == 2 ? M.ddz : f == 3 ? M.rkdt : M.dkapa;
            zm[f][kk - k0] = src[kk];
        }
    }
    if constexpr (DEN) {
        for (int q = threadIdx.x; q < KCMAX + 2; q += blockDim.x) {
            const int kk = k0 - 1 + q;  // planes k0-1 .. k1+1 (within 0 .. npz+1)
            if (kk <= k1 + 1) {
                zx[ZX_DEN][q] = M.rom[kk];
                zx[ZX_DEN + 4][q] = M.ttm[kk];
            }
            if (q < KCMAX && k0 + q <= k1) {
                zx[ZX_DEN + 1][q] = M.dth[k0 + q];
                zx[ZX_DEN + 2][q] = M.dtv[k0 + q];
                zx[ZX_DEN + 3][q] = M.ftc[k0 + q];
                zx[ZX_DEN + 5][q] = M.hrad[k0 + q];
            }
        }
    }
    if (STAGE == 2 && a.screen && threadIdx.x == 0) {
        screen_consts(C, a.scal->lim_prev, a.scal->um_prev, scr_s);
    }
    if constexpr (HEAT) {
        for (int q = threadIdx.x; q < KCMAX; q += blockDim.x)
            if (k0 + q <= k1) zx[ZX_HEAT][q] = M.hgz[k0 + q];
        for (int q = threadIdx.x; q < TW + 2 + ROWS; q += blockDim.x) {
            if (q < TW + 2) {  // column q: global i0 - 1 + q, wrapped as the lanes read it
                const int gi2 = i0 - 1 + q;
                hgx_s[q] = M.hgx[gi2 == 0 ? g.npx : (gi2 > g.npx ? 1 : gi2)];
            } else {
                const int gj2 = j0 - 1 + (q - TW - 2);
                hgy_s[q - TW - 2] = M.hgy[gj2 == 0 ? g.npy : (gj2 > g.npy ? 1 : gj2)];
            }
        }
    }
    __syncthreads();

    const double cg = C.gam[STAGE] * dt;
    const double cz = (STAGE < 2) ? C.zeta[STAGE < 2 ? STAGE : 0] * dt : 0.0;

    // x/y metric factors of this column (src/fluxes.cpp:90, 250-254)
    const double dxx = M.dxx[li], d2xx = M.d2xx[li];
    const double dyy = M.dyy[lj], d2yy = M.d2yy[lj];
    if (STAGE == 2 && heavy) {
        ddxy_s[r][c] = dmin(M.ddx[li], M.ddy[lj]);
        if (ln == 0) {
            pmw_s[0][fw] = 0.0;
            for (int q = 1; q < NPM; ++q) pmw_s[q][fw] = 1e300;
        }
    }

    Col cm = derive(load_raw<MAG>(a.xin, g.pp, k0 - 1, voff));
    Col cc = derive(load_raw<MAG>(a.xin, g.pp, k0, voff));
    double rw_m2 = ldp(a.xin.q[QRW], g.pp, max(k0 - 2, 0), voff);
    Raw rn = load_raw<MAG>(a.xin, g.pp, min(k0 + 1, g.npz + 1), voff);

    double pm[NPART] = {0.0, 1e300, 1e300, 1e300, 1e300};

    // The z loop is instantiated twice: for the flux rows (HEAVY) and for the
    // halo waves, which only load and publish.  Both run one barrier per
    // plane over the same planes.  With one loop for both, the compiler's
    // wait counts at the loop head would be those of the halo path, which
    // issues no stores: a flux wave would then wait there for all of the
    // previous plane's stores to complete.
    const int r_all = r, c_all = c;
    auto zloop = [&](auto heavy_c) {
    constexpr bool HEAVY = decltype(heavy_c)::value;
    // with 64-column tiles a flux row's LDS row is its (uniform) wave index
    // and its column the lane + 1: no per-lane row index has to stay live in
    // that loop
    const int r = HEAVY && TW == 64 ? wv : r_all;
    const int c = HEAVY && TW == 64 ? ln + 1 : c_all;
    // Every load issued before the loop has landed before it starts.  The
    // compiler's wait counts at the loop head merge the entry path with the
    // back edge: a load still in flight on entry would otherwise make every
    // plane wait, right after its barrier, for all of the previous plane's
    // stores (vector memory operations retire in issue order).
    __builtin_amdgcn_s_waitcnt(0);
    for (int k = k0; k <= k1; ++k) {
        const int buf = k & 1;
        // Keep the lane offsets opaque to loop-invariant code motion: every
        // access is then (uniform plane base) + zext(voff), the saddr form,
        // instead of one hoisted 64-bit per-lane address per field.
        asm volatile("" : "+v"(voff));
        asm volatile("" : "+v"(soff));
        // This plane's Z is issued before the next-but-one X plane, in the
        // order the flux blocks consume it (FR first): the first RK update
        // then waits for the Z of RO only, not for the X loads it does not use.
        double zc[NF];
        if constexpr (STAGE > 0 && HEAVY) {
            zc[QRO] = ldp_stream(a.z.q[QRO], g.pp, k, voff);
            zc[QRU] = ldp_stream(a.z.q[QRU], g.pp, k, voff);
            zc[QRV] = ldp_stream(a.z.q[QRV], g.pp, k, voff);
            zc[QRW] = ldp_stream(a.z.q[QRW], g.pp, k, voff);
            zc[QTT] = ldp_stream(a.z.q[QTT], g.pp, k, voff);
            for (int f = QBX; f < NF; ++f) zc[f] = ldp_stream(a.z.q[f], g.pp, k, voff);
            __builtin_amdgcn_sched_barrier(0);
        }
        const Raw rnn = load_raw<MAG>(a.xin, g.pp, min(k + 2, g.npz + 1), voff);
        const Col cp = derive(rn);
        const int kl = k - k0;
        const double dzz = zm[0][kl], d2zz = zm[1][kl];
        // vertical derivatives of this column at plane k
        // (WW3/WW1/WW2 of src/fluxes.cpp:423, 464, 469)
        const double wz = (cp.w - cm.w) * C.hz * dzz;
        const double uz = (cp.u - cm.u) * C.hz * dzz;
        const double vz = (cp.v - cm.v) * C.hz * dzz;
        if (HEAVY || act) {
            sh[buf][SH_RU][r][c] = cc.ru;
            sh[buf][SH_RV][r][c] = cc.rv;
            sh[buf][SH_RW][r][c] = cc.rw;
            sh[buf][SH_TT][r][c] = cc.tt;
            sh[buf][SH_P][r][c] = cc.p;
            sh[buf][SH_U][r][c] = cc.u;
            sh[buf][SH_V][r][c] = cc.v;
            sh[buf][SH_W][r][c] = cc.w;
            sh[buf][SH_WZ][r][c] = wz;
            sh[buf][SH_UZ][r][c] = uz;
            sh[buf][SH_VZ][r][c] = vz;
            if constexpr (MAG) {
                sh[buf][SH_BX][r][c] = cc.bx;
                sh[buf][SH_BY][r][c] = cc.by;
                sh[buf][SH_BZ][r][c] = cc.bz;
            }
            if constexpr (DEN) sh[buf][SH_RO][r][c] = cc.ro;
        }
        __syncthreads();

        if constexpr (HEAVY) {
#define SH(q, rr, cc_) sh[buf][q][rr][cc_]
            // One flux component at a time: each block reads only its own LDS
            // operands, accumulates its terms in the reference's order, applies
            // the RK update and stores, so that only finished values stay live
            // (one scheduling barrier, after FU, in the hydrodynamic kernels).
            // Without barriers between the other blocks the compiler issues a
            // block's LDS reads during the previous block, which hides their
            // latency, and merges the derivatives that several blocks
            // recompute from the same operands (same doubles).  Measured per
            // stage launch: the MHD kernels (two waves per SIMD) 7% faster
            // with no barrier at all; the hydrodynamic ones (three waves per
            // SIMD) 3% faster with the one after FU than with one after every
            // block, and 35-50% slower with none (their registers grow and
            // the loads of the first blocks are scheduled late).
            // Derivatives that several components need (e.g. d(UU)/dx) are
            // recomputed from the same operands, which gives the same doubles.
            const bool bplane = (k == 1) || (k == g.npz);
            // y metric products, recomputed per plane (cheaper than holding them)
            const double tmpy = C.hy * dyy;
            const double tmpy1 = C.hy * d2yy;
            const double tmpy2 = C.h2y * dyy * dyy;
            if (STAGE == 0) {
                zc[0] = cc.ru;  // stage 1: Z = X (save_state, src/step.cpp:38-50)
                zc[1] = cc.rv;
                zc[2] = cc.rw;
                zc[3] = cc.ro;
                zc[4] = cc.tt;
                if constexpr (MAG) {
                    zc[5] = cc.bx;
                    zc[6] = cc.by;
                    zc[7] = cc.bz;
                }
            }
            const double tmpz = C.hz * dzz;
            const double tmpz1 = C.hz * d2zz;
            const double tmpz2 = C.h2z * dzz * dzz;
            const double hxd = C.hx;
            double xv[NF];
            // RK update of component f (src/step.cpp:53-80), its stores and the
            // next stage's mix (src/step.cpp:83-110) or the saved state.
            // The stores are unconditional so that the components form basic
            // blocks the scheduler can interleave; a heavy lane that is not an
            // output cell stores into its row's padding.
            auto finish = [&](int f, double fval, double cur) {
                const double x = zc[f] + cg * fval;
                xv[f] = x;
                stp(a.xout.q[f], g.pp, k, soff, x);
                if (ZMODE == 1) {
                    stp(a.z.q[f], g.pp, k, soff, cur);
                } else if (ZMODE == 0 && STAGE < 2) {
                    // on the two BCON planes Z' holds F until the boundary
                    // kernel mixes it with the post-BCON state
                    stp(a.z.q[f], g.pp, k, soff, bplane ? fval : x + cz * fval);
                }
            };
#define LAPX(e, ww, c2) ((e - ww) * hxd * d2xx + (e - c2 + ww) * C.h2x * dxx * dxx)
// LMAG derivatives with the Fortran's left-to-right metric products
// ((diff * h) * metric), golden/fortran/3dmhdsub.f:2150-2651
#define DXS(q) ((SH(q, r, c + 1) - SH(q, r, c - 1)) * hxd * dxx)
#define DYS(q) ((SH(q, r + 1, c) - SH(q, r - 1, c)) * C.hy * dyy)
#define DZC(p_, m_) ((p_ - m_) * C.hz * dzz)
#define LAPY(n_, s_, c2) ((n_ - s_) * C.hy * d2yy + (n_ - c2 + s_) * C.h2y * dyy * dyy)
#define LAPZ(p_, m_, c2) ((p_ - m_) * C.hz * d2zz + (p_ - c2 + m_) * C.h2z * dzz * dzz)
            {  // ---- FR: continuity (src/fluxes.cpp:87-121) ----
                const double w1top = (4.0 * cp.rw - rnn.rw) * tmpz;  // one-sided, top
                const double w1bot = (rw_m2 - 4.0 * cm.rw) * tmpz;   // one-sided, bottom
                const double w1mid = (cp.rw - cm.rw) * tmpz;
                const double w1 = k == 1 ? w1top : (k == g.npz ? w1bot : w1mid);
                double fr = (SH(SH_RU, r, c - 1) - SH(SH_RU, r, c + 1)) * hxd * dxx;
                fr = fr - (SH(SH_RV, r + 1, c) - SH(SH_RV, r - 1, c)) * tmpy;
                fr = fr - w1;
                if constexpr (DEN) {
                    // diffusion of density perturbations WW1 = RO - ROM(K),
                    // golden/fortran/3dmhdsub.f:1316-1380 (operation order kept)
                    const double rm = zx[ZX_DEN][kl + 1];
                    const double w0 = cc.ro - rm;
                    if (C.den_h) {
                        const double wE = SH(SH_RO, r, c + 1) - rm, wW = SH(SH_RO, r, c - 1) - rm;
                        const double wN = SH(SH_RO, r + 1, c) - rm, wS = SH(SH_RO, r - 1, c) - rm;
                        const double ww2 =
                            (wE - wW) * hxd * d2xx + (wE - 2.0 * w0 + wW) * C.h2x * dxx * dxx;
                        const double ww3 = (wN - wS) * tmpy1 + (wN - 2.0 * w0 + wS) * tmpy2;
                        fr = fr + (ww2 + ww3) * zx[ZX_DEN + 1][kl];
                    }
                    if (C.den_v) {
                        const double wU = cp.ro - zx[ZX_DEN][kl + 2];
                        const double wD = cm.ro - zx[ZX_DEN][kl];
                        const double ww2 = (wU - wD) * tmpz1 + (wU - 2.0 * w0 + wD) * tmpz2;
                        fr = fr + ww2 * zx[ZX_DEN + 2][kl];
                    }
                }
                finish(QRO, fr, cc.ro);
            }
            
            {  // ---- FU (src/fluxes.cpp:149, 177-185, 289-297, 367, 449, 523) ----
                const double uE = SH(SH_U, r, c + 1), uW = SH(SH_U, r, c - 1);
                const double uN = SH(SH_U, r + 1, c), uS = SH(SH_U, r - 1, c);
                const double u2 = 2.0 * cc.u;
                double fu = (SH(SH_P, r, c - 1) - SH(SH_P, r, c + 1)) * hxd * dxx;
                fu = fu - (SH(SH_RU, r, c + 1) * uE - SH(SH_RU, r, c - 1) * uW) * hxd * dxx;
                fu = fu - (SH(SH_RU, r + 1, c) * SH(SH_V, r + 1, c) -
                           SH(SH_RU, r - 1, c) * SH(SH_V, r - 1, c)) * tmpy;
                fu = fu - (cp.ruw - cm.ruw) * tmpz;
                fu = fu + C.ore * (C.c43 * LAPX(uE, uW, u2) + (uN - uS) * tmpy1 +
                                   (uN - u2 + uS) * tmpy2 + (cp.u - cm.u) * tmpz1 +
                                   (cp.u - u2 + cm.u) * tmpz2);
                const double vxN = (SH(SH_V, r + 1, c + 1) - SH(SH_V, r + 1, c - 1)) * hxd * dxx;
                const double vxS = (SH(SH_V, r - 1, c + 1) - SH(SH_V, r - 1, c - 1)) * hxd * dxx;
                fu = fu + C.t2 * (vxN - vxS) * C.hy * dyy;
                fu = fu + C.t2 * (SH(SH_WZ, r, c + 1) - SH(SH_WZ, r, c - 1)) * hxd * dxx;
                if (C.lrot) fu = fu + C.omz * cc.rv;
                if constexpr (MAG) {  // Lorentz force, 3dmhdsub.f:1190, 1340
                    const double bxy = DYS(SH_BX), byx = DXS(SH_BY), bzx = DXS(SH_BZ);
                    const double bxz = DZC(cp.bx, cm.bx);
                    fu = fu + C.obeta * cc.by * (bxy - byx) - C.obeta * cc.bz * bzx;
                    fu = fu + C.obeta * cc.bz * bxz;
                }
                finish(QRU, fu, cc.ru);
            }
            
            if constexpr (!MAG) __builtin_amdgcn_sched_barrier(0);
            {  // ---- FV (src/fluxes.cpp:148, 195-203, 298-307, 373, 455, 528) ----
                const double vE = SH(SH_V, r, c + 1), vW = SH(SH_V, r, c - 1);
                const double vN = SH(SH_V, r + 1, c), vS = SH(SH_V, r - 1, c);
                const double v2 = 2.0 * cc.v;
                double fv = (SH(SH_P, r - 1, c) - SH(SH_P, r + 1, c)) * tmpy;
                fv = fv - (SH(SH_RV, r, c + 1) * SH(SH_U, r, c + 1) -
                           SH(SH_RV, r, c - 1) * SH(SH_U, r, c - 1)) * hxd * dxx;
                fv = fv - (SH(SH_RV, r + 1, c) * vN - SH(SH_RV, r - 1, c) * vS) * tmpy;
                fv = fv - (cp.rvw - cm.rvw) * tmpz;
                fv = fv + C.ore * (LAPX(vE, vW, v2) +
                                   C.c43 * ((vN - vS) * tmpy1 + (vN - v2 + vS) * tmpy2) +
                                   (cp.v - cm.v) * tmpz1 + (cp.v - v2 + cm.v) * tmpz2);
                const double uxN = (SH(SH_U, r + 1, c + 1) - SH(SH_U, r + 1, c - 1)) * hxd * dxx;
                const double uxS = (SH(SH_U, r - 1, c + 1) - SH(SH_U, r - 1, c - 1)) * hxd * dxx;
                fv = fv + C.t2 * (uxN - uxS) * C.hy * dyy;
                fv = fv + C.t2 * (SH(SH_WZ, r + 1, c) - SH(SH_WZ, r - 1, c)) * C.hy * dyy;
                if (C.lrot) fv = fv + (-C.omz * cc.ru + C.omx * cc.rw);
                if constexpr (MAG) {  // Lorentz force, 3dmhdsub.f:1200, 1460
                    const double byx = DXS(SH_BY), bxy = DYS(SH_BX);
                    const double byz = DZC(cp.by, cm.by), bzy = DYS(SH_BZ);
                    fv = fv + C.obeta * cc.bx * (byx - bxy);
                    fv = fv + C.obeta * cc.bz * (byz - bzy);
                }
                finish(QRV, fv, cc.rv);
            }
            
            {  // ---- FW (src/fluxes.cpp:129, 147, 213-221, 308-317, 496, 532) ----
                const double wE = SH(SH_W, r, c + 1), wW = SH(SH_W, r, c - 1);
                const double wN = SH(SH_W, r + 1, c), wS = SH(SH_W, r - 1, c);
                const double w2 = 2.0 * cc.w;
                double fw = C.grav * cc.ro;
                fw = fw - (cp.p - cm.p) * tmpz;
                fw = fw - (SH(SH_RW, r, c + 1) * SH(SH_U, r, c + 1) -
                           SH(SH_RW, r, c - 1) * SH(SH_U, r, c - 1)) * hxd * dxx;
                fw = fw - (SH(SH_RW, r + 1, c) * SH(SH_V, r + 1, c) -
                           SH(SH_RW, r - 1, c) * SH(SH_V, r - 1, c)) * tmpy;
                fw = fw - (cp.rww - cm.rww) * tmpz;
                fw = fw + C.ore * (LAPX(wE, wW, w2) + (wN - wS) * tmpy1 + (wN - w2 + wS) * tmpy2 +
                                   C.c43 * ((cp.w - cm.w) * tmpz1 + (cp.w - w2 + cm.w) * tmpz2));
                fw = fw + C.t2 * ((SH(SH_UZ, r, c + 1) - SH(SH_UZ, r, c - 1)) * hxd * dxx +
                                  (SH(SH_VZ, r + 1, c) - SH(SH_VZ, r - 1, c)) * C.hy * dyy);
                if (C.lrot) fw = fw + -C.omx * cc.rv;
                if constexpr (MAG) {  // Lorentz force, 3dmhdsub.f:1210, 1350, 1470
                    const double bzx = DXS(SH_BZ), bxz = DZC(cp.bx, cm.bx);
                    const double byz = DZC(cp.by, cm.by), bzy = DYS(SH_BZ);
                    fw = fw + C.obeta * cc.bx * bzx;
                    fw = fw - C.obeta * cc.bx * bxz;
                    fw = fw - C.obeta * cc.by * (byz - bzy);
                }
                finish(QRW, fw, cc.rw);
            }
            
            {  // ---- FT (src/fluxes.cpp:235-272, 349-512) ----
                const double rk = zm[3][kl], dk = zm[4][kl];
                const double tt0 = cc.tt, ttU = cp.tt, ttD = cm.tt;
                const double ttE = SH(SH_TT, r, c + 1), ttW = SH(SH_TT, r, c - 1);
                const double ttN = SH(SH_TT, r + 1, c), ttS = SH(SH_TT, r - 1, c);
                const double t2c = 2.0 * tt0;
                double ft = -1.0 * cc.u * (ttE - ttW) * hxd * dxx;
                ft = ft - cc.v * (ttN - ttS) * tmpy;
                ft = ft - cc.w * (ttU - ttD) * tmpz;
                bool lrem_diff = false;
                if constexpr (DEN) lrem_diff = C.lrem;
                if (!lrem_diff) {
                    ft = ft + ((ttE - ttW) * hxd * d2xx + (ttE - t2c + ttW) * C.h2x * dxx * dxx) *
                                  cc.ri * C.tk * rk;
                    ft = ft + ((ttN - ttS) * tmpy1 + (ttN - t2c + ttS) * tmpy2) * cc.ri * C.tk * rk;
                    ft = ft + ((ttU - ttD) * tmpz1 + (ttU - t2c + ttD) * tmpz2) * cc.ri * C.tk * rk;
                    ft = ft + (ttU - ttD) * C.hz * dzz * cc.ri * C.tk * dk;
                }
                if constexpr (DEN) {
                    if (C.lrem) {
                        // LREM: diffusion of WW1 = TT - TTM(K) with unit conductivity,
                        // then RO*TMP*HRAD(K) (golden/fortran/3dmhdsub.f:1551-1586)
                        const double tm = zx[ZX_DEN + 4][kl + 1];
                        const double tmU = zx[ZX_DEN + 4][kl + 2], tmD = zx[ZX_DEN + 4][kl];
                        const double w0 = tt0 - tm, wE = ttE - tm, wW = ttW - tm;
                        const double wN = ttN - tm, wS = ttS - tm;
                        const double wU = ttU - tmU, wD = ttD - tmD;
                        ft = ft + ((wE - wW) * hxd * d2xx + (wE - 2.0 * w0 + wW) * C.h2x * dxx * dxx) *
                                      cc.ri * C.tk;
                        ft = ft + ((wN - wS) * tmpy1 + (wN - 2.0 * w0 + wS) * tmpy2) * cc.ri * C.tk;
                        ft = ft + ((wU - wD) * tmpz1 + (wU - 2.0 * w0 + wD) * tmpz2) * cc.ri * C.tk;
                        ft = ft + cc.ri * C.tk * zx[ZX_DEN + 5][kl];
                    }
                    // flux relaxation, golden/fortran/3dmhdsub.f:1765-1771
                    if (C.rlx) ft = ft + zx[ZX_DEN + 3][kl];
                }
                if constexpr (HEAT) {
                    // localized embedded heat loss, golden/fortran/3dmhdsub.f:1799-1822
                    // (RO there already holds 1/RO, :1405), same operation order
                    ft = ft - C.hloss * cc.ri * hgx_s[c] / C.hh * hgy_s[r] / C.hh *
                                  zx[ZX_HEAT][kl] / C.hh;
                }
                const double ux = (SH(SH_U, r, c + 1) - SH(SH_U, r, c - 1)) * hxd * dxx;
                const double vx = (SH(SH_V, r, c + 1) - SH(SH_V, r, c - 1)) * hxd * dxx;
                const double wx = (SH(SH_W, r, c + 1) - SH(SH_W, r, c - 1)) * hxd * dxx;
                ft = ft + (C.c43 * ux * ux + vx * vx + wx * wx) * cc.ri * C.t1;
                ft = ft - C.ocv * ux * tt0;
                const double uy = (SH(SH_U, r + 1, c) - SH(SH_U, r - 1, c)) * C.hy * dyy;
                ft = ft + (uy * uy + 2.0 * uy * vx) * cc.ri * C.t1;
                const double vy = (SH(SH_V, r + 1, c) - SH(SH_V, r - 1, c)) * C.hy * dyy;
                ft = ft + vy * vy * cc.ri * C.t3;
                ft = ft - C.ocv * vy * tt0;
                ft = ft + (wz * wz - ux * wz - ux * vy - vy * wz) * cc.ri * C.t3;
                ft = ft - C.ocv * wz * tt0;
                const double wy = (SH(SH_W, r + 1, c) - SH(SH_W, r - 1, c)) * C.hy * dyy;
                ft = ft + (uz * uz + vz * vz + wy * wy + 2.0 * vz * wy) * cc.ri * C.t1;
                ft = ft + 2.0 * uz * wx * cc.ri * C.t1;
                if constexpr (MAG) {  // Ohmic heating, 3dmhdsub.f:1220, 1360, 1480
                    const double bxy = DYS(SH_BX), byx = DXS(SH_BY), bzx = DXS(SH_BZ);
                    const double bxz = DZC(cp.bx, cm.bx);
                    const double byz = DZC(cp.by, cm.by), bzy = DYS(SH_BZ);
                    ft = ft + (bxy * bxy + byx * byx + bzx * bzx - 2.0 * byx * bxy) * cc.ri * C.ohm;
                    ft = ft + (bxz * bxz - 2.0 * bxz * bzx) * cc.ri * C.ohm;
                    ft = ft + (byz * byz + bzy * bzy - 2.0 * byz * bzy) * cc.ri * C.ohm;
                }
                finish(QTT, ft, cc.tt);
            }
            if constexpr (MAG) {
                // d(WW)/dz and the z second derivatives of BX, BY are one-sided
                // on the walls (3dmhdsub.f:1125-1126, 1391-1392, 1511-1513); the
                // bottom wall only when npez > 1 (ELSE IF).  k is uniform over
                // the block, so these are scalar branches; the extra planes
                // they need are read directly.
                const bool wtop = k == 1, wbot = k == g.npz && C.mag_bot;
                {  // ---- WW1, flux of BX (3dmhdsub.f:1080, 1130, 1230, 1280, 1370, 1410) ----
                    const double b2c = 2.0 * cc.bx;
                    double wzm = DZC(cp.w, cm.w);
                    double lbz = LAPZ(cp.bx, cm.bx, b2c);
                    if (wtop) {
                        const double b3 = ldp(a.xin.q[QBX], g.pp, 4, voff);
                        wzm = (4.0 * cp.w - rnn.rw * (1.0 / rnn.ro)) * C.hz * dzz;
                        lbz = (-3.0 * cc.bx + 4.0 * cp.bx - rnn.bx) * C.hz * d2zz +
                              (2.0 * cc.bx - 5.0 * cp.bx + 4.0 * rnn.bx - b3) * C.h2z * dzz * dzz;
                    } else if (wbot) {
                        const double ro2 = ldp(a.xin.q[QRO], g.pp, k - 2, voff);
                        const double b2 = ldp(a.xin.q[QBX], g.pp, k - 2, voff);
                        const double b3 = ldp(a.xin.q[QBX], g.pp, k - 3, voff);
                        wzm = (rw_m2 * (1.0 / ro2) - 4.0 * cm.w) * C.hz * dzz;
                        lbz = (3.0 * cc.bx - 4.0 * cm.bx + b2) * C.hz * d2zz +
                              (2.0 * cc.bx - 5.0 * cm.bx + 4.0 * b2 - b3) * C.h2z * dzz * dzz;
                    }
                    const double bE = SH(SH_BX, r, c + 1), bW = SH(SH_BX, r, c - 1);
                    const double bN = SH(SH_BX, r + 1, c), bS = SH(SH_BX, r - 1, c);
                    double w1 = cc.by * DYS(SH_U) - cc.bx * DYS(SH_V);
                    w1 = w1 + cc.bz * DZC(cp.u, cm.u) - cc.bx * wzm;
                    w1 = w1 - cc.u * DXS(SH_BX) - cc.v * DYS(SH_BX);
                    w1 = w1 + C.orm * (LAPX(bE, bW, b2c) + LAPY(bN, bS, b2c));
                    w1 = w1 - cc.w * DZC(cp.bx, cm.bx);
                    w1 = w1 + C.orm * lbz;
                    finish(QBX, w1, cc.bx);
                }
                {  // ---- WW2, flux of BY (3dmhdsub.f:1030, 1140, 1240, 1290, 1380, 1420, 1490, 1530) ----
                    const double b2c = 2.0 * cc.by;
                    double wzm = DZC(cp.w, cm.w);
                    double lbz = LAPZ(cp.by, cm.by, b2c);
                    if (wtop) {
                        const double b3 = ldp(a.xin.q[QBY], g.pp, 4, voff);
                        wzm = (4.0 * cp.w - rnn.rw * (1.0 / rnn.ro)) * C.hz * dzz;
                        lbz = (-3.0 * cc.by + 4.0 * cp.by - rnn.by) * C.hz * d2zz +
                              (2.0 * cc.by - 5.0 * cp.by + 4.0 * rnn.by - b3) * C.h2z * dzz * dzz;
                    } else if (wbot) {
                        const double ro2 = ldp(a.xin.q[QRO], g.pp, k - 2, voff);
                        const double b2 = ldp(a.xin.q[QBY], g.pp, k - 2, voff);
                        const double b3 = ldp(a.xin.q[QBY], g.pp, k - 3, voff);
                        wzm = (rw_m2 * (1.0 / ro2) - 4.0 * cm.w) * C.hz * dzz;
                        lbz = (3.0 * cc.by - 4.0 * cm.by + b2) * C.hz * d2zz +
                              (2.0 * cc.by - 5.0 * cm.by + 4.0 * b2 - b3) * C.h2z * dzz * dzz;
                    }
                    const double bE = SH(SH_BY, r, c + 1), bW = SH(SH_BY, r, c - 1);
                    const double bN = SH(SH_BY, r + 1, c), bS = SH(SH_BY, r - 1, c);
                    double w2 = cc.bx * DXS(SH_V) - cc.by * DXS(SH_U);
                    w2 = w2 + cc.bz * DZC(cp.v, cm.v) - cc.by * wzm;
                    w2 = w2 - cc.u * DXS(SH_BY);
                    w2 = w2 + C.orm * LAPX(bE, bW, b2c);
                    w2 = w2 - cc.v * DYS(SH_BY);
                    w2 = w2 + C.orm * LAPY(bN, bS, b2c);
                    w2 = w2 - cc.w * DZC(cp.by, cm.by);
                    w2 = w2 + C.orm * lbz;
                    finish(QBY, w2, cc.by);
                }
                {  // ---- WW3, flux of BZ (3dmhdsub.f:1040, 1090, 1250, 1310, 1500-1580) ----
                    const double b2c = 2.0 * cc.bz;
                    const double bE = SH(SH_BZ, r, c + 1), bW = SH(SH_BZ, r, c - 1);
                    const double bN = SH(SH_BZ, r + 1, c), bS = SH(SH_BZ, r - 1, c);
                    double w3 = cc.bx * DXS(SH_W) - cc.bz * DXS(SH_U);
                    w3 = w3 + cc.by * DYS(SH_W) - cc.bz * DYS(SH_V);
                    w3 = w3 - cc.u * DXS(SH_BZ);
                    w3 = w3 + C.orm * LAPX(bE, bW, b2c);
                    w3 = w3 - cc.v * DYS(SH_BZ);
                    w3 = w3 + C.orm * LAPY(bN, bS, b2c);
                    w3 = w3 - cc.w * DZC(cp.bz, cm.bz);
                    w3 = w3 + C.orm * LAPZ(cp.bz, cm.bz, b2c);
                    finish(QBZ, w3, cc.bz);
                }
            }
#undef LAPX
#undef DXS
#undef DYS
#undef DZC
#undef LAPY
#undef LAPZ
#undef SH
            if (STAGE == 2 && a.want_dt) {
                double q[NPART] = {0.0, 1e300, 1e300, 1e300, 1e300};
                asm volatile("" ::: "memory");  // keeps the LDS values in LDS
                const double dd = dmin(ddxy_s[r][c], zm[2][kl]);
                double bxv = 0.0, byv = 0.0, bzv = 0.0;
                if constexpr (MAG) {
                    bxv = xv[QBX];
                    byv = xv[QBY];
                    bzv = xv[QBZ];
                }
                const double rk = zm[3][kl];
                // the exact limits on output cells that pass the screen (all
                // with a.screen = 0); the others leave the extrema as they are
                bool hit = false;
                if (outp && !bplane &&
                    (!a.screen || dt_screen<MAG>(C, xv[QRU], xv[QRV], xv[QRW], xv[QRO], xv[QTT],
                                                 bxv, byv, bzv, dd, rk, scr_s))) {
                    dt_terms<MAG>(C, xv[QRU], xv[QRV], xv[QRW], xv[QRO], xv[QTT], bxv, byv, bzv,
                                  dd, rk, q);
                    hit = true;
                }
                // the row's extrema, on planes where some lane evaluated
                // (lanes that did not hold the identities)
                if (__ballot(hit)) {
                    const double v0 = wave_max(q[0]);
                    double vm[NPM];
                    for (int v = 1; v < NPM; ++v) vm[v] = wave_min(q[v]);
                    if (ln == 0) {
                        pmw_s[0][fw] = dmax(pmw_s[0][fw], v0);
                        for (int v = 1; v < NPM; ++v)
                            pmw_s[v][fw] = dmin(pmw_s[v][fw], vm[v]);
                    }
                }
            }
        }

        rw_m2 = cm.rw;
        cm = cc;
        cc = cp;
        rn = rnn;
    }
    };
    if (heavy)
        zloop(BoolConst<true>{});
    else
        zloop(BoolConst<false>{});

    if (STAGE == 2 && a.want_dt) {
        if (heavy)
            for (int q = 0; q < NPM; ++q) pm[q] = pmw_s[q][fw];
        block_partials<NW>(pm, red, a.partial, t);
    }
}

// One RK stage of one run: block -> tile with the XCD-aware order.
template <int TY, int STAGE, int ZMODE, bool MAG, bool HEAT = false, bool DEN = false, int TW = 64>
__global__ __launch_bounds__((TileShape<TY, TW>::NW * 64)) void k_stage(const StageArgs a,
                                                                        const Consts C,
                                                                        const Metrics M) {
    __shared__ StageLds<TY, STAGE, MAG, HEAT, DEN, TW> lds;
    const int t = logical_block();
    if (t >= a.nblocks) return;
    stage_tile<TY, STAGE, ZMODE, MAG, HEAT, DEN, TW>(a, C, M, t, lds);
}

}  // namespace R3D_KNS

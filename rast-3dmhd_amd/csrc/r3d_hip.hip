// r3d_hip.hip - host side of the gfx950 time-step library: the C ABI of
// include/r3d_hip.h over the kernels in r3d_stage.hpp / r3d_aux.hpp.
//
// Built with -ffp-contract=off (bit-exact variant).  Every HIP call is
// checked; failures are returned as R3D_E* codes with a thread-local message,
// never as process exits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "r3d_ctx.hpp"

using namespace r3d;

namespace {
thread_local std::string g_err;

// Scratch of r3d_hip_step_noalloc: the dt partials of its steps live in this
// static device array instead of the context's allocation.  Each GPU the
// library runs on holds its own copy until the process ends.  In host code
// the array's name is not a device address: the address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t kNoallocScratchDoubles = size_t(1) << 17;  // 1 MiB
__device__ __align__(256) double scratch[kNoallocScratchDoubles];
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device ordinal -> address of scratch there
}  // namespace

int r3d::set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

// Gather (dir = 1) or scatter (dir = 0) between the global array and one
// virtual rank's local array (nx*ny*nz, x fastest).  Local (i, j, k) is global
// (i, gj0 + j, gk0 + k).  Scatter writes only local j in [jlo, jhi] and k in
// [klo, khi]; gather with interior_only leaves zeros outside the interior.
__global__ void k_rank_copy(double* glob, double* loc, Geo g, int nx, int ny, int nz, int gj0,
                            int gk0, int dir, int jlo, int jhi, int klo, int khi,
                            int interior_only) {
    const long total = (long)nx * ny * nz;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx);
        const long rest = q / nx;
        const int j = (int)(rest % ny);
        const int k = (int)(rest / ny);
        const long n = g.at(i, gj0 + j, gk0 + k);
        if (dir == 0) {
            if (j >= jlo && j <= jhi && k >= klo && k <= khi) glob[n] = loc[q];
        } else {
            const bool in = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2;
            loc[q] = (interior_only && !in) ? 0.0 : glob[n];
        }
    }
}

// Padded global field <-> dense (npx+2)(npy+2)(npz+2) array, x fastest.
__global__ void k_global_copy(double* padded, double* dense, Geo g, int to_dense) {
    const long W = g.npx + 2, H = g.npy + 2;
    const long total = W * H * (g.npz + 2);
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long)gridDim.x * blockDim.x) {
        const int i = (int)(q % W);
        const long rest = q / W;
        const long n = g.at(i, (int)(rest % H), (int)(rest / H));
        if (to_dense)
            dense[q] = padded[n];
        else
            padded[n] = dense[q];
    }
}

// Kernels of one run over the device functions of r3d_aux.hpp.
__global__ __launch_bounds__(BOUND_THREADS) void k_bound(const StageArgs a, const Consts C,
                                                         const Metrics M, int partial_base) {
    bound_block(a, C, M, partial_base, blockIdx.x, gridDim.x, blockIdx.y);
}

__global__ __launch_bounds__(DT_THREADS) void k_dt(const Geo g, const State s, const Consts C,
                                                   const Metrics M, double* partial) {
    dt_block(g, s, C, M, partial, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(1024) void k_finish(DevScal* sc, const double* partial, int npart,
                                                 const Consts C, int reduce, int advance,
                                                 int screened) {
    finish_block(sc, partial, npart, C, reduce, advance, screened);
}

// The exact dt pass after a screened reduction that could not prove its
// extrema (DevScal::dt_redo); every block returns at once otherwise.
__global__ __launch_bounds__(DT_THREADS) void k_dt_redo(const Geo g, const State s,
                                                        const Consts C, const Metrics M,
                                                        double* partial, const DevScal* sc) {
    if (!sc->dt_redo) return;
    dt_block(g, s, C, M, partial, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(1024) void k_finish_redo(DevScal* sc, const double* partial,
                                                      int npart, const Consts C) {
    // all threads read the flag before the reduction's first barrier; thread
    // 0 clears it after the last
    if (!sc->dt_redo) return;
    finish_block(sc, partial, npart, C, 1, 0, 0);
}

// grid = (perimeter blocks, npz): planes 1..npz
__global__ __launch_bounds__(HALO_THREADS) void k_halo(const Geo g, const State x, int nf) {
    halo_thread(g, x, nf, blockIdx.x * HALO_THREADS + threadIdx.x, 1 + blockIdx.y);
}

// Density diffusion, before each stage (golden/fortran/3dmhdsub.f:1311-1328).
// The outer RO planes take their mirror points, RO(:,:,1) = RO(:,:,3) and
// RO(:,:,NZ) = RO(:,:,NZ-2), over the whole plane: the x/y ghost cells of
// the source plane are its periodic images (the reference's halo exchange
// keeps them current; here they are read through the wrap).
constexpr int DEN_THREADS = 256;
__global__ __launch_bounds__(DEN_THREADS) void k_den_mirror(const Geo g, double* ro) {
    const int W = g.npx + 2, H = g.npy + 2;
    const int q = blockIdx.x * DEN_THREADS + threadIdx.x;
    if (q >= W * H) return;
    const int i = q % W, j = q / W;
    const int si = i == 0 ? g.npx : (i == g.npx + 1 ? 1 : i);
    const int sj = j == 0 ? g.npy : (j == g.npy + 1 ? 1 : j);
    ro[g.at(i, j, 0)] = ro[g.at(si, sj, 2)];
    ro[g.at(i, j, g.npz + 1)] = ro[g.at(si, sj, g.npz - 1)];
}

// Horizontal mean, first half (src/communicate.cpp:193-199): thread
// (K, q) sums y rank q's interior rows of global plane K, j outer and i
// inner, in one sequential chain, and divides by npx*npy.  The loads of a
// row run ahead of the additions, which keep their order.  There are only
// (npz+2)*npey chains, so blocks are one wave each and spread over every CU:
// the speed is then the number of loads each lane keeps in flight.
constexpr int DEN_PARTIAL_THREADS = 64;
__device__ __forceinline__ void hmean_partial_chain(const Geo& g, const double* ro, double* part,
                                                    int npey, double cnt) {
    const int t = blockIdx.x * DEN_PARTIAL_THREADS + threadIdx.x;
    if (t >= (g.npz + 2) * npey) return;
    const int K = t / npey, q = t % npey, nry = g.npy / npey;
    // 16-byte loads (rows start on 128-byte boundaries), 32 in flight
    constexpr int U2 = 32, U = 16;
    double sum = 0.0;
    for (int j = q * nry + 1; j <= (q + 1) * nry; ++j) {
        const double* row = ro + g.at(1, j, K);
        int i = 0;
        for (; i + 2 * U2 <= g.npx; i += 2 * U2) {
            double2 v[U2];
#pragma unroll
            for (int u = 0; u < U2; ++u) v[u] = *(const double2*)(row + i + 2 * u);
#pragma unroll
            for (int u = 0; u < U2; ++u) {
                sum += v[u].x;
                sum += v[u].y;
            }
        }
        for (; i + U <= g.npx; i += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = row[i + u];
#pragma unroll
            for (int u = 0; u < U; ++u) sum += v[u];
        }
        for (; i < g.npx; ++i) sum += row[i];
    }
    part[t] = sum / cnt;
}
__global__ __launch_bounds__(DEN_PARTIAL_THREADS) void k_den_partial(const Geo g,
                                                                     const double* ro,
                                                                     double* part, int npey,
                                                                     double cnt) {
    hmean_partial_chain(g, ro, part, npey, cnt);
}

// The convective flux WW2 = -RW*(2.5*(TT-TTM(K)) + 0.5*(UU*UU+VV*VV+WW*WW)),
// UU = RU*(1/RO) (golden/fortran/3dmhdsub.f:1656-1666), of every interior
// cell of planes 0..npz+1, into a field laid out like the state, so that
// k_den_partial sums it with the reference's order.  One thread per cell:
// the per-cell work (a division among it) runs at full occupancy instead of
// inside the sequential sums.  grid = (row blocks, npy, npz + 2).
__global__ __launch_bounds__(DEN_THREADS) void k_ww2_field(const Geo g, const State x,
                                                           const double* ttm, double* w) {
    const int i = 1 + blockIdx.x * DEN_THREADS + threadIdx.x, j = 1 + blockIdx.y, K = blockIdx.z;
    if (i > g.npx) return;
    const long n = g.at(i, j, K);
    const double rw = x.q[QRW][n], ri = 1.0 / x.q[QRO][n];
    const double uu = x.q[QRU][n] * ri, vv = x.q[QRV][n] * ri, ww = rw * ri;
    w[n] = -rw * (2.5 * (x.q[QTT][n] - ttm[K]) + 0.5 * (uu * uu + vv * vv + ww * ww));
}

// The y ranks' partial means of one plane added in rank order.
__global__ __launch_bounds__(DEN_THREADS) void k_hmean_final(const double* part, double* out,
                                                             int nk, int npey) {
    const int K = blockIdx.x * DEN_THREADS + threadIdx.x;
    if (K >= nk) return;
    double acc = part[(long)K * npey];
    for (int q = 1; q < npey; ++q) acc += part[(long)K * npey + q];
    out[K] = acc;
}

// LREM radiative heating of the mean profile per plane K = 1..npz
// (golden/fortran/3dmhdsub.f:1574-1580): HRAD = (d2 TTM/dz2 terms) RKAPA
// + dTTM/dz DKAPA, in the Fortran's operation order.
__global__ __launch_bounds__(DEN_THREADS) void k_lrem_hrad(const double* ttm, double* hrad,
                                                           const Metrics M, int npz, double hz,
                                                           double h2z) {
    const int K = 1 + blockIdx.x * DEN_THREADS + threadIdx.x;
    if (K > npz) return;
    const double dz = M.dzz[K];
    hrad[K] = ((ttm[K + 1] - 2.0 * ttm[K] + ttm[K - 1]) * h2z * dz * dz +
               (ttm[K + 1] - ttm[K - 1]) * hz * M.d2zz[K]) * M.rkapa[K] +
              (ttm[K + 1] - ttm[K - 1]) * hz * dz * M.dkapa[K];
}

// Flux relaxation, one thread (golden/fortran/3dmhdsub.f:1625-1771, the
// ITC=0 / IZC=1 branch): the TFAC update from the bottom-wall TTM, then per z
// rank (nrz planes, the rank's lower ghost plane being the next rank's first)
// FRADM, ALPHA and the trapezoid integral CORRECT, offset by the sum of the
// ranks above's end values, and ftc = RLAX*TFAC*CORRECT for the interior.
// One block.  ALPHA depends on the plane only (a rank's lower ghost plane
// takes the next rank's FRADM, which is that plane's own centred value), so
// the block computes it for planes 1..npz+1 first; thread 0 then runs the
// chain, reading ALPHA and ZEE eight planes ahead of the in-order sums.
constexpr int RLX_THREADS = 256;
__global__ __launch_bounds__(RLX_THREADS) void k_rlax_chain(
    const DevScal* sc, int stage, double* st, const double* ttm, const double* fconm, double* alpha,
    double* ftc, const Metrics M, int npz, int npez, double hz, double repr, double theta,
    double rlax, double tstart, double toff) {
    const double ftot = theta / repr;
    for (int G = 1 + threadIdx.x; G <= npz + 1; G += RLX_THREADS) {
        double fr;
        if (G == npz + 1) {
            fr = 0.0;  // the last rank's, never assigned in the Fortran; only in its unused end value
        } else if (G == 1) {
            fr = (-3.0 * ttm[1] + 4.0 * ttm[2] - ttm[3]) * hz * M.dzz[1] * M.rkapa[1] / repr;
        } else if (G == npz) {
            fr = (3.0 * ttm[npz] - 4.0 * ttm[npz - 1] + ttm[npz - 2]) * hz * M.dzz[npz] *
                 M.rkapa[npz] / repr;
        } else {
            fr = (ttm[G + 1] - ttm[G - 1]) * hz * M.dzz[G] * M.rkapa[G] / repr;
        }
        double a = 1.0 - (fconm[G] + fr) / ftot;
        alpha[G] = a / (1.0 + 4.0 * fabs(a));
    }
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double dt = stage == 0 ? sc->dt_next : sc->dt;  // the step's dt
    const double timt = sc->timc + sc->timi;
    double tfac = st[0];
    const bool on = timt > tstart && tfac > 1e-4;
    if (on) {
        const double tbot = ttm[npz];
        if (tbot <= st[1]) {
            tfac = tfac - dt / (3.0 * toff);
        } else {
            st[1] = tbot;
            tfac = tfac + dt / (3.0 * toff);
            tfac = fmin(tfac, 2.0);
        }
        st[0] = tfac;
    }
    const double rt = rlax * tfac;
    if (!on) {
        for (int G = 1; G <= npz; ++G) ftc[G] = rt * 0.0;
        return;
    }
    const int nrz = npz / npez;
    double sum = 0.0;  // end values of the ranks above, added in rank order
    for (int pz = 0; pz < npez; ++pz) {
        const int G0 = pz * nrz;
        double corr = 0.0;
        // CORRECT(k) = CORRECT(k-1) + 0.5*(ALPHA(k)+ALPHA(k-1))*(ZEE(k)-ZEE(k-1)),
        // k = 2..nrz+1 of the rank (global G0+k)
        int k = 2;
        constexpr int U = 8;
        for (; k + U - 1 <= nrz + 1; k += U) {
            double ap[U + 1], zp[U + 1];
#pragma unroll
            for (int u = 0; u <= U; ++u) {
                ap[u] = alpha[G0 + k - 1 + u];
                zp[u] = M.zee[G0 + k - 1 + u];
            }
            if (k == 2) ftc[G0 + 1] = rt * (corr + sum);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                corr = corr + 0.5 * (ap[u + 1] + ap[u]) * (zp[u + 1] - zp[u]);
                if (k + u <= nrz) ftc[G0 + k + u] = rt * (corr + sum);
            }
        }
        if (k == 2) ftc[G0 + 1] = rt * (corr + sum);
        for (; k <= nrz + 1; ++k) {
            const int G = G0 + k;
            corr = corr + 0.5 * (alpha[G] + alpha[G - 1]) * (M.zee[G] - M.zee[G - 1]);
            if (k <= nrz) ftc[G] = rt * (corr + sum);
        }
        sum = sum + corr;  // this rank's end value, CORRECT(NZ)
    }
}


// Second half (src/communicate.cpp:205-211): the y ranks' partial means
// added in rank order; then the per-plane coefficients TMP = DH*ORE*WWZ(K)
// and DV*ORE*WWZ(K), WWZ = 1/ROM (ID=1) or the wall profile (ID=2).
__global__ __launch_bounds__(DEN_THREADS) void k_den_final(const double* part, double* rom,
                                                           double* dth, double* dtv,
                                                           const double* wwz2, int nk, int npey,
                                                           int id, double dhore, double dvore) {
    const int K = blockIdx.x * DEN_THREADS + threadIdx.x;
    if (K >= nk) return;
    double acc = part[(long)K * npey];
    for (int q = 1; q < npey; ++q) acc += part[(long)K * npey + q];
    rom[K] = acc;
    const double wz = id == 1 ? 1.0 / acc : wwz2[K];
    dth[K] = dhore * wz;
    dtv[K] = dvore * wz;
}

}  // namespace

int r3d::dev_alloc(r3d_hip_ctx* c, double** p, size_t n) {
    hipError_t e = hipMalloc((void**)p, n * sizeof(double));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(R3D_ENOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(double),
                       hipGetErrorString(e));
    }
    c->bytes += n * sizeof(double);
    return R3D_OK;
}

namespace {
// Address of the static scratch on device dev, taken once per device.
int noalloc_scratch(int dev, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return R3D_OK;
    }
    HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_err(R3D_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                       hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
    *out = static_cast<double*>(p);
    return R3D_OK;
}
}  // namespace

namespace {

// ITCON=2 heat-loss amplitude at time timt: TP*OCV, ramped by
// 0.5*(1+TANH(LOG(3.0E00)/QFH*(TIMT-TC))) when TC != 0, with the Fortran's
// single-precision LOG(3.0E00) (golden/fortran/3dmhdsub.f:1801-1820)
double heat_amplitude(const r3d_hip_cfg& f, double timt) {
    if (f.itcon != 2) return 0.0;
    if (f.tc == 0.0) return f.tp * f.ocv;
    return f.tp * f.ocv * 0.5 * (1.0 + tanh((double)logf(3.0f) / f.qfh * (timt - f.tc)));
}

void fill_consts(r3d_hip_ctx* c) {
    const r3d_hip_cfg& f = c->cfg;
    Consts& C = c->C;
    C.dtfix = 0.0;
    C.c13 = f.c13;
    C.c23 = f.c23;
    C.c43 = f.c43;
    C.ore = f.ore;
    C.ocv = f.ocv;
    C.grav = f.grav;
    C.hx = f.hx;
    C.h2x = f.h2x;
    C.hy = f.hy;
    C.h2y = f.h2y;
    C.hz = f.hz;
    C.h2z = f.h2z;
    C.omx = f.omx;
    C.omz = f.omz;
    C.tk = f.ocv / f.repr;
    C.t1 = f.ocv * f.ore;
    C.t2 = f.c13 * f.ore;
    C.t3 = f.c43 * f.ocv * f.ore;
    C.og = 1.0 / f.gamma;
    C.gamma = f.gamma;
    C.re = f.re;
    C.repr = f.repr;
    C.cv = f.cv;
    C.sff = f.sff;
    C.rm = f.rm;
    C.obeta = f.obeta;
    C.orm = f.orm;
    C.theta = f.theta;
    C.tb = f.tb;
    C.hh = f.hh;
    C.tp = f.tp;
    C.tpr = f.tp;
    C.hloss = heat_amplitude(f, 0.0);
    C.den_h = f.id != 0 && f.dh != 0.0;
    C.den_v = f.id != 0 && f.dv != 0.0;
    C.rlx = f.rlax > 0.0;
    C.lrem = f.lrem != 0;
    C.tu = f.tu;
    C.dztu = f.dztu;
    C.dztb = f.dztb;
    const double margin = f.dt_screen != 0.0 ? f.dt_screen : 0.05;
    C.scr_up = 1.0 + margin;
    C.scr_dn = 1.0 - margin;
    C.gam[0] = f.gam1;
    C.gam[1] = f.gam2;
    C.gam[2] = f.gam3;
    C.zeta[0] = f.zeta1;
    C.zeta[1] = f.zeta2;
    C.lrot = f.lrot;
    C.lmag = f.lmag;
    C.izcon = f.izcon;
    C.bz_zero = f.lmag && f.ibcon == 0;
    C.ohm = f.obeta * f.orm * f.ocv;
    C.mag_bot = f.npez > 1;
    // top boundary branch (reference src/bcon.cpp:43-94)
    C.half_t = 0;
    if (f.tp != 0.0 && (f.itcon == 0 || f.itcon == 1)) {
        C.top_mode = f.itcon == 0 ? TOP_GAUSS_T : TOP_GAUSS_FLUX;
    } else if (f.itcon == 0 || f.itcon == 2 || f.itcon == 4) {
        C.top_mode = TOP_CONST_T;
        C.half_t = f.itcon == 4;
    } else if (f.itcon == 1) {
        C.top_mode = TOP_CONST_FLUX;
    } else {
        C.top_mode = TOP_NONE;
    }
}

// Time the launch made by fn() when timing is on.
template <class Fn>
int timed(r3d_hip_ctx* c, int cls, Fn fn) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, c->stream));
    }
    fn();
    HIPCHK(hipGetLastError());
    if (c->timing) {
        HIPCHK(hipEventRecord(e1, c->stream));
        c->pending.push_back({cls, e0});
        c->pending_end.push_back(e1);
    }
    return R3D_OK;
}

int drain_timing(r3d_hip_ctx* c) {
    for (size_t q = 0; q < c->pending.size(); ++q) {
        HIPCHK(hipEventSynchronize(c->pending_end[q]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->pending[q].second, c->pending_end[q]));
        c->launches[c->pending[q].first] += 1;
        c->total_ms[c->pending[q].first] += ms;
        (void)hipEventDestroy(c->pending[q].second);
        (void)hipEventDestroy(c->pending_end[q]);
    }
    c->pending.clear();
    c->pending_end.clear();
    return R3D_OK;
}

StageArgs stage_args(r3d_hip_ctx* c, int stage, int zmode, int want_dt) {
    StageArgs a{};
    a.g = c->g;
    for (int f = 0; f < NXF; ++f) {
        a.xin.q[f] = c->xs[c->cur][f];
        a.xout.q[f] = c->xs[c->cur ^ 1][f];
        a.z.q[f] = c->z[f];
    }
    a.partial = c->partial;
    a.scal = c->scal;
    a.stage = stage;
    a.zmode = zmode;
    a.want_dt = want_dt;
    a.screen = stage == 2 && want_dt;
    a.ntx = c->ntx;
    a.nty = c->nty;
    a.nkc = c->nkc;
    a.kc = c->kc;
    a.nblocks = c->nblocks;
    return a;
}

// The five (stage, z mode) kernels of one instantiation family.
template <int TY, bool MG, bool HT = false, bool DN = false, int TW = 64>
void launch_stage_kernel(r3d_hip_ctx* c, const StageArgs& a, int stage, int zmode) {
    const dim3 blk(TileShape<TY, TW>::NW * 64), grid(c->grid);
#define R3D_STAGE(S, Z) \
    hipLaunchKernelGGL((k_stage<TY, S, Z, MG, HT, DN, TW>), grid, blk, 0, c->stream, a, c->C, c->M)
    if (stage == 0 && zmode == 0) R3D_STAGE(0, 0);
    else if (stage == 0) R3D_STAGE(0, 1);
    else if (stage == 1 && zmode == 0) R3D_STAGE(1, 0);
    else if (stage == 1) R3D_STAGE(1, 2);
    else R3D_STAGE(2, 2);
#undef R3D_STAGE
}

// Density diffusion set-up of this stage's input: mirrored outer RO planes,
// horizontal mean ROM and the per-plane coefficients.
int launch_density_prep(r3d_hip_ctx* c) {
    const r3d_hip_cfg& f = c->cfg;
    double* ro = c->xs[c->cur][QRO];
    const int nk = f.npz + 2, cells = (f.npx + 2) * (f.npy + 2);
    hipLaunchKernelGGL(k_den_mirror, dim3((cells + DEN_THREADS - 1) / DEN_THREADS),
                       dim3(DEN_THREADS), 0, c->stream, c->g, ro);
    hipLaunchKernelGGL(k_den_partial,
                       dim3((nk * f.npey + DEN_PARTIAL_THREADS - 1) / DEN_PARTIAL_THREADS),
                       dim3(DEN_PARTIAL_THREADS), 0, c->stream, c->g, (const double*)ro, c->den_buf,
                       f.npey, (double)(f.npx * f.npy));
    const size_t nkz = (size_t)nk;
    double* rom = c->den_buf + nkz * f.npey;
    hipLaunchKernelGGL(k_den_final, dim3((nk + DEN_THREADS - 1) / DEN_THREADS), dim3(DEN_THREADS),
                       0, c->stream, (const double*)c->den_buf, rom, rom + nkz, rom + 2 * nkz,
                       (const double*)(rom + 3 * nkz), nk, f.npey, f.id, f.dh * f.ore,
                       f.dv * f.ore);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

// Flux relaxation set-up of this stage's input: TTM, FCONM, then TFAC and
// RLAX*TFAC*CORRECT per plane (after the density set-up, whose mirrored RO
// planes the lower ghost plane's convective flux reads, as in the Fortran).
int launch_relax_prep(r3d_hip_ctx* c, int stage) {
    const r3d_hip_cfg& f = c->cfg;
    const int nk = f.npz + 2;
    const size_t nkz = (size_t)nk;
    double* part = c->den_buf;
    double* ttm = c->den_buf + nkz * (f.npey + 4);
    double* fconm = ttm + nkz;
    double* ftc = fconm + nkz;
    double* alpha = ftc + nkz;
    double* st = c->den_buf + nkz * (f.npey + 9);
    State s;
    for (int q = 0; q < NXF; ++q) s.q[q] = c->xs[c->cur][q];
    const double cnt = (double)(f.npx * f.npy);
    const dim3 gp((nk * f.npey + DEN_PARTIAL_THREADS - 1) / DEN_PARTIAL_THREADS);
    const dim3 gf((nk + DEN_THREADS - 1) / DEN_THREADS);
    hipLaunchKernelGGL(k_den_partial, gp, dim3(DEN_PARTIAL_THREADS), 0, c->stream, c->g,
                       (const double*)s.q[QTT], part, f.npey, cnt);
    hipLaunchKernelGGL(k_hmean_final, gf, dim3(DEN_THREADS), 0, c->stream, (const double*)part, ttm,
                       nk, f.npey);
    hipLaunchKernelGGL(k_ww2_field, dim3((f.npx + DEN_THREADS - 1) / DEN_THREADS, f.npy, nk),
                       dim3(DEN_THREADS), 0, c->stream, c->g, s, (const double*)ttm, c->ww2_buf);
    hipLaunchKernelGGL(k_den_partial, gp, dim3(DEN_PARTIAL_THREADS), 0, c->stream, c->g,
                       (const double*)c->ww2_buf, part, f.npey, cnt);
    hipLaunchKernelGGL(k_hmean_final, gf, dim3(DEN_THREADS), 0, c->stream, (const double*)part,
                       fconm, nk, f.npey);
    hipLaunchKernelGGL(k_rlax_chain, dim3(1), dim3(RLX_THREADS), 0, c->stream,
                       (const DevScal*)c->scal, stage, st, (const double*)ttm, (const double*)fconm,
                       alpha, ftc, c->M, f.npz, f.npez, f.hz, f.repr, f.theta, f.rlax, f.tstart,
                       f.toff);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

// LREM set-up of this stage's input: TTM = the horizontal mean of TT in the
// reference's order, then HRAD per plane (golden/fortran/3dmhdsub.f:1540, 1574-1580).
int launch_lrem_prep(r3d_hip_ctx* c) {
    const r3d_hip_cfg& f = c->cfg;
    const int nk = f.npz + 2;
    const size_t nkz = (size_t)nk;
    double* part = c->den_buf;
    double* ttm = c->den_buf + nkz * (f.npey + 4);
    double* hrad = c->den_buf + nkz * (f.npey + 8);
    const dim3 gp((nk * f.npey + DEN_PARTIAL_THREADS - 1) / DEN_PARTIAL_THREADS);
    const dim3 gf((nk + DEN_THREADS - 1) / DEN_THREADS);
    hipLaunchKernelGGL(k_den_partial, gp, dim3(DEN_PARTIAL_THREADS), 0, c->stream, c->g,
                       (const double*)c->xs[c->cur][QTT], part, f.npey,
                       (double)(f.npx * f.npy));
    hipLaunchKernelGGL(k_hmean_final, gf, dim3(DEN_THREADS), 0, c->stream, (const double*)part, ttm,
                       nk, f.npey);
    hipLaunchKernelGGL(k_lrem_hrad, gf, dim3(DEN_THREADS), 0, c->stream, (const double*)ttm, hrad,
                       c->M, f.npz, f.hz, f.h2z);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

int launch_stage(r3d_hip_ctx* c, int stage, int zmode, int want_dt) {
    const StageArgs a = stage_args(c, stage, zmode, want_dt);
    hipError_t fma_err = hipSuccess;
    int rc;
    // RO's mirrored outer planes with density diffusion or LREM
    // (golden/fortran/3dmhdsub.f:1316); then TTM and HRAD with LREM
    if ((c->cfg.id != 0 || c->cfg.lrem) && (rc = launch_density_prep(c))) return rc;
    if (c->cfg.rlax > 0.0 && (rc = launch_relax_prep(c, stage))) return rc;
    if (c->cfg.lrem && (rc = launch_lrem_prep(c))) return rc;
    rc = timed(c, 0, [&] {
        const bool ht = c->cfg.itcon == 2,
                   dn = c->cfg.id != 0 || c->cfg.rlax > 0.0 || c->cfg.lrem;
        if (c->cfg.variant == 1) {
            fma_err = launch_stage_fma(a, c->C, c->M, c->ty, c->cfg.lmag != 0, c->grid,
                                       c->stream);
        } else if (c->cfg.lmag) {
            constexpr int W = TILE_W_MAG;
            if (ht && dn) launch_stage_kernel<TILE_Y_MAG, true, true, true, W>(c, a, stage, zmode);
            else if (ht) launch_stage_kernel<TILE_Y_MAG, true, true, false, W>(c, a, stage, zmode);
            else if (dn) launch_stage_kernel<TILE_Y_MAG, true, false, true, W>(c, a, stage, zmode);
            else if (c->ty == TILE_Y_MAG64) launch_stage_kernel<TILE_Y_MAG64, true>(c, a, stage, zmode);
            else launch_stage_kernel<TILE_Y_MAG, true, false, false, W>(c, a, stage, zmode);
        } else if (ht || dn) {  // these families are built with 8-row tiles
            if (ht && dn) launch_stage_kernel<8, false, true, true>(c, a, stage, zmode);
            else if (ht) launch_stage_kernel<8, false, true, false>(c, a, stage, zmode);
            else launch_stage_kernel<8, false, false, true>(c, a, stage, zmode);
        } else if (c->ty == 8) {
            launch_stage_kernel<8, false>(c, a, stage, zmode);
        } else {
            launch_stage_kernel<TILE_Y, false>(c, a, stage, zmode);
        }
    });
    if (rc) return rc;
    if (fma_err != hipSuccess)
        return set_err(R3D_EHIP, "FMA stage launch: %s", hipGetErrorString(fma_err));
    rc = timed(c, 1, [&] {
        hipLaunchKernelGGL(k_bound, dim3(c->nbb, 2), dim3(BOUND_THREADS), 0, c->stream, a, c->C,
                           c->M, c->nblocks);
    });
    if (rc) return rc;
    c->cur ^= 1;
    return R3D_OK;
}

int launch_dt(r3d_hip_ctx* c) {
    State s;
    for (int f = 0; f < NXF; ++f) s.q[f] = c->xs[c->cur][f];
    int rc = timed(c, 2, [&] {
        hipLaunchKernelGGL(k_dt, dim3(DT_BLOCKS), dim3(DT_THREADS), 0, c->stream, c->g, s, c->C,
                           c->M, c->partial);
    });
    if (rc) return rc;
    return timed(c, 3, [&] {
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, c->stream, c->scal, c->partial,
                           DT_BLOCKS, c->C, 1, 0, 0);
    });
}

}  // namespace

int r3d::global_copy(r3d_hip_ctx* c, double* field, double* host, bool to_host) {
    const r3d_hip_cfg& f = c->cfg;
    const size_t n = (size_t)(f.npx + 2) * (f.npy + 2) * (f.npz + 2);
    int rc;
    if (!c->gstage && (rc = dev_alloc(c, &c->gstage, n))) return rc;
    const dim3 grid((unsigned)std::min<size_t>(4096, (n + 255) / 256)), blk(256);
    if (to_host) {
        hipLaunchKernelGGL(k_global_copy, grid, blk, 0, c->stream, field, c->gstage, c->g, 1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(host, c->gstage, n * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(c->gstage, host, n * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        hipLaunchKernelGGL(k_global_copy, grid, blk, 0, c->stream, field, c->gstage, c->g, 0);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return R3D_OK;
}

int r3d::global_copy_device(r3d_hip_ctx* c, double* field, double* dense, bool to_dense) {
    const r3d_hip_cfg& f = c->cfg;
    const size_t n = (size_t)(f.npx + 2) * (f.npy + 2) * (f.npz + 2);
    const dim3 grid((unsigned)std::min<size_t>(4096, (n + 255) / 256)), blk(256);
    hipLaunchKernelGGL(k_global_copy, grid, blk, 0, c->stream, field, dense, c->g,
                       to_dense ? 1 : 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return R3D_OK;
}

int r3d::fill_ghosts(r3d_hip_ctx* c) {
    State s;
    for (int f = 0; f < NXF; ++f) s.q[f] = c->xs[c->cur][f];
    const int per = 2 * (c->cfg.npx + 2) + 2 * c->cfg.npy;
    hipLaunchKernelGGL(k_halo, dim3((per + HALO_THREADS - 1) / HALO_THREADS, c->cfg.npz),
                       dim3(HALO_THREADS), 0, c->stream, c->g, s, c->nf);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

int r3d::check_ready(r3d_hip_ctx* c) {
    if (!c->metrics_set) return set_err(R3D_ESTATE, "metrics not uploaded");
    for (size_t r = 0; r < c->uploaded.size(); ++r)
        if (!c->uploaded[r]) return set_err(R3D_ESTATE, "virtual rank %zu not uploaded", r);
    return R3D_OK;
}

int r3d::step_one(r3d_hip_ctx* c, int nstages, double* const* keep, double* const* outs) {
    int rc;
    if (c->cfg.tc != 0.0) {
        // time-dependent plume temperature (src/bcon.cpp:45) from the host libm
        DevScal d{};
        HIPCHK(hipMemcpyAsync(&d, c->scal, sizeof d, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const double timt = d.timc + d.timi;
        c->C.tpr = c->cfg.tp + (1.0 - c->cfg.tp) * exp(-timt / c->cfg.tc);
        c->C.hloss = heat_amplitude(c->cfg, timt);
    }
    if (!c->next_valid && (rc = launch_dt(c))) return rc;
    const size_t fe = field_elems(c->g);
    for (int s = 0; s < nstages; ++s) {
        const int last = s == nstages - 1;
        const int zmode = !last ? 0 : (s == 0 ? 1 : 2);
        if (outs && outs[s])
            for (int f = 0; f < c->nf; ++f) c->xs[c->cur ^ 1][f] = outs[s] + f * fe;
        if ((rc = launch_stage(c, s, zmode, nstages == 3 && s == 2))) return rc;
        if (keep && s < 2 && keep[s])
            for (int f = 0; f < c->nf; ++f)
                HIPCHK(hipMemcpyAsync(keep[s] + f * fe, c->xs[c->cur][f], fe * sizeof(double),
                                      hipMemcpyDeviceToDevice, c->stream));
    }
    rc = timed(c, 3, [&] {
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, c->stream, c->scal, c->partial,
                           c->nblocks + 2 * c->nbb, c->C, nstages == 3 ? 1 : 0, 1,
                           nstages == 3 ? 1 : 0);
    });
    if (rc) return rc;
    if (nstages == 3) {
        // the screened reduction's fallback: an exact pass over the new state
        // when its extrema were not proven (both kernels exit at once if so)
        State s;
        for (int f = 0; f < NXF; ++f) s.q[f] = c->xs[c->cur][f];
        rc = timed(c, 2, [&] {
            hipLaunchKernelGGL(k_dt_redo, dim3(DT_BLOCKS), dim3(DT_THREADS), 0, c->stream, c->g,
                               s, c->C, c->M, c->partial, (const DevScal*)c->scal);
        });
        if (rc) return rc;
        rc = timed(c, 3, [&] {
            hipLaunchKernelGGL(k_finish_redo, dim3(1), dim3(1024), 0, c->stream, c->scal,
                               c->partial, DT_BLOCKS, c->C);
        });
        if (rc) return rc;
    }
    c->next_valid = nstages == 3;
    return R3D_OK;
}

extern "C" {

const char* r3d_hip_last_error(void) { return g_err.c_str(); }

int r3d_hip_create(const r3d_hip_cfg* cfg, r3d_hip_ctx** out) {
    if (!cfg || !out) return set_err(R3D_EINVAL, "null argument");
    *out = nullptr;
    const r3d_hip_cfg& f = *cfg;
    if (f.npx < 3 || f.npy < 3 || f.npz < 6)
        return set_err(R3D_EINVAL, "grid %dx%dx%d too small (need npx,npy >= 3, npz >= 6)", f.npx,
                       f.npy, f.npz);
    if (f.npey < 1 || f.npez < 1 || f.npy % f.npey || f.npz % f.npez)
        return set_err(R3D_EINVAL, "npey=%d / npez=%d must divide npy=%d / npz=%d", f.npey, f.npez,
                       f.npy, f.npz);
    if (f.npz / f.npez < 3) return set_err(R3D_EINVAL, "each virtual z rank needs >= 3 planes");
    if (f.izcon != 0 && f.izcon != 1) return set_err(R3D_EINVAL, "BCON: invalid IZCON %d", f.izcon);
    // ITCON=3 is only named in the Fortran (3dmhd.f:645, 1179), never implemented
    if (f.itcon == 3)
        return set_err(R3D_EUNSUP, "FLUXES: unsupported feature enabled: ITCON=3 embedded thermal");
    if (f.itcon == 2 && f.variant != 0)
        return set_err(R3D_EUNSUP, "ITCON=2 heat loss is built in the bit-exact variant only");
    if (f.itcon == 2 && f.tc != 0.0 && f.qfh == 0.0)
        return set_err(R3D_EINVAL, "ITCON=2 with tc != 0 needs qfh != 0 (the heat-loss ramp time)");
    if (f.id != 0 && f.id != 1 && f.id != 2)
        return set_err(R3D_EUNSUP, "density diffusion ID=%d: 1 and 2 are defined", f.id);
    if (f.id != 0 && f.variant != 0)
        return set_err(R3D_EUNSUP, "density diffusion is built in the bit-exact variant only");
    if (f.rlax > 0.0) {
        if (f.variant != 0)
            return set_err(R3D_EUNSUP, "flux relaxation is built in the bit-exact variant only");
        if (!(f.itcon == 0 && f.izcon == 1))
            return set_err(R3D_EUNSUP, "flux relaxation (RLAX) is defined for ITCON=0 with IZCON=1 only");
        if (f.npez < 2 || !(f.toff > 0.0))
            return set_err(R3D_EINVAL, "flux relaxation needs npez >= 2 and toff > 0");
    }
    if (f.variant != 0 && f.variant != 1)
        return set_err(R3D_EUNSUP, "variant %d: 0 (bit-exact) and 1 (FMA) are built", f.variant);
    if (f.lrem && f.variant != 0)
        return set_err(R3D_EUNSUP, "LREM is built in the bit-exact variant only");
    if (f.lrem && f.rlax > 0.0)
        return set_err(R3D_EUNSUP, "flux relaxation with LREM (FTOT = 8.07*0.4*THETA) is not built");
    const bool feature = f.itcon == 2 || f.id != 0 || f.rlax > 0.0 || f.lrem;
    if (f.tile_rows != 0) {
        const bool ok = f.lmag ? f.tile_rows == TILE_Y_MAG ||
                                     (f.tile_rows == TILE_Y_MAG64 && !feature && f.variant == 0)
                               : feature ? f.tile_rows == 8 : (f.tile_rows == 8 || f.tile_rows == TILE_Y);
        if (!ok)
            return set_err(R3D_EINVAL, "tile_rows %d: built are %s", f.tile_rows,
                           f.lmag ? "12 (LMAG; 32 columns) and 5 (LMAG without itcon 2, id, rlax, "
                                    "variant 0; 64 columns)"
                           : feature ? "8 (itcon 2, id, rlax)" : "8 and 9");
    }
    if (!(f.dt_screen > -1.0))
        return set_err(R3D_EINVAL, "dt_screen %g: the margin must be > -1", f.dt_screen);
    if (f.zchunk != 0 && f.zchunk != 8 && f.zchunk != 16 && f.zchunk != 32 && f.zchunk != 64 &&
        f.zchunk != 128)
        return set_err(R3D_EINVAL, "zchunk %d: one of 8, 16, 32, 64, 128", f.zchunk);
    HIPCHK(hipSetDevice(f.device));

    auto* c = new r3d_hip_ctx();
    c->cfg = f;
    c->nf = f.lmag ? 8 : 5;
    c->g.npx = f.npx;
    c->g.npy = f.npy;
    c->g.npz = f.npz;
    c->g.px = ((long)f.npx + 17 + 15) / 16 * 16;
    c->g.pp = c->g.px * (f.npy + 2);
    c->nry = f.npy / f.npey;
    c->nrz = f.npz / f.npez;
    c->nx = f.npx + 2;
    c->ny = c->nry + 2;
    c->nz = c->nrz + 2;
    c->uploaded.assign((size_t)f.npey * f.npez, 0);
    fill_consts(c);

    // stage tiling: 64 x ty columns, z chunks sized for >= ~8 blocks per CU
    c->ty = f.lmag ? TILE_Y_MAG : feature ? 8 : hydro_tile_rows(f.npy);
    if (f.tile_rows) c->ty = f.tile_rows;
    c->tw = tile_width(f.lmag != 0, c->ty);
    c->ntx = (f.npx + c->tw - 1) / c->tw;
    c->nty = (f.npy + c->ty - 1) / c->ty;
    c->kc = 8;
    for (int kc : {128, 64, 32, 16, 8}) {
        if ((long)c->ntx * c->nty * ((f.npz + kc - 1) / kc) >= 2048) {
            c->kc = kc;
            break;
        }
    }
    if (f.zchunk) c->kc = f.zchunk;
    c->nkc = (f.npz + c->kc - 1) / c->kc;
    c->nblocks = c->ntx * c->nty * c->nkc;
    c->grid = (c->nblocks + 7) / 8 * 8;
    c->nbb = ((f.npx + 2) * (f.npy + 2) + BOUND_THREADS - 1) / BOUND_THREADS;

    int rc = R3D_OK;
    auto fail = [&](int code) {
        r3d_hip_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(R3D_EHIP, "hipStreamCreate failed"));
    const size_t fe = field_elems(c->g);
    // The state's fields are allocated after a first set of the same buffers
    // has been allocated and released.  Measured on MI355X: the first context
    // of a process otherwise tends to get device memory on which the stage
    // kernels run about 5% slower (504x504x2048: 15.3 against 14.5 ms per
    // stage launch, mostly in stage 2); a context allocated after such a
    // release runs at the faster rate.  No memory is held twice: the first
    // set is freed before the state is allocated, and an allocation failure
    // here only ends this pass early.
    {
        std::vector<void*> first;
        for (int i = 0; i < 3 * c->nf; ++i) {
            void* p = nullptr;
            if (hipMalloc(&p, fe * sizeof(double)) != hipSuccess) {
                (void)hipGetLastError();
                break;
            }
            first.push_back(p);
        }
        for (void* p : first) (void)hipFree(p);
    }
    for (int f2 = 0; f2 < c->nf; ++f2) {
        if ((rc = dev_alloc(c, &c->xs[0][f2], fe))) return fail(rc);
        if ((rc = dev_alloc(c, &c->xs[1][f2], fe))) return fail(rc);
        if ((rc = dev_alloc(c, &c->z[f2], fe))) return fail(rc);
        if (hipMemsetAsync(c->z[f2], 0, fe * sizeof(double), c->stream) != hipSuccess)
            return fail(set_err(R3D_EHIP, "hipMemset failed"));
    }
    const size_t npart = std::max<size_t>((size_t)c->nblocks + 2 * c->nbb, DT_BLOCKS);
    if ((rc = dev_alloc(c, &c->partial, npart * NPART))) return fail(rc);
    {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(DevScal)) != hipSuccess)
            return fail(set_err(R3D_ENOMEM, "hipMalloc scalars failed"));
        c->scal = (DevScal*)p;
        c->bytes += sizeof(DevScal);
        DevScal z{};
        if (hipMemcpy(c->scal, &z, sizeof z, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(R3D_EHIP, "hipMemcpy scalars failed"));
    }
    const size_t nmet = 3 * (size_t)(f.npx + 2) + 3 * (size_t)(f.npy + 2) + 5 * (size_t)(f.npz + 2) +
                        2 * ((size_t)(f.npx + 2) + (size_t)(f.npy + 2)) + 3 * (size_t)(f.npz + 2);
    if ((rc = dev_alloc(c, &c->metric_buf, nmet))) return fail(rc);
    if (f.id != 0 || f.rlax > 0.0 || f.lrem) {
        // rows of npz+2: npey partial means, ROM, DH*ORE*WWZ, DV*ORE*WWZ, the ID=2
        // profile, TTM, FCONM, RLAX*TFAC*CORRECT, ALPHA, HRAD; then TFAC, TTM_MAX
        const size_t nk = (size_t)f.npz + 2, n = nk * ((size_t)f.npey + 9) + 2;
        if ((rc = dev_alloc(c, &c->den_buf, n))) return fail(rc);
        // zero rows the run does not compute; TFAC = 1, TTM_MAX = -1e99
        // (DATA statements, golden/fortran/3dmhdsub.f:1258-1259)
        std::vector<double> init(n, 0.0);
        init[n - 2] = 1.0;
        init[n - 1] = -1e99;
        if (hipMemcpy(c->den_buf, init.data(), n * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess)
            return fail(set_err(R3D_EHIP, "hipMemcpy of the per-plane buffer failed"));
        c->M.rom = c->den_buf + nk * f.npey;
        c->M.dth = c->M.rom + nk;
        c->M.dtv = c->M.dth + nk;
        c->M.ftc = c->M.rom + 6 * nk;
        c->M.ttm = c->M.rom + 4 * nk;
        c->M.hrad = c->M.rom + 8 * nk;
    }
    if (f.rlax > 0.0 && (rc = dev_alloc(c, &c->ww2_buf, field_elems(c->g)))) return fail(rc);
    const size_t stage_elems = (size_t)std::max<long>((long)c->nx * c->ny * c->nz,
                                                      (long)(f.npx + 2) * (f.npy + 2));
    if ((rc = dev_alloc(c, &c->stage_buf, stage_elems))) return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(set_err(R3D_EHIP, "stream sync failed"));
    *out = c;
    return R3D_OK;
}

int r3d_hip_destroy(r3d_hip_ctx* c) {
    if (!c) return R3D_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) (void)hipEventDestroy(p.second);
    for (auto& e : c->pending_end) (void)hipEventDestroy(e);
    for (int s = 0; s < 2; ++s)
        for (int f = 0; f < NXF; ++f)
            if (c->xs[s][f]) (void)hipFree(c->xs[s][f]);
    for (int f = 0; f < NXF; ++f)
        if (c->z[f]) (void)hipFree(c->z[f]);
    if (c->partial) (void)hipFree(c->partial);
    if (c->scal) (void)hipFree(c->scal);
    if (c->metric_buf) (void)hipFree(c->metric_buf);
    if (c->den_buf) (void)hipFree(c->den_buf);
    if (c->ww2_buf) (void)hipFree(c->ww2_buf);
    if (c->stage_buf) (void)hipFree(c->stage_buf);
    if (c->gstage) (void)hipFree(c->gstage);
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    if (c->snap_ready) (void)hipEventDestroy(c->snap_ready);
    if (c->snap_done) (void)hipEventDestroy(c->snap_done);
    if (c->snap_dev) (void)hipFree(c->snap_dev);
    if (c->snap_host) (void)hipHostFree(c->snap_host);
    if (c->snap_scal) (void)hipHostFree(c->snap_scal);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return R3D_OK;
}

int r3d_hip_set_metrics(r3d_hip_ctx* c, const r3d_hip_metrics* m) {
    if (!c || !m) return set_err(R3D_EINVAL, "null argument");
    const r3d_hip_cfg& f = c->cfg;
    const int X = f.npx + 2, Y = f.npy + 2, Z = f.npz + 2;
    const double* need[] = {m->exx, m->dxxdx, m->d2xxdx2, m->ddx, m->wyy, m->dyydy, m->d2yydy2,
                            m->ddy, m->dzzdz, m->d2zzdz2, m->ddz, m->rkapa, m->dkapa};
    for (const double* p : need)
        if (!p) return set_err(R3D_EINVAL, "metric array missing");
    // plume Gaussians, evaluated with the host libm exactly as the reference
    // does per boundary cell (src/bcon.cpp:52-55 for ITCON=0, :67-68 otherwise)
    const double cln = (double)(-4.0f * logf(2.0f)) / (f.hh * f.hh);
    std::vector<double> e1(X), e2(Y);
    for (int i = 0; i < X; ++i) {
        const double dx = m->exx[i] - f.xp;
        e1[i] = c->C.top_mode == TOP_GAUSS_T ? exp(cln * dx * dx) : exp(cln * (dx * dx));
    }
    for (int j = 0; j < Y; ++j) {
        const double dy = m->wyy[j] - f.yp;
        e2[j] = c->C.top_mode == TOP_GAUSS_T ? exp(cln * dy * dy) : exp(cln * (dy * dy));
    }
    // ITCON=2 heat-loss Gaussians, CLN=-4.0E00*LOG(2.0E00)/HH/HH and
    // EXP(CLN*(EXX(I)-XP)**2) as golden/fortran/3dmhdsub.f:1800-1820 evaluates them
    std::vector<double> hg(X + Y + Z, 0.0);
    if (f.itcon == 2) {
        if (!m->zee) return set_err(R3D_EINVAL, "metric array zee missing (ITCON=2)");
        const double hcln = (double)(-4.0f * logf(2.0f)) / f.hh / f.hh;
        for (int i = 0; i < X; ++i) hg[i] = exp(hcln * ((m->exx[i] - f.xp) * (m->exx[i] - f.xp)));
        for (int j = 0; j < Y; ++j)
            hg[X + j] = exp(hcln * ((m->wyy[j] - f.yp) * (m->wyy[j] - f.yp)));
        for (int k = 0; k < Z; ++k)
            hg[X + Y + k] = exp(hcln * ((m->zee[k] - f.zp) * (m->zee[k] - f.zp)));
    }
    std::vector<double> h;
    h.reserve(3 * X + 3 * Y + 5 * Z + 2 * (X + Y) + 2 * Z);
    if (f.rlax > 0.0 && !m->zee) return set_err(R3D_EINVAL, "metric array zee missing (RLAX)");
    std::vector<double> zee(Z, 0.0);
    if (m->zee) zee.assign(m->zee, m->zee + Z);
    auto put = [&](const double* p, int n) -> const double* {
        const size_t off = h.size();
        h.insert(h.end(), p, p + n);
        return c->metric_buf + off;
    };
    Metrics& M = c->M;
    M.dxx = put(m->dxxdx, X);
    M.d2xx = put(m->d2xxdx2, X);
    M.ddx = put(m->ddx, X);
    M.dyy = put(m->dyydy, Y);
    M.d2yy = put(m->d2yydy2, Y);
    M.ddy = put(m->ddy, Y);
    M.dzz = put(m->dzzdz, Z);
    M.d2zz = put(m->d2zzdz2, Z);
    M.ddz = put(m->ddz, Z);
    M.rkapa = put(m->rkapa, Z);
    M.dkapa = put(m->dkapa, Z);
    M.e1 = put(e1.data(), X);
    M.e2 = put(e2.data(), Y);
    M.hgx = put(hg.data(), X);
    M.hgy = put(hg.data() + X, Y);
    M.hgz = put(hg.data() + X + Y, Z);
    M.zee = put(zee.data(), Z);
    // the thermal time-step limit's conductivity: 1 + RKAPA with LREM
    // (golden/fortran/3dmhdsub.f:845-851), RKAPA otherwise (src/step.cpp:197)
    std::vector<double> rkdt(m->rkapa, m->rkapa + Z);
    if (f.lrem)
        for (int k = 0; k < Z; ++k) rkdt[k] = 1.0 + m->rkapa[k];
    M.rkdt = put(rkdt.data(), Z);
    M.dz_top = f.dzf / m->dzzdz[1];
    M.dz_bot = f.dzf / m->dzzdz[f.npz];
    // ID=2 density diffusion profile WWZ = EXP(CLN*ZEE**2) + EXP(CLN*(ZEE-ZMAX)**2),
    // SGM = 0.1E00 (REAL*8 from a single literal), CLN = -4.0E00*LOG(2.0E00)/SGM/SGM
    // (golden/fortran/3dmhdsub.f:1334-1338)
    std::vector<double> wwz2(Z, 0.0);
    if (f.id == 2) {
        if (!m->zee) return set_err(R3D_EINVAL, "metric array zee missing (ID=2)");
        const double sgm = (double)0.1f;
        const double wcln = (double)(-4.0f * logf(2.0f)) / sgm / sgm;
        for (int k = 0; k < Z; ++k)
            wwz2[k] = exp(wcln * (m->zee[k] * m->zee[k])) +
                      exp(wcln * ((m->zee[k] - f.zmax) * (m->zee[k] - f.zmax)));
    }
    if (c->den_buf)
        HIPCHK(hipMemcpyAsync(c->den_buf + (size_t)Z * (f.npey + 3), wwz2.data(), Z * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->metric_buf, h.data(), h.size() * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->metrics_set = true;
    c->next_valid = false;
    return R3D_OK;
}

int r3d_hip_upload_rank(r3d_hip_ctx* c, int rank, const double* const* fields) {
    if (!c || !fields) return set_err(R3D_EINVAL, "null argument");
    const r3d_hip_cfg& f = c->cfg;
    if (rank < 0 || rank >= f.npey * f.npez) return set_err(R3D_EINVAL, "rank %d out of range", rank);
    const int py = rank % f.npey, pz = rank / f.npey;
    const int jlo = py == 0 ? 0 : 1, jhi = py == f.npey - 1 ? c->ny - 1 : c->ny - 2;
    const int klo = pz == 0 ? 0 : 1, khi = pz == f.npez - 1 ? c->nz - 1 : c->nz - 2;
    const size_t n = (size_t)c->nx * c->ny * c->nz;
    for (int q = 0; q < c->nf; ++q) {
        if (!fields[q]) return set_err(R3D_EINVAL, "field %d missing", q);
        HIPCHK(hipMemcpyAsync(c->stage_buf, fields[q], n * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        for (int s = 0; s < 2; ++s)
            hipLaunchKernelGGL(k_rank_copy, dim3(1024), dim3(256), 0, c->stream, c->xs[s][q],
                               c->stage_buf, c->g, c->nx, c->ny, c->nz, py * c->nry, pz * c->nrz,
                               0, jlo, jhi, klo, khi, 0);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->uploaded[rank] = 1;
    c->next_valid = false;
    return R3D_OK;
}

int r3d_hip_upload_global(r3d_hip_ctx* c, const double* const* fields) {
    if (!c || !fields) return set_err(R3D_EINVAL, "null argument");
    for (int q = 0; q < c->nf; ++q) {
        if (!fields[q]) return set_err(R3D_EINVAL, "field %d missing", q);
        for (int s = 0; s < 2; ++s) {
            int rc = global_copy(c, c->xs[s][q], const_cast<double*>(fields[q]), false);
            if (rc) return rc;
        }
    }
    std::fill(c->uploaded.begin(), c->uploaded.end(), 1);
    c->next_valid = false;
    return R3D_OK;
}

int r3d_hip_download_rank(r3d_hip_ctx* c, int rank, double* const* fields) {
    if (!c || !fields) return set_err(R3D_EINVAL, "null argument");
    const r3d_hip_cfg& f = c->cfg;
    if (rank < 0 || rank >= f.npey * f.npez) return set_err(R3D_EINVAL, "rank %d out of range", rank);
    const int py = rank % f.npey, pz = rank / f.npey;
    const size_t n = (size_t)c->nx * c->ny * c->nz;
    int rc = fill_ghosts(c);
    if (rc) return rc;
    for (int q = 0; q < R3D_NFIELDS; ++q) {
        if (!fields[q]) continue;
        const bool isz = q >= R3D_ZRU;
        const int fi = isz ? q - R3D_ZRU : q;
        if (fi >= c->nf) {
            std::memset(fields[q], 0, n * sizeof(double));
            continue;
        }
        double* src = isz ? c->z[fi] : c->xs[c->cur][fi];
        hipLaunchKernelGGL(k_rank_copy, dim3(1024), dim3(256), 0, c->stream, src, c->stage_buf,
                           c->g, c->nx, c->ny, c->nz, py * c->nry, pz * c->nrz, 1, 0, 0, 0, 0,
                           isz ? 1 : 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(fields[q], c->stage_buf, n * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return R3D_OK;
}

int r3d_hip_download_global(r3d_hip_ctx* c, double* const* fields) {
    if (!c || !fields) return set_err(R3D_EINVAL, "null argument");
    const r3d_hip_cfg& f = c->cfg;
    const size_t W = (size_t)(f.npx + 2), rows = (size_t)(f.npy + 2) * (f.npz + 2);
    int rc = fill_ghosts(c);
    if (rc) return rc;
    for (int q = 0; q < R3D_NFIELDS; ++q) {
        if (!fields[q]) continue;
        const bool isz = q >= R3D_ZRU;
        const int fi = isz ? q - R3D_ZRU : q;
        if (fi >= c->nf) {
            std::memset(fields[q], 0, W * rows * sizeof(double));
            continue;
        }
        double* src = isz ? c->z[fi] : c->xs[c->cur][fi];
        if ((rc = global_copy(c, src, fields[q], true))) return rc;
    }
    return R3D_OK;
}

int r3d_hip_snapshot_begin(r3d_hip_ctx* c) {
    if (!c) return set_err(R3D_EINVAL, "null context");
    int rc = check_ready(c);
    if (rc) return rc;
    const r3d_hip_cfg& f = c->cfg;
    const size_t n = (size_t)(f.npx + 2) * (f.npy + 2) * (f.npz + 2);
    const size_t nsf = 2 * (size_t)c->nf;  // X and Z
    if (!c->snap_done) {  // first use: all of these, or none
        auto undo = [&] {
            if (c->snap_dev) {
                (void)hipFree(c->snap_dev);
                c->bytes -= nsf * n * sizeof(double);
            }
            if (c->snap_host) (void)hipHostFree(c->snap_host);
            if (c->snap_scal) (void)hipHostFree(c->snap_scal);
            if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
            if (c->snap_ready) (void)hipEventDestroy(c->snap_ready);
            c->snap_dev = c->snap_host = nullptr;
            c->snap_scal = nullptr;
            c->copy_stream = nullptr;
            c->snap_ready = nullptr;
        };
        if ((rc = dev_alloc(c, &c->snap_dev, nsf * n))) {
            undo();
            return rc;
        }
        if (hipHostMalloc((void**)&c->snap_host, nsf * n * sizeof(double)) != hipSuccess ||
            hipHostMalloc((void**)&c->snap_scal, sizeof(DevScal)) != hipSuccess) {
            undo();
            return set_err(R3D_ENOMEM, "snapshot: pinned host allocation of %zu bytes failed",
                           nsf * n * sizeof(double));
        }
        if (hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming) != hipSuccess) {
            undo();
            c->snap_done = nullptr;
            return set_err(R3D_EHIP, "snapshot: stream or event creation failed");
        }
    }
    // the previous snapshot's copy must be done before its buffer is reused
    if (c->snap_taken) HIPCHK(hipStreamWaitEvent(c->stream, c->snap_done, 0));
    if ((rc = fill_ghosts(c))) return rc;
    const dim3 grid((unsigned)std::min<size_t>(4096, (n + 255) / 256)), blk(256);
    for (int q = 0; q < c->nf; ++q) {
        hipLaunchKernelGGL(k_global_copy, grid, blk, 0, c->stream, c->xs[c->cur][q],
                           c->snap_dev + q * n, c->g, 1);
        hipLaunchKernelGGL(k_global_copy, grid, blk, 0, c->stream, c->z[q],
                           c->snap_dev + (c->nf + q) * n, c->g, 1);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->snap_scal, c->scal, sizeof(DevScal), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipEventRecord(c->snap_ready, c->stream));
    HIPCHK(hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
    HIPCHK(hipMemcpyAsync(c->snap_host, c->snap_dev, nsf * n * sizeof(double),
                          hipMemcpyDeviceToHost, c->copy_stream));
    HIPCHK(hipEventRecord(c->snap_done, c->copy_stream));
    c->snap_taken = true;
    return R3D_OK;
}

int r3d_hip_snapshot_rank(r3d_hip_ctx* c, int rank, double* const* fields, r3d_hip_scalars* s) {
    if (!c || !fields) return set_err(R3D_EINVAL, "null argument");
    const r3d_hip_cfg& f = c->cfg;
    if (rank < 0 || rank >= f.npey * f.npez) return set_err(R3D_EINVAL, "rank %d out of range", rank);
    if (!c->snap_taken) return set_err(R3D_ESTATE, "no snapshot has been taken");
    HIPCHK(hipEventSynchronize(c->snap_done));
    const size_t W = (size_t)(f.npx + 2), H = (size_t)(f.npy + 2);
    const size_t n = W * H * (f.npz + 2);
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const int gj0 = (rank % f.npey) * c->nry, gk0 = (rank / f.npey) * c->nrz;
    for (int q = 0; q < R3D_NFIELDS; ++q) {
        if (!fields[q]) continue;
        const bool isz = q >= R3D_ZRU;
        const int fi = isz ? q - R3D_ZRU : q;
        if (fi >= c->nf) {
            std::memset(fields[q], 0, (size_t)nx * ny * nz * sizeof(double));
            continue;
        }
        // the rank's local box of the global field, as r3d_hip_download_rank
        // returns it: primaries with every ghost cell, accumulators interior only
        const double* g = c->snap_host + (size_t)(isz ? c->nf + fi : fi) * n;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                double* dst = fields[q] + ((size_t)k * ny + j) * nx;
                if (isz && (k == 0 || k == nz - 1 || j == 0 || j == ny - 1)) {
                    std::memset(dst, 0, nx * sizeof(double));
                    continue;
                }
                std::memcpy(dst, g + ((size_t)(gk0 + k) * H + (gj0 + j)) * W, nx * sizeof(double));
                if (isz) dst[0] = dst[nx - 1] = 0.0;
            }
    }
    if (s) {
        s->dt = c->snap_scal->dt;
        s->timc = c->snap_scal->timc;
        s->timt = c->snap_scal->timt;
        s->timi = c->snap_scal->timi;
        s->umach = c->snap_scal->umach;
        s->nit = c->snap_scal->nit;
    }
    return R3D_OK;
}

int r3d_hip_set_scalars(r3d_hip_ctx* c, const r3d_hip_scalars* s) {
    if (!c || !s) return set_err(R3D_EINVAL, "null argument");
    DevScal d{};
    HIPCHK(hipMemcpyAsync(&d, c->scal, sizeof d, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    d.dt = s->dt;
    d.timc = s->timc;
    d.timt = s->timt;
    d.timi = s->timi;
    d.umach = s->umach;
    d.nit = s->nit;
    HIPCHK(hipMemcpyAsync(c->scal, &d, sizeof d, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return R3D_OK;
}

int r3d_hip_get_scalars(r3d_hip_ctx* c, r3d_hip_scalars* s) {
    if (!c || !s) return set_err(R3D_EINVAL, "null argument");
    DevScal d{};
    HIPCHK(hipMemcpyAsync(&d, c->scal, sizeof d, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    s->dt = d.dt;
    s->timc = d.timc;
    s->timt = d.timt;
    s->timi = d.timi;
    s->umach = d.umach;
    s->nit = d.nit;
    return R3D_OK;
}

int r3d_hip_step(r3d_hip_ctx* c, int nsteps, int nstages) {
    if (!c) return set_err(R3D_EINVAL, "null context");
    if (nstages < 1 || nstages > 3) return set_err(R3D_EINVAL, "nstages must be 1..3");
    if (nsteps < 0 || (nstages < 3 && nsteps > 1))
        return set_err(R3D_EINVAL, "a truncated step (nstages < 3) advances exactly one step");
    int rc = check_ready(c);
    if (rc) return rc;
    for (int st = 0; st < nsteps; ++st)
        if ((rc = step_one(c, nstages, nullptr))) return rc;
    if (c->timing) return drain_timing(c);
    return R3D_OK;
}

int r3d_hip_step_noalloc(r3d_hip_ctx* c, int nsteps, int nstages) {
    if (!c) return set_err(R3D_EINVAL, "null context");
    if (nstages < 1 || nstages > 3) return set_err(R3D_EINVAL, "nstages must be 1..3");
    if (nsteps < 0 || (nstages < 3 && nsteps > 1))
        return set_err(R3D_EINVAL, "a truncated step (nstages < 3) advances exactly one step");
    if (c->timing)
        return set_err(R3D_ESTATE, "step_noalloc: kernel timing is on (it creates events)");
    int rc = check_ready(c);
    if (rc) return rc;
    // every partial slot a step reads is written earlier in the same step
    // (stage-3 tiles, both boundary sides, or the dt pass), so the array's
    // contents from an earlier call never reach a result
    const size_t need = std::max<size_t>((size_t)c->nblocks + 2 * c->nbb, DT_BLOCKS) * NPART;
    if (need > kNoallocScratchDoubles)
        return set_err(R3D_EINVAL,
                       "step_noalloc: this grid needs %zu doubles of dt partials; the static "
                       "scratch holds %zu",
                       need, kNoallocScratchDoubles);
    double* sp = nullptr;
    if ((rc = noalloc_scratch(c->cfg.device, &sp))) return rc;
    // the kernels take the partials' address at launch: point the context at
    // the static scratch for this call's launches only
    double* const own = c->partial;
    c->partial = sp;
    for (int st = 0; st < nsteps && rc == R3D_OK; ++st) rc = step_one(c, nstages, nullptr);
    c->partial = own;
    return rc;
}

int r3d_hip_sync(r3d_hip_ctx* c) {
    if (!c) return set_err(R3D_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return c->timing ? drain_timing(c) : R3D_OK;
}

int r3d_hip_set_timing(r3d_hip_ctx* c, int enable) {
    if (!c) return set_err(R3D_EINVAL, "null context");
    int rc = drain_timing(c);
    if (rc) return rc;
    c->timing = enable != 0;
    for (int q = 0; q < 4; ++q) {
        c->launches[q] = 0;
        c->total_ms[q] = 0.0;
    }
    return R3D_OK;
}

int r3d_hip_kernel_stats(r3d_hip_ctx* c, int cls, long* launches, double* total_ms) {
    if (!c || cls < 0 || cls > 3) return set_err(R3D_EINVAL, "bad argument");
    int rc = drain_timing(c);
    if (rc) return rc;
    if (launches) *launches = c->launches[cls];
    if (total_ms) *total_ms = c->total_ms[cls];
    return R3D_OK;
}

int r3d_hip_memory(r3d_hip_ctx* c, unsigned long long* bytes) {
    if (!c || !bytes) return set_err(R3D_EINVAL, "null argument");
    *bytes = c->bytes;
    return R3D_OK;
}

int r3d_hip_tiling(r3d_hip_ctx* c, int out[5]) {
    if (!c || !out) return set_err(R3D_EINVAL, "null argument");
    out[0] = c->ty;
    out[1] = c->kc;
    out[2] = c->ntx;
    out[3] = c->nty;
    out[4] = c->nkc;
    return R3D_OK;
}

}  // extern "C"

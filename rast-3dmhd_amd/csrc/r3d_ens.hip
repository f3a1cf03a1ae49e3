// r3d_ens.hip - batched ensemble of independent runs (include/r3d_ens.h).
//
// Every member is an ordinary run context (r3d_hip_ctx: its own buffers,
// metrics, constants and scalars).  The ensemble adds a device table with one
// descriptor per member and launches each RK stage once for all members: the
// global tile index of a block is mapped to (member, tile) through the
// members' first-tile offsets, so a stage of 64 small runs fills the GPU the
// way one large run does.  The boundary pass, the dt reductions and the
// end-of-step bookkeeping are likewise one launch each, indexed by member.
// A step is captured in a HIP graph per X parity and replayed.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "r3d_ctx.hpp"
#include "r3d_ens.h"

using namespace r3d;

namespace {

// Device-side description of one member.
struct EnsMember {
    Geo g;
    State x[2];   // X ping-pong
    State z;
    double* partial;
    DevScal* scal;
    int ntx, nty, nkc, kc, nblocks, nbb, tile0, npart_dt;
    int bnd0, nbbp;  // persistent launch: first boundary unit, boundary blocks per side
    Consts C;
    Metrics M;
};

constexpr int ENS_DT_BLOCKS = 64;  // blocks per member of the standalone dt pass

__device__ __forceinline__ StageArgs member_args(const EnsMember& e, int cur, int stage,
                                                 int zmode, int want_dt) {
    StageArgs a;
    a.g = e.g;
    a.xin = e.x[cur];
    a.xout = e.x[cur ^ 1];
    a.z = e.z;
    a.partial = e.partial;
    a.scal = e.scal;
    a.stage = stage;
    a.zmode = zmode;
    a.want_dt = want_dt;
    a.screen = stage == 2 && want_dt;  // the time-step screen, as in a single run
    a.ntx = e.ntx;
    a.nty = e.nty;
    a.nkc = e.nkc;
    a.kc = e.kc;
    a.nblocks = e.nblocks;
    return a;
}

// One stage of a full step for every member.  Tiles of one member are
// contiguous in the global order, and the XCD-aware block order keeps each
// member's tiles together on one XCD where it can.
template <int TY, int STAGE, bool MAG, int TW>
__global__ __launch_bounds__((TileShape<TY, TW>::NW * 64)) void k_stage_ens(
    const EnsMember* __restrict__ mem, int nmem, int ntiles, int cur) {
    __shared__ StageLds<TY, STAGE, MAG, false, false, TW> lds;
    const int T = logical_block();
    if (T >= ntiles) return;
    int lo = 0, hi = nmem - 1;  // last member whose first tile is <= T
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mem[mid].tile0 <= T) lo = mid;
        else hi = mid - 1;
    }
    const EnsMember& e = mem[lo];
    constexpr int ZMODE = STAGE < 2 ? 0 : 2;
    const StageArgs a = member_args(e, cur, STAGE, ZMODE, STAGE == 2);
    stage_tile<TY, STAGE, ZMODE, MAG, false, false, TW>(a, e.C, e.M, T - e.tile0, lds);
}

// Last member m with key(mem[m]) <= v (keys ascending from mem[0] = 0).
template <class Key>
__device__ __forceinline__ int find_member(const EnsMember* mem, int nmem, int v, Key key) {
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (key(mem[mid]) <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The persistent form: nsteps full steps of every member in one cooperative
// launch of NT-thread blocks, all resident (at most one per CU at this LDS
// size).  Each phase is a loop of its blocks over work units, followed by a
// grid-wide barrier (cooperative groups: every block writes its L2 back and
// invalidates it, so a phase sees the previous phase's stores on every XCD):
//   per stage: the stage tiles, then the boundary pass;
//   per step: each member's dt reduction with the screen's proof, then, only
//   when some member's proof failed, the exact pass and its reduction for
//   those members.
// Stage tiles are taken dynamically: the global tile order is cut into eight
// contiguous ranges, one per XCD (blocks b with b % 8 = x, as the per-stage
// launches place them), and a block takes the next tile of its range from a
// counter (work[x]).  Counters are never reset during the launch: each block
// takes exactly one tile past the end of its range per stage, so the next
// stage's range starts per_xcd further on.  work[8] is set by a member whose
// proof failed and read by every block after the barrier that follows, so
// all blocks agree on whether the exact pass runs; block 0 clears it one
// barrier later.  Every block runs the same sequence of barriers and the work
// loops contain none.  Results are those of the per-stage launches bit for
// bit (the same tile and cell code; the min / max reductions do not depend on
// how the partials are grouped).  work[0..8] must be zero at launch.
template <int TY, bool MAG, int TW>
__global__ __launch_bounds__((TileShape<TY, TW>::NW * 64)) void k_ens_run(
    const EnsMember* __restrict__ mem, int nmem, int ntiles, int nbunits, int cur, int nsteps,
    int* work) {
    constexpr int NT = TileShape<TY, TW>::NW * 64;
    __shared__ union {
        StageLds<TY, 0, MAG, false, false, TW> s0;
        StageLds<TY, 1, MAG, false, false, TW> s1;
        StageLds<TY, 2, MAG, false, false, TW> s2;
    } lds;
    __shared__ int grab;
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int nblk = gridDim.x;  // a multiple of 8
    const int xcd = blockIdx.x & 7, per_xcd = nblk >> 3;
    const int tiles_xcd = (ntiles + 7) >> 3;
    const int n_x = max(0, min(tiles_xcd, ntiles - xcd * tiles_xcd));  // this XCD's tiles
    int base = 0;  // counter value of this XCD's first tile in the current stage
    for (int st = 0; st < nsteps; ++st) {
        for (int s = 0; s < 3; ++s) {
            if (st + s > 0 && blockIdx.x == 0 && threadIdx.x == 0 && s == 0)
                work[8] = 0;  // every block read it before the last barrier
            for (;;) {
                if (threadIdx.x == 0) grab = atomicAdd(&work[xcd], 1) - base;
                __syncthreads();
                const int i = grab;
                __syncthreads();  // grab is rewritten by the next take
                if (i >= n_x) break;
                const int T = xcd * tiles_xcd + i;
                const EnsMember& e =
                    mem[find_member(mem, nmem, T, [](const EnsMember& x) { return x.tile0; })];
                const StageArgs a = member_args(e, cur, s, s < 2 ? 0 : 2, s == 2);
                if (s == 0)
                    stage_tile<TY, 0, 0, MAG, false, false, TW>(a, e.C, e.M, T - e.tile0, lds.s0);
                else if (s == 1)
                    stage_tile<TY, 1, 0, MAG, false, false, TW>(a, e.C, e.M, T - e.tile0, lds.s1);
                else
                    stage_tile<TY, 2, 2, MAG, false, false, TW>(a, e.C, e.M, T - e.tile0, lds.s2);
                __syncthreads();  // the next tile reuses the LDS
            }
            base += n_x + per_xcd;
            grid.sync();
            for (int u = blockIdx.x; u < nbunits; u += nblk) {
                const EnsMember& e =
                    mem[find_member(mem, nmem, u, [](const EnsMember& x) { return x.bnd0; })];
                const int l = u - e.bnd0;
                const StageArgs a = member_args(e, cur, s, s < 2 ? 0 : 2, s == 2);
                bound_block<NT>(a, e.C, e.M, e.nblocks, l % e.nbbp, e.nbbp, l / e.nbbp);
                __syncthreads();
            }
            grid.sync();
            cur ^= 1;
        }
        for (int m = blockIdx.x; m < nmem; m += nblk) {
            const EnsMember& e = mem[m];
            finish_block<NT>(e.scal, e.partial, e.nblocks + 2 * e.nbbp, e.C, 1, 1, 1);
            if (threadIdx.x == 0 && e.scal->dt_redo) work[8] = 1;
            __syncthreads();
        }
        grid.sync();
        if (work[8]) {  // the same value in every block
            for (int u = blockIdx.x; u < nmem * ENS_DT_BLOCKS; u += nblk) {
                const EnsMember& e = mem[u / ENS_DT_BLOCKS];
                if (e.scal->dt_redo) {
                    dt_block<NT>(e.g, e.x[cur], e.C, e.M, e.partial, u % ENS_DT_BLOCKS,
                                 ENS_DT_BLOCKS);
                    __syncthreads();
                }
            }
            grid.sync();
            // all threads read the flag before the reduction's barrier;
            // thread 0 clears it after
            for (int m = blockIdx.x; m < nmem; m += nblk) {
                const EnsMember& e = mem[m];
                if (e.scal->dt_redo) {
                    finish_block<NT>(e.scal, e.partial, e.npart_dt, e.C, 1, 0, 0);
                    __syncthreads();
                }
            }
            grid.sync();
        }
    }
}

// grid = (max boundary blocks, 2 sides, members)
__global__ __launch_bounds__(BOUND_THREADS) void k_bound_ens(const EnsMember* __restrict__ mem,
                                                             int cur, int stage) {
    const EnsMember& e = mem[blockIdx.z];
    if ((int)blockIdx.x >= e.nbb) return;
    const StageArgs a = member_args(e, cur, stage, stage < 2 ? 0 : 2, stage == 2);
    bound_block(a, e.C, e.M, e.nblocks, blockIdx.x, e.nbb, blockIdx.y);
}

// grid = (ENS_DT_BLOCKS, members)
__global__ __launch_bounds__(DT_THREADS) void k_dt_ens(const EnsMember* __restrict__ mem,
                                                       int cur) {
    const EnsMember& e = mem[blockIdx.y];
    dt_block(e.g, e.x[cur], e.C, e.M, e.partial, blockIdx.x, gridDim.x);
}

// grid = (members); fused: the screened partials of the last stage, else
// those of k_dt_ens
__global__ __launch_bounds__(1024) void k_finish_ens(const EnsMember* __restrict__ mem,
                                                     int fused, int advance) {
    const EnsMember& e = mem[blockIdx.x];
    finish_block(e.scal, e.partial, fused ? e.nblocks + 2 * e.nbb : e.npart_dt, e.C, 1, advance,
                 fused);
}

// The exact dt pass of the members whose screened extrema were not proven
// (DevScal::dt_redo); the blocks of every other member return at once.
// grid = (ENS_DT_BLOCKS, members)
__global__ __launch_bounds__(DT_THREADS) void k_dt_redo_ens(const EnsMember* __restrict__ mem,
                                                            int cur) {
    const EnsMember& e = mem[blockIdx.y];
    if (!e.scal->dt_redo) return;
    dt_block(e.g, e.x[cur], e.C, e.M, e.partial, blockIdx.x, gridDim.x);
}
// grid = (members): all threads read the flag before the reduction's first
// barrier; thread 0 clears it after the last
__global__ __launch_bounds__(1024) void k_finish_redo_ens(const EnsMember* __restrict__ mem) {
    const EnsMember& e = mem[blockIdx.x];
    if (!e.scal->dt_redo) return;
    finish_block(e.scal, e.partial, e.npart_dt, e.C, 1, 0, 0);
}

// Stage tiles of the ensemble: ty rows, kc planes (0: 32), tw columns (64,
// or 32 with two rows per wave for members whose width leaves much of a
// 64-column tile idle).
struct Tiling {
    int ty, kc, tw;
};

}  // namespace

struct r3d_ens_ctx {
    std::vector<r3d_hip_ctx*> mem;
    std::vector<double*> part;     // per-member partial buffers sized for any tiling
    std::vector<EnsMember> desc;
    EnsMember* d_desc = nullptr;
    int* d_work = nullptr;  // persistent launch: tile counters per XCD, exact-pass flag
    hipStream_t stream = nullptr;
    int device = 0, lmag = 0;
    Tiling tile{8, 0, 64};
    int ntiles = 0, grid = 0, nbb_max = 0;
    int cur = 0;
    bool use_graph = true;
    bool persistent = false;  // one cooperative launch per r3d_ens_step call
    int nbunits = 0;          // persistent launch: boundary units of all members
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    unsigned long long bytes = 0;
};

namespace {

int member_kc(const r3d_ens_ctx* e, const r3d_hip_ctx* c) {
    if (e->tile.kc > 0) return std::min(e->tile.kc, c->cfg.npz);
    return std::min(32, c->cfg.npz);
}

void drop_graphs(r3d_ens_ctx* e) {
    for (auto& gr : e->graph)
        if (gr) {
            (void)hipGraphExecDestroy(gr);
            gr = nullptr;
        }
}

// Threads of a persistent-launch block at the current tiling.
int wave_count(int ty, int tw) { return tw == 64 ? ty + 3 : ty / (64 / tw) + 2; }
int persistent_threads(const r3d_ens_ctx* e) { return wave_count(e->tile.ty, e->tile.tw) * 64; }

// Rebuild the member table for the current tiling and parity.
int build_table(r3d_ens_ctx* e) {
    int tile0 = 0, bnd0 = 0;
    const int npt = persistent_threads(e);
    e->nbb_max = 0;
    for (size_t m = 0; m < e->mem.size(); ++m) {
        r3d_hip_ctx* c = e->mem[m];
        EnsMember& d = e->desc[m];
        d.g = c->g;
        for (int s = 0; s < 2; ++s)
            for (int f = 0; f < NXF; ++f) d.x[s].q[f] = c->xs[s][f];
        for (int f = 0; f < NXF; ++f) d.z.q[f] = c->z[f];
        d.partial = e->part[m];
        d.scal = c->scal;
        d.ntx = (c->cfg.npx + e->tile.tw - 1) / e->tile.tw;
        d.nty = (c->cfg.npy + e->tile.ty - 1) / e->tile.ty;
        d.kc = member_kc(e, c);
        d.nkc = (c->cfg.npz + d.kc - 1) / d.kc;
        d.nblocks = d.ntx * d.nty * d.nkc;
        d.nbb = c->nbb;
        d.tile0 = tile0;
        d.npart_dt = ENS_DT_BLOCKS;
        d.nbbp = ((c->cfg.npx + 2) * (c->cfg.npy + 2) + npt - 1) / npt;
        d.bnd0 = bnd0;
        bnd0 += 2 * d.nbbp;
        d.C = c->C;
        d.M = c->M;
        tile0 += d.nblocks;
        e->nbb_max = std::max(e->nbb_max, d.nbb);
    }
    e->ntiles = tile0;
    e->nbunits = bnd0;
    e->grid = (tile0 + 7) / 8 * 8;
    HIPCHK(hipMemcpyAsync(e->d_desc, e->desc.data(), e->desc.size() * sizeof(EnsMember),
                          hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    drop_graphs(e);
    return R3D_OK;
}

template <int TY, bool MAG, int TW = 64>
void launch_stages_ty(r3d_ens_ctx* e, int stage, int cur) {
    const dim3 blk(TileShape<TY, TW>::NW * 64);
    const int n = (int)e->mem.size();
    if (stage == 0)
        hipLaunchKernelGGL((k_stage_ens<TY, 0, MAG, TW>), dim3(e->grid), blk, 0, e->stream,
                           e->d_desc, n, e->ntiles, cur);
    else if (stage == 1)
        hipLaunchKernelGGL((k_stage_ens<TY, 1, MAG, TW>), dim3(e->grid), blk, 0, e->stream,
                           e->d_desc, n, e->ntiles, cur);
    else
        hipLaunchKernelGGL((k_stage_ens<TY, 2, MAG, TW>), dim3(e->grid), blk, 0, e->stream,
                           e->d_desc, n, e->ntiles, cur);
}

// The launches of one full step starting from parity cur.
int enqueue_step(r3d_ens_ctx* e, int cur) {
    const int n = (int)e->mem.size();
    for (int s = 0; s < 3; ++s) {
        if (e->lmag) launch_stages_ty<TILE_Y_MAG, true, TILE_W_MAG>(e, s, cur);
        else if (e->tile.tw == 32 && e->tile.ty == 8) launch_stages_ty<8, false, 32>(e, s, cur);
        else if (e->tile.tw == 32) launch_stages_ty<16, false, 32>(e, s, cur);
        else if (e->tile.tw == 16 && e->tile.ty == 16) launch_stages_ty<16, false, 16>(e, s, cur);
        else if (e->tile.tw == 16) launch_stages_ty<28, false, 16>(e, s, cur);
        else if (e->tile.ty == 4) launch_stages_ty<4, false>(e, s, cur);
        else if (e->tile.ty == 8) launch_stages_ty<8, false>(e, s, cur);
        else launch_stages_ty<9, false>(e, s, cur);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_bound_ens, dim3(e->nbb_max, 2, n), dim3(BOUND_THREADS), 0, e->stream,
                           e->d_desc, cur, s);
        HIPCHK(hipGetLastError());
        cur ^= 1;
    }
    hipLaunchKernelGGL(k_finish_ens, dim3(n), dim3(1024), 0, e->stream, e->d_desc, 1, 1);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_dt_redo_ens, dim3(ENS_DT_BLOCKS, n), dim3(DT_THREADS), 0, e->stream,
                       e->d_desc, cur);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_finish_redo_ens, dim3(n), dim3(1024), 0, e->stream, e->d_desc);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

// nsteps full steps in one cooperative launch (k_ens_run): as many blocks as
// can be resident at once (a multiple of 8, at most one per tile), each
// looping over its share of every phase.
template <int TY, bool MAG, int TW = 64>
int launch_persistent_ty(r3d_ens_ctx* e, int nsteps) {
    constexpr int NT = TileShape<TY, TW>::NW * 64;
    const void* fn = reinterpret_cast<const void*>(&k_ens_run<TY, MAG, TW>);
    int coop = 0, ncu = 0, per_cu = 0;
    HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, e->device));
    if (!coop) return set_err(R3D_EUNSUP, "device %d has no cooperative launch", e->device);
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, e->device));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, 0));
    int nblk = std::min(per_cu * ncu, (e->ntiles + 7) / 8 * 8) / 8 * 8;
    if (nblk < 8)
        return set_err(R3D_EUNSUP, "persistent ensemble: %d resident blocks of %d threads", nblk,
                       NT);
    const EnsMember* d = e->d_desc;
    int nmem = (int)e->mem.size(), ntiles = e->ntiles, nbu = e->nbunits, cur = e->cur, ns = nsteps;
    int* work = e->d_work;
    HIPCHK(hipMemsetAsync(work, 0, 16 * sizeof(int), e->stream));
    void* args[] = {(void*)&d, &nmem, &ntiles, &nbu, &cur, &ns, (void*)&work};
    HIPCHK(hipLaunchCooperativeKernel(fn, dim3(nblk), dim3(NT), args, 0, e->stream));
    return R3D_OK;
}

int launch_persistent(r3d_ens_ctx* e, int nsteps) {
    if (e->lmag) return launch_persistent_ty<TILE_Y_MAG, true, TILE_W_MAG>(e, nsteps);
    if (e->tile.tw == 32 && e->tile.ty == 8) return launch_persistent_ty<8, false, 32>(e, nsteps);
    if (e->tile.tw == 32) return launch_persistent_ty<16, false, 32>(e, nsteps);
    if (e->tile.tw == 16 && e->tile.ty == 16) return launch_persistent_ty<16, false, 16>(e, nsteps);
    if (e->tile.tw == 16) return launch_persistent_ty<28, false, 16>(e, nsteps);
    if (e->tile.ty == 4) return launch_persistent_ty<4, false>(e, nsteps);
    if (e->tile.ty == 8) return launch_persistent_ty<8, false>(e, nsteps);
    return launch_persistent_ty<9, false>(e, nsteps);
}

int enqueue_dt(r3d_ens_ctx* e) {
    const int n = (int)e->mem.size();
    hipLaunchKernelGGL(k_dt_ens, dim3(ENS_DT_BLOCKS, n), dim3(DT_THREADS), 0, e->stream,
                       e->d_desc, e->cur);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_finish_ens, dim3(n), dim3(1024), 0, e->stream, e->d_desc, 0, 0);
    HIPCHK(hipGetLastError());
    return R3D_OK;
}

int capture(r3d_ens_ctx* e, int cur) {
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step(e, cur);
    const hipError_t ec = hipStreamEndCapture(e->stream, &gr);
    if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
    }
    if (ec != hipSuccess) return set_err(R3D_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&e->graph[cur], gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ei != hipSuccess) return set_err(R3D_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    return R3D_OK;
}

int check_member(r3d_ens_ctx* e, int m) {
    if (!e) return set_err(R3D_EINVAL, "null ensemble");
    if (m < 0 || m >= (int)e->mem.size()) return set_err(R3D_EINVAL, "member %d out of range", m);
    return R3D_OK;
}

// Members' own bookkeeping follows the ensemble's parity.
void sync_members(r3d_ens_ctx* e, bool next_valid) {
    for (r3d_hip_ctx* c : e->mem) {
        c->cur = e->cur;
        c->next_valid = next_valid;
    }
}

bool all_next_valid(const r3d_ens_ctx* e) {
    for (const r3d_hip_ctx* c : e->mem)
        if (!c->next_valid) return false;
    return true;
}

}  // namespace

extern "C" {

int r3d_ens_create(int nmembers, const r3d_hip_cfg* cfgs, r3d_ens_ctx** out) {
    if (!cfgs || !out || nmembers < 1) return set_err(R3D_EINVAL, "bad arguments");
    *out = nullptr;
    for (int m = 0; m < nmembers; ++m) {
        if (cfgs[m].device != cfgs[0].device || cfgs[m].lmag != cfgs[0].lmag)
            return set_err(R3D_EINVAL, "member %d: all members need the same device and lmag", m);
        if (cfgs[m].tc != 0.0)
            return set_err(R3D_EUNSUP, "member %d: tc != 0 is not supported in an ensemble", m);
        if (cfgs[m].id != 0)
            return set_err(R3D_EUNSUP, "member %d: density diffusion is not built in an ensemble", m);
        if (cfgs[m].rlax > 0.0)
            return set_err(R3D_EUNSUP, "member %d: flux relaxation is not built in an ensemble", m);
        if (cfgs[m].itcon == 2)
            return set_err(R3D_EUNSUP, "member %d: ITCON=2 heat loss is not built in an ensemble", m);
        if (cfgs[m].lrem)
            return set_err(R3D_EUNSUP, "member %d: LREM is not built in an ensemble", m);
    }
    auto* e = new r3d_ens_ctx();
    e->device = cfgs[0].device;
    e->lmag = cfgs[0].lmag;
    e->tile = {e->lmag ? TILE_Y_MAG : TILE_Y, 0, e->lmag ? TILE_W_MAG : 64};
    auto fail = [&](int code) {
        r3d_ens_destroy(e);
        return code;
    };
    for (int m = 0; m < nmembers; ++m) {
        r3d_hip_ctx* c = nullptr;
        const int rc = r3d_hip_create(&cfgs[m], &c);
        if (rc) {
            const std::string msg = std::string("member ") + std::to_string(m) + ": " +
                                    r3d_hip_last_error();
            r3d_ens_destroy(e);
            return set_err(rc, "%s", msg.c_str());
        }
        e->mem.push_back(c);
        // partial slots for the finest tiling (4 rows x 64 columns, 8 x 32 or
        // 16 x 16, 8 planes) or the dt pass
        const long cols = std::max<long>(
            std::max<long>((long)((cfgs[m].npx + 63) / 64) * ((cfgs[m].npy + 3) / 4),
                           (long)((cfgs[m].npx + 31) / 32) * ((cfgs[m].npy + 7) / 8)),
            (long)((cfgs[m].npx + 15) / 16) * ((cfgs[m].npy + 15) / 16));
        const long worst = cols * ((cfgs[m].npz + 7) / 8) + 2L * c->nbb;
        double* p = nullptr;
        const size_t np = (size_t)std::max<long>(worst, ENS_DT_BLOCKS) * NPART;
        if (hipMalloc((void**)&p, np * sizeof(double)) != hipSuccess)
            return fail(set_err(R3D_ENOMEM, "member %d: partial buffer", m));
        e->part.push_back(p);
        e->bytes += np * sizeof(double) + c->bytes;
    }
    e->desc.resize(nmembers);
    if (hipMalloc((void**)&e->d_desc, nmembers * sizeof(EnsMember)) != hipSuccess)
        return fail(set_err(R3D_ENOMEM, "member table"));
    if (hipMalloc((void**)&e->d_work, 16 * sizeof(int)) != hipSuccess)
        return fail(set_err(R3D_ENOMEM, "persistent work counters"));
    if (hipSetDevice(e->device) != hipSuccess ||
        hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(R3D_EHIP, "ensemble stream"));
    *out = e;
    return R3D_OK;
}

int r3d_ens_destroy(r3d_ens_ctx* e) {
    if (!e) return R3D_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    for (double* p : e->part) (void)hipFree(p);
    for (r3d_hip_ctx* c : e->mem) r3d_hip_destroy(c);
    if (e->d_desc) (void)hipFree(e->d_desc);
    if (e->d_work) (void)hipFree(e->d_work);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return R3D_OK;
}

int r3d_ens_set_metrics(r3d_ens_ctx* e, int m, const r3d_hip_metrics* mt) {
    int rc = check_member(e, m);
    if (rc) return rc;
    if ((rc = r3d_hip_set_metrics(e->mem[m], mt))) return rc;
    e->ntiles = 0;  // constants and metric pointers go into the member table
    return R3D_OK;
}

int r3d_ens_upload_global(r3d_ens_ctx* e, int m, const double* const* fields) {
    int rc = check_member(e, m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    r3d_hip_ctx* c = e->mem[m];
    // the upload writes both X buffers; the member follows the ensemble parity
    if ((rc = r3d_hip_upload_global(c, fields))) return rc;
    c->cur = e->cur;
    return R3D_OK;
}

int r3d_ens_download_global(r3d_ens_ctx* e, int m, double* const* fields) {
    int rc = check_member(e, m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    return r3d_hip_download_global(e->mem[m], fields);
}

int r3d_ens_download_rank(r3d_ens_ctx* e, int m, int rank, double* const* fields) {
    int rc = check_member(e, m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    return r3d_hip_download_rank(e->mem[m], rank, fields);
}

int r3d_ens_set_scalars(r3d_ens_ctx* e, int m, const r3d_hip_scalars* s) {
    int rc = check_member(e, m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    return r3d_hip_set_scalars(e->mem[m], s);
}

int r3d_ens_get_scalars(r3d_ens_ctx* e, int m, r3d_hip_scalars* s) {
    int rc = check_member(e, m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    return r3d_hip_get_scalars(e->mem[m], s);
}

int r3d_ens_set_tile_shape(r3d_ens_ctx* e, int ty, int kc, int tw) {
    if (!e) return set_err(R3D_EINVAL, "null ensemble");
    bool ok;
    if (e->lmag) ok = tw == TILE_W_MAG && ty == TILE_Y_MAG;
    else if (tw == 64) ok = ty == 4 || ty == 8 || ty == 9;
    else if (tw == 32) ok = ty == 8 || ty == 16;
    else ok = tw == 16 && (ty == 16 || ty == 28);
    if (!ok)
        return set_err(R3D_EINVAL, "tile of %d rows x %d columns not built for this ensemble",
                       ty, tw);
    if (kc != 0 && (kc < 8 || kc > 128)) return set_err(R3D_EINVAL, "kc %d outside 8..128", kc);
    e->tile = {ty, kc, tw};
    e->ntiles = 0;
    return R3D_OK;
}

int r3d_ens_set_tiles(r3d_ens_ctx* e, int ty, int kc) {
    return r3d_ens_set_tile_shape(e, ty, kc, e && e->lmag ? TILE_W_MAG : 64);
}

int r3d_ens_get_tile_shape(r3d_ens_ctx* e, int* ty, int* kc, int* tw, int* ntiles) {
    if (!e) return set_err(R3D_EINVAL, "null ensemble");
    if (!e->ntiles) {
        const int rc = build_table(e);
        if (rc) return rc;
    }
    if (ty) *ty = e->tile.ty;
    if (kc) *kc = e->tile.kc;
    if (tw) *tw = e->tile.tw;
    if (ntiles) *ntiles = e->ntiles;
    return R3D_OK;
}

int r3d_ens_get_tiles(r3d_ens_ctx* e, int* ty, int* kc, int* ntiles) {
    return r3d_ens_get_tile_shape(e, ty, kc, nullptr, ntiles);
}

int r3d_ens_set_graph(r3d_ens_ctx* e, int enable) {
    if (!e) return set_err(R3D_EINVAL, "null ensemble");
    HIPCHK(hipStreamSynchronize(e->stream));
    e->use_graph = enable != 0;
    return R3D_OK;
}

int r3d_ens_set_persistent(r3d_ens_ctx* e, int enable) {
    if (!e) return set_err(R3D_EINVAL, "null ensemble");
    HIPCHK(hipStreamSynchronize(e->stream));
    e->persistent = enable != 0;
    return R3D_OK;
}

int r3d_ens_step(r3d_ens_ctx* e, int nsteps) {
    if (!e) return set_err(R3D_EINVAL, "null ensemble");
    if (nsteps < 0) return set_err(R3D_EINVAL, "nsteps < 0");
    int rc;
    for (r3d_hip_ctx* c : e->mem)
        if ((rc = check_ready(c))) return rc;
    if (!e->ntiles && (rc = build_table(e))) return rc;
    if (nsteps == 0) return R3D_OK;
    // members may have used their own streams (uploads, scalars)
    for (r3d_hip_ctx* c : e->mem) HIPCHK(hipStreamSynchronize(c->stream));
    if (!all_next_valid(e) && (rc = enqueue_dt(e))) return rc;
    if (e->persistent) {
        if ((rc = launch_persistent(e, nsteps))) return rc;
        e->cur ^= nsteps & 1;  // three stages per step flip the X parity
        sync_members(e, true);
        return R3D_OK;
    }
    for (int st = 0; st < nsteps; ++st) {
        if (e->use_graph) {
            if (!e->graph[e->cur] && (rc = capture(e, e->cur))) return rc;
            HIPCHK(hipGraphLaunch(e->graph[e->cur], e->stream));
        } else if ((rc = enqueue_step(e, e->cur))) {
            return rc;
        }
        e->cur ^= 1;  // three stages flip the X parity
    }
    sync_members(e, true);
    return R3D_OK;
}

int r3d_ens_sync(r3d_ens_ctx* e) {
    if (!e) return set_err(R3D_EINVAL, "null ensemble");
    HIPCHK(hipStreamSynchronize(e->stream));
    return R3D_OK;
}

int r3d_ens_autotune(r3d_ens_ctx* e, int trial_steps, double* best_ms) {
    if (!e || trial_steps < 1) return set_err(R3D_EINVAL, "bad arguments");
    int rc;
    for (r3d_hip_ctx* c : e->mem)
        if ((rc = check_ready(c))) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    for (r3d_hip_ctx* c : e->mem) HIPCHK(hipStreamSynchronize(c->stream));
    // save every member's current X and scalars
    std::vector<double*> save;
    std::vector<DevScal> scal(e->mem.size());
    auto release = [&] {
        for (double* p : save) (void)hipFree(p);
    };
    const int cur0 = e->cur;
    std::vector<char> valid0;
    for (size_t m = 0; m < e->mem.size(); ++m) {
        r3d_hip_ctx* c = e->mem[m];
        valid0.push_back(c->next_valid);
        const size_t fe = field_elems(c->g);
        double* p = nullptr;
        if (hipMalloc((void**)&p, fe * c->nf * sizeof(double)) != hipSuccess) {
            release();
            return set_err(R3D_ENOMEM, "autotune: state backup");
        }
        save.push_back(p);
        for (int f = 0; f < c->nf; ++f)
            HIPCHK(hipMemcpyAsync(p + f * fe, c->xs[cur0][f], fe * sizeof(double),
                                  hipMemcpyDeviceToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(&scal[m], c->scal, sizeof(DevScal), hipMemcpyDeviceToHost,
                              e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    auto restore = [&]() -> int {
        e->cur = cur0;
        for (size_t m = 0; m < e->mem.size(); ++m) {
            r3d_hip_ctx* c = e->mem[m];
            const size_t fe = field_elems(c->g);
            for (int f = 0; f < c->nf; ++f)
                HIPCHK(hipMemcpyAsync(c->xs[cur0][f], save[m] + f * fe, fe * sizeof(double),
                                      hipMemcpyDeviceToDevice, e->stream));
            HIPCHK(hipMemcpyAsync(c->scal, &scal[m], sizeof(DevScal), hipMemcpyHostToDevice,
                                  e->stream));
            c->cur = cur0;
            c->next_valid = valid0[m];
        }
        HIPCHK(hipStreamSynchronize(e->stream));
        return R3D_OK;
    };
    std::vector<Tiling> cand;
    if (e->lmag) {
        for (int kc : {8, 16, 32, 64}) cand.push_back({TILE_Y_MAG, kc, TILE_W_MAG});
    } else {
        for (int ty : {4, 8, 9})
            for (int kc : {8, 16, 32, 64}) cand.push_back({ty, kc, 64});
        for (int ty : {8, 16})
            for (int kc : {16, 32, 64}) cand.push_back({ty, kc, 32});
        for (int ty : {16, 28})
            for (int kc : {16, 32, 64}) cand.push_back({ty, kc, 16});
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    Tiling best = e->tile;
    double best_t = 1e300;
    rc = R3D_OK;
    for (const Tiling& t : cand) {
        e->tile = t;
        if ((rc = build_table(e)) || (rc = r3d_ens_step(e, 1))) break;
        if (hipEventRecord(t0, e->stream) != hipSuccess) { rc = set_err(R3D_EHIP, "event"); break; }
        if ((rc = r3d_ens_step(e, trial_steps))) break;
        if (hipEventRecord(t1, e->stream) != hipSuccess ||
            hipEventSynchronize(t1) != hipSuccess) { rc = set_err(R3D_EHIP, "event"); break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        if (ms < best_t) {
            best_t = ms;
            best = t;
        }
        if ((rc = restore())) break;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    const int rc2 = restore();
    release();
    if (rc) return rc;
    if (rc2) return rc2;
    e->tile = best;
    if ((rc = build_table(e))) return rc;
    if (best_ms) *best_ms = best_t / trial_steps;
    return R3D_OK;
}

int r3d_ens_memory(r3d_ens_ctx* e, unsigned long long* bytes) {
    if (!e || !bytes) return set_err(R3D_EINVAL, "null argument");
    *bytes = e->bytes;
    return R3D_OK;
}

}  // extern "C"

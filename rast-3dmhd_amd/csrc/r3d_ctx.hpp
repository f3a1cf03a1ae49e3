// r3d_ctx.hpp - host-side internals shared by the single-run library
// (r3d_hip.hip) and the ensemble (r3d_ens.hip): the device context of one run
// and the helpers that build and drive it.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "r3d_aux.hpp"
#include "r3d_hip.h"

namespace r3d {

constexpr int TILE_Y = 9;  // the larger hydro stage tile (waves 1..9 compute fluxes)
// Hydro tile rows for a grid npy cells high: 8 (11 waves; the 8 flux waves
// sit two per SIMD, which measures faster than 9 rows with three flux waves on
// one SIMD) unless 9 leaves fewer rows of the last tile idle.
inline int hydro_tile_rows(int npy) {
    const int w9 = (npy + 8) / 9 * 9 - npy, w8 = (npy + 7) / 8 * 8 - npy;
    return w9 < w8 ? TILE_Y : 8;
}
// MHD tiles: 8 waves per workgroup (two per SIMD), so that the larger MHD
// working set gets 256 VGPRs per lane instead of 168: 12 rows of 32 columns
// (six flux waves of two rows each), or 5 rows of 64 columns (five flux
// waves, the first tiling; 4% slower at 252x252x512, kept as an override)
constexpr int TILE_Y_MAG = 12, TILE_W_MAG = 32;
constexpr int TILE_Y_MAG64 = 5;
// Columns of a stage tile of ty rows.
inline int tile_width(bool lmag, int ty) { return lmag && ty == TILE_Y_MAG ? TILE_W_MAG : 64; }
constexpr int DT_BLOCKS = 1024;

// Set the calling thread's error message and return code.
int set_err(int code, const char* fmt, ...);

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::r3d::set_err(R3D_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

}  // namespace r3d

// The device state of one run on one GPU.
struct r3d_hip_ctx {
    r3d_hip_cfg cfg;
    r3d::Geo g;
    int nf = 5;
    int nx = 0, ny = 0, nz = 0, nry = 0, nrz = 0;
    hipStream_t stream = nullptr;
    double* xs[2][r3d::NXF] = {};  // X ping-pong; xs[cur] is the current state
    double* z[r3d::NXF] = {};
    int cur = 0;
    double* partial = nullptr;
    r3d::DevScal* scal = nullptr;
    double* metric_buf = nullptr;
    // per-plane extras (density diffusion, flux relaxation): per-(plane, y
    // rank) partial means, then ROM, the two diffusion coefficient rows, the
    // ID=2 wall profile, TTM, FCONM and RLAX*TFAC*CORRECT, (npz+2) each, and
    // the relaxation state TFAC, TTM_MAX
    double* den_buf = nullptr;
    double* ww2_buf = nullptr;  // flux relaxation: the convective flux WW2, one field
    double* stage_buf = nullptr;
    double* gstage = nullptr;  // one global field, (npx+2)(npy+2)(npz+2), allocated on first use
    // asynchronous snapshot (r3d_hip_snapshot_begin): X and Z as dense global
    // fields on the device and in pinned host memory, the scalars, a copy
    // stream and its events; allocated on first use
    double* snap_dev = nullptr;
    double* snap_host = nullptr;
    r3d::DevScal* snap_scal = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    bool snap_taken = false;
    r3d::Metrics M{};
    r3d::Consts C{};
    bool metrics_set = false;
    bool next_valid = false;  // scal->dt_next holds the dt of the current state
    std::vector<char> uploaded;
    int ty = r3d::TILE_Y, tw = 64, ntx = 0, nty = 0, nkc = 0, kc = 0, nblocks = 0, grid = 0, nbb = 0;
    unsigned long long bytes = 0;
    bool timing = false;
    long launches[4] = {0, 0, 0, 0};
    double total_ms[4] = {0, 0, 0, 0};
    std::vector<std::pair<int, hipEvent_t>> pending;  // (class, start event), end follows
    std::vector<hipEvent_t> pending_end;
};

namespace r3d {

// Elements of one padded global field (plus slack for the row padding).
inline size_t field_elems(const Geo& g) { return (size_t)g.pp * (g.npz + 2) + 16; }

// hipMalloc with the error mapped to R3D_ENOMEM; counts bytes on c.
int dev_alloc(r3d_hip_ctx* c, double** p, size_t n);
// Copy one padded global field (device, at its allocation base) to or from a
// host array of (npx+2)(npy+2)(npz+2) doubles, x fastest: packed on the device
// and moved in one transfer.  Synchronises c->stream.
int global_copy(r3d_hip_ctx* c, double* field, double* host, bool to_host);
// The same between a padded field and a dense global array in device memory
// (no staging); returns after the copy has completed.
int global_copy_device(r3d_hip_ctx* c, double* field, double* dense, bool to_dense);
// Write the periodic ghost images of c's current X (downloads read them).
int fill_ghosts(r3d_hip_ctx* c);
// All virtual ranks and the metrics uploaded.
int check_ready(r3d_hip_ctx* c);
// One step of nstages stages (the body of r3d_hip_step).  keep[s] (s = 0, 1),
// if given, receives the state after stage s+1 and its boundary pass (nf
// fields at a stride of field_elems).
// One step.  keep[s] (s < 2, may be null): copy stage s's output there.
// outs[s] (may be null): stage s writes its output into that state buffer
// (nf fields of field_elems each) instead of the context's own; the caller
// restores c->xs afterwards.
int step_one(r3d_hip_ctx* c, int nstages, double* const* keep, double* const* outs = nullptr);
// The FMA-contracted stage kernel (r3d_fma.hip); ty must be the run's tile rows.
hipError_t launch_stage_fma(const StageArgs& a, const Consts& C, const Metrics& M, int ty,
                            bool mag, unsigned grid, hipStream_t stream);

}  // namespace r3d

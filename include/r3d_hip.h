/* r3d_hip.h - C ABI of the MI355X (gfx950) Rast-3DMHD time-step library.
 *
 * One context holds the whole global domain of one run on one GPU.  The
 * reference's npey x npez MPI decomposition is kept as a set of *virtual*
 * ranks: upload and download move the reference's per-rank arrays (local
 * extents nx = npx+2, ny = npy/npey+2, nz = npz/npez+2, x fastest, ghosts
 * included), while the device keeps one padded global array per field.
 *
 * Every call returns 0 on success or a negative R3D_E* code; the message of
 * the last failure on the calling thread is in r3d_hip_last_error().  No call
 * exits the process.  A context is not thread-safe: callers serialise calls
 * on it.  All work is queued on the context's own HIP stream.
 *
 * Reference interfaces replaced (reference include/gpu.hpp:35-49):
 *   gpu_alloc(GpuSim&, const Derived&)        -> r3d_hip_create
 *   gpu_upload(GpuSim&, const SimState&)       -> r3d_hip_set_metrics + r3d_hip_upload_rank
 *   gpu_download(SimState&, GpuSim&)           -> r3d_hip_download_rank + r3d_hip_get_scalars
 *   gpu_free(GpuSim&)                          -> r3d_hip_destroy
 *   step_gpu(MPI_Comm, ..., SimState& scalars) -> r3d_hip_step (nstages = 3)
 *   step_gpu(int nstages, MPI_Comm, ...)       -> r3d_hip_step (nstages = 1..3)
 * The C++ mirror of those signatures is include/r3d_gpu.hpp.
 */
#ifndef R3D_HIP_H
#define R3D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    R3D_OK = 0,
    R3D_EINVAL = -1,   /* invalid argument or configuration */
    R3D_EHIP = -2,     /* HIP runtime error (message has the HIP error string) */
    R3D_ENOMEM = -3,   /* device allocation failed */
    R3D_ESTATE = -4,   /* call not valid in the context's current state */
    R3D_EUNSUP = -5    /* feature the reference rejects (src/fluxes.cpp:46-51) */
};

/* Field ids used by upload/download (gstate order of the primaries and the
 * RK accumulators, reference include/fields.hpp:77-104). */
enum {
    R3D_RU = 0, R3D_RV, R3D_RW, R3D_RO, R3D_TT, R3D_BX, R3D_BY, R3D_BZ,
    R3D_ZRU, R3D_ZRV, R3D_ZRW, R3D_ZRO, R3D_ZTT, R3D_ZBX, R3D_ZBY, R3D_ZBZ,
    R3D_NFIELDS
};

/* Run configuration: the Params/Derived scalars the step reads
 * (reference include/params.hpp:24-145, src/params.cpp:14-66). */
typedef struct r3d_hip_cfg {
    int npx, npy, npz;          /* global interior points */
    int npey, npez;             /* virtual rank grid (npy%npey == 0, npz%npez == 0) */
    int lmag, lrot;             /* physics switches */
    int itcon, izcon, ibcon;    /* boundary selectors (reference src/bcon.cpp); itcon 2 also
                                   adds the Fortran's embedded heat loss to the energy flux
                                   (bit-exact variant only), itcon 3 gives R3D_EUNSUP */
    int variant;                /* 0 = bit-exact (no FMA contraction), 1 = FMA-contracted */
    int device;                 /* HIP device ordinal */
    double c13, c23, c43;
    double gam1, gam2, gam3, zeta1, zeta2;
    double hx, h2x, hy, h2y, hz, h2z;
    double re, repr, ore, cv, ocv, gamma, grav, theta, sff;
    double rm, orm, obeta, omx, omz;
    double tp, tc, hh, xp, yp;  /* plume boundary condition */
    double dzf;                 /* f4(1/(npz-1)), the BCON flux spacing numerator */
    double tb;                  /* bottom temperature from the static profile */
    double zp, qfh;             /* ITCON=2 embedded heat loss: centre z (with xp, yp, hh, tp)
                                   and ramp time scale (golden/fortran/3dmhdsub.f:1797-1822) */
    double dh, dv, zmax;        /* density diffusion coefficients and domain depth */
    int id;                     /* density diffusion: 0 off, 1 WWZ = 1/ROM, 2 wall Gaussians
                                   (golden/fortran/3dmhdsub.f:1306-1380; needs NGRID=0) */
    double rlax, tstart, toff;  /* flux relaxation (golden/fortran/3dmhdsub.f:1625-1771):
                                   rlax > 0 turns it on; ITCON=0 with IZCON=1, NGRID=0,
                                   npez >= 2 and toff > 0 only */
    /* Stage-kernel tiling (results do not depend on it; 0 = chosen by the
     * library).  tile_rows: y rows per workgroup tile, 8 or 9 for hydro (8
     * with itcon 2, id or rlax; 64 columns), 12 with lmag (32 columns, two
     * rows per wave) or 5 with lmag (64 columns; not with itcon 2, id, rlax or
     * variant 1).  zchunk: z planes per
     * workgroup, 8, 16, 32, 64 or 128.  No reference counterpart: the
     * reference GPU launches one thread per cell (src/gpu.cpp). */
    int tile_rows, zchunk;
    /* LREM (M. Rempel's stratification, golden/fortran/3dmhdsub.f:437-490,
     * 817-851, 1316, 1534-1596, 2852-2952; the reference C++ refuses it,
     * src/static_ic.cpp:27): lrem != 0 diffuses only the temperature
     * perturbation about the horizontal mean, adds the mean profile's
     * radiative heating, takes 1 + rkapa in the thermal time-step limit and
     * the boundary values tu (constant top temperature), dztu and dztb (flux
     * conditions) from the static set-up.  Bit-exact variant, NGRID=0, no RLAX. */
    int lrem;
    double tu, dztu, dztb;
    /* Time-step screen of the last stage (results do not depend on it; no
     * reference counterpart).  A cell's exact time-step limits are evaluated
     * only where a division-free bound says one of them may come within
     * this relative margin of the previous step's extrema; the step then
     * proves that the extrema it found are the exact ones, or redoes them
     * with an exact pass over every cell.  0 = 0.05; a large value (1e9)
     * evaluates every cell; a value in (-1, 0) puts the thresholds inside
     * the previous extrema, so that every step takes the exact pass. */
    double dt_screen;
} r3d_hip_cfg;

/* Global 1-D metrics.  x arrays have npx+2 entries, y arrays npy+2, z arrays
 * npz+2, indexed like the reference's arrays of a single rank covering the
 * whole domain (interior 1..np, ghosts 0 and np+1).  Only interior entries
 * are read by the time step. */
typedef struct r3d_hip_metrics {
    const double *exx, *dxxdx, *d2xxdx2, *ddx;
    const double *wyy, *dyydy, *d2yydy2, *ddy;
    const double *zee, *dzzdz, *d2zzdz2, *ddz, *rkapa, *dkapa;
} r3d_hip_metrics;

/* Run scalars (the SimState scalar mirror, reference src/main.cpp:242-249). */
typedef struct r3d_hip_scalars {
    double dt, timc, timt, timi, umach;
    int nit;
} r3d_hip_scalars;

typedef struct r3d_hip_ctx r3d_hip_ctx;

/* Allocate the device state for cfg on cfg->device. */
int r3d_hip_create(const r3d_hip_cfg* cfg, r3d_hip_ctx** out);
int r3d_hip_destroy(r3d_hip_ctx* ctx);

/* Upload the global metrics (copied; host arrays only borrowed). */
int r3d_hip_set_metrics(r3d_hip_ctx* ctx, const r3d_hip_metrics* m);

/* Upload one virtual rank's primaries (fields[R3D_RU..R3D_BZ]; B entries are
 * ignored unless lmag), each nx*ny*nz doubles with ghosts.  Every rank must be
 * uploaded before the first step; ghost cells on the global boundary (periodic
 * images and the outer z planes) are taken from the ranks that own them. */
int r3d_hip_upload_rank(r3d_hip_ctx* ctx, int rank, const double* const* fields);

/* Upload the whole global state at once: fields[R3D_RU..R3D_BZ], each
 * (npx+2)*(npy+2)*(npz+2) doubles, x fastest. */
int r3d_hip_upload_global(r3d_hip_ctx* ctx, const double* const* fields);

/* Download one virtual rank: fields[R3D_RU..R3D_ZBZ] (NULL entries skipped),
 * each nx*ny*nz doubles.  Primaries come back with every ghost cell the
 * reference's halo exchange would hold; RK accumulators hold the interior
 * (ghost cells zero, as in a reference dump). */
int r3d_hip_download_rank(r3d_hip_ctx* ctx, int rank, double* const* fields);
int r3d_hip_download_global(r3d_hip_ctx* ctx, double* const* fields);

/* Asynchronous snapshot of the current state, for checkpoint dumps that
 * overlap stepping (the reference's dumps stop the run for a blocking
 * download, src/main.cpp:258-264).  begin copies X
 * and Z (dense global fields) and the scalars on the device, ordered after
 * every step issued so far, then to pinned host memory on a second stream;
 * it returns at once, and steps issued afterwards run while the copy is in
 * flight.  snapshot_rank waits for that copy and returns one virtual rank as
 * r3d_hip_download_rank would have at the snapshot's point (fields[R3D_RU..
 * R3D_ZBZ], NULL entries skipped) plus its scalars (s may be NULL).  The
 * buffers (2 x nfields global fields on the device and pinned on the host)
 * are allocated on the first begin.  A later begin reuses them: finish the
 * snapshot_rank calls of one snapshot before the next begin.  snapshot_rank
 * may be called from another host thread than the one stepping. */
int r3d_hip_snapshot_begin(r3d_hip_ctx* ctx);
int r3d_hip_snapshot_rank(r3d_hip_ctx* ctx, int rank, double* const* fields,
                          r3d_hip_scalars* s);

int r3d_hip_set_scalars(r3d_hip_ctx* ctx, const r3d_hip_scalars* s);
int r3d_hip_get_scalars(r3d_hip_ctx* ctx, r3d_hip_scalars* s);

/* Advance nsteps RK3 steps (nstages = 3), or one step truncated after
 * nstages = 1 or 2 stages (reference src/step.cpp:114-257).  Asynchronous on
 * the context stream except when tc != 0 (the plume time factor is taken
 * from timc on the host between steps). */
int r3d_hip_step(r3d_hip_ctx* ctx, int nsteps, int nstages);

/* r3d_hip_step that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * kernels' scratch, the per-workgroup dt partials that r3d_hip_step keeps in
 * the context's allocation, lives in a static device array of 1 MiB per GPU
 * (131072 doubles; enough for about 26000 stage tiles, e.g. 504x504x2048
 * takes 10066 x 5).  Larger grids get R3D_EINVAL.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (two contexts on
 * different streams) race.  Results equal r3d_hip_step's bit for bit.
 * Refused (R3D_ESTATE) while kernel timing is on, whose events are created
 * per launch; with tc != 0 it synchronises the stream once per step, as
 * r3d_hip_step does. */
int r3d_hip_step_noalloc(r3d_hip_ctx* ctx, int nsteps, int nstages);

/* Block until the context stream is idle. */
int r3d_hip_sync(r3d_hip_ctx* ctx);

/* Kernel timing: when enabled, each stage launch is bracketed by HIP events
 * on the context stream; r3d_hip_kernel_stats returns, per kernel class
 * (0 = fused stage kernel, 1 = boundary kernel, 2 = dt kernel, 3 = small
 * kernels), the number of launches and their total milliseconds since the
 * last reset. */
int r3d_hip_set_timing(r3d_hip_ctx* ctx, int enable);
int r3d_hip_kernel_stats(r3d_hip_ctx* ctx, int cls, long* launches, double* total_ms);

/* Device bytes held by the context. */
int r3d_hip_memory(r3d_hip_ctx* ctx, unsigned long long* bytes);

/* Stage-kernel tiling in use: out = {tile rows, z chunk, x tiles, y tiles,
 * z chunks}.  No reference counterpart (the reference GPU launches one
 * thread per cell). */
int r3d_hip_tiling(r3d_hip_ctx* ctx, int out[5]);

/* Thread-local message of the last failed call. */
const char* r3d_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

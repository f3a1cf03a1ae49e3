/* r3d_host.h - C ABI of the host-side setup for the gfx950 time step.
 *
 * Builds a run from the reference's command line (src/params.cpp:83-228):
 * parameters and derived constants (src/params.cpp:14-66), the global grid
 * metrics (src/grid.cpp), the conductivity profile (src/main.cpp:135-146) and
 * the initial condition of STATIC (src/static_ic.cpp) as one global padded
 * state, i.e. what the reference's ranks hold after static_ic + communicate
 * (src/main.cpp:148-154), laid out for r3d_hip_upload_global().
 *
 * Flags beyond the reference's: --npez N (virtual z ranks; the reference
 * derives npez from the MPI size), --seed S (ran2 seed for the temperature
 * perturbation; the reference hard-codes -62659, src/static_ic.cpp:119).
 * Errors return NULL / non-zero with r3d_host_last_error() set; the library
 * never exits the process.
 */
#ifndef R3D_HOST_H
#define R3D_HOST_H

#include "r3d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct r3d_host r3d_host;

/* Parse flags, derive constants, build metrics and the initial state (the
 * static initial condition, or with --nstart N > 0 the state of the dumps
 * --dumpdir/gstate.step.<rank>.<N>, --dumpdir defaulting to "."). */
r3d_host* r3d_host_create(int argc, const char* const* argv);
void r3d_host_free(r3d_host* h);
const char* r3d_host_last_error(void);

/* Device configuration for r3d_hip_create (device ordinal from --gpu). */
int r3d_host_cfg(const r3d_host* h, r3d_hip_cfg* out);
/* Global metrics (pointers stay valid while h lives). */
int r3d_host_metrics(const r3d_host* h, r3d_hip_metrics* out);
/* Global initial state: fields[R3D_RU..R3D_BZ], each (npx+2)(npy+2)(npz+2)
 * doubles, x fastest (B entries are NULL when lmag is off). */
int r3d_host_state(const r3d_host* h, const double* fields[8]);
/* dims = {npx, npy, npz, npey, npez, lmag, ntotal, nstep0}. */
void r3d_host_dims(const r3d_host* h, int dims[8]);
/* tb (bottom temperature of the static profile). */
double r3d_host_tb(const r3d_host* h);

/* Per-virtual-rank 1-D metrics exactly as the reference computes them for
 * that rank (src/grid.cpp:148-284), for gstate dumps: out[0..13] in gstate
 * order, lengths nx,nx,nx,nx, ny x4, nz x6. */
int r3d_host_rank_metrics(const r3d_host* h, int rank, double* const* out);

/* Write one virtual rank's gstate dump (reference format, src/main.cpp:44-93)
 * from downloaded fields[R3D_RU..R3D_ZBZ] (NULL = zeros) and the scalars. */
int r3d_host_write_gstate(const r3d_host* h, int rank, const char* path,
                          const double* const* fields, const r3d_hip_scalars* s);

/* Read one virtual rank's gstate dump (the same format; a dump of the
 * reference CPU build or of r3d_host_write_gstate) into fields[R3D_RU..
 * R3D_ZBZ] (NULL entries skipped, each nx*ny*nz doubles) and s.  Fails
 * unless the header describes this rank of this run (local extents, rank
 * coordinates, LMAG, boundary switches) and every array has its length.
 * The reference has no reader: its restart stops with "restart (NSTART!=0)
 * is not yet ported" (src/main.cpp:148-153). */
int r3d_host_read_gstate(const r3d_host* h, int rank, const char* path,
                         double* const* fields, r3d_hip_scalars* s);

/* Starting scalars of the run: zero for a fresh start; with --nstart N
 * (restart from --dumpdir/gstate.step.<rank>.<N>, every rank's dump) those
 * of the dumps, which also supply the global state. */
int r3d_host_scalars(const r3d_host* h, r3d_hip_scalars* out);

#ifdef __cplusplus
}
#endif
#endif

/* r3d_ens.h - batched ensemble of independent Rast-3DMHD runs on one GPU.
 *
 * Extension with no counterpart in the reference: many independent runs
 * (members) of uneven grid sizes, parameters and seeds advance together.
 * Each RK stage of every member is one launch over a tile table spanning all
 * members; each member keeps its own dt (its own CFL reduction), scalars and
 * state, and its results are bit-identical to the same run stepped alone
 * through r3d_hip_step.  A full step (three stages, the boundary passes and
 * the per-member dt reductions) is captured in a HIP graph and replayed.
 * Tile shapes are chosen by r3d_ens_autotune, which times the candidates on
 * the members themselves and restores their state afterwards.
 *
 * Members are described by r3d_hip_cfg (one per member, all on the same
 * device and all with the same lmag); everything else (metrics, state,
 * scalars) moves per member exactly as for a single run.  Errors follow
 * r3d_hip.h: negative R3D_E* codes, message in r3d_hip_last_error().
 */
#ifndef R3D_ENS_H
#define R3D_ENS_H

#include "r3d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct r3d_ens_ctx r3d_ens_ctx;

/* Allocate nmembers runs from cfgs[0..nmembers) (same device, same lmag,
 * tc = 0: the time-dependent plume needs a host value per member and step). */
int r3d_ens_create(int nmembers, const r3d_hip_cfg* cfgs, r3d_ens_ctx** out);
int r3d_ens_destroy(r3d_ens_ctx* e);

/* Per-member data movement, as r3d_hip_set_metrics / upload_global /
 * download_global / download_rank / set_scalars / get_scalars. */
int r3d_ens_set_metrics(r3d_ens_ctx* e, int member, const r3d_hip_metrics* m);
int r3d_ens_upload_global(r3d_ens_ctx* e, int member, const double* const* fields);
int r3d_ens_download_global(r3d_ens_ctx* e, int member, double* const* fields);
int r3d_ens_download_rank(r3d_ens_ctx* e, int member, int rank, double* const* fields);
int r3d_ens_set_scalars(r3d_ens_ctx* e, int member, const r3d_hip_scalars* s);
int r3d_ens_get_scalars(r3d_ens_ctx* e, int member, r3d_hip_scalars* s);

/* Tile shape: ty rows per tile (4, 8 or 9 for hydro, 5 for MHD) and kc planes
 * per z chunk (8..128; 0 = per-member default), 64 columns wide.  ntiles =
 * tiles per stage launch over all members. */
int r3d_ens_set_tiles(r3d_ens_ctx* e, int ty, int kc);
int r3d_ens_get_tiles(r3d_ens_ctx* e, int* ty, int* kc, int* ntiles);

/* The same with the tile width tw: 64 (the above), or, for hydro ensembles,
 * 32 columns with two rows per wave (ty 8 or 16) or 16 columns with four
 * (ty 16 or 28), which leave fewer lanes idle on members whose width is not
 * a multiple of 64 (80 and 96 wide: 83% and 100% of the lanes busy with 32
 * columns, 100% with 16, against 62.5% and 75%).  Results do not depend on
 * the tile shape. */
int r3d_ens_set_tile_shape(r3d_ens_ctx* e, int ty, int kc, int tw);
int r3d_ens_get_tile_shape(r3d_ens_ctx* e, int* ty, int* kc, int* tw, int* ntiles);

/* Time every candidate tile shape over trial_steps full steps (after one
 * untimed step), keep the fastest, and restore every member's state and
 * scalars to what they were before the call. */
int r3d_ens_autotune(r3d_ens_ctx* e, int trial_steps, double* best_ms_per_step);

/* enable != 0: steps replay captured HIP graphs (default); 0: plain launches. */
int r3d_ens_set_graph(r3d_ens_ctx* e, int enable);

/* enable != 0: each r3d_ens_step call runs all its steps in one persistent
 * cooperative launch (every block resident, grid-wide barriers between the
 * phases of a stage and of a step) instead of 7 launches per step; results
 * are bit-identical.  R3D_EUNSUP from r3d_ens_step where the device has no
 * cooperative launch. */
int r3d_ens_set_persistent(r3d_ens_ctx* e, int enable);

/* Advance every member by nsteps full RK3 steps (asynchronous). */
int r3d_ens_step(r3d_ens_ctx* e, int nsteps);
int r3d_ens_sync(r3d_ens_ctx* e);

/* Device bytes held by the ensemble (members included). */
int r3d_ens_memory(r3d_ens_ctx* e, unsigned long long* bytes);

#ifdef __cplusplus
}
#endif
#endif

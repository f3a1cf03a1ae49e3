/* r3d_adj.h - discrete adjoint of the Rast-3DMHD time step on one GPU.
 *
 * Extension with no counterpart in the reference.  For a run of nsteps full
 * RK3 steps from an initial state X0 it computes the objective
 *     J = sum over interior cells of vol * (ru^2 + rv^2 + rw^2) / (2 ro),
 *     vol = 1 / (dxxdx[i] * dyydy[j] * dzzdz[k])   (cell volume in units of
 *                                                   the computational cell),
 * at the final time, and the exact gradient of that discrete J (the same
 * floating-point program the forward path runs, differentiated by hand) with
 * respect to every primary of X0, the Reynolds number Re and the temperature
 * gradient theta (through the boundary heat-flux conditions).  The CFL step
 * dt depends on the state; its sensitivity enters through the cell and limit
 * that attain the minimum.
 *
 * Memory: the forward keeps a checkpoint every `checkpoint_every` steps; the
 * backward pass recomputes the states of one segment at a time from its
 * checkpoint, keeping the two intermediate stage states of every step.  The
 * last segment's states are kept by the forward itself (its stages write
 * into the segment buffers), so the first segment the backward transposes
 * needs no recompute; the buffers hold 3 * checkpoint_every + 1 states
 * either way.
 *
 * Hydrodynamic and MHD runs (with LMAG the state and the gradient have the
 * eight primaries R3D_RU..R3D_BZ); tc = 0.  Errors follow r3d_hip.h.
 */
#ifndef R3D_ADJ_H
#define R3D_ADJ_H

#include "r3d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct r3d_adj_ctx r3d_adj_ctx;

int r3d_adj_create(const r3d_hip_cfg* cfg, int checkpoint_every, r3d_adj_ctx** out);
int r3d_adj_destroy(r3d_adj_ctx* a);
int r3d_adj_set_metrics(r3d_adj_ctx* a, const r3d_hip_metrics* m);
/* X0: fields[R3D_RU..R3D_TT] (..R3D_BZ with LMAG), each (npx+2)(npy+2)(npz+2)
 * doubles, x fastest.  As with r3d_hip_upload_global, the outer z planes
 * (k = 0, npz+1) stay as given for the whole run and are read by the wall
 * cells' fluxes: they must hold finite values with RO > 0, e.g. those of a
 * downloaded state or of the initial condition. */
int r3d_adj_set_initial(r3d_adj_ctx* a, const double* const* fields);
/* Run nsteps steps from X0, keeping the checkpoints; J = objective. */
int r3d_adj_forward(r3d_adj_ctx* a, int nsteps, double* J);
/* Gradient of the last forward's J: grad[R3D_RU..R3D_TT] (..R3D_BZ with LMAG;
 * global arrays as X0; ghost entries zero), dJ/dRe, dJ/dtheta.  NULL outputs
 * are skipped. */
int r3d_adj_backward(r3d_adj_ctx* a, double* const* grad, double* dJ_dre, double* dJ_dtheta);
/* dJ/d(tb) of the last backward: the sensitivity to the fixed bottom
 * temperature of the default bottom condition (izcon = 0), the wall value
 * the boundary pass imposes (0 with a bottom heat-flux condition).  tb is
 * the static profile's bottom temperature, a function of theta: with dJ_dtheta
 * (theta in the heat-flux conditions, tb and X0 held) and grad (X0) it gives
 * the derivative through the initial profile as well. */
int r3d_adj_dJ_dtb(r3d_adj_ctx* a, double* dJ_dtb);
/* dt > 0: every step of later forward runs takes this dt instead of the CFL
 * step, and J is then a smooth function of X0, Re and theta (the CFL step is
 * a minimum over cells and limits, differentiable only where one of them
 * attains it).  dt = 0 restores the CFL step.  Invalidates the last forward. */
int r3d_adj_set_fixed_dt(r3d_adj_ctx* a, double dt);
/* Wall milliseconds of the last forward and backward calls (backward: up to
 * the gradient's copy to the host, which is not counted). */
int r3d_adj_timing(r3d_adj_ctx* a, double* forward_ms, double* backward_ms);
/* Final state of the last forward (as r3d_hip_download_global). */
int r3d_adj_download_final(r3d_adj_ctx* a, double* const* fields);
/* r3d_adj_set_initial and r3d_adj_backward with the fields and gradients in
 * device memory of the context's GPU (same dense global layout), for callers
 * that hold the state on the GPU already, e.g. a torch binding: no host copy
 * is made.  The caller's writes to the inputs must be complete (its stream
 * synchronized) before the call; the outputs are complete when it returns. */
int r3d_adj_set_initial_device(r3d_adj_ctx* a, const double* const* fields);
int r3d_adj_backward_device(r3d_adj_ctx* a, double* const* grad, double* dJ_dre,
                            double* dJ_dtheta);

#ifdef __cplusplus
}
#endif
#endif

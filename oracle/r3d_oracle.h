/* r3d_oracle.h - CPU oracle for the Rast-3DMHD RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (rast-3dmhd_amd/) never links or
 * calls it.
 *
 * It restates the reference CPU algorithm sweep-for-sweep (so even the
 * scratch arrays UU..WW3 of a gstate dump come out identical) with the
 * reference's per-rank arrays and ghost layers, but runs all npey x npez
 * ranks of a decomposition inside one process: the MPI halo exchange becomes
 * plain copies between the rank arrays and the dt reductions a loop over ranks.
 * Every function cites the reference file:line it follows.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   params / derive ........ reference KATs tests/params_test.cpp:11-51,84-142
 *   ran2 ................... tests/_ran2_ref.inc (every 1000th draw)
 *   static profile ......... tests/ran2_ode_test.cpp:56-61 + monotonicity
 *   mkgrid ................. closed forms (uniform grid; arctan grid end points)
 *   whole pipeline ......... byte-identical gstate dumps of the real reference
 *                            CPU build (oracle/_ref) for hydro configurations,
 *                            several decompositions and boundary branches
 *                            (fixtures under tests/golden/)
 *   LMAG (magnetic) ........ parity unpinned: the reference C++ aborts on LMAG
 *                            (src/fluxes.cpp:539-541); the oracle restates the
 *                            Fortran (golden/fortran/3dmhdsub.f) and is checked
 *                            only by invariants (B=0 reduces to hydro).
 *   ITCON=2 heat loss ...... parity unpinned: the reference C++ refuses it
 *                            (src/fluxes.cpp:50); the term restates
 *                            golden/fortran/3dmhdsub.f:1797-1822 and is checked
 *                            by closed forms (Gaussian sink of FWHM HH on a
 *                            fluid at rest, exact 1/RO scaling, the TANH ramp
 *                            at 1/10, TP=0 equal to the pinned ITCON=0 run).
 *   horizontal mean ........ bit-identical to the reference's own
 *                            horizontal_mean (src/communicate.cpp:184-223) run
 *                            from its sources (oracle/_ref/hmean, several
 *                            decompositions)
 *   density diffusion ID ... parity unpinned: the reference C++ refuses ID
 *                            (src/fluxes.cpp:49); the terms restate
 *                            golden/fortran/3dmhdsub.f:1306-1380 and are checked
 *                            by discrete-Laplacian closed forms (cosine modes,
 *                            quadratic z profiles, the mirrored wall planes,
 *                            1/ROM and the ID=2 wall profile).
 *   flux relaxation RLAX ... parity unpinned: the reference C++ refuses RLAX
 *                            (src/fluxes.cpp:48); restated from
 *                            golden/fortran/3dmhdsub.f:1534-1547, 1625-1771 for
 *                            ITC=0 / IZC=1 and checked by closed forms (linear
 *                            T: a linear correction across z ranks; FCONM of
 *                            cosine and uniform flows; TFAC steps and cap).
 */
#ifndef R3D_ORACLE_H
#define R3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct o_sim o_sim;

/* Build a simulation from reference-style CLI flags (src/params.cpp:83-228).
 * Extensions: --npez N (virtual z ranks; default 2), --seed S (ran2 seed,
 * default -62659).  --npey gives the virtual y ranks.  Returns NULL and sets
 * o_last_error() on invalid input. */
o_sim* o_create(int argc, const char* const* argv);
const char* o_last_error(void);
void o_free(o_sim* s);

/* Grid metrics, conductivity, STATIC initial condition, initial halo
 * exchange and the pre-loop fast-mode speed (src/main.cpp:133-171). */
int o_init(o_sim* s);

/* One RK3 step, or only the first nstages (1..3) stages (src/step.cpp:114). */
int o_step(o_sim* s, int nstages);
/* The time-step limits of STEP alone (reference src/step.cpp:121-229): UU,
 * VV (LMAG), WW1..WW3 over every rank's interior, umach and dt, without the
 * RK stages.  o_step begins with it.  Pinned against the reference's own
 * STEP with LMAG (tests/test_lmag_pins.py). */
int o_step_dt(o_sim* s);

/* LREM stratification (golden/fortran/3dmhdsub.f:437-490, 4442-4650;
 * golden/fortran/3dmhdset.f:154-170): per plane T, rho, the scaled
 * conductivity RK and its derivative DRK, and bnd = {TU, DZTU, DZTB}; and the
 * background T, rho at depth z alone.  Parity unpinned (no Fortran here):
 * pinned by the ODE's own identities (tests/test_lrem.py). */
int o_lrem_profile(const o_sim* s, double* t, double* r, double* rk, double* drk, double bnd[3]);
int o_lrem_background(double z, double out[2]);

/* Standalone pieces for invariant tests. */
int o_fluxes(o_sim* s, int rank);      /* src/fluxes.cpp:44 */
int o_bcon(o_sim* s, int rank);        /* src/bcon.cpp:12 */
/* Density diffusion (ID != 0): mirror RO's outer planes, then o_hmean of RO
 * (golden/fortran/3dmhdsub.f:1311-1328).  o_step calls it before FLUXES. */
int o_density_prep(o_sim* s);
/* Horizontal mean of field fid into every rank's o_rom() (src/communicate.cpp:184). */
int o_hmean(o_sim* s, int fid);
double* o_rom(o_sim* s, int rank);
/* Flux relaxation before FLUXES (RLAX > 0): TTM, TFAC, FCONM, CORRECT
 * (golden/fortran/3dmhdsub.f:1534-1547, 1625-1771).  o_step calls it. */
int o_relax_prep(o_sim* s);
void o_get_relax(const o_sim* s, double v[2]); /* TFAC, TTM_MAX */
double* o_correct(o_sim* s, int rank);         /* CORRECT per local plane */
void o_communicate(o_sim* s);          /* src/communicate.cpp:171 */

/* Introspection.  dims = {nx, ny, nz, npey, npez, nranks, npx, npy, npz, lmag}. */
void o_dims(const o_sim* s, int dims[10]);
/* Field ids follow the gstate order: ru rv rw ro tt uu vv ww fu fv fw fr ft
 * zru zrv zrw zro ztt ww1 ww2 ww3 bx by bz zbx zby zbz (0..26). */
double* o_field(o_sim* s, int rank, int fid);
/* Metric ids: exx dxxdx d2xxdx2 ddx wyy dyydy d2yydy2 ddy zee dzzdz d2zzdz2
 * ddz rkapa dkapa (0..13). */
double* o_metric(o_sim* s, int rank, int mid);
/* scal = {dt, timc, timt, timi, umach, tb, nit}. */
void o_get_scalars(const o_sim* s, double scal[7]);
void o_set_scalars(o_sim* s, const double scal[7]);
/* Derived constants in this order: c13 c23 c43 gam1 gam2 gam3 zeta1 zeta2
 * hx h2x hy h2y hz h2z repr ore cv ocv rkapst orm obeta omx omz,
 * then params pr gamma ampt sf (27 values). */
void o_get_derived(const o_sim* s, double out[27]);

/* gstate dump of one rank in the reference format (src/main.cpp:44-93). */
int o_write_gstate(const o_sim* s, int rank, const char* path);

/* Known-answer helpers. */
void o_ran2_sequence(int seed, int n, double* out);
int o_static_profile(const o_sim* s, double* tprof, double* rprof); /* len npz */
void o_mkgrid(const o_sim* s, double scode, int icoord, double out[3]);

#ifdef __cplusplus
}
#endif
#endif

/* r3d_oracle.c - CPU oracle for the Rast-3DMHD RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY (see r3d_oracle.h): a plain-C restatement of the
 * reference CPU path that the HIP product is checked against.  Built with
 * -O2 -ffp-contract=off so that every expression rounds exactly as the
 * reference build (-O2 -ffp-contract=off, reference makefile:104) does.
 *
 * Layout: one array set per virtual rank, x fastest, element (i,j,k) at
 * i + nx*(j + ny*k) (reference include/fields.hpp:29-33), ghosts included.
 */
#include "r3d_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#define O_PAR _Pragma("omp parallel for schedule(static)")
#else
#define O_PAR
#endif

static char g_err[512];
const char* o_last_error(void) { return g_err; }
static int fail(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

/* Fortran single-precision literal widened to double
 * (reference include/numcompat.hpp:17-21). */
static double w4(float x) { return (double)x; }

/* ------------------------------------------------------------------------ */
/* Parameters: reference include/params.hpp:24-120 (defaults),              */
/* src/params.cpp:83-228 (CLI).                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    int npx, npy, npz, npey, npez, ix, iy, ilap, ngrid, id, ncol, nrow;
    int ixcon, iycon, izcon, itcon, ibcon;
    int lrot, lmag, lpot, lrem, lshr;
    int ncase, ncasep, ntotal, nstep0, nstart, ntube, gpu_id, backend_gpu, seed;
    double re, pr, theta, grav, r_y, ang, rm, beta, pzp, sigma, polys, tp, tc;
    double xp, yp, zp, xmax, ymax, zmax, ampt, ampb, bfh, bzp, qfh, gamma, hh;
    double dh, dv, tstart, toff, rlax;
    double xx1, xx2, yy1, yy2, zz1, zz2;
    double xa, xb, xc, xd, ya, yb, yc, yd, za, zb, zc, zd;
    double xcent, zcent, rmax, cmt, a_tube, lambda, vz0, sf;
} o_params;

static void params_default(o_params* p) {
    memset(p, 0, sizeof *p);
    p->npx = 504; p->npy = 504; p->npz = 2048; p->npey = 8; p->npez = 2;
    p->ix = 2; p->iy = 2; p->ilap = 2; p->ngrid = 1;
    p->itcon = 1;
    p->ncase = 1; p->ncasep = 1; p->ntotal = 1000; p->nstep0 = 100;
    p->seed = -62659;
    p->re = 2.0; p->pr = w4(5.0e-2f); p->theta = 2.5e-1; p->grav = 6.25e-1;
    p->tp = 2.0; p->xp = 10.0; p->yp = 10.0;
    p->xmax = 20.0; p->ymax = 20.0; p->zmax = 40.0;
    p->ampt = w4(5.0e-3f); p->gamma = w4(5.0f / 3.0f); p->hh = 1.0;
    p->xx1 = -7.0; p->xx2 = 7.0; p->yy1 = -7.0; p->yy2 = 7.0;
    p->zz1 = -2.0; p->zz2 = 1.0e-2;
    p->xa = 0.5; p->xb = 0.48; p->xc = 28.0; p->xd = 0.025;
    p->ya = 0.5; p->yb = 0.48; p->yc = 28.0; p->yd = 0.025;
    p->za = 0.65; p->zb = 1.0; p->zc = 5.0; p->zd = 0.2;
    p->sf = w4(0.315f);
}

enum { T_INT, T_DBL, T_BOOL };
typedef struct { const char* name; int type; size_t off; } flagdef;
#define FI(n, f) {n, T_INT, offsetof(o_params, f)}
#define FD(n, f) {n, T_DBL, offsetof(o_params, f)}
#define FB(n, f) {n, T_BOOL, offsetof(o_params, f)}
static const flagdef k_flags[] = {
    FI("--npx", npx), FI("--npy", npy), FI("--npz", npz), FI("--npey", npey),
    FI("--npez", npez), FI("--ix", ix), FI("--iy", iy), FI("--ilap", ilap),
    FI("--ngrid", ngrid), FI("--id", id), FI("--ncol", ncol), FI("--nrow", nrow),
    FI("--ixcon", ixcon), FI("--iycon", iycon), FI("--izcon", izcon),
    FI("--itcon", itcon), FI("--ibcon", ibcon), FI("--ncase", ncase),
    FI("--ncasep", ncasep), FI("--ntotal", ntotal), FI("--nstep0", nstep0),
    FI("--nstart", nstart), FI("--ntube", ntube), FI("--gpu", gpu_id),
    FI("--seed", seed),
    FD("--re", re), FD("--pr", pr), FD("--theta", theta), FD("--grav", grav),
    FD("--ry", r_y), FD("--ang", ang), FD("--rm", rm), FD("--beta", beta),
    FD("--pzp", pzp), FD("--sigma", sigma), FD("--polys", polys), FD("--tp", tp),
    FD("--tc", tc), FD("--xp", xp), FD("--yp", yp), FD("--zp", zp),
    FD("--xmax", xmax), FD("--ymax", ymax), FD("--zmax", zmax), FD("--ampt", ampt),
    FD("--ampb", ampb), FD("--bfh", bfh), FD("--bzp", bzp), FD("--qfh", qfh),
    FD("--gamma", gamma), FD("--hh", hh), FD("--dh", dh), FD("--dv", dv),
    FD("--tstart", tstart), FD("--toff", toff), FD("--rlax", rlax),
    FD("--xx1", xx1), FD("--xx2", xx2), FD("--yy1", yy1), FD("--yy2", yy2),
    FD("--zz1", zz1), FD("--zz2", zz2), FD("--xa", xa), FD("--xb", xb),
    FD("--xc", xc), FD("--xd", xd), FD("--ya", ya), FD("--yb", yb), FD("--yc", yc),
    FD("--yd", yd), FD("--za", za), FD("--zb", zb), FD("--zc", zc), FD("--zd", zd),
    FD("--xcent", xcent), FD("--zcent", zcent), FD("--rmax", rmax), FD("--cmt", cmt),
    FD("--a", a_tube), FD("--lambda", lambda), FD("--vz0", vz0), FD("--sff", sf),
    FB("--lrot", lrot), FB("--lmag", lmag), FB("--lpot", lpot), FB("--lrem", lrem),
    FB("--lshr", lshr),
};

static int parse_args(int argc, const char* const* argv, o_params* p) {
    for (int a = 0; a < argc; ++a) {
        const char* arg = argv[a];
        const char* val = NULL;
        const flagdef* hit = NULL;
        for (size_t f = 0; f < sizeof k_flags / sizeof k_flags[0]; ++f) {
            size_t L = strlen(k_flags[f].name);
            if (strncmp(arg, k_flags[f].name, L) == 0 && (arg[L] == 0 || arg[L] == '=')) {
                hit = &k_flags[f];
                val = arg[L] == '=' ? arg + L + 1 : NULL;
                break;
            }
        }
        if (!hit && strncmp(arg, "--backend", 9) == 0 && (arg[9] == 0 || arg[9] == '=')) {
            const char* v = arg[9] == '=' ? arg + 10 : (a + 1 < argc ? argv[++a] : NULL);
            if (!v) return fail("missing value for --backend");
            if (strcmp(v, "cpu") == 0) p->backend_gpu = 0;
            else if (strcmp(v, "gpu") == 0) p->backend_gpu = 1;
            else return fail("--backend must be 'cpu' or 'gpu'");
            continue;
        }
        if (!hit) return fail("unknown option '%s'", arg);
        if (!val) {
            if (a + 1 >= argc) return fail("missing value for %s", hit->name);
            val = argv[++a];
        }
        char* base = (char*)p + hit->off;
        if (hit->type == T_INT) *(int*)base = atoi(val);
        else if (hit->type == T_DBL) *(double*)base = atof(val);
        else *(int*)base = (!strcmp(val, "1") || !strcmp(val, "true") ||
                            !strcmp(val, "yes") || !strcmp(val, "on"));
    }
    return 0;
}

/* Derived constants: reference src/params.cpp:14-66. */
typedef struct {
    int nry, nrz, nx, ny, nz;
    double c13, c23, c43, gam1, gam2, gam3, zeta1, zeta2;
    double hx, h2x, hy, h2y, hz, h2z;
    double repr, ore, cv, ocv, rkapst, orm, obeta, omx, omz;
} o_derived;

static void derive(const o_params* p, o_derived* d) {
    d->nry = p->npy / p->npey;
    d->nrz = p->npz / p->npez;
    d->nx = p->npx + p->ix;
    d->ny = d->nry + p->iy;
    d->nz = d->nrz + p->ilap;
    d->c13 = w4(1.0f / 3.0f);
    d->c23 = 2.0 * d->c13;
    d->c43 = 4.0 * d->c13;
    d->gam1 = w4(8.0f / 15.0f);
    d->gam2 = w4(5.0f / 12.0f);
    d->gam3 = w4(3.0f / 4.0f);
    d->zeta1 = w4(-17.0f / 60.0f);
    d->zeta2 = w4(-5.0f / 12.0f);
    d->hx = w4(0.5f * (float)(p->npx - 1));
    d->h2x = w4((float)(p->npx - 1) * (float)(p->npx - 1));
    d->hy = w4(0.5f * (float)(p->npy - 1));
    d->h2y = w4((float)(p->npy - 1) * (float)(p->npy - 1));
    d->hz = w4(0.5f * (float)(p->npz - 1));
    d->h2z = w4((float)(p->npz - 1) * (float)(p->npz - 1));
    d->repr = p->re * p->pr;
    d->ore = 1.0 / p->re;
    d->cv = 1.0 / (p->gamma - 1.0);
    d->ocv = 1.0 / d->cv;
    d->rkapst = (p->grav == 0.0) ? 1.0 : (p->polys + 1.0) * p->theta / p->grav;
    d->orm = p->lmag ? 1.0 / p->rm : 0.0;
    d->obeta = p->lmag ? 2.0 / p->beta : 0.0;
    d->omx = p->lrot ? sin(p->ang) / p->r_y : 0.0;
    d->omz = p->lrot ? cos(p->ang) / p->r_y : 0.0;
}

/* ------------------------------------------------------------------------ */
/* Grid mapping: reference src/grid.cpp:28-146 (MKGRID).                    */
/* ------------------------------------------------------------------------ */
static void mkgrid(const o_params* p, double sc, int ic, double* sphys, double* dsds,
                   double* d2sds2) {
    const double pi = 2.0 * asin(1.0);
    const double smax = ic == 1 ? p->xmax : ic == 2 ? p->ymax : p->zmax;
    if (p->ngrid == 0) {
        *sphys = sc * smax;
        *dsds = 1.0 / smax;
        *d2sds2 = 0.0;
        return;
    }
    if (p->ngrid == 1) {
        const double s1 = ic == 1 ? p->xx1 : ic == 2 ? p->yy1 : p->zz1;
        const double s2 = ic == 1 ? p->xx2 : ic == 2 ? p->yy2 : p->zz2;
        const double a1 = (s2 - s1) / smax;
        const double a3 = atan(s1);
        const double a2 = atan(s2) - a3;
        double x;
        if (sc == 0.0) x = 0.0;
        else if (sc == 1.0) x = smax;
        else x = (tan(a2 * sc + a3) - s1) / a1;
        *sphys = x;
        *dsds = a1 / a2 / (1.0 + (s1 + a1 * x) * (s1 + a1 * x));
        const double t = s1 + a1 * x;
        const double q = 1.0 + t * t;
        *d2sds2 = -2.0 * a1 * a1 * t / a2 / (q * q);
        return;
    }
    /* ngrid == 2: two-arctan stretching (reference src/grid.cpp:78-141). */
    const double a = ic == 1 ? p->xa : ic == 2 ? p->ya : p->za;
    const double b = ic == 1 ? p->xb : ic == 2 ? p->yb : p->zb;
    const double c = ic == 1 ? p->xc : ic == 2 ? p->yc : p->zc;
    const double dp = ic == 1 ? p->xd : ic == 2 ? p->yd : p->zd;
    if (c <= 1.0e-9) {
        *sphys = sc * smax;
        *dsds = 1.0 / smax;
        *d2sds2 = 0.0;
        return;
    }
    const double amb = a - b, apb = a + b;
    const double sh = (1.0 - dp) * pi /
                      (dp * atan(amb * c) + 2.0 * atan(b * c) - dp * atan(apb * c));
    const double sk =
        2.0 * c * pi * smax /
        (2.0 * c * pi +
         sh * (2.0 * c *
                   (atan(amb * c) * amb - atan(apb * c) * apb + atan((a - b - 1.0) * c) +
                    atan((a + b - 1.0) * c) * (a + b - 1.0) + atan((1.0 - a + b) * c) * amb) -
               log(1.0 + amb * amb * c * c) + log(1.0 + apb * apb * c * c) -
               log(1.0 + (a + b - 1.0) * (a + b - 1.0) * c * c) +
               log(1.0 + (1.0 - a + b) * (1.0 - a + b) * c * c)));
    const double ds = 1.0 / (sk * (1.0 + sh / pi * (atan(c * (sc - a - b)) + atan(c * (a - sc - b)))));
    *dsds = ds;
    *d2sds2 = -ds * ds * ds * c * sh * sk / pi *
              (1.0 / (1.0 + c * c * (a + b - sc) * (a + b - sc)) -
               1.0 / (1.0 + c * c * (a - b - sc) * (a - b - sc)));
    double x = sk / (2.0 * c * pi) *
               (2.0 * c * pi * sc +
                sh * (2.0 * c *
                          (atan(amb * c) * amb - atan(apb * c) * apb + atan((a - b - sc) * c) * sc +
                           atan((a + b - sc) * c) * (a + b - sc) + atan((sc - a + b) * c) * amb) -
                      log(1.0 + amb * amb * c * c) + log(1.0 + apb * apb * c * c) -
                      log(1.0 + (a + b - sc) * (a + b - sc) * c * c) +
                      log(1.0 + (sc - a + b) * (sc - a + b) * c * c)));
    if (sc == 0.0) x = 0.0;
    if (sc == 1.0) x = smax;
    *sphys = x;
}

/* Fill one metric quadruple (coordinate, d/ds, d2/ds2, spacing) for the local
 * index range of one direction.  `code[m]` is the computational coordinate of
 * local index m; the three-branch structure of the reference (first rank,
 * last rank, middle ranks: src/grid.cpp:148-276) reduces to choosing `code`. */
static void metric_line(const o_params* p, int ic, int n, const double* code, double spacing,
                        double* s, double* ds, double* d2s, double* dd) {
    for (int m = 0; m < n; ++m) mkgrid(p, code[m], ic, &s[m], &ds[m], &d2s[m]);
    for (int m = 0; m < n; ++m) dd[m] = spacing * (1.0 / ds[m]);
}

/* Computational coordinates seen by local index m of rank `myp` out of `np`
 * ranks along a decomposed direction with `npt` global points, nr local
 * interior points and half-width h (reference src/grid.cpp:169-276).  The
 * reference reads one element past the end of its coordinate vector for the
 * last ghost when a direction has a single rank (npey == 1); that ghost metric
 * is never used by the time step, and here it continues the uniform sequence. */
static void comp_coords(int myp, int np, int npt, int nr, int h, int n, double orv, double* code) {
    double* g = malloc(sizeof(double) * (size_t)(npt + 2 * h + 1));
    for (int m = 0; m < npt + 2 * h + 1; ++m) g[m] = m * orv;
    g[0] = 0.0;
    g[npt - 1] = 1.0;
    if (myp == 0) {
        for (int m = 0; m < h; ++m) code[m] = 0.0;
        for (int m = h; m < nr + 2 * h; ++m) code[m] = g[m - h];
    } else if (myp == np - 1) {
        for (int m = 0; m < nr + h; ++m) code[m] = g[m + npt - nr - h];
        for (int m = nr + h; m < nr + 2 * h; ++m) code[m] = 1.0;
    } else {
        for (int m = 0; m < n; ++m) code[m] = g[m + myp * nr - h];
    }
    free(g);
}

/* ------------------------------------------------------------------------ */
/* RAN2: reference src/ran2.cpp:17-37 (Numerical Recipes ran2 with a 97-entry */
/* shuffle table).                                                          */
/* ------------------------------------------------------------------------ */
typedef struct { int idum, iiy, iir[97]; } o_ran2;

static double ran2(o_ran2* r) {
    const int M = 714025, IA = 1366, IC = 150889;
    const double RM = w4(1.0f / 714025.0f);
    if (r->idum < 0) {
        r->idum = (IC - r->idum) % M;
        for (int j = 0; j < 97; ++j) {
            r->idum = (IA * r->idum + IC) % M;
            r->iir[j] = r->idum;
        }
        r->idum = (IA * r->idum + IC) % M;
        r->iiy = r->idum;
    }
    const int j = (97 * r->iiy) / M; /* 0-based slot */
    if (j < 0 || j > 96) abort();
    r->iiy = r->iir[j];
    const double v = (double)r->iiy * RM;
    r->idum = (IA * r->idum + IC) % M;
    r->iir[j] = r->idum;
    return v;
}

void o_ran2_sequence(int seed, int n, double* out) {
    o_ran2 r;
    memset(&r, 0, sizeof r);
    r.idum = seed;
    for (int i = 0; i < n; ++i) out[i] = ran2(&r);
}

/* ------------------------------------------------------------------------ */
/* Static stratification: Bulirsch-Stoer over the arctan z map.             */
/* Reference src/ode.cpp:23-174 (DERIVS, RZEXTR, MMID, BSSTEP, STATIC).      */
/* ------------------------------------------------------------------------ */
static void derivs(const o_params* p, double zz, const double* y, double* dydx) {
    double z, dz, d2z;
    mkgrid(p, zz, 3, &z, &dz, &d2z);
    const double rkapst = (p->grav == 0.0) ? 1.0 : (p->polys + 1.0) * p->theta / p->grav;
    const double rkap =
        (p->pzp == 0.0) ? 1.0 : 1.0 + (rkapst - 1.0) / 2.0 * (1.0 + tanh((z - p->pzp) / p->sigma));
    dydx[0] = p->theta / rkap / dz;
    dydx[1] = p->grav / (1.0 + y[0]) / dz;
}

enum { BS_NUSE = 7, BS_IMAX = 11 };
static const int k_nseq[BS_IMAX] = {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96};

static void rat_extrap(int iest, double xest, const double* yest, double* yz,
                       double tab[2][BS_NUSE], double* dy, double* xs) {
    xs[iest - 1] = xest;
    if (iest == 1) {
        for (int v = 0; v < 2; ++v) yz[v] = tab[v][0] = dy[v] = yest[v];
        return;
    }
    const int m1 = iest < BS_NUSE ? iest : BS_NUSE;
    double fx[BS_NUSE + 1];
    for (int q = 1; q < m1; ++q) fx[q] = xs[iest - q - 1] / xest;
    for (int v = 0; v < 2; ++v) {
        double yy = yest[v], prev = tab[v][0], c = yy, ddy = 0.0;
        tab[v][0] = yy;
        for (int q = 1; q < m1; ++q) {
            const double b1 = fx[q] * prev;
            double b = b1 - c;
            if (b != 0.0) {
                b = (c - prev) / b;
                ddy = c * b;
                c = b1 * b;
            } else {
                ddy = prev;
            }
            prev = tab[v][q];
            tab[v][q] = ddy;
            yy += ddy;
        }
        dy[v] = ddy;
        yz[v] = yy;
    }
}

static void midpoint(const o_params* p, const double* y, const double* dydx, double xs,
                     double htot, int nstep, double* yout) {
    const double h = htot / nstep;
    double ym[2], yn[2], dv[2];
    for (int v = 0; v < 2; ++v) {
        ym[v] = y[v];
        yn[v] = y[v] + h * dydx[v];
    }
    double x = xs + h;
    derivs(p, x, yn, dv);
    const double h2 = 2.0 * h;
    for (int n = 1; n < nstep; ++n) {
        for (int v = 0; v < 2; ++v) {
            const double t = ym[v] + h2 * dv[v];
            ym[v] = yn[v];
            yn[v] = t;
        }
        x += h;
        derivs(p, x, yn, dv);
    }
    for (int v = 0; v < 2; ++v) yout[v] = 0.5 * (ym[v] + yn[v] + h * dv[v]);
}

/* Returns the step actually taken. */
static double bs_step(const o_params* p, double* y, const double* dydx, double* x, double htry,
                      double eps, double* xs) {
    double h = htry;
    const double x0 = *x;
    double y0[2] = {y[0], y[1]}, d0[2] = {dydx[0], dydx[1]};
    double yseq[2], yerr[2], tab[2][BS_NUSE];
    for (;;) {
        for (int i = 1; i <= BS_IMAX; ++i) {
            midpoint(p, y0, d0, x0, h, k_nseq[i - 1], yseq);
            const double xest = (h / k_nseq[i - 1]) * (h / k_nseq[i - 1]);
            rat_extrap(i, xest, yseq, y, tab, yerr, xs);
            double emax = 0.0;
            for (int v = 0; v < 2; ++v) {
                const double e = fabs(yerr[v] / y[v]);
                if (e > emax) emax = e;
            }
            emax /= eps;
            if (emax < 1.0) {
                *x += h;
                return h;
            }
        }
        h = 0.25 * h / pow(2.0, (double)((BS_IMAX - BS_NUSE) / 2));
        if (*x + h == *x) abort();
    }
}

static int static_profile(const o_params* p, int npz, double* tp, double* rp) {
    const double dzz = w4(1.0f / (float)(npz - 1));
    double y[2] = {0.0, 0.0}, zz = 0.0, xs[BS_IMAX];
    tp[0] = 1.0;
    rp[0] = 1.0;
    for (int k = 1; k < npz; ++k) {
        double dydx[2];
        derivs(p, zz, y, dydx);
        const double did = bs_step(p, y, dydx, &zz, dzz, 1.0e-12, xs);
        if (did != dzz) return fail("static profile: Bulirsch-Stoer step shrank at k=%d", k);
        tp[k] = 1.0 + y[0];
        rp[k] = exp(y[1]) / (1.0 + y[0]);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* LREM stratification (M. Rempel): golden/fortran/3dmhdsub.f:437-490       */
/* (STATIC), 4442-4650 (GETBACKGROUND, DERIVS1, ODEINT, RKQC, RK4) and      */
/* golden/fortran/3dmhdset.f:154-170 (KAPPA).  Not in the reference C++,    */
/* which refuses LREM (src/static_ic.cpp:27).  The Fortran is compiled      */
/* without -fdefault-real-8 (golden/build_golden.sh:52-55), so its          */
/* single-precision literals are widened floats here (w4).                  */
/* ------------------------------------------------------------------------ */
static double lrem_kappa(double z, double rho) {
    double krad = rho;
    krad = krad + (1.25 - krad) * exp(-(5.0 * z));
    return krad;
}

static void lrem_derivs(double x, const double* yt, double* dyt) {
    const double nad = w4(0.41f), os = 1.0;
    const double t = yt[0], rho = yt[1] / yt[0];
    const double nrad = w4(0.4f) / lrem_kappa(x, rho);
    (void)t;
    dyt[0] = nrad <= os * nad ? nrad : nad;
    dyt[1] = yt[1] / yt[0];
}

static void lrem_rk4(const double* y, const double* dydx, double x, double h, double* yout) {
    double yt[2], dyt[2], dym[2];
    const double hh = h * 0.5, h6 = h / 6.0, xh = x + hh;
    for (int i = 0; i < 2; ++i) yt[i] = y[i] + hh * dydx[i];
    lrem_derivs(xh, yt, dyt);
    for (int i = 0; i < 2; ++i) yt[i] = y[i] + hh * dyt[i];
    lrem_derivs(xh, yt, dym);
    for (int i = 0; i < 2; ++i) {
        yt[i] = y[i] + h * dym[i];
        dym[i] = dyt[i] + dym[i];
    }
    lrem_derivs(x + h, yt, dyt);
    for (int i = 0; i < 2; ++i) yout[i] = y[i] + h6 * (dydx[i] + dyt[i] + 2.0 * dym[i]);
}

/* step doubling with error control (RKQC); returns -1 where the Fortran PAUSEs */
static int lrem_rkqc(double* y, double* dydx, double* x, double htry, double eps,
                     const double* yscal, double* hnext) {
    const double fcor = 1.0 / 15.0, safety = w4(0.9f), errcon = 6.0e-4;
    const double pgrow = w4(-0.20f), pshrnk = -0.25;
    const double xsav = *x;
    const double ysav[2] = {y[0], y[1]}, dysav[2] = {dydx[0], dydx[1]};
    double ytemp[2], h = htry;
    for (;;) {
        const double hh = 0.5 * h;
        lrem_rk4(ysav, dysav, xsav, hh, ytemp);
        *x = xsav + hh;
        lrem_derivs(*x, ytemp, dydx);
        lrem_rk4(ytemp, dydx, *x, hh, y);
        *x = xsav + h;
        if (*x == xsav) return fail("LREM background: step size not significant");
        lrem_rk4(ysav, dysav, xsav, h, ytemp);
        double errmax = 0.0;
        for (int i = 0; i < 2; ++i) {
            ytemp[i] = y[i] - ytemp[i];
            const double e = fabs(ytemp[i] / yscal[i]);
            errmax = errmax > e ? errmax : e;
        }
        errmax = errmax / eps;
        if (errmax > 1.0) {
            h = safety * h * pow(errmax, pshrnk);
            continue;
        }
        *hnext = errmax > errcon ? safety * h * pow(errmax, pgrow) : 4.0 * h;
        break;
    }
    for (int i = 0; i < 2; ++i) y[i] = y[i] + ytemp[i] * fcor;
    return 0;
}

/* ODEINT from x1 to x2 keeping only the end point (KMAX = 1) */
static int lrem_odeint(const double* ystart, double x1, double x2, double eps, double h1,
                       double hmin, double* yend) {
    const double tiny = w4(1.0e-30f);
    double x = x1, h = copysign(fabs(h1), x2 - x1);
    double y[2] = {ystart[0], ystart[1]}, dydx[2] = {0.0, 0.0}, yscal[2];
    if (x2 == x1) { /* the Fortran's first step would be 0 and PAUSE */
        yend[0] = y[0];
        yend[1] = y[1];
        return 0;
    }
    for (int nstp = 0; nstp < 1000; ++nstp) {
        lrem_derivs(x, y, dydx);
        for (int i = 0; i < 2; ++i) yscal[i] = fabs(y[i]) + fabs(h * dydx[i]) + tiny;
        if ((x + h - x2) * (x + h - x1) > 0.0) h = x2 - x;
        double hnext = 0.0;
        if (lrem_rkqc(y, dydx, &x, h, eps, yscal, &hnext)) return -1;
        if ((x - x2) * (x2 - x1) >= 0.0) {
            yend[0] = y[0];
            yend[1] = y[1];
            return 0;
        }
        if (fabs(hnext) < hmin) return fail("LREM background: step size smaller than minimum");
        h = hnext;
    }
    return fail("LREM background: too many steps");
}

/* T(z), rho(z): integrate from z = 2 (T = 1, p = rho T = 1) to z */
static int lrem_background(double z, double* t, double* rho) {
    const double ys[2] = {1.0, 1.0};
    double y[2] = {0.0, 0.0};
    if (lrem_odeint(ys, 2.0, z, 1.0e-8, 1.0e-4, 1.0e-50, y)) return -1;
    *t = y[0];
    *rho = y[1] / y[0];
    return 0;
}

/* STATIC's LREM profile on the npz planes: T, R, the conductivity RK scaled
 * by 8.07*THETA*REPR and its derivative DRK (one-sided at the ends), and the
 * boundary values TU = T(1), DZTU, DZTB (golden/fortran/3dmhdsub.f:437-490) */
static int lrem_profile(const o_params* p, const o_derived* d, int npz, double* T, double* R,
                        double* RK, double* DRK, double bnd[3]) {
    for (int k = 0; k < npz; ++k) {
        const double szz = w4((float)k / (float)(npz - 1));
        double sz, sd, sd2;
        mkgrid(p, szz, 3, &sz, &sd, &sd2);
        double tz, rz;
        if (lrem_background(sz, &tz, &rz)) return -1;
        T[k] = tz;
        R[k] = rz;
        RK[k] = lrem_kappa(sz, rz);
    }
    const double c807 = w4(8.07f);
    for (int k = 0; k < npz; ++k) RK[k] = RK[k] * c807 * p->theta * d->repr;
    for (int k = 1; k < npz - 1; ++k) {
        double sz, sd, sd2;
        mkgrid(p, w4((float)k / (float)(npz - 1)), 3, &sz, &sd, &sd2);
        DRK[k] = (RK[k + 1] - RK[k - 1]) * d->hz * sd;
    }
    double sz, sd, sd2;
    mkgrid(p, 0.0, 3, &sz, &sd, &sd2);
    DRK[0] = (-3.0 * RK[0] + 4.0 * RK[1] - RK[2]) * d->hz * sd;
    mkgrid(p, 1.0, 3, &sz, &sd, &sd2);
    DRK[npz - 1] = (3.0 * RK[npz - 1] - 4.0 * RK[npz - 2] + RK[npz - 3]) * d->hz * sd;
    const double f43 = w4(4.0f / 3.0f), f13 = w4(1.0f / 3.0f);
    bnd[0] = T[0];
    bnd[1] = T[0] - f43 * T[1] + f13 * T[2];
    bnd[2] = T[npz - 1] - f43 * T[npz - 2] + f13 * T[npz - 3];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Simulation state.                                                        */
/* ------------------------------------------------------------------------ */
enum {
    F_RU, F_RV, F_RW, F_RO, F_TT, F_UU, F_VV, F_WW, F_FU, F_FV, F_FW, F_FR, F_FT,
    F_ZRU, F_ZRV, F_ZRW, F_ZRO, F_ZTT, F_W1, F_W2, F_W3, F_BX, F_BY, F_BZ,
    F_ZBX, F_ZBY, F_ZBZ, NF3
};
enum { M_EXX, M_DXX, M_D2XX, M_DDX, M_WYY, M_DYY, M_D2YY, M_DDY, M_ZEE, M_DZZ, M_D2ZZ,
       M_DDZ, M_RKAPA, M_DKAPA, NM1 };

typedef struct {
    int mype, mypey, mypez;
    double* f[NF3];
    double* m[NM1];
    double* rom; /* horizontal mean of RO per local plane (density diffusion) */
    double *ttm, *fconm, *correct; /* flux relaxation, per local plane */
} o_rank;

struct o_sim {
    o_params p;
    o_derived d;
    int nranks;
    o_rank* r;
    double dt, timc, timt, timi, umach, tb;
    double tu, dztu, dztb; /* LREM boundary values of STATIC (COMMON/SPECIALBOUND/) */
    double tfac, ttm_max; /* flux relaxation state kept across calls (SAVE / DATA) */
    int nit;
};

static long ncell(const o_sim* s) { return (long)s->d.nx * s->d.ny * s->d.nz; }

o_sim* o_create(int argc, const char* const* argv) {
    o_params p;
    params_default(&p);
    if (parse_args(argc, argv, &p)) return NULL;
    if (p.ix != 2 || p.iy != 2 || p.ilap != 2) {
        fail("only ix=iy=ilap=2 is supported");
        return NULL;
    }
    if (p.npey < 1 || p.npez < 1 || p.npy % p.npey || p.npz % p.npez) {
        fail("npey=%d/npez=%d must divide npy=%d/npz=%d", p.npey, p.npez, p.npy, p.npz);
        return NULL;
    }
    if (p.npz / p.npez < 3 || p.npy / p.npey < 1 || p.npx < 3) {
        fail("subdomains too small");
        return NULL;
    }
    o_sim* s = calloc(1, sizeof *s);
    s->p = p;
    derive(&p, &s->d);
    s->nranks = p.npey * p.npez;
    s->r = calloc((size_t)s->nranks, sizeof(o_rank));
    const long n = ncell(s);
    for (int r = 0; r < s->nranks; ++r) {
        o_rank* R = &s->r[r];
        R->mype = r;
        R->mypey = r % p.npey; /* reference src/topology.cpp:9-14 */
        R->mypez = r / p.npey;
        for (int f = 0; f < NF3; ++f) R->f[f] = calloc((size_t)n, sizeof(double));
        const int len[NM1] = {s->d.nx, s->d.nx, s->d.nx, s->d.nx, s->d.ny, s->d.ny, s->d.ny,
                              s->d.ny, s->d.nz, s->d.nz, s->d.nz, s->d.nz, s->d.nz, s->d.nz};
        for (int m = 0; m < NM1; ++m) R->m[m] = calloc((size_t)len[m], sizeof(double));
        R->rom = calloc((size_t)s->d.nz, sizeof(double));
        R->ttm = calloc((size_t)s->d.nz, sizeof(double));
        R->fconm = calloc((size_t)s->d.nz, sizeof(double));
        R->correct = calloc((size_t)s->d.nz, sizeof(double));
    }
    s->tb = 1.0;
    s->tfac = 1.0;      /* DATA TFAC/1D0/, golden/fortran/3dmhdsub.f:1259 */
    s->ttm_max = -1e99; /* DATA TTM_MAX/-1D99/, :1258 */
    return s;
}

void o_free(o_sim* s) {
    if (!s) return;
    for (int r = 0; r < s->nranks; ++r) {
        for (int f = 0; f < NF3; ++f) free(s->r[r].f[f]);
        for (int m = 0; m < NM1; ++m) free(s->r[r].m[m]);
        free(s->r[r].rom);
        free(s->r[r].ttm);
        free(s->r[r].fconm);
        free(s->r[r].correct);
    }
    free(s->r);
    free(s);
}

void o_dims(const o_sim* s, int dims[10]) {
    dims[0] = s->d.nx; dims[1] = s->d.ny; dims[2] = s->d.nz;
    dims[3] = s->p.npey; dims[4] = s->p.npez; dims[5] = s->nranks;
    dims[6] = s->p.npx; dims[7] = s->p.npy; dims[8] = s->p.npz; dims[9] = s->p.lmag;
}
double* o_field(o_sim* s, int rank, int fid) {
    if (rank < 0 || rank >= s->nranks || fid < 0 || fid >= NF3) return NULL;
    return s->r[rank].f[fid];
}
void o_get_relax(const o_sim* s, double v[2]) {
    v[0] = s->tfac;
    v[1] = s->ttm_max;
}
double* o_correct(o_sim* s, int rank) {
    if (rank < 0 || rank >= s->nranks) return NULL;
    return s->r[rank].correct;
}
double* o_rom(o_sim* s, int rank) {
    if (rank < 0 || rank >= s->nranks) return NULL;
    return s->r[rank].rom;
}
double* o_metric(o_sim* s, int rank, int mid) {
    if (rank < 0 || rank >= s->nranks || mid < 0 || mid >= NM1) return NULL;
    return s->r[rank].m[mid];
}
void o_get_scalars(const o_sim* s, double v[7]) {
    v[0] = s->dt; v[1] = s->timc; v[2] = s->timt; v[3] = s->timi; v[4] = s->umach;
    v[5] = s->tb; v[6] = s->nit;
}
void o_set_scalars(o_sim* s, const double v[7]) {
    s->dt = v[0]; s->timc = v[1]; s->timt = v[2]; s->timi = v[3]; s->umach = v[4];
    s->tb = v[5]; s->nit = (int)v[6];
}
void o_get_derived(const o_sim* s, double o[27]) {
    const o_derived* d = &s->d;
    const double v[27] = {d->c13, d->c23, d->c43, d->gam1, d->gam2, d->gam3, d->zeta1,
                          d->zeta2, d->hx, d->h2x, d->hy, d->h2y, d->hz, d->h2z, d->repr,
                          d->ore, d->cv, d->ocv, d->rkapst, d->orm, d->obeta, d->omx, d->omz,
                          s->p.pr, s->p.gamma, s->p.ampt, s->p.sf};
    memcpy(o, v, sizeof v);
}
void o_mkgrid(const o_sim* s, double sc, int ic, double out[3]) {
    mkgrid(&s->p, sc, ic, &out[0], &out[1], &out[2]);
}
int o_static_profile(const o_sim* s, double* t, double* r) {
    return static_profile(&s->p, s->p.npz, t, r);
}

int o_lrem_profile(const o_sim* s, double* t, double* r, double* rk, double* drk, double bnd[3]) {
    return lrem_profile(&s->p, &s->d, s->p.npz, t, r, rk, drk, bnd);
}

int o_lrem_background(double z, double out[2]) { return lrem_background(z, &out[0], &out[1]); }

#include "r3d_oracle_step.inc"

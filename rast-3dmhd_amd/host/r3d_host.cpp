// r3d_host.cpp - host-side setup of a run (C ABI in include/r3d_host.h).
//
// Runs once per run, on the host: parameters, derived constants, global grid
// metrics, the static stratification and the seeded temperature perturbation.
// Built with -ffp-contract=off so every value rounds exactly as in the
// reference CPU build (reference makefile:104); the arithmetic of each
// function follows the reference routine it cites.
#include "r3d_host.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <sys/mman.h>

namespace {

thread_local std::string g_err;

int fail(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

// Fortran single-precision literal widened to double (reference
// include/numcompat.hpp:17-21).
inline double single(float x) { return static_cast<double>(x); }

// ---------------------------------------------------------------- params ----
// Defaults: reference include/params.hpp:24-120.
struct Params {
    int npx = 504, npy = 504, npz = 2048, npey = 8, npez = 2;
    int ix = 2, iy = 2, ilap = 2, ngrid = 1, id = 0, ncol = 0, nrow = 0;
    int ixcon = 0, iycon = 0, izcon = 0, itcon = 1, ibcon = 0;
    int lrot = 0, lmag = 0, lpot = 0, lrem = 0, lshr = 0;
    int ncase = 1, ncasep = 1, ntotal = 1000, nstep0 = 100, nstart = 0, ntube = 0;
    int gpu = 0, backend_gpu = 0, seed = -62659;
    double re = 2.0, pr = single(5.0e-2f), theta = 2.5e-1, grav = 6.25e-1, ry = 0.0;
    double ang = 0.0, rm = 0.0, beta = 0.0, pzp = 0.0, sigma = 0.0, polys = 0.0;
    double tp = 2.0, tc = 0.0, xp = 10.0, yp = 10.0, zp = 0.0;
    double xmax = 20.0, ymax = 20.0, zmax = 40.0, ampt = single(5.0e-3f), ampb = 0.0;
    double bfh = 0.0, bzp = 0.0, qfh = 0.0, gamma = single(5.0f / 3.0f), hh = 1.0;
    double dh = 0.0, dv = 0.0, tstart = 0.0, toff = 0.0, rlax = 0.0;
    double xx1 = -7.0, xx2 = 7.0, yy1 = -7.0, yy2 = 7.0, zz1 = -2.0, zz2 = 1.0e-2;
    double xa = 0.5, xb = 0.48, xc = 28.0, xd = 0.025;
    double ya = 0.5, yb = 0.48, yc = 28.0, yd = 0.025;
    double za = 0.65, zb = 1.0, zc = 5.0, zd = 0.2;
    double xcent = 0.0, zcent = 0.0, rmax = 0.0, cmt = 0.0, a_tube = 0.0;
    double lambda = 0.0, vz0 = 0.0, sff = single(0.315f);
    // restart (--nstart N): directory of the gstate.step.<rank>.<N> dumps
    std::string dumpdir = ".";
};

struct Flag {
    const char* name;
    char kind;  // 'i' int, 'd' double, 'b' bool
    void* ptr;
};

std::vector<Flag> flag_table(Params& p) {
    return {
        {"--npx", 'i', &p.npx}, {"--npy", 'i', &p.npy}, {"--npz", 'i', &p.npz},
        {"--npey", 'i', &p.npey}, {"--npez", 'i', &p.npez}, {"--ix", 'i', &p.ix},
        {"--iy", 'i', &p.iy}, {"--ilap", 'i', &p.ilap}, {"--ngrid", 'i', &p.ngrid},
        {"--id", 'i', &p.id}, {"--ncol", 'i', &p.ncol}, {"--nrow", 'i', &p.nrow},
        {"--ixcon", 'i', &p.ixcon}, {"--iycon", 'i', &p.iycon}, {"--izcon", 'i', &p.izcon},
        {"--itcon", 'i', &p.itcon}, {"--ibcon", 'i', &p.ibcon}, {"--ncase", 'i', &p.ncase},
        {"--ncasep", 'i', &p.ncasep}, {"--ntotal", 'i', &p.ntotal},
        {"--nstep0", 'i', &p.nstep0}, {"--nstart", 'i', &p.nstart}, {"--ntube", 'i', &p.ntube},
        {"--gpu", 'i', &p.gpu}, {"--seed", 'i', &p.seed},
        {"--re", 'd', &p.re}, {"--pr", 'd', &p.pr}, {"--theta", 'd', &p.theta},
        {"--grav", 'd', &p.grav}, {"--ry", 'd', &p.ry}, {"--ang", 'd', &p.ang},
        {"--rm", 'd', &p.rm}, {"--beta", 'd', &p.beta}, {"--pzp", 'd', &p.pzp},
        {"--sigma", 'd', &p.sigma}, {"--polys", 'd', &p.polys}, {"--tp", 'd', &p.tp},
        {"--tc", 'd', &p.tc}, {"--xp", 'd', &p.xp}, {"--yp", 'd', &p.yp}, {"--zp", 'd', &p.zp},
        {"--xmax", 'd', &p.xmax}, {"--ymax", 'd', &p.ymax}, {"--zmax", 'd', &p.zmax},
        {"--ampt", 'd', &p.ampt}, {"--ampb", 'd', &p.ampb}, {"--bfh", 'd', &p.bfh},
        {"--bzp", 'd', &p.bzp}, {"--qfh", 'd', &p.qfh}, {"--gamma", 'd', &p.gamma},
        {"--hh", 'd', &p.hh}, {"--dh", 'd', &p.dh}, {"--dv", 'd', &p.dv},
        {"--tstart", 'd', &p.tstart}, {"--toff", 'd', &p.toff}, {"--rlax", 'd', &p.rlax},
        {"--xx1", 'd', &p.xx1}, {"--xx2", 'd', &p.xx2}, {"--yy1", 'd', &p.yy1},
        {"--yy2", 'd', &p.yy2}, {"--zz1", 'd', &p.zz1}, {"--zz2", 'd', &p.zz2},
        {"--xa", 'd', &p.xa}, {"--xb", 'd', &p.xb}, {"--xc", 'd', &p.xc}, {"--xd", 'd', &p.xd},
        {"--ya", 'd', &p.ya}, {"--yb", 'd', &p.yb}, {"--yc", 'd', &p.yc}, {"--yd", 'd', &p.yd},
        {"--za", 'd', &p.za}, {"--zb", 'd', &p.zb}, {"--zc", 'd', &p.zc}, {"--zd", 'd', &p.zd},
        {"--xcent", 'd', &p.xcent}, {"--zcent", 'd', &p.zcent}, {"--rmax", 'd', &p.rmax},
        {"--cmt", 'd', &p.cmt}, {"--a", 'd', &p.a_tube}, {"--lambda", 'd', &p.lambda},
        {"--vz0", 'd', &p.vz0}, {"--sff", 'd', &p.sff},
        {"--lrot", 'b', &p.lrot}, {"--lmag", 'b', &p.lmag}, {"--lpot", 'b', &p.lpot},
        {"--lrem", 'b', &p.lrem}, {"--lshr", 'b', &p.lshr},
    };
}

// "--key value" and "--key=value"; atoi/atof like the reference.
int parse(int argc, const char* const* argv, Params& p) {
    const std::vector<Flag> flags = flag_table(p);
    for (int a = 0; a < argc; ++a) {
        const std::string arg = argv[a];
        const auto eq = arg.find('=');
        const std::string key = arg.substr(0, eq);
        auto value = [&](const char* name) -> const char* {
            if (eq != std::string::npos) return argv[a] + eq + 1;
            if (a + 1 >= argc) {
                fail("missing value for %s", name);
                return nullptr;
            }
            return argv[++a];
        };
        if (key == "--dumpdir") {
            const char* v = value("--dumpdir");
            if (!v) return -1;
            p.dumpdir = v;
            continue;
        }
        if (key == "--backend") {
            const char* v = value("--backend");
            if (!v) return -1;
            if (!strcmp(v, "cpu")) p.backend_gpu = 0;
            else if (!strcmp(v, "gpu")) p.backend_gpu = 1;
            else return fail("--backend must be 'cpu' or 'gpu'");
            continue;
        }
        const Flag* f = nullptr;
        for (const Flag& q : flags)
            if (key == q.name) f = &q;
        if (!f) return fail("unknown option '%s'", argv[a]);
        const char* v = value(f->name);
        if (!v) return -1;
        if (f->kind == 'i') *static_cast<int*>(f->ptr) = atoi(v);
        else if (f->kind == 'd') *static_cast<double*>(f->ptr) = atof(v);
        else *static_cast<int*>(f->ptr) = !strcmp(v, "1") || !strcmp(v, "true") ||
                                          !strcmp(v, "yes") || !strcmp(v, "on");
    }
    return 0;
}

// Derived constants: reference src/params.cpp:14-66.
struct Derived {
    double c13, c23, c43, gam1, gam2, gam3, zeta1, zeta2;
    double hx, h2x, hy, h2y, hz, h2z;
    double repr, ore, cv, ocv, rkapst, orm, obeta, omx, omz;
};

Derived derive(const Params& p) {
    Derived d;
    d.c13 = single(1.0f / 3.0f);
    d.c23 = 2.0 * d.c13;
    d.c43 = 4.0 * d.c13;
    d.gam1 = single(8.0f / 15.0f);
    d.gam2 = single(5.0f / 12.0f);
    d.gam3 = single(3.0f / 4.0f);
    d.zeta1 = single(-17.0f / 60.0f);
    d.zeta2 = single(-5.0f / 12.0f);
    const float fx = (float)(p.npx - 1), fy = (float)(p.npy - 1), fz = (float)(p.npz - 1);
    d.hx = single(0.5f * fx);
    d.h2x = single(fx * fx);
    d.hy = single(0.5f * fy);
    d.h2y = single(fy * fy);
    d.hz = single(0.5f * fz);
    d.h2z = single(fz * fz);
    d.repr = p.re * p.pr;
    d.ore = 1.0 / p.re;
    d.cv = 1.0 / (p.gamma - 1.0);
    d.ocv = 1.0 / d.cv;
    d.rkapst = (p.grav == 0.0) ? 1.0 : (p.polys + 1.0) * p.theta / p.grav;
    d.orm = p.lmag ? 1.0 / p.rm : 0.0;
    d.obeta = p.lmag ? 2.0 / p.beta : 0.0;
    d.omx = p.lrot ? sin(p.ang) / p.ry : 0.0;
    d.omz = p.lrot ? cos(p.ang) / p.ry : 0.0;
    return d;
}

// ------------------------------------------------------------------ grid ----
// Coordinate map s -> (x(s), dx/ds... as ds/dx, d2s/dx2) of one direction:
// reference src/grid.cpp:28-146 (NGRID 0 uniform, 1 arctan, 2 two-arctan).
struct MapOut {
    double x, dsdx, d2s;
};

MapOut grid_map(const Params& p, double s, int dir) {
    const double pi = 2.0 * asin(1.0);
    const double smax = dir == 1 ? p.xmax : dir == 2 ? p.ymax : p.zmax;
    MapOut o{};
    if (p.ngrid == 0) {
        o.x = s * smax;
        o.dsdx = 1.0 / smax;
        o.d2s = 0.0;
        return o;
    }
    if (p.ngrid == 1) {
        const double s1 = dir == 1 ? p.xx1 : dir == 2 ? p.yy1 : p.zz1;
        const double s2 = dir == 1 ? p.xx2 : dir == 2 ? p.yy2 : p.zz2;
        const double a1 = (s2 - s1) / smax;
        const double a3 = atan(s1);
        const double a2 = atan(s2) - a3;
        o.x = s == 0.0 ? 0.0 : s == 1.0 ? smax : (tan(a2 * s + a3) - s1) / a1;
        o.dsdx = a1 / a2 / (1.0 + (s1 + a1 * o.x) * (s1 + a1 * o.x));
        const double t = s1 + a1 * o.x;
        const double q = 1.0 + t * t;
        o.d2s = -2.0 * a1 * a1 * t / a2 / (q * q);
        return o;
    }
    const double a = dir == 1 ? p.xa : dir == 2 ? p.ya : p.za;
    const double b = dir == 1 ? p.xb : dir == 2 ? p.yb : p.zb;
    const double c = dir == 1 ? p.xc : dir == 2 ? p.yc : p.zc;
    const double dp = dir == 1 ? p.xd : dir == 2 ? p.yd : p.zd;
    if (c <= 1.0e-9) {
        o.x = s * smax;
        o.dsdx = 1.0 / smax;
        o.d2s = 0.0;
        return o;
    }
    const double sh =
        (1.0 - dp) * pi / (dp * atan((a - b) * c) + 2.0 * atan(b * c) - dp * atan((a + b) * c));
    const double sk =
        2.0 * c * pi * smax /
        (2.0 * c * pi +
         sh * (2.0 * c *
                   (atan((a - b) * c) * (a - b) - atan((a + b) * c) * (a + b) +
                    atan((a - b - 1.0) * c) + atan((a + b - 1.0) * c) * (a + b - 1.0) +
                    atan((1.0 - a + b) * c) * (a - b)) -
               log(1.0 + (a - b) * (a - b) * c * c) + log(1.0 + (a + b) * (a + b) * c * c) -
               log(1.0 + (a + b - 1.0) * (a + b - 1.0) * c * c) +
               log(1.0 + (1.0 - a + b) * (1.0 - a + b) * c * c)));
    o.dsdx = 1.0 / (sk * (1.0 + sh / pi * (atan(c * (s - a - b)) + atan(c * (a - s - b)))));
    o.d2s = -o.dsdx * o.dsdx * o.dsdx * c * sh * sk / pi *
            (1.0 / (1.0 + c * c * (a + b - s) * (a + b - s)) -
             1.0 / (1.0 + c * c * (a - b - s) * (a - b - s)));
    o.x = sk / (2.0 * c * pi) *
          (2.0 * c * pi * s +
           sh * (2.0 * c *
                     (atan((a - b) * c) * (a - b) - atan((a + b) * c) * (a + b) +
                      atan((a - b - s) * c) * s + atan((a + b - s) * c) * (a + b - s) +
                      atan((s - a + b) * c) * (a - b)) -
                 log(1.0 + (a - b) * (a - b) * c * c) + log(1.0 + (a + b) * (a + b) * c * c) -
                 log(1.0 + (a + b - s) * (a + b - s) * c * c) +
                 log(1.0 + (s - a + b) * (s - a + b) * c * c)));
    if (s == 0.0) o.x = 0.0;
    if (s == 1.0) o.x = smax;
    return o;
}

// Computational coordinate of global index m (0..n+1) along a direction with
// n points: interior m in 1..n maps to (m-1)/(n-1) with exact end points
// (reference src/grid.cpp:152-156, 176-180, 230-234).  spacing = f4(1/(n-1)).
double comp_coord(int m, int n, double spacing) {
    if (m <= 1) return 0.0;
    if (m >= n) return 1.0;
    return (m - 1) * spacing;
}

// ---------------------------------------------------------------- ran2 ----
// Numerical Recipes ran2 with a 97-entry shuffle table: reference
// src/ran2.cpp:17-37.
struct Ran2 {
    int idum, iy, tab[97];
    explicit Ran2(int seed) : idum(seed), iy(0) { memset(tab, 0, sizeof tab); }
    double next() {
        constexpr int M = 714025, A = 1366, C = 150889;
        const double scale = single(1.0f / 714025.0f);
        if (idum < 0) {
            idum = (C - idum) % M;
            for (int q = 0; q < 97; ++q) {
                idum = (A * idum + C) % M;
                tab[q] = idum;
            }
            idum = (A * idum + C) % M;
            iy = idum;
        }
        const int slot = (97 * iy) / M;
        iy = tab[slot];
        const double v = (double)iy * scale;
        idum = (A * idum + C) % M;
        tab[slot] = idum;
        return v;
    }
};

// ------------------------------------------------------- static profile ----
// T(z), rho(z) of the polytrope: Bulirsch-Stoer integration of
// dT/dzz = theta/kappa/(dz/dzz), dln p/dzz = grav/T/(dz/dzz) at fixed step
// 1/(npz-1): reference src/ode.cpp:23-174.
class Profile {
   public:
    explicit Profile(const Params& p) : p_(p) {}

    int build(int npz, std::vector<double>& T, std::vector<double>& R) {
        T.assign(npz, 0.0);
        R.assign(npz, 0.0);
        const double h = single(1.0f / (float)(npz - 1));
        double y[2] = {0.0, 0.0}, z = 0.0;
        T[0] = 1.0;
        R[0] = 1.0;
        for (int k = 1; k < npz; ++k) {
            double dy[2];
            rhs(z, y, dy);
            if (step(y, dy, z, h) != h) return fail("static profile: step shrank at k=%d", k);
            T[k] = 1.0 + y[0];
            R[k] = exp(y[1]) / (1.0 + y[0]);
        }
        return 0;
    }

   private:
    static constexpr int kUse = 7, kMax = 11;
    const Params& p_;
    double xs_[kMax];

    void rhs(double zz, const double* y, double* dy) const {
        const MapOut m = grid_map(p_, zz, 3);
        const double rkapst = (p_.grav == 0.0) ? 1.0 : (p_.polys + 1.0) * p_.theta / p_.grav;
        const double rkap = p_.pzp == 0.0
                                ? 1.0
                                : 1.0 + (rkapst - 1.0) / 2.0 * (1.0 + tanh((m.x - p_.pzp) / p_.sigma));
        dy[0] = p_.theta / rkap / m.dsdx;
        dy[1] = p_.grav / (1.0 + y[0]) / m.dsdx;
    }

    // modified midpoint rule over one macro step
    void midpoint(const double* y, const double* dy, double x0, double H, int n,
                  double* out) const {
        const double h = H / n;
        double a[2], b[2], d[2];
        for (int v = 0; v < 2; ++v) {
            a[v] = y[v];
            b[v] = y[v] + h * dy[v];
        }
        double x = x0 + h;
        rhs(x, b, d);
        const double h2 = 2.0 * h;
        for (int s = 1; s < n; ++s) {
            for (int v = 0; v < 2; ++v) {
                const double t = a[v] + h2 * d[v];
                a[v] = b[v];
                b[v] = t;
            }
            x += h;
            rhs(x, b, d);
        }
        for (int v = 0; v < 2; ++v) out[v] = 0.5 * (a[v] + b[v] + h * d[v]);
    }

    // diagonal rational extrapolation to zero step
    void extrapolate(int i, double xe, const double* ye, double* yz, double tab[2][kUse],
                     double* err) {
        xs_[i - 1] = xe;
        if (i == 1) {
            for (int v = 0; v < 2; ++v) yz[v] = tab[v][0] = err[v] = ye[v];
            return;
        }
        const int m = i < kUse ? i : kUse;
        double f[kUse + 1];
        for (int q = 1; q < m; ++q) f[q] = xs_[i - q - 1] / xe;
        for (int v = 0; v < 2; ++v) {
            double acc = ye[v], prev = tab[v][0], cc = acc, dd = 0.0;
            tab[v][0] = acc;
            for (int q = 1; q < m; ++q) {
                const double b1 = f[q] * prev;
                double b = b1 - cc;
                if (b != 0.0) {
                    b = (cc - prev) / b;
                    dd = cc * b;
                    cc = b1 * b;
                } else {
                    dd = prev;
                }
                prev = tab[v][q];
                tab[v][q] = dd;
                acc += dd;
            }
            err[v] = dd;
            yz[v] = acc;
        }
    }

    double step(double* y, const double* dy, double& x, double htry) {
        static const int seq[kMax] = {2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96};
        double h = htry;
        const double x0 = x;
        const double y0[2] = {y[0], y[1]}, d0[2] = {dy[0], dy[1]};
        double ys[2], err[2], tab[2][kUse];
        for (;;) {
            for (int i = 1; i <= kMax; ++i) {
                midpoint(y0, d0, x0, h, seq[i - 1], ys);
                const double xe = (h / seq[i - 1]) * (h / seq[i - 1]);
                extrapolate(i, xe, ys, y, tab, err);
                double emax = 0.0;
                for (int v = 0; v < 2; ++v) {
                    const double e = std::fabs(err[v] / y[v]);
                    if (e > emax) emax = e;
                }
                if (emax / 1.0e-12 < 1.0) {
                    x += h;
                    return h;
                }
            }
            h = 0.25 * h / std::pow(2.0, double((kMax - kUse) / 2));
            if (x + h == x) return -1.0;
        }
    }
};

// ------------------------------------------------------- LREM profile ----
// M. Rempel's stratification (golden/fortran/3dmhdsub.f:437-490, STATIC;
// 4442-4650, GETBACKGROUND / DERIVS1 / ODEINT / RKQC / RK4;
// golden/fortran/3dmhdset.f:154-170, KAPPA).  The reference C++ refuses LREM
// (src/static_ic.cpp:27).  The Fortran is built without -fdefault-real-8, so
// its single-precision literals are widened floats here.  T(z) and
// p = rho*T are integrated from z = 2 (T = p = 1) to each plane's depth by
// Runge-Kutta with step doubling (tolerance 1e-8): dT/dz = min-like switch
// between the radiative gradient 0.4/kappa and the adiabatic 0.41,
// dp/dz = p/T, kappa = rho + (1.25 - rho) exp(-5 z).
class LremProfile {
   public:
    LremProfile(const Params& p, const Derived& d) : p_(p), d_(d) {}

    // T, R, the conductivity RK = kappa*8.07*theta*repr and its derivative
    // DRK per plane; bnd = {TU, DZTU, DZTB}
    int build(int npz, std::vector<double>& T, std::vector<double>& R, std::vector<double>& RK,
              std::vector<double>& DRK, double bnd[3]) const {
        T.assign(npz, 0.0);
        R.assign(npz, 0.0);
        RK.assign(npz, 0.0);
        DRK.assign(npz, 0.0);
        for (int k = 0; k < npz; ++k) {
            const MapOut m = grid_map(p_, single((float)k / (float)(npz - 1)), 3);
            double t = 0.0, rho = 0.0;
            if (background(m.x, t, rho)) return -1;
            T[k] = t;
            R[k] = rho;
            RK[k] = kappa(m.x, rho);
        }
        const double c807 = single(8.07f);
        for (int k = 0; k < npz; ++k) RK[k] = RK[k] * c807 * p_.theta * d_.repr;
        for (int k = 1; k < npz - 1; ++k) {
            const MapOut m = grid_map(p_, single((float)k / (float)(npz - 1)), 3);
            DRK[k] = (RK[k + 1] - RK[k - 1]) * d_.hz * m.dsdx;
        }
        const MapOut m0 = grid_map(p_, 0.0, 3), m1 = grid_map(p_, 1.0, 3);
        DRK[0] = (-3.0 * RK[0] + 4.0 * RK[1] - RK[2]) * d_.hz * m0.dsdx;
        DRK[npz - 1] = (3.0 * RK[npz - 1] - 4.0 * RK[npz - 2] + RK[npz - 3]) * d_.hz * m1.dsdx;
        const double f43 = single(4.0f / 3.0f), f13 = single(1.0f / 3.0f);
        bnd[0] = T[0];
        bnd[1] = T[0] - f43 * T[1] + f13 * T[2];
        bnd[2] = T[npz - 1] - f43 * T[npz - 2] + f13 * T[npz - 3];
        return 0;
    }

   private:
    const Params& p_;
    const Derived& d_;

    static double kappa(double z, double rho) {
        double krad = rho;
        krad = krad + (1.25 - krad) * std::exp(-(5.0 * z));
        return krad;
    }

    static void rhs(double x, const double* y, double* dy) {
        const double nad = single(0.41f), os = 1.0;
        const double rho = y[1] / y[0];
        const double nrad = single(0.4f) / kappa(x, rho);
        dy[0] = nrad <= os * nad ? nrad : nad;
        dy[1] = y[1] / y[0];
    }

    static void rk4(const double* y, const double* dy, double x, double h, double* out) {
        double yt[2], dyt[2], dym[2];
        const double hh = h * 0.5, h6 = h / 6.0, xh = x + hh;
        for (int i = 0; i < 2; ++i) yt[i] = y[i] + hh * dy[i];
        rhs(xh, yt, dyt);
        for (int i = 0; i < 2; ++i) yt[i] = y[i] + hh * dyt[i];
        rhs(xh, yt, dym);
        for (int i = 0; i < 2; ++i) {
            yt[i] = y[i] + h * dym[i];
            dym[i] = dyt[i] + dym[i];
        }
        rhs(x + h, yt, dyt);
        for (int i = 0; i < 2; ++i) out[i] = y[i] + h6 * (dy[i] + dyt[i] + 2.0 * dym[i]);
    }

    // one accepted step with error control by step doubling
    static int step(double* y, double* dy, double& x, double htry, double eps,
                    const double* yscal, double& hnext) {
        const double fcor = 1.0 / 15.0, safety = single(0.9f), errcon = 6.0e-4;
        const double pgrow = single(-0.20f), pshrnk = -0.25;
        const double x0 = x, y0[2] = {y[0], y[1]}, d0[2] = {dy[0], dy[1]};
        double err[2], h = htry;
        for (;;) {
            const double hh = 0.5 * h;
            rk4(y0, d0, x0, hh, err);
            x = x0 + hh;
            rhs(x, err, dy);
            rk4(err, dy, x, hh, y);
            x = x0 + h;
            if (x == x0) return fail("LREM background: step size not significant");
            rk4(y0, d0, x0, h, err);
            double emax = 0.0;
            for (int i = 0; i < 2; ++i) {
                err[i] = y[i] - err[i];
                const double e = std::fabs(err[i] / yscal[i]);
                emax = emax > e ? emax : e;
            }
            emax = emax / eps;
            if (emax > 1.0) {
                h = safety * h * std::pow(emax, pshrnk);
                continue;
            }
            hnext = emax > errcon ? safety * h * std::pow(emax, pgrow) : 4.0 * h;
            break;
        }
        for (int i = 0; i < 2; ++i) y[i] = y[i] + err[i] * fcor;
        return 0;
    }

    static int background(double z, double& t, double& rho) {
        const double x1 = 2.0, x2 = z, eps = 1.0e-8, hmin = 1.0e-50, tiny = single(1.0e-30f);
        double x = x1, h = std::copysign(1.0e-4, x2 - x1);
        double y[2] = {1.0, 1.0}, dy[2] = {0.0, 0.0}, ys[2];
        if (x2 != x1) {  // a zero-length integration: the start values (the Fortran pauses)
            int n = 0;
            for (;; ++n) {
                if (n == 1000) return fail("LREM background: too many steps");
                rhs(x, y, dy);
                for (int i = 0; i < 2; ++i) ys[i] = std::fabs(y[i]) + std::fabs(h * dy[i]) + tiny;
                if ((x + h - x2) * (x + h - x1) > 0.0) h = x2 - x;
                double hnext = 0.0;
                if (step(y, dy, x, h, eps, ys, hnext)) return -1;
                if ((x - x2) * (x2 - x1) >= 0.0) break;
                if (std::fabs(hnext) < hmin)
                    return fail("LREM background: step size smaller than minimum");
                h = hnext;
            }
        }
        t = y[0];
        rho = y[1] / y[0];
        return 0;
    }
};

}  // namespace

// Storage of the large global fields: not zero-initialised on resize (every
// cell is then written by the builder, in parallel, so that the first touch
// of the pages is spread over threads), 2 MiB aligned and advised for
// transparent huge pages.  Zero-filling a 504x504x2048 state on one thread
// spent 24 of 30 s in page faults.
template <class T>
struct BigAlloc {
    using value_type = T;
    BigAlloc() = default;
    template <class U>
    BigAlloc(const BigAlloc<U>&) {}
    T* allocate(size_t n) {
        constexpr size_t kAlign = 2u << 20;
        const size_t bytes = (n * sizeof(T) + kAlign - 1) / kAlign * kAlign;
        void* p = std::aligned_alloc(kAlign, bytes);
        if (!p) throw std::bad_alloc();
        madvise(p, bytes, MADV_HUGEPAGE);  // advisory; ignored where unsupported
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { std::free(p); }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        if constexpr (sizeof...(A) > 0) ::new ((void*)p) U(std::forward<A>(a)...);
        else ::new ((void*)p) U;  // default-initialise: no write, no page touch
    }
    template <class U>
    bool operator==(const BigAlloc<U>&) const { return true; }
    template <class U>
    bool operator!=(const BigAlloc<U>&) const { return false; }
};
using BigVec = std::vector<double, BigAlloc<double>>;

// fn(lo, hi) over [0, n) in contiguous ranges on up to 16 threads
template <class F>
void parallel_ranges(long n, F fn) {
    const long nt = std::max(1L, std::min({16L, (long)std::thread::hardware_concurrency(), n}));
    std::vector<std::thread> th;
    for (long t = 1; t < nt; ++t) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
    fn(0L, n / nt);
    for (auto& x : th) x.join();
}

struct r3d_host {
    Params p;
    Derived d;
    double tb = 1.0;
    double tu = 1.0, dztu = 0.0, dztb = 0.0;  // LREM boundary values of STATIC
    std::vector<double> m1[14];  // global metrics, gstate order
    BigVec f[8];                 // global initial primaries
    r3d_hip_scalars sc{};        // starting scalars (from the dumps on a restart)
};

namespace {

// Global metrics of the whole domain: interior entries are those of the
// reference's rank that owns the point; ghost entries follow the first/last
// rank branches (reference src/grid.cpp:148-284).
void build_metrics(r3d_host* h) {
    const Params& p = h->p;
    const int X = p.npx + 2, Y = p.npy + 2, Z = p.npz + 2;
    for (int q = 0; q < 14; ++q) h->m1[q].assign(q < 4 ? X : q < 8 ? Y : Z, 0.0);
    const double orx = single(1.0f / (float)(p.npx - 1));
    for (int i = 1; i <= p.npx; ++i) {  // x ghosts stay 0 (src/grid.cpp:148-167)
        const MapOut m = grid_map(p, comp_coord(i, p.npx, orx), 1);
        h->m1[0][i] = m.x;
        h->m1[1][i] = m.dsdx;
        h->m1[2][i] = m.d2s;
        h->m1[3][i] = orx * (1.0 / m.dsdx);
    }
    const double ory = single(1.0f / (float)(p.npy - 1));
    for (int j = 0; j < Y; ++j) {
        const MapOut m = grid_map(p, comp_coord(j, p.npy, ory), 2);
        h->m1[4][j] = m.x;
        h->m1[5][j] = m.dsdx;
        h->m1[6][j] = m.d2s;
        h->m1[7][j] = ory * (1.0 / m.dsdx);
    }
    const double orz = single(1.0f / (float)(p.npz - 1));
    for (int k = 0; k < Z; ++k) {
        const MapOut m = grid_map(p, comp_coord(k, p.npz, orz), 3);
        h->m1[8][k] = m.x;
        h->m1[9][k] = m.dsdx;
        h->m1[10][k] = m.d2s;
        h->m1[11][k] = orz * (1.0 / m.dsdx);
    }
    for (int k = 0; k < Z; ++k) {  // conductivity, src/main.cpp:135-146
        if (p.pzp == 0.0) {
            h->m1[12][k] = 1.0;
            h->m1[13][k] = 0.0;
        } else {
            const double xe = h->m1[8][k];
            h->m1[12][k] = 1.0 + (h->d.rkapst - 1.0) / 2.0 * (1.0 + tanh((xe - p.pzp) / p.sigma));
            h->m1[13][k] = (h->d.rkapst - 1.0) / 2.0 / p.sigma / cosh((xe - p.pzp) / p.sigma) /
                           cosh((xe - p.pzp) / p.sigma);
        }
    }
}

// The static profile T, rho per plane: the polytrope, or with LREM M.
// Rempel's stratification, which also sets the conductivity RKAPA, DKAPA of
// every plane (golden/fortran/3dmhd.f:267-269; ghost planes take the end
// values, golden/fortran/3dmhdsub.f:530-560) and the boundary values TU,
// DZTU, DZTB.  tb = T at the bottom either way.
int static_profile(r3d_host* h, std::vector<double>& T, std::vector<double>& R) {
    const Params& p = h->p;
    if (!p.lrem) {
        Profile prof(p);
        if (prof.build(p.npz, T, R)) return -1;
    } else {
        std::vector<double> RK, DRK;
        double bnd[3];
        LremProfile prof(p, h->d);
        if (prof.build(p.npz, T, R, RK, DRK, bnd)) return -1;
        for (int k = 0; k < p.npz + 2; ++k) {
            const int src = k == 0 ? 0 : k <= p.npz ? k - 1 : p.npz - 1;
            h->m1[12][k] = RK[src];
            h->m1[13][k] = DRK[src];
        }
        h->tu = bnd[0];
        h->dztu = bnd[1];
        h->dztb = bnd[2];
    }
    h->tb = T[p.npz - 1];
    return 0;
}

// Global initial state of STATIC (reference src/static_ic.cpp:24-161) after
// the first halo exchange: uniform planes of the profile, the optional
// magnetic layer, the ran2 perturbation in global (k, j, i) order, then the
// periodic x/y ghost images.
int build_state(r3d_host* h) {
    const Params& p = h->p;
    const int X = p.npx + 2, Y = p.npy + 2, Z = p.npz + 2;
    const size_t plane = (size_t)X * Y, n = plane * Z;
    const int nf = p.lmag ? 8 : 5;  // B fields only exist with LMAG
    for (int q = 0; q < nf; ++q) h->f[q].resize(n);  // every cell is written below
    std::vector<double> T, R;
    if (static_profile(h, T, R)) return -1;
    const double rb = R[p.npz - 1];
    double* tt = h->f[R3D_TT].data();
    double* ro = h->f[R3D_RO].data();
    auto at = [&](int i, int j, int k) { return (size_t)k * plane + (size_t)j * X + i; };
    // src/static_ic.cpp:94-108 (magnetic layer)
    const double cln = p.lmag ? single(-4.0f * logf(2.0f)) / (p.bfh * p.bfh) : 0.0;
    // uniform planes of the profile (and the layer), one z plane per task
    parallel_ranges(Z, [&](long ka, long kb) {
        for (int k = (int)ka; k < (int)kb; ++k) {
            const double tv = k == 0 ? T[0] : k <= p.npz ? T[k - 1] : h->tb;
            const double rv = k == 0 ? R[0] : k <= p.npz ? R[k - 1] : rb;
            std::fill(tt + k * plane, tt + (k + 1) * plane, tv);
            std::fill(ro + k * plane, ro + (k + 1) * plane, rv);
            for (int q : {R3D_RU, R3D_RV, R3D_RW})
                std::fill(h->f[q].data() + k * plane, h->f[q].data() + (k + 1) * plane, 0.0);
            if (!p.lmag) continue;
            double* bx = h->f[R3D_BX].data();
            for (int q : {R3D_BY, R3D_BZ})
                std::fill(h->f[q].data() + k * plane, h->f[q].data() + (k + 1) * plane, 0.0);
            if (k < 1 || k > p.npz) {
                std::fill(bx + k * plane, bx + (k + 1) * plane, 0.0);
                continue;
            }
            const double zk = h->m1[8][k];
            const double bfac = p.ampb * std::exp(cln * (zk - p.bzp) * (zk - p.bzp));
            std::fill(bx + k * plane, bx + (k + 1) * plane, bfac);
            for (int j = 1; j <= p.npy; ++j)
                for (int i = 1; i <= p.npx; ++i) {
                    const size_t q = at(i, j, k);
                    ro[q] = ro[q] - bx[q] * bx[q] / tt[q] * h->d.obeta;
                }
        }
    });
    if (p.ampt != 0.0) {
        // src/static_ic.cpp:114-159: one stream in global (k, j, i) order
        Ran2 g(p.seed);
        for (int k = 1; k <= p.npz; ++k)
            for (int j = 1; j <= p.npy; ++j)
                for (int i = 1; i <= p.npx; ++i) {
                    const double w = 1.0 + p.ampt * (g.next() - 0.5);
                    tt[at(i, j, k)] *= w;
                }
    }
    // periodic x/y images of every plane (src/communicate.cpp:105-168)
    parallel_ranges(Z, [&](long ka, long kb) {
        for (int q = 0; q < nf; ++q) {
            double* a = h->f[q].data();
            for (int k = (int)ka; k < (int)kb; ++k) {
                for (int i = 0; i < X; ++i) {
                    a[at(i, 0, k)] = a[at(i, p.npy, k)];
                    a[at(i, p.npy + 1, k)] = a[at(i, 1, k)];
                }
                for (int j = 0; j < Y; ++j) {
                    a[at(0, j, k)] = a[at(p.npx, j, k)];
                    a[at(p.npx + 1, j, k)] = a[at(1, j, k)];
                }
            }
        }
    });
    return 0;
}

// gstate field order (src/main.cpp:44-93) -> R3D field id (-1: scratch array,
// not part of the state)
constexpr int kGstateMap[27] = {R3D_RU, R3D_RV, R3D_RW, R3D_RO, R3D_TT, -1, -1, -1, -1, -1,
                                -1, -1, -1, R3D_ZRU, R3D_ZRV, R3D_ZRW, R3D_ZRO, R3D_ZTT, -1,
                                -1, -1, R3D_BX, R3D_BY, R3D_BZ, R3D_ZBX, R3D_ZBY, R3D_ZBZ};
constexpr int32_t kGstateMagic = 1129928788;

// Read one virtual rank's gstate dump (the format src/main.cpp:44-93 writes)
// into fields[R3D_RU..R3D_ZBZ] (NULL entries skipped; each nx*ny*nz) and s.
// The header must describe this rank of this run: magic, 27 + 14 arrays, the
// local extents, the rank's coordinates, LMAG and the boundary switches.
int read_gstate(const r3d_host* h, int rank, const char* path, double* const* fields,
                r3d_hip_scalars* s) {
    const Params& p = h->p;
    if (rank < 0 || rank >= p.npey * p.npez) return fail("rank %d out of range", rank);
    const int nry = p.npy / p.npey, nrz = p.npz / p.npez;
    const int nx = p.npx + 2, ny = nry + 2, nz = nrz + 2;
    FILE* f = fopen(path, "rb");
    if (!f) return fail("cannot open dump '%s'", path);
    struct Closer {
        FILE* f;
        ~Closer() { fclose(f); }
    } closer{f};
    fseek(f, 0, SEEK_END);
    const long fsize = ftell(f);
    fseek(f, 0, SEEK_SET);
    auto rd = [&](void* v, size_t bytes) { return fread(v, 1, bytes, f) == bytes; };
    // skip bytes; false if that passes the end of the file
    auto skip = [&](int64_t bytes) {
        return fseek(f, (long)bytes, SEEK_CUR) == 0 && ftell(f) <= fsize;
    };
    int32_t head[3], hdr[13];
    double sc[27];
    if (!rd(head, sizeof head) || head[0] != kGstateMagic || head[1] != 27 || head[2] != 14)
        return fail("'%s' is not a gstate dump", path);
    if (!rd(hdr, sizeof hdr) || !rd(sc, sizeof sc)) return fail("'%s': truncated header", path);
    const int32_t want[13] = {nx, ny, nz, rank, rank % p.npey, rank / p.npey, hdr[6], p.lmag,
                              p.ixcon, p.iycon, p.izcon, p.itcon, p.ibcon};
    static const char* names[13] = {"nx", "ny", "nz", "rank", "mypey", "mypez", "nit", "lmag",
                                    "ixcon", "iycon", "izcon", "itcon", "ibcon"};
    for (int q = 0; q < 13; ++q)
        if (hdr[q] != want[q])
            return fail("'%s': %s = %d, this run has %d", path, names[q], hdr[q], want[q]);
    const int64_t cnt = (int64_t)nx * ny * nz;
    for (int q = 0; q < 27; ++q) {
        int64_t c;
        if (!rd(&c, 8)) return fail("'%s': truncated before field %d", path, q);
        if (c != cnt) return fail("'%s': field %d has a wrong length", path, q);
        double* dst = kGstateMap[q] >= 0 ? fields[kGstateMap[q]] : nullptr;
        if (dst ? !rd(dst, (size_t)cnt * 8) : !skip(cnt * 8))
            return fail("'%s': truncated field %d", path, q);
    }
    for (int q = 0; q < 14; ++q) {
        int64_t c;
        const int64_t len = q < 4 ? nx : q < 8 ? ny : nz;
        if (!rd(&c, 8)) return fail("'%s': truncated before metric %d", path, q);
        if (c != len) return fail("'%s': metric %d has a wrong length", path, q);
        if (!skip(c * 8)) return fail("'%s': truncated metric %d", path, q);
    }
    int64_t end;
    if (!rd(&end, 8) || end != -1) return fail("'%s': missing end marker", path);
    s->dt = sc[0];
    s->timc = sc[1];
    s->timt = sc[2];
    s->timi = sc[3];
    s->umach = sc[4];
    s->nit = hdr[6];
    return 0;
}

// Restart (--nstart N): the global primaries and the scalars from the dumps
// gstate.step.<rank>.<N> in --dumpdir, one per virtual rank.  A rank supplies
// its interior and the ghost cells on the global boundary (the periodic
// images and the outer z planes, whose RO is as the dumped step left it).
// The static profile still gives tb, the fixed bottom temperature.
int restart_state(r3d_host* h) {
    const Params& p = h->p;
    const int X = p.npx + 2, Y = p.npy + 2, Z = p.npz + 2;
    const size_t plane = (size_t)X * Y, n = plane * Z;
    const int nf = p.lmag ? 8 : 5;
    for (int q = 0; q < nf; ++q) h->f[q].resize(n);
    parallel_ranges(Z, [&](long ka, long kb) {  // parallel first touch
        for (int q = 0; q < nf; ++q)
            std::fill(h->f[q].data() + ka * plane, h->f[q].data() + kb * plane, 0.0);
    });
    std::vector<double> T, R;
    if (static_profile(h, T, R)) return -1;  // tb (and with LREM the conductivity)
    const int nry = p.npy / p.npey, nrz = p.npz / p.npez;
    const int nx = X, ny = nry + 2, nz = nrz + 2;
    const size_t cnt = (size_t)nx * ny * nz;
    std::vector<double> loc[R3D_NFIELDS];
    double* ptr[R3D_NFIELDS] = {};
    for (int q = 0; q < nf; ++q) {
        loc[q].assign(cnt, 0.0);
        ptr[q] = loc[q].data();
    }
    for (int r = 0; r < p.npey * p.npez; ++r) {
        char path[4096];
        snprintf(path, sizeof path, "%s/gstate.step.%03d.%04d", p.dumpdir.c_str(), r, p.nstart);
        r3d_hip_scalars s{};
        if (read_gstate(h, r, path, ptr, &s)) return -1;
        if (s.nit != p.nstart) return fail("'%s' holds step %d, not %d", path, s.nit, p.nstart);
        if (r == 0)
            h->sc = s;
        else if (memcmp(&s, &h->sc, sizeof s))
            return fail("'%s': scalars differ from rank 0's dump", path);
        const int gj0 = (r % p.npey) * nry, gk0 = (r / p.npey) * nrz;
        for (int k = 0; k < nz; ++k) {
            const int gk = gk0 + k;
            if (!(k >= 1 && k <= nz - 2) && gk != 0 && gk != p.npz + 1) continue;
            for (int j = 0; j < ny; ++j) {
                const int gj = gj0 + j;
                if (!(j >= 1 && j <= ny - 2) && gj != 0 && gj != p.npy + 1) continue;
                for (int q = 0; q < nf; ++q)
                    memcpy(&h->f[q][(size_t)gk * plane + (size_t)gj * X],
                           &loc[q][((size_t)k * ny + j) * nx], sizeof(double) * nx);
            }
        }
    }
    return 0;
}

int validate(const Params& p) {
    if (p.ix != 2 || p.iy != 2 || p.ilap != 2) return fail("only ix=iy=ilap=2 is supported");
    if (p.npey < 1 || p.npez < 1 || p.npy % p.npey || p.npz % p.npez)
        return fail("NPY=%d / NPZ=%d not divisible by NPEY=%d / NPEZ=%d", p.npy, p.npz, p.npey,
                    p.npez);
    if (p.npx < 3 || p.npy < 3 || p.npz < 6) return fail("grid too small");
    if (p.nstart < 0) return fail("NSTART must be >= 0");
    // the reference's feature guards (src/fluxes.cpp:46-51, src/static_ic.cpp:26-29)
    if (p.lshr) return fail("unsupported feature enabled: LSHR");
    // LREM (golden/fortran/3dmhdsub.f:437-490, 817-851, 1316, 1534-1596,
    // 2852-2952) is built; its mean temperature TTM uses the reference's
    // horizontal_mean, NGRID=0 only, and its flux relaxation variant is not
    if (p.lrem && p.ngrid != 0)
        return fail("LREM needs NGRID=0 (horizontal_mean: only NGRID=0 is supported)");
    if (p.lrem && p.rlax > 0.0)
        return fail("flux relaxation with LREM (FTOT = 8.07*0.4*THETA) is not built");
    // flux relaxation (golden/fortran/3dmhdsub.f:1625-1771) is built for the
    // ITC=0 / IZC=1 branch; in the ITC=1 / IZC=0 branch the last z rank's
    // never-assigned FRADM(NZ) reaches every rank's correction (:1697-1760),
    // and other boundary choices stop the Fortran ("Relax: check boundary
    // temperature")
    if (p.rlax > 0.0) {
        if (!(p.itcon == 0 && p.izcon == 1))
            return fail("flux relaxation (RLAX) is defined for ITCON=0 with IZCON=1 only");
        if (p.ngrid != 0)
            return fail("flux relaxation needs NGRID=0 (horizontal_mean: only NGRID=0 is supported)");
        if (p.toff <= 0.0) return fail("flux relaxation needs TOFF > 0");
        if (p.npez < 2) return fail("flux relaxation needs npez >= 2 (its z-rank exchanges)");
    }
    // density diffusion (golden/fortran/3dmhdsub.f:1306-1380) is built for ID
    // 1 and 2; its horizontal mean exists for NGRID=0 only (src/communicate.cpp:190)
    if (p.id != 0 && p.id != 1 && p.id != 2)
        return fail("density diffusion ID=%d: 1 and 2 are defined", p.id);
    if (p.id != 0 && p.ngrid != 0)
        return fail("density diffusion needs NGRID=0 (horizontal_mean: only NGRID=0 is supported)");
    // ITCON=2 (embedded heat loss, golden/fortran/3dmhdsub.f:1797-1822) is
    // built; ITCON=3 has only a banner line in the Fortran (3dmhd.f:645, 1179)
    if (p.itcon == 3) return fail("unsupported feature enabled: ITCON=3 embedded thermal");
    // the ramp divides by QFH (golden/fortran/3dmhdsub.f:1806): without it the
    // loss is inf or NaN from the first step on
    if (p.itcon == 2 && p.tc != 0.0 && p.qfh == 0.0)
        return fail("ITCON=2 with TC != 0 needs QFH != 0 (the heat-loss ramp time)");
    if (p.ixcon != 0 || p.iycon != 0) return fail("unsupported feature enabled: non-periodic horizontal boundaries");
    if (p.lmag && p.ntube != 0) return fail("unsupported feature enabled: TUBE models");
    if (p.ampb != 0.0 && !p.lmag) return fail("AMPB magnetic layer requires LMAG");
    if (p.izcon != 0 && p.izcon != 1) return fail("BCON: invalid IZCON %d", p.izcon);
    return 0;
}

}  // namespace

extern "C" {

const char* r3d_host_last_error(void) { return g_err.c_str(); }

r3d_host* r3d_host_create(int argc, const char* const* argv) {
    auto* h = new r3d_host();
    if (parse(argc, argv, h->p) || validate(h->p)) {
        delete h;
        return nullptr;
    }
    h->d = derive(h->p);
    build_metrics(h);
    if (h->p.nstart > 0 ? restart_state(h) : build_state(h)) {
        delete h;
        return nullptr;
    }
    return h;
}

int r3d_host_scalars(const r3d_host* h, r3d_hip_scalars* out) {
    if (!h || !out) return fail("null argument");
    *out = h->sc;
    return 0;
}

int r3d_host_read_gstate(const r3d_host* h, int rank, const char* path, double* const* fields,
                         r3d_hip_scalars* s) {
    if (!h || !path || !fields || !s) return fail("null argument");
    return read_gstate(h, rank, path, fields, s);
}

void r3d_host_free(r3d_host* h) { delete h; }

int r3d_host_cfg(const r3d_host* h, r3d_hip_cfg* o) {
    if (!h || !o) return fail("null argument");
    const Params& p = h->p;
    const Derived& d = h->d;
    memset(o, 0, sizeof *o);
    o->npx = p.npx;
    o->npy = p.npy;
    o->npz = p.npz;
    o->npey = p.npey;
    o->npez = p.npez;
    o->lmag = p.lmag;
    o->lrot = p.lrot;
    o->itcon = p.itcon;
    o->izcon = p.izcon;
    o->ibcon = p.ibcon;
    o->variant = 0;
    o->device = p.gpu;
    o->c13 = d.c13;
    o->c23 = d.c23;
    o->c43 = d.c43;
    o->gam1 = d.gam1;
    o->gam2 = d.gam2;
    o->gam3 = d.gam3;
    o->zeta1 = d.zeta1;
    o->zeta2 = d.zeta2;
    o->hx = d.hx;
    o->h2x = d.h2x;
    o->hy = d.hy;
    o->h2y = d.h2y;
    o->hz = d.hz;
    o->h2z = d.h2z;
    o->re = p.re;
    o->repr = d.repr;
    o->ore = d.ore;
    o->cv = d.cv;
    o->ocv = d.ocv;
    o->gamma = p.gamma;
    o->grav = p.grav;
    o->theta = p.theta;
    o->sff = p.sff;
    o->rm = p.rm;
    o->orm = d.orm;
    o->obeta = d.obeta;
    o->omx = d.omx;
    o->omz = d.omz;
    o->tp = p.tp;
    o->tc = p.tc;
    o->hh = p.hh;
    o->xp = p.xp;
    o->yp = p.yp;
    o->dzf = single(1.0f / (float)(p.npz - 1));
    o->tb = h->tb;
    o->zp = p.zp;
    o->qfh = p.qfh;
    o->dh = p.dh;
    o->dv = p.dv;
    o->zmax = p.zmax;
    o->id = p.id;
    o->rlax = p.rlax;
    o->tstart = p.tstart;
    o->toff = p.toff;
    o->lrem = p.lrem;
    o->tu = h->tu;
    o->dztu = h->dztu;
    o->dztb = h->dztb;
    return 0;
}

int r3d_host_metrics(const r3d_host* h, r3d_hip_metrics* o) {
    if (!h || !o) return fail("null argument");
    const double** dst[14] = {&o->exx, &o->dxxdx, &o->d2xxdx2, &o->ddx, &o->wyy, &o->dyydy,
                              &o->d2yydy2, &o->ddy, &o->zee, &o->dzzdz, &o->d2zzdz2, &o->ddz,
                              &o->rkapa, &o->dkapa};
    for (int q = 0; q < 14; ++q) *dst[q] = h->m1[q].data();
    return 0;
}

int r3d_host_state(const r3d_host* h, const double* fields[8]) {
    if (!h || !fields) return fail("null argument");
    for (int q = 0; q < 8; ++q) fields[q] = h->f[q].empty() ? nullptr : h->f[q].data();
    return 0;
}

void r3d_host_dims(const r3d_host* h, int dims[8]) {
    const Params& p = h->p;
    const int v[8] = {p.npx, p.npy, p.npz, p.npey, p.npez, p.lmag, p.ntotal, p.nstep0};
    memcpy(dims, v, sizeof v);
}

double r3d_host_tb(const r3d_host* h) { return h->tb; }

int r3d_host_rank_metrics(const r3d_host* h, int rank, double* const* out) {
    if (!h || !out) return fail("null argument");
    const Params& p = h->p;
    if (rank < 0 || rank >= p.npey * p.npez) return fail("rank %d out of range", rank);
    const int py = rank % p.npey, pz = rank / p.npey;
    const int nry = p.npy / p.npey, nrz = p.npz / p.npez;
    const int nx = p.npx + 2, ny = nry + 2, nz = nrz + 2;
    for (int q = 0; q < 4; ++q)
        if (out[q]) memcpy(out[q], h->m1[q].data(), sizeof(double) * nx);
    // rank-local y/z metrics: local index m is global py*nry + m (z alike);
    // on the first/last rank the outer ghost uses coordinate 0 / 1, which is
    // what the global arrays hold at 0 / np+1
    for (int q = 4; q < 8; ++q)
        if (out[q])
            for (int j = 0; j < ny; ++j) out[q][j] = h->m1[q][py * nry + j];
    for (int q = 8; q < 14; ++q)
        if (out[q])
            for (int k = 0; k < nz; ++k) out[q][k] = h->m1[q][pz * nrz + k];
    return 0;
}

int r3d_host_write_gstate(const r3d_host* h, int rank, const char* path,
                          const double* const* fields, const r3d_hip_scalars* s) {
    if (!h || !path || !fields || !s) return fail("null argument");
    const Params& p = h->p;
    const Derived& d = h->d;
    if (rank < 0 || rank >= p.npey * p.npez) return fail("rank %d out of range", rank);
    const int nry = p.npy / p.npey, nrz = p.npz / p.npez;
    const int nx = p.npx + 2, ny = nry + 2, nz = nrz + 2;
    FILE* f = fopen(path, "wb");
    if (!f) return fail("cannot open dump '%s'", path);
    const int32_t head[3] = {kGstateMagic, 27, 14};
    const int32_t hdr[13] = {nx, ny, nz, rank, rank % p.npey, rank / p.npey, s->nit, p.lmag,
                             p.ixcon, p.iycon, p.izcon, p.itcon, p.ibcon};
    const double sc[27] = {s->dt, s->timc, s->timt, s->timi, s->umach, p.re, d.repr, d.cv,
                           d.ocv, d.ore, p.theta, p.grav, d.rkapst, d.orm, p.rm, d.obeta,
                           d.omx, d.omz, p.sff, p.ampt, 0.0, d.hx, d.h2x, d.hy, d.h2y, d.hz,
                           d.h2z};
    fwrite(head, 4, 3, f);
    fwrite(hdr, 4, 13, f);
    fwrite(sc, 8, 27, f);
    // gstate order: ru rv rw ro tt uu vv ww fu fv fw fr ft zru zrv zrw zro ztt
    // ww1 ww2 ww3 bx by bz zbx zby zbz; scratch arrays are not part of the
    // device state and are written as zeros
    const int* map = kGstateMap;
    const int64_t cnt = (int64_t)nx * ny * nz;
    std::vector<double> zero((size_t)cnt, 0.0);
    for (int q = 0; q < 27; ++q) {
        fwrite(&cnt, 8, 1, f);
        const double* src = map[q] >= 0 && fields[map[q]] ? fields[map[q]] : zero.data();
        fwrite(src, 8, (size_t)cnt, f);
    }
    std::vector<std::vector<double>> m(14);
    double* mp[14];
    for (int q = 0; q < 14; ++q) {
        m[q].assign(q < 4 ? nx : q < 8 ? ny : nz, 0.0);
        mp[q] = m[q].data();
    }
    r3d_host_rank_metrics(h, rank, mp);
    for (int q = 0; q < 14; ++q) {
        const int64_t c = (int64_t)m[q].size();
        fwrite(&c, 8, 1, f);
        fwrite(m[q].data(), 8, m[q].size(), f);
    }
    const int64_t end = -1;
    fwrite(&end, 8, 1, f);
    return fclose(f) ? fail("write error on '%s'", path) : 0;
}

}  // extern "C"

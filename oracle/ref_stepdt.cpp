// ref_stepdt.cpp - runs the reference's own STEP time-step limits, LMAG
// included (CHECKER ONLY).
//
// Test infrastructure: linked by oracle/ref.mk against the reference sources
// where they lie into oracle/_ref/stepdt.  The product never builds, links or
// runs it.
//
//   mpirun -np <npey*npez> oracle/_ref/stepdt <outdir> <flow> <reference flags...>
//
// STEP (src/step.cpp:121-229) first forms, per interior cell, UU = |m|^2/rho^2
// and then the fast speed UU + gamma*T (+ b2 with LMAG) + 2*sqrt(...), the
// limits WW1 = dd/sqrt(UU), WW2 (thermal), WW3 (viscous) and, with LMAG, the
// resistive VV = 0.5*dd*dd*Rm; umach and dt = sff*min are reduced over the
// ranks; then it saves X into Z and calls FLUXES.  The reference C++ cannot
// take an LMAG step (FLUXES stops, src/fluxes.cpp:539-541), but everything
// before that first FLUXES call is its own code.  This driver builds the
// state as src/main.cpp does (flags, topology, derived constants, grid
// metrics, STATIC, halo exchange; the conductivity profile only for pzp = 0),
// optionally replaces the interior velocities and transverse field by a
// deterministic flow, and calls r3d::step(1, ...).  The link wraps
// r3d::fluxes (ld --wrap, see oracle/ref.mk), so the first FLUXES call lands
// in this file instead: it writes what STEP computed and ends the run.
//
//   <outdir>/stepdt.<rank>.bin:
//     int32 nx, ny, nz; double dt, umach;
//     then nx*ny*nz doubles each of ru rv rw ro tt bx by bz uu vv ww1 ww2 ww3
//
// flow = 1 sets, for every interior cell (i, j, k) of a rank, with global
// ghost-inclusive indices ig = i, jg = mypey*nry + j, kg = mypez*nrz + k and
//   h(c) = (ig*7919 + jg*104729 + kg*1299709 + c*15485863) % 10007,
//   w(c) = h(c) / 10007.0 - 0.5
//   ru = ro * (w(0) * 0.4), rv = ro * (w(1) * 0.4), rw = ro * (w(2) * 0.4),
//   and with LMAG by = w(3) * 4.0, bz = w(4) * 4.0
// (each product rounded once, as written).  flow = 0 keeps STATIC's state
// (fluid at rest, with LMAG the magnetic layer).
#include <mpi.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "communicate.hpp"
#include "fields.hpp"
#include "grid.hpp"
#include "params.hpp"
#include "static_ic.hpp"
#include "step.hpp"
#include "topology.hpp"

namespace {

std::string g_out;

void write_state(const r3d::Derived& d, const r3d::Topology& t, const r3d::SimState& s) {
    const std::string path = g_out + "/stepdt." + std::to_string(t.mype) + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "stepdt: cannot open %s\n", path.c_str());
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const int dims[3] = {d.nx, d.ny, d.nz};
    std::fwrite(dims, sizeof(int), 3, f);
    std::fwrite(&s.dt, sizeof(double), 1, f);
    std::fwrite(&s.umach, sizeof(double), 1, f);
    const r3d::Field* fl[13] = {&s.ru, &s.rv, &s.rw, &s.ro,  &s.tt,  &s.bx, &s.by,
                                &s.bz, &s.uu, &s.vv, &s.ww1, &s.ww2, &s.ww3};
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    for (const r3d::Field* a : fl) std::fwrite(a->data(), sizeof(double), n, f);
    if (std::fclose(f) != 0) MPI_Abort(MPI_COMM_WORLD, 1);
}

}  // namespace

// The first FLUXES call of STEP (r3d::fluxes, src/fluxes.cpp:44) is linked
// here: STEP's limits are complete and nothing has overwritten them yet.
extern "C" void __wrap__ZN3r3d6fluxesERKNS_6ParamsERKNS_7DerivedERKNS_8TopologyEPKdSA_RNS_8SimStateE(
    const r3d::Params&, const r3d::Derived& d, const r3d::Topology& t, const double*,
    const double*, r3d::SimState& s) {
    write_state(d, t, s);
    std::fflush(nullptr);
    MPI_Barrier(MPI_COMM_WORLD);
    MPI_Finalize();
    _exit(0);
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int npe = 1, mype = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &npe);
    MPI_Comm_rank(MPI_COMM_WORLD, &mype);
    if (argc < 3) {
        if (mype == 0) std::fprintf(stderr, "usage: stepdt outdir flow(0|1) [flags...]\n");
        MPI_Finalize();
        return 2;
    }
    g_out = argv[1];
    const int flow = std::atoi(argv[2]);
    // the reference parser sees argv[0] followed by the flags
    std::vector<char*> av;
    av.push_back(argv[0]);
    for (int a = 3; a < argc; ++a) av.push_back(argv[a]);
    r3d::Params p;
    if (r3d::parse_args((int)av.size(), av.data(), p) != 0) MPI_Abort(MPI_COMM_WORLD, 1);
    if (p.pzp != 0.0) {
        if (mype == 0) std::fprintf(stderr, "stepdt: only pzp = 0 (uniform conductivity)\n");
        MPI_Finalize();
        return 2;
    }
    const r3d::Topology t = r3d::resolve_topology(npe, mype, p);
    const r3d::Derived d = r3d::derive(p, npe, t.npey, t.npez);
    r3d::SimState s(d);
    r3d::build_grid_metrics(p, t, s);
    std::fill(s.rkapa.begin(), s.rkapa.end(), 1.0);
    std::fill(s.dkapa.begin(), s.dkapa.end(), 0.0);
    r3d::static_ic(MPI_COMM_WORLD, p, d, t, s);
    s.timi = 0.0;
    r3d::communicate(MPI_COMM_WORLD, p, t, s);

    if (flow) {
        for (int k = 1; k <= d.nz - 2; ++k)
            for (int j = 1; j <= d.ny - 2; ++j)
                for (int i = 1; i <= d.nx - 2; ++i) {
                    const long ig = i, jg = (long)t.mypey * d.nry + j,
                               kg = (long)t.mypez * d.nrz + k;
                    double w[5];
                    for (int c = 0; c < 5; ++c) {
                        const long h =
                            (ig * 7919 + jg * 104729 + kg * 1299709 + c * 15485863L) % 10007;
                        w[c] = (double)h / 10007.0 - 0.5;
                    }
                    const double ro = s.ro.at(i, j, k);
                    s.ru.at(i, j, k) = ro * (w[0] * 0.4);
                    s.rv.at(i, j, k) = ro * (w[1] * 0.4);
                    s.rw.at(i, j, k) = ro * (w[2] * 0.4);
                    if (p.lmag) {
                        s.by.at(i, j, k) = w[3] * 4.0;
                        s.bz.at(i, j, k) = w[4] * 4.0;
                    }
                }
    }

    r3d::step(1, MPI_COMM_WORLD, p, d, t, s);
    if (mype == 0) std::fprintf(stderr, "stepdt: FLUXES was not reached\n");
    MPI_Finalize();
    return 3;
}

// ref_hmean.cpp - runs the reference's own horizontal_mean (CHECKER ONLY).
//
// Test infrastructure: linked by oracle/ref.mk against the reference sources
// where they lie (src/communicate.cpp:184-223 and what it needs) into
// oracle/_ref/hmean.  The product never builds, links or runs it.
//
//   mpirun -np <npey*npez> oracle/_ref/hmean <npx> <npy> <npz> <npey>
//
// Every rank fills its local field, ghosts included, from the global cell
// position (ig, jg, kg) with
//   v = 1 + ((ig*7919 + jg*104729 + kg*1299709) % 10007) / 10007 / 2
// (jg = mypey*nry + j, kg = mypez*nrz + k, global ghost-inclusive indices),
// calls r3d::horizontal_mean with NGRID=0 and, on the ranks with mypey = 0,
// prints "HMEAN <mypez> <k> <%a of varm[k]>" for every local plane k.
#include <mpi.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "communicate.hpp"
#include "fields.hpp"
#include "params.hpp"
#include "topology.hpp"

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int npe = 1, mype = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &npe);
    MPI_Comm_rank(MPI_COMM_WORLD, &mype);
    if (argc != 5) {
        if (mype == 0) std::fprintf(stderr, "usage: hmean npx npy npz npey\n");
        MPI_Finalize();
        return 2;
    }
    r3d::Params p;
    p.npx = std::atoi(argv[1]);
    p.npy = std::atoi(argv[2]);
    p.npz = std::atoi(argv[3]);
    p.npey = std::atoi(argv[4]);
    p.ngrid = 0;
    if (p.npey < 1 || npe % p.npey || p.npy % p.npey || p.npz % (npe / p.npey)) {
        if (mype == 0) std::fprintf(stderr, "hmean: bad decomposition\n");
        MPI_Finalize();
        return 2;
    }
    // rank -> (mypey, mypez) as the reference maps it (src/topology.cpp:9-14)
    r3d::Topology t;
    t.npe = npe;
    t.mype = mype;
    t.npey = p.npey;
    t.npez = npe / p.npey;
    t.mypey = mype % t.npey;
    t.mypez = mype / t.npey;
    const r3d::Derived d = r3d::derive(p, npe, t.npey, t.npez);
    r3d::SimState s(d);
    r3d::Field f(r3d::Grid3{d.nx, d.ny, d.nz}, 0.0);
    for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i) {
                const long ig = i, jg = (long)t.mypey * d.nry + j, kg = (long)t.mypez * d.nrz + k;
                const long h = (ig * 7919 + jg * 104729 + kg * 1299709) % 10007;
                f.at(i, j, k) = 1.0 + (double)h / 10007.0 / 2.0;
            }
    std::vector<double> varm(d.nz, 0.0);
    r3d::horizontal_mean(MPI_COMM_WORLD, p, t, s, f, varm.data());
    if (t.mypey == 0)
        for (int k = 0; k < d.nz; ++k) std::printf("HMEAN %d %d %a\n", t.mypez, k, varm[k]);
    std::fflush(stdout);
    MPI_Finalize();
    return 0;
}

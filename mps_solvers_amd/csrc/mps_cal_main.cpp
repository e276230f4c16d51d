// mps_cal -- command-line driver with the reference's main.cpp behaviour
// (MPS_CAL_v4.0/main.cpp:14-37): read ./input.dat and ./GEO.dat, run to TT,
// write ./out/%04d.out every TOUT.
//
//   mps_cal [dir] [--semi] [--fp64] [--rebin-every-move] [--steps N] [--learned] [--vtk]
// --vtk also writes ./out/part%04d.vtk (the files MPS_POST would make).
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/mps_solvers.hpp"

int main(int argc, char** argv)
{
    std::string dir = ".";
    bool semi = false, fp64 = false, every = false, learned = false, vtk = false;
    long steps = -1;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--semi")) semi = true;
        else if (!strcmp(argv[i], "--fp64")) fp64 = true;
        else if (!strcmp(argv[i], "--rebin-every-move")) every = true;
        else if (!strcmp(argv[i], "--learned")) learned = true;
        else if (!strcmp(argv[i], "--vtk")) vtk = true;
        else if (!strcmp(argv[i], "--steps") && i + 1 < argc) steps = atol(argv[++i]);
        else dir = argv[i];
    }
    mkdir((dir + "/out").c_str(), 0755);
    mps_amd::MPS_GPU_EXPL expl(fp64 ? 1 : 0, every ? 1 : 0);
    mps_amd::MPS_CPU_EXPL impl(fp64 ? 1 : 0, every ? 1 : 0, 1e-8, learned ? 3 : 1);
    mps_amd::MPS_GPU_EXPL& m = semi ? (mps_amd::MPS_GPU_EXPL&)impl : expl;
    m.setWriteVtk(vtk);
    int rc = m.Initial(dir);
    if (rc) {
        fprintf(stderr, "mps_cal: initialisation failed with status %d\n", rc);
        return rc;
    }
    rc = m.mps_cal(steps);
    if (rc) fprintf(stderr, "mps_cal: stopped with status %d after %ld steps\n", rc, m.steps());
    printf("mps_cal: %ld steps, t = %.6f s, %d particles\n", m.steps(), m.time(), m.getNp());
    return rc;
}

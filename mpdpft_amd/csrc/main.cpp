// main.cpp -- mpdpft_hip: drop-in for ./mpDPFT (task types 0, 1 and 8).  Runs
// in the given directory (default: the current one), reading mpDPFT.input and
// writing mpDPFT_Den.dat, mpDPFT_V.dat, mpDPFT_Control.dat and
// mpDPFT_ControlTask.dat.
#include <cstdio>

#include "mpdpft.h"

int main(int argc, char **argv) {
    const char *dir = argc > 1 ? argv[1] : ".";
    mpdpft_result r;
    int rc = mpdpft_run_dir(dir, -1, &r);
    if (rc == MPDPFT_OK || rc == MPDPFT_ENOTCONV) {
        std::printf("SCiterations = %d; TotalEnergy = %.16g; converged = %d; solve_seconds = %.6g\n", r.SCcount - 1,
                    r.TotalEnergy, r.converged, r.solve_seconds);
    } else {
        std::fprintf(stderr, "mpdpft_hip: error %d (see mpDPFT_Control.dat)\n", rc);
    }
    return rc == MPDPFT_OK ? 0 : (rc == MPDPFT_ENOTCONV ? 3 : 1);
}

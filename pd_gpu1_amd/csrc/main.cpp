// main.cpp -- `pd_parallel [device] [mode] [dx] [gpu]`, the positional command
// line of main.cu:11-50.  device 2 = GPU (the CPU path, device 1, is not part
// of this engine); mode 1 plate tensile, 2 crack growth, 3 Kalthoff-Winkler
// (default).  dx is parsed as a float (main.cu:23 truncates it with atoi).
// The optional 4th argument is the GPU ordinal (main.cu:16 hard-codes 1), the
// optional 5th a path for the final state in the binary checkpoint format.
#include <cstdio>
#include <cstdlib>

#include "../../include/pd_solvers.h"

int main(int argc, char **argv)
{
    Dselect d{2, 3, 1, 3};
    int mode = 3;
    double dx = 0.0;
    int gpu = 0;
    const char *ckpt = nullptr;
    if (argc > 1) {
        d.device = std::atoi(argv[1]);
        if (argc > 2) mode = std::atoi(argv[2]);
        if (argc > 3) dx = std::atof(argv[3]);
        if (argc > 4) gpu = std::atoi(argv[4]);
        if (argc > 5) ckpt = argv[5];
    }
    int rc = 0;
    switch (mode) {
    case 1: rc = solve_plate_tensile(&d, dx, gpu, "disp.txt", ckpt); break;
    case 2: rc = solve_plate_crack_grow(&d, dx, gpu, "disp.txt", ckpt); break;
    case 3: rc = solve_kalthoff_winkler(&d, dx, gpu, "disp.txt", -1, ckpt); break;
    default: break;
    }
    return rc;
}

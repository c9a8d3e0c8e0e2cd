// eyelid_ref_main.cpp -- runs the reference's eyelid-conditioning driver
// (Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementbase.cpp and
// ecmanagementdelay.cpp, compiled unchanged from the reference tree against
// this library's drop-in headers and linked to libcbm_core) for a few trials
// of the delay protocol, then writes its data file.  Only
// tests/test_reference_driver_links.py builds it, where the reference tree is
// present.
//
//   eyelid_ref_main <conParams> <actParams> <trials> <iti ms> <out prefix>
#include <cstdio>
#include <cstdlib>
#include <exception>

#include <ectrial/ecmanagementdelay.h>

int main(int argc, char **argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: eyelid_ref_main conParams actParams trials iti outPrefix\n");
        return 2;
    }
    try {
        // the delay protocol's arguments as Eyelid_Sim_CBM_Dev's main passes them:
        // CS 200-700 ms of the trial, phasic until 240, CS from trial 1, data for every trial
        const int trials = std::atoi(argv[3]), iti = std::atoi(argv[4]);
        ECManagementDelay sim(argv[1], argv[2], 1, trials, iti, 200, 700, 240, 1, 0, trials, 0.03f, 0.02f,
                              0.1f, 1, 1, 30, 90, 200, 10, 5, 60, 110, 250, argv[5]);
        int steps = 0;
        while (sim.runStep()) steps++;
        sim.writeDataToFile();
        std::printf("eyelid_ref_main: %d steps\n", steps);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "eyelid_ref_main: %s\n", e.what());
        return 1;
    }
    return 0;
}

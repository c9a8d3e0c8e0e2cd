// ecmfpop_driver.cpp -- writes ECMFPopulation's file (writeToFile) for the
// seeded constructor or for recorded trials, through the C++ API the eyelid
// driver uses (Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementdelay.cpp:20,42,61).
//
//   ecmfpop_driver seeded <out> numMF seed fracCST fracCSP fracCtxt bgMin csBGMin ctxtMin csTMin csPMin
//                                    bgMax csBGMax ctxtMax csTMax csPMax
//   ecmfpop_driver trials <trials.dat> <out>   (ECTrialsData(fstream&) -> ECMFPopulation(ECTrialsData*))
//   ecmfpop_driver copy <in> <out>             (ECMFPopulation(fstream&) -> writeToFile)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include <CBMToolsInclude/ecmfpopulation.h>

int main(int argc, char **argv)
{
    if (argc < 4) return 2;
    const std::string mode = argv[1];
    ECMFPopulation *pop = nullptr;
    std::string out;
    if (mode == "seeded" && argc == 18) {
        float f[13];
        for (int i = 0; i < 13; i++) f[i] = (float)std::atof(argv[5 + i]);
        pop = new ECMFPopulation(std::atoi(argv[3]), std::atoi(argv[4]), f[0], f[1], f[2], f[3], f[4], f[5], f[6],
                                 f[7], f[8], f[9], f[10], f[11], f[12]);
        out = argv[2];
    } else if (mode == "trials" && argc == 4) {
        std::fstream in(argv[2], std::ios::in | std::ios::binary);
        ECTrialsData data(in);
        pop = new ECMFPopulation(&data);
        out = argv[3];
    } else if (mode == "copy" && argc == 4) {
        std::fstream in(argv[2], std::ios::in | std::ios::binary);
        pop = new ECMFPopulation(in);
        out = argv[3];
    } else {
        return 2;
    }
    std::fstream o(out, std::ios::out | std::ios::binary | std::ios::trunc);
    pop->writeToFile(o);
    delete pop;
    return 0;
}

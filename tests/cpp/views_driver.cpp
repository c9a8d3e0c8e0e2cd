// views_driver.cpp -- prints what the CBMState parameter and state views
// return (CBM_STATE_LIB_DEV/CBMStateInclude/interfaces/cbmstate.h:47-53), for
// comparison with the state's arrays.
//
//   views_driver <actParams> <conParams> <seed>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <algorithm>
#include <iostream>
#include <sstream>
#include <vector>

#include <CBMStateInclude/interfaces/cbmstate.h>

static void list(const char *name, const std::vector<ct_uint32_t> &v)
{
    std::printf("%s", name);
    for (ct_uint32_t x : v) std::printf(" %u", x);
    std::printf("\n");
}

int main(int argc, char **argv)
{
    if (argc != 4) return 2;
    std::fstream act(argv[1], std::ios::in), con(argv[2], std::ios::in);
    const int seed = std::atoi(argv[3]);
    int crSeed = seed + 101, arSeed = seed + 1001;
    CBMState state(act, con, 1, seed, &crSeed, &arSeed);
    IConnectivityParams *cp = state.getConnectivityParams();
    IActivityParams *ap = state.getActivityParams();
    std::printf("sizes %u %u %u %u %u %u %u %u %u %u %u %u %u\n", cp->getNumGR(), cp->getNumGO(), cp->getNumMF(),
                cp->getNumSC(), cp->getNumBC(), cp->getNumPC(), cp->getNumNC(), cp->getNumIO(), cp->getGRX(),
                cp->getGRY(), cp->getGOX(), cp->getGOY(), cp->getNumGL());
    std::printf("con grXP2 %u\n", cp->getParamCopy()["grXP2"]);
    std::printf("act msPerTimeStep %u msLTDStartAPIO %g copies %zu\n", ap->getMSPerTimeStep(),
                ap->getParam("msLTDStartAPIO"), ap->getParamCopy().size());
    std::ostringstream shown;
    ap->showParams(shown);
    cp->showParams(shown);
    const std::string txt = shown.str();
    std::printf("shown lines %zu\n", (size_t)std::count(txt.begin(), txt.end(), '\n'));
    const bool unknown = ap->setParam("noSuchParam", 1.0f);
    const bool known = ap->setParam("msLTDStartAPIO", -150.0f);
    std::printf("setParam unknown %d known %d now %g\n", (int)unknown, (int)known, ap->getParam("msLTDStartAPIO"));
    IInNetConState *in = state.getInnetConState();
    list("grMF5", in->getpGRfromMFtoGR(5));
    list("grGO5", in->getpGRfromGOtoGR(5));
    list("goGR3", in->getpGOfromGRtoGOCon(3));
    list("goGOGR3", in->getpGOfromGOtoGRCon(3));
    list("mfGO2", in->getpMFfromMFtoGOCon(2));
    list("goMF4", in->getpGOfromMFtoGOCon(4));
    list("goOut1", in->getpGOOutGOGOCon(1));
    list("goIn1", in->getpGOInGOGOCon(1));
    const std::vector<ct_uint32_t> out1 = in->getpGOOutGOGOCon(1);
    const int victim = out1.empty() ? 0 : (int)out1[0];
    const bool del = in->deleteGOGOConPair(1, victim);
    const bool delAgain = in->deleteGOGOConPair(1, victim);
    const bool add = in->addGOGOConPair(1, victim);
    const bool addAgain = in->addGOGOConPair(1, victim);
    std::printf("gogo %d %d %d %d\n", (int)del, (int)delAgain, (int)add, (int)addAgain);
    list("goOut1after", in->getpGOOutGOGOCon(1));
    std::printf("pop %zu %zu %zu\n", in->getpGRPopfromMFtoGR().size(), in->getpGRPopfromGOtoGR().size(),
                in->getpGOPopfromMFtoGOCon().size());
    IMZoneActState *z = state.getMZoneActState(0);
    std::vector<float> w = z->getGRPCSynWeightLinear();
    std::printf("w %zu %g\n", w.size(), w.empty() ? 0.0 : (double)w[0]);
    z->resetGRPCSynWeight();
    w = z->getGRPCSynWeightLinear();
    std::printf("wreset %g %g\n", (double)w.front(), (double)w.back());
    return 0;
}

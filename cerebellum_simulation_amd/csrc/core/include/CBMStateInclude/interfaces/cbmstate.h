// cbmstate.h -- CBMState with the constructors and accessors drivers use
// (CBM_STATE_LIB_DEV/CBMStateInclude/interfaces/cbmstate.h:34-81), built on
// the network-state C-ABI (include/cbm_state.h).
#ifndef CBM_CBMSTATE_H
#define CBM_CBMSTATE_H

#include <fstream>
#include <map>
#include <string>

#include <CXXToolsInclude/stdDefinitions/pstdint.h>

#include <cbm_state.h>

#include "iactivityparams.h"
#include "iconnectivityparams.h"
#include "iinnetconstate.h"
#include "imzoneactstate.h"

class CBMState
{
public:
    // Parameter files and zone count; seeds drawn from time(0) as in cbmstate.cpp:35-58.
    // A conParams file with a glomerulus layer (glXP2, ... as CBM_Params/conParams.txt
    // has) is wired by the reference's generator (CBM_GEN_REFERENCE); one without
    // it by the synthetic generator with 4 MF and 3 GO inputs per GR.
    CBMState(std::fstream &actPFile, std::fstream &conPFile, unsigned int nZones);
    // Explicit seeds (cbmstate.cpp:60-64).  mzoneARSeed seeds each zone's
    // generator for the NC->IO release draws.
    CBMState(std::fstream &actPFile, std::fstream &conPFile, unsigned int nZones, int innetCRSeed,
             int *mzoneCRSeed, int *mzoneARSeed);
    // A saved state (cbmstate.cpp:12-33): read from the stream's position; the
    // stream is left after the state.
    CBMState(std::fstream &infile);
    virtual ~CBMState();

    // Save the state (cbmstate.cpp:86-103).  As in the reference, this sets
    // the GO synaptic-depression arrays to 1 and 0 (innetactivitystate.cpp:247-248).
    void writeState(std::fstream &outfile);

    ct_uint32_t getNumZones();

    // Views of the parameters and state (cbmstate.h:47-53).  They stay valid
    // while the CBMState lives; they read and write its arrays, which
    // CBMSimCore::writeToState refreshes from a running simulation.
    IConnectivityParams *getConnectivityParams();
    IActivityParams *getActivityParams();
    IInNetConState *getInnetConState();
    IMZoneActState *getMZoneActState(unsigned int zoneN);

    // The underlying state handle and its arrays (CBMState layout).
    cbm_state *handle() { return st; }
    const cbm_sizes &sizes() const { return sz; }
    const cbm_act_params &params() const { return ap; }
    void *field(int f, int zone, size_t *bytes);

private:
    CBMState();
    void build(const std::string &act, const std::string &con, unsigned int nZones, int innetCRSeed,
               const int *mzoneCRSeed, const int *mzoneARSeed);
    cbm_state *st;
    cbm_sizes sz;
    cbm_act_params ap;
    IConnectivityParams *conView = nullptr;
    IActivityParams *actView = nullptr;
    IInNetConState *innetView = nullptr;
    IMZoneActState *zoneView[8] = {};
    void deleteViews();
    friend class CBMStateActView;
};

#endif

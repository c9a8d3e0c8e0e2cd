// cbmsimcore.h -- the simulation core drivers link against
// (CBM_CORE_LIB_MGPU_DEV/CBMCoreInclude/interface/cbmsimcore.h:32-75).
//
// One CBMSimCore runs its network on one GPU through the cbm_* C-ABI
// (include/cbm_hip.h).  gpuIndStart selects the device; numGPUP2 is accepted
// for source compatibility and must describe one GPU (-1 or 0).
#ifndef CBM_CBMSIMCORE_H
#define CBM_CBMSIMCORE_H

#include <fstream>

#include <CBMStateInclude/interfaces/cbmstate.h>
#include <CXXToolsInclude/stdDefinitions/pstdint.h>

#include "innetinterface.h"
#include "mzoneinterface.h"

#include <cbm_hip.h>

class CBMInNetView;
class CBMZoneView;

class CBMSimCore
{
public:
    CBMSimCore(CBMState *state, int gpuIndStart = -1, int numGPUP2 = -1);
    CBMSimCore(CBMState *state, int *mzoneRSeed, int gpuIndStart = -1, int numGPUP2 = -1);
    // Not in the reference: a core over a simulation handle made elsewhere
    // (e.g. by a Python host); it does not own the handle, and it has no
    // CBMState, so writeToState throws.
    CBMSimCore(cbm_handle *handle, const cbm_sizes &sizes);

    ~CBMSimCore();

    void calcActivity();

    void updateMFInput(const ct_uint8_t *mfIn);
    void updateErrDrive(unsigned int zoneN, float errDriveRelative);

    // Copy the device state back into the CBMState arrays.
    void writeToState();
    // ... and save it (cbmsimcore.cpp:75-80).  The reference's step reads the
    // GO synaptic-depression arrays from the state object that writeState
    // resets, so the running network is reset with it here too.
    void writeToState(std::fstream &outfile);

    InNetInterface *getInputNet();
    MZoneInterface **getMZoneList();

    // The C-ABI handle (for callers that mix both interfaces).
    cbm_handle *handle() { return h; }

private:
    CBMSimCore();
    void construct(CBMState *state, const int *mzoneRSeed, int gpuIndStart, int numGPUP2);

    void makeViews(const cbm_sizes &s);

    CBMState *simState;
    cbm_handle *h;
    bool ownsHandle;
    unsigned int numZones;
    CBMInNetView *inputNet;
    CBMZoneView **zones;
    MZoneInterface **zoneList;
};

#endif

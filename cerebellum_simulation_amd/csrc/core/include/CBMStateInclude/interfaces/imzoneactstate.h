// imzoneactstate.h -- a microzone's PF->PC weights (CBM_STATE_LIB_DEV/
// CBMStateInclude/interfaces/imzoneactstate.h; returned by
// CBMState::getMZoneActState).
#ifndef CBM_IMZONEACTSTATE_H
#define CBM_IMZONEACTSTATE_H

#include <vector>

#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class IMZoneActState
{
public:
    virtual ~IMZoneActState() {}

    // PF->PC weights of the zone in GR order
    virtual std::vector<float> getGRPCSynWeightLinear() = 0;
    // every weight back to pfSynWInitPC
    virtual void resetGRPCSynWeight() = 0;
};

#endif

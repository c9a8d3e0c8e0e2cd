// mzoneinterface.h -- microzone export interface
// (CBM_CORE_LIB_MGPU_DEV/CBMCoreInclude/interface/mzoneinterface.h:13-35).
#ifndef CBM_MZONEINTERFACE_H
#define CBM_MZONEINTERFACE_H
#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class MZoneInterface
{
public:
    virtual ~MZoneInterface() {}

    virtual void setGRPCPlastSteps(float ltdStep, float ltpStep) = 0;
    virtual void resetGRPCPlastSteps() = 0;

    virtual const ct_uint8_t *exportAPBC() = 0;
    virtual const ct_uint8_t *exportAPPC() = 0;
    virtual const ct_uint8_t *exportAPIO() = 0;
    virtual const ct_uint8_t *exportAPNC() = 0;

    virtual const ct_uint32_t *exportAPBufBC() = 0;
    virtual const ct_uint32_t *exportAPBufPC() = 0;
    virtual const ct_uint32_t *exportAPBufIO() = 0;
    virtual const ct_uint32_t *exportAPBufNC() = 0;

    virtual const float *exportVmBC() = 0;
    virtual const float *exportVmPC() = 0;
    virtual const float *exportVmNC() = 0;
    virtual const float *exportVmIO() = 0;
};
#endif

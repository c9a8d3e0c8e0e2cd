// innetinterface.h -- input-net export interface
// (CBM_CORE_LIB_MGPU_DEV/CBMCoreInclude/interface/innetinterface.h:13-42).
#ifndef CBM_INNETINTERFACE_H
#define CBM_INNETINTERFACE_H
#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class InNetInterface
{
public:
    virtual ~InNetInterface() {}

    virtual void setGIncGRtoGO(float inc) = 0;
    virtual void resetGIncGRtoGO() = 0;

    virtual const ct_uint8_t *exportAPSC() = 0;
    virtual const ct_uint8_t *exportAPGO() = 0;
    virtual const ct_uint8_t *exportAPGR() = 0;

    virtual const ct_uint8_t *exportHistMF() = 0;

    virtual const ct_uint32_t *exportAPBufMF() = 0;
    virtual const ct_uint32_t *exportAPBufGR() = 0;
    virtual const ct_uint32_t *exportAPBufGO() = 0;
    virtual const ct_uint32_t *exportAPBufSC() = 0;

    virtual const float *exportVmGR() = 0;
    virtual const float *exportVmGO() = 0;
    virtual const float *exportVmSC() = 0;

    virtual const float *exportGESumGR() = 0;
    virtual const float *exportGISumGR() = 0;

    virtual const ct_uint32_t *exportSumGRInputGO() = 0;
    virtual const float *exportSumGOInputGO() = 0;
    virtual const float *exportGOOutSynScaleGOGO() = 0;
};
#endif

// iactivityparams.h -- the activity-parameter interface drivers read
// parameters through (CBM_STATE_LIB_DEV/CBMStateInclude/interfaces/
// iactivityparams.h; returned by CBMState::getActivityParams).
#ifndef CBM_IACTIVITYPARAMS_H
#define CBM_IACTIVITYPARAMS_H

#include <iostream>
#include <map>
#include <string>

#include <CXXToolsInclude/stdDefinitions/pstdint.h>

class IActivityParams
{
public:
    virtual ~IActivityParams() {}

    virtual unsigned int getMSPerTimeStep() = 0;
    virtual void showParams(std::ostream &outSt) = 0;
    // the actParams file's key-value map
    virtual std::map<std::string, float> getParamCopy() = 0;
    // a value of that map (0 for a key the file does not give)
    virtual float getParam(std::string paramName) = 0;
    // change a value the file gives and re-derive the constants; false for
    // a key it does not give
    virtual bool setParam(std::string paramName, float value) = 0;
};

#endif

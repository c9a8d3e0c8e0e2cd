// randomc.h -- the random-generator declarations tools include
// (CXX_TOOLS_LIB's randomc.h, as vendored in CBM_Cart_On_Cart/includes): here
// the one generator class the drivers and tools use, CRandomSFMT0.
#ifndef CBM_CXXTOOLS_RANDOMC_H
#define CBM_CXXTOOLS_RANDOMC_H

#include "sfmt.h"

#endif

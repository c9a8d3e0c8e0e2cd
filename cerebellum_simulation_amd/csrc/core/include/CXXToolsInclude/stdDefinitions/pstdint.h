// pstdint.h -- the fixed-width integer names the reference's public headers use
// (ct_uint8_t ... ct_int64_t).  The original comes from the CXX_TOOLS_LIB
// sibling library, which is not part of the reference tree; these are plain
// aliases of <stdint.h>.
#ifndef CBM_PSTDINT_H
#define CBM_PSTDINT_H
#include <stdint.h>
typedef uint8_t ct_uint8_t;
typedef int8_t ct_int8_t;
typedef uint16_t ct_uint16_t;
typedef int16_t ct_int16_t;
typedef uint32_t ct_uint32_t;
typedef int32_t ct_int32_t;
typedef uint64_t ct_uint64_t;
typedef int64_t ct_int64_t;
#endif

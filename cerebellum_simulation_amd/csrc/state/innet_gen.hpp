// innet_gen.hpp -- the reference's input-net connectivity generator: the
// glomerulus layer and the GR, GO and MF lists built through it
// (InNetConnectivityState, CBM_STATE_LIB_DEV/src/state/innetconnectivitystate.cpp:
// connectGRGL 716-732, connectGOGL 734-753, connectMFGL 755-796, translateMFGL
// 798-829, translateGOGL 831-838, connectGRGO 840-850, connectCommon 964-1083,
// translateCommon 1085-1122), drawing from one SFMT stream in the reference's
// order.  GO-GO wiring and the delay masks, which follow on the same stream,
// are the state library's.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "sfmt.hpp"

namespace cbm {

struct InnetGenParams {
    uint32_t glX, glY, grX, grY, goX, goY, numMF;
    uint32_t maxGLtoGR, lowGLtoGR, maxGLtoGO, maxGLfromGO;     // per glomerulus
    uint32_t maxGRfromGL, maxGRfromGO, maxGRfromMF, maxGRtoGO;  // per GR
    uint32_t maxGOfromGR, maxGOfromGL, maxGOtoGL;               // per GO
    uint32_t spanGRDenX, spanGRDenY, spanGODecX, spanGODecY, spanGOAxonX, spanGOAxonY, spanGOAscX, spanGOAscY;
};

// A connection list per cell: a count and `cols` slots (UINT32_MAX when empty).
struct ConList {
    uint32_t rows = 0, cols = 0;
    std::vector<int32_t> num;
    std::vector<uint32_t> p;
    void init(uint32_t r, uint32_t c)
    {
        rows = r;
        cols = c;
        num.assign(r, 0);
        p.assign((size_t)r * c, UINT32_MAX);
    }
    uint32_t *row(uint32_t i) { return &p[(size_t)i * cols]; }
};

struct InnetGen {
    std::vector<uint8_t> hasGLfromMF;  // haspGLfromMFtoGL
    std::vector<uint32_t> glFromMF;    // pGLfromMFtoGL
    ConList glToGO, glFromGO, glToGR;  // pGLfromGLtoGO, pGLfromGOtoGL, pGLfromGLtoGR
    ConList mfToGL, mfToGR, mfToGO;    // pMFfromMFtoGL, pMFfromMFtoGR, pMFfromMFtoGO
    ConList goFromGL, goToGL, goFromMF, goToGR, goFromGR;  // pGOfromGLtoGO, pGOfromGOtoGL, ...
    ConList grFromGL, grToGO, grFromGO, grFromMF;          // pGRfromGLtoGR, pGRfromGRtoGO, ...
    uint64_t droppedConnections = 0;  // connections given up after the attempt limit
};

// Returns an empty string, or why the parameters cannot be generated.
std::string generateInnet(const InnetGenParams &prm, Sfmt &rng, InnetGen &out);

}  // namespace cbm

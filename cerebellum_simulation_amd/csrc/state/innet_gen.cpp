// innet_gen.cpp -- see innet_gen.hpp.
//
// The spatial connector (connectCommon): in each of srcMax passes every
// source cell, in an order drawn by IRandom with rejection of cells already
// done, tries up to maxAttempts destinations around its position scaled onto
// the destination grid -- the scaled position rounded, plus a uniform offset
// round((Random() - 0.5) * span) on each axis, wrapped onto the grid -- and
// connects to the first one with fewer than `lim` inputs (the normal limit
// for the first normAttempts tries, then the maximum).  Kept as the reference
// has them: the Y position is scaled by the X ratio of the grids
// (innetconnectivitystate.cpp:1033); the wrap takes the remainder of the
// position converted to unsigned (exact for the power-of-two grids); a
// source that finds no destination in maxAttempts tries leaves its slot of
// that pass empty and is counted in droppedConnections (1075-1078).
#include "innet_gen.hpp"

#include <algorithm>
#include <cmath>

namespace cbm {
namespace {

// One connector call; `unique` refuses a destination the source already has.
void connect(ConList &src, ConList &dst, uint32_t srcMax, uint32_t numSrc, uint32_t dstMax, uint32_t dstNorm,
             uint32_t srcGX, uint32_t srcGY, uint32_t dstGX, uint32_t dstGY, uint32_t spanX, uint32_t spanY,
             bool unique, Sfmt &rng, uint64_t &dropped)
{
    (void)srcGY;
    const uint32_t normAttempts = 20000, maxAttempts = 50000;
    const float scale = (float)srcGX / (float)dstGX;
    std::vector<uint8_t> done(numSrc);
    for (uint32_t pass = 0; pass < srcMax; pass++) {
        std::fill(done.begin(), done.end(), 0);
        uint32_t connected = 0;
        while (connected < numSrc) {
            const int s = rng.IRandom(0, (int)numSrc - 1);
            if (done[s]) continue;
            done[s] = 1;
            connected++;
            const int px = s % (int)srcGX, py = s / (int)srcGX;
            int32_t lim = (int32_t)dstNorm;
            uint32_t a = 0;
            for (; a < maxAttempts; a++) {
                if (a == normAttempts) lim = (int32_t)dstMax;
                int x = (int)std::round((float)px / scale);
                int y = (int)std::round((float)py / scale);
                x = (int)((double)x + std::round((rng.Random() - 0.5) * spanX));
                y = (int)((double)y + std::round((rng.Random() - 0.5) * spanY));
                x = (int)(((uint32_t)x % dstGX + dstGX) % dstGX);
                y = (int)(((uint32_t)y % dstGY + dstGY) % dstGY);
                const uint32_t d = (uint32_t)y * dstGX + (uint32_t)x;
                if (unique) {
                    const uint32_t *have = src.row((uint32_t)s);
                    bool dup = false;
                    for (uint32_t j = 0; j < pass && !dup; j++) dup = have[j] == d;
                    if (dup) continue;
                }
                if (dst.num[d] < lim) {
                    dst.row(d)[dst.num[d]] = (uint32_t)s;
                    dst.num[d]++;
                    src.row((uint32_t)s)[pass] = d;
                    src.num[s]++;
                    break;
                }
            }
            if (a == maxAttempts) dropped++;
        }
    }
}

// Relay lists through the glomeruli (translateCommon): for each presynaptic
// cell, every postsynaptic cell of every glomerulus it reaches, in order,
// appended to both lists.  An empty slot of the presynaptic list (a dropped
// connection) is skipped.
void relay(ConList &preGL, ConList &glPost, ConList &preOut, ConList &postIn, uint32_t numPre)
{
    for (uint32_t i = 0; i < numPre; i++) {
        preOut.num[i] = 0;
        for (int32_t j = 0; j < preGL.num[i]; j++) {
            const uint32_t gl = preGL.row(i)[j];
            if (gl == UINT32_MAX) continue;
            for (int32_t k = 0; k < glPost.num[gl]; k++) {
                const uint32_t post = glPost.row(gl)[k];
                preOut.row(i)[preOut.num[i]++] = post;
                postIn.row(post)[postIn.num[post]++] = i;
            }
        }
    }
}

}  // namespace

std::string generateInnet(const InnetGenParams &q, Sfmt &rng, InnetGen &g)
{
    const uint32_t numGL = q.glX * q.glY, numGR = q.grX * q.grY, numGO = q.goX * q.goY, numMF = q.numMF;
    if (!numGL || !numGR || !numGO || !numMF) return "empty population";
    if (numGL % numMF) return "numGL must be a multiple of numMF";
    const uint32_t mfGL = numGL / numMF;
    // list widths as InNetConnectivityState::allocateMemory makes them
    // (innetconnectivitystate.cpp:462-527); what the passes write must fit
    if (q.maxGRfromGL > q.maxGRfromGO) return "maxnumpGRfromGLtoGR exceeds the maxnumpGRfromGOtoGR columns it is stored in";
    if (q.maxGRfromGL > q.maxGRfromMF)
        return "a GR's MF inputs (one per glomerulus) can exceed maxnumpGRfromMFtoGR";
    if (q.maxGRfromGL * q.maxGLfromGO > q.maxGRfromGO)
        return "a GR's GO inputs (GOs per glomerulus x glomeruli) can exceed maxnumpGRfromGOtoGR";
    if (q.lowGLtoGR > q.maxGLtoGR) return "lownumpGLfromGLtoGR exceeds maxnumpGLfromGLtoGR";
    if (q.maxGRfromGL > numGL || q.maxGOfromGR > numGR) return "more connections than destinations";

    g.hasGLfromMF.assign(numGL, 0);
    g.glFromMF.assign(numGL, UINT32_MAX);
    g.glToGO.init(numGL, q.maxGLtoGO);
    g.glFromGO.init(numGL, q.maxGLfromGO);
    g.glToGR.init(numGL, q.maxGLtoGR);
    g.mfToGL.init(numMF, mfGL);
    g.mfToGR.init(numMF, mfGL * q.maxGLtoGR);
    g.mfToGO.init(numMF, mfGL * q.maxGLtoGO);
    g.goFromGL.init(numGO, q.maxGOfromGL);
    g.goToGL.init(numGO, q.maxGOtoGL);
    g.goFromMF.init(numGO, q.maxGOfromGL);
    g.goToGR.init(numGO, q.maxGOtoGL * q.maxGLtoGR);
    g.goFromGR.init(numGO, q.maxGOfromGR);
    g.grFromGL.init(numGR, q.maxGRfromGO);
    g.grToGO.init(numGR, q.maxGRtoGO);
    g.grFromGO.init(numGR, q.maxGRfromGO);
    g.grFromMF.init(numGR, q.maxGRfromGO);
    g.droppedConnections = 0;

    // connectGRGL: GR dendrites onto glomeruli, distinct per GR
    connect(g.grFromGL, g.glToGR, q.maxGRfromGL, numGR, q.maxGLtoGR, q.lowGLtoGR, q.grX, q.grY, q.glX, q.glY,
            q.spanGRDenX, q.spanGRDenY, true, rng, g.droppedConnections);
    // connectGOGL: GO basolateral dendrites, then GO axons
    connect(g.goFromGL, g.glToGO, q.maxGOfromGL, numGO, q.maxGLtoGO, q.maxGLtoGO, q.goX, q.goY, q.glX, q.glY,
            q.spanGODecX, q.spanGODecY, false, rng, g.droppedConnections);
    connect(g.goToGL, g.glFromGO, q.maxGOtoGL, numGO, q.maxGLfromGO, q.maxGLfromGO, q.goX, q.goY, q.glX, q.glY,
            q.spanGOAxonX, q.spanGOAxonY, false, rng, g.droppedConnections);

    // connectMFGL: each MF but the last draws mfGL free glomeruli; the last
    // takes the rest without counting them (its count stays 0)
    for (uint32_t i = 0; i + 1 < numMF; i++)
        for (uint32_t j = 0; j < mfGL; j++)
            for (;;) {
                const uint32_t gl = (uint32_t)rng.IRandom(0, (int)numGL - 1);
                if (g.hasGLfromMF[gl]) continue;
                g.glFromMF[gl] = i;
                g.hasGLfromMF[gl] = 1;
                g.mfToGL.row(i)[j] = gl;
                g.mfToGL.num[i]++;
                break;
            }
    uint32_t last = 0;
    for (uint32_t gl = 0; gl < numGL; gl++)
        if (!g.hasGLfromMF[gl]) {
            g.hasGLfromMF[gl] = 1;
            g.glFromMF[gl] = numMF - 1;
            g.mfToGL.row(numMF - 1)[last++] = gl;
        }

    // translateMFGL, translateGOGL
    relay(g.mfToGL, g.glToGR, g.mfToGR, g.grFromMF, numMF);
    relay(g.mfToGL, g.glToGO, g.mfToGO, g.goFromMF, numMF);
    relay(g.goToGL, g.glToGR, g.goToGR, g.grFromGO, numGO);

    // connectGRGO: GO ascending dendrites onto GRs
    connect(g.goFromGR, g.grToGO, q.maxGOfromGR, numGO, q.maxGRtoGO, q.maxGRtoGO, q.goX, q.goY, q.grX, q.grY,
            q.spanGOAscX, q.spanGOAscY, false, rng, g.droppedConnections);
    return std::string();
}

}  // namespace cbm

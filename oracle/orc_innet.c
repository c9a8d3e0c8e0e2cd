/* orc_innet.c -- see orc_innet.h.  TEST INFRASTRUCTURE. */
#include "orc_innet.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_sfmt.h"

typedef struct {
    const char *name;
    uint32_t rows, cols;
    int32_t *num;
    uint32_t *p;
} olist;

enum {
    L_GL_TO_GO, L_GL_FROM_GO, L_GL_TO_GR, L_MF_TO_GL, L_MF_TO_GR, L_MF_TO_GO, L_GO_FROM_GL, L_GO_TO_GL,
    L_GO_FROM_MF, L_GO_TO_GR, L_GO_FROM_GR, L_GR_FROM_GL, L_GR_TO_GO, L_GR_FROM_GO, L_GR_FROM_MF, L_COUNT
};
static const char *const kNames[L_COUNT] = {
    "pGLfromGLtoGO", "pGLfromGOtoGL", "pGLfromGLtoGR", "pMFfromMFtoGL", "pMFfromMFtoGR", "pMFfromMFtoGO",
    "pGOfromGLtoGO", "pGOfromGOtoGL", "pGOfromMFtoGO", "pGOfromGOtoGR", "pGOfromGRtoGO", "pGRfromGLtoGR",
    "pGRfromGRtoGO", "pGRfromGOtoGR", "pGRfromMFtoGR"};

struct orc_innet {
    uint32_t numGL, numMF, numGO, numGR;
    uint8_t *has;   /* haspGLfromMFtoGL */
    uint32_t *own;  /* pGLfromMFtoGL */
    olist l[L_COUNT];
    uint64_t dropped;
};

static void make(olist *x, const char *name, uint32_t rows, uint32_t cols)
{
    x->name = name;
    x->rows = rows;
    x->cols = cols;
    x->num = (int32_t *)calloc(rows ? rows : 1, sizeof(int32_t));
    x->p = (uint32_t *)malloc((size_t)(rows ? rows : 1) * (cols ? cols : 1) * sizeof(uint32_t));
    for (size_t i = 0; i < (size_t)rows * cols; i++) x->p[i] = 0xFFFFFFFFu;
}

/* connectCommon with the reference's parameters, in its order */
static void connect_common(olist *srcCon, olist *destCon, uint32_t srcMaxNumCon, uint32_t numSrcCells,
                           uint32_t destMaxNumCon, uint32_t destNormNumCon, uint32_t srcGridX, uint32_t srcGridY,
                           uint32_t destGridX, uint32_t destGridY, uint32_t spanX, uint32_t spanY,
                           uint32_t normConAttempts, uint32_t maxConAttempts, int needUnique, orc_sfmt *rng,
                           uint64_t *dropped)
{
    float gridXScaleStoD = (float)srcGridX / (float)destGridX;
    (void)srcGridY;
    char *srcConnected = (char *)malloc(numSrcCells);
    for (uint32_t i = 0; i < srcMaxNumCon; i++) {
        uint32_t srcNumConnected = 0;
        memset(srcConnected, 0, numSrcCells);
        while (srcNumConnected < numSrcCells) {
            int srcInd = orc_sfmt_irandom(rng, 0, (int)numSrcCells - 1);
            if (srcConnected[srcInd]) continue;
            srcConnected[srcInd] = 1;
            srcNumConnected++;
            int srcPosX = srcInd % (int)srcGridX;
            int srcPosY = srcInd / (int)srcGridX;
            int lim = (int)destNormNumCon;
            uint32_t attempts;
            for (attempts = 0; attempts < maxConAttempts; attempts++) {
                if (attempts == normConAttempts) lim = (int)destMaxNumCon;
                int tx = (int)roundf((float)srcPosX / gridXScaleStoD);
                int ty = (int)roundf((float)srcPosY / gridXScaleStoD); /* the X scale, as the reference */
                tx = (int)((double)tx + round((orc_sfmt_random(rng) - 0.5) * spanX));
                ty = (int)((double)ty + round((orc_sfmt_random(rng) - 0.5) * spanY));
                tx = (int)(((uint32_t)tx % destGridX + destGridX) % destGridX);
                ty = (int)(((uint32_t)ty % destGridY + destGridY) % destGridY);
                uint32_t derived = (uint32_t)ty * destGridX + (uint32_t)tx;
                if (needUnique) {
                    int unique = 1;
                    for (uint32_t j = 0; j < i; j++)
                        if (srcCon->p[(size_t)srcInd * srcCon->cols + j] == derived) {
                            unique = 0;
                            break;
                        }
                    if (!unique) continue;
                }
                if (destCon->num[derived] < lim) {
                    destCon->p[(size_t)derived * destCon->cols + destCon->num[derived]] = (uint32_t)srcInd;
                    destCon->num[derived]++;
                    srcCon->p[(size_t)srcInd * srcCon->cols + i] = derived;
                    srcCon->num[srcInd]++;
                    break;
                }
            }
            if (attempts == maxConAttempts) (*dropped)++;
        }
    }
    free(srcConnected);
}

/* translateCommon; a UINT_MAX glomerulus slot is skipped */
static int translate_common(olist *preGL, olist *glPost, olist *preCon, olist *postCon, uint32_t numPre)
{
    for (uint32_t i = 0; i < numPre; i++) {
        preCon->num[i] = 0;
        for (int32_t j = 0; j < preGL->num[i]; j++) {
            uint32_t gl = preGL->p[(size_t)i * preGL->cols + j];
            if (gl == 0xFFFFFFFFu) continue;
            for (int32_t k = 0; k < glPost->num[gl]; k++) {
                uint32_t post = glPost->p[(size_t)gl * glPost->cols + k];
                if ((uint32_t)preCon->num[i] >= preCon->cols || (uint32_t)postCon->num[post] >= postCon->cols) return -1;
                preCon->p[(size_t)i * preCon->cols + preCon->num[i]] = post;
                preCon->num[i]++;
                postCon->p[(size_t)post * postCon->cols + postCon->num[post]] = i;
                postCon->num[post]++;
            }
        }
    }
    return 0;
}

int orc_innet_generate(const orc_innet_params *c, int seed, orc_innet **out)
{
    orc_innet *g = (orc_innet *)calloc(1, sizeof(orc_innet));
    const uint32_t glX = 1u << c->glXP2, glY = 1u << c->glYP2, grX = 1u << c->grXP2, grY = 1u << c->grYP2;
    const uint32_t goX = 1u << c->goXP2, goY = 1u << c->goYP2;
    g->numGL = glX * glY;
    g->numMF = 1u << c->numMFP2;
    g->numGO = goX * goY;
    g->numGR = grX * grY;
    const uint32_t numpMFfromMFtoGL = g->numGL / g->numMF;
    g->has = (uint8_t *)calloc(g->numGL, 1);
    g->own = (uint32_t *)malloc(g->numGL * sizeof(uint32_t));
    for (uint32_t i = 0; i < g->numGL; i++) g->own[i] = 0xFFFFFFFFu;
    make(&g->l[L_GL_TO_GO], kNames[L_GL_TO_GO], g->numGL, c->maxnumpGLfromGLtoGO);
    make(&g->l[L_GL_FROM_GO], kNames[L_GL_FROM_GO], g->numGL, c->maxnumpGLfromGOtoGL);
    make(&g->l[L_GL_TO_GR], kNames[L_GL_TO_GR], g->numGL, c->maxnumpGLfromGLtoGR);
    make(&g->l[L_MF_TO_GL], kNames[L_MF_TO_GL], g->numMF, numpMFfromMFtoGL);
    make(&g->l[L_MF_TO_GR], kNames[L_MF_TO_GR], g->numMF, numpMFfromMFtoGL * c->maxnumpGLfromGLtoGR);
    make(&g->l[L_MF_TO_GO], kNames[L_MF_TO_GO], g->numMF, numpMFfromMFtoGL * c->maxnumpGLfromGLtoGO);
    make(&g->l[L_GO_FROM_GL], kNames[L_GO_FROM_GL], g->numGO, c->maxnumpGOfromGLtoGO);
    make(&g->l[L_GO_TO_GL], kNames[L_GO_TO_GL], g->numGO, c->maxnumpGOfromGOtoGL);
    make(&g->l[L_GO_FROM_MF], kNames[L_GO_FROM_MF], g->numGO, c->maxnumpGOfromGLtoGO);
    make(&g->l[L_GO_TO_GR], kNames[L_GO_TO_GR], g->numGO, c->maxnumpGOfromGOtoGL * c->maxnumpGLfromGLtoGR);
    make(&g->l[L_GO_FROM_GR], kNames[L_GO_FROM_GR], g->numGO, c->maxnumpGOfromGRtoGO);
    /* the GR's glomerulus and MF lists are as wide as its GO list (allocateMemory) */
    make(&g->l[L_GR_FROM_GL], kNames[L_GR_FROM_GL], g->numGR, c->maxnumpGRfromGOtoGR);
    make(&g->l[L_GR_TO_GO], kNames[L_GR_TO_GO], g->numGR, c->maxnumpGRfromGRtoGO);
    make(&g->l[L_GR_FROM_GO], kNames[L_GR_FROM_GO], g->numGR, c->maxnumpGRfromGOtoGR);
    make(&g->l[L_GR_FROM_MF], kNames[L_GR_FROM_MF], g->numGR, c->maxnumpGRfromGOtoGR);
    *out = g;
    if (c->maxnumpGRfromGLtoGR > c->maxnumpGRfromGOtoGR || g->numGL % g->numMF) return -1;

    orc_sfmt *rng = (orc_sfmt *)malloc(sizeof(orc_sfmt));
    orc_sfmt_init(rng, seed);
    /* connectGRGL */
    connect_common(&g->l[L_GR_FROM_GL], &g->l[L_GL_TO_GR], c->maxnumpGRfromGLtoGR, g->numGR, c->maxnumpGLfromGLtoGR,
                   c->lownumpGLfromGLtoGR, grX, grY, glX, glY, c->spanGRDenOnGLX, c->spanGRDenOnGLY, 20000, 50000, 1,
                   rng, &g->dropped);
    /* connectGOGL */
    connect_common(&g->l[L_GO_FROM_GL], &g->l[L_GL_TO_GO], c->maxnumpGOfromGLtoGO, g->numGO, c->maxnumpGLfromGLtoGO,
                   c->maxnumpGLfromGLtoGO, goX, goY, glX, glY, c->spanGODecDenOnGLX, c->spanGODecDenOnGLY, 20000,
                   50000, 0, rng, &g->dropped);
    connect_common(&g->l[L_GO_TO_GL], &g->l[L_GL_FROM_GO], c->maxnumpGOfromGOtoGL, g->numGO, c->maxnumpGLfromGOtoGL,
                   c->maxnumpGLfromGOtoGL, goX, goY, glX, glY, c->spanGOAxonOnGLX, c->spanGOAxonOnGLY, 20000, 50000,
                   0, rng, &g->dropped);
    /* connectMFGL */
    olist *mfgl = &g->l[L_MF_TO_GL];
    for (uint32_t i = 0; i + 1 < g->numMF; i++)
        for (uint32_t j = 0; j < numpMFfromMFtoGL; j++)
            for (;;) {
                int gl = orc_sfmt_irandom(rng, 0, (int)g->numGL - 1);
                if (g->has[gl]) continue;
                g->own[gl] = i;
                g->has[gl] = 1;
                mfgl->p[(size_t)i * mfgl->cols + j] = (uint32_t)gl;
                mfgl->num[i]++;
                break;
            }
    uint32_t lastCount = 0;
    for (uint32_t i = 0; i < g->numGL; i++)
        if (!g->has[i]) {
            g->has[i] = 1;
            g->own[i] = g->numMF - 1;
            mfgl->p[(size_t)(g->numMF - 1) * mfgl->cols + lastCount] = i;
            lastCount++;
        }
    /* translateMFGL, translateGOGL */
    int rc = translate_common(mfgl, &g->l[L_GL_TO_GR], &g->l[L_MF_TO_GR], &g->l[L_GR_FROM_MF], g->numMF);
    if (!rc) rc = translate_common(mfgl, &g->l[L_GL_TO_GO], &g->l[L_MF_TO_GO], &g->l[L_GO_FROM_MF], g->numMF);
    if (!rc)
        rc = translate_common(&g->l[L_GO_TO_GL], &g->l[L_GL_TO_GR], &g->l[L_GO_TO_GR], &g->l[L_GR_FROM_GO], g->numGO);
    /* connectGRGO */
    if (!rc)
        connect_common(&g->l[L_GO_FROM_GR], &g->l[L_GR_TO_GO], c->maxnumpGOfromGRtoGO, g->numGO,
                       c->maxnumpGRfromGRtoGO, c->maxnumpGRfromGRtoGO, goX, goY, grX, grY, c->spanGOAscDenOnGRX,
                       c->spanGOAscDenOnGRY, 20000, 50000, 0, rng, &g->dropped);
    free(rng);
    return rc;
}

void orc_innet_free(orc_innet *g)
{
    if (!g) return;
    for (int k = 0; k < L_COUNT; k++) {
        free(g->l[k].num);
        free(g->l[k].p);
    }
    free(g->has);
    free(g->own);
    free(g);
}

static olist *find(orc_innet *g, const char *name)
{
    for (int k = 0; k < L_COUNT; k++)
        if (g->l[k].name && !strcmp(g->l[k].name, name)) return &g->l[k];
    return NULL;
}

int orc_innet_shape(orc_innet *g, const char *name, uint32_t *rows, uint32_t *cols)
{
    if (!strcmp(name, "haspGLfromMFtoGL") || !strcmp(name, "pGLfromMFtoGL")) {
        *rows = g->numGL;
        *cols = 1;
        return 1;
    }
    olist *x = find(g, name);
    if (!x) return 0;
    *rows = x->rows;
    *cols = x->cols;
    return 1;
}

int orc_innet_copy(orc_innet *g, const char *name, int32_t *num, void *slots)
{
    if (!strcmp(name, "haspGLfromMFtoGL")) {
        memcpy(slots, g->has, g->numGL);
        return 1;
    }
    if (!strcmp(name, "pGLfromMFtoGL")) {
        memcpy(slots, g->own, g->numGL * sizeof(uint32_t));
        return 1;
    }
    olist *x = find(g, name);
    if (!x) return 0;
    if (num) memcpy(num, x->num, x->rows * sizeof(int32_t));
    if (slots) memcpy(slots, x->p, (size_t)x->rows * x->cols * sizeof(uint32_t));
    return 1;
}

uint64_t orc_innet_dropped(orc_innet *g) { return g->dropped; }

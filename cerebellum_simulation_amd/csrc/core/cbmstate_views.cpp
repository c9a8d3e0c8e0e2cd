// cbmstate_views.cpp -- the parameter and state interfaces of CBMState
// (CBM_STATE_LIB_DEV/CBMStateInclude/interfaces/cbmstate.h:47-53) as views of
// the state's host arrays and parsed parameter files (include/cbm_state.h).
//
//   IConnectivityParams  sizes of cbm_sizes, grid sides of the conParams
//                        file, its key-value map
//   IActivityParams      the actParams file's map; setParam re-derives the
//                        constants (cbm_state_set_act_param)
//   IInNetConState       the input-net lists, forward rows as stored, reverse
//                        lists gathered in increasing source order, GO->GO
//                        GABA edits on the stored out-lists
//   IMZoneActState       a zone's PF->PC weights
#include <CBMStateInclude/interfaces/cbmstate.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include <cbm_hip.h>

namespace {

std::map<std::string, double> paramMap(cbm_state *st, int which)
{
    std::map<std::string, double> m;
    const char *name = nullptr;
    double v = 0;
    for (size_t i = 0; cbm_state_param_at(st, which, i, &name, &v) == CBM_OK; i++) m[name] = v;
    return m;
}

// a [rows][cols] u32 list field with its per-row counts (nullptr: every slot
// up to the first UINT32_MAX)
struct Lists {
    const uint32_t *p = nullptr;
    const int32_t *n = nullptr;
    size_t rows = 0, cols = 0;
    std::vector<ct_uint32_t> row(size_t r) const
    {
        std::vector<ct_uint32_t> out;
        if (r >= rows) return out;
        const size_t k = n ? (size_t)std::max(0, std::min<int32_t>(n[r], (int32_t)cols)) : cols;
        for (size_t j = 0; j < k; j++) {
            const uint32_t v = p[r * cols + j];
            if (v == UINT32_MAX) {
                if (!n) break;
                continue;
            }
            out.push_back(v);
        }
        return out;
    }
    // rows whose list holds `target`, ascending (a row listing it twice
    // appears twice)
    std::vector<ct_uint32_t> holders(uint32_t target) const
    {
        std::vector<ct_uint32_t> out;
        for (size_t r = 0; r < rows; r++)
            for (ct_uint32_t v : row(r))
                if (v == target) out.push_back((ct_uint32_t)r);
        return out;
    }
};

class ConView : public IConnectivityParams
{
public:
    explicit ConView(CBMState *s) : s(s) {}
    void showParams(std::ostream &o) override
    {
        for (const auto &kv : paramMap(s->handle(), 1)) o << kv.first << " " << (ct_uint32_t)kv.second << std::endl;
        const cbm_sizes &z = s->sizes();
        o << "numGR " << z.numGR << std::endl << "numGO " << z.numGO << std::endl << "numMF " << z.numMF << std::endl
          << "numSC " << z.numSC << std::endl << "numBC " << z.numBC << std::endl << "numPC " << z.numPC << std::endl
          << "numNC " << z.numNC << std::endl << "numIO " << z.numIO << std::endl << "numGL " << getNumGL()
          << std::endl;
    }
    ct_uint32_t getGOX() override { return s->sizes().goX; }
    ct_uint32_t getGOY() override { return s->sizes().goY; }
    ct_uint32_t getGRX() override { return s->sizes().grX; }
    ct_uint32_t getGRY() override { return s->sizes().grY; }
    // the glomerulus grid of a conParams file with a glomerulus layer (0 without)
    ct_uint32_t getGLX() override { return side("glXP2"); }
    ct_uint32_t getGLY() override { return side("glYP2"); }
    ct_uint32_t getNumMF() override { return s->sizes().numMF; }
    ct_uint32_t getNumGO() override { return s->sizes().numGO; }
    ct_uint32_t getNumGR() override { return s->sizes().numGR; }
    ct_uint32_t getNumGL() override { return getGLX() * getGLY(); }
    ct_uint32_t getNumSC() override { return s->sizes().numSC; }
    ct_uint32_t getNumBC() override { return s->sizes().numBC; }
    ct_uint32_t getNumPC() override { return s->sizes().numPC; }
    ct_uint32_t getNumNC() override { return s->sizes().numNC; }
    ct_uint32_t getNumIO() override { return s->sizes().numIO; }
    std::map<std::string, ct_uint32_t> getParamCopy() override
    {
        std::map<std::string, ct_uint32_t> m;
        for (const auto &kv : paramMap(s->handle(), 1)) m[kv.first] = (ct_uint32_t)kv.second;
        return m;
    }

private:
    ct_uint32_t side(const char *key)
    {
        const auto m = paramMap(s->handle(), 1);
        const auto it = m.find(key);
        return it == m.end() ? 0u : 1u << (ct_uint32_t)it->second;
    }
    CBMState *s;
};

class InNetView : public IInNetConState
{
public:
    explicit InNetView(CBMState *s) : s(s) {}
    // The glomerulus-layer lists are not part of the step's state arrays
    // (the synthetic generator has no glomeruli): empty.
    std::vector<ct_uint32_t> getpGOfromGOtoGLCon(int) override { return {}; }
    std::vector<ct_uint32_t> getpGOfromGLtoGOCon(int) override { return {}; }
    std::vector<ct_uint32_t> getpMFfromMFtoGLCon(int) override { return {}; }
    std::vector<ct_uint32_t> getpGLfromGLtoGRCon(int) override { return {}; }

    std::vector<ct_uint32_t> getpGRfromMFtoGR(int grN) override { return lists(CBM_F_NUMP_GR_FROM_MF_TO_GR, CBM_F_P_GR_FROM_MF_TO_GR).row(idx(grN)); }
    std::vector<std::vector<ct_uint32_t> > getpGRPopfromMFtoGR() override { return all(lists(CBM_F_NUMP_GR_FROM_MF_TO_GR, CBM_F_P_GR_FROM_MF_TO_GR)); }
    std::vector<ct_uint32_t> getpGRfromGOtoGR(int grN) override { return lists(CBM_F_NUMP_GR_FROM_GO_TO_GR, CBM_F_P_GR_FROM_GO_TO_GR).row(idx(grN)); }
    std::vector<std::vector<ct_uint32_t> > getpGRPopfromGOtoGR() override { return all(lists(CBM_F_NUMP_GR_FROM_GO_TO_GR, CBM_F_P_GR_FROM_GO_TO_GR)); }
    // GRs that project to GO goN / receive from GO goN / receive from MF mfN
    std::vector<ct_uint32_t> getpGOfromGRtoGOCon(int goN) override { return lists(CBM_F_NUMP_GR_FROM_GR_TO_GO, CBM_F_P_GR_FROM_GR_TO_GO).holders(idx(goN)); }
    std::vector<ct_uint32_t> getpGOfromGOtoGRCon(int goN) override { return lists(CBM_F_NUMP_GR_FROM_GO_TO_GR, CBM_F_P_GR_FROM_GO_TO_GR).holders(idx(goN)); }
    std::vector<ct_uint32_t> getpMFfromMFtoGRCon(int mfN) override { return lists(CBM_F_NUMP_GR_FROM_MF_TO_GR, CBM_F_P_GR_FROM_MF_TO_GR).holders(idx(mfN)); }
    std::vector<ct_uint32_t> getpMFfromMFtoGOCon(int mfN) override { return lists(CBM_F_NUMP_MF_FROM_MF_TO_GO, CBM_F_P_MF_FROM_MF_TO_GO).row(idx(mfN)); }
    // MFs that project to GO goN
    std::vector<ct_uint32_t> getpGOfromMFtoGOCon(int goN) override { return lists(CBM_F_NUMP_MF_FROM_MF_TO_GO, CBM_F_P_MF_FROM_MF_TO_GO).holders(idx(goN)); }
    std::vector<std::vector<ct_uint32_t> > getpGOPopfromMFtoGOCon() override
    {
        const Lists l = lists(CBM_F_NUMP_MF_FROM_MF_TO_GO, CBM_F_P_MF_FROM_MF_TO_GO);
        std::vector<std::vector<ct_uint32_t> > out(s->sizes().numGO);
        for (size_t mf = 0; mf < l.rows; mf++)
            for (ct_uint32_t go : l.row(mf))
                if (go < out.size()) out[go].push_back((ct_uint32_t)mf);
        return out;
    }
    std::vector<ct_uint32_t> getpGOOutGOGOCon(int goN) override { return lists(CBM_F_NUMP_GO_GABA_OUT_GOGO, CBM_F_P_GO_GABA_OUT_GOGO).row(idx(goN)); }
    // GOs whose GABA out-list holds goN
    std::vector<ct_uint32_t> getpGOInGOGOCon(int goN) override { return lists(CBM_F_NUMP_GO_GABA_OUT_GOGO, CBM_F_P_GO_GABA_OUT_GOGO).holders(idx(goN)); }
    // GOs with fewer GR inputs than the most-connected GO
    std::vector<ct_uint32_t> getGOIncompIndfromGRtoGO() override
    {
        const Lists l = lists(CBM_F_NUMP_GR_FROM_GR_TO_GO, CBM_F_P_GR_FROM_GR_TO_GO);
        std::vector<size_t> cnt(s->sizes().numGO, 0);
        for (size_t gr = 0; gr < l.rows; gr++)
            for (ct_uint32_t go : l.row(gr))
                if (go < cnt.size()) cnt[go]++;
        const size_t full = cnt.empty() ? 0 : *std::max_element(cnt.begin(), cnt.end());
        std::vector<ct_uint32_t> out;
        for (size_t g = 0; g < cnt.size(); g++)
            if (cnt[g] < full) out.push_back((ct_uint32_t)g);
        return out;
    }
    // GRs with fewer GO targets than maxnumpGRfromGRtoGO
    std::vector<ct_uint32_t> getGRIncompIndfromGRtoGO() override
    {
        const Lists l = lists(CBM_F_NUMP_GR_FROM_GR_TO_GO, CBM_F_P_GR_FROM_GR_TO_GO);
        std::vector<ct_uint32_t> out;
        for (size_t gr = 0; gr < l.rows; gr++)
            if ((size_t)l.n[gr] < l.cols) out.push_back((ct_uint32_t)gr);
        return out;
    }
    // Remove destGON from srcGON's GABA out-list (false: not there).
    bool deleteGOGOConPair(int srcGON, int destGON) override
    {
        uint32_t *p;
        int32_t *n;
        size_t cols;
        if (!gaba(srcGON, &p, &n, &cols)) return false;
        for (int32_t j = 0; j < *n; j++)
            if (p[j] == (uint32_t)destGON) {
                for (int32_t k = j; k + 1 < *n; k++) p[k] = p[k + 1];
                p[--*n] = UINT32_MAX;
                return true;
            }
        return false;
    }
    // Append destGON to srcGON's GABA out-list (false: already there, list
    // full or an index out of range).
    bool addGOGOConPair(int srcGON, int destGON) override
    {
        uint32_t *p;
        int32_t *n;
        size_t cols;
        if (!gaba(srcGON, &p, &n, &cols) || destGON < 0 || (uint32_t)destGON >= s->sizes().numGO) return false;
        if ((size_t)*n >= cols) return false;
        for (int32_t j = 0; j < *n; j++)
            if (p[j] == (uint32_t)destGON) return false;
        p[(*n)++] = (uint32_t)destGON;
        return true;
    }

private:
    static size_t idx(int i) { return i < 0 ? SIZE_MAX : (size_t)i; }
    Lists lists(int numField, int listField)
    {
        Lists l;
        size_t bn = 0, bp = 0;
        l.n = static_cast<const int32_t *>(s->field(numField, 0, &bn));
        l.p = static_cast<const uint32_t *>(s->field(listField, 0, &bp));
        l.rows = bn / 4;
        l.cols = l.rows ? bp / 4 / l.rows : 0;
        return l;
    }
    static std::vector<std::vector<ct_uint32_t> > all(const Lists &l)
    {
        std::vector<std::vector<ct_uint32_t> > out(l.rows);
        for (size_t r = 0; r < l.rows; r++) out[r] = l.row(r);
        return out;
    }
    bool gaba(int src, uint32_t **p, int32_t **n, size_t *cols)
    {
        if (src < 0 || (uint32_t)src >= s->sizes().numGO) return false;
        size_t bn = 0, bp = 0;
        int32_t *nn = static_cast<int32_t *>(s->field(CBM_F_NUMP_GO_GABA_OUT_GOGO, 0, &bn));
        uint32_t *pp = static_cast<uint32_t *>(s->field(CBM_F_P_GO_GABA_OUT_GOGO, 0, &bp));
        *cols = bp / 4 / s->sizes().numGO;
        *p = pp + (size_t)src * *cols;
        *n = nn + src;
        return true;
    }
    CBMState *s;
};

class ZoneView : public IMZoneActState
{
public:
    ZoneView(CBMState *s, unsigned int z) : s(s), z(z) {}
    std::vector<float> getGRPCSynWeightLinear() override
    {
        size_t b = 0;
        const float *w = static_cast<const float *>(s->field(CBM_F_PFPC_SYNW, (int)z, &b));
        return std::vector<float>(w, w + b / 4);
    }
    void resetGRPCSynWeight() override
    {
        size_t b = 0;
        float *w = static_cast<float *>(s->field(CBM_F_PFPC_SYNW, (int)z, &b));
        const auto m = paramMap(s->handle(), 0);
        const auto it = m.find("pfSynWInitPC");
        std::fill(w, w + b / 4, it == m.end() ? 0.0f : (float)it->second);
    }

private:
    CBMState *s;
    unsigned int z;
};

}  // namespace

class CBMStateActView : public IActivityParams
{
public:
    explicit CBMStateActView(CBMState *s) : s(s) {}
    unsigned int getMSPerTimeStep() override { return (unsigned int)s->params().msPerTimeStep; }
    void showParams(std::ostream &o) override
    {
        for (const auto &kv : paramMap(s->handle(), 0)) o << kv.first << " " << (float)kv.second << std::endl;
    }
    std::map<std::string, float> getParamCopy() override
    {
        std::map<std::string, float> m;
        for (const auto &kv : paramMap(s->handle(), 0)) m[kv.first] = (float)kv.second;
        return m;
    }
    float getParam(std::string name) override
    {
        const auto m = paramMap(s->handle(), 0);
        const auto it = m.find(name);
        return it == m.end() ? 0.0f : (float)it->second;
    }
    bool setParam(std::string name, float value) override
    {
        if (cbm_state_set_act_param(s->handle(), name.c_str(), value) != CBM_OK) return false;
        cbm_state_params(s->handle(), &s->ap);
        return true;
    }

private:
    CBMState *s;
};

IConnectivityParams *CBMState::getConnectivityParams()
{
    if (!conView) conView = new ConView(this);
    return conView;
}

IActivityParams *CBMState::getActivityParams()
{
    if (!actView) actView = new CBMStateActView(this);
    return actView;
}

IInNetConState *CBMState::getInnetConState()
{
    if (!innetView) innetView = new InNetView(this);
    return innetView;
}

IMZoneActState *CBMState::getMZoneActState(unsigned int zoneN)
{
    if (zoneN >= sz.numZones) throw std::out_of_range("CBMState::getMZoneActState: zone out of range");
    if (!zoneView[zoneN]) zoneView[zoneN] = new ZoneView(this, zoneN);
    return zoneView[zoneN];
}

void CBMState::deleteViews()
{
    delete conView;
    delete actView;
    delete innetView;
    for (IMZoneActState *&v : zoneView) {
        delete v;
        v = nullptr;
    }
    conView = nullptr;
    actView = nullptr;
    innetView = nullptr;
}

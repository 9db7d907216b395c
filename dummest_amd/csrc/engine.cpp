// Host side of the C ABI: device buffers, the length-sorted work queue, and
// kernel launches on the engine's HIP stream.
//
// Work queue (replaces findFinalSimilaritiesBatched's per-profile lane
// packing, dummest-search.hh:76-285): every pair gets the kernel
// configuration (G lanes per pair, R rows per lane) with the least estimated
// work; pairs of one configuration and one profile are sorted by length and
// cut into units of 64/G pairs; units run longest first on a persistent grid
// that pulls them from an atomic counter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "dm_layout.h"
#include "dm_host.h"
#include "../../include/dummest_amd.h"

using namespace dm;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max(bytes, (size_t)256);
        hipError_t r = hipMalloc(&p, want);
        if (r == hipSuccess) cap = want;
        return r;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return (T *)p; }
};

constexpr int kChunkPairs = 1 << 18;          // pairs per queue chunk
constexpr int64_t kScratchBudget = 8LL << 30;  // bytes of W1 scratch per launch

}  // namespace

struct dm_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    double rates[6] = {0.0171, 0.0018, 0.0328, 0.0083, 0.0005, 0.046875};
    std::vector<HostProfile> profiles;
    bool profDirty = true;
    DevBuf dProfs, dParams, dEmis, dBg;
    std::vector<uint8_t> codes;
    std::vector<int64_t> seqOff;
    std::vector<int32_t> seqLen;
    bool seqDirty = true;
    DevBuf dCodes;
    DevBuf dPairs, dBrec, dFrec, dPre, dSuf, dUnits, dCounter, dWbuf, dW0buf, dEbuf, dFbuf, dOut;
    std::vector<int> cfgG, cfgR;
    bool plain = false;
    int gridOpt = 0;
    int numCUs = 256;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double lastMs = 0;
    int64_t lastCells = 0;

    float bgfs1() const { return (float)rates[0] + (float)rates[3]; }  // INSERT1 + DELETE2 (Float)
    float bgfs2() const { return (float)rates[1] + (float)rates[2]; }  // INSERT2 + DELETE1 (Float)
};

static int fail(dm_engine *e, int code, const std::string &msg) {
    if (e) e->err = msg;
    return code;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t _r = (expr);                                                          \
        if (_r != hipSuccess)                                                            \
            return fail(e, DM_EHIP, std::string(#expr) + ": " + hipGetErrorString(_r)); \
    } while (0)

// ---------------------------------------------------------------------------
// uploads
// ---------------------------------------------------------------------------
static int syncProfiles(dm_engine *e) {
    if (!e->profDirty) return DM_OK;
    std::vector<ProfInfo> info;
    std::vector<float> params, emis, bg;
    for (const HostProfile &p : e->profiles) {
        ProfInfo pi{};
        pi.M = p.M;
        pi.parOff = (int64_t)params.size();
        pi.emOff = (int64_t)emis.size();
        params.insert(params.end(), p.params.begin(), p.params.end());
        for (int i = 0; i <= p.M; i++)
            for (int c = 0; c < kEmStride; c++) emis.push_back(c < kNCode ? p.values[(size_t)i * kWidth + 4 + c] : 0.f);
        for (int c = 0; c < 32; c++) bg.push_back(c < kNCode ? p.bg[4 + c] : 0.f);
        info.push_back(pi);
    }
    HIPCHK(e->dProfs.ensure(sizeof(ProfInfo) * std::max<size_t>(1, info.size())));
    HIPCHK(e->dParams.ensure(sizeof(float) * std::max<size_t>(1, params.size())));
    HIPCHK(e->dEmis.ensure(sizeof(float) * std::max<size_t>(1, emis.size())));
    HIPCHK(e->dBg.ensure(sizeof(float) * std::max<size_t>(1, bg.size())));
    if (!info.empty()) {
        HIPCHK(hipMemcpy(e->dProfs.p, info.data(), sizeof(ProfInfo) * info.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->dParams.p, params.data(), sizeof(float) * params.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->dEmis.p, emis.data(), sizeof(float) * emis.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->dBg.p, bg.data(), sizeof(float) * bg.size(), hipMemcpyHostToDevice));
    }
    e->profDirty = false;
    return DM_OK;
}

static int syncSequences(dm_engine *e) {
    if (!e->seqDirty) return DM_OK;
    HIPCHK(e->dCodes.ensure(std::max<size_t>(1, e->codes.size())));
    if (!e->codes.empty())
        HIPCHK(hipMemcpy(e->dCodes.p, e->codes.data(), e->codes.size(), hipMemcpyHostToDevice));
    e->seqDirty = false;
    return DM_OK;
}

// ---------------------------------------------------------------------------
// configuration choice
// ---------------------------------------------------------------------------
// Estimated wave-steps for one pair: (n + G - 1) steps of R rows, shared by
// 64/G pairs per wave.
static int chooseConfig(const dm_engine *e, int M, int n, bool forcePlain) {
    int best = -1;
    double bestCost = 0;
    for (size_t c = 0; c < e->cfgG.size(); c++) {
        const int G = e->cfgG[c], R = e->cfgR[c];
        if (G * R < M + 1) continue;
        if (forcePlain && G != 64) continue;
        const double cost = (double)(n + G - 1) * R * G / 64.0;
        if (best < 0 || cost < bestCost - 1e-9) {
            best = (int)c;
            bestCost = cost;
        }
    }
    return best;
}

struct Chunk {
    std::vector<int64_t> order;  // caller pair index per chunk position
    std::vector<PairInfo> pairs;
    std::vector<Unit> units;
    std::vector<int> unitCfg;    // cfg per unit (units grouped by cfg)
    std::vector<int> cfgFirstUnit, cfgUnits, cfgMaxN;
    int64_t records = 0;
};

// Build the queue for caller pairs [sorted positions b, e).
static void buildChunk(const dm_engine *E, const std::vector<int64_t> &sorted, const std::vector<int> &cfgOf,
                       const int32_t *prof, const int32_t *seq, size_t b, size_t en, Chunk &C) {
    const int ncfg = (int)E->cfgG.size();
    C = Chunk();
    C.cfgFirstUnit.assign(ncfg, 0);
    C.cfgUnits.assign(ncfg, 0);
    C.cfgMaxN.assign(ncfg, 0);
    // units per cfg
    std::vector<std::vector<std::vector<int64_t>>> unitsByCfg(ncfg);
    size_t i = b;
    while (i < en) {
        const int64_t p0 = sorted[i];
        const int cfg = cfgOf[p0];
        const int P = 64 / E->cfgG[cfg];
        std::vector<int64_t> u;
        u.push_back(p0);
        size_t j = i + 1;
        while (j < en && (int)u.size() < P && cfgOf[sorted[j]] == cfg && prof[sorted[j]] == prof[p0]) {
            u.push_back(sorted[j]);
            j++;
        }
        unitsByCfg[cfg].push_back(std::move(u));
        i = j;
    }
    for (int cfg = 0; cfg < ncfg; cfg++) {
        auto &us = unitsByCfg[cfg];
        const int G = E->cfgG[cfg];
        // longest first (units hold length-sorted pairs, so pair 0 is the longest)
        std::stable_sort(us.begin(), us.end(), [&](const std::vector<int64_t> &x, const std::vector<int64_t> &y) {
            return E->seqLen[seq[x[0]]] > E->seqLen[seq[y[0]]];
        });
        C.cfgFirstUnit[cfg] = (int)C.units.size();
        C.cfgUnits[cfg] = (int)us.size();
        for (auto &u : us) {
            Unit U{};
            U.prof = prof[u[0]];
            U.first = (int)C.pairs.size();
            U.count = (int)u.size();
            U.maxN = 0;
            for (int64_t pidx : u) {
                PairInfo pi{};
                pi.prof = prof[pidx];
                pi.seq = seq[pidx];
                pi.n = E->seqLen[pi.seq];
                pi.seqOff = E->seqOff[pi.seq];
                pi.recOff = C.records + kRPad;
                C.records += pi.n + 2 * kRPad;
                U.maxN = std::max(U.maxN, pi.n);
                C.pairs.push_back(pi);
                C.order.push_back(pidx);
            }
            C.cfgMaxN[cfg] = std::max(C.cfgMaxN[cfg], U.maxN);
            C.units.push_back(U);
            C.unitCfg.push_back(cfg);
        }
        (void)G;
    }
}

static int uploadChunk(dm_engine *e, Chunk &C) {
    const size_t np = C.pairs.size();
    HIPCHK(e->dPairs.ensure(sizeof(PairInfo) * np));
    HIPCHK(e->dUnits.ensure(sizeof(Unit) * C.units.size()));
    HIPCHK(e->dBrec.ensure(sizeof(BRec) * C.records));
    HIPCHK(e->dFrec.ensure(sizeof(FRec) * C.records));
    HIPCHK(e->dPre.ensure(sizeof(ES) * C.records));
    HIPCHK(e->dSuf.ensure(sizeof(ES) * C.records));
    HIPCHK(e->dCounter.ensure(sizeof(int) * std::max<size_t>(64, e->cfgG.size())));
    HIPCHK(hipMemcpyAsync(e->dPairs.p, C.pairs.data(), sizeof(PairInfo) * np, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->dUnits.p, C.units.data(), sizeof(Unit) * C.units.size(), hipMemcpyHostToDevice,
                          e->stream));
    NullArgs na{};
    na.codes = e->dCodes.as<uint8_t>();
    na.bg = e->dBg.as<float>();
    na.pairs = e->dPairs.as<PairInfo>();
    na.brec = e->dBrec.as<BRec>();
    na.frec = e->dFrec.as<FRec>();
    na.scratchPre = e->dPre.as<ES>();
    na.scratchSuf = e->dSuf.as<ES>();
    na.npairs = (int)np;
    na.bgfs1 = e->bgfs1();
    na.bgfs2 = e->bgfs2();
    HIPCHK(launch_null(na, e->stream));
    return DM_OK;
}

static DPArgs baseArgs(dm_engine *e) {
    DPArgs a{};
    a.profs = e->dProfs.as<ProfInfo>();
    a.params = e->dParams.as<float>();
    a.emis = e->dEmis.as<float>();
    a.pairs = e->dPairs.as<PairInfo>();
    a.brec = e->dBrec.as<BRec>();
    a.frec = e->dFrec.as<FRec>();
    return a;
}

// Run filter or score mode over all caller pairs.
static int runPairs(dm_engine *e, int mode, int64_t npairs, const int32_t *prof, const int32_t *seq, int wantBwd,
                    float *filterOut, PairScore *scoreOut) {
    if (npairs < 0 || (npairs > 0 && (!prof || !seq))) return fail(e, DM_EINVAL, "bad pair arrays");
    int rc = syncProfiles(e);
    if (rc) return rc;
    rc = syncSequences(e);
    if (rc) return rc;
    const int nprof = (int)e->profiles.size();
    const int nseq = (int)e->seqLen.size();
    std::vector<int> cfgOf(npairs);
    int64_t cells = 0;
    for (int64_t i = 0; i < npairs; i++) {
        if (prof[i] < 0 || prof[i] >= nprof || seq[i] < 0 || seq[i] >= nseq)
            return fail(e, DM_EINVAL, "pair references an unknown profile or sequence");
        const int M = e->profiles[prof[i]].M, n = e->seqLen[seq[i]];
        if (n < 3) return fail(e, DM_EINVAL, "sequences must have at least 3 letters (dummest-dp.hh:84)");
        cfgOf[i] = chooseConfig(e, M, n, e->plain);
        if (cfgOf[i] < 0) return fail(e, DM_EUNSUP, "profile longer than 1023 positions is not supported");
        cells += (int64_t)(M + 1) * n;
    }
    std::vector<int64_t> sorted(npairs);
    std::iota(sorted.begin(), sorted.end(), 0);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int64_t a, int64_t b) {
        if (cfgOf[a] != cfgOf[b]) return cfgOf[a] < cfgOf[b];
        if (prof[a] != prof[b]) return prof[a] < prof[b];
        return e->seqLen[seq[a]] > e->seqLen[seq[b]];
    });
    double totalMs = 0;
    for (size_t b = 0; b < (size_t)npairs;) {
        const size_t en = std::min((size_t)npairs, b + (e->plain ? 1 : (size_t)kChunkPairs));
        Chunk C;
        buildChunk(e, sorted, cfgOf, prof, seq, b, en, C);
        rc = uploadChunk(e, C);
        if (rc) return rc;
        const size_t np = C.pairs.size();
        if (mode == kFilter) HIPCHK(e->dOut.ensure(sizeof(float) * np));
        else HIPCHK(e->dOut.ensure(sizeof(PairScore) * np));
        HIPCHK(hipMemsetAsync(e->dCounter.p, 0, sizeof(int) * e->cfgG.size(), e->stream));
        HIPCHK(hipEventRecord(e->ev0, e->stream));
        for (size_t cfg = 0; cfg < e->cfgG.size(); cfg++) {
            const int nu = C.cfgUnits[cfg];
            if (nu == 0) continue;
            const int G = e->cfgG[cfg], R = e->cfgR[cfg];
            const int P = 64 / G, ROWS = G * R;
            DPArgs a = baseArgs(e);
            a.units = e->dUnits.as<Unit>() + C.cfgFirstUnit[cfg];
            a.nUnits = nu;
            a.unitCounter = e->dCounter.as<int>() + cfg;
            a.wantBwd = wantBwd;
            int grid = e->gridOpt > 0 ? e->gridOpt : e->numCUs * 8;
            grid = std::min(grid, nu);
            if (mode == kScore) {
                const int64_t perPair = (int64_t)(C.cfgMaxN[cfg] + G - 1) * ROWS;
                a.waveFloats = perPair * P;
                const int64_t perWaveBytes = a.waveFloats * (int64_t)sizeof(float);
                grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, kScratchBudget / perWaveBytes));
                HIPCHK(e->dWbuf.ensure((size_t)(a.waveFloats * grid) * sizeof(float)));
                HIPCHK(e->dEbuf.ensure((size_t)(a.waveFloats * grid / ROWS + 1) * sizeof(int)));
                a.wbuf = e->dWbuf.as<float>();
                a.ebuf = e->dEbuf.as<int>();
                a.scoreOut = e->dOut.as<PairScore>();
            } else {
                a.filterOut = e->dOut.as<float>();
            }
            HIPCHK(launch_dp(G, R, mode, a, grid, e->stream));
        }
        HIPCHK(hipEventRecord(e->ev1, e->stream));
        std::vector<char> host(np * (mode == kFilter ? sizeof(float) : sizeof(PairScore)));
        HIPCHK(hipMemcpyAsync(host.data(), e->dOut.p, host.size(), hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        totalMs += ms;
        for (size_t k = 0; k < np; k++) {
            const int64_t dst = C.order[k];
            if (mode == kFilter) filterOut[dst] = ((const float *)host.data())[k];
            else scoreOut[dst] = ((const PairScore *)host.data())[k];
        }
        b = en;
    }
    e->lastMs = totalMs;
    e->lastCells = cells;
    return DM_OK;
}

// Dump the DP matrices of one pair (dump mode, one unit, one wave).
struct PairDump {
    int G, R, M, n;
    std::vector<float> w0, W1;      // [(n+G-1) * G*R]
    std::vector<int> ef, eb;        // [(n+G-1)]
    std::vector<BRec> brec;         // columns [-kRPad, n+kRPad)
    float s1, s2, sC;
};

static int dumpPair(dm_engine *e, int32_t prof, int32_t seq, PairDump &D) {
    if (prof < 0 || prof >= (int)e->profiles.size() || seq < 0 || seq >= (int)e->seqLen.size())
        return fail(e, DM_EINVAL, "unknown profile or sequence");
    int rc = syncProfiles(e);
    if (rc) return rc;
    rc = syncSequences(e);
    if (rc) return rc;
    const int M = e->profiles[prof].M, n = e->seqLen[seq];
    if (n < 3) return fail(e, DM_EINVAL, "sequences must have at least 3 letters (dummest-dp.hh:84)");
    const int cfg = chooseConfig(e, M, n, false);
    if (cfg < 0) return fail(e, DM_EUNSUP, "profile longer than 1023 positions is not supported");
    std::vector<int64_t> sorted{0};
    std::vector<int> cfgOf{cfg};
    Chunk C;
    buildChunk(e, sorted, cfgOf, &prof, &seq, 0, 1, C);
    const int G = e->cfgG[cfg], R = e->cfgR[cfg], ROWS = G * R;
    const int64_t T = n + G - 1;
    C.pairs[0].bufOff = 0;
    rc = uploadChunk(e, C);
    if (rc) return rc;
    HIPCHK(e->dWbuf.ensure(sizeof(float) * T * ROWS));
    HIPCHK(e->dW0buf.ensure(sizeof(float) * T * ROWS));
    HIPCHK(e->dEbuf.ensure(sizeof(int) * T));
    HIPCHK(e->dFbuf.ensure(sizeof(int) * T));
    HIPCHK(hipMemsetAsync(e->dCounter.p, 0, sizeof(int), e->stream));
    DPArgs a = baseArgs(e);
    a.units = e->dUnits.as<Unit>();
    a.nUnits = 1;
    a.unitCounter = e->dCounter.as<int>();
    a.wbuf = e->dWbuf.as<float>();
    a.w0buf = e->dW0buf.as<float>();
    a.ebuf = e->dEbuf.as<int>();
    a.fbuf = e->dFbuf.as<int>();
    HIPCHK(launch_dp(G, R, kDump, a, 1, e->stream));
    D.G = G;
    D.R = R;
    D.M = M;
    D.n = n;
    D.w0.resize(T * ROWS);
    D.W1.resize(T * ROWS);
    D.ef.resize(T);
    D.eb.resize(T);
    D.brec.resize(n + 2 * kRPad);
    PairInfo pi{};
    HIPCHK(hipMemcpyAsync(D.w0.data(), e->dW0buf.p, sizeof(float) * T * ROWS, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(D.W1.data(), e->dWbuf.p, sizeof(float) * T * ROWS, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(D.ef.data(), e->dFbuf.p, sizeof(int) * T, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(D.eb.data(), e->dEbuf.p, sizeof(int) * T, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(D.brec.data(), e->dBrec.as<BRec>(), sizeof(BRec) * D.brec.size(), hipMemcpyDeviceToHost,
                          e->stream));
    HIPCHK(hipMemcpyAsync(&pi, e->dPairs.p, sizeof(PairInfo), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    D.s1 = pi.s1;
    D.s2 = pi.s2;
    D.sC = pi.sC;
    return DM_OK;
}

static inline ES esFromFloat(float v, int off) {
    if (!(v > 0.f)) return ES{0.f, 0};
    int ex;
    float m = std::frexp(v, &ex);
    return ES{m, ex + off};
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int dm_engine_create(int device, dm_engine **out) {
    if (!out) return DM_EINVAL;
    *out = nullptr;
    dm_engine *e = new dm_engine();
    e->device = device;
    hipError_t r = hipSetDevice(device);
    if (r == hipSuccess) r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (r == hipSuccess) r = hipEventCreate(&e->ev0);
    if (r == hipSuccess) r = hipEventCreate(&e->ev1);
    if (r != hipSuccess) {
        delete e;
        return DM_EHIP;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        e->numCUs = cus;
    int gs[64], rs[64];
    const int nc = dp_configs(gs, rs, 64);
    for (int i = 0; i < nc && i < 64; i++) {
        e->cfgG.push_back(gs[i]);
        e->cfgR.push_back(rs[i]);
    }
    *out = e;
    return DM_OK;
}

void dm_engine_destroy(dm_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    for (DevBuf *b : {&e->dProfs, &e->dParams, &e->dEmis, &e->dBg, &e->dCodes, &e->dPairs, &e->dBrec, &e->dFrec,
                      &e->dPre, &e->dSuf, &e->dUnits, &e->dCounter, &e->dWbuf, &e->dW0buf, &e->dEbuf, &e->dFbuf,
                      &e->dOut})
        b->release();
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *dm_engine_error(const dm_engine *e) { return e ? e->err.c_str() : "no engine"; }

int dm_set_rates(dm_engine *e, const double rates[6]) {
    if (!e || !rates) return DM_EINVAL;
    for (int i = 0; i < 6; i++)
        if (!(rates[i] >= 0 && rates[i] < 1)) return fail(e, DM_EINVAL, "rates must lie in [0,1)");
    std::memcpy(e->rates, rates, sizeof(e->rates));
    return DM_OK;
}

int dm_read_hmm(dm_engine *e, const char *path, int bg_type, int32_t *first_id, int32_t *count) {
    if (!e || !path) return DM_EINVAL;
    std::vector<HostProfile> ps;
    std::string msg;
    if (!readHmmFile(path, bg_type, e->rates, ps, msg)) return fail(e, DM_EFORMAT, msg);
    if (first_id) *first_id = (int32_t)e->profiles.size();
    if (count) *count = (int32_t)ps.size();
    for (auto &p : ps) e->profiles.push_back(std::move(p));
    e->profDirty = true;
    return DM_OK;
}

int dm_parse_hmm_text(const char *text, int bg_type, const double rates[6], int32_t index, int32_t *M, float *values,
                      float *params, float *bg, int32_t *count) {
    if (!text || !rates) return DM_EINVAL;
    std::vector<HostProfile> ps;
    std::string msg;
    if (!readHmmText(text, bg_type, rates, ps, msg)) return DM_EFORMAT;
    if (count) *count = (int32_t)ps.size();
    if (index < 0 || index >= (int)ps.size()) return ps.empty() ? DM_EFORMAT : DM_OK;
    const HostProfile &p = ps[index];
    if (M) *M = p.M;
    if (values) std::memcpy(values, p.values.data(), sizeof(float) * p.values.size());
    if (params) std::memcpy(params, p.params.data(), sizeof(float) * p.params.size());
    if (bg) std::memcpy(bg, p.bg.data(), sizeof(float) * p.bg.size());
    return DM_OK;
}

int dm_add_profile(dm_engine *e, int32_t M, const float *values, const float *params, const float *bg, int32_t *id) {
    if (!e || M < 2 || !values || !params || !bg) return DM_EINVAL;
    HostProfile p;
    p.M = M;
    p.values.assign(values, values + (size_t)(M + 1) * kWidth);
    p.params.assign(params, params + (size_t)(M + 2) * 11);
    p.bg.assign(bg, bg + kWidth);
    p.name = "profile" + std::to_string(e->profiles.size());
    p.consensus.assign(M, 0);
    if (id) *id = (int32_t)e->profiles.size();
    e->profiles.push_back(std::move(p));
    e->profDirty = true;
    return DM_OK;
}

int dm_get_profile(const dm_engine *e, int32_t id, int32_t *M, float *values, float *params, float *bg) {
    if (!e || id < 0 || id >= (int)e->profiles.size()) return DM_EINVAL;
    const HostProfile &p = e->profiles[id];
    if (M) *M = p.M;
    if (values) std::memcpy(values, p.values.data(), sizeof(float) * p.values.size());
    if (params) std::memcpy(params, p.params.data(), sizeof(float) * p.params.size());
    if (bg) std::memcpy(bg, p.bg.data(), sizeof(float) * p.bg.size());
    return DM_OK;
}

int dm_profile_count(const dm_engine *e) { return e ? (int)e->profiles.size() : 0; }

int dm_decode(const char *dna, int64_t n, uint8_t *codes) {
    if (n < 0 || (n > 0 && (!dna || !codes))) return DM_EINVAL;
    decodeDna(dna, n, codes);
    return DM_OK;
}

int dm_add_sequences(dm_engine *e, int32_t nseq, const uint8_t *codes, const int64_t *offsets, const int32_t *lengths,
                     int32_t *first_id) {
    if (!e || nseq < 0 || (nseq > 0 && (!codes || !offsets || !lengths))) return DM_EINVAL;
    if (first_id) *first_id = (int32_t)e->seqLen.size();
    for (int32_t i = 0; i < nseq; i++) {
        if (lengths[i] < 0) return fail(e, DM_EINVAL, "negative sequence length");
        for (int32_t j = 0; j < lengths[i]; j++)
            if (codes[offsets[i] + j] >= kNCode) return fail(e, DM_EINVAL, "codon code above 24");
        e->seqOff.push_back((int64_t)e->codes.size());
        e->seqLen.push_back(lengths[i]);
        e->codes.insert(e->codes.end(), codes + offsets[i], codes + offsets[i] + lengths[i]);
    }
    e->seqDirty = true;
    return DM_OK;
}

int dm_clear_sequences(dm_engine *e) {
    if (!e) return DM_EINVAL;
    e->codes.clear();
    e->seqOff.clear();
    e->seqLen.clear();
    e->seqDirty = true;
    return DM_OK;
}

int dm_filter_scores(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, float *best) {
    if (!e || (npairs > 0 && !best)) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    return runPairs(e, kFilter, npairs, prof, seq, 0, best, nullptr);
}

int dm_fwdbwd_scores(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, int want_bwd, float *mid_m,
                     int32_t *mid_e, float *bwd_m, int32_t *bwd_e) {
    if (!e || (npairs > 0 && (!mid_m || !mid_e))) return DM_EINVAL;
    if (want_bwd && npairs > 0 && (!bwd_m || !bwd_e)) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    std::vector<PairScore> out(npairs);
    int rc = runPairs(e, kScore, npairs, prof, seq, want_bwd, nullptr, out.data());
    if (rc) return rc;
    for (int64_t i = 0; i < npairs; i++) {
        mid_m[i] = out[i].midM;
        mid_e[i] = out[i].midE;
        if (want_bwd) {
            bwd_m[i] = out[i].bwdM;
            bwd_e[i] = out[i].bwdE;
        }
    }
    return DM_OK;
}

int dm_pair_matrices(dm_engine *e, int32_t prof, int32_t seq, float *w0_m, int32_t *w0_e, float *W1_m, int32_t *W1_e) {
    if (!e) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    PairDump D;
    int rc = dumpPair(e, prof, seq, D);
    if (rc) return rc;
    const int ROWS = D.G * D.R;
    for (int i = 0; i <= D.M; i++) {
        const int k = i / D.R, q = i - k * D.R;
        for (int j = 0; j < D.n; j++) {
            const int d = j + k;
            const size_t src = (size_t)d * ROWS + (size_t)k * D.R + q;
            const size_t dst = (size_t)i * D.n + j;
            ES a = esFromFloat(D.w0[src], D.ef[d]);
            ES b = esFromFloat(D.W1[src], D.eb[d]);
            if (w0_m) { w0_m[dst] = a.m; w0_e[dst] = a.e; }
            if (W1_m) { W1_m[dst] = b.m; W1_e[dst] = b.e; }
        }
    }
    return DM_OK;
}

int dm_pair_null(dm_engine *e, int32_t prof, int32_t seq, float *prefix_seed, float *suffix_seed, float *scales) {
    if (!e) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    PairDump D;
    int rc = dumpPair(e, prof, seq, D);
    if (rc) return rc;
    for (int j = 0; j < D.n; j++) {
        if (prefix_seed) prefix_seed[j] = D.brec[j + kRPad].ps;
        if (suffix_seed) suffix_seed[j] = D.brec[j + kRPad].ss;
    }
    if (scales) {
        scales[0] = D.s1;
        scales[1] = D.s2;
        scales[2] = D.sC;
    }
    return DM_OK;
}

int dm_pair_hits(dm_engine *e, int32_t prof, int32_t seq, double min_prob_ratio, double *hits, int32_t hit_cap,
                 int32_t *segs, int32_t seg_cap, int32_t *nhits, int32_t *nsegs) {
    if (!e || !hits || !segs || !nhits || !nsegs) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    PairDump D;
    int rc = dumpPair(e, prof, seq, D);
    if (rc) return rc;
    const HostProfile &P = e->profiles[prof];
    PostInput in;
    in.M = D.M;
    in.n = D.n;
    in.R = D.R;
    in.G = D.G;
    in.w0 = D.w0.data();
    in.W1 = D.W1.data();
    in.ef = D.ef.data();
    in.eb = D.eb.data();
    in.codes = e->codes.data() + e->seqOff[seq];
    in.prefixSeed.resize(D.n);
    for (int j = 0; j < D.n; j++) in.prefixSeed[j] = D.brec[j + kRPad].ps;
    in.s1 = D.s1;
    in.s2 = D.s2;
    in.sC = D.sC;
    in.bgfs1 = e->bgfs1();
    in.bgfs2 = e->bgfs2();
    std::vector<Hit> out;
    postProcess(P, in, (float)min_prob_ratio, out);
    int32_t nh = 0, ns = 0;
    for (const Hit &h : out) {
        if (nh >= hit_cap) return fail(e, DM_EINVAL, "hit capacity exceeded");
        double *o = hits + 6 * nh;
        o[0] = h.lnratio;
        o[1] = h.pair.m;
        o[2] = h.pair.e;
        o[3] = h.anchor1;
        o[4] = h.anchor2;
        o[5] = (double)h.segs.size();
        for (const Seg &s : h.segs) {
            if (ns >= seg_cap) return fail(e, DM_EINVAL, "segment capacity exceeded");
            segs[3 * ns] = s.start1;
            segs[3 * ns + 1] = s.start2;
            segs[3 * ns + 2] = s.length;
            ns++;
        }
        nh++;
    }
    *nhits = nh;
    *nsegs = ns;
    return DM_OK;
}

int dm_set_option(dm_engine *e, const char *name, int64_t value) {
    if (!e || !name) return DM_EINVAL;
    if (!std::strcmp(name, "plain")) e->plain = value != 0;
    else if (!std::strcmp(name, "grid")) e->gridOpt = (int)value;
    else return fail(e, DM_EINVAL, std::string("unknown option ") + name);
    return DM_OK;
}

double dm_last_dp_ms(const dm_engine *e) { return e ? e->lastMs : 0.0; }
int64_t dm_last_cells(const dm_engine *e) { return e ? e->lastCells : 0; }

}  // extern "C"

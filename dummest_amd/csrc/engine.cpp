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
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "bw.h"
#include "dm_layout.h"
#include "grad.h"
#include "postgate.h"
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

// Pinned host staging, grown on demand.
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = std::max(bytes, (size_t)4096);
        hipError_t r = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (r == hipSuccess) cap = want;
        return r;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return (T *)p; }
};

constexpr size_t kStaticScratchBytes = (size_t)64 << 20;  // dm_filter_scores_static's scratch (64 MiB)

constexpr int kChunkPairs = 1 << 22;          // pairs per queue chunk (at most)
constexpr int64_t kChunkBytes = 24LL << 30;   // record + null-scratch bytes per chunk (at most)
constexpr int64_t kScratchMin = 8LL << 30;    // bytes of W1 scratch per plan, at least
constexpr int64_t kScratchMax = 64LL << 30;   // and at most (a third of the free memory in between)

}  // namespace

// Device scratch of dm_filter_scores_static (pair and unit descriptors,
// column records, null-model scratch, strip boundaries, counters, staged
// results): one array per GPU, held until the process ends, shared by every
// call on that GPU. Host code never uses this name as an address: it takes
// the device address from hipGetSymbolAddress (scratchAddress below).
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

// The scratch array's device address on each GPU, looked up once per device
// with that device current.
static std::mutex g_scratchMu;
static std::vector<void *> g_scratchAddr;
static hipError_t scratchAddress(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (device < (int)g_scratchAddr.size() && g_scratchAddr[device]) {
        *out = g_scratchAddr[device];
        return hipSuccess;
    }
    void *ptr = nullptr;
    const hipError_t r = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (r != hipSuccess) {
        (void)hipGetLastError();
        return r;
    }
    if (device >= (int)g_scratchAddr.size()) g_scratchAddr.resize(device + 1, nullptr);
    g_scratchAddr[device] = ptr;
    *out = ptr;
    return hipSuccess;
}

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
    int64_t seqEpoch = 0;  // dm_clear_sequences count: sequence ids of an earlier epoch are gone
    DevBuf dCodes;
    DevBuf dPairs, dRec, dNull, dUnits, dCounter, dWbuf, dW0buf, dEbuf, dFbuf, dOut;
    std::vector<int> cfgG, cfgR;
    std::vector<char> cfgStrips;  // row-strip configurations (G = 64, any profile length)
    std::vector<double> cfgStepNs;  // autotuned device ns per wave-step (0: not measured)
    int forceCfg = -1;              // autotuning: every pair runs this configuration
    DevBuf dBnd;
    bool plain = false;
    int gridOpt = 0;
    int64_t chunkPairs = kChunkPairs;  // pairs per plan chunk (option "chunk_pairs")
    int numCUs = 256;
    void *scratchDev = nullptr;    // this device's address of the static scratch array
    std::string scratchErr;        // why it could not be looked up
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double lastMs = 0;
    int64_t lastCells = 0;
    std::vector<double> cfgMs;
    std::vector<int64_t> cfgCells, cfgLaunches;
    double lastTantanMs = 0;  // device time of the last dm_tantan_dna kernel
    int maskBits = 0;       // -m: 1 = profiles, 2 = sequences
    float maskThr = 0.5f;   // --tantan-threshold (TANTAN_MASK_THRESHOLD, a Float)
    double filterSd = 0;    // -d / -D: Gaussian letter filter of profiles read afterwards (0 = off)
    bool keepNonvarying = false;
    DevBuf dTanBases, dTanProbs, dTanScales, dTanMeta, dTanB2f;
    // profile builder's Baum-Welch (dm_bw_prepare / dm_bw_estep)
    int bwL = -1, bwA = 0, bwN = 0;
    std::vector<int> bwLen;         // sequence lengths
    std::vector<int> bwBatch;       // first sequence of each batch, then nseq
    int64_t bwMatCells = 0;         // cells of the largest batch (per matrix)
    int64_t bwBudget = (int64_t)8 << 30;  // bytes of forward/backward matrices per batch
    DevBuf dBwLetters, dBwMeta, dBwTables, dBwMat, dBwRaw, dBwSeqW, dBwOut;
    double lastBwMs = 0;
    // gradient of the summed log-likelihood (dm_grad)
    int64_t gradBudget = (int64_t)8 << 30;  // bytes of FP64 matrices per launch batch
    DevBuf dGradCodon, dGradMat, dGradBnd, dGradMeta, dGradRaw, dGradW, dGradOut, dGradExp;
    std::vector<ProfInfo> profInfo;  // device offsets of each profile's tables (last full upload)
    // registered sequence sets (dm_grad_set_create): consecutive sequences of
    // the store and their codon indices on the device
    struct GradSet {
        int32_t s0 = -1, n = 0;
        size_t codes0 = 0;
        DevBuf dCodon;
    };
    std::vector<GradSet> gradSets;
    double lastGradMs = 0;
    int64_t lastGradCells = 0;
    int64_t lastGradSegments = 0;  // strip segments recomputed from edge rows by the last gradient call
    // hits of the last dm_pair_hits call, kept so that a call repeated with
    // larger buffers after DM_ERANGE does not run the pair again
    struct HitCache {
        int32_t prof = -1, seq = -1;
        double ratio = 0;
        int64_t seqEpoch = -1, profGen = -1;
        std::vector<Hit> hits;
    } hitCache;
    // hits of the last dm_pairs_hits call (the same, for a whole pair list)
    struct HitsCache {
        std::vector<int32_t> prof, seq;
        std::vector<double> ratio;
        int64_t seqEpoch = -1, profGen = -1;
        std::vector<std::vector<Hit>> hits;
    } hitsCache;
    // batched post-gate path (dm_pairs_hits): dump matrices, tables, walk output
    DevBuf dPgW0, dPgW1, dPgEf, dPgEb, dPgPairs, dPgRes, dPgMat, dPgCol, dPgInt, dPgSeg, dPgHit, dPgCand;
    HostBuf hPgCand, hPgHit, hPgSeg;  // their host staging
    double lastHitsMs = 0;    // device time of the last dm_pairs_hits call
    int64_t lastHitsPairs = 0;
    int64_t profGen = 0;  // profile edits (a cached hit list of an older generation is stale)

    float profileMaskThr() const { return (maskBits & 1) ? maskThr : -1.f; }

    float bgfs1() const { return (float)rates[0] + (float)rates[3]; }  // INSERT1 + DELETE2 (Float)
    float bgfs2() const { return (float)rates[1] + (float)rates[2]; }  // INSERT2 + DELETE1 (Float)
};

static int fail(dm_engine *e, int code, const std::string &msg) {
    if (e) e->err = msg;
    return code;
}

// A failing HIP call also clears the runtime's last error: launches report
// hipGetLastError(), which would otherwise hand the same error (say, an
// allocation that did not fit) to the engine's next, unrelated launch.
#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t _r = (expr);                                                          \
        if (_r != hipSuccess) {                                                          \
            (void)hipGetLastError();                                                     \
            return fail(e, DM_EHIP, std::string(#expr) + ": " + hipGetErrorString(_r)); \
        }                                                                                \
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
    e->profInfo = info;
    e->profDirty = false;
    return DM_OK;
}

static int syncSequences(dm_engine *e) {
    if (!e->seqDirty) return DM_OK;
    // The byte after the codes holds the zero-probability code: the null
    // kernel points its loads of columns outside a sequence there.
    HIPCHK(e->dCodes.ensure(e->codes.size() + 1));
    if (!e->codes.empty())
        HIPCHK(hipMemcpy(e->dCodes.p, e->codes.data(), e->codes.size(), hipMemcpyHostToDevice));
    const uint8_t zero = kZeroIdx;
    HIPCHK(hipMemcpy(e->dCodes.as<uint8_t>() + e->codes.size(), &zero, 1, hipMemcpyHostToDevice));
    e->seqDirty = false;
    return DM_OK;
}

// ---------------------------------------------------------------------------
// configuration choice
// ---------------------------------------------------------------------------
// Estimated wave-steps for one pair: (n + G - 1) steps of R rows, shared by
// 64/G pairs per wave.
static int stripsOf(const dm_engine *e, int cfg, int M);

// Cost of one pair under a configuration: wave-steps (n + G - 1 per row
// strip) shared by the 64/G pairs of a wave, times the autotuned time of one
// wave-step when it has been measured, else the rows a lane handles (R).
static double pairCost(const dm_engine *e, int c, int M, int n) {
    const int G = e->cfgG[c], R = e->cfgR[c];
    const double steps = (double)(n + G - 1) * stripsOf(e, c, M) * G / 64.0;
    const double t = (c < (int)e->cfgStepNs.size() && e->cfgStepNs[c] > 0) ? e->cfgStepNs[c] : (double)R;
    return steps * t;
}

// skip: configurations not to choose (none when null); -1 when nothing else fits.
static int chooseConfig(const dm_engine *e, int M, int n, bool forcePlain, const std::vector<char> *skip = nullptr) {
    // a forced configuration (autotuning, the "config" option) applies where its rows fit
    if (e->forceCfg >= 0 && (e->cfgStrips[e->forceCfg] || e->cfgG[e->forceCfg] * e->cfgR[e->forceCfg] >= M + 1))
        return e->forceCfg;
    int best = -1;
    double bestCost = 0;
    const bool tuned = !e->cfgStepNs.empty();
    for (size_t c = 0; c < e->cfgG.size(); c++) {
        const int G = e->cfgG[c], R = e->cfgR[c];
        if (skip && (*skip)[c]) continue;
        if (e->cfgStrips[c] && (!tuned || forcePlain)) continue;  // strips: only when measured (or needed)
        if (!e->cfgStrips[c] && G * R < M + 1) continue;
        if (forcePlain && G != 64) continue;
        const double cost = pairCost(e, (int)c, M, n);
        if (best < 0 || cost < bestCost * (1 - 1e-9)) {
            best = (int)c;
            bestCost = cost;
        }
    }
    if (best < 0)  // profiles longer than one group: the cheapest row-strip kernel
        for (size_t c = 0; c < e->cfgG.size(); c++) {
            if (!e->cfgStrips[c] || (skip && (*skip)[c])) continue;
            const double cost = pairCost(e, (int)c, M, n);
            if (best < 0 || cost < bestCost * (1 - 1e-9)) {
                best = (int)c;
                bestCost = cost;
            }
        }
    return best;
}

static int stripsOf(const dm_engine *e, int cfg, int M) {
    return e->cfgStrips[cfg] ? (M + e->cfgG[cfg] * e->cfgR[cfg]) / (e->cfgG[cfg] * e->cfgR[cfg]) : 1;
}

struct Chunk {
    std::vector<int64_t> order;  // caller pair index per chunk position
    std::vector<PairInfo> pairs;
    std::vector<Unit> units;
    std::vector<int> unitCfg;    // cfg per unit (units grouped by cfg)
    std::vector<int> cfgFirstUnit, cfgUnits, cfgMaxN, cfgMaxS;
    int64_t records = 0;
    int64_t nullScratch = 0;  // ES entries of the null kernel's interleaved scratch
};

// Null-kernel scratch: the pairs of each 64-pair wave share an interleaved
// block of 64 x (longest n + 4) ExpScores, so the suffix values are written
// and read with coalesced accesses.
static void layoutNullScratch(Chunk &C) {
    int64_t off = 0;
    for (size_t w = 0; w < C.pairs.size(); w += 64) {
        const size_t en = std::min(C.pairs.size(), w + 64);
        int maxN = 0;
        for (size_t i = w; i < en; i++) maxN = std::max(maxN, C.pairs[i].n);
        for (size_t i = w; i < en; i++) C.pairs[i].nullOff = off + (int64_t)(i - w);
        off += 64 * (int64_t)(maxN + 4);
    }
    C.nullScratch = std::max<int64_t>(off, 64);
}

// Build the queue for caller pairs [sorted positions b, e). Sequence lengths
// and code offsets come from the engine's store, or from lens / offs (indexed
// by the pairs' sequence numbers) when given.
static void buildChunk(const dm_engine *E, int mode, const std::vector<int64_t> &sorted, const std::vector<int> &cfgOf,
                       const int32_t *prof, const int32_t *seq, size_t b, size_t en, Chunk &C,
                       const int32_t *lens = nullptr, const int64_t *offs = nullptr) {
    if (!lens) {
        lens = E->seqLen.data();
        offs = E->seqOff.data();
    }
    const int ncfg = (int)E->cfgG.size();
    C = Chunk();
    C.cfgFirstUnit.assign(ncfg, 0);
    C.cfgUnits.assign(ncfg, 0);
    C.cfgMaxN.assign(ncfg, 0);
    C.cfgMaxS.assign(ncfg, 1);
    // units per cfg
    std::vector<std::vector<std::vector<int64_t>>> unitsByCfg(ncfg);
    size_t i = b;
    while (i < en) {
        const int64_t p0 = sorted[i];
        const int cfg = cfgOf[p0];
        // a dump unit holds one pair: the wave's scale exponents then follow
        // that pair alone, as when it is dumped on its own
        const int P = mode == kDump ? 1 : 64 / E->cfgG[cfg];
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
        const int WG = dp_wg(G, E->cfgR[cfg], E->cfgStrips[cfg] != 0, mode);
        auto lenOf = [&](const std::vector<int64_t> &x) { return x.empty() ? -1 : lens[seq[x[0]]]; };
        if (WG == 1) {
            // longest first (units hold length-sorted pairs, so pair 0 is the longest)
            std::stable_sort(us.begin(), us.end(), [&](const std::vector<int64_t> &x, const std::vector<int64_t> &y) {
                return lenOf(x) > lenOf(y);
            });
        } else {
            // items of WG units of one profile (the kernel's waves share its
            // emission table), longest items first; a profile's last item is
            // padded with empty units. Units of a profile arrive consecutive
            // and length-sorted, so an item's units have similar lengths.
            std::vector<std::vector<std::vector<int64_t>>> items;
            for (size_t a0 = 0; a0 < us.size();) {
                std::vector<std::vector<int64_t>> it;
                const int32_t pr = prof[us[a0][0]];
                while (a0 < us.size() && (int)it.size() < WG && prof[us[a0][0]] == pr) it.push_back(std::move(us[a0++]));
                items.push_back(std::move(it));
            }
            std::stable_sort(items.begin(), items.end(), [&](const auto &x, const auto &y) {
                return lenOf(x[0]) > lenOf(y[0]);
            });
            us.clear();
            for (auto &it : items) {
                const size_t real = it.size();
                for (size_t w = 0; w < (size_t)WG; w++) us.push_back(w < real ? std::move(it[w]) : std::vector<int64_t>());
            }
        }
        C.cfgFirstUnit[cfg] = (int)C.units.size();
        C.cfgUnits[cfg] = (int)us.size();
        for (size_t ui = 0; ui < us.size(); ui++) {
            auto &u = us[ui];
            Unit U{};
            if (u.empty()) {  // padding of a work item: the partner unit's profile, no pairs
                U = C.units.back();
                U.count = 0;
                U.maxN = 0;
                C.units.push_back(U);
                C.unitCfg.push_back(cfg);
                continue;
            }
            U.prof = prof[u[0]];
            U.first = (int)C.pairs.size();
            U.count = (int)u.size();
            U.maxN = 0;
            for (int64_t pidx : u) {
                PairInfo pi{};
                pi.prof = prof[pidx];
                pi.seq = seq[pidx];
                pi.n = lens[pi.seq];
                pi.seqOff = offs[pi.seq];
                pi.recOff = C.records + kRPad;
                pi.outIdx = (int)pidx;  // plans hold fewer than 2^31 pairs
                C.records += recSpan(pi.n);
                U.maxN = std::max(U.maxN, pi.n);
                C.pairs.push_back(pi);
                C.order.push_back(pidx);
            }
            C.cfgMaxN[cfg] = std::max(C.cfgMaxN[cfg], U.maxN);
            C.cfgMaxS[cfg] = std::max(C.cfgMaxS[cfg], stripsOf(E, cfg, E->profiles[U.prof].M));
            C.units.push_back(U);
            C.unitCfg.push_back(cfg);
        }
        (void)G;
    }
    layoutNullScratch(C);
}

static int uploadChunk(dm_engine *e, Chunk &C) {
    const size_t np = C.pairs.size();
    HIPCHK(e->dPairs.ensure(sizeof(PairInfo) * np));
    HIPCHK(e->dUnits.ensure(sizeof(Unit) * C.units.size()));
    HIPCHK(e->dRec.ensure(sizeof(Rec) * C.records));
    HIPCHK(e->dNull.ensure(sizeof(ES) * C.nullScratch));
    HIPCHK(e->dCounter.ensure(sizeof(int) * std::max<size_t>(64, e->cfgG.size())));
    HIPCHK(hipMemcpyAsync(e->dPairs.p, C.pairs.data(), sizeof(PairInfo) * np, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->dUnits.p, C.units.data(), sizeof(Unit) * C.units.size(), hipMemcpyHostToDevice,
                          e->stream));
    NullArgs na{};
    na.codes = e->dCodes.as<uint8_t>();
    na.zeroCode = e->dCodes.as<uint8_t>() + e->codes.size();
    na.bg = e->dBg.as<float>();
    na.pairs = e->dPairs.as<PairInfo>();
    na.rec = e->dRec.as<Rec>();
    na.scratch = e->dNull.as<ES>();
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
    a.rec = e->dRec.as<Rec>();
    a.bg = e->dBg.as<float>();
    return a;
}

// A plan: the work queue of a pair list, built once and kept on the device.
// Executing it runs the null-model kernel and the DP kernels chunk by chunk.
struct Launch {
    int cfg;
    bool strips;
    int grid;
    DPArgs args;
    int64_t cells;
};

struct DevChunk {
    DevBuf pairs, units;
    int npairs = 0;
    int64_t records = 0;
    std::vector<Launch> launches;
};

// A plan: the work queue of a pair list, built once; its device arrays and
// scratch belong to the plan, so an execution only launches kernels (and can
// be captured into a hipGraph when the plan has one chunk).
struct dm_plan {
    int mode = kScore;
    int wantBwd = 0;
    int64_t npairs = 0;
    int64_t cells = 0;
    int64_t maxChunk = 0;
    int64_t scratchBudget = 0;  // bytes of W1 scratch the plan's launches may use
    std::vector<DevChunk> chunks;
    DevBuf out, rec, nullScratch, counter, wbuf, ebuf, bnd;
    void *hostOut = nullptr;  // pinned staging for results
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    // what the plan was built against: the sequence epoch (its sequence ids)
    // and the length of every profile it uses (configurations, scratch sizes)
    int64_t seqEpoch = 0;
    std::vector<std::pair<int32_t, int32_t>> profM;
    // what a captured graph holds by value: the device tables, the position of
    // the zero-code byte after the codes, the background frameshift rates
    const void *capTabs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t capCodes = 0;
    float capFs1 = 0.f, capFs2 = 0.f;
    void dropGraph() {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        graphExec = nullptr;
        graph = nullptr;
    }
    void release() {
        for (auto &c : chunks) {
            c.pairs.release();
            c.units.release();
        }
        for (DevBuf *b : {&out, &rec, &nullScratch, &counter, &wbuf, &ebuf, &bnd}) b->release();
        dropGraph();
        if (hostOut) (void)hipHostFree(hostOut);
        hostOut = nullptr;
    }
};

static int buildPlan(dm_engine *e, int mode, int64_t npairs, const int32_t *prof, const int32_t *seq, int wantBwd,
                     dm_plan &P) {
    if (npairs < 0 || (npairs > 0 && (!prof || !seq))) return fail(e, DM_EINVAL, "bad pair arrays");
    if (npairs > INT32_MAX) return fail(e, DM_EINVAL, "a plan holds at most 2^31 - 1 pairs");
    int rc = syncProfiles(e);
    if (rc) return rc;
    rc = syncSequences(e);
    if (rc) return rc;
    const int nprof = (int)e->profiles.size();
    const int nseq = (int)e->seqLen.size();
    std::vector<int> cfgOf(npairs);
    P.mode = mode;
    P.wantBwd = wantBwd;
    P.npairs = npairs;
    P.cells = 0;
    P.seqEpoch = e->seqEpoch;
    std::vector<char> used(nprof, 0);
    for (int64_t i = 0; i < npairs; i++) {
        if (prof[i] < 0 || prof[i] >= nprof || seq[i] < 0 || seq[i] >= nseq)
            return fail(e, DM_EINVAL, "pair references an unknown profile or sequence");
        const int M = e->profiles[prof[i]].M, n = e->seqLen[seq[i]];
        if (n < 3) return fail(e, DM_EINVAL, "sequences must have at least 3 letters (dummest-dp.hh:84)");
        cfgOf[i] = chooseConfig(e, M, n, e->plain);
        P.cells += (int64_t)(M + 1) * n;
        if (!used[prof[i]]) {
            used[prof[i]] = 1;
            P.profM.push_back({prof[i], M});
        }
    }
    // A configuration whose pairs would fill fewer than four persistent grids
    // (the waves the device holds at once) leaves SIMDs idle for much of its
    // launch, the last round of units half empty on average: its pairs take
    // the cheapest configuration that already has such a launch, where one fits. Decided over the whole pair list, so every
    // chunk of the plan runs a pair with the same kernel.
    if (!e->plain && e->forceCfg < 0 && !e->cfgStepNs.empty() && npairs > 0) {
        const size_t nc = e->cfgG.size();
        std::vector<int64_t> cnt(nc, 0);
        for (int64_t i = 0; i < npairs; i++) cnt[cfgOf[i]]++;
        const int64_t full = (int64_t)e->numCUs * 8 * 4;
        std::vector<char> skip(nc, 0);
        bool small = false, kept = false;
        for (size_t c = 0; c < nc; c++) {
            const int64_t waves = (cnt[c] * e->cfgG[c] + 63) / 64;
            skip[c] = waves < full;
            if (cnt[c] > 0) (skip[c] ? small : kept) = true;
        }
        if (small && kept)
            for (int64_t i = 0; i < npairs; i++) {
                if (!skip[cfgOf[i]]) continue;
                const int c = chooseConfig(e, e->profiles[prof[i]].M, e->seqLen[seq[i]], false, &skip);
                if (c >= 0) cfgOf[i] = c;
            }
    }
    // the score kernels read a row strip of a wave's W1 blocks through one
    // buffer resource: (64/G) groups x (n + G - 1) steps x G R floats < 2^31 bytes
    if (mode == kScore)
        for (int64_t i = 0; i < npairs; i++)
            if (256LL * e->cfgR[cfgOf[i]] * (e->seqLen[seq[i]] + e->cfgG[cfgOf[i]] - 1) > INT32_MAX)
                return fail(e, DM_EUNSUP, "a sequence is too long for the score pass (at most 2^23 / R - G + 1 "
                                          "letters for a configuration of G lanes and R rows per lane: about 1M at R = 8)");
    std::vector<int64_t> sorted(npairs);
    std::iota(sorted.begin(), sorted.end(), 0);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int64_t a, int64_t b) {
        if (cfgOf[a] != cfgOf[b]) return cfgOf[a] < cfgOf[b];
        if (prof[a] != prof[b]) return prof[a] < prof[b];
        return e->seqLen[seq[a]] > e->seqLen[seq[b]];
    });
    int64_t maxRecords = 1, maxW = 1, maxE = 1, maxB = 1, maxNull = 64;
    const size_t ncfg = e->cfgG.size();
    // W1 scratch: a third of the device's free memory, within [8, 64] GiB. A
    // launch's persistent grid shrinks to what fits, so too small a budget
    // leaves SIMDs without waves (row-strip kernels size every wave for the
    // longest profile of the launch).
    {
        size_t freeB = 0, totalB = 0;
        int64_t budget = kScratchMin;
        if (hipMemGetInfo(&freeB, &totalB) == hipSuccess)
            budget = std::max<int64_t>(kScratchMin, std::min<int64_t>(kScratchMax, (int64_t)(freeB / 3)));
        else
            (void)hipGetLastError();
        P.scratchBudget = budget;
    }
    for (size_t b = 0, en = 0; b < (size_t)npairs; b = en) {
        // chunk: up to chunkPairs pairs whose records and null scratch fit kChunkBytes
        en = b + 1;
        if (!e->plain) {
            int64_t bytes = 0;
            for (en = b; en < (size_t)npairs && en - b < (size_t)e->chunkPairs; en++) {
                const int64_t nn = e->seqLen[seq[sorted[en]]];
                bytes += recSpan((int)nn) * (int64_t)sizeof(Rec) + (nn + 4) * (int64_t)sizeof(ES);
                if (bytes > kChunkBytes && en > b) break;
            }
        }
        Chunk C;
        buildChunk(e, mode, sorted, cfgOf, prof, seq, b, en, C);
        DevChunk D;
        D.npairs = (int)C.pairs.size();
        D.records = C.records;
        std::vector<int64_t> cfgCells(ncfg, 0);
        for (size_t u = 0; u < C.units.size(); u++)
            for (int k = 0; k < C.units[u].count; k++) {
                const PairInfo &pi = C.pairs[C.units[u].first + k];
                cfgCells[C.unitCfg[u]] += (int64_t)(e->profiles[pi.prof].M + 1) * pi.n;
            }
        HIPCHK(D.pairs.ensure(sizeof(PairInfo) * C.pairs.size()));
        HIPCHK(D.units.ensure(sizeof(Unit) * std::max<size_t>(1, C.units.size())));
        HIPCHK(hipMemcpy(D.pairs.p, C.pairs.data(), sizeof(PairInfo) * C.pairs.size(), hipMemcpyHostToDevice));
        if (!C.units.empty())
            HIPCHK(hipMemcpy(D.units.p, C.units.data(), sizeof(Unit) * C.units.size(), hipMemcpyHostToDevice));
        for (size_t cfg = 0; cfg < ncfg; cfg++) {
            const int nu = C.cfgUnits[cfg];
            if (nu == 0) continue;
            Launch L{};
            L.cfg = (int)cfg;
            L.strips = e->cfgStrips[cfg] != 0;
            L.cells = cfgCells[cfg];
            const int G = e->cfgG[cfg], R = e->cfgR[cfg];
            const int Pp = 64 / G, ROWS = G * R;
            DPArgs &a = L.args;
            a.units = D.units.as<Unit>() + C.cfgFirstUnit[cfg];
            a.nUnits = nu;
            a.wantBwd = wantBwd;
            a.stripSteps = C.cfgMaxN[cfg] + G - 1;
            // grid in workgroups of WG waves; scratch per wave
            const int WG = dp_wg(G, R, L.strips, mode);
            int grid = e->gridOpt > 0 ? e->gridOpt : e->numCUs * 8 / WG;
            grid = std::min(grid, nu / WG);
            if (mode == kScore) {
                a.waveFloats = (int64_t)C.cfgMaxS[cfg] * a.stripSteps * ROWS * Pp;
                // exponents: per group and strip, stripSteps entries between 4 guard ints on each side
                a.escGroup = (int64_t)C.cfgMaxS[cfg] * (a.stripSteps + 8);
                const int64_t perWaveBytes = a.waveFloats * (int64_t)sizeof(float) * WG;
                grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, P.scratchBudget / perWaveBytes));
                maxW = std::max(maxW, a.waveFloats * grid * WG);
                maxE = std::max(maxE, a.escGroup * Pp * grid * WG);
            }
            if (L.strips) {
                a.bndCols = C.cfgMaxN[cfg] + 2;
                maxB = std::max(maxB, a.bndCols * grid * WG * Pp);
            }
            L.grid = grid;
            D.launches.push_back(L);
        }
        maxRecords = std::max(maxRecords, C.records);
        maxNull = std::max(maxNull, C.nullScratch);
        P.maxChunk = std::max<int64_t>(P.maxChunk, D.npairs);
        P.chunks.push_back(std::move(D));
    }
    const size_t outBytes = (size_t)std::max<int64_t>(1, npairs) * sizeof(PairScore);
    HIPCHK(P.out.ensure(outBytes));
    HIPCHK(hipHostMalloc(&P.hostOut, outBytes, hipHostMallocDefault));
    HIPCHK(P.rec.ensure(sizeof(Rec) * maxRecords));
    HIPCHK(P.nullScratch.ensure(sizeof(ES) * maxNull));
    HIPCHK(P.counter.ensure(sizeof(int) * std::max<size_t>(64, ncfg)));
    if (mode == kScore) {
        HIPCHK(P.wbuf.ensure(sizeof(float) * maxW));
        // the score kernels read only the W1 runs their backward stored in the
        // same pass; the scratch is zeroed anyway so that no slot can ever hold
        // a non-finite value
        HIPCHK(hipMemsetAsync(P.wbuf.p, 0, P.wbuf.cap, e->stream));
        HIPCHK(P.ebuf.ensure(sizeof(int) * maxE));
    }
    HIPCHK(P.bnd.ensure(sizeof(Bnd) * 3 * maxB));
    for (DevChunk &D : P.chunks)
        for (Launch &L : D.launches) {
            DPArgs &a = L.args;
            a.profs = e->dProfs.as<ProfInfo>();
            a.params = e->dParams.as<float>();
            a.emis = e->dEmis.as<float>();
            a.pairs = D.pairs.as<PairInfo>();
            a.rec = P.rec.as<Rec>();
            a.bg = e->dBg.as<float>();
            a.unitCounter = P.counter.as<int>() + L.cfg;
            a.wbuf = P.wbuf.as<float>();
            a.ebuf = P.ebuf.as<int>();
            a.bnd = P.bnd.as<Bnd>();
            a.scoreOut = P.out.as<PairScore>();
            a.filterOut = P.out.as<float>();
        }
    return DM_OK;
}

// Enqueue one chunk: null model, counter reset, DP launches, result copy.
static int enqueueChunk(dm_engine *e, dm_plan &P, DevChunk &D, std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> *evs) {
    NullArgs na{};
    na.codes = e->dCodes.as<uint8_t>();
    na.zeroCode = e->dCodes.as<uint8_t>() + e->codes.size();
    na.bg = e->dBg.as<float>();
    na.pairs = D.pairs.as<PairInfo>();
    na.rec = P.rec.as<Rec>();
    na.scratch = P.nullScratch.as<ES>();
    na.npairs = D.npairs;
    na.bgfs1 = e->bgfs1();
    na.bgfs2 = e->bgfs2();
    HIPCHK(launch_null(na, e->stream));
    HIPCHK(launch_zero_ints(P.counter.as<int>(), (int)e->cfgG.size(), e->stream));
    for (Launch &L : D.launches) {
        hipEvent_t b0 = nullptr, b1 = nullptr;
        if (evs) {
            HIPCHK(hipEventCreate(&b0));
            HIPCHK(hipEventCreate(&b1));
            HIPCHK(hipEventRecord(b0, e->stream));
        }
        HIPCHK(launch_dp(e->cfgG[L.cfg], e->cfgR[L.cfg], P.mode, L.strips, L.args, L.grid, e->stream));
        if (evs) {
            HIPCHK(hipEventRecord(b1, e->stream));
            evs->push_back({L.cfg, {b0, b1}});
        }
    }
    return DM_OK;
}

// After the last chunk: the results, which the kernels wrote at each pair's
// index in the caller's list, into the pinned staging in one copy.
static int enqueueResults(dm_engine *e, dm_plan &P) {
    const size_t bytes = (size_t)P.npairs * (P.mode == kFilter ? sizeof(float) : sizeof(PairScore));
    if (bytes) HIPCHK(hipMemcpyAsync(P.hostOut, P.out.p, bytes, hipMemcpyDeviceToHost, e->stream));
    return DM_OK;
}

// The staged results of an execution, in the caller's pair order.
static void copyResults(const dm_plan &P, float *filterOut, PairScore *scoreOut) {
    if (P.mode == kFilter && filterOut) std::memcpy(filterOut, P.hostOut, sizeof(float) * (size_t)P.npairs);
    if (P.mode != kFilter && scoreOut) std::memcpy(scoreOut, P.hostOut, sizeof(PairScore) * (size_t)P.npairs);
}

static int executePlan(dm_engine *e, dm_plan &P, float *filterOut, PairScore *scoreOut) {
    const size_t ncfg = e->cfgG.size();
    e->cfgMs.assign(ncfg, 0.0);
    e->cfgCells.assign(ncfg, 0);
    e->cfgLaunches.assign(ncfg, 0);
    double totalMs = 0;
    if (P.graphExec) {
        HIPCHK(hipEventRecord(e->ev0, e->stream));
        HIPCHK(hipGraphLaunch(P.graphExec, e->stream));
        HIPCHK(hipEventRecord(e->ev1, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        totalMs = ms;
        copyResults(P, filterOut, scoreOut);
        e->lastMs = totalMs;
        e->lastCells = P.cells;
        return DM_OK;
    }
    // every chunk is enqueued before the one synchronization: chunks share the
    // device buffers but run in stream order, and each copies its results to
    // its own slots of the host staging
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evs;
    for (DevChunk &D : P.chunks) {
        int rc = enqueueChunk(e, P, D, &evs);
        if (rc) return rc;
    }
    {
        int rc = enqueueResults(e, P);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    for (auto &ev : evs) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev.second.first, ev.second.second));
        e->cfgMs[ev.first] += ms;
        totalMs += ms;
        (void)hipEventDestroy(ev.second.first);
        (void)hipEventDestroy(ev.second.second);
    }
    for (DevChunk &D : P.chunks) {
        for (const Launch &L : D.launches) {
            e->cfgCells[L.cfg] += L.cells;
            e->cfgLaunches[L.cfg] += 1;
        }
    }
    copyResults(P, filterOut, scoreOut);
    e->lastMs = totalMs;
    e->lastCells = P.cells;
    return DM_OK;
}

// The device tables a graph holds by value, as they are now.
static void captureKey(const dm_engine *e, const void *tabs[5], size_t &codes, float &fs1, float &fs2) {
    tabs[0] = e->dProfs.p;
    tabs[1] = e->dParams.p;
    tabs[2] = e->dEmis.p;
    tabs[3] = e->dBg.p;
    tabs[4] = e->dCodes.p;
    codes = e->codes.size();
    fs1 = e->bgfs1();
    fs2 = e->bgfs2();
}

static int capturePlan(dm_engine *e, dm_plan &P);

// Before a plan runs (the engine's tables in sync): a plan whose sequences were
// cleared, or one of whose profiles changed length, no longer describes its
// pairs (DM_EINVAL); otherwise the launches take the current table addresses
// (an upload after the plan was built may have moved them), and a captured
// graph recorded against other tables or rates is captured again.
static int refreshPlan(dm_engine *e, dm_plan &P) {
    if (P.seqEpoch != e->seqEpoch)
        return fail(e, DM_EINVAL, "the engine's sequences were cleared after the plan was created");
    for (const auto &pm : P.profM)
        if (pm.first >= (int32_t)e->profiles.size() || e->profiles[pm.first].M != pm.second)
            return fail(e, DM_EINVAL, "a profile of the plan changed length after the plan was created");
    for (DevChunk &D : P.chunks)
        for (Launch &L : D.launches) {
            L.args.profs = e->dProfs.as<ProfInfo>();
            L.args.params = e->dParams.as<float>();
            L.args.emis = e->dEmis.as<float>();
            L.args.bg = e->dBg.as<float>();
        }
    if (P.graphExec) {
        const void *tabs[5];
        size_t codes;
        float fs1, fs2;
        captureKey(e, tabs, codes, fs1, fs2);
        if (std::memcmp(tabs, P.capTabs, sizeof(tabs)) != 0 || codes != P.capCodes || fs1 != P.capFs1 ||
            fs2 != P.capFs2) {
            P.dropGraph();
            return capturePlan(e, P);
        }
    }
    return DM_OK;
}

// Capture a plan's launches (all chunks, in stream order) into a hipGraph;
// later executions replay the graph.
static int capturePlan(dm_engine *e, dm_plan &P) {
    if (P.graphExec) return DM_OK;
    captureKey(e, P.capTabs, P.capCodes, P.capFs1, P.capFs2);
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = DM_OK;
    for (DevChunk &D : P.chunks) {
        rc = enqueueChunk(e, P, D, nullptr);
        if (rc) break;
    }
    if (rc == DM_OK) rc = enqueueResults(e, P);
    hipGraph_t g = nullptr;
    hipError_t r = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (r != hipSuccess) return fail(e, DM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(r));
    P.graph = g;
    HIPCHK(hipGraphInstantiate(&P.graphExec, g, nullptr, nullptr, 0));
    return DM_OK;
}

static int runPairs(dm_engine *e, int mode, int64_t npairs, const int32_t *prof, const int32_t *seq, int wantBwd,
                    float *filterOut, PairScore *scoreOut) {
    dm_plan P;
    int rc = buildPlan(e, mode, npairs, prof, seq, wantBwd, P);
    if (rc == DM_OK) rc = executePlan(e, P, filterOut, scoreOut);
    P.release();
    return rc;
}

// Dump the DP matrices of one pair (dump mode, one unit, one wave).
struct PairDump {
    int G, R, M, n, T;
    std::vector<float> w0, W1;      // [(n+G-1) * G*R]
    std::vector<int> ef, eb;        // [(n+G-1)]
    std::vector<Rec> rec;           // columns [-kRPad, n+kRPad)
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
    std::vector<int64_t> sorted{0};
    std::vector<int> cfgOf{cfg};
    Chunk C;
    buildChunk(e, kDump, sorted, cfgOf, &prof, &seq, 0, 1, C);
    const int G = e->cfgG[cfg], R = e->cfgR[cfg], ROWS = G * R;
    const int S = stripsOf(e, cfg, M);
    const int64_t T = n + G - 1;
    C.pairs[0].bufOff = 0;
    rc = uploadChunk(e, C);
    if (rc) return rc;
    HIPCHK(e->dWbuf.ensure(sizeof(float) * S * T * ROWS));
    HIPCHK(e->dW0buf.ensure(sizeof(float) * S * T * ROWS));
    HIPCHK(e->dEbuf.ensure(sizeof(int) * S * T));
    HIPCHK(e->dFbuf.ensure(sizeof(int) * S * T));
    HIPCHK(launch_zero_ints(e->dCounter.as<int>(), 1, e->stream));
    DPArgs a = baseArgs(e);
    a.units = e->dUnits.as<Unit>();
    a.nUnits = 1;
    a.unitCounter = e->dCounter.as<int>();
    a.wbuf = e->dWbuf.as<float>();
    a.w0buf = e->dW0buf.as<float>();
    a.ebuf = e->dEbuf.as<int>();
    a.fbuf = e->dFbuf.as<int>();
    a.stripSteps = T;
    const bool strips = e->cfgStrips[cfg] != 0;
    if (strips) {
        a.bndCols = n + 2;
        HIPCHK(e->dBnd.ensure((size_t)(3 * a.bndCols * (64 / G)) * sizeof(Bnd)));
        a.bnd = e->dBnd.as<Bnd>();
    }
    HIPCHK(launch_dp(G, R, kDump, strips, a, 1, e->stream));
    D.G = G;
    D.R = R;
    D.M = M;
    D.n = n;
    D.T = (int)T;
    D.w0.resize(S * T * ROWS);
    D.W1.resize(S * T * ROWS);
    D.ef.resize(S * T);
    D.eb.resize(S * T);
    D.rec.resize(n + 2 * kRPad);
    PairInfo pi{};
    HIPCHK(hipMemcpyAsync(D.w0.data(), e->dW0buf.p, sizeof(float) * D.w0.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(D.W1.data(), e->dWbuf.p, sizeof(float) * D.W1.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(D.ef.data(), e->dFbuf.p, sizeof(int) * D.ef.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(D.eb.data(), e->dEbuf.p, sizeof(int) * D.eb.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(D.rec.data(), e->dRec.as<Rec>(), sizeof(Rec) * D.rec.size(), hipMemcpyDeviceToHost,
                          e->stream));
    HIPCHK(hipMemcpyAsync(&pi, e->dPairs.p, sizeof(PairInfo), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    D.s1 = pi.s1;
    D.s2 = pi.s2;
    D.sC = pi.sC;
    return DM_OK;
}

// Cell (i, j) of a dump: strip i / ROWS, lane k = (i % ROWS) / R, step D = j + k.
static inline size_t dumpIndex(const PairDump &D, int i, int j, size_t *step) {
    const int ROWS = D.G * D.R;
    const int s = i / ROWS, ii = i - s * ROWS;
    const int k = ii / D.R, q = ii - k * D.R;
    const size_t st = (size_t)s * D.T + (size_t)(j + k);
    *step = st;
    return st * ROWS + (size_t)k * D.R + q;
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
    {
        const hipError_t sr = scratchAddress(device, &e->scratchDev);
        if (sr != hipSuccess) e->scratchErr = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(sr);
    }
    int gs[64], rs[64];
    const int nc = dp_configs(gs, rs, 64);
    for (int i = 0; i < nc && i < 64; i++) {
        e->cfgG.push_back(gs[i]);
        e->cfgR.push_back(rs[i]);
        e->cfgStrips.push_back(0);
    }
    const int ns = dp_strip_configs(gs, rs, 64);
    for (int i = 0; i < ns && i < 64; i++) {
        e->cfgG.push_back(gs[i]);
        e->cfgR.push_back(rs[i]);
        e->cfgStrips.push_back(1);
    }
    *out = e;
    return DM_OK;
}

void dm_engine_destroy(dm_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    e->dBnd.release();
    for (auto &g : e->gradSets) g.dCodon.release();
    for (DevBuf *b : {&e->dProfs, &e->dParams, &e->dEmis, &e->dBg, &e->dCodes, &e->dPairs, &e->dRec,
                      &e->dNull, &e->dUnits, &e->dCounter, &e->dWbuf, &e->dW0buf, &e->dEbuf, &e->dFbuf,
                      &e->dOut, &e->dTanBases, &e->dTanProbs, &e->dTanScales, &e->dTanMeta, &e->dTanB2f,
                      &e->dBwLetters, &e->dBwMeta, &e->dBwTables, &e->dBwMat, &e->dBwRaw,
                      &e->dBwSeqW, &e->dBwOut, &e->dGradCodon, &e->dGradMat, &e->dGradBnd, &e->dGradMeta,
                      &e->dGradRaw, &e->dGradW, &e->dGradOut, &e->dGradExp, &e->dPgW0, &e->dPgW1,
                      &e->dPgEf, &e->dPgEb, &e->dPgPairs, &e->dPgRes, &e->dPgMat, &e->dPgCol, &e->dPgInt,
                      &e->dPgSeg, &e->dPgHit, &e->dPgCand})
        b->release();
    for (HostBuf *b : {&e->hPgCand, &e->hPgHit, &e->hPgSeg}) b->release();
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
    if (!readHmmFile(path, bg_type, e->rates, ps, msg, e->profileMaskThr(), e->filterSd, e->keepNonvarying))
        return fail(e, DM_EFORMAT, msg);
    if (first_id) *first_id = (int32_t)e->profiles.size();
    if (count) *count = (int32_t)ps.size();
    for (auto &p : ps) e->profiles.push_back(std::move(p));
    e->profDirty = true;
    e->profGen++;
    return DM_OK;
}

int dm_parse_hmm_text(const char *text, int bg_type, const double rates[6], double mask_threshold,
                      double filter_std_dev, int keep_nonvarying, int32_t index, int32_t *M, float *values,
                      float *params, float *bg, int32_t *count, int32_t *masked) {
    if (!text || !rates || mask_threshold > 1) return DM_EINVAL;
    if (filter_std_dev != 0 && !(filter_std_dev >= 2 && filter_std_dev <= 1000)) return DM_EINVAL;
    std::vector<HostProfile> ps;
    std::string msg;
    if (!readHmmText(text, bg_type, rates, ps, msg, mask_threshold >= 0 ? (float)mask_threshold : -1.f,
                     filter_std_dev, keep_nonvarying != 0))
        return DM_EFORMAT;
    if (count) *count = (int32_t)ps.size();
    if (index < 0 || index >= (int)ps.size()) return ps.empty() ? DM_EFORMAT : DM_OK;
    const HostProfile &p = ps[index];
    if (M) *M = p.M;
    if (values) std::memcpy(values, p.values.data(), sizeof(float) * p.values.size());
    if (params) std::memcpy(params, p.params.data(), sizeof(float) * p.params.size());
    if (bg) std::memcpy(bg, p.bg.data(), sizeof(float) * p.bg.size());
    if (masked) *masked = p.masked;
    return DM_OK;
}

int dm_add_raw_profiles(dm_engine *e, int32_t nprof, const int32_t *M, const float *raw, int bg_type,
                        int32_t *first_id) {
    if (!e || nprof < 0 || (nprof > 0 && (!M || !raw))) return DM_EINVAL;
    if (first_id) *first_id = (int32_t)e->profiles.size();
    size_t off = 0;
    for (int32_t k = 0; k < nprof; k++) {
        if (M[k] < 2) return fail(e, DM_EINVAL, "profile shorter than 2 positions");
        HostProfile p;
        p.M = M[k];
        p.name = "profile" + std::to_string(e->profiles.size());
        p.values.assign(raw + off, raw + off + (size_t)(M[k] + 1) * kWidth);
        off += (size_t)(M[k] + 1) * kWidth;
        p.consensus.resize(M[k]);
        for (int i = 0; i < M[k]; i++) {
            const float *r = &p.values[(size_t)i * kWidth + 4];
            p.consensus[i] = (char)(std::max_element(r, r + 20) - r);
        }
        std::string msg;
        if (!finalizeRaw(p, bg_type, e->rates, msg, e->profileMaskThr(), e->filterSd, e->keepNonvarying))
            return fail(e, DM_EFORMAT, msg);
        e->profiles.push_back(std::move(p));
    }
    e->profDirty = true;
    e->profGen++;
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
    e->profGen++;
    return DM_OK;
}

int dm_set_profile(dm_engine *e, int32_t id, int32_t M, const float *values, const float *params, const float *bg) {
    if (!e || M < 2 || !values || !params || !bg) return DM_EINVAL;
    if (id < 0 || id >= (int)e->profiles.size()) return fail(e, DM_EINVAL, "unknown profile");
    HostProfile &p = e->profiles[id];
    const bool inPlace = !e->profDirty && p.M == M && id < (int)e->profInfo.size();
    p.M = M;
    p.values.assign(values, values + (size_t)(M + 1) * kWidth);
    p.params.assign(params, params + (size_t)(M + 2) * 11);
    p.bg.assign(bg, bg + kWidth);
    p.consensus.assign(M, 0);
    p.masked = 0;
    e->profGen++;
    // the device tables count as stale until every copy below has succeeded: a
    // failed copy leaves them dirty, so the next call re-uploads them all
    e->profDirty = true;
    if (!inPlace) return DM_OK;
    // same length, tables in sync: rewrite this profile's slices only (a
    // gradient loop updates one profile per evaluation among many)
    HIPCHK(hipSetDevice(e->device));
    const ProfInfo &pi = e->profInfo[id];
    std::vector<float> emis((size_t)(M + 1) * kEmStride, 0.f), bgv(32, 0.f);
    for (int i = 0; i <= M; i++)
        for (int c = 0; c < kNCode; c++) emis[(size_t)i * kEmStride + c] = p.values[(size_t)i * kWidth + 4 + c];
    for (int c = 0; c < kNCode; c++) bgv[c] = p.bg[4 + c];
    HIPCHK(hipMemcpy(e->dParams.as<float>() + pi.parOff, p.params.data(), sizeof(float) * p.params.size(),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->dEmis.as<float>() + pi.emOff, emis.data(), sizeof(float) * emis.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->dBg.as<float>() + (int64_t)id * 32, bgv.data(), sizeof(float) * 32, hipMemcpyHostToDevice));
    e->profDirty = false;
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

int dm_set_masking(dm_engine *e, int mask, double threshold) {
    if (!e || mask < 0 || mask > 3 || !(threshold >= 0 && threshold <= 1)) return DM_EINVAL;
    e->maskBits = mask;
    e->maskThr = (float)threshold;
    return DM_OK;
}

int dm_set_profile_filter(dm_engine *e, double std_dev, int keep_nonvarying) {
    if (!e) return DM_EINVAL;
    if (std_dev != 0 && !(std_dev >= 2 && std_dev <= 1000))
        return fail(e, DM_EINVAL, "the filter's standard deviation must lie in [2, 1000] (dummest.cc:185-195)");
    e->filterSd = std_dev;
    e->keepNonvarying = keep_nonvarying != 0;
    return DM_OK;
}

int dm_profile_masked(const dm_engine *e, int32_t id) {
    if (!e || id < 0 || id >= (int)e->profiles.size()) return -1;
    return e->profiles[id].masked;
}

// tantan's letter numbers for DNA (makeMaskedSequence, dummest-sequence.hh:206-230):
// A, C, G, T (either case) are 0..3, every other letter counts as A.
static inline uint8_t tantanLetter(char c) {
    switch (c) {
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 0;
    }
}

int dm_tantan_dna(dm_engine *e, int32_t nseq, const char *dna, const int64_t *offsets, const int32_t *lengths,
                  float *probs) {
    if (!e || nseq < 0 || (nseq > 0 && (!dna || !offsets || !lengths || !probs))) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    // sequences of a wave run in lockstep: order them by length, longest first
    std::vector<int32_t> ord(nseq);
    int64_t total = 0, nscale = 0;
    for (int32_t i = 0; i < nseq; i++) {
        if (lengths[i] < 0 || offsets[i] < 0) return fail(e, DM_EINVAL, "negative sequence length or offset");
        ord[i] = i;
        total += lengths[i];
        nscale += lengths[i] / 16;
    }
    if (total == 0) return DM_OK;
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return lengths[a] > lengths[b]; });
    // packed device layout: sequence k (in length order) at pOff[k]
    std::vector<int64_t> meta(3 * (size_t)nseq);  // offsets, scale offsets, lengths (as int64 pairs)
    std::vector<uint8_t> bases(total);
    int64_t at = 0, sat = 0;
    std::vector<int32_t> len32(nseq);
    for (int32_t k = 0; k < nseq; k++) {
        const int32_t i = ord[k];
        meta[k] = at;
        meta[nseq + k] = sat;
        len32[k] = lengths[i];
        for (int32_t j = 0; j < lengths[i]; j++) bases[at + j] = tantanLetter(dna[offsets[i] + j]);
        at += lengths[i];
        sat += lengths[i] / 16;
    }
    const TantanParams tp = tantanParams(false);
    HIPCHK(e->dTanBases.ensure(total));
    HIPCHK(e->dTanProbs.ensure(sizeof(float) * total));
    HIPCHK(e->dTanScales.ensure(sizeof(double) * std::max<int64_t>(1, nscale)));
    HIPCHK(e->dTanMeta.ensure(sizeof(int64_t) * 2 * nseq + sizeof(int32_t) * nseq));
    HIPCHK(e->dTanB2f.ensure(sizeof(double) * kTantanK));
    HIPCHK(hipMemcpyAsync(e->dTanBases.p, bases.data(), total, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->dTanMeta.p, meta.data(), sizeof(int64_t) * 2 * nseq, hipMemcpyHostToDevice, e->stream));
    int32_t *dLen = (int32_t *)((int64_t *)e->dTanMeta.p + 2 * nseq);
    HIPCHK(hipMemcpyAsync(dLen, len32.data(), sizeof(int32_t) * nseq, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->dTanB2f.p, tp.b2f, sizeof(double) * kTantanK, hipMemcpyHostToDevice, e->stream));
    TantanArgs ta{};
    ta.bases = e->dTanBases.as<uint8_t>();
    ta.offsets = e->dTanMeta.as<int64_t>();
    ta.scaleOff = e->dTanMeta.as<int64_t>() + nseq;
    ta.lengths = dLen;
    ta.probs = e->dTanProbs.as<float>();
    ta.scales = e->dTanScales.as<double>();
    ta.b2f = e->dTanB2f.as<double>();
    ta.b2b = tp.b2b;
    ta.f2b = tp.f2b;
    ta.f2f0 = tp.f2f0;
    ta.nseq = nseq;
    if (!e->ev0) HIPCHK(hipEventCreate(&e->ev0));
    if (!e->ev1) HIPCHK(hipEventCreate(&e->ev1));
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    HIPCHK(launch_tantan_dna(ta, e->stream));
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    std::vector<float> packed(total);
    HIPCHK(hipMemcpyAsync(packed.data(), e->dTanProbs.p, sizeof(float) * total, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    float tms = 0.f;
    HIPCHK(hipEventElapsedTime(&tms, e->ev0, e->ev1));
    e->lastTantanMs = tms;
    for (int32_t k = 0; k < nseq; k++) {
        const int32_t i = ord[k];
        std::memcpy(probs + offsets[i], packed.data() + meta[k], sizeof(float) * lengths[i]);
    }
    return DM_OK;
}

int dm_add_dna(dm_engine *e, int32_t nseq, const char *dna, const int64_t *offsets, const int32_t *lengths,
               int32_t *first_id) {
    if (!e || nseq < 0 || (nseq > 0 && (!dna || !offsets || !lengths))) return DM_EINVAL;
    std::vector<int64_t> poff(nseq);
    int64_t total = 0;
    for (int32_t i = 0; i < nseq; i++) {
        if (lengths[i] < 0 || offsets[i] < 0) return fail(e, DM_EINVAL, "negative sequence length or offset");
        poff[i] = total;
        total += lengths[i];
    }
    std::vector<char> letters(std::max<int64_t>(1, total));
    for (int32_t i = 0; i < nseq; i++) std::memcpy(letters.data() + poff[i], dna + offsets[i], lengths[i]);
    if (e->maskBits & 2) {
        // -m 2: letters whose tantan repeat probability reaches the threshold
        // become a non-ACGT letter, so every codon over them decodes to '?'
        // (makeMaskedSequence + decodeSequence, dummest.cc:557-573)
        std::vector<float> probs(std::max<int64_t>(1, total));
        const int rc = dm_tantan_dna(e, nseq, letters.data(), poff.data(), lengths, probs.data());
        if (rc != DM_OK) return rc;
        for (int64_t j = 0; j < total; j++)
            if (!(probs[j] < e->maskThr)) letters[j] = 'N';
    }
    std::vector<uint8_t> codes(std::max<int64_t>(1, total));
    for (int32_t i = 0; i < nseq; i++) decodeDna(letters.data() + poff[i], lengths[i], codes.data() + poff[i]);
    return dm_add_sequences(e, nseq, codes.data(), poff.data(), lengths, first_id);
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
    e->seqEpoch++;
    // registered gradient sets live in the store: they go with it
    if (!e->gradSets.empty()) (void)hipSetDevice(e->device);
    for (auto &g : e->gradSets) {
        g.dCodon.release();
        g.s0 = -1;
    }
    return DM_OK;
}

int dm_filter_scores(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, float *best) {
    if (!e || (npairs > 0 && !best)) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    return runPairs(e, kFilter, npairs, prof, seq, 0, best, nullptr);
}

int dm_filter_scores_static(dm_engine *e, int32_t prof, int32_t nseq, const uint8_t *codes, const int64_t *offsets,
                            const int32_t *lengths, float *best) {
    if (!e || nseq < 0) return DM_EINVAL;
    if (prof < 0 || prof >= (int)e->profiles.size()) return fail(e, DM_EINVAL, "unknown profile");
    if (nseq == 0) return DM_OK;
    if (!codes || !offsets || !lengths || !best) return fail(e, DM_EINVAL, "null array");
    if (!e->scratchDev) return fail(e, DM_EHIP, e->scratchErr);
    if (e->profDirty)
        return fail(e, DM_EINVAL, "the profile tables are not on the device yet (any scoring call or plan uploads "
                                  "them); this call allocates nothing and does not upload them itself");
    for (int32_t s = 0; s < nseq; s++) {
        if (lengths[s] < 3) return fail(e, DM_EINVAL, "sequences must have at least 3 letters (dummest-dp.hh:84)");
        if (offsets[s] < 0) return fail(e, DM_EINVAL, "negative sequence offset");
    }
    HIPCHK(hipSetDevice(e->device));
    // the work queue, as a filter plan builds it, over the caller's sequences
    const int M = e->profiles[prof].M;
    std::vector<int32_t> pp(nseq, prof), ss(nseq);
    std::vector<int> cfgOf(nseq);
    for (int32_t s = 0; s < nseq; s++) {
        ss[s] = s;
        cfgOf[s] = chooseConfig(e, M, lengths[s], false);
    }
    std::vector<int64_t> sorted(nseq);
    std::iota(sorted.begin(), sorted.end(), 0);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int64_t a, int64_t b) {
        if (cfgOf[a] != cfgOf[b]) return cfgOf[a] < cfgOf[b];
        return lengths[a] > lengths[b];
    });
    Chunk C;
    buildChunk(e, kFilter, sorted, cfgOf, pp.data(), ss.data(), 0, (size_t)nseq, C, lengths, offsets);
    const int np = (int)C.pairs.size();
    for (int k = 0; k < np; k++) C.pairs[k].outIdx = k;  // results staged by queue position, scattered below
    // scratch layout, every part on a 256-byte boundary
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t ncfg = e->cfgG.size();
    const size_t oZero = take(1), oCounter = take(sizeof(int) * std::max<size_t>(64, ncfg));
    const size_t oPairs = take(sizeof(PairInfo) * np), oUnits = take(sizeof(Unit) * C.units.size());
    const size_t oOrder = take(sizeof(int) * np), oStage = take(sizeof(float) * np);
    const size_t oRec = take(sizeof(Rec) * C.records), oNull = take(sizeof(ES) * C.nullScratch);
    const size_t oBnd = at;
    if (oBnd > kStaticScratchBytes)
        return fail(e, DM_EUNSUP, "the sequences need " + std::to_string(oBnd) + " bytes of scratch; the static "
                                  "scratch holds " + std::to_string(kStaticScratchBytes) + " (split the call)");
    unsigned char *base = (unsigned char *)e->scratchDev;
    std::vector<int> order(np);
    for (int k = 0; k < np; k++) order[k] = (int)C.order[k];
    static const uint8_t zeroCode = kZeroIdx;
    HIPCHK(hipMemcpyAsync(base + oZero, &zeroCode, 1, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(base + oPairs, C.pairs.data(), sizeof(PairInfo) * np, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(base + oUnits, C.units.data(), sizeof(Unit) * C.units.size(), hipMemcpyHostToDevice,
                          e->stream));
    HIPCHK(hipMemcpyAsync(base + oOrder, order.data(), sizeof(int) * np, hipMemcpyHostToDevice, e->stream));
    NullArgs na{};
    na.codes = codes;
    na.zeroCode = base + oZero;
    na.bg = e->dBg.as<float>();
    na.pairs = (PairInfo *)(base + oPairs);
    na.rec = (Rec *)(base + oRec);
    na.scratch = (ES *)(base + oNull);
    na.npairs = np;
    na.bgfs1 = e->bgfs1();
    na.bgfs2 = e->bgfs2();
    HIPCHK(launch_null(na, e->stream));
    int *counter = (int *)(base + oCounter);
    HIPCHK(launch_zero_ints(counter, (int)ncfg, e->stream));
    for (size_t cfg = 0; cfg < ncfg; cfg++) {
        const int nu = C.cfgUnits[cfg];
        if (nu == 0) continue;
        const int G = e->cfgG[cfg], R = e->cfgR[cfg];
        const bool strips = e->cfgStrips[cfg] != 0;
        const int WG = dp_wg(G, R, strips, kFilter), Pp = 64 / G;
        DPArgs a = baseArgs(e);
        a.pairs = na.pairs;
        a.rec = na.rec;
        a.units = (const Unit *)(base + oUnits) + C.cfgFirstUnit[cfg];
        a.nUnits = nu;
        a.unitCounter = counter + cfg;
        a.filterOut = (float *)(base + oStage);
        a.stripSteps = C.cfgMaxN[cfg] + G - 1;
        int grid = std::max(1, std::min(e->numCUs * 8 / WG, nu / WG));
        if (strips) {
            // boundary records per group: as many waves as the rest of the scratch holds
            a.bndCols = C.cfgMaxN[cfg] + 2;
            const size_t perGrid = sizeof(Bnd) * 3 * (size_t)a.bndCols * WG * Pp;
            const size_t fit = (kStaticScratchBytes - oBnd) / perGrid;
            if (fit < 1)
                return fail(e, DM_EUNSUP, "a sequence is too long for the static scratch's row-strip boundaries");
            grid = (int)std::min<size_t>((size_t)grid, fit);
            a.bnd = (Bnd *)(base + oBnd);
        }
        HIPCHK(launch_dp(G, R, kFilter, strips, a, grid, e->stream));
    }
    HIPCHK(launch_scatter_floats((const float *)(base + oStage), (const int *)(base + oOrder), best, np, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return DM_OK;
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
    for (int i = 0; i <= D.M; i++) {
        for (int j = 0; j < D.n; j++) {
            size_t d;
            const size_t src = dumpIndex(D, i, j, &d);
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
        if (prefix_seed) prefix_seed[j] = D.rec[j + kRPad].ps;
        if (suffix_seed) suffix_seed[j] = D.rec[j + kRPad].ss;
    }
    if (scales) {
        scales[0] = D.s1;
        scales[1] = D.s2;
        scales[2] = D.sC;
    }
    return DM_OK;
}

static int writeHits(dm_engine *e, const std::vector<Hit> &out, double *hits, int32_t hit_cap, int32_t *segs,
                     int32_t seg_cap, int32_t *nhits, int32_t *nsegs);

int dm_post_process(int32_t M, const float *values, const float *params, const float *bg, int32_t n,
                    const uint8_t *codes, const float *w0_m, const int32_t *w0_e, const float *W1_m,
                    const int32_t *W1_e, const float *prefix_seed, const float *scales, const double rates[6],
                    double min_prob_ratio, double *hits, int32_t hit_cap, int32_t *segs, int32_t seg_cap,
                    int32_t *nhits, int32_t *nsegs) {
    if (M < 2 || n < 3 || !values || !params || !bg || !codes || !w0_m || !w0_e || !W1_m || !W1_e ||
        !prefix_seed || !scales || !rates || !hits || !segs || !nhits || !nsegs)
        return DM_EINVAL;
    HostProfile P;
    P.M = M;
    P.values.assign(values, values + (size_t)(M + 1) * kWidth);
    P.params.assign(params, params + (size_t)(M + 2) * 11);
    P.bg.assign(bg, bg + kWidth);
    PostInput in;
    in.M = M;
    in.n = n;
    const size_t cells = (size_t)(M + 1) * n;
    in.w0.resize(cells);
    in.W1.resize(cells);
    for (size_t c = 0; c < cells; c++) {
        in.w0[c] = ES{w0_m[c], w0_e[c]};
        in.W1[c] = ES{W1_m[c], W1_e[c]};
    }
    in.codes = codes;
    in.prefixSeed.assign(prefix_seed, prefix_seed + n);
    in.s1 = scales[0];
    in.s2 = scales[1];
    in.sC = scales[2];
    in.bgfs1 = (float)rates[0] + (float)rates[3];
    in.bgfs2 = (float)rates[1] + (float)rates[2];
    std::vector<Hit> out;
    postProcess(P, in, (float)min_prob_ratio, out);
    return writeHits(nullptr, out, hits, hit_cap, segs, seg_cap, nhits, nsegs);
}

int dm_pair_hits(dm_engine *e, int32_t prof, int32_t seq, double min_prob_ratio, double *hits, int32_t hit_cap,
                 int32_t *segs, int32_t seg_cap, int32_t *nhits, int32_t *nsegs) {
    if (!e || !hits || !segs || !nhits || !nsegs || hit_cap < 0 || seg_cap < 0) return DM_EINVAL;
    auto &hc = e->hitCache;
    if (hc.prof == prof && hc.seq == seq && hc.ratio == min_prob_ratio && hc.seqEpoch == e->seqEpoch &&
        hc.profGen == e->profGen)
        return writeHits(e, hc.hits, hits, hit_cap, segs, seg_cap, nhits, nsegs);
    HIPCHK(hipSetDevice(e->device));
    PairDump D;
    int rc = dumpPair(e, prof, seq, D);
    if (rc) return rc;
    const HostProfile &P = e->profiles[prof];
    PostInput in;
    in.M = D.M;
    in.n = D.n;
    in.w0.resize((size_t)(D.M + 1) * D.n);
    in.W1.resize((size_t)(D.M + 1) * D.n);
    for (int i = 0; i <= D.M; i++) {
        for (int j = 0; j < D.n; j++) {
            size_t d;
            const size_t src = dumpIndex(D, i, j, &d);
            in.w0[(size_t)i * D.n + j] = esFromStored(D.w0[src], D.ef[d]);
            in.W1[(size_t)i * D.n + j] = esFromStored(D.W1[src], D.eb[d]);
        }
    }
    in.codes = e->codes.data() + e->seqOff[seq];
    in.prefixSeed.resize(D.n);
    for (int j = 0; j < D.n; j++) in.prefixSeed[j] = D.rec[j + kRPad].ps;
    in.s1 = D.s1;
    in.s2 = D.s2;
    in.sC = D.sC;
    in.bgfs1 = e->bgfs1();
    in.bgfs2 = e->bgfs2();
    hc.hits.clear();
    postProcess(P, in, (float)min_prob_ratio, hc.hits);
    hc.prof = prof;
    hc.seq = seq;
    hc.ratio = min_prob_ratio;
    hc.seqEpoch = e->seqEpoch;
    hc.profGen = e->profGen;
    return writeHits(e, hc.hits, hits, hit_cap, segs, seg_cap, nhits, nsegs);
}

// The post-gate path of many pairs at once (postgate.hip): per batch of pairs
// whose device buffers fit the budget, the dump-mode DP of every pair (one
// launch per kernel configuration), the per-column pass, flank totals and MEA
// tables on the device; the anchors ranked on the host (rankAnchors); the
// tracebacks on the device. The same numbers as dm_pair_hits pair by pair.
static int pairsHitsRun(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, const double *ratio,
                        std::vector<std::vector<Hit>> &out) {
    out.assign((size_t)npairs, {});
    const int nprof = (int)e->profiles.size(), nseq = (int)e->seqLen.size();
    for (int64_t k = 0; k < npairs; k++) {
        if (prof[k] < 0 || prof[k] >= nprof || seq[k] < 0 || seq[k] >= nseq)
            return fail(e, DM_EINVAL, "pair references an unknown profile or sequence");
        if (e->seqLen[seq[k]] < 3) return fail(e, DM_EINVAL, "sequences must have at least 3 letters (dummest-dp.hh:84)");
    }
    int rc = syncProfiles(e);
    if (rc) return rc;
    rc = syncSequences(e);
    if (rc) return rc;
    const int ncfg = (int)e->cfgG.size();
    std::vector<int> cfgOf((size_t)npairs);
    std::vector<int64_t> need((size_t)npairs);  // device bytes of a pair (dump, tables, columns, walk)
    for (int64_t k = 0; k < npairs; k++) {
        const int M = e->profiles[prof[k]].M, n = e->seqLen[seq[k]];
        const int cfg = chooseConfig(e, M, n, false);
        if (!post_rows_supported(e->cfgR[cfg])) return fail(e, DM_EUNSUP, "post-gate pass: unsupported rows per lane");
        cfgOf[k] = cfg;
        const int64_t ROWS = (int64_t)e->cfgG[cfg] * e->cfgR[cfg], S = stripsOf(e, cfg, M);
        const int64_t steps = S * (n + e->cfgG[cfg] - 1);
        need[k] = steps * ROWS * 8 + steps * 8 + 3 * (int64_t)(M + 2) * (n + 7) * 8 + (int64_t)kPostCols * (n + 8) * 8 +
                  (2 * (int64_t)(n + 8) + M + 2) * 4 + (3 * (int64_t)n + 3 * (M + 4)) * 12 + (int64_t)n * 40;
    }
    int64_t budget = (int64_t)2 << 30;
    {
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) == hipSuccess)
            budget = std::max<int64_t>(budget, std::min<int64_t>((int64_t)32 << 30, (int64_t)(freeB / 4)));
        else
            (void)hipGetLastError();
    }
    if (!e->ev0) HIPCHK(hipEventCreate(&e->ev0));
    if (!e->ev1) HIPCHK(hipEventCreate(&e->ev1));
    double ms = 0;
    // DM_TIMING in the environment: wall time of the phases on stderr
    const bool timing = std::getenv("DM_TIMING") != nullptr;
    double tPrep = 0, tTables = 0, tRank = 0, tWalk = 0, tOut = 0;
    auto now = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
    double t0 = now();
    for (int64_t b = 0, en = 0; b < npairs; b = en) {
        int64_t bytes = 0;
        for (en = b; en < npairs; en++) {
            if (en > b && bytes + need[en] > budget) break;
            bytes += need[en];
        }
        // the batch sorted as a plan sorts its pairs (configuration, profile, longest first)
        std::vector<int64_t> sorted;
        for (int64_t k = b; k < en; k++) sorted.push_back(k);
        std::stable_sort(sorted.begin(), sorted.end(), [&](int64_t x, int64_t y) {
            if (cfgOf[x] != cfgOf[y]) return cfgOf[x] < cfgOf[y];
            if (prof[x] != prof[y]) return prof[x] < prof[y];
            return e->seqLen[seq[x]] > e->seqLen[seq[y]];
        });
        Chunk C;
        buildChunk(e, kDump, sorted, cfgOf, prof, seq, 0, sorted.size(), C);
        const int np = (int)C.pairs.size();
        // dump regions: per configuration, each pair S x T x ROWS values (T = the
        // configuration's longest sequence + G - 1) and S x T exponents
        std::vector<int64_t> wBase(ncfg, 0), eBase(ncfg, 0), cfgT(ncfg, 0);
        std::vector<PostPair> pp(np);
        int64_t wTot = 0, eTot = 0, matTot = 0, colTot = 0, intTot = 0, segTot = 0, hitTot = 0, maxMat = 1;
        int maxN = 3, maxM = 1;
        for (int cfg = 0; cfg < ncfg; cfg++) {
            if (C.cfgUnits[cfg] == 0) continue;
            const int G = e->cfgG[cfg], R = e->cfgR[cfg], ROWS = G * R;
            cfgT[cfg] = C.cfgMaxN[cfg] + G - 1;
            wTot = (wTot + 63) / 64 * 64;
            wBase[cfg] = wTot;
            eBase[cfg] = eTot;
            int64_t off = 0;
            for (int u = C.cfgFirstUnit[cfg]; u < C.cfgFirstUnit[cfg] + C.cfgUnits[cfg]; u++) {
                for (int q = 0; q < C.units[u].count; q++) {
                    const int c = C.units[u].first + q;
                    PairInfo &pi = C.pairs[c];
                    const int M = e->profiles[pi.prof].M;
                    pi.bufOff = off;
                    PostPair &P = pp[c];
                    P.M = M;
                    P.n = pi.n;
                    P.G = G;
                    P.R = R;
                    P.T = (int)cfgT[cfg];
                    P.prof = pi.prof;
                    P.pairIdx = c;
                    P.minRatio = (float)ratio[C.order[c]];
                    P.wOff = wBase[cfg] + off;
                    P.eOff = eBase[cfg] + off / ROWS;
                    off += (int64_t)stripsOf(e, cfg, M) * cfgT[cfg] * ROWS;
                }
            }
            wTot += off;
            eTot += off / ROWS;
        }
        for (int c = 0; c < np; c++) {
            PostPair &P = pp[c];
            P.matOff = matTot;
            P.colOff = colTot;
            P.intOff = intTot;
            P.segCap = 2 * P.n + 2 * (P.M + 4);
            P.segOff = segTot;
            P.hitOff = hitTot;
            P.candOff = hitTot;
            P.ncand = 0;
            const int64_t mc = 3 * (int64_t)(P.M + 2) * (P.n + 7);
            matTot += mc;
            maxMat = std::max(maxMat, mc);
            colTot += (int64_t)kPostCols * (P.n + 8);
            intTot += 2 * (int64_t)(P.n + 8) + P.M + 2;
            segTot += P.segCap;
            hitTot += P.n;
            maxN = std::max(maxN, P.n);
            maxM = std::max(maxM, P.M);
        }
        HIPCHK(e->dPgW0.ensure(sizeof(float) * (size_t)std::max<int64_t>(wTot, 64)));
        HIPCHK(e->dPgW1.ensure(sizeof(float) * (size_t)std::max<int64_t>(wTot, 64)));
        HIPCHK(e->dPgEf.ensure(sizeof(int) * (size_t)std::max<int64_t>(eTot, 64)));
        HIPCHK(e->dPgEb.ensure(sizeof(int) * (size_t)std::max<int64_t>(eTot, 64)));
        HIPCHK(e->dPgPairs.ensure(sizeof(PostPair) * (size_t)np));
        HIPCHK(e->dPgRes.ensure(sizeof(PostRes) * (size_t)np));
        HIPCHK(e->dPgMat.ensure(sizeof(ES) * (size_t)matTot));
        HIPCHK(e->dPgCol.ensure(sizeof(ES) * (size_t)colTot));
        HIPCHK(e->dPgInt.ensure(sizeof(int) * (size_t)intTot));
        HIPCHK(e->dPgSeg.ensure(sizeof(PostSeg) * (size_t)segTot));
        HIPCHK(e->dPgHit.ensure(sizeof(PostHit) * (size_t)hitTot));
        HIPCHK(e->dPgCand.ensure(sizeof(PostCand) * (size_t)hitTot));
        rc = uploadChunk(e, C);  // pair records (with the dump offsets) and the null model
        if (rc) return rc;
        tPrep += now() - t0;
        t0 = now();
        HIPCHK(hipEventRecord(e->ev0, e->stream));
        HIPCHK(launch_zero_ints(e->dCounter.as<int>(), ncfg, e->stream));
        for (int cfg = 0; cfg < ncfg; cfg++) {
            const int nu = C.cfgUnits[cfg];
            if (nu == 0) continue;
            const int G = e->cfgG[cfg], R = e->cfgR[cfg];
            DPArgs a = baseArgs(e);
            a.units = e->dUnits.as<Unit>() + C.cfgFirstUnit[cfg];
            a.nUnits = nu;
            a.unitCounter = e->dCounter.as<int>() + cfg;
            a.wbuf = e->dPgW1.as<float>() + wBase[cfg];
            a.w0buf = e->dPgW0.as<float>() + wBase[cfg];
            a.ebuf = e->dPgEb.as<int>() + eBase[cfg];
            a.fbuf = e->dPgEf.as<int>() + eBase[cfg];
            a.stripSteps = cfgT[cfg];
            const int grid = std::max(1, std::min(e->numCUs * 8, nu));
            const bool strips = e->cfgStrips[cfg] != 0;
            if (strips) {
                a.bndCols = C.cfgMaxN[cfg] + 2;
                HIPCHK(e->dBnd.ensure((size_t)(3 * a.bndCols * (64 / G)) * grid * sizeof(Bnd)));
                a.bnd = e->dBnd.as<Bnd>();
            }
            HIPCHK(launch_dp(G, R, kDump, strips, a, grid, e->stream));
        }
        HIPCHK(hipMemcpyAsync(e->dPgPairs.p, pp.data(), sizeof(PostPair) * np, hipMemcpyHostToDevice, e->stream));
        PostArgs pa{};
        pa.npairs = np;
        pa.pp = e->dPgPairs.as<PostPair>();
        pa.res = e->dPgRes.as<PostRes>();
        pa.profs = e->dProfs.as<ProfInfo>();
        pa.params = e->dParams.as<float>();
        pa.emis = e->dEmis.as<float>();
        pa.bg = e->dBg.as<float>();
        pa.pairs = e->dPairs.as<PairInfo>();
        pa.rec = e->dRec.as<Rec>();
        pa.codes = e->dCodes.as<uint8_t>();
        pa.w0buf = e->dPgW0.as<float>();
        pa.wbuf = e->dPgW1.as<float>();
        pa.fbuf = e->dPgEf.as<int>();
        pa.ebuf = e->dPgEb.as<int>();
        pa.mat = e->dPgMat.as<ES>();
        pa.col = e->dPgCol.as<ES>();
        pa.ints = e->dPgInt.as<int>();
        pa.segs = e->dPgSeg.as<PostSeg>();
        pa.hits = e->dPgHit.as<PostHit>();
        pa.cands = e->dPgCand.as<PostCand>();
        pa.bgfs1 = e->bgfs1();
        pa.bgfs2 = e->bgfs2();
        HIPCHK(launch_post_tables(pa, maxN, maxM, maxMat, e->stream));
        // each column's best cell, to the host for ranking
        std::vector<PostRes> res(np);
        HIPCHK(e->hPgCand.ensure(sizeof(PostCand) * (size_t)hitTot));
        PostCand *cb = e->hPgCand.as<PostCand>();
        HIPCHK(hipMemcpyAsync(res.data(), pa.res, sizeof(PostRes) * np, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipMemcpyAsync(cb, pa.cands, sizeof(PostCand) * hitTot, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipEventRecord(e->ev1, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        float ms1 = 0.f;
        HIPCHK(hipEventElapsedTime(&ms1, e->ev0, e->ev1));
        ms += ms1;
        tTables += now() - t0;
        t0 = now();
        // anchors ranked per pair (independent), on a pool of host threads
        std::vector<std::vector<Hit>> bestOnly(np);
        {
            const int nthr = hostThreads();
            auto rankRange = [&](int c0, int c1) {
                std::vector<ES> colBest;
                std::vector<int> colRow;
                std::vector<Anchor> ranked;
                for (int c = c0; c < c1; c++) {
                    PostPair &P = pp[c];
                    const PostCand *cc = cb + P.candOff;
                    colBest.resize(P.n);
                    colRow.resize(P.n);
                    for (int j = 0; j < P.n; j++) {
                        colBest[j] = ES{cc[j].m, cc[j].e};
                        colRow[j] = cc[j].a1;
                    }
                    if (P.minRatio < 0) {  // the best anchor, no alignment
                        rankAnchors(colBest.data(), colRow.data(), P.n, P.minRatio, ranked);
                        Hit h;
                        h.pair = ranked.front().pair;
                        h.lnratio = h.pair.m == 0.f ? -INFINITY
                                                    : std::log(h.pair.m) + (double)h.pair.e * 0.6931471805599453;
                        h.anchor1 = ranked.front().a1;
                        h.anchor2 = ranked.front().a2;
                        bestOnly[c].push_back(h);
                        continue;
                    }
                    if (!res[c].live) continue;  // below the gate
                    rankAnchors(colBest.data(), colRow.data(), P.n, P.minRatio, ranked);
                    P.ncand = (int)ranked.size();
                    for (size_t k = 0; k < ranked.size(); k++)
                        cb[P.candOff + k] = PostCand{ranked[k].pair.m, ranked[k].pair.e, ranked[k].a1, ranked[k].a2};
                }
            };
            if (nthr == 1 || np < 64) {
                rankRange(0, np);
            } else {
                std::vector<std::thread> pool;
                for (int w = 0; w < nthr; w++) pool.emplace_back(rankRange, (int)((int64_t)np * w / nthr),
                                                                 (int)((int64_t)np * (w + 1) / nthr));
                for (auto &th : pool) th.join();
            }
        }
        bool walk = false;
        for (int c = 0; c < np; c++) {
            if (!bestOnly[c].empty()) out[C.order[c]] = std::move(bestOnly[c]);
            if (pp[c].ncand) walk = true;
        }
        tRank += now() - t0;
        t0 = now();
        if (!walk) continue;
        HIPCHK(hipEventRecord(e->ev0, e->stream));
        HIPCHK(hipMemcpyAsync(e->dPgPairs.p, pp.data(), sizeof(PostPair) * np, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(e->dPgCand.p, cb, sizeof(PostCand) * hitTot, hipMemcpyHostToDevice, e->stream));
        HIPCHK(launch_post_walk(pa, e->stream));
        HIPCHK(e->hPgHit.ensure(sizeof(PostHit) * (size_t)hitTot));
        HIPCHK(e->hPgSeg.ensure(sizeof(PostSeg) * (size_t)segTot));
        PostHit *hb = e->hPgHit.as<PostHit>();
        PostSeg *sb = e->hPgSeg.as<PostSeg>();
        HIPCHK(hipMemcpyAsync(res.data(), pa.res, sizeof(PostRes) * np, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipMemcpyAsync(hb, pa.hits, sizeof(PostHit) * hitTot, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipMemcpyAsync(sb, pa.segs, sizeof(PostSeg) * segTot, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipEventRecord(e->ev1, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipEventElapsedTime(&ms1, e->ev0, e->ev1));
        ms += ms1;
        tWalk += now() - t0;
        t0 = now();
        for (int c = 0; c < np; c++) {
            const PostPair &P = pp[c];
            if (P.ncand == 0) continue;
            if (res[c].overflow) return fail(e, DM_EUNSUP, "post-gate segment buffer overflow");
            for (int h = 0; h < res[c].nhits; h++) {
                const PostHit &ph = hb[P.hitOff + h];
                Hit H;
                H.pair = ES{ph.m, ph.e};
                H.lnratio = std::log(ph.m) + (double)ph.e * 0.6931471805599453;
                H.anchor1 = ph.a1;
                H.anchor2 = ph.a2;
                for (int g = 0; g < ph.nseg; g++) {
                    const PostSeg &q = sb[P.segOff + ph.seg0 + g];
                    H.segs.push_back(Seg{q.start1, q.start2, q.length});
                }
                out[C.order[c]].push_back(std::move(H));
            }
        }
        tOut += now() - t0;
        t0 = now();
    }
    if (timing)
        std::fprintf(stderr, "dm_pairs_hits: %lld pairs: prepare %.3f s, dumps+tables %.3f s, ranking %.3f s, "
                             "tracebacks %.3f s, output %.3f s; device %.1f ms\n",
                     (long long)npairs, tPrep, tTables, tRank, tWalk, tOut, ms);
    e->lastHitsMs = ms;
    e->lastHitsPairs = npairs;
    return DM_OK;
}

int dm_pairs_hits(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, const double *min_prob_ratio,
                  double *hits, int64_t hit_cap, int32_t *segs, int64_t seg_cap, int64_t *first_hit, int64_t *nhits,
                  int64_t *nsegs) {
    if (!e || npairs < 0 || hit_cap < 0 || seg_cap < 0 || !first_hit || !nhits || !nsegs ||
        (npairs > 0 && (!prof || !seq || !min_prob_ratio)) || (hit_cap > 0 && !hits) || (seg_cap > 0 && !segs))
        return DM_EINVAL;
    auto &hc = e->hitsCache;
    const bool cached = hc.seqEpoch == e->seqEpoch && hc.profGen == e->profGen && (int64_t)hc.prof.size() == npairs &&
                        std::equal(hc.prof.begin(), hc.prof.end(), prof) && std::equal(hc.seq.begin(), hc.seq.end(), seq) &&
                        std::equal(hc.ratio.begin(), hc.ratio.end(), min_prob_ratio);
    if (!cached) {
        HIPCHK(hipSetDevice(e->device));
        hc.seqEpoch = -1;
        int rc = pairsHitsRun(e, npairs, prof, seq, min_prob_ratio, hc.hits);
        if (rc) return rc;
        hc.prof.assign(prof, prof + npairs);
        hc.seq.assign(seq, seq + npairs);
        hc.ratio.assign(min_prob_ratio, min_prob_ratio + npairs);
        hc.seqEpoch = e->seqEpoch;
        hc.profGen = e->profGen;
    }
    int64_t needH = 0, needS = 0;
    for (const auto &v : hc.hits) {
        needH += (int64_t)v.size();
        for (const Hit &h : v) needS += (int64_t)h.segs.size();
    }
    *nhits = needH;
    *nsegs = needS;
    if (needH > hit_cap || needS > seg_cap)
        return fail(e, DM_ERANGE, "hit buffers too small: " + std::to_string(needH) + " hits and " +
                                      std::to_string(needS) + " segments needed");
    int64_t nh = 0, ns = 0;
    for (int64_t k = 0; k < npairs; k++) {
        first_hit[k] = nh;
        for (const Hit &h : hc.hits[(size_t)k]) {
            double *o = hits + 6 * nh;
            o[0] = h.lnratio;
            o[1] = h.pair.m;
            o[2] = h.pair.e;
            o[3] = h.anchor1;
            o[4] = h.anchor2;
            o[5] = (double)h.segs.size();
            for (const Seg &g : h.segs) {
                segs[3 * ns] = g.start1;
                segs[3 * ns + 1] = g.start2;
                segs[3 * ns + 2] = g.length;
                ns++;
            }
            nh++;
        }
    }
    first_hit[npairs] = nh;
    return DM_OK;
}

double dm_last_hits_ms(const dm_engine *e) { return e ? e->lastHitsMs : 0.0; }

// Hits and their segments into the caller's arrays. When they do not fit,
// *nhits / *nsegs receive the counts needed and nothing is written (DM_ERANGE).
static int writeHits(dm_engine *e, const std::vector<Hit> &out, double *hits, int32_t hit_cap, int32_t *segs,
                     int32_t seg_cap, int32_t *nhits, int32_t *nsegs) {
    int64_t needH = (int64_t)out.size(), needS = 0;
    for (const Hit &h : out) needS += (int64_t)h.segs.size();
    if (needH > hit_cap || needS > seg_cap) {
        *nhits = (int32_t)std::min<int64_t>(needH, INT32_MAX);
        *nsegs = (int32_t)std::min<int64_t>(needS, INT32_MAX);
        return fail(e, DM_ERANGE, "hit buffers too small: " + std::to_string(needH) + " hits and " +
                                      std::to_string(needS) + " segments needed");
    }
    int32_t nh = 0, ns = 0;
    for (const Hit &h : out) {
        double *o = hits + 6 * nh;
        o[0] = h.lnratio;
        o[1] = h.pair.m;
        o[2] = h.pair.e;
        o[3] = h.anchor1;
        o[4] = h.anchor2;
        o[5] = (double)h.segs.size();
        for (const Seg &s : h.segs) {
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

int dm_plan_create(dm_engine *e, int mode, int64_t npairs, const int32_t *prof, const int32_t *seq, int want_bwd,
                   dm_plan **out) {
    if (!e || !out || (mode != kFilter && mode != kScore)) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    dm_plan *P = new dm_plan();
    int rc = buildPlan(e, mode, npairs, prof, seq, want_bwd, *P);
    if (rc) {
        P->release();
        delete P;
        return rc;
    }
    *out = P;
    return DM_OK;
}

int dm_plan_execute(dm_engine *e, dm_plan *P, float *best, float *mid_m, int32_t *mid_e, float *bwd_m, int32_t *bwd_e) {
    if (!e || !P) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    int rc = syncProfiles(e);
    if (rc) return rc;
    rc = syncSequences(e);
    if (rc) return rc;
    rc = refreshPlan(e, *P);
    if (rc) return rc;
    if (P->mode == kFilter) {
        if (P->npairs > 0 && !best) return DM_EINVAL;
        return executePlan(e, *P, best, nullptr);
    }
    if (P->npairs > 0 && (!mid_m || !mid_e)) return DM_EINVAL;
    rc = executePlan(e, *P, nullptr, nullptr);
    if (rc) return rc;
    const PairScore *out = (const PairScore *)P->hostOut;
    for (int64_t i = 0; i < P->npairs; i++) {
        mid_m[i] = out[i].midM;
        mid_e[i] = out[i].midE;
        if (P->wantBwd && bwd_m && bwd_e) {
            bwd_m[i] = out[i].bwdM;
            bwd_e[i] = out[i].bwdE;
        }
    }
    return DM_OK;
}

int dm_plan_capture(dm_engine *e, dm_plan *P) {
    if (!e || !P) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    int rc = syncProfiles(e);
    if (rc) return rc;
    rc = syncSequences(e);
    if (rc) return rc;
    rc = refreshPlan(e, *P);
    if (rc) return rc;
    return capturePlan(e, *P);
}

void dm_plan_destroy(dm_engine *e, dm_plan *P) {
    if (!P) return;
    if (e) (void)hipSetDevice(e->device);
    P->release();
    delete P;
}

int dm_kernel_stats(const dm_engine *e, int32_t cap, int32_t *G, int32_t *R, int32_t *strips, double *ms,
                    int64_t *cells, int64_t *launches) {
    if (!e) return DM_EINVAL;
    int n = 0;
    for (size_t c = 0; c < e->cfgMs.size(); c++) {
        if (e->cfgLaunches[c] == 0) continue;
        if (n < cap) {
            if (G) G[n] = e->cfgG[c];
            if (R) R[n] = e->cfgR[c];
            if (strips) strips[n] = e->cfgStrips[c];
            if (ms) ms[n] = e->cfgMs[c];
            if (cells) cells[n] = e->cfgCells[c];
            if (launches) launches[n] = e->cfgLaunches[c];
        }
        n++;
    }
    return n;
}

int dm_set_option(dm_engine *e, const char *name, int64_t value) {
    if (!e || !name) return DM_EINVAL;
    if (!std::strcmp(name, "plain")) e->plain = value != 0;
    else if (!std::strcmp(name, "grid")) e->gridOpt = (int)value;
    else if (!std::strcmp(name, "chunk_pairs")) {
        if (value < 1 || value > kChunkPairs) return fail(e, DM_EINVAL, "chunk_pairs must lie in [1, 2^22]");
        e->chunkPairs = value;
    }
    else if (!std::strcmp(name, "grad_budget")) {
        if (value < 1) return fail(e, DM_EINVAL, "grad_budget must be positive");
        e->gradBudget = value;
    }
    else if (!std::strcmp(name, "bw_budget")) {
        if (value < 1) return fail(e, DM_EINVAL, "bw_budget must be positive");
        e->bwBudget = value;
    }
    else if (!std::strcmp(name, "config")) {
        if (value < -1 || value >= (int64_t)e->cfgG.size()) return fail(e, DM_EINVAL, "configuration index out of range");
        e->forceCfg = (int)value;
    }
    else return fail(e, DM_EINVAL, std::string("unknown option ") + name);
    return DM_OK;
}

// The calibration trials of nprof profiles (trial t of a profile seeded with
// 5489 + t, dummest-calibration.hh:436-482), decoded into the engine's
// sequence store after its current sequences: pair i = (profile i / trials,
// sequence seqs0 + i). dropTrials removes them again.
static int checkCalibArgs(dm_engine *e, int32_t nprof, const int32_t *prof, int32_t seq_len, int32_t border,
                          int32_t trials) {
    if (seq_len < 1 || border < 0 || trials < 2 || (int64_t)seq_len + border < 3 ||
        (int64_t)seq_len + border > (1 << 30))
        return fail(e, DM_EINVAL, "calibration needs seq_len >= 1, border >= 0, trials >= 2");
    for (int32_t k = 0; k < nprof; k++)
        if (prof[k] < 0 || prof[k] >= (int32_t)e->profiles.size()) return fail(e, DM_EINVAL, "unknown profile");
    return DM_OK;
}

static void addTrials(dm_engine *e, int32_t nprof, const int32_t *prof, int32_t seq_len, int32_t border,
                      int32_t trials, std::vector<int32_t> &pp, std::vector<int32_t> &ss) {
    const size_t codes0 = e->codes.size(), seqs0 = e->seqLen.size();
    const int len = seq_len + border;
    const int64_t npairs = (int64_t)nprof * trials;
    pp.assign(npairs, 0);
    ss.assign(npairs, 0);
    e->codes.resize(codes0 + (size_t)npairs * len);
    for (int64_t i = 0; i < npairs; i++) {
        pp[i] = prof[i / trials];
        ss[i] = (int32_t)(seqs0 + (size_t)i);
        e->seqOff.push_back((int64_t)(codes0 + (size_t)i * len));
        e->seqLen.push_back(len);
    }
    // the trials are independent (each seeds its own generator): host threads
    // generate and decode them into their slots
    const int nth = (int)std::max<int64_t>(1, std::min<int64_t>(hostThreads(), (npairs + 63) / 64));
    std::vector<std::thread> pool;
    for (int th = 0; th < nth; th++)
        pool.emplace_back([&, th]() {
            std::vector<char> dna(len);
            for (int64_t i = th; i < npairs; i += nth) {
                const HostProfile &P = e->profiles[pp[i]];
                calibSequence(P.bg.data() + 4, (float)e->rates[5], e->bgfs1(), e->bgfs2(), seq_len, border,
                              (int)(i % trials), dna.data());
                decodeDna(dna.data(), len, e->codes.data() + codes0 + (size_t)i * len);
            }
        });
    for (auto &t : pool) t.join();
    e->seqDirty = true;
}

static void dropTrials(dm_engine *e, size_t codes0, size_t seqs0) {
    e->codes.resize(codes0);
    e->seqOff.resize(seqs0);
    e->seqLen.resize(seqs0);
    e->seqDirty = true;
}

int dm_calibrate(dm_engine *e, int32_t nprof, const int32_t *prof, int32_t seq_len, int32_t border, int32_t trials,
                 double *mid_scores, double *flt_scores, double *fits) {
    if (!e || nprof < 1 || !prof || !fits) return DM_EINVAL;
    int rc = checkCalibArgs(e, nprof, prof, seq_len, border, trials);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->device));
    // The trial sequences go into the engine's store for the duration of the
    // call and are removed afterwards.
    const size_t codes0 = e->codes.size(), seqs0 = e->seqLen.size();
    const int64_t npairs = (int64_t)nprof * trials;
    std::vector<int32_t> pp, ss;
    addTrials(e, nprof, prof, seq_len, border, trials, pp, ss);
    std::vector<PairScore> out(npairs);
    // the whole sweep (null model and DP launches of every chunk) is captured
    // as one hipGraph and replayed
    {
        dm_plan P;
        rc = buildPlan(e, kScore, npairs, pp.data(), ss.data(), 1, P);
        if (rc == DM_OK) rc = capturePlan(e, P);
        if (rc == DM_OK) rc = executePlan(e, P, nullptr, out.data());
        P.release();
    }
    dropTrials(e, codes0, seqs0);
    if (rc) return rc;
    auto lnOf = [](float m, int ex) {  // scoreToLog, dummest-score.hh:201-206
        if (m == 0.0f) return -INFINITY * 1.0;
        return (double)std::log(m) + (double)ex * 0.6931471805599453;
    };
    std::vector<double> mid(trials), flt(trials);
    for (int32_t k = 0; k < nprof; k++) {
        for (int32_t t = 0; t < trials; t++) {
            const PairScore &s = out[(size_t)k * trials + t];
            mid[t] = lnOf(s.midM, s.midE);
            flt[t] = lnOf(s.bwdM, s.bwdE);
            if (mid_scores) mid_scores[(size_t)k * trials + t] = mid[t];
            if (flt_scores) flt_scores[(size_t)k * trials + t] = flt[t];
        }
        gumbelFits(mid, (double)seq_len, fits + (size_t)k * 16);
        gumbelFits(flt, (double)seq_len, fits + (size_t)k * 16 + 8);
    }
    return DM_OK;
}

int dm_calibrate_anchors(dm_engine *e, int32_t nprof, const int32_t *prof, int32_t seq_len, int32_t border,
                         int32_t trials, int32_t *anchors, double *mid_scores) {
    if (!e || nprof < 1 || !prof || !anchors) return DM_EINVAL;
    int rc = checkCalibArgs(e, nprof, prof, seq_len, border, trials);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->device));
    const size_t codes0 = e->codes.size(), seqs0 = e->seqLen.size();
    const int64_t npairs = (int64_t)nprof * trials;
    std::vector<int32_t> pp, ss;
    addTrials(e, nprof, prof, seq_len, border, trials, pp, ss);
    // a negative ratio reports the best anchor of each pair and no alignment
    // (dummest-dp-forward.hh:536-543), from the same dumped matrices as hits
    std::vector<double> ratio((size_t)npairs, -2.0);
    std::vector<std::vector<Hit>> out;
    rc = pairsHitsRun(e, npairs, pp.data(), ss.data(), ratio.data(), out);
    dropTrials(e, codes0, seqs0);
    if (rc) return rc;
    for (int64_t i = 0; i < npairs; i++) {
        const bool has = !out[i].empty();
        anchors[2 * i] = has ? out[i][0].anchor1 : 0;
        anchors[2 * i + 1] = has ? out[i][0].anchor2 : 0;
        if (mid_scores) mid_scores[i] = has ? out[i][0].lnratio : -INFINITY;
    }
    return DM_OK;
}

int dm_calib_sequence(const float *bg, const double rates[6], int32_t L, int32_t border, int32_t trial, char *out) {
    if (!bg || !rates || !out || L < 1 || border < 0) return DM_EINVAL;
    const float bgfs1 = (float)rates[0] + (float)rates[3], bgfs2 = (float)rates[1] + (float)rates[2];
    calibSequence(bg + 4, (float)rates[5], bgfs1, bgfs2, L, border, trial, out);
    return DM_OK;
}

int dm_gumbel_fits(const double *scores, int32_t n, double seq_len, double out[8]) {
    if (!scores || !out || n < 2) return DM_EINVAL;
    gumbelFits(std::vector<double>(scores, scores + n), seq_len, out);
    return DM_OK;
}

int dm_autotune(dm_engine *e, int32_t seq_len, double *ns_per_step, int32_t cap) {
    if (!e || seq_len < 3 || seq_len > (1 << 20)) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    const size_t ncfg = e->cfgG.size();
    const size_t prof0 = e->profiles.size(), codes0 = e->codes.size(), seqs0 = e->seqLen.size();
    // Synthetic inputs: per configuration one profile that fills its row
    // group (two strips for the strip kernel) with flat emissions, and random
    // DNA of seq_len letters.
    std::vector<int32_t> cfgProf(ncfg, -1);
    std::string msg;
    for (size_t c = 0; c < ncfg; c++) {
        const int rows = e->cfgG[c] * e->cfgR[c] * (e->cfgStrips[c] ? 2 : 1);
        HostProfile P;
        P.name = "autotune";
        P.M = rows - 1;
        P.values.assign((size_t)rows * kWidth, 0.f);
        for (int i = 0; i < rows; i++) {
            float *v = &P.values[(size_t)i * kWidth];
            v[0] = 0.02f;
            v[1] = 0.5f;
            v[2] = (i < rows - 1) ? 0.02f : 0.f;
            v[3] = (i < rows - 1) ? 0.5f : 0.f;
            if (i < rows - 1)
                for (int a = 0; a < 20; a++) v[4 + a] = 0.03f + 0.004f * (float)((a * 7 + i) % 10);
        }
        P.consensus.assign(rows - 1, 0);
        if (!finalizeRaw(P, 'G', e->rates, msg)) {
            e->profiles.resize(prof0);
            return fail(e, DM_EINVAL, "autotune profile: " + msg);
        }
        cfgProf[c] = (int32_t)e->profiles.size();
        e->profiles.push_back(std::move(P));
    }
    e->profDirty = true;
    const int nseq = 64;
    std::vector<char> dna(seq_len);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (int k = 0; k < nseq; k++) {
        for (int j = 0; j < seq_len; j++) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            dna[j] = "ACGT"[x & 3];
        }
        const size_t at = e->codes.size();
        e->codes.resize(at + seq_len);
        decodeDna(dna.data(), seq_len, e->codes.data() + at);
        e->seqOff.push_back((int64_t)at);
        e->seqLen.push_back(seq_len);
    }
    e->seqDirty = true;
    std::vector<double> tuned(ncfg, 0.0);
    int rc = DM_OK;
    for (size_t c = 0; c < ncfg && rc == DM_OK; c++) {
        const int G = e->cfgG[c];
        const int S = e->cfgStrips[c] ? 2 : 1;
        const int units = e->numCUs * 16;
        const int64_t np = (int64_t)units * (64 / G);
        std::vector<int32_t> pp(np, cfgProf[c]), ss(np);
        for (int64_t i = 0; i < np; i++) ss[i] = (int32_t)(seqs0 + (size_t)(i % nseq));
        e->forceCfg = (int)c;
        dm_plan P;
        rc = buildPlan(e, kScore, np, pp.data(), ss.data(), 0, P);
        std::vector<PairScore> out(np);
        if (rc == DM_OK) rc = executePlan(e, P, nullptr, out.data());  // warm-up
        // median of three timed runs: one slow or fast outlier would move
        // every pair of a length class to the wrong configuration
        double ms[3] = {0, 0, 0};
        for (int r = 0; r < 3 && rc == DM_OK; r++) {
            rc = executePlan(e, P, nullptr, out.data());
            ms[r] = e->cfgMs[c];
        }
        if (rc == DM_OK) {
            std::sort(ms, ms + 3);
            const double waveSteps = (double)units * (double)(seq_len + G - 1) * S;
            tuned[c] = ms[1] * 1e6 / waveSteps;
        }
        P.release();
    }
    e->forceCfg = -1;
    e->profiles.resize(prof0);
    e->profDirty = true;
    e->codes.resize(codes0);
    e->seqOff.resize(seqs0);
    e->seqLen.resize(seqs0);
    e->seqDirty = true;
    if (rc) return rc;
    e->cfgStepNs = tuned;
    if (ns_per_step)
        for (size_t c = 0; c < ncfg && (int32_t)c < cap; c++) ns_per_step[c] = tuned[c];
    return (int)ncfg;
}

int dm_set_config_costs(dm_engine *e, int32_t n, const double *ns_per_step) {
    if (!e || (n > 0 && !ns_per_step)) return DM_EINVAL;
    if (n == 0) {
        e->cfgStepNs.clear();
        return DM_OK;
    }
    if (n != (int32_t)e->cfgG.size()) return fail(e, DM_EINVAL, "cost table size differs from the configuration table");
    for (int32_t c = 0; c < n; c++)
        if (!(ns_per_step[c] > 0)) return fail(e, DM_EINVAL, "costs must be positive");
    e->cfgStepNs.assign(ns_per_step, ns_per_step + n);
    return DM_OK;
}

int dm_config_table(const dm_engine *e, int32_t cap, int32_t *G, int32_t *R, int32_t *strips, double *ns_per_step) {
    if (!e) return DM_EINVAL;
    const size_t ncfg = e->cfgG.size();
    for (size_t c = 0; c < ncfg && (int32_t)c < cap; c++) {
        if (G) G[c] = e->cfgG[c];
        if (R) R[c] = e->cfgR[c];
        if (strips) strips[c] = e->cfgStrips[c];
        if (ns_per_step) ns_per_step[c] = c < e->cfgStepNs.size() ? e->cfgStepNs[c] : 0.0;
    }
    return (int)ncfg;
}

int dm_copy_probe(dm_engine *e, int64_t bytes, double *ms) {
    if (!e || bytes < 16 || (bytes & 15) || !ms) return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    DevBuf src, dst;
    const int rc = [&]() -> int {
        HIPCHK(src.ensure((size_t)bytes));
        HIPCHK(dst.ensure((size_t)bytes));
        HIPCHK(launch_copy_probe(src.as<float>(), dst.as<float>(), bytes / 16, e->stream));  // warm-up
        HIPCHK(hipEventRecord(e->ev0, e->stream));
        HIPCHK(launch_copy_probe(src.as<float>(), dst.as<float>(), bytes / 16, e->stream));
        HIPCHK(hipEventRecord(e->ev1, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e->ev0, e->ev1));
        *ms = t;
        return DM_OK;
    }();
    src.release();
    dst.release();
    return rc;
}

#ifndef DM_BUILD_ID
#define DM_BUILD_ID "unknown"
#endif
const char *dm_build_id(void) { return DM_BUILD_ID; }

double dm_last_dp_ms(const dm_engine *e) { return e ? e->lastMs : 0.0; }
double dm_last_tantan_ms(const dm_engine *e) { return e ? e->lastTantanMs : 0.0; }
double dm_last_bw_ms(const dm_engine *e) { return e ? e->lastBwMs : 0.0; }
double dm_last_grad_ms(const dm_engine *e) { return e ? e->lastGradMs : 0.0; }
int64_t dm_last_grad_cells(const dm_engine *e) { return e ? e->lastGradCells : 0; }
int64_t dm_last_grad_segments(const dm_engine *e) { return e ? e->lastGradSegments : 0; }

// ------------------------------------- gradient of the summed log-likelihood
// Appends DNA sequences to the engine's store (decoded codes for the null model
// and the DP) and returns their codon indices for the emission counts (16 b0 +
// 4 b1 + b2 over ACGT, 64 otherwise), one byte per letter.
static void appendGradSequences(dm_engine *e, int32_t nseq, const char *dna, const int64_t *offsets,
                                const int32_t *lengths, std::vector<uint8_t> &codon) {
    const size_t codes0 = e->codes.size();
    int64_t total = 0;
    for (int32_t s = 0; s < nseq; s++) total += lengths[s];
    e->codes.resize(codes0 + (size_t)total);
    codon.assign((size_t)total, 64);
    auto base4 = [](char c) -> int {
        switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return -1;
        }
    };
    int64_t at = 0;
    for (int32_t s = 0; s < nseq; s++) {
        const char *d = dna + offsets[s];
        const int n = lengths[s];
        decodeDna(d, n, e->codes.data() + codes0 + at);
        for (int j = 0; j + 2 < n; j++) {
            const int x = base4(d[j]), y = base4(d[j + 1]), z = base4(d[j + 2]);
            if (x >= 0 && y >= 0 && z >= 0) codon[at + j] = (uint8_t)(16 * x + 4 * y + z);
        }
        e->seqOff.push_back((int64_t)(codes0 + at));
        e->seqLen.push_back(n);
        at += n;
    }
    e->seqDirty = true;
}

static int checkGradLengths(dm_engine *e, int32_t nseq, const int64_t *offsets, const int32_t *lengths) {
    for (int32_t s = 0; s < nseq; s++) {
        if (offsets[s] < 0 || lengths[s] < 3)
            return fail(e, DM_EINVAL, "sequences must have at least 3 letters (dummest-dp.hh:84)");
        if (lengths[s] > kGradMaxN)
            return fail(e, DM_EUNSUP, "the gradient takes sequences of up to " + std::to_string(kGradMaxN) + " letters");
    }
    return DM_OK;
}

// The gradient over the nseq consecutive sequences of the store from id s0
// (codes contiguous from codes0), their codon indices on the device at dCodon.
static int gradRun(dm_engine *e, int32_t prof, int32_t nseq, int32_t s0, size_t codes0, const uint8_t *dCodon,
                   const double *weights, double *lnZ, double *counts) {
    const int M = e->profiles[prof].M;
    const int WR = kGradRaw;
    std::vector<int32_t> lengthsV(nseq), pp(nseq, prof), ss(nseq);
    for (int32_t s = 0; s < nseq; s++) {
        ss[s] = s0 + s;
        lengthsV[s] = e->seqLen[(size_t)(s0 + s)];
    }
    const int32_t *lengths = lengthsV.data();
    int r = syncSequences(e);
    if (r) return r;
    // Batches of consecutive sequences whose six matrices fit the budget
    // together. A sequence whose matrices alone pass it forms a batch of its
    // own, run in segments of seg strips (at least one) recomputed from the
    // strips' edge rows.
    struct Batch {
        size_t b, e;
        int seg;  // strips per segment; S = the matrices held in full
    };
    const int S = gradStrips(M);
    std::vector<Batch> batches;
    int64_t maxCells = 1, maxPairs = 1, maxBnd = 1, maxExp = 1;
    {
        int64_t cells = 0, bnd = 0, nexp = 0;
        size_t first = 0;
        auto close = [&](size_t end) {
            if (end > first) batches.push_back({first, end, S});
            first = end;
            cells = bnd = nexp = 0;
        };
        for (int32_t s = 0; s < nseq; s++) {
            const int n = lengths[s];
            const int64_t c = gradSlots(M, n), perStrip = (int64_t)(n + 63) * 64;
            if (c * 6 * 8 > e->gradBudget) {
                close((size_t)s);
                const int seg = (int)std::clamp<int64_t>(e->gradBudget / (6 * 8 * perStrip), 1, S);
                batches.push_back({(size_t)s, (size_t)s + 1, seg});
                first = (size_t)s + 1;
                maxCells = std::max(maxCells, seg * perStrip);
                maxBnd = std::max(maxBnd, (int64_t)S * n);
                maxExp = std::max(maxExp, gradExps(M, n));
                continue;
            }
            if (cells > 0 && (cells + c) * 6 * 8 > e->gradBudget) close((size_t)s);
            cells += c;
            bnd += (int64_t)S * n;
            nexp += gradExps(M, n);
            maxCells = std::max(maxCells, cells);
            maxBnd = std::max(maxBnd, bnd);
            maxExp = std::max(maxExp, nexp);
            maxPairs = std::max<int64_t>(maxPairs, (int64_t)s + 1 - (int64_t)first);
        }
        close((size_t)nseq);
    }
    HIPCHK(e->dGradMat.ensure(sizeof(double) * 6 * (size_t)maxCells));
    // edge rows: five doubles and two exponents per strip and column
    HIPCHK(e->dGradBnd.ensure((sizeof(double) * 5 + sizeof(int) * 2) * (size_t)maxBnd));
    HIPCHK(e->dGradExp.ensure(sizeof(int) * 2 * (size_t)maxExp));
    HIPCHK(e->dGradMeta.ensure((sizeof(int64_t) * 3 + sizeof(int)) * (size_t)maxPairs));
    HIPCHK(e->dGradRaw.ensure(sizeof(double) * (size_t)nseq * (M + 1) * kGradRaw));
    // per sequence: weight, ln Z, factor, Z as mantissa and exponent
    HIPCHK(e->dGradW.ensure(sizeof(double) * 5 * (size_t)nseq));
    HIPCHK(e->dGradOut.ensure(sizeof(double) * (size_t)(M + 1) * WR));
    std::vector<double> w(nseq, 1.0);
    if (weights) std::copy(weights, weights + nseq, w.begin());
    HIPCHK(hipMemcpyAsync(e->dGradW.p, w.data(), sizeof(double) * nseq, hipMemcpyHostToDevice, e->stream));
    std::vector<int64_t> sorted(nseq);
    std::iota(sorted.begin(), sorted.end(), 0);
    const std::vector<int> cfgOf(nseq, 0);  // only the pair records and the null model are used
    double *mat = e->dGradMat.as<double>();
    const int64_t C6 = maxCells;
    if (!e->ev0) HIPCHK(hipEventCreate(&e->ev0));
    if (!e->ev1) HIPCHK(hipEventCreate(&e->ev1));
    float msAll = 0.f;
    int64_t segments = 0;
    for (const Batch &bt : batches) {
        Chunk C;
        buildChunk(e, kScore, sorted, cfgOf, pp.data(), ss.data(), bt.b, bt.e, C);
        int r2 = uploadChunk(e, C);  // pair records and the null model
        if (r2) return r2;
        const int np = (int)C.pairs.size();
        std::vector<int64_t> meta(3 * (size_t)np);  // cell, edge and exponent offsets
        std::vector<int> rawIdx(np);
        int64_t co = 0, bo = 0, eo = 0;
        int maxN = 3;
        for (int k = 0; k < np; k++) {
            const int n = C.pairs[k].n;
            meta[k] = co;
            meta[np + k] = bo;
            meta[2 * np + k] = eo;
            rawIdx[k] = (int)C.order[k];
            co += gradSlots(M, n);
            bo += (int64_t)S * n;
            eo += gradExps(M, n);
            maxN = std::max(maxN, n);
        }
        int64_t *dm = e->dGradMeta.as<int64_t>();
        int *dIdx = (int *)(dm + 3 * np);
        HIPCHK(hipMemcpyAsync(dm, meta.data(), sizeof(int64_t) * 3 * np, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(dIdx, rawIdx.data(), sizeof(int) * np, hipMemcpyHostToDevice, e->stream));
        GradArgs ga{};
        ga.prof = e->dProfs.as<ProfInfo>() + prof;
        ga.params = e->dParams.as<float>();
        ga.emis = e->dEmis.as<float>();
        ga.bg = e->dBg.as<float>() + (int64_t)prof * 32;
        ga.pairs = e->dPairs.as<PairInfo>();
        ga.rec = e->dRec.as<Rec>();
        ga.codon = dCodon - codes0;  // indexed by the pairs' seqOff
        ga.pairBase = 0;
        ga.cellOff = dm;
        ga.bndOff = dm + np;
        ga.expOff = dm + 2 * np;
        ga.rawIdx = dIdx;
        ga.w0 = mat;
        ga.Y = mat + C6;
        ga.Z0 = mat + 2 * C6;
        ga.W1 = mat + 3 * C6;
        ga.Yb = mat + 4 * C6;
        ga.Zb = mat + 5 * C6;
        ga.fwdA = e->dGradBnd.as<double>();
        ga.fwdB = ga.fwdA + maxBnd;
        ga.bwdA = ga.fwdA + 2 * maxBnd;
        ga.bwdB = ga.fwdA + 3 * maxBnd;
        ga.bwdC = ga.fwdA + 4 * maxBnd;
        ga.fwdE = (int *)(ga.fwdA + 5 * maxBnd);
        ga.bwdE = ga.fwdE + maxBnd;
        ga.expF = e->dGradExp.as<int>();
        ga.expB = ga.expF + maxExp;
        ga.zInfo = e->dGradW.as<double>() + 3 * (size_t)nseq;
        ga.raw = e->dGradRaw.as<double>();
        HIPCHK(hipEventRecord(e->ev0, e->stream));
        ga.stripLo = 0;
        ga.stripHi = S - 1;
        ga.store = bt.seg >= S;
        // both passes over every strip (the matrices, or the edge rows alone), Z
        HIPCHK(launch_grad_forward(ga, np, maxN, e->stream));
        HIPCHK(launch_grad_backward(ga, np, maxN, e->stream));
        HIPCHK(launch_grad_z(ga, np, e->stream));
        if (bt.seg >= S) {
            HIPCHK(launch_grad_counts(ga, np, M, maxN, e->stream));
        } else {
            // per segment: its strips again from their edge rows, with the matrices, then counted
            ga.store = 1;
            for (int lo = 0; lo < S; lo += bt.seg) {
                ga.stripLo = lo;
                ga.stripHi = std::min(lo + bt.seg, S) - 1;
                HIPCHK(launch_grad_forward(ga, np, maxN, e->stream));
                HIPCHK(launch_grad_backward(ga, np, maxN, e->stream));
                HIPCHK(launch_grad_counts(ga, np, M, maxN, e->stream));
                segments++;
            }
        }
        HIPCHK(hipEventRecord(e->ev1, e->stream));
        HIPCHK(hipEventSynchronize(e->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        msAll += ms;
    }
    GradReduceArgs ra{};
    ra.M = M;
    ra.npairs = nseq;
    ra.raw = e->dGradRaw.as<double>();
    ra.zInfo = e->dGradW.as<double>() + 3 * (size_t)nseq;
    ra.weight = e->dGradW.as<double>();
    ra.lnZ = e->dGradW.as<double>() + nseq;
    ra.counts = e->dGradOut.as<double>();
    HIPCHK(launch_grad_reduce(ra, e->dGradW.as<double>() + 2 * (size_t)nseq, e->stream));
    HIPCHK(hipMemcpyAsync(counts, ra.counts, sizeof(double) * (size_t)(M + 1) * WR, hipMemcpyDeviceToHost,
                          e->stream));
    if (lnZ) HIPCHK(hipMemcpyAsync(lnZ, ra.lnZ, sizeof(double) * nseq, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->lastGradMs = msAll;
    e->lastGradSegments = segments;
    e->lastGradCells = 0;
    for (int32_t s = 0; s < nseq; s++) e->lastGradCells += gradCells(M, lengths[s]);
    return DM_OK;
}

int dm_grad(dm_engine *e, int32_t prof, int32_t nseq, const char *dna, const int64_t *offsets, const int32_t *lengths,
            const double *weights, double *lnZ, double *counts) {
    if (!e || nseq < 0 || !counts || (nseq > 0 && (!dna || !offsets || !lengths))) return DM_EINVAL;
    if (prof < 0 || prof >= (int)e->profiles.size()) return fail(e, DM_EINVAL, "unknown profile");
    int rc = checkGradLengths(e, nseq, offsets, lengths);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->device));
    rc = syncProfiles(e);
    if (rc) return rc;
    const int M = e->profiles[prof].M;
    if (nseq == 0) {
        std::fill(counts, counts + (size_t)(M + 1) * kGradRaw, 0.0);
        return DM_OK;
    }
    // the sequences stay in the engine's store for the duration of the call
    const size_t codes0 = e->codes.size(), seqs0 = e->seqLen.size();
    std::vector<uint8_t> codon;
    appendGradSequences(e, nseq, dna, offsets, lengths, codon);
    rc = [&]() -> int {
        HIPCHK(e->dGradCodon.ensure(codon.size()));
        HIPCHK(hipMemcpyAsync(e->dGradCodon.p, codon.data(), codon.size(), hipMemcpyHostToDevice, e->stream));
        return gradRun(e, prof, nseq, (int32_t)seqs0, codes0, e->dGradCodon.as<uint8_t>(), weights, lnZ, counts);
    }();
    e->codes.resize(codes0);
    e->seqOff.resize(seqs0);
    e->seqLen.resize(seqs0);
    e->seqDirty = true;
    return rc;
}

int dm_grad_set_create(dm_engine *e, int32_t nseq, const char *dna, const int64_t *offsets, const int32_t *lengths,
                       int32_t *set) {
    if (!e || nseq < 0 || !set || (nseq > 0 && (!dna || !offsets || !lengths))) return DM_EINVAL;
    int rc = checkGradLengths(e, nseq, offsets, lengths);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->device));
    dm_engine::GradSet g;
    g.s0 = (int32_t)e->seqLen.size();
    g.n = nseq;
    g.codes0 = e->codes.size();
    std::vector<uint8_t> codon;
    appendGradSequences(e, nseq, dna, offsets, lengths, codon);
    if (!codon.empty()) {
        HIPCHK(g.dCodon.ensure(codon.size()));
        HIPCHK(hipMemcpy(g.dCodon.p, codon.data(), codon.size(), hipMemcpyHostToDevice));
    }
    e->gradSets.push_back(g);
    *set = (int32_t)e->gradSets.size() - 1;
    return DM_OK;
}

int dm_grad_set(dm_engine *e, int32_t prof, int32_t set, const double *weights, double *lnZ, double *counts) {
    if (!e || !counts) return DM_EINVAL;
    if (prof < 0 || prof >= (int)e->profiles.size()) return fail(e, DM_EINVAL, "unknown profile");
    if (set < 0 || set >= (int)e->gradSets.size() || e->gradSets[set].s0 < 0)
        return fail(e, DM_EINVAL, "unknown sequence set");
    HIPCHK(hipSetDevice(e->device));
    int rc = syncProfiles(e);
    if (rc) return rc;
    const auto &g = e->gradSets[set];
    if ((size_t)g.s0 + (size_t)g.n > e->seqLen.size() ||
        (g.n > 0 && (size_t)e->seqOff[(size_t)g.s0] != g.codes0) || g.codes0 > e->codes.size())
        return fail(e, DM_EINVAL, "the sequence set's sequences are no longer in the engine's store");
    if (g.n == 0) {
        std::fill(counts, counts + (size_t)(e->profiles[prof].M + 1) * kGradRaw, 0.0);
        return DM_OK;
    }
    return gradRun(e, prof, g.n, g.s0, g.codes0, g.dCodon.as<uint8_t>(), weights, lnZ, counts);
}

int dm_grad_set_destroy(dm_engine *e, int32_t set) {
    if (!e) return DM_EINVAL;
    if (set < 0 || set >= (int)e->gradSets.size() || e->gradSets[set].s0 < 0)
        return fail(e, DM_EINVAL, "unknown sequence set");
    HIPCHK(hipSetDevice(e->device));
    e->gradSets[set].dCodon.release();
    e->gradSets[set].s0 = -1;  // the sequences stay in the store (other ids keep their meaning)
    return DM_OK;
}

// ------------------------------------------------ profile builder: Baum-Welch
int dm_bw_prepare(dm_engine *e, int32_t A, int32_t L, int32_t nseq, const uint8_t *letters, const int64_t *offsets,
                  const int32_t *lengths, const double *weights) {
    if (!e || (A != 4 && A != 20) || L < 0 || nseq < 0 ||
        (nseq > 0 && (!letters || !offsets || !lengths || !weights)))
        return DM_EINVAL;
    HIPCHK(hipSetDevice(e->device));
    e->bwL = -1;
    // letters packed back to back, one zero pad after the last (the recurrences
    // read one letter past a sequence and multiply it by zero)
    std::vector<int64_t> loff(nseq), coff(nseq);
    std::vector<int32_t> len(lengths, lengths + nseq);
    int64_t total = 0;
    for (int32_t s = 0; s < nseq; s++) {
        if (lengths[s] < 0 || offsets[s] < 0) return fail(e, DM_EINVAL, "negative sequence length or offset");
        if (lengths[s] > kBwMaxLen)
            return fail(e, DM_EUNSUP, "Baum-Welch takes sequences of up to " + std::to_string(kBwMaxLen) + " letters");
        loff[s] = total;
        total += lengths[s];
    }
    std::vector<uint8_t> packed((size_t)total + 1, 0);
    for (int32_t s = 0; s < nseq; s++)
        for (int32_t j = 0; j < lengths[s]; j++) {
            const uint8_t x = letters[offsets[s] + j];
            if (x > A) return fail(e, DM_EINVAL, "letter code above the alphabet size");
            packed[loff[s] + j] = x;
        }
    // batches whose six matrices fit the budget
    e->bwBatch.clear();
    e->bwMatCells = 0;
    int64_t cells = 0;
    for (int32_t s = 0; s < nseq; s++) {
        const int64_t c = bwCells(L, lengths[s]);
        if (e->bwBatch.empty() || (cells > 0 && (cells + c) * 6 * 8 > e->bwBudget)) {
            e->bwBatch.push_back(s);
            cells = 0;
        }
        coff[s] = cells;
        cells += c;
        e->bwMatCells = std::max(e->bwMatCells, cells);
    }
    e->bwBatch.push_back(nseq);
    e->bwLen.assign(lengths, lengths + nseq);
    const size_t metaBytes = sizeof(int64_t) * 2 * nseq + sizeof(int32_t) * nseq;
    HIPCHK(e->dBwLetters.ensure(packed.size()));
    HIPCHK(e->dBwMeta.ensure(metaBytes));
    HIPCHK(e->dBwMat.ensure(sizeof(double) * 6 * std::max<int64_t>(1, e->bwMatCells)));
    HIPCHK(e->dBwRaw.ensure(sizeof(double) * (size_t)std::max(1, nseq) * (L + 1) * (kBwRaw + A)));
    HIPCHK(e->dBwSeqW.ensure(sizeof(double) * 3 * std::max(1, nseq)));
    HIPCHK(e->dBwTables.ensure(sizeof(double) * (size_t)(L + 1) * (4 + A + 1)));
    HIPCHK(e->dBwOut.ensure(sizeof(double) * (size_t)(L + 1) * (7 + A)));
    HIPCHK(hipMemcpyAsync(e->dBwLetters.p, packed.data(), packed.size(), hipMemcpyHostToDevice, e->stream));
    int64_t *m = e->dBwMeta.as<int64_t>();
    if (nseq > 0) {
        HIPCHK(hipMemcpyAsync(m, loff.data(), sizeof(int64_t) * nseq, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(m + nseq, coff.data(), sizeof(int64_t) * nseq, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(m + 2 * nseq, len.data(), sizeof(int32_t) * nseq, hipMemcpyHostToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(e->dBwSeqW.p, weights, sizeof(double) * nseq, hipMemcpyHostToDevice, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    e->bwL = L;
    e->bwA = A;
    e->bwN = nseq;
    return DM_OK;
}

int dm_bw_estep(dm_engine *e, const double *probs, double *counts, double *totals) {
#pragma clang fp contract(off)
    if (!e || !probs || !counts) return DM_EINVAL;
    if (e->bwL < 0) return fail(e, DM_EINVAL, "dm_bw_prepare first");
    HIPCHK(hipSetDevice(e->device));
    const int L = e->bwL, A = e->bwA, N = e->bwN, width = 7 + A;
    // per profile row: a' = ins (1 - ext), ext, d' = del (1 - next dext),
    // e' = dext (1 - next dext) / (1 - dext) (0 when not finite), and the
    // match likelihood ratio (1 - ins - del) p[x] / bg[x] (0 when not finite,
    // and for letters outside the alphabet) -- dummest-build.cc:566-601
    std::vector<double> tab((size_t)(L + 1) * (4 + A + 1));
    double *rates = tab.data(), *ratio = tab.data() + (size_t)(L + 1) * 4;
    const double *bg = probs + (size_t)L * width + 7;
    for (int r = 0; r <= L; r++) {
        const double *p = probs + (size_t)r * width;
        const double nextDext = (r == L) ? 0.0 : probs[(size_t)(r + 1) * width + 6];
        rates[4 * r] = p[1] * (1.0 - p[4]);
        rates[4 * r + 1] = p[4];
        rates[4 * r + 2] = p[2] * (1 - nextDext);
        double ep = p[6] * (1 - nextDext) / (1 - p[6]);
        rates[4 * r + 3] = std::isfinite(ep) ? ep : 0.0;
        for (int x = 0; x < A; x++) {
            const double v = (1 - p[1] - p[2]) * p[7 + x] / bg[x];
            ratio[(size_t)r * (A + 1) + x] = std::isfinite(v) ? v : 0.0;
        }
        ratio[(size_t)r * (A + 1) + A] = 0.0;
    }
    HIPCHK(hipMemcpyAsync(e->dBwTables.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, e->stream));
    if (!e->ev0) HIPCHK(hipEventCreate(&e->ev0));
    if (!e->ev1) HIPCHK(hipEventCreate(&e->ev1));
    const int64_t *m = e->dBwMeta.as<int64_t>();
    double *mat = e->dBwMat.as<double>();
    const int64_t C = std::max<int64_t>(1, e->bwMatCells);
    BWArgs a{};
    a.L = L;
    a.A = A;
    a.letters = e->dBwLetters.as<uint8_t>();
    a.letterOff = m;
    a.cellOff = m + N;
    a.len = (const int *)(m + 2 * N);
    a.rates = e->dBwTables.as<double>();
    a.ratio = e->dBwTables.as<double>() + (size_t)(L + 1) * 4;
    a.X = mat;
    a.Y = mat + C;
    a.Z = mat + 2 * C;
    a.Wb = mat + 3 * C;
    a.Yb = mat + 4 * C;
    a.Zb = mat + 5 * C;
    a.weights = e->dBwSeqW.as<double>();
    a.total = e->dBwSeqW.as<double>() + std::max(1, N);
    a.wt = e->dBwSeqW.as<double>() + 2 * std::max(1, N);
    a.raw = e->dBwRaw.as<double>();
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    for (size_t b = 0; b + 1 < e->bwBatch.size(); b++) {
        a.seqBase = e->bwBatch[b];
        const int n = e->bwBatch[b + 1] - e->bwBatch[b];
        const int maxLen = *std::max_element(e->bwLen.begin() + e->bwBatch[b], e->bwLen.begin() + e->bwBatch[b + 1]);
        HIPCHK(launch_bw_forward(a, n, maxLen, e->stream));
        HIPCHK(launch_bw_backward(a, n, maxLen, e->stream));
        HIPCHK(launch_bw_counts(a, n, e->stream));
    }
    BWReduceArgs ra{};
    ra.L = L;
    ra.A = A;
    ra.nseq = N;
    ra.wt = a.wt;
    ra.raw = a.raw;
    ra.counts = e->dBwOut.as<double>();
    HIPCHK(launch_bw_reduce(ra, e->stream));
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    HIPCHK(hipMemcpyAsync(counts, ra.counts, sizeof(double) * (size_t)(L + 1) * width, hipMemcpyDeviceToHost,
                          e->stream));
    if (totals && N > 0)
        HIPCHK(hipMemcpyAsync(totals, a.total, sizeof(double) * N, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->lastBwMs = ms;
    return DM_OK;
}
int64_t dm_last_cells(const dm_engine *e) { return e ? e->lastCells : 0; }

}  // extern "C"

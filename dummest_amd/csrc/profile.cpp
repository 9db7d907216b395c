// HMMER3/f profile reader and transform, and DNA decoding, on the host.
//
// Reader: the line state machine of readProfiles (dummest-profile.hh:235-316):
// transitions m->i, i->i, m->d, d->d and 20 match emissions per node, stored
// as probabilities exp(-x) ("*" = 0, probFromText dummest-profile.hh:14-19).
// Row r holds node r's transitions and node r+1's emissions.
// Transform: finalizeProfile (dummest-profile.hh:105-233), with the profile
// masking of -m 1 (tantan on the consensus) when a threshold is given, or the
// Gaussian letter filter of -d/-D when a standard deviation is. The
// arithmetic keeps the reference's Float/double mix (Float = float for the
// FP32 engine): rate products are formed in float, the per-row Params in
// double and narrowed to float.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <thread>

#include <sched.h>

#include "dm_host.h"

namespace dm {

namespace {

const char *kAlpha = "ACDEFGHIKLMNPQRSTVWYUO?*";

int synonymousCodons(char aa) {  // codon counts of the standard code (dummest-sequence.hh:103-124)
    switch (aa) {
    case 'A': case 'G': case 'P': case 'T': case 'V': return 4;
    case 'R': case 'L': case 'S': return 6;
    case 'I': return 3;
    case 'M': case 'W': return 1;
    default: return 2;
    }
}

double probFromWord(const std::string &w) {
    if (!w.empty() && w[0] == '*') return 0.0;
    return std::exp(-std::strtod(w.c_str(), nullptr));
}

// myMean, dummest-profile.hh:37-62: positions whose repeat probability reaches
// the threshold are left out (profile masking).
double backgroundMean(const float *v, int length, int step, int type, const float *rep = nullptr,
                      float thr = 0.f) {
    double mean = 0;
    int n = 0;
    std::vector<float> med;
    for (int i = 0; i < length; ++i) {
        if (rep && rep[i] >= thr) continue;
        const double x = v[(size_t)i * step];
        if (type == 'G') mean += std::log(std::max(x, 1e-6));
        else if (type == 'A') mean += x;
        else med.push_back((float)x);
        ++n;
    }
    if (n == 0) return NAN;  // the reference asserts n > 0 (dummest-profile.hh:53)
    if (type == 'G') return std::exp(mean / n);
    if (type == 'A') return mean / n;
    std::sort(med.begin(), med.end());
    return med[n / 2];
}

// The Gaussian letter filter of -d/-D (filterLetterProbabilities,
// dummest-profile.hh:64-103): for each letter, the log probability less the
// position's mean log probability over the 20 letters is a signal along the
// profile; its Gaussian-smoothed version (standard deviation sd positions,
// tails cut at 8 sd, the profile taken as circular) is subtracted, which
// keeps the part that varies from position to position. keepNonvarying (-D)
// adds the signal's mean back. Each position is then renormalized. Doubles
// throughout, results narrowed to float, in the reference's order of
// operations (its profiles must come out bit for bit); the Gaussian weights
// are tabulated once per call, the same doubles the reference recomputes.
void gaussianLetterFilter(float *lp, int length, int step, double sd, bool keepNonvarying) {
    const int A = 20;
    const double inv2var = 0.5 / (sd * sd);
    const double norm = sd * 2.5066282746310005;  // sd * sqrt(2 pi)
    const int reach = (int)std::ceil(sd * 8);
    std::vector<double> gw(2 * (size_t)reach + 1);
    for (int d = -reach; d <= reach; ++d) gw[d + reach] = std::exp(-1.0 * d * d * inv2var);
    std::vector<double> posMeanLog(length), sig(length);
    for (int i = 0; i < length; ++i) {
        double prod = 1;
        for (int k = 0; k < A; ++k) prod *= lp[(size_t)i * step + k];
        posMeanLog[i] = std::log(prod) / A;
    }
    for (int k = 0; k < A; ++k) {
        for (int i = 0; i < length; ++i) sig[i] = std::log((double)lp[(size_t)i * step + k]) - posMeanLog[i];
        double level = 0;
        if (keepNonvarying) {
            for (int i = 0; i < length; ++i) level += sig[i];
            level /= length;
        }
        for (int i = 0; i < length; ++i) {
            double smooth = 0;
            for (int d = -reach; d <= reach; ++d) {
                int x = (i + d) % length;
                if (x < 0) x += length;
                smooth += sig[x] * gw[d + reach];
            }
            smooth /= norm;
            lp[(size_t)i * step + k] = (float)std::exp(sig[i] - smooth + level);
        }
    }
    for (int i = 0; i < length; ++i) {
        float *row = lp + (size_t)i * step;
        double total = 0;
        for (int k = 0; k < A; ++k) total += row[k];
        for (int k = 0; k < A; ++k) row[k] = (float)(row[k] / total);
    }
}

bool finalizeImpl(HostProfile &p, int bgType, const double rates[6], std::string &msg, float maskThr, double filterSd,
                  bool keepNonvarying) {
    const float ins1 = (float)rates[0], ins2 = (float)rates[1], del1 = (float)rates[2], del2 = (float)rates[3];
    const float stop = (float)rates[4], bgStop = (float)rates[5];
    const int M = p.M;
    float *v = p.values.data();
    float *end = v + (size_t)kWidth * M;
    if (end[3] <= 0) end[3] = (float)backgroundMean(v + kWidth + 3, M - 1, kWidth, 'G');
    // The Gaussian filter takes the place of profile masking
    // (dummest-profile.hh:119-126; dummest.cc also drops -m's profile bit).
    if (filterSd > 0) {
        gaussianLetterFilter(v + 4, M, kWidth, filterSd, keepNonvarying);
        maskThr = -1.f;
    }
    // Profile masking (-m 1, dummest-profile.hh:122-125): tantan repeat
    // probabilities of the consensus (the first most probable letter of each
    // node); positions at or above the threshold leave the background means
    // and emit like the background.
    std::vector<float> rep;
    p.masked = 0;
    if (maskThr >= 0.f) {
        if ((int)p.consensus.size() != M) {
            p.consensus.assign(M, 0);
            for (int i = 0; i < M; ++i) {
                const float *e = v + (size_t)i * kWidth + 4;
                p.consensus[i] = (char)(std::max_element(e, e + 20) - e);
            }
        }
        rep.resize(M);
        tantanRepeatProbs(reinterpret_cast<const uint8_t *>(p.consensus.data()), M, true, rep.data());
        for (int i = 0; i < M; ++i)
            if (rep[i] >= maskThr) {
                ++p.masked;
                p.consensus[i] |= 32;  // printed in lowercase (dummest-profile.hh:225-226)
            }
    }
    const float *rp = rep.empty() ? nullptr : rep.data();
    double sum = 0;
    for (int k = 4; k < 24; ++k) {
        const double m = backgroundMean(v + k, M, kWidth, bgType, rp, maskThr);
        if (!(m == m)) {
            msg = "every profile position is masked";
            return false;
        }
        end[k] = (float)m;
        sum += m;
    }
    p.bg.assign(kWidth, 0.f);
    for (int k = 4; k < 24; ++k) {
        end[k] = (float)((double)end[k] / sum);
        end[k] *= (1 - bgStop) / (float)synonymousCodons(kAlpha[k - 4]);
        p.bg[k] = end[k];
    }
    p.bg[26] = (float)(1.0 / 64.0);
    p.bg[27] = (float)(bgStop / 3.0);
    p.params.assign((size_t)(M + 2) * 11, 0.f);
    const float keep = 1 - ins1 - ins2 - del1 - del2;
    for (int i = 0;; ++i) {
        float *pp = &p.params[(size_t)i * 11];
        float *row = v + (size_t)i * kWidth;
        const double alpha = row[0] * keep;
        const double beta = row[1];
        pp[0] = (float)(alpha * (1 - beta));
        pp[1] = ins1;
        pp[2] = ins2;
        pp[3] = (float)beta;
        const double delta = row[2] * keep;
        const double eps = row[3];
        if (i == M) break;
        const double eps1 = row[kWidth + 3];
        pp[6] = (float)(delta * (1 - eps1));
        pp[7] = del1;
        pp[8] = del2;
        pp[9] = (float)(eps * (1 - eps1) / (1 - eps));
        pp[10] = (float)(1 - alpha - (double)ins1 - (double)ins2 - delta - (double)del1 - (double)del2);
        if (eps >= 1) {
            msg = "profile has a delete-to-delete probability of 1";
            return false;
        }
        row[2] = (float)delta;
        row[3] = (float)eps;
        const bool masked = rp && rp[i] >= maskThr;
        for (int k = 4; k < 24; ++k) {
            const double x = row[k];
            row[k] = masked ? end[k] : (float)((1 - stop) * x / synonymousCodons(kAlpha[k - 4]));
        }
        row[24] = row[5];   // U = C
        row[25] = row[12];  // O = K
        row[26] = (float)(1.0 / 64.0);
        row[27] = (float)(stop / 3.0);
        row[28] = 0.f;
    }
    return true;
}

}  // namespace

int hostThreads() {
    int n = 0;
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(16, n));
}

namespace {

// The HMMER3 reader's state machine over text[b, e): profiles are appended to
// out (finalized); *endState receives the state at the end (0 between
// profiles). Returns false with msg on malformed data.
bool readHmmRange(const std::string &text, size_t b, size_t e, int bgType, const double rates[6],
                  std::vector<HostProfile> &out, std::string &msg, float maskThr, double filterSd, bool keepNonvarying,
                  int *endState) {
    std::istringstream in(text.substr(b, e - b));
    std::string line;
    int state = 0;
    HostProfile cur;
    std::vector<float> vals;
    auto bad = [&](const char *m) {
        msg = m;
        return false;
    };
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string w;
        ls >> w;
        if (state == 0) {
            if (w == "NAME") ls >> cur.name;
            else if (w == "HMM") state = 1;
        } else if (state == 1) {
            state = 2;
        } else if (state == 2) {
            if (w != "COMPO") state = 3;
        } else if (state == 3) {
            std::string mi, md, im, ii, dm_, dd;
            ls >> mi >> md >> im >> ii >> dm_ >> dd;
            if (!ls) return bad("bad transition line in profile");
            const double MtoI = probFromWord(mi), MtoD = probFromWord(md), ItoI = probFromWord(ii), DtoD = probFromWord(dd);
            if (MtoI > 1 || MtoD > 1 || ItoI > 1 || DtoD > 1) return bad("transition probability above 1");
            vals.push_back((float)MtoI);
            vals.push_back((float)ItoI);
            vals.push_back((float)MtoD);
            vals.push_back((float)DtoD);
            state = 4;
        } else {
            if (w == "//") {
                if (cur.M < 2) return bad("profile shorter than 2 positions");
                vals.insert(vals.end(), kWidth - 4, 0.f);
                cur.values = vals;
                if (!finalizeImpl(cur, bgType, rates, msg, maskThr, filterSd, keepNonvarying)) return false;
                out.push_back(std::move(cur));
                cur = HostProfile();
                vals.clear();
                state = 0;
            } else {
                int k = 0;
                float best = -1;
                int bestK = 0;
                while (ls >> w && w.find('.') != std::string::npos) {
                    const double pr = probFromWord(w);
                    if (pr > 1) return bad("emission probability above 1");
                    if ((float)pr > best) {
                        best = (float)pr;
                        bestK = k;
                    }
                    vals.push_back((float)pr);
                    ++k;
                }
                if (k != 20) return bad("only 20-letter protein profiles are supported");
                vals.insert(vals.end(), 5, 0.f);
                cur.M += 1;
                cur.consensus.push_back((char)bestK);
                state = 2;
            }
        }
    }
    if (endState) *endState = state;
    if (state != 0) return bad("truncated profile data");
    return true;
}

}  // namespace

// Profiles are independent records ending in a "//" line, so they are parsed
// and finalized on a pool of host threads, one record at a time, into their
// places in file order. A record that fails, or whose "//" does not end a
// profile, sends the whole text through the serial reader instead, so that
// malformed data gives exactly the serial reader's result and message.
bool readHmmText(const std::string &text, int bgType, const double rates[6], std::vector<HostProfile> &out,
                 std::string &msg, float maskThr, double filterSd, bool keepNonvarying) {
    std::vector<size_t> ends;  // one past each "//" line
    for (size_t pos = 0; pos < text.size();) {
        size_t nl = text.find('\n', pos);
        const size_t le = nl == std::string::npos ? text.size() : nl + 1;
        size_t f = pos;
        while (f < le && (text[f] == ' ' || text[f] == '\t')) f++;
        if (f + 1 < le && text[f] == '/' && text[f + 1] == '/' &&
            (f + 2 == le || std::isspace((unsigned char)text[f + 2])))
            ends.push_back(le);
        pos = le;
    }
    const size_t nrec = ends.size();
    const size_t nth = std::min<size_t>((size_t)hostThreads(), nrec);
    if (nrec >= 8 && nth > 1) {
        std::vector<std::vector<HostProfile>> got(nrec);
        std::vector<char> ok(nrec, 0);
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nth; t++)
            pool.emplace_back([&, t]() {
                std::string m;
                for (size_t r = t; r < nrec; r += nth) {
                    int st = -1;
                    ok[r] = readHmmRange(text, r ? ends[r - 1] : 0, ends[r], bgType, rates, got[r], m, maskThr, filterSd,
                                         keepNonvarying, &st) &&
                            st == 0 && got[r].size() == 1;
                }
            });
        for (auto &th : pool) th.join();
        int tail = -1;
        std::vector<HostProfile> rest;
        std::string m;
        const bool tailOk = readHmmRange(text, nrec ? ends[nrec - 1] : 0, text.size(), bgType, rates, rest, m, maskThr,
                                         filterSd, keepNonvarying, &tail) &&
                            rest.empty();
        if (tailOk && std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; })) {
            for (auto &g : got) out.push_back(std::move(g[0]));
            return true;
        }
    }
    return readHmmRange(text, 0, text.size(), bgType, rates, out, msg, maskThr, filterSd, keepNonvarying, nullptr);
}

bool finalizeRaw(HostProfile &p, int bgType, const double rates[6], std::string &msg, float maskThr, double filterSd,
                 bool keepNonvarying) {
    return finalizeImpl(p, bgType, rates, msg, maskThr, filterSd, keepNonvarying);
}

bool readHmmFile(const char *path, int bgType, const double rates[6], std::vector<HostProfile> &out, std::string &msg,
                 float maskThr, double filterSd, bool keepNonvarying) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        msg = std::string("can't open file: ") + path;
        return false;
    }
    f.seekg(0, std::ios::end);
    const std::streamoff size = f.tellg();
    f.seekg(0, std::ios::beg);
    std::string text;
    if (size > 0) {
        text.resize((size_t)size);
        f.read(&text[0], size);
        text.resize((size_t)f.gcount());
    } else {  // not seekable: read the stream as it comes
        f.clear();
        std::stringstream ss;
        ss << f.rdbuf();
        text = ss.str();
    }
    return readHmmText(text, bgType, rates, out, msg, maskThr, filterSd, keepNonvarying);
}

// Codon table in TCAG order; non-ACGT letters make the codon '?'.
void decodeDna(const char *dna, int64_t n, uint8_t *codes) {
    // lookup tables, built once (thread-safe static initialisation)
    struct Tables {
        int8_t base[256];
        uint8_t num[256];
        Tables() {
            std::memset(base, -1, sizeof(base));
            base['T'] = base['t'] = 0;
            base['C'] = base['c'] = 1;
            base['A'] = base['a'] = 2;
            base['G'] = base['g'] = 3;
            for (int c = 0; c < 256; c++) num[c] = 22;
            for (int i = 0; kAlpha[i]; i++) num[(unsigned char)kAlpha[i]] = (uint8_t)i;
        }
    };
    static const Tables T;
    static const char *aa = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
    for (int64_t i = 0; i < n; i++) codes[i] = 0;
    for (int64_t i = 0; i + 2 < n; i++) {
        const int b0 = T.base[(unsigned char)dna[i]], b1 = T.base[(unsigned char)dna[i + 1]],
                  b2 = T.base[(unsigned char)dna[i + 2]];
        codes[i] = (b0 < 0 || b1 < 0 || b2 < 0) ? 22 : T.num[(unsigned char)aa[b0 * 16 + b1 * 4 + b2]];
    }
}


// ---------------------------------------------------------------------------
// tantan repeat probabilities on the host (profile masking): the gap-free
// forward-backward of tantan/tantan.cc:40-151, 222-397 with the parameters of
// calcTantanProbabilities (tantan-wrapper.hh:13-64). Four interleaved partial
// sums over the repeat offsets reduced as (s0+s2)+(s1+s3), and the fused
// multiply-adds written out, as the reference's x86-64-v3 build evaluates
// them (the DNA kernel in kernels.hip does the same on the device).
// ---------------------------------------------------------------------------
namespace {
const signed char kBlosum62[20][20] = {
    {4, 0, -2, -1, -2, 0, -2, -1, -1, -1, -1, -2, -1, -1, -1, 1, 0, 0, -3, -2},
    {0, 9, -3, -4, -2, -3, -3, -1, -3, -1, -1, -3, -3, -3, -3, -1, -1, -1, -2, -2},
    {-2, -3, 6, 2, -3, -1, -1, -3, -1, -4, -3, 1, -1, 0, -2, 0, -1, -3, -4, -3},
    {-1, -4, 2, 5, -3, -2, 0, -3, 1, -3, -2, 0, -1, 2, 0, 0, -1, -2, -3, -2},
    {-2, -2, -3, -3, 6, -3, -1, 0, -3, 0, 0, -3, -4, -3, -3, -2, -2, -1, 1, 3},
    {0, -3, -1, -2, -3, 6, -2, -4, -2, -4, -3, 0, -2, -2, -2, 0, -2, -3, -2, -3},
    {-2, -3, -1, 0, -1, -2, 8, -3, -1, -3, -2, 1, -2, 0, 0, -1, -2, -3, -2, 2},
    {-1, -1, -3, -3, 0, -4, -3, 4, -3, 2, 1, -3, -3, -3, -3, -2, -1, 3, -3, -1},
    {-1, -3, -1, 1, -3, -2, -1, -3, 5, -2, -1, 0, -1, 1, 2, 0, -1, -2, -3, -2},
    {-1, -1, -4, -3, 0, -4, -3, 2, -2, 4, 2, -3, -3, -2, -2, -2, -1, 1, -2, -1},
    {-1, -1, -3, -2, 0, -3, -2, 1, -1, 2, 5, -2, -2, 0, -1, -1, -1, 1, -1, -1},
    {-2, -3, 1, 0, -3, 0, 1, -3, 0, -3, -2, 6, -2, 0, 0, 1, 0, -3, -4, -2},
    {-1, -3, -1, -1, -4, -2, -2, -3, -1, -3, -2, -2, 7, -1, -2, -1, -1, -2, -4, -3},
    {-1, -3, 0, 2, -3, -2, 0, -3, 1, -2, 0, 0, -1, 5, 1, 0, -1, -2, -2, -1},
    {-1, -3, -2, 0, -3, -2, 0, -3, 2, -2, -1, 0, -2, 1, 5, -1, -1, -3, -3, -2},
    {1, -1, 0, 0, -2, 0, -1, -2, 0, -2, -1, 1, -1, 0, -1, 4, 1, -2, -3, -2},
    {0, -1, -1, -1, -2, -2, -2, -1, -1, -1, -1, 0, -1, -1, -1, 1, 5, 0, -2, -2},
    {0, -1, -3, -2, -1, -3, -3, 3, -2, 1, 1, -3, -2, -2, -3, -2, 0, 4, -3, -1},
    {-3, -2, -4, -3, 1, -2, -2, -3, -3, -2, -1, -4, -4, -2, -3, -3, -2, -3, 11, 2},
    {-2, -2, -3, -2, 3, -3, 2, -1, -2, -1, -1, -2, -3, -1, -2, -2, -2, -1, 2, 7},
};
}  // namespace

TantanParams tantanParams(bool isProtein) {
    TantanParams t;
    t.K = isProtein ? 50 : 100;
    for (int i = 0; i < 22; ++i)
        for (int j = 0; j < 22; ++j)
            t.ratio[i][j] = isProtein ? std::exp(0.32403216734804408 * ((i < 20 && j < 20) ? kBlosum62[i][j] : -4))
                                      : ((i < 4 && i == j) ? 3.0 : 1.0 / 3.0);
    const double repeatProb = 0.005, repeatEndProb = 0.05, decay = 0.9;
    t.b2b = 1 - repeatProb;
    t.f2b = repeatEndProb;
    t.f2f0 = 1 - repeatEndProb;
    // firstRepeatOffsetProb (tantan.cc:25-30) for the decaying offset weights
    const double first = (decay < 1 || decay > 1) ? (1 - decay) / (1 - std::pow(decay, t.K)) : 1.0 / t.K;
    double q = repeatProb * first;
    for (int i = 0; i < t.K; ++i) {
        t.b2f[i] = q;
        q *= decay;
    }
    return t;
}

void tantanRepeatProbs(const uint8_t *seq, int n, bool isProtein, float *out) {
    const TantanParams t = tantanParams(isProtein);
    const int K = t.K;
    std::vector<double> fp(K, 0.0), scales(n / 16);
    double bg = 1.0;
    for (int pos = 0; pos < n; ++pos) {  // forward, with a rescale every 16 letters
        const double b = bg;
        const double *lr = t.ratio[seq[pos]];
        const int m = std::min(pos, K);
        double s4[4] = {0, 0, 0, 0};
        int i = 0;
        for (; i + 4 <= m; i += 4)
            for (int l = 0; l < 4; ++l) {
                const double f = fp[i + l];
                s4[l] += f;
                fp[i + l] = std::fma(b, t.b2f[i + l], f * t.f2f0) * lr[seq[pos - i - l - 1]];
            }
        double sum = (s4[0] + s4[2]) + (s4[1] + s4[3]);
        for (; i < m; ++i) {
            const double f = fp[i];
            sum += f;
            fp[i] = std::fma(b, t.b2f[i], f * t.f2f0) * lr[seq[pos - i - 1]];
        }
        bg = std::fma(b, t.b2b, sum * t.f2b);
        if (pos % 16 == 15) {
            const double sc = 1 / bg;
            scales[pos / 16] = sc;
            bg *= sc;
            for (double &f : fp) f *= sc;
        }
        out[pos] = (float)bg;
    }
    double all = 0.0;
    for (double f : fp) all += f;
    const double z = std::fma(all, t.f2b, bg * t.b2b);
    bg = t.b2b;
    std::fill(fp.begin(), fp.end(), t.f2b);
    for (int pos = n - 1; pos >= 0; --pos) {  // backward: posterior of the non-repeat state
        const double nonRepeat = (double)out[pos] * bg / z;
        out[pos] = 1 - (float)nonRepeat;
        if (pos % 16 == 15) {
            const double sc = scales[pos / 16];
            bg *= sc;
            for (double &f : fp) f *= sc;
        }
        const double toBg = t.f2b * bg;
        const double *lr = t.ratio[seq[pos]];
        const int m = std::min(pos, K);
        double s4[4] = {0, 0, 0, 0};
        int i = 0;
        for (; i + 4 <= m; i += 4)
            for (int l = 0; l < 4; ++l) {
                const double f = fp[i + l] * lr[seq[pos - i - l - 1]];
                s4[l] = std::fma(t.b2f[i + l], f, s4[l]);
                fp[i + l] = std::fma(f, t.f2f0, toBg);
            }
        double tf = (s4[0] + s4[2]) + (s4[1] + s4[3]);
        for (; i < m; ++i) {
            const double f = fp[i] * lr[seq[pos - i - 1]];
            tf = std::fma(t.b2f[i], f, tf);
            fp[i] = std::fma(f, t.f2f0, toBg);
        }
        bg = std::fma(t.b2b, bg, tf);
    }
}

}  // namespace dm

// HMMER3/f profile reader and transform, and DNA decoding, on the host.
//
// Reader: the line state machine of readProfiles (dummest-profile.hh:235-316):
// transitions m->i, i->i, m->d, d->d and 20 match emissions per node, stored
// as probabilities exp(-x) ("*" = 0, probFromText dummest-profile.hh:14-19).
// Row r holds node r's transitions and node r+1's emissions.
// Transform: finalizeProfile (dummest-profile.hh:105-233) without tantan
// masking. The arithmetic keeps the reference's Float/double mix (Float =
// float for the FP32 engine): rate products are formed in float, the
// per-row Params in double and narrowed to float.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

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

double backgroundMean(const float *v, int length, int step, int type) {  // myMean, dummest-profile.hh:37-62
    double mean = 0;
    std::vector<float> med;
    for (int i = 0; i < length; ++i) {
        const double x = v[(size_t)i * step];
        if (type == 'G') mean += std::log(std::max(x, 1e-6));
        else if (type == 'A') mean += x;
        else med.push_back((float)x);
    }
    if (type == 'G') return std::exp(mean / length);
    if (type == 'A') return mean / length;
    std::sort(med.begin(), med.end());
    return med[length / 2];
}

bool finalizeImpl(HostProfile &p, int bgType, const double rates[6], std::string &msg) {
    const float ins1 = (float)rates[0], ins2 = (float)rates[1], del1 = (float)rates[2], del2 = (float)rates[3];
    const float stop = (float)rates[4], bgStop = (float)rates[5];
    const int M = p.M;
    float *v = p.values.data();
    float *end = v + (size_t)kWidth * M;
    if (end[3] <= 0) end[3] = (float)backgroundMean(v + kWidth + 3, M - 1, kWidth, 'G');
    double sum = 0;
    for (int k = 4; k < 24; ++k) {
        const double m = backgroundMean(v + k, M, kWidth, bgType);
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
        for (int k = 4; k < 24; ++k) {
            const double x = row[k];
            row[k] = (float)((1 - stop) * x / synonymousCodons(kAlpha[k - 4]));
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

bool readHmmText(const std::string &text, int bgType, const double rates[6], std::vector<HostProfile> &out,
                 std::string &msg) {
    std::istringstream in(text);
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
                if (!finalizeImpl(cur, bgType, rates, msg)) return false;
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
    if (state != 0) return bad("truncated profile data");
    return true;
}

bool finalizeRaw(HostProfile &p, int bgType, const double rates[6], std::string &msg) {
    return finalizeImpl(p, bgType, rates, msg);
}

bool readHmmFile(const char *path, int bgType, const double rates[6], std::vector<HostProfile> &out, std::string &msg) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        msg = std::string("can't open file: ") + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return readHmmText(ss.str(), bgType, rates, out, msg);
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

}  // namespace dm

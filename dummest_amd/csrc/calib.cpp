// Calibration on the host side of the GPU engine: the random frameshifting
// codon DNA of each trial and the Gumbel fits of the scores.
//
// Sequences follow estimateK's generator (dummest-calibration.hh:363-376 for
// the amino-acid weights, 436-482 for the sequence): mt19937_64 seeded with
// 5489 + trial, a 0-2 base random offset, then per codon slot either a 1- or
// 2-base background frameshift or an amino acid drawn from the background
// (times its synonymous-codon count) with a uniformly chosen codon, and the
// first `border` bases copied to the end. The draws use libstdc++'s
// distributions, as the reference does, so the same trial gives the same DNA.
// The synonymous-codon lists keep the reference's order
// (dummest-sequence.hh:103-124) because the codon draw indexes into them.
//
// Fits: estimateGumbel (dummest-calibration.hh:7-104): method of moments,
// maximum likelihood (bisection) and L-moments on the sorted scores.
#include <algorithm>
#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "dm_host.h"

namespace dm {

namespace {

const char *kAa = "ACDEFGHIKLMNPQRSTVWY";

const std::vector<std::string> &codonsOf(char aa) {
    static const std::vector<std::string> A{"GCT", "GCC", "GCA", "GCG"}, R{"CGT", "CGC", "CGA", "CGG", "AGA", "AGG"},
        N{"AAT", "AAC"}, D{"GAT", "GAC"}, C{"TGT", "TGC"}, Q{"CAA", "CAG"}, E{"GAA", "GAG"},
        G{"GGT", "GGC", "GGA", "GGG"}, H{"CAT", "CAC"}, I{"ATT", "ATC", "ATA"},
        L{"TTA", "TTG", "CTT", "CTC", "CTA", "CTG"}, K{"AAA", "AAG"}, Mc{"ATG"}, F{"TTT", "TTC"},
        P{"CCT", "CCC", "CCA", "CCG"}, S{"TCT", "TCC", "TCA", "TCG", "AGT", "AGC"}, T{"ACT", "ACC", "ACA", "ACG"},
        W{"TGG"}, Y{"TAT", "TAC"}, V{"GTT", "GTC", "GTA", "GTG"}, Stop{"TAA", "TAG", "TGA"};
    switch (aa) {
    case 'A': return A;
    case 'R': return R;
    case 'N': return N;
    case 'D': return D;
    case 'C': return C;
    case 'Q': return Q;
    case 'E': return E;
    case 'G': return G;
    case 'H': return H;
    case 'I': return I;
    case 'L': return L;
    case 'K': return K;
    case 'M': return Mc;
    case 'F': return F;
    case 'P': return P;
    case 'S': return S;
    case 'T': return T;
    case 'W': return W;
    case 'Y': return Y;
    case 'V': return V;
    default: return Stop;
    }
}

double meanOf(const std::vector<double> &x) {
    double s = 0;
    for (double v : x) s += v;
    return s / (double)x.size();
}

double mlZero(const std::vector<double> &s, double lam) {
    double x = 0, y = 0, z = 0;
    for (double v : s) {
        x += v;
        y += std::exp(-lam * v);
        z += v * std::exp(-lam * v);
    }
    return 1 / lam - x / (double)s.size() + z / y;
}

double mlLambda(const std::vector<double> &s) {
    double lo = 1, hi = 1, x, y;
    int it = 0;
    do {
        if (++it > 1000) return 1.0;
        lo /= 2;
        hi *= 2;
        x = mlZero(s, lo);
        y = mlZero(s, hi);
    } while ((x < 0 && y < 0) || (x > 0 && y > 0));
    double gap = hi - lo;
    for (;;) {
        gap /= 2;
        const double mid = lo + gap;
        if (mid <= lo) return lo;
        const double z = mlZero(s, mid);
        if ((x < 0 && z <= 0) || (x > 0 && z >= 0)) lo = mid;
    }
}

}  // namespace

void calibSequence(const float *letterFreqs20, float bgStop, float bgfs1, float bgfs2, int L, int border, int trial,
                   char *out) {
    std::vector<double> w(21);
    for (int k = 0; k < 20; k++) w[k] = letterFreqs20[k] * (int)codonsOf(kAa[k]).size();  // Float * int
    w[20] = bgStop;
    std::discrete_distribution<> aaDist(w.begin(), w.end());
    std::mt19937_64 rng(5489 + trial);
    static const char bases[] = {'A', 'C', 'G', 'T'};
    std::uniform_int_distribution<int> baseDist(0, 3), offsetDist(0, 2);
    std::vector<char> buf((size_t)L + border + 16, 'A');
    const int offset = offsetDist(rng);
    for (int j = 0; j < offset; j++) buf[j] = bases[baseDist(rng)];
    const float both = bgfs1 + bgfs2;
    for (int j = offset; j <= L; j += 3) {
        const double r = std::generate_canonical<double, 10>(rng);
        if (r < bgfs1) {  // 1-base frameshift
            if (j <= L) buf[j] = bases[baseDist(rng)];
            j -= 2;
            continue;
        }
        if (r < both) {  // 2-base frameshift
            if (j <= L) buf[j] = bases[baseDist(rng)];
            if (j + 1 <= L) buf[j + 1] = bases[baseDist(rng)];
            j -= 1;
            continue;
        }
        const int x = aaDist(rng);
        const std::vector<std::string> &cod = codonsOf(x < 20 ? kAa[x] : '*');
        std::uniform_int_distribution<int> pick(0, (int)cod.size() - 1);
        const std::string &c = cod[pick(rng)];
        for (int k = 0; k < 3; k++)
            if (j + k <= L) buf[j + k] = c[k];
    }
    for (int j = 0; j < border; j++) buf[(size_t)L + j] = buf[j];
    std::copy(buf.begin(), buf.begin() + L + border, out);
}

void gumbelFits(std::vector<double> s, double seqLen, double out[8]) {
    std::sort(s.begin(), s.end());
    const double n = (double)s.size();
    const double mu = meanOf(s);
    double v = 0;
    for (double x : s) v += (x - mu) * (x - mu);
    const double pi = 3.1415926535897932, euler = 0.57721566490153286;
    out[0] = pi / std::sqrt(6 * (v / n));                  // MM lambda
    out[1] = std::exp(out[0] * mu - euler) / seqLen;       // MM K
    out[2] = std::exp(mu - euler) / seqLen;                // MM K with lambda 1
    out[3] = mlLambda(s);                                  // ML lambda
    double a = 0;
    for (double x : s) a += std::exp(-out[3] * x);
    out[4] = n / (a * seqLen);                             // ML K
    a = 0;
    for (double x : s) a += std::exp(-x);
    out[5] = n / (a * seqLen);                             // ML K with lambda 1
    double lm = 0;
    for (size_t i = 0; i < s.size(); i++) lm += (double)i * s[i];
    out[6] = std::log(2.0) / (lm / (0.5 * n * (n - 1)) - mu);  // L-moments lambda
    out[7] = std::exp(out[6] * mu - euler) / seqLen;
}

}  // namespace dm

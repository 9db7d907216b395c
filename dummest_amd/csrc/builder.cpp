// Profile builder: Stockholm multiple alignments to the HMMER3/f-style
// profiles the search reads (dummest-build, dummest-build.cc:1-1167).
//
// Per alignment:
//   1. read it (readMultipleAlignment, :297-365), decide protein or DNA
//      (isProteinAlignment, :367-389), map letters and gaps, mark end gaps;
//   2. position-based sequence weights (makeSequenceWeights, :406-444);
//   3. count match/insert/delete events over the columns kept as positions,
//      splitting adjacent gaps into extensions and restarts by a fixed point
//      (countEvents, :446-544);
//   4. Baum-Welch over the ungapped sequences with the profile's own
//      forward/backward (baumWelch, :748-834), unless --countonly;
//   5. counts to probabilities with the Dirichlet mixture and gap priors,
//      with the total count capped so the relative entropy meets its target
//      (countsToProbs, :177-203; entropyWeight, :220-245);
//   6. print (printProfile, :862-914).
// The arithmetic follows the reference's order of operations so that the
// printed values agree with its output.
//
// Letters outside the alphabet (X, B, Z, ...) get a zero likelihood ratio in
// Baum-Welch; the reference reads past its probability table for them.
// Baum-Welch's expected counts run on the GPU (bw.hip); without a HIP device
// only --countonly works.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dummest_amd.h"
#include "dm_host.h"

namespace dm {
namespace {

// HMMER 3.4's gap pseudocounts (priors/gap-priors.hh): match, insStart,
// delStart, insEnd, insExtend, delEnd, delExtend.
const BuildGapPrior kAminoGaps = {0.7939, 0.0278, 0.0135, 0.1551, 0.1331, 0.9002, 0.5630};
const BuildGapPrior kNucleotideGaps = {2, 0.1, 0.1, 0.12, 0.4, 0.5, 1};

// Dirichlet mixtures as HMMER 3.4 ships them: per component the mixture
// coefficient, then one alpha per letter. blocks9: Sjolander et al. 1996
// (priors/blocks9.hh); wheeler4: nhmmer's DNA mixture (priors/wheeler4.hh).
const double kBlocks9[9 * 21] = {
    0.178091,  0.270671, 0.039848, 0.017576, 0.016415, 0.014268, 0.131916, 0.012391, 0.022599, 0.020358, 0.030727,
    0.015315,  0.048298, 0.053803, 0.020662, 0.023612, 0.216147, 0.147226, 0.065438, 0.003758, 0.009621,
    0.056591,  0.021465, 0.010300, 0.011741, 0.010883, 0.385651, 0.016416, 0.076196, 0.035329, 0.013921, 0.093517,
    0.022034,  0.028593, 0.013086, 0.023011, 0.018866, 0.029156, 0.018153, 0.036100, 0.071770, 0.419641,
    0.0960191, 0.561459, 0.045448, 0.438366, 0.764167, 0.087364, 0.259114, 0.214940, 0.145928, 0.762204, 0.247320,
    0.118662,  0.441564, 0.174822, 0.530840, 0.465529, 0.583402, 0.445586, 0.227050, 0.029510, 0.121090,
    0.0781233, 0.070143, 0.011140, 0.019479, 0.094657, 0.013162, 0.048038, 0.077000, 0.032939, 0.576639, 0.072293,
    0.028240,  0.080372, 0.037661, 0.185037, 0.506783, 0.073732, 0.071587, 0.042532, 0.011254, 0.028723,
    0.0834977, 0.041103, 0.014794, 0.005610, 0.010216, 0.153602, 0.007797, 0.007175, 0.299635, 0.010849, 0.999446,
    0.210189,  0.006127, 0.013021, 0.019798, 0.014509, 0.012049, 0.035799, 0.180085, 0.012744, 0.026466,
    0.0904123, 0.115607, 0.037381, 0.012414, 0.018179, 0.051778, 0.017255, 0.004911, 0.796882, 0.017074, 0.285858,
    0.075811,  0.014548, 0.015092, 0.011382, 0.012696, 0.027535, 0.088333, 0.944340, 0.004373, 0.016741,
    0.114468,  0.093461, 0.004737, 0.387252, 0.347841, 0.010822, 0.105877, 0.049776, 0.014963, 0.094276, 0.027761,
    0.010040,  0.187869, 0.050018, 0.110039, 0.038668, 0.119471, 0.065802, 0.025430, 0.003215, 0.018742,
    0.0682132, 0.452171, 0.114613, 0.062460, 0.115702, 0.284246, 0.140204, 0.100358, 0.550230, 0.143995, 0.700649,
    0.276580,  0.118569, 0.097470, 0.126673, 0.143634, 0.278983, 0.358482, 0.661750, 0.061533, 0.199373,
    0.234585,  0.005193, 0.004039, 0.006722, 0.006121, 0.003468, 0.016931, 0.003647, 0.002184, 0.005019, 0.005990,
    0.001473,  0.004158, 0.009055, 0.003630, 0.006583, 0.003172, 0.003690, 0.002967, 0.002772, 0.002686,
};
const double kWheeler4[4 * 5] = {
    0.24, 0.16, 0.45, 0.12, 0.39,  //
    0.26, 0.09, 0.03, 0.09, 0.04,  //
    0.08, 1.29, 0.40, 6.58, 0.51,  //
    0.42, 1.74, 1.49, 1.57, 1.95,
};

// HMMER 3.4 defaults (dummest-build.cc:26-34)
constexpr double kEreAmino = 0.59, kEreNucleotide = 0.62;

struct Mixture {
    const double *params;
    int components;
};

struct Alignment {
    std::string name;
    int nseq = 0, ncol = 0;
    std::vector<unsigned char> cells;  // nseq x ncol, row-major
    std::string rf;                    // RF annotation, empty when absent
};

[[noreturn]] void fatal(const char *what) { throw BuildError(what); }

bool isGapChar(char c) { return c == '.' || c == '-' || c == '_' || c == '~'; }

double sumOf(const double *x, int n) { return std::accumulate(x, x + n, 0.0); }

void normalizeInPlace(double *x, int n) {
    const double s = sumOf(x, n);
    for (int i = 0; i < n; i++) x[i] /= s;
}

// Checks a finished block of sequence lines (checkAlignmentBlock, :275-295).
void endBlock(Alignment &a, size_t lineTotal, const std::vector<std::string> &rf, size_t lastLen, int blockLines) {
    if (blockLines) {
        if (a.nseq && blockLines != a.nseq) fatal("alignment blocks should have same number of sequences");
        a.nseq = blockLines;
    }
    if (!rf.empty()) {
        const size_t blocks = a.nseq ? lineTotal / a.nseq : 0;
        if (rf.size() != blocks) fatal("bad RF lines");
        if (rf.back().size() != lastLen) fatal("RF length != aligned sequence length");
    }
}

// One Stockholm alignment; false at the end of the input.
bool readAlignment(std::istream &in, Alignment &a) {
    std::vector<std::string> lines, rfLines;
    std::string line, word;
    int blockLines = 0;
    a = Alignment();
    a.name = "Anon";
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        if (!(ls >> word)) {
            endBlock(a, lines.size(), rfLines, lines.empty() ? 0 : lines.back().size(), blockLines);
            blockLines = 0;
            continue;
        }
        if (word == "#=GF") {
            if ((ls >> word) && word == "ID") ls >> a.name;
        } else if (word == "//") {
            break;
        } else if (word == "#=GC") {
            if (!(ls >> word) || word != "RF") continue;
            if (!(ls >> word)) fatal("bad RF line");
            if (a.nseq && rfLines.empty()) fatal("bad RF lines");
            rfLines.push_back(word);
        } else if (word[0] != '#') {
            if (!(ls >> word)) fatal("bad alignment file");
            if (blockLines) {
                if (word.size() != lines.back().size()) fatal("aligned sequences should have the same length");
            } else {
                a.ncol += (int)word.size();
            }
            lines.push_back(word);
            ++blockLines;
        }
    }
    const bool more = (bool)in;
    endBlock(a, lines.size(), rfLines, lines.empty() ? 0 : lines.back().size(), blockLines);
    a.cells.resize((size_t)a.nseq * a.ncol);
    size_t k = 0;
    for (int s = 0; s < a.nseq; s++)
        for (size_t b = s; b < lines.size(); b += a.nseq)
            for (char c : lines[b]) a.cells[k++] = (unsigned char)c;
    for (const std::string &r : rfLines) a.rf += r;
    return more;
}

// Protein unless at least 90% of the letters are nucleotide-like (:367-389).
bool looksLikeProtein(const Alignment &a) {
    long counts[256] = {0};
    for (unsigned char c : a.cells) ++counts[c];
    long letters = 0, nucleotide = 0;
    for (int c = 'A'; c <= 'Z'; c++) letters += counts[c] + counts[std::tolower(c)];
    for (const char *p = "ACGTNU"; *p; ++p) nucleotide += counts[(unsigned char)*p] + counts[std::tolower(*p)];
    return 1.0 * nucleotide / letters < 0.9;
}

// Columns that become profile positions (a "match" column).
inline bool isMatchColumn(bool hand, const Alignment &a, int col, double nonGap, double total, double symfrac) {
    return hand ? !isGapChar(a.rf[col]) : (nonGap > 0 && nonGap >= symfrac * total);
}

// Henikoff position-based weights, scaled so the largest is 1 (:406-444).
std::vector<double> sequenceWeights(const Alignment &a, int A, bool hand, double symfrac) {
    const int midGap = A + 1, endGap = A + 2;
    std::vector<double> w(a.nseq, 0.0);
    std::vector<int> used(a.nseq, 0);
    for (int col = 0; col < a.ncol; col++) {
        int n[32] = {0};
        for (int s = 0; s < a.nseq; s++) ++n[a.cells[(size_t)s * a.ncol + col]];
        const int total = a.nseq - n[endGap];
        const int nonGap = total - n[midGap];
        // (this test has no nonGap > 0 condition, unlike the event counts)
        if (hand ? isGapChar(a.rf[col]) : !(nonGap >= symfrac * total)) continue;
        int kinds = 0;
        for (int k = 0; k < A; k++) kinds += n[k] > 0;
        for (int s = 0; s < a.nseq; s++) {
            const int x = a.cells[(size_t)s * a.ncol + col];
            if (x < A) {
                w[s] += 1.0 * nonGap / (kinds * n[x]);
                ++used[s];
            }
        }
    }
    for (int s = 0; s < a.nseq; s++)
        if (used[s]) w[s] /= used[s];
    const double top = *std::max_element(w.begin(), w.end());
    if (top)
        for (double &x : w) x /= top;
    return w;
}

// (count1 r + pseudo1) / ((count1 + count2) r + pseudo1 + pseudo2), where r
// caps the total count at maxTotal (getProb, :107-114).
double cappedRatio(double c1, double c2, double p1, double p2, double maxTotal) {
    const double s = c1 + c2;
    if (s + (p1 + p2) == 0.0) return 0.5;
    const double r = (s > maxTotal) ? maxTotal / s : 1.0;
    return (c1 * r + p1) / (s * r + (p1 + p2));
}

// Event counts per position, width 7 + A (countEvents, :446-544): match,
// not-insert, delete start, insert end (= insert start), insert extend, delete
// end, delete extend, then the letter counts; one more row for the end.
std::vector<double> eventCounts(const Alignment &a, int A, bool hand, double symfrac, const std::vector<double> &w,
                                double wsum, const BuildGapPrior &gp, std::vector<int> &columns) {
    const int midGap = A + 1, endGap = A + 2;
    std::vector<double> out;
    std::vector<char> state(a.nseq, 0);
    double notIns = 0, insBeg = 0, insExt = 0, delEnd = 0;
    for (int col = 0; col < a.ncol; col++) {
        double n[32] = {0};
        for (int s = 0; s < a.nseq; s++) n[a.cells[(size_t)s * a.ncol + col]] += w[s];
        const double total = wsum - n[endGap];
        const double nonGap = total - n[midGap];
        if (isMatchColumn(hand, a, col, nonGap, total, symfrac)) {
            columns.push_back(col);
            double delBeg = 0, delExt = 0;
            for (int s = 0; s < a.nseq; s++) {
                const int x = a.cells[(size_t)s * a.ncol + col];
                if (x <= A) {
                    if (state[s] <= 'd') notIns += w[s];
                    if (state[s] == 'd') delEnd += w[s];
                    state[s] = 'm';
                } else if (x == midGap) {
                    if (state[s] != 'd') delBeg += w[s];
                    if (state[s] == 'd') delExt += w[s];
                    state[s] = 'd';
                }
            }
            // Adjacent gap letters count partly as restarts: iterate the
            // extension fractions to their fixed point.
            double iFrac = 1, dFrac = 1, iFrac0, dFrac0;
            double iExt, dExt, iBeg, dBeg, dEnd, nIns;
            const double gpNotIns = gp.match + gp.delStart;
            do {
                iExt = insExt * iFrac;
                dExt = delExt * dFrac;
                iBeg = insBeg + (insExt - iExt);
                dBeg = delBeg + (delExt - dExt);
                dEnd = delEnd + (delExt - dExt);
                nIns = notIns + (delExt - dExt);
                const double pa = cappedRatio(iBeg, nIns, gp.insStart, gpNotIns, 1e37);
                const double pd = cappedRatio(dBeg, nonGap, gp.delStart, gp.match, 1e37);
                const double pb = cappedRatio(iExt, iBeg, gp.insExt, gp.insEnd, 1e37);
                const double pe = cappedRatio(dExt, dEnd, gp.delExt, gp.delEnd, 1e37);
                iFrac0 = iFrac;
                dFrac0 = dFrac;
                if (insExt > 0) iFrac = pb / (pb + pa * (1 - pb));
                if (delExt > 0) dFrac = pe / (pe + (1 - pa) * pd * (1 - pe));
            } while (std::fabs(iFrac - iFrac0) > 1e-6 || std::fabs(dFrac - dFrac0) > 1e-6);
            const double row[7] = {nonGap, nIns, dBeg, iBeg, iExt, dEnd, dExt};
            out.insert(out.end(), row, row + 7);
            out.insert(out.end(), n, n + A);
            insBeg = insExt = delEnd = 0;
            notIns = nonGap;
        } else {
            for (int s = 0; s < a.nseq; s++) {
                const int x = a.cells[(size_t)s * a.ncol + col];
                if (x <= A) {
                    if (state[s] <= 'd') notIns += w[s];
                    if (state[s] == 'd') delEnd += w[s];
                    if (state[s] != 'i') insBeg += w[s];
                    if (state[s] == 'i') insExt += w[s];
                    state[s] = 'i';
                }
            }
        }
    }
    out.push_back(0.0);
    out.push_back(notIns);
    out.insert(out.end(), 5 + A, 0.0);
    return out;
}

// Posterior mean of a Dirichlet mixture given letter counts, with the total
// count capped at maxTotal (applyDirichletMixture, :116-147).
void mixtureEstimate(const Mixture &mix, int A, double maxTotal, const double *counts, double *out) {
    const int stride = 1 + A;
    const double total = sumOf(counts, A);
    const double r = (total > maxTotal) ? maxTotal / total : 1.0;
    std::vector<double> logw(mix.components, 0.0);
    std::fill_n(out, A, 0.0);
    if (mix.components > 1) {
        for (int c = 0; c < mix.components; c++) {
            const double *alpha = mix.params + c * stride + 1;
            const double alphaSum = sumOf(alpha, A);
            logw[c] = std::lgamma(alphaSum) - std::lgamma(total * r + alphaSum + 1);
            for (int k = 0; k < A; k++) logw[c] += std::lgamma(counts[k] * r + alpha[k]) - std::lgamma(alpha[k]);
        }
    }
    const double top = *std::max_element(logw.begin(), logw.end());
    for (int c = 0; c < mix.components; c++) {
        const double *alpha = mix.params + c * stride + 1;
        const double m = mix.params[c * stride] * std::exp(logw[c] - top);
        for (int k = 0; k < A; k++) out[k] += m * (counts[k] * r + alpha[k]);
    }
    normalizeInPlace(out, A);
}

// Transition probabilities of one position from its event counts
// (gapCountsToProbs, :149-175).
void gapProbs(const BuildGapPrior &gp, double maxTotal, const double *c, double *p) {
    const double match = c[0], notIns = c[1], delBeg = c[2], insBeg = c[3], insExt = c[4], delEnd = c[5],
                 delExt = c[6];
    const double a = cappedRatio(insBeg, notIns, gp.insStart, gp.match + gp.delStart, maxTotal);
    const double d = cappedRatio(delBeg, match, gp.delStart, gp.match, maxTotal);
    p[1] = a;
    p[0] = (1 - a) * (1 - d);
    p[2] = (1 - a) * d;
    const double b = cappedRatio(insExt, insBeg, gp.insExt, gp.insEnd, maxTotal);
    p[3] = 1 - b;
    p[4] = b;
    const double e = cappedRatio(delExt, delEnd, gp.delExt, gp.delEnd, maxTotal);
    p[5] = 1 - e;
    p[6] = e;
}

// All probabilities from counts (countsToProbs, :177-203): the background
// row is the normalized geometric mean of the letter probabilities.
void probabilities(const Mixture &mix, const BuildGapPrior &gp, int A, double maxTotal, int L, const double *counts,
                   double *probs) {
    const int width = 7 + A;
    for (int i = 0;; i++) {
        gapProbs(gp, maxTotal, counts + i * width, probs + i * width);
        if (i == L) break;
        mixtureEstimate(mix, A, maxTotal, counts + i * width + 7, probs + i * width + 7);
    }
    probs[5] = 1;
    probs[6] = 0;
    double *end = probs + (size_t)L * width;
    end[0] = 1 - end[1];
    end[2] = 0;
    for (int k = 0; k < A; k++) {
        double s = 0.0;
        for (int i = 0; i < L; i++) s += std::log(std::max(probs[i * width + 7 + k], 1e-6));
        end[7 + k] = std::exp(s / L);
    }
    normalizeInPlace(end + 7, A);
}

// Relative entropy (bits) of the letter probabilities to the background.
double relativeEntropyBits(const double *probs, int A, int L) {
    const int width = 7 + A;
    const double *bg = probs + (size_t)L * width + 7;
    double r = 0;
    for (int i = 0; i < L; i++) {
        const double *p = probs + i * width + 7;
        for (int k = 0; k < A; k++)
            if (p[k] > 0) r += p[k] * std::log2(p[k] / bg[k]);
    }
    return r;
}

// Bisection on the total-count cap until the relative entropy meets the
// target; returns the effective sequence number (entropyWeight, :220-245).
double capForEntropy(const Mixture &mix, const BuildGapPrior &gp, int A, double rangeEnd, double target, int L,
                     const double *counts, double *probs, int verbosity, std::ostream &log) {
    double lo = 0, len = rangeEnd, cap = 1e37;
    for (;;) {
        probabilities(mix, gp, A, cap, L, counts, probs);
        if (verbosity > 1) {
            const double *bg = probs + (size_t)L * (7 + A) + 7;
            log << "Background letter probabilities:" << std::setprecision(3);
            for (int k = 0; k < A; k++) log << " " << bg[k];
            log << std::setprecision(6) << "\n";
        }
        const double r = relativeEntropyBits(probs, A, L);
        if (verbosity) log << "Max total weight: " << cap << "  Relative entropy: " << r << "\n";
        if (cap >= rangeEnd && r <= target) return rangeEnd;
        if (len < 0.01) return cap;
        if (r < target) lo = cap;
        len /= 2;
        cap = lo + len;
    }
}

// ----------------------------------------------------------- Baum-Welch
// Iterates expected counts -> probabilities until no probability moves by
// more than maxDiff (baumWelch, :748-834). The expected counts of each
// iteration (forward, backward and calculateTransitionCounts over every
// sequence, :561-736) come from the gfx950 kernels (dm_bw_estep); the delete
// starts, the priors and the convergence test stay here.
void baumWelch(std::vector<double> &counts, const Alignment &a, int A, const Mixture &mix, const BuildGapPrior &gp,
               int L, const std::vector<double> &weights, double maxIter, double maxDiff, dm_engine *E,
               std::ostream &log) {
    const int width = 7 + A;
    std::vector<double> oldP(counts.size()), newP(counts.size());
    std::vector<uint8_t> letters;
    std::vector<int64_t> offsets(a.nseq);
    std::vector<int32_t> lengths(a.nseq, 0);
    for (int s = 0; s < a.nseq; s++) {
        offsets[s] = (int64_t)letters.size();
        for (int c = 0; c < a.ncol; c++) {
            const unsigned char x = a.cells[(size_t)s * a.ncol + c];
            if (x <= A) {
                letters.push_back(x);
                lengths[s]++;
            }
        }
    }
    if (dm_bw_prepare(E, A, L, a.nseq, letters.data(), offsets.data(), lengths.data(), weights.data()))
        throw BuildError((std::string("Baum-Welch on the GPU: ") + dm_engine_error(E)).c_str(), false);
    std::vector<double> totals(std::max(1, a.nseq));
    probabilities(mix, gp, A, 1e37, L, counts.data(), oldP.data());
    bool converged = false;
    int iter = 0;
    do {
        if (dm_bw_estep(E, oldP.data(), counts.data(), totals.data()))
            throw BuildError((std::string("Baum-Welch on the GPU: ") + dm_engine_error(E)).c_str(), false);
        for (int s = 0; s < a.nseq; s++)
            if (!std::isfinite(totals[s])) throw BuildError("numbers overflowed to infinity in Baum-Welch: quitting", false);
        // delete starts from the delete ends and extensions around them
        for (int i = 1; i <= L; i++) {
            const double d = counts[(size_t)i * width + 5] + counts[(size_t)i * width + 6] -
                             counts[(size_t)(i - 1) * width + 6];
            counts[(size_t)(i - 1) * width + 2] = std::max(d, 0.0);
        }
        probabilities(mix, gp, A, 1e37, L, counts.data(), newP.data());
        double moved = 0.0;
        for (size_t k = 0; k < oldP.size(); k++) moved = std::max(moved, std::fabs(oldP[k] - newP[k]));
        converged = moved <= maxDiff;
        oldP.swap(newP);
        log << "\rIteration " << ++iter << " / " << maxIter << ": " << (converged ? "    converged" : "not converged");
    } while (!converged && iter < maxIter);
    log << "\nError tolerance: " << maxDiff << ", iterations: " << iter << ", "
        << (converged ? "Baum-Welch converged" : "Baum-Welch not converged") << "\n";
}

// Most probable letter; uppercase when it holds at least half (protein) or
// 90% (DNA) of the position's probability (getConsensusAt, :843-850).
char consensusLetter(const double *p, int A, const char *alphabet) {
    const double thr = (A > 4) ? 0.5 : 0.9;
    const int k = (int)(std::max_element(p, p + A) - p);
    const double s = sumOf(p, A);
    return (p[k] >= thr * s) ? (char)std::toupper(alphabet[k]) : (char)std::tolower(alphabet[k]);
}

void putValue(std::ostream &out, bool counts, double v) {
    if (counts) out << "  " << std::left << std::setw(8) << v << std::right;
    else if (v > 0) out << "  " << std::fabs(std::log(v));
    else out << "        *";
}

void writeProfile(std::ostream &out, const double *p, const std::vector<int> &columns, const char *alphabet, int L,
                  const Alignment &a, double neff, bool counts, const BuildOptions &o) {
    const int A = (int)std::strlen(alphabet), width = 7 + A;
    const double *bg = p + (size_t)L * width + 7;
    out << "HMMER3/f [" << o.versionText << "]\n";
    out << "NAME  " << a.name << "\n";
    out << "LENG  " << L << "\n";
    out << "ALPH  " << (A > 4 ? "amino" : "DNA") << "\n";
    out << "RF    " << (a.rf.empty() ? "no" : "yes") << "\n";
    out << "CONS  yes\nMAP   yes\n";
    out << "NSEQ  " << a.nseq << "\n";
    out << "EFFN  " << neff << "\n";
    out << "BM   " << o.bmText << "\n";
    out << "HMM     ";
    for (int k = 0; k < A; k++) out << "     " << alphabet[k] << (counts ? "    " : "   ");
    out << "\n";
    out << "           match   "
        << (counts ? " notIns    delStart  insEnd    insExt    delEnd    delExt\n"
                   : "insStart delStart insEnd   insExt   delEnd   delExt\n");
    if (counts) out << std::defaultfloat;
    out.precision(counts ? 3 : 5);
    for (int i = 0;; i++) {
        const double *row = p + (size_t)i * width;
        out << "        ";
        for (int k = 0; k < A; k++) putValue(out, counts, bg[k]);
        out << "\n        ";
        for (int k = 0; k < 7; k++) putValue(out, counts, row[k]);
        out << "\n";
        if (i == L) break;
        out << std::setw(7) << i + 1 << " ";
        for (int k = 0; k < A; k++) putValue(out, counts, row[7 + k]);
        out << std::setw(7) << columns[i] + 1 << " " << consensusLetter(row + 7, A, alphabet) << " "
            << (a.rf.empty() ? '-' : a.rf[columns[i]]) << " - -\n";
    }
    out.precision(6);
    out << "//\n";
}

}  // namespace

// -------------------------------------------------------------- options
bool readGapPrior(std::istream &in, BuildGapPrior &gp) {
    in >> gp.match >> gp.insStart >> gp.delStart >> gp.insEnd >> gp.insExt >> gp.delEnd >> gp.delExt;
    return (bool)in;
}

bool readMixture(std::istream &in, std::vector<double> &params, int &components) {
    int A = 0;
    components = 0;
    in >> A >> components;
    double d = 0;
    for (int c = 0; c < components; c++) {
        in >> d;
        params.push_back(d);
        for (int k = 0; k < A; k++) {
            in >> d;
            params.push_back(d);
        }
    }
    if (components < 1) in.setstate(std::ios::failbit);
    return (bool)in;
}

int buildProfiles(std::istream &in, const BuildOptions &o, std::ostream &out, std::ostream &log) {
    Alignment a;
    unsigned long count = 0;
    // the GPU is opened when the first Baum-Welch starts (--countonly needs none)
    std::unique_ptr<dm_engine, void (*)(dm_engine *)> engine(nullptr, dm_engine_destroy);
    for (;;) {
        const bool more = readAlignment(in, a);
        if (!more) break;
        out << std::fixed;
        log << "MSA #" << ++count << ": " << a.name << std::endl;
        if (o.hand && a.rf.empty()) {
            log << "missing RF annotation line: needed for --hand\n";
            return 1;
        }
        const bool protein = looksLikeProtein(a);
        const char *alphabet = protein ? "ACDEFGHIKLMNPQRSTVWY" : "ACGT";
        const int A = (int)std::strlen(alphabet);
        unsigned char code[256];
        std::memset(code, A, sizeof code);
        auto mapLetters = [&](const char *letters) {
            for (int k = 0; letters[k]; k++)
                code[std::toupper(letters[k])] = code[std::tolower(letters[k])] = (unsigned char)k;
        };
        mapLetters(alphabet);
        mapLetters(protein ? "AUDEFGHIO" : "ACGU");  // selenocysteine as C, pyrrolysine as K; U as T
        code['-'] = code['.'] = code['_'] = code['~'] = (unsigned char)(A + 1);
        const double ere = (o.ere > 0) ? o.ere : protein ? kEreAmino : kEreNucleotide;
        Mixture mix;
        std::vector<double> flat;
        if (!o.mixture.empty()) {
            if ((int)o.mixture.size() / o.mixtureComponents - 1 != A) {
                log << "the Dirichlet mixture file has wrong alphabet size\n";
                return 1;
            }
            mix = {o.mixture.data(), o.mixtureComponents};
        } else if (o.pnone) {
            flat.assign(1 + A, 0.0);
            flat[0] = 1;
            mix = {flat.data(), 1};
        } else {
            mix = protein ? Mixture{kBlocks9, 9} : Mixture{kWheeler4, 4};
        }
        const BuildGapPrior &gp = (o.hasGapPrior || o.pnone) ? o.gapPrior : protein ? kAminoGaps : kNucleotideGaps;
        if (o.verbosity) log << "Alignment length: " << a.ncol << "\n";
        for (unsigned char &c : a.cells) c = code[c];
        // leading and trailing gaps of each row are end gaps (A + 2)
        for (int s = 0; s < a.nseq; s++) {
            unsigned char *row = a.cells.data() + (size_t)s * a.ncol;
            for (int c = 0; c < a.ncol && row[c] == A + 1; c++) row[c] = (unsigned char)(A + 2);
            for (int c = a.ncol; c > 0 && row[c - 1] == A + 1; c--) row[c - 1] = (unsigned char)(A + 2);
        }
        std::vector<double> weights = sequenceWeights(a, A, o.hand, o.symfrac);
        const double wsum = std::accumulate(weights.begin(), weights.end(), 0.0);
        if (o.verbosity > 1) {
            log << "Sequence weights:\n";
            log.precision(3);
            for (double x : weights) log << x << "\n";
            log.precision(6);
        }
        if (o.verbosity) log << "Total weight: " << wsum << "\n";
        std::vector<int> columns;
        std::vector<double> counts = eventCounts(a, A, o.hand, o.symfrac, weights, wsum, gp, columns);
        const int L = (int)columns.size();
        if (!o.countOnly) {
            if (!engine) {
                dm_engine *E = nullptr;
                if (dm_engine_create(o.device, &E)) throw BuildError("no usable HIP device for Baum-Welch", false);
                engine.reset(E);
            }
            baumWelch(counts, a, A, mix, gp, L, weights, o.maxIter, o.maxDiff, engine.get(), log);
        }
        if (o.counts) {
            writeProfile(out, counts.data(), columns, alphabet, L, a, wsum, true, o);
            continue;
        }
        std::vector<double> probs(counts.size());
        const double target = std::max(o.esigma, ere * L);
        if (o.verbosity) log << "Target relative entropy: " << target << "\n";
        const double neff = capForEntropy(mix, gp, A, wsum, target, L, counts.data(), probs.data(), o.verbosity, log);
        writeProfile(out, probs.data(), columns, alphabet, L, a, neff, false, o);
    }
    return 0;
}

}  // namespace dm

// dummest-amd: the engine executable, a drop-in for the reference's
// dummestl_core (dummest.cc:18-660) with the DP on the GPU.
//
// Same argv, stdin/FASTA input (pipeline headers included) and output records:
//   1. read and finalize the profiles (host, dummest-profile.hh:235-329);
//   2. calibrate every profile in one batched GPU pass (dm_calibrate;
//      estimateK, dummest-calibration.hh:338-600) and print the same
//      per-profile header lines (dummest.cc:384-413, dummest-calibration.hh:603-649);
//   3. read the sequences in contigs (readContig, dummest-sequence.hh:6-89),
//      total length first (dummest.cc:508-528), then in chunks of --batch;
//   4. per chunk: score-only filter wave on the GPU and the E-value test of
//      the filter (dummest-search.hh:146-179), then the forward+backward score
//      wave for the survivors and the anchor gate (dummest-dp-forward.hh:252-274),
//      then the matrices of each gated pair and the host traceback
//      (dm_pair_hits), and the records (dummest-search.hh:89-105,
//      dummest-output.hh:54-107).
//
// tantan masking (-m, default 3): profiles are masked on the host when they are
// read (finalizeProfile, dummest-profile.hh:105-233), sequences per chunk on
// the GPU (dm_tantan_dna) before they are decoded (dummest.cc:557-573); masked
// letters print in lowercase (dummest-output.hh:6-30).
//
// Calibration cache (ProfileCache, dummest-calibration.hh:191-336): the same
// file and keys as the reference's, so a profile calibrated once is looked up
// instead of calibrated again (unless -S); only the misses go to the GPU.
//
// The Gaussian profile filter (-d/-D) is applied on the host when the
// profiles are read (filterLetterProbabilities, dummest-profile.hh:64-103).
#include <getopt.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <numeric>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "dm_host.h"
#include "../../include/dummest_amd.h"

using namespace dm;

namespace {

const char *kAlphabet = "ACDEFGHIKLMNPQRSTVWYUO?*";  // getAlphabet(20), dummest-core.hh:350-355

const double kDefaults[6] = {0.0171, 0.0018, 0.0328, 0.0083, 0.0005, 0.046875};

int badOpt() {
    std::cerr << "dummest-amd: bad option value\n";
    return 1;
}

int intFromText(const char *s) {
    char *end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (!end || *end || v < INT32_MIN || v > INT32_MAX) return -1;
    return (int)v;
}

// Decimal log-score text: fixed, 6 decimals, trailing zeros trimmed
// (scoreStr, dummest-output.hh:48-66).
std::string scoreStr(double v) {
    if (v == 0.0) v = 0.0;
    std::ostringstream os;
    if (!std::isfinite(v)) {
        os << v;
        return os.str();
    }
    os << std::fixed << std::setprecision(6) << v;
    std::string s = os.str();
    if (s.find('.') != std::string::npos) {
        s.erase(s.find_last_not_of('0') + 1);
        if (s.back() == '.') s.pop_back();
    }
    return s;
}

double natsToLog2(double ln) { return ln / 0.6931471805599453; }

int numOfDigits(int x) {
    int n = 0;
    do ++n;
    while (x /= 10);
    return n;
}

// ---------------------------------------------------------------------------
// Sequence reading (readContig, dummest-sequence.hh:6-89, PIPELINE_MODE on).
// The record state persists from one record to the next as the reference's
// reused Sequence object does; is_plus starts true.
// ---------------------------------------------------------------------------
struct SeqState {
    std::string name;
    int length = 0;  // letters of the record read so far (contigs and other symbols)
    size_t wStart = 0, wEnd = 0, trueLength = 0;
    std::string targetProfile;
    bool isPlus = true;
    bool hasPipelineFields = false;
};

struct Contig {
    int start = 0;
    int length = 0;
};

class ContigReader {
public:
    explicit ContigReader(std::istream &in) : in_(in) {
        std::memset(map_, 127, sizeof map_);
        std::memset(map_, 125, ' ' + 1);
        map_[(unsigned char)'>'] = 126;
        for (int i = 0; kAlphabet[i]; i++) {
            map_[(unsigned char)std::toupper(kAlphabet[i])] = (char)i;
            map_[(unsigned char)std::tolower(kAlphabet[i])] = (char)i;
        }
        for (const char *c = "RYSWKMBDHVN"; *c; c++) {  // IUPAC ambiguity (dummest.cc:445-456)
            map_[(unsigned char)*c] = 22;
            map_[(unsigned char)std::tolower(*c)] = 22;
        }
    }
    // Next contig into `letters` (alphabet indices); false at the end or on a
    // format error (err set).
    bool next(SeqState &s, Contig &c, std::vector<char> &letters, std::string &err) {
        letters.clear();
        std::streambuf *buf = in_.rdbuf();
        if (c.length == 0) {
            char x;
            if (!(in_ >> x)) return false;
            if (x != '>') {
                err = "bad sequence data: no '>'";
                return false;
            }
            std::string line, word;
            std::getline(in_, line);
            std::istringstream iss(line);
            if (!(iss >> word)) {
                err = "bad sequence data: no name";
                return false;
            }
            const size_t slash = word.find('/');
            if (slash != std::string::npos) {
                const std::string chr = word.substr(0, slash), range = word.substr(slash + 1);
                const size_t dash = range.find('-');
                if (dash != std::string::npos) {
                    s.wStart = std::stoul(range.substr(0, dash));
                    s.wEnd = std::stoul(range.substr(dash + 1));
                }
                std::string tok, length, profile, strand;
                bool haveStrand = false;
                while (iss >> tok) {
                    if (tok.rfind("length=", 0) == 0) length = tok;
                    else if (tok.rfind("profile=", 0) == 0) profile = tok;
                    else if (tok.rfind("bitscore=", 0) == 0) continue;
                    else if (!haveStrand) {
                        strand = tok;
                        haveStrand = true;
                    }
                }
                if (!length.empty() && !profile.empty() && haveStrand) {
                    s.trueLength = std::stoull(length.substr(length.find('=') + 1));
                    s.targetProfile = profile.substr(profile.find('=') + 1);
                    s.isPlus = strand.find("plus_strand") != std::string::npos;
                }
                s.hasPipelineFields = true;
                word = chr;
            }
            s.name = word;
            s.length = 0;
        }
        int ch = buf->sgetc();
        while (ch != EOF && ch != '>') {
            if (map_[ch & 255] < 125) break;
            if (ch > ' ') ++s.length;
            ch = buf->snextc();
        }
        while (ch != EOF && map_[ch & 255] < 126) {
            if (ch > ' ') letters.push_back(map_[ch & 255]);
            ch = buf->snextc();
        }
        if (letters.size() > (size_t)INT32_MAX - 64) {
            err = "sequence is too long!";
            return false;
        }
        c.start = s.length;
        c.length = (int)letters.size();
        s.length += c.length;
        return true;
    }

private:
    std::istream &in_;
    char map_[256];
};

// One contig searched as one strand (strandNum even: PIPELINE_MODE fixes the
// strand option to 1, dummest.cc:460-462, so only s = 0 runs).
struct Target {
    SeqState rec;   // record state when the contig was read
    Contig contig;
    std::string letters;  // alphabet letters (uppercase) of the contig
    std::vector<uint8_t> nums;  // the reader's letter numbers (alphabet indices, 22 = IUPAC)
    std::vector<char> masked;   // tantan-masked positions (-m 2)
    std::vector<uint8_t> codes;
};

// Aligned-row letters (dummest-output.hh:6-30): masked profile positions carry
// bit 32 of the consensus code; a masked sequence letter prints in lowercase
// when the mask code (22) is above its own code.
char profLetter(char code) { return (char)(kAlphabet[code & 31] + (code & 32)); }
char seqLetter(const Target &t, int pos) {
    const bool low = !t.masked.empty() && t.masked[pos] && t.nums[pos] < 22;
    return (char)(t.letters[pos] + (low ? 32 : 0));
}

struct Options {
    double evalue = 10;
    long long totOverride = -1;
    int mask = 3;
    int trials = 1000, length = 5000, border = 100;
    int batch = 0;
    int bgType = 'G';
    const char *scoresFile = nullptr;
    bool warmup = false;
    bool maxMode = false;
    double filterEvalue = -1;
    bool filterExplicit = false;
    int threads = 1;
    int verbosity = 0;
    double rates[6];
    double tantan = 0.5;
    int device = 0;
};

int die(const std::string &msg) {
    std::cerr << "dummest-amd: " << msg << "\n";
    return 1;
}

// The process's calibration cache, as the reference's global ProfileCache
// (dummest-calibration.hh:191-336): read on first use, written back when the
// process ends. The reference prints the file name when it first reads it
// (in the first profile's calibration step, or at exit when no profile got
// that far); announce() prints it at that point here, since the lookups of
// all profiles happen before the one batched calibration pass.
class CalibrationCache {
public:
    ~CalibrationCache() {
        load();
        announce();
        std::vector<std::pair<std::string, CacheEntry>> v(entries_.begin(), entries_.end());
        writeCalibrationCache(path_, v);  // an unwritable file is skipped silently, as in save()
    }
    void load() {
        if (loaded_) return;
        loaded_ = true;
        path_ = calibrationCachePath();
        std::vector<std::pair<std::string, CacheEntry>> v;
        std::string err;
        if (readCalibrationCache(path_, v, err)) {
            for (auto &kv : v) entries_.emplace(std::move(kv.first), kv.second);  // the first of equal keys stays
        } else {
            std::cerr << "# Cache Load Error: " << err << ".\n";
            std::error_code ec;
            std::filesystem::remove(path_, ec);
        }
    }
    void announce() {
        if (announced_) return;
        announced_ = true;
        std::cout << "# Cache file: " << std::filesystem::path(path_) << std::endl;
    }
    bool find(const std::string &key, CacheEntry &e) const {
        const auto it = entries_.find(key);
        if (it == entries_.end()) return false;
        e = it->second;
        return true;
    }
    void store(const std::string &key, const CacheEntry &e) { entries_[key] = e; }
    const std::string &binaryHashHex() {
        if (binHash_.empty()) binHash_ = binaryHash();
        return binHash_;
    }

private:
    bool loaded_ = false, announced_ = false;
    std::string path_, binHash_;
    std::unordered_map<std::string, CacheEntry> entries_;
} gCache;

// The cache entry of one calibration: the mid-anchored fits f[0..7] and the
// filter's method-of-moments fits f[8..10] of dm_calibrate, and the fits x of
// the all -inf end- and start-anchored samples (estimateK's entry,
// dummest-calibration.hh:584-599).
CacheEntry entryFromFits(const double *f, const double *x) {
    CacheEntry e;
    for (int g = 0; g < 8; g++) {
        e.v[3 * g] = x[g];
        e.v[3 * g + 1] = x[g];
        e.v[3 * g + 2] = f[g];
    }
    e.v[24] = f[8];
    e.v[25] = f[9];
    e.v[26] = f[10];
    return e;
}

int engineFail(dm_engine *e, const char *what, int rc) {
    std::cerr << "dummest-amd: " << what << " failed (" << rc << "): " << dm_engine_error(e) << "\n";
    return 1;
}

}  // namespace

// DM_TIMING set in the environment: the run's wall time and the part spent
// in the post-gate step (dm_pairs_hits), on stderr at the end.
static double wallNow() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
static double hitsWall = 0, hitsDevice = 0;
// wall time per phase of the run (DM_TIMING)
enum { kTProfiles, kTCalib, kTInput, kTPrepare, kTFilter, kTScores, kTOutput, kTPhases };
static double phaseWall[kTPhases] = {};
static const char *const phaseName[kTPhases] = {"profiles", "calibration", "input", "mask+decode+upload",
                                                "filter", "scores", "records"};

int main(int argc, char *argv[]) {
    const double runStart = wallNow();
    Options o;
    std::copy(kDefaults, kDefaults + 6, o.rates);
    {
        const unsigned hc = std::thread::hardware_concurrency();
        o.threads = hc ? (int)hc : 1;
    }
    double filterStdDev = 0;
    bool keepNonvarying = false;
    const char help[] =
        "usage: dummest-amd profiles.hmm [sequences.fa]\n"
        "\n"
        "Find similarities between DNA sequences and protein profiles, allowing\n"
        "frameshifts; the dynamic programming runs on the GPU.\n"
        "\n"
        "Options:\n"
        "  -h, --help        show this help message and exit\n"
        "  -V, --version     show version and exit\n"
        "  -v, --verbose     show progress messages\n"
        "  -T N, --threads N host threads (default: CPU cores)\n"
        "  -e E, --evalue E  find similarities with E-value <= this (default: 10)\n"
        "  -N N, --total-length N  override total search space size (default: sum of input lengths)\n"
        "  -s S, --strand S  accepted for compatibility; each record is searched as given\n"
        "  -m M, --mask M    mask simple regions of: 0=neither, 1=profile, 2=sequence, 3=both (default: 3)\n"
        "  -d D, --dev D     standard deviation of a Gaussian high-pass filter on position-specific\n"
        "                    letter probabilities\n"
        "  -D D, --Dev D     same as above, but keep the non-varying component\n"
        "  -t T, --trials T  random sequences for calibration (default: 1000)\n"
        "  -l L, --length L  length of each random sequence (default: 5000)\n"
        "  -b B, --border B  border added to each random sequence (default: 100)\n"
        "  --batch N         stream sequences in chunks of N\n"
        "  -S F, --scores-file F  write bit scores of the random sequences to F and exit\n"
        "  --warmup          calibrate cache-missing profiles and exit\n"
        "  --barithmetic / --bgeometric (default) / --bmedian  background probabilities\n"
        "  --insert1 F --insert2 F --delete1 F --delete2 F  frameshift rates per base\n"
        "  --stop-codon-prob F  --bg-stop-codon-prob F  --tantan-threshold F\n"
        "  --max             skip the filter pass\n"
        "  -W E, --filter-evalue E  filter-pass E-value threshold (default: value of -e)\n"
        "  --device D        HIP device (default: 0)\n";
    enum { OI1 = 256, OI2, OD1, OD2, OST, OBG, OTT, OMAX, OBATCH, OWARM, ODEV };
    static struct option lopts[] = {{"help", no_argument, 0, 'h'},
                                    {"version", no_argument, 0, 'V'},
                                    {"verbose", no_argument, 0, 'v'},
                                    {"threads", required_argument, 0, 'T'},
                                    {"evalue", required_argument, 0, 'e'},
                                    {"total-length", required_argument, 0, 'N'},
                                    {"strand", required_argument, 0, 's'},
                                    {"mask", required_argument, 0, 'm'},
                                    {"dev", required_argument, 0, 'd'},
                                    {"Dev", required_argument, 0, 'D'},
                                    {"trials", required_argument, 0, 't'},
                                    {"length", required_argument, 0, 'l'},
                                    {"border", required_argument, 0, 'b'},
                                    {"barithmetic", no_argument, 0, 'A'},
                                    {"bgeometric", no_argument, 0, 'G'},
                                    {"bmedian", no_argument, 0, 'M'},
                                    {"insert1", required_argument, 0, OI1},
                                    {"insert2", required_argument, 0, OI2},
                                    {"delete1", required_argument, 0, OD1},
                                    {"delete2", required_argument, 0, OD2},
                                    {"stop-codon-prob", required_argument, 0, OST},
                                    {"bg-stop-codon-prob", required_argument, 0, OBG},
                                    {"tantan-threshold", required_argument, 0, OTT},
                                    {"max", no_argument, 0, OMAX},
                                    {"batch", required_argument, 0, OBATCH},
                                    {"filter-evalue", required_argument, 0, 'W'},
                                    {"scores-file", required_argument, 0, 'S'},
                                    {"warmup", no_argument, 0, OWARM},
                                    {"device", required_argument, 0, ODEV},
                                    {0, 0, 0, 0}};
    auto rate01 = [](const char *s, double &v) {
        v = std::strtod(s, nullptr);
        return v >= 0 && v < 1;
    };
    int c;
    while ((c = getopt_long(argc, argv, "hVve:N:s:m:d:D:t:l:b:T:S:W:", lopts, nullptr)) != -1) {
        switch (c) {
        case 'h': std::cout << help; return 0;
        case 'V': std::cout << "DUMMEST-AMD 1\n"; return 0;
        case 'v': ++o.verbosity; break;
        case 'T':
            o.threads = intFromText(optarg);
            if (o.threads < 1) return badOpt();
            break;
        case 'e':
            o.evalue = std::strtod(optarg, nullptr);
            if (o.evalue < 0) return badOpt();
            break;
        case 'N':
            o.totOverride = std::atoll(optarg);
            if (o.totOverride < 0) return badOpt();
            break;
        case 's': {
            const int s = intFromText(optarg);
            if (s < 0 || s > 2) return badOpt();
            break;
        }
        case 'm':
            o.mask = intFromText(optarg);
            if (o.mask < 0 || o.mask > 3) return badOpt();
            break;
        case 'd':
        case 'D':
            // too low: a coarse discretized Gaussian; too high: slow (dummest.cc:185-195)
            filterStdDev = std::strtod(optarg, nullptr);
            if (filterStdDev < 2 || filterStdDev > 1000) return badOpt();
            keepNonvarying = (c == 'D');
            break;
        case 't':
            o.trials = intFromText(optarg);
            if (o.trials < 1) return badOpt();
            break;
        case 'l':
            o.length = intFromText(optarg);
            if (o.length < 1 || o.length > INT32_MAX - 64) return badOpt();
            break;
        case 'b':
            o.border = intFromText(optarg);
            if (o.border < 0) return badOpt();
            break;
        case 'S': o.scoresFile = optarg; break;
        case OWARM: o.warmup = true; break;
        case 'A': o.bgType = 'A'; break;
        case 'G': o.bgType = 'G'; break;
        case 'M': o.bgType = 'M'; break;
        case OI1: if (!rate01(optarg, o.rates[0])) return badOpt(); break;
        case OI2: if (!rate01(optarg, o.rates[1])) return badOpt(); break;
        case OD1: if (!rate01(optarg, o.rates[2])) return badOpt(); break;
        case OD2: if (!rate01(optarg, o.rates[3])) return badOpt(); break;
        case OST: if (!rate01(optarg, o.rates[4])) return badOpt(); break;
        case OBG: if (!rate01(optarg, o.rates[5])) return badOpt(); break;
        case OTT:
            o.tantan = std::strtod(optarg, nullptr);
            if (!(o.tantan >= 0 && o.tantan <= 1)) return badOpt();
            break;
        case 'W':
            o.filterEvalue = std::strtod(optarg, nullptr);
            if (o.filterEvalue < 0) return badOpt();
            o.filterExplicit = true;
            break;
        case OMAX: o.maxMode = true; break;
        case OBATCH:
            o.batch = intFromText(optarg);
            if (o.batch < 1) return badOpt();
            break;
        case ODEV:
            o.device = intFromText(optarg);
            if (o.device < 0) return badOpt();
            break;
        default: std::cerr << help; return 1;
        }
    }
    const float bgfs1 = (float)o.rates[0] + (float)o.rates[3], bgfs2 = (float)o.rates[1] + (float)o.rates[2];
    if (!(o.rates[0] + o.rates[1] + o.rates[2] + o.rates[3] < 1)) return badOpt();
    if (!(bgfs1 + bgfs2 < 1)) return badOpt();
    if (!o.maxMode && !o.filterExplicit) o.filterEvalue = o.evalue;
    const int nargs = argc - optind;
    if (nargs < 1 || nargs > 2) {
        std::cerr << help;
        return 1;
    }
    if (filterStdDev > 0) o.mask &= 2;  // filtering turns off profile masking (dummest.cc:284-285)
    if (o.border > INT32_MAX - 64 - o.length) return die("sequence + border is too big");

    // ------------------------------------------------------------- profiles
    double phaseStart = wallNow();
    std::vector<HostProfile> profiles;
    {
        std::string msg;
        if (!readHmmFile(argv[optind], o.bgType, o.rates, profiles, msg, (o.mask & 1) ? (float)o.tantan : -1.f,
                         filterStdDev, keepNonvarying))
            return die(msg.empty() ? "can't read the profile data" : msg);
    }
    dm_engine *E = nullptr;
    int rc = dm_engine_create(o.device, &E);
    if (rc) return die("no usable HIP device");
    std::unique_ptr<dm_engine, void (*)(dm_engine *)> guard(E, dm_engine_destroy);
    if ((rc = dm_set_rates(E, o.rates))) return engineFail(E, "dm_set_rates", rc);
    std::vector<int32_t> pids(profiles.size());
    for (size_t i = 0; i < profiles.size(); i++) {
        const HostProfile &p = profiles[i];
        if ((rc = dm_add_profile(E, p.M, p.values.data(), p.params.data(), p.bg.data(), &pids[i])))
            return engineFail(E, "dm_add_profile", rc);
    }

    phaseWall[kTProfiles] += wallNow() - phaseStart;
    std::cout << "# DUMMEST-AMD 1\n";
    std::cout << "# Bytes per floating-point number: 4\n";
    if (filterStdDev > 0)
        std::cout << "# Filtering position-specific letter probabilities: std dev " << filterStdDev << "\n";
    std::cout << "# Background letter probabilities: "
              << (o.bgType == 'A' ? "arithmetic mean" : o.bgType == 'G' ? "geometric mean" : "median")
              << " of foreground probabilities\n";
    std::cout << "# Random sequences: trials=" << o.trials << " length=" << o.length << " border=" << o.border
              << "\n";
    std::cout << "# Frameshift rates: insert1=" << o.rates[0] << " insert2=" << o.rates[1]
              << " delete1=" << o.rates[2] << " delete2=" << o.rates[3] << "\n";
    std::cout << "# Stop codon probs: stop=" << o.rates[4] << " bg-stop=" << o.rates[5]
              << " tantan-threshold=" << o.tantan << "\n";
    if (o.maxMode) std::cout << "# Max mode: filter pass disabled\n";
    else if (nargs > 1) std::cout << "# Filter E-value <= " << o.filterEvalue << "\n";
    if (o.mask & 1) std::cout << "# Masking simple regions in profiles\n";
    if (nargs > 1 && (o.mask & 2)) std::cout << "# Masking simple regions in sequences\n";
    if (nargs > 1 && o.evalue > 0) std::cout << "# E-value <= " << o.evalue << "\n";

    // ----------------------------------------------------------- calibration
    // Profiles are looked up in profile order (not with -S), so a profile with
    // the key of an earlier one calibrated in this run is a hit as in the
    // reference's one-by-one loop; the misses are calibrated in one GPU pass.
    const int np = (int)profiles.size();
    std::vector<std::string> keys(np);
    std::vector<CacheEntry> entry(np);
    std::vector<int> from(np, -1);  // -1: calibrated here, -2: cached, j >= 0: the entry of profile j
    std::vector<int32_t> calPids;
    std::vector<int> calIdx;
    if (np > 0) {
        const char *ign = std::getenv("DUMMEST_CACHE_IGNORE_BINARY_HASH");
        const bool ignoreBin = ign && *ign && std::strcmp(ign, "0") != 0;
        const std::string bin = ignoreBin ? std::string() : gCache.binaryHashHex();
        gCache.load();
        std::unordered_map<std::string, int> firstHere;
        for (int i = 0; i < np; i++) {
            keys[i] = calibrationKey(bin, profiles[i], o.length, o.border, o.trials, o.rates, (float)o.tantan);
            if (!o.scoresFile) {
                if (gCache.find(keys[i], entry[i])) {
                    from[i] = -2;
                    continue;
                }
                const auto it = firstHere.find(keys[i]);
                if (it != firstHere.end()) {
                    from[i] = it->second;
                    continue;
                }
                firstHere.emplace(keys[i], i);
            }
            calIdx.push_back(i);
            calPids.push_back(pids[i]);
        }
    }
    const int nc = (int)calIdx.size();
    std::vector<double> fits((size_t)nc * 16), midS, fltS;
    if (o.scoresFile) {
        midS.resize((size_t)nc * o.trials);
        fltS.resize((size_t)nc * o.trials);
    }
    const int printVerbosity = (nargs < 2) * 2 + (o.evalue <= 0);
    // verbose calibration prints each trial's best anchor (dummest-calibration.hh:511-519)
    std::vector<int32_t> trialAnchor;
    std::vector<double> trialMid;
    phaseStart = wallNow();
    if (nc > 0) {
        if (o.trials < 2) return die("calibration needs at least 2 trials");
        rc = dm_calibrate(E, nc, calPids.data(), o.length, o.border, o.trials, o.scoresFile ? midS.data() : nullptr,
                          o.scoresFile ? fltS.data() : nullptr, fits.data());
        if (rc) return engineFail(E, "dm_calibrate", rc);
        if (printVerbosity > 1) {
            trialAnchor.resize((size_t)nc * o.trials * 2);
            trialMid.resize((size_t)nc * o.trials);
            rc = dm_calibrate_anchors(E, nc, calPids.data(), o.length, o.border, o.trials, trialAnchor.data(),
                                      trialMid.data());
            if (rc) return engineFail(E, "dm_calibrate_anchors", rc);
        }
    }
    phaseWall[kTCalib] += wallNow() - phaseStart;
    std::ofstream scores;
    if (o.scoresFile) {
        scores.open(o.scoresFile);
        if (!scores) return die("can't open scores file");
        scores << "mode\tprofile_name\ttrial\tanchor_type\tln_score\n";
    }
    // The reference's end- and start-anchored samples are all -inf
    // (dummest-dp-forward.hh:536-543); their fits are those of such a sample.
    double endFits[8];
    gumbelFits(std::vector<double>(std::max(2, o.trials), -INFINITY), (double)o.length, endFits);
    for (int k = 0; k < nc; k++) entry[calIdx[k]] = entryFromFits(&fits[(size_t)k * 16], endFits);
    for (int i = 0; i < np; i++)
        if (from[i] >= 0) entry[i] = entry[from[i]];
    // Stats block of one cache entry (printEntryStats, dummest-calibration.hh:605-638).
    auto printEntry = [&](const CacheEntry &c) {
        const double *e = c.v;
        if (printVerbosity > 1) {
            static const char *rows[8] = {"#lamMM", "#kMM", "#kMM1", "#lamML", "#kML", "#kML1", "#lamLM", "#kLM"};
            std::cout << "#\tend-\tstart-\tmid-anchored\n";
            for (int g = 0; g < 8; g++)
                std::cout << rows[g] << "\t" << e[3 * g] << "\t" << e[3 * g + 1] << "\t" << e[3 * g + 2] << "\n";
        } else if (printVerbosity > 0) {
            std::cout << "# K: " << e[6] << " " << e[7] << " " << e[8] << "\n";
        } else {
            std::cout << "# K: " << e[8] << "\n";
        }
        std::cout << "# Lambda: " << e[2] << "\n";
    };
    std::vector<double> K(np), lambda(np), Kf(np), lambdaF(np);
    int warmHits = 0, warmCalibrated = 0, calK = 0;
    for (int i = 0; i < np; i++) {
        const HostProfile &p = profiles[i];
        const CacheEntry &c = entry[i];
        K[i] = c.v[5];
        lambda[i] = c.v[2];
        Kf[i] = c.v[25];
        lambdaF[i] = c.v[24];
        std::cout << "\n# Profile name: " << p.name << "\n# Profile length: " << p.M << "\n";
        if (o.mask & 1) std::cout << "# Positions masked by tantan: " << p.masked << "\n";
        std::cout << "# Background letter probabilities:";
        for (int j = 0; j < 20; j++) std::cout << " " << p.bg[4 + j];
        std::cout << std::endl;
        const bool cached = from[i] != -1;
        if (!o.scoresFile) gCache.announce();  // the lookup reads the cache
        if (cached) {
            if (printVerbosity > 1) std::cout << "# Warning: using cached results\n";
        } else {
            // sum of the amino-acid weights of the trial generator (dummest-calibration.hh:362-377)
            float sum = 0;
            static const int ncod[20] = {4, 2, 2, 2, 2, 4, 2, 3, 2, 6, 1, 2, 4, 2, 6, 6, 4, 4, 1, 2};
            for (int j = 0; j < 20; j++) sum += (float)(p.bg[4 + j] * ncod[j]);
            sum += (float)o.rates[5];
            std::cout << "# sum is " << sum << std::endl;
            if (printVerbosity > 1) {
                // the trial table, in trial order; the end- and start-anchored
                // entries repeat the mid anchor with score -inf (dummest-dp-forward.hh:536-543)
                std::cout << "#trial\tend-anchored\t\tstart-anchored\t\tmid-anchored\n"
                             "#\tprofPos\tseqPos\tscore\tprofPos\tseqPos\tscore\tprofPos\tseqPos\tscore"
                          << std::endl;
                const std::string ninf = scoreStr(-INFINITY);
                for (int t = 0; t < o.trials; t++) {
                    const size_t k = (size_t)calK * o.trials + t;
                    const int a1 = trialAnchor[2 * k], a2 = trialAnchor[2 * k + 1];
                    std::cout << t + 1 << "\t" << a1 << "\t" << a2 << "\t" << ninf << "\t" << a1 << "\t" << a2 << "\t"
                              << ninf << "\t" << a1 << "\t" << a2 << "\t" << scoreStr(natsToLog2(trialMid[k]))
                              << "\n";
                }
                std::cout.flush();
            }
            gCache.announce();  // with -S the store is the first use
            gCache.store(keys[i], c);
        }
        printEntry(c);
        CacheEntry flt = c;
        for (int j = 0; j < 3; j++) {
            flt.v[j] = c.v[24];
            flt.v[3 + j] = c.v[25];
            flt.v[6 + j] = c.v[26];
        }
        printEntry(flt);
        if (o.scoresFile) {
            for (int t = 0; t < o.trials; t++) {
                const size_t k = (size_t)calK * o.trials + t;
                scores << "full\t" << p.name << "\t" << t + 1 << "\tend\t-inf\n";
                scores << "full\t" << p.name << "\t" << t + 1 << "\tstart\t-inf\n";
                scores << "full\t" << p.name << "\t" << t + 1 << "\tmid\t" << scoreStr(natsToLog2(midS[k])) << "\n";
                scores << "filter\t" << p.name << "\t" << t + 1 << "\tall\t" << scoreStr(natsToLog2(fltS[k]))
                       << "\n";
            }
        }
        if (!cached) calK++;
        if (o.warmup) {
            std::cout << "# Warmup: " << p.name << (cached ? " cache hit" : " calibrated") << "\n";
            (cached ? warmHits : warmCalibrated)++;
        }
    }
    if (o.warmup)
        std::cout << "# Warmup: " << np << " profiles, " << warmHits << " cache hits, " << warmCalibrated
                  << " calibrated\n";
    if (o.scoresFile || o.warmup) return 0;
    if (nargs < 2 || np < 1) return 0;
    std::cout << std::endl;

    // -------------------------------------------------------------- sequences
    const std::string seqPath = argv[optind + 1];
    std::ifstream file;
    std::string spool;
    std::istream *in = nullptr;
    const bool needTotal = o.totOverride < 0;
    if (seqPath == "-" && needTotal) {
        spool = std::string(std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp") + "/dummest-amd-stdin-" +
                std::to_string((long)getpid()) + ".fa";
        {
            std::ofstream out(spool, std::ios::binary);
            out << std::cin.rdbuf();
        }
        file.open(spool, std::ios::binary);
        if (!file) return 1;
        in = &file;
    } else if (seqPath == "-") {
        in = &std::cin;
    } else {
        file.open(seqPath);
        if (!file) return die("can't open file: " + seqPath);
        in = &file;
    }
    size_t totLen = needTotal ? 0 : (size_t)o.totOverride;
    std::string err;
    if (needTotal) {
        ContigReader rd(*in);
        SeqState s;
        Contig ct;
        std::vector<char> letters;
        while (rd.next(s, ct, letters, err))
            if (ct.length) totLen += ct.length;
        if (!err.empty()) return die(err);
        file.clear();
        file.seekg(0);
    }
    std::cout << "# Total sequence length: " << totLen << "\n";
    std::cout.precision(3);

    const int batch = o.batch > 0 ? o.batch : o.threads * 25000;
    size_t minProbRatioTotal = needTotal ? 0 : totLen;
    std::vector<Target> targets;
    std::vector<std::pair<int, float>> reqProfRatio;  // per request: profile, minProbRatio
    std::vector<int> reqTarget;
    int64_t nFilterPairs = 0, nFullPairs = 0, nGated = 0, maxWidth = 0;

    auto flush = [&]() -> int {
        if (targets.empty()) return 0;
        double t0 = wallNow();
        // Records with the same letters (a pipeline lists a window once per
        // candidate profile) get the same masks and codes: each distinct
        // sequence is masked, decoded and uploaded once, and the records keep
        // their own headers and coordinates.
        std::vector<int> uniq, seqOf(targets.size());
        {
            std::unordered_map<std::string_view, int> seen;
            seen.reserve(targets.size());
            for (size_t k = 0; k < targets.size(); k++) {
                const auto ins = seen.emplace(std::string_view(targets[k].letters), (int)uniq.size());
                if (ins.second) uniq.push_back((int)k);
                seqOf[k] = ins.first->second;
            }
        }
        // -m 2: tantan on the chunk's contigs on the GPU; letters at or above
        // the threshold (a Float, as TANTAN_MASK_THRESHOLD) decode as non-ACGT
        if (o.mask & 2) {
            std::string all;
            std::vector<int64_t> toff;
            std::vector<int32_t> tlen;
            for (const int u : uniq) {
                toff.push_back((int64_t)all.size());
                tlen.push_back((int32_t)targets[u].letters.size());
                all += targets[u].letters;
            }
            std::vector<float> probs(std::max<size_t>(1, all.size()));
            int r = dm_tantan_dna(E, (int32_t)uniq.size(), all.data(), toff.data(), tlen.data(), probs.data());
            if (r) return engineFail(E, "dm_tantan_dna", r);
            const float thr = (float)o.tantan;
            for (size_t q = 0; q < uniq.size(); q++) {
                Target &t = targets[uniq[q]];
                t.masked.assign(t.letters.size(), 0);
                for (size_t j = 0; j < t.letters.size(); j++) t.masked[j] = !(probs[toff[q] + j] < thr);
            }
            for (size_t k = 0; k < targets.size(); k++)
                if (uniq[seqOf[k]] != (int)k) targets[k].masked = targets[uniq[seqOf[k]]].masked;
        }
        for (const int u : uniq) {
            Target &t = targets[u];
            std::string dec = t.letters;
            for (size_t j = 0; j < t.masked.size(); j++)
                if (t.masked[j]) dec[j] = 'N';
            t.codes.resize(dec.size());
            decodeDna(dec.data(), (int64_t)dec.size(), t.codes.data());
        }
        std::vector<uint8_t> codes;
        std::vector<int64_t> offs;
        std::vector<int32_t> lens;
        for (const int u : uniq) {
            const Target &t = targets[u];
            offs.push_back((int64_t)codes.size());
            lens.push_back((int32_t)t.codes.size());
            codes.insert(codes.end(), t.codes.begin(), t.codes.end());
            maxWidth = std::max<int64_t>(maxWidth, (int64_t)t.codes.size());
        }
        int r = dm_clear_sequences(E);
        int32_t first = 0;
        if (!r) r = dm_add_sequences(E, (int32_t)uniq.size(), codes.data(), offs.data(), lens.data(), &first);
        if (r) return engineFail(E, "dm_add_sequences", r);
        // requests on sequences shorter than 3 cannot be scored (dummest-dp.hh:84)
        std::vector<int32_t> pp, ss;
        std::vector<size_t> rq;
        for (size_t k = 0; k < reqTarget.size(); k++) {
            const int u = seqOf[reqTarget[k]];
            if (targets[uniq[u]].codes.size() < 3) continue;
            pp.push_back(pids[reqProfRatio[k].first]);
            ss.push_back(first + u);
            rq.push_back(k);
        }
        phaseWall[kTPrepare] += wallNow() - t0;
        t0 = wallNow();
        if (!o.maxMode && !pp.empty()) {
            std::vector<float> best(pp.size());
            r = dm_filter_scores(E, (int64_t)pp.size(), pp.data(), ss.data(), best.data());
            if (r) return engineFail(E, "dm_filter_scores", r);
            nFilterPairs += (int64_t)pp.size();
            std::vector<int32_t> p2, s2;
            std::vector<size_t> q2;
            for (size_t k = 0; k < pp.size(); k++) {
                if (!(best[k] > 0)) continue;
                const int pi = reqProfRatio[rq[k]].first;
                const double lp = std::log((double)best[k]);
                const double pr = std::exp(lp * lambdaF[pi]);
                const double ev = Kf[pi] * (double)totLen / pr;
                if (ev <= o.filterEvalue || !std::isfinite(lp)) {
                    p2.push_back(pp[k]);
                    s2.push_back(ss[k]);
                    q2.push_back(rq[k]);
                }
            }
            pp.swap(p2);
            ss.swap(s2);
            rq.swap(q2);
        }
        phaseWall[kTFilter] += wallNow() - t0;
        t0 = wallNow();
        std::ostringstream os;
        if (!pp.empty()) {
            std::vector<float> mm(pp.size()), bm(pp.size());
            std::vector<int32_t> me(pp.size()), be(pp.size());
            r = dm_fwdbwd_scores(E, (int64_t)pp.size(), pp.data(), ss.data(), 0, mm.data(), me.data(), bm.data(),
                                 be.data());
            if (r) return engineFail(E, "dm_fwdbwd_scores", r);
            nFullPairs += (int64_t)pp.size();
            phaseWall[kTScores] += wallNow() - t0;
            // gate: best anchored score >= minProbRatio (dummest-dp-forward.hh:263-271).
            // The score pass rounds differently from the reference (fused
            // multiply-adds, another summation order), so here a pair is
            // dropped only when it is clearly below the threshold; the
            // post-gate step (dm_pairs_hits) works on matrices that equal the
            // reference's and applies the exact gate itself.
            std::vector<int32_t> gp, gs;
            std::vector<double> gr;
            std::vector<size_t> gk;
            for (size_t k = 0; k < pp.size(); k++) {
                const float ratio = reqProfRatio[rq[k]].second;
                if (ratio < 0) continue;  // calibration-style lanes report no alignments
                if (ratio > 0) {
                    if (mm[k] == 0.f) continue;
                    const double lnGot = std::log((double)mm[k]) + (double)me[k] * 0.6931471805599453;
                    if (lnGot < std::log((double)ratio) - 1e-5) continue;
                }
                gp.push_back(pp[k]);
                gs.push_back(ss[k]);
                gr.push_back((double)ratio);
                gk.push_back(k);
            }
            nGated += (int64_t)gk.size();
            // the hits of every gated pair in one batched call
            std::vector<double> hits(6 * std::max<size_t>(4096, 4 * gk.size()));
            std::vector<int32_t> segs(3 * std::max<size_t>(1 << 20, 64 * gk.size()));
            std::vector<int64_t> first(gk.size() + 1, 0);
            int64_t nh = 0, ns = 0;
            const double hitsStart = wallNow();
            r = dm_pairs_hits(E, (int64_t)gk.size(), gp.data(), gs.data(), gr.data(), hits.data(),
                              (int64_t)(hits.size() / 6), segs.data(), (int64_t)(segs.size() / 3), first.data(), &nh,
                              &ns);
            if (r == DM_ERANGE) {  // more hits than the buffers hold: grow them to the reported need
                hits.resize(6 * (size_t)nh);
                segs.resize(3 * (size_t)ns);
                r = dm_pairs_hits(E, (int64_t)gk.size(), gp.data(), gs.data(), gr.data(), hits.data(), nh,
                                  segs.data(), ns, first.data(), &nh, &ns);
            }
            if (r) return engineFail(E, "dm_pairs_hits", r);
            hitsWall += wallNow() - hitsStart;
            hitsDevice += dm_last_hits_ms(E) * 1e-3;
            t0 = wallNow();
            int64_t so = 0;
            for (size_t gi = 0; gi < gk.size(); gi++) {
                const size_t k = gk[gi];
                const int pi = reqProfRatio[rq[k]].first;
                const Target &T = targets[reqTarget[rq[k]]];
                const HostProfile &P = profiles[pi];
                const SeqState &S = T.rec;
                for (int64_t h = first[gi]; h < first[gi + 1]; h++) {
                    const double lnr = hits[6 * h];
                    const int a1 = (int)hits[6 * h + 3], a2 = (int)hits[6 * h + 4], nseg = (int)hits[6 * h + 5];
                    const int32_t *sg = &segs[3 * so];
                    so += nseg;
                    if (nseg == 0) continue;  // empty alignments are not records (dummest-search.hh:36-41)
                    // aligned rows (addAlignedProfile / addAlignedSequence, dummest-output.hh:10-46)
                    std::string row1, row2;
                    int pos1 = sg[0], pos2 = sg[1];
                    for (int g = 0; g < nseg; g++) {
                        const int s1 = sg[3 * g], s2 = sg[3 * g + 1], ln = sg[3 * g + 2];
                        for (; pos1 < s1; ++pos1) row1.push_back(profLetter(P.consensus[pos1]));
                        row1.append(s2 - pos2, '-');
                        for (; pos1 < s1 + ln; ++pos1) row1.push_back(profLetter(P.consensus[pos1]));
                        pos2 = s2 + ln;
                    }
                    pos1 = sg[0];
                    pos2 = sg[1];
                    for (int g = 0; g < nseg; g++) {
                        const int s1 = sg[3 * g], s2 = sg[3 * g + 1], ln = sg[3 * g + 2];
                        row2.append(s1 - pos1, '-');
                        for (; pos2 < s2 + ln; ++pos2) row2.push_back(seqLetter(T, pos2));
                        pos1 = s1 + ln;
                    }
                    const double ev = K[pi] * (double)totLen * std::exp(-lambda[pi] * lnr);
                    if (o.evalue > 0 && ev > o.evalue) continue;
                    if (std::isnan(ev)) continue;
                    // printSimilarity, dummest-output.hh:68-107
                    const int len = (int)row1.size();
                    const int span1 = len - (int)std::count(row1.begin(), row1.end(), '-');
                    const int span2 = len - (int)std::count(row2.begin(), row2.end(), '-');
                    int start2 = T.contig.start + sg[1];
                    if (S.hasPipelineFields)
                        start2 = S.isPlus ? (int)(S.wStart - 1 + start2) : (int)(S.trueLength - S.wEnd + start2);
                    const int reportLen = S.hasPipelineFields ? (int)S.trueLength : S.length;
                    const int anchor2 = T.contig.start + a2;
                    const int start1 = sg[0];
                    const int w1 = (int)std::max(P.name.size(), S.name.size());
                    const int w2 = std::max(numOfDigits(start1), numOfDigits(start2));
                    const int w3 = std::max(numOfDigits(span1), numOfDigits(span2));
                    const int w4 = std::max(numOfDigits(P.M), numOfDigits(reportLen));
                    os << "a score=" << scoreStr(natsToLog2(lnr)) << " E=" << ev << " anchor=" << a1 << ","
                       << anchor2 << "\n";
                    os << "s " << std::left << std::setw(w1) << P.name << " " << std::right << std::setw(w2)
                       << start1 << " " << std::setw(w3) << span1 << " +"
                       << " " << std::setw(w4) << P.M << " " << row1 << "\n";
                    os << "s " << std::left << std::setw(w1) << S.name << " " << std::right << std::setw(w2)
                       << start2 << " " << std::setw(w3) << span2 << " " << (S.isPlus ? '+' : '-') << " "
                       << std::setw(w4) << reportLen << " " << row2 << "\n\n";
                }
            }
        }
        const std::string out = os.str();
        std::fwrite(out.data(), 1, out.size(), stdout);
        if (!pp.empty()) phaseWall[kTOutput] += wallNow() - t0;
        targets.clear();
        reqProfRatio.clear();
        reqTarget.clear();
        return 0;
    };

    // profiles by name, in index order (a pipeline record's profile= picks them)
    std::unordered_map<std::string, std::vector<int>> byName;
    for (int j = 0; j < np; j++) byName[profiles[j].name].push_back(j);
    std::vector<int> allProfiles(np);
    std::iota(allProfiles.begin(), allProfiles.end(), 0);
    const std::vector<int> noProfiles;
    {
        ContigReader rd(*in);
        SeqState s;
        Contig ct;
        std::vector<char> letters;
        double inStart = wallNow();
        while (rd.next(s, ct, letters, err)) {
            if (ct.length == 0) continue;
            if (needTotal) minProbRatioTotal += ct.length;
            Target T;
            T.rec = s;
            T.contig = ct;
            T.letters.resize(letters.size());
            T.nums.resize(letters.size());
            for (size_t j = 0; j < letters.size(); j++) {
                T.letters[j] = kAlphabet[(int)letters[j]];
                T.nums[j] = (uint8_t)letters[j];
            }
            const int ti = (int)targets.size();
            const std::vector<int> *req = &allProfiles;
            if (!s.targetProfile.empty()) {
                const auto it = byName.find(s.targetProfile);
                req = it == byName.end() ? &noProfiles : &it->second;
            }
            for (const int j : *req) {
                const float ratio =
                    o.evalue > 0 ? (float)std::pow(K[j] * (double)minProbRatioTotal / o.evalue, 1.0 / lambda[j]) : -1.f;
                reqProfRatio.push_back({j, ratio});
                reqTarget.push_back(ti);
            }
            targets.push_back(std::move(T));
            if ((int)targets.size() >= batch) {
                phaseWall[kTInput] += wallNow() - inStart;
                if ((rc = flush())) return rc;
                inStart = wallNow();
            }
        }
        phaseWall[kTInput] += wallNow() - inStart;
        if (!err.empty()) return die(err);
        if ((rc = flush())) return rc;
    }
    std::cout << "# Batches: fast_batches=" << nFullPairs << " calib_batches=" << (int64_t)np * o.trials << "\n";
    std::cout << "# Filter stats: filter_batches=" << nFilterPairs << " discarded_empty=" << (nFilterPairs - nFullPairs)
              << " max_dp_width=" << maxWidth << "\n";
    if (!spool.empty()) std::remove(spool.c_str());
    if (std::getenv("DM_TIMING")) {
        const double total = wallNow() - runStart;
        std::fprintf(stderr, "timing: run %.3f s, post-gate hits %.3f s (%.1f%%; %.3f s on the device) for %lld gated pairs\n",
                     total, hitsWall, 100.0 * hitsWall / total, hitsDevice, (long long)nGated);
        std::fprintf(stderr, "timing:");
        for (int k = 0; k < kTPhases; k++) std::fprintf(stderr, " %s %.3f s", phaseName[k], phaseWall[k]);
        std::fprintf(stderr, "\n");
    }
    return 0;
}

// dummest-amd-build: the profile builder executable, a drop-in for the
// reference's dummest-build (dummest-build.cc:916-1167): the same options,
// Stockholm input ("-" = stdin), profiles on stdout, progress on stderr.
#include <getopt.h>

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "dm_host.h"

using namespace dm;

namespace {

const char kHelp[] =
    "usage: dummest-amd-build alignments.stk\n"
    "\n"
    "Read multiple sequence alignments: write position-specific (-log) probabilities\n"
    "of letters, insertions, and deletions.\n"
    "\n"
    "Options:\n"
    "  -h, --help     show this help message and exit\n"
    "  -V, --version  show version and exit\n"
    "  -v, --verbose  show progress messages\n"
    "  --counts       output weighted counts instead of -log probs (implies --enone)\n"
    "\n"
    "Options for defining non-insert positions:\n"
    "  --hand         use alignment RF column annotation lines (ignore symfrac)\n"
    "  --symfrac F    minimum (weighted) non-gap fraction to define a non-insert\n"
    "                     position (default: 0.50)\n"
    "\n"
    "Options for effective sequence number:\n"
    "  --enone        ignore relative entropy: set maximum sequence weight to 1\n"
    "  --ere E        aim for this relative entropy per position (default:\n"
    "                     0.59 for protein, 0.62 for nucleotide)\n"
    "  --esigma E     aim for this relative entropy (default: 45.0)\n"
    "\n"
    "Baum-Welch options:\n"
    "  --maxiter N    maximum number of Baum-Welch iterations (default: 1000)\n"
    "  --maxdiff X    convergence threshold for Baum-Welch (default: 1e-6)\n"
    "  --countonly    only count events, do not run Baum-Welch\n"
    "\n"
    "Prior probability options:\n"
    "  --dmix         Dirichlet mixture file (esl-mixdchlet format)\n"
    "  --gapprior     file with 7 gap pseudocounts:\n"
    "                     match insStart delStart insEnd insExtend delEnd delExtend\n"
    "  --pnone        don't use any prior probabilities (implies --enone)\n";

int badOpt() {
    std::cerr << "bad option value\n";
    return 1;
}

std::istream &openInput(std::ifstream &file, const char *name) {
    if (std::string(name) == "-") return std::cin;
    file.open(name);
    if (!file) std::cerr << "can't open file: " << name << "\n";
    return file;
}

}  // namespace

int main(int argc, char *argv[]) {
    BuildOptions o;
    const char *mixFile = nullptr, *gapFile = nullptr;
    static struct option longOpts[] = {{"help", no_argument, 0, 'h'},          {"version", no_argument, 0, 'V'},
                                       {"verbose", no_argument, 0, 'v'},       {"symfrac", required_argument, 0, 's'},
                                       {"counts", no_argument, 0, 'C'},        {"hand", no_argument, 0, 'H'},
                                       {"enone", no_argument, 0, 'n'},         {"ere", required_argument, 0, 'p'},
                                       {"esigma", required_argument, 0, 'e'},  {"maxiter", required_argument, 0, 'N'},
                                       {"maxdiff", required_argument, 0, 'X'}, {"countonly", no_argument, 0, 'c'},
                                       {"dmix", required_argument, 0, 'D'},    {"gapprior", required_argument, 0, 'G'},
                                       {"pnone", no_argument, 0, 'P'},         {0, 0, 0, 0}};
    int c, idx;
    while ((c = getopt_long(argc, argv, "hVv", longOpts, &idx)) != -1) {
        switch (c) {
        case 'h': std::cout << kHelp; return 0;
        case 'V': std::cout << o.versionText << "\n"; return 0;
        case 'v': ++o.verbosity; break;
        case 's':
            o.symfrac = std::strtod(optarg, nullptr);
            if (o.symfrac < 0 || o.symfrac > 1) return badOpt();
            break;
        case 'C': o.counts = true; break;
        case 'H': o.hand = true; break;
        case 'n': o.esigma = 1e37; break;
        case 'p':
            o.ere = std::strtod(optarg, nullptr);
            if (o.ere <= 0) return badOpt();
            break;
        case 'e':
            o.esigma = std::strtod(optarg, nullptr);
            if (o.esigma <= 0) return badOpt();
            break;
        case 'N':
            o.maxIter = (double)std::strtol(optarg, nullptr, 0);
            if (o.maxIter <= 0) return badOpt();
            break;
        case 'X':
            o.maxDiff = std::strtod(optarg, nullptr);
            if (o.maxDiff < 0) return badOpt();
            break;
        case 'c': o.countOnly = true; break;
        case 'D': mixFile = optarg; break;
        case 'G': gapFile = optarg; break;
        case 'P':
            o.pnone = true;
            o.esigma = 1e37;
            break;
        default: std::cerr << kHelp; return 1;
        }
    }
    if (argc - optind != 1) {
        std::cerr << kHelp;
        return 1;
    }
    // the BM line lists the command without its last argument
    for (int k = 0; k + 1 < argc; k++) o.bmText += std::string(" ") + argv[k];
    if (mixFile) {
        std::ifstream f;
        std::istream &in = openInput(f, mixFile);
        if (!f || !readMixture(in, o.mixture, o.mixtureComponents)) {
            if (f) std::cerr << "can't read the Dirichlet mixture file\n";
            return 1;
        }
    }
    if (gapFile) {
        std::ifstream f;
        std::istream &in = openInput(f, gapFile);
        if (!f || !readGapPrior(in, o.gapPrior)) {
            if (f) std::cerr << "can't read the gap pseudocounts\n";
            return 1;
        }
        o.hasGapPrior = true;
    }
    std::ifstream f;
    std::istream &in = openInput(f, argv[optind]);
    if (!f) return 1;
    try {
        return buildProfiles(in, o, std::cout, std::cerr);
    } catch (const BuildError &e) {
        if (e.fatal) {
            std::cerr << "FATAL: " << e.what() << "\n";
            std::abort();
        }
        std::cerr << e.what() << "\n";
        return 1;
    }
}

// tables.cpp - host-side energy tables of the sampler.
//
// OrigamiPotential::calc_energies and its helpers
// (origami_potential.cpp:1057-1221) with the nearest-neighbour
// thermodynamics of nearest_neighbour.cpp:34-177 (SantaLucia 2004 unified
// parameters, Kilchherr 2016 stacking).  Tables are dense over the domain
// identity range and are built once per temperature-ladder slot; the device
// only reads them.  The arithmetic follows the reference operation by
// operation so the FP64 tables are bit-identical to the reference's.
#include "tables.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace ldo {

namespace {

constexpr double kR = 8.3144598;   // J/K/mol
constexpr double kJPerCal = 4.184;

int nuc(char b) {
    switch (b) {
    case 'A':
        return 0;
    case 'C':
        return 1;
    case 'G':
        return 2;
    case 'T':
        return 3;
    default:
        throw std::invalid_argument("sequence holds a character that is not A, C, G or T");
    }
}

// nearest-neighbour pair parameters by 5'->3' dinucleotide (row: first
// base, column: second base, order A C G T)
struct PairTable {
    double dh[4][4];
    double ds[4][4];
    double stack[4][4];
};

const PairTable kPairs = {
        {{-7.6, -8.4, -7.8, -7.2},
         {-8.5, -8.0, -10.6, -7.8},
         {-8.2, -9.8, -8.0, -8.4},
         {-7.2, -8.2, -8.5, -7.6}},
        {{-0.0213, -0.0224, -0.0210, -0.0204},
         {-0.0227, -0.0199, -0.0272, -0.0210},
         {-0.0222, -0.0244, -0.0199, -0.0224},
         {-0.0213, -0.0222, -0.0227, -0.0213}},
        {{-1.36, -2.03, -1.60, -2.35},
         {-0.81, -1.64, -2.06, -1.39},
         {-1.39, -3.42, -1.64, -0.81},
         {-1.01, -1.60, -2.03, -1.36}}};

const double kInitDH = 0.2, kInitDS = -0.0057;
const double kTermDH = 2.2, kTermDS = 0.0069;
const double kSymDS = -0.0014;

char comp_base(char b) { return "TGCA"[nuc(b)]; }

std::string revcomp(const std::string& s) {
    std::string r(s.rbegin(), s.rend());
    for (auto& c: r) {
        c = comp_base(c);
    }
    return r;
}

// calc_hybridization_H_and_S (nearest_neighbour.cpp:66-115), kcal/mol
void duplex(const std::string& seq, double cation, double& h, double& s) {
    double hst = 0, sst = 0;
    for (size_t i = 0; i + 1 < seq.size(); i++) {
        int a = nuc(seq[i]), b = nuc(seq[i + 1]);
        hst += kPairs.dh[a][b];
        sst += kPairs.ds[a][b];
    }
    int term = (seq.front() == 'A' || seq.front() == 'T') + (seq.back() == 'A' || seq.back() == 'T');
    double sym = (revcomp(seq) == seq) ? kSymDS : 0.0;
    h = kInitDH + hst + kTermDH * term;
    s = kInitDS + sym + sst + kTermDS * term;
    s += 0.368 * seq.size() * std::log(cation) / 1000;
}

// find_longest_contig_complement (nearest_neighbour.cpp:117-160)
std::vector<std::string> longest_complements(const std::string& a, const std::string& b) {
    bool a_short = a.size() <= b.size();
    const std::string& fivep = a_short ? a : b;
    std::string threep = revcomp(a_short ? b : a);
    std::vector<std::string> out;
    for (size_t n = threep.size(); n >= 1; n--) {
        for (size_t i = 0; i + n <= threep.size(); i++) {
            std::string sub = threep.substr(i, n);
            for (size_t at = fivep.find(sub); at != std::string::npos; at = fivep.find(sub, at + 1)) {
                out.push_back(sub);
            }
        }
        if (!out.empty()) {
            break;
        }
    }
    return out;
}

} // namespace

TableSet build_tables(const SystemSpec& sys, double temp, double stacking_mult) {
    TableSet t;
    size_t n = (size_t)sys.n_ident;
    t.hyb_e.assign(n * n, 0.0);
    t.hyb_h.assign(n * n, 0.0);
    t.hyb_s.assign(n * n, 0.0);
    t.stack.assign(n * n, 0.0);
    // calc_unitless_init_thermo (nearest_neighbour.cpp:57-64)
    t.init_h = kInitDH * kJPerCal * 1000 / kR / temp;
    t.init_s = kInitDS * kJPerCal * 1000 / kR;
    t.init_e = t.init_h - t.init_s;
    double stacking_ene = sys.stacking_ene * stacking_mult;
    const double ln6 = std::log(6);
    for (size_t ci = 0; ci < sys.idents.size(); ci++) {
        for (size_t cj = 0; cj < sys.idents.size(); cj++) {
            for (size_t di = 0; di < sys.idents[ci].size(); di++) {
                for (size_t dj = 0; dj < sys.idents[cj].size(); dj++) {
                    int ii = sys.idents[ci][di], ij = sys.idents[cj][dj];
                    size_t k = (size_t)(ii - sys.ident_lo) * n + (size_t)(ij - sys.ident_lo);
                    bool comp = ii == -ij;
                    double h = 0, s = 0;
                    if (sys.hybridization_pot == 0) {
                        const std::string& a = sys.seqs[ci][di];
                        const std::string& b = sys.seqs[cj][dj];
                        std::vector<std::string> cs = longest_complements(a, b);
                        if (!cs.empty()) {
                            for (const std::string& c: cs) {
                                double ch, cs_;
                                duplex(c, sys.cation_M, ch, cs_);
                                h += ch * kJPerCal * 1000 / kR / temp - t.init_h;
                                s += cs_ * kJPerCal * 1000 / kR - t.init_s;
                            }
                            h /= (int)cs.size();
                            s /= (int)cs.size();
                            s += ln6;
                            bool full = cs[0].size() == a.size() && a.size() == b.size();
                            if (comp && cs[0].size() != a.size() && a.size() == b.size()) {
                                throw std::runtime_error("sequences that should be complementary are not");
                            }
                            if (!comp && full) {
                                throw std::runtime_error("sequences that should not be complementary are");
                            }
                        }
                    }
                    else {
                        if (comp) {
                            if (sys.hybridization_pot == 1) {
                                h = sys.binding_h / temp;
                                s = sys.binding_s;
                            }
                            else {
                                int idx = std::abs(ii) - 1;
                                h = sys.spec_h.at(idx) / temp;
                                s = sys.spec_s.at(idx);
                            }
                        }
                        else {
                            h = sys.misbinding_h / temp;
                            s = sys.misbinding_s;
                        }
                        s += ln6;
                    }
                    t.hyb_h[k] = h;
                    t.hyb_s[k] = s;
                    t.hyb_e[k] = h - s;
                    if (sys.stacking_pot == 1) {
                        const std::string& a = sys.seqs[ci][di];
                        const std::string& b = sys.seqs[cj][dj];
                        t.stack[k] = kPairs.stack[nuc(a.back())][nuc(b.front())];
                    }
                    else {
                        t.stack[k] = stacking_ene / temp;
                    }
                }
            }
        }
    }
    return t;
}

} // namespace ldo

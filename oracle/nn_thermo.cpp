// nn_thermo.cpp - oracle restatement of nearest_neighbour.cpp.
//
// TEST INFRASTRUCTURE ONLY (see ldo_oracle.h).
//
// The SantaLucia (2004) nearest-neighbour parameters and the Kilchherr (2016)
// stacking free energies are physical data; they are laid out here as
// dinucleotide-indexed arrays (index = 4*base(5') + base(3'), bases ordered
// A,C,G,T) instead of the reference's string-keyed maps
// (nearest_neighbour.hpp:32-98).
#include "nn_thermo.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace orc_nn {

namespace {

int base_index(char b) {
    switch (b) {
    case 'A':
        return 0;
    case 'C':
        return 1;
    case 'G':
        return 2;
    case 'T':
        return 3;
    }
    throw std::invalid_argument(std::string("not a DNA base: ") + b);
}

char base_complement(char b) {
    switch (b) {
    case 'A':
        return 'T';
    case 'T':
        return 'A';
    case 'G':
        return 'C';
    case 'C':
        return 'G';
    }
    throw std::invalid_argument(std::string("not a DNA base: ") + b);
}

// Duplex enthalpy (kcal/mol) of a 5'XY3' top strand dinucleotide.
// Rows X = A,C,G,T; columns Y = A,C,G,T.
const double kDinucH[16] = {
        -7.6, -8.4, -7.8, -7.2,  // AA AC AG AT
        -8.5, -8.0, -10.6, -7.8, // CA CC CG CT
        -8.2, -9.8, -8.0, -8.4,  // GA GC GG GT
        -7.2, -8.2, -8.5, -7.6   // TA TC TG TT
};
// Duplex entropy (kcal/mol/K)
const double kDinucS[16] = {
        -0.0213, -0.0224, -0.0210, -0.0204, // AA AC AG AT
        -0.0227, -0.0199, -0.0272, -0.0210, // CA CC CG CT
        -0.0222, -0.0244, -0.0199, -0.0224, // GA GC GG GT
        -0.0213, -0.0222, -0.0227, -0.0213  // TA TC TG TT
};
// Coaxial stacking free energy (kcal/mol) of 5'XY3' across a nick
const double kDinucStack[16] = {
        -1.36, -2.03, -1.60, -2.35, // AA AC AG AT
        -0.81, -1.64, -2.06, -1.39, // CA CC CG CT
        -1.39, -3.42, -1.64, -0.81, // GA GC GG GT
        -1.01, -1.60, -2.03, -1.36  // TA TC TG TT
};
constexpr double kInitH = 0.2;
constexpr double kInitS = -0.0057;
constexpr double kTerminalATH = 2.2;
constexpr double kTerminalATS = 0.0069;
constexpr double kSymmetryS = -0.0014;

bool palindromic(const std::string& seq) {
    std::string rc = complement(seq);
    std::reverse(rc.begin(), rc.end());
    return rc == seq;
}

} // namespace

std::string complement(const std::string& seq) {
    std::string out(seq.size(), ' ');
    for (size_t i = 0; i < seq.size(); i++) {
        out[i] = base_complement(seq[i]);
    }
    return out;
}

HS duplex_h_s(const std::string& seq, double cation_M) {
    double h_stack = 0;
    double s_stack = 0;
    for (size_t i = 0; i + 1 < seq.size(); i++) {
        int k = 4 * base_index(seq[i]) + base_index(seq[i + 1]);
        h_stack += kDinucH[k];
        s_stack += kDinucS[k];
    }
    int terminal_at = 0;
    if (seq.front() == 'A' || seq.front() == 'T') {
        terminal_at += 1;
    }
    if (seq.back() == 'A' || seq.back() == 'T') {
        terminal_at += 1;
    }
    double s_sym = palindromic(seq) ? kSymmetryS : 0.0;
    double h = kInitH + h_stack + kTerminalATH * terminal_at;
    double s = kInitS + s_sym + s_stack + kTerminalATS * terminal_at;
    s += 0.368 * seq.size() * std::log(cation_M) / 1000;
    return {h, s};
}

HS reduced_duplex(const std::string& seq, double temp, double cation_M) {
    HS hs = duplex_h_s(seq, cation_M);
    return {hs.h * kJoulePerCal * 1000 / kGasR / temp,
            hs.s * kJoulePerCal * 1000 / kGasR};
}

HS reduced_initiation(double temp) {
    return {kInitH * kJoulePerCal * 1000 / kGasR / temp,
            kInitS * kJoulePerCal * 1000 / kGasR};
}

double reduced_duplex_energy(const std::string& seq, double temp, double cation_M) {
    HS hs = reduced_duplex(seq, temp, cation_M);
    return hs.h - hs.s;
}

std::vector<std::string> longest_contig_complements(const std::string& a,
                                                    const std::string& b) {
    // The shorter strand (the first one on ties) is searched for reverse
    // complements of every substring of the other, longest first.
    const std::string& shorter = (a.size() <= b.size()) ? a : b;
    const std::string& longer = (a.size() <= b.size()) ? b : a;
    std::string target(longer.rbegin(), longer.rend());
    target = complement(target);
    std::vector<std::string> found;
    for (size_t len = target.size(); len > 0; len--) {
        for (size_t start = 0; start + len <= target.size(); start++) {
            std::string sub = target.substr(start, len);
            size_t at = shorter.find(sub);
            while (at != std::string::npos) {
                found.push_back(sub);
                at = shorter.find(sub, at + 1);
            }
        }
        if (!found.empty()) {
            return found;
        }
    }
    return found;
}

double seq_stacking_energy(const std::string& a, const std::string& b) {
    return kDinucStack[4 * base_index(a.back()) + base_index(b.front())];
}

} // namespace orc_nn

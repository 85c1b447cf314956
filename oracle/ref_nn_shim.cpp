// ref_nn_shim.cpp - C entry points around the reference's own
// nearest-neighbour module (src/nearest_neighbour.cpp, compiled from the
// reference tree by oracle/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE
// ONLY: used to cross-check the oracle's restatement in nn_thermo.cpp.
#include <cstring>
#include <string>
#include <vector>

#include "LatticeDNAOrigami/nearest_neighbour.hpp"

extern "C" {

int ref_nn_hybridization_thermo(const char* seq, double temp, double cation_M,
                                double* out_h, double* out_s) {
    nearestNeighbour::ThermoOfHybrid t =
            nearestNeighbour::calc_unitless_hybridization_thermo(seq, temp, cation_M);
    *out_h = t.enthalpy;
    *out_s = t.entropy;
    return 0;
}

int ref_nn_init_thermo(double temp, double* out_h, double* out_s) {
    nearestNeighbour::ThermoOfHybrid t = nearestNeighbour::calc_unitless_init_thermo(temp);
    *out_h = t.enthalpy;
    *out_s = t.entropy;
    return 0;
}

int ref_nn_longest_contig_complement(const char* a, const char* b, char* out, int cap) {
    std::vector<std::string> v = nearestNeighbour::find_longest_contig_complement(a, b);
    std::string joined;
    for (size_t i = 0; i < v.size(); i++) {
        if (i) {
            joined += ' ';
        }
        joined += v[i];
    }
    if ((int)joined.size() + 1 > cap) {
        return 1;
    }
    std::memcpy(out, joined.c_str(), joined.size() + 1);
    return 0;
}

double ref_nn_stacking_energy(const char* a, const char* b) {
    return nearestNeighbour::calc_seq_spec_stacking_energy(a, b, 0, 0);
}
}

// nn_thermo.hpp - oracle restatement of the reference's nearest-neighbour
// hybridization thermodynamics (src/nearest_neighbour.cpp:19-177,
// include/LatticeDNAOrigami/nearest_neighbour.hpp:22-98).
//
// TEST INFRASTRUCTURE ONLY (see ldo_oracle.h).
#ifndef ORACLE_NN_THERMO_HPP
#define ORACLE_NN_THERMO_HPP

#include <string>
#include <vector>

namespace orc_nn {

// R (J/K/mol) and J per cal used by the reference (nearest_neighbour.hpp:21-25)
constexpr double kGasR = 8.3144598;
constexpr double kJoulePerCal = 4.184;

struct HS {
    double h;
    double s;
};

// kcal/mol and kcal/mol/K, before conversion (nearest_neighbour.cpp:66-115)
HS duplex_h_s(const std::string& seq, double cation_M);
// H/(RT) and S/R of a duplex (nearest_neighbour.cpp:34-45)
HS reduced_duplex(const std::string& seq, double temp, double cation_M);
// initiation term only (nearest_neighbour.cpp:57-64)
HS reduced_initiation(double temp);
double reduced_duplex_energy(const std::string& seq, double temp, double cation_M);
// all longest contiguous complementary stretches (nearest_neighbour.cpp:117-160)
std::vector<std::string> longest_contig_complements(const std::string& a,
                                                    const std::string& b);
std::string complement(const std::string& seq);
// sequence-specific stacking between the 3' end of a and the 5' end of b
// (nearest_neighbour.cpp:19-32)
double seq_stacking_energy(const std::string& a, const std::string& b);

} // namespace orc_nn

#endif

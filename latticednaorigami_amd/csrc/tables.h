// tables.h - host-side energy tables of the sampler.
#ifndef LDO_TABLES_H
#define LDO_TABLES_H

#include <string>
#include <vector>

namespace ldo {

struct SystemSpec {
    std::vector<std::vector<int>> idents;
    std::vector<std::vector<std::string>> seqs;
    std::vector<double> spec_h, spec_s;
    int ident_lo = 0;
    int n_ident = 0;
    int hybridization_pot = 0; // 0 NN, 1 Uniform, 2 Specified
    int stacking_pot = 0;      // 0 Constant, 1 SequenceSpecific
    double binding_h = 0, binding_s = 0, misbinding_h = 0, misbinding_s = 0;
    double stacking_ene = 0, cation_M = 1;
};

struct TableSet {
    std::vector<double> hyb_e, hyb_h, hyb_s, stack;
    double init_h = 0, init_s = 0, init_e = 0;
};

TableSet build_tables(const SystemSpec& sys, double temp, double stacking_mult);

} // namespace ldo

#endif

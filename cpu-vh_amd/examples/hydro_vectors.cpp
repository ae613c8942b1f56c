// Runs HYDRO (include/cpuvh_hydro.hpp) the way a framework that owns the
// initial state does: the vectors are handed over in memory, then
// run_hydro_no_cli() reads rhic-conf/ and writes cpu-vh_output/ under the
// working directory (initialConditionType 13 or 14 in rhic-conf/ic.properties).
//
//   hydro_vectors <file> [--reference-ut]
//
// <file>: one line per cell (x fastest), either one column, the energy
// density (initialize_from_vector), or sixteen, e ut ux uy un pitt pitx pity
// pitn pixx pixy pixn piyy piyn pinn Pi (initialize_from_vectors); energy
// density and pi in GeV/fm^3.  --reference-ut: type 14 stores the
// reference's u^tau = 1 + ux^2 + uy^2 + un^2 (HYDRO::reference_ut).
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cpuvh_hydro.hpp"

int main(int argc, char **argv) {
    const bool reference_ut = argc == 3 && std::string(argv[2]) == "--reference-ut";
    if (argc != 2 && !reference_ut) {
        std::fprintf(stderr, "usage: %s <vectors file> [--reference-ut]\n", argv[0]);
        return 64;
    }
    std::ifstream in(argv[1]);
    if (!in) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 66;
    }
    std::vector<std::vector<double>> col;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::vector<double> row;
        double v;
        while (ls >> v) row.push_back(v);
        if (row.empty()) continue;
        if (col.empty()) col.resize(row.size());
        if (row.size() != col.size() || (row.size() != 1 && row.size() != 16)) {
            std::fprintf(stderr, "%s: every line holds 1 or 16 numbers\n", argv[1]);
            return 65;
        }
        for (size_t c = 0; c < row.size(); ++c) col[c].push_back(row[c]);
    }
    HYDRO hydro;
    hydro.reference_ut = reference_ut;
    if (col.size() == 1)
        hydro.initialize_from_vector(col[0]);
    else
        hydro.initialize_from_vectors(col[0], col[1], col[2], col[3], col[4], col[5], col[6], col[7], col[8], col[9],
                                      col[10], col[11], col[12], col[13], col[14], col[15]);
    return hydro.run_hydro_no_cli();
}

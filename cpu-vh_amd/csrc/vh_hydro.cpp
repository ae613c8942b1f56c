// HYDRO (include/cpuvh_hydro.hpp): the reference's wrapper class
// (rhic/src/HydroWrapper.cpp:20-136) over the driver.
#include "../../include/cpuvh_hydro.hpp"

#include <utility>

void HYDRO::initialize_from_vector(std::vector<double> e) { initial_energy_density = std::move(e); }

void HYDRO::initialize_from_vectors(std::vector<double> e, std::vector<double> ut, std::vector<double> ux,
                                    std::vector<double> uy, std::vector<double> un, std::vector<double> pitt,
                                    std::vector<double> pitx, std::vector<double> pity, std::vector<double> pitn,
                                    std::vector<double> pixx, std::vector<double> pixy, std::vector<double> pixn,
                                    std::vector<double> piyy, std::vector<double> piyn, std::vector<double> pinn,
                                    std::vector<double> Pi) {
    initial_energy_density = std::move(e);
    initial_ut = std::move(ut);
    initial_ux = std::move(ux);
    initial_uy = std::move(uy);
    initial_un = std::move(un);
    initial_pitt = std::move(pitt);
    initial_pitx = std::move(pitx);
    initial_pity = std::move(pity);
    initial_pitn = std::move(pitn);
    initial_pixx = std::move(pixx);
    initial_pixy = std::move(pixy);
    initial_pixn = std::move(pixn);
    initial_piyy = std::move(piyy);
    initial_piyn = std::move(piyn);
    initial_pinn = std::move(pinn);
    initial_Pi = std::move(Pi);
}

cpuvh::InitialTmunu HYDRO::vectors() const {
    cpuvh::InitialTmunu v;
    v.e = initial_energy_density;
    v.ut = initial_ut; v.ux = initial_ux; v.uy = initial_uy; v.un = initial_un;
    v.pitt = initial_pitt; v.pitx = initial_pitx; v.pity = initial_pity; v.pitn = initial_pitn;
    v.pixx = initial_pixx; v.pixy = initial_pixy; v.pixn = initial_pixn;
    v.piyy = initial_piyy; v.piyn = initial_piyn; v.pinn = initial_pinn;
    v.Pi = initial_Pi;
    v.reference_ut = reference_ut;
    return v;
}

int HYDRO::run_hydro(int argc, char **argv) {
    const cpuvh::InitialTmunu v = vectors();
    return cpuvh::run_main(argc, argv, &v);
}

int HYDRO::run_hydro_no_cli() {
    const cpuvh::InitialTmunu v = vectors();
    return cpuvh::run_no_cli(&v);
}

/* The reference's C++ entry class HYDRO (rhic/include/HydroWrapper.h:28-91,
 * rhic/src/HydroWrapper.cpp:20-136) over this library: the same members and
 * methods, so a caller that drives the reference through HYDRO (e.g. a
 * JETSCAPE module) switches by including this header and linking
 * libcpuvh_driver.so and libcpuvh_amd.so.
 *
 *   run_hydro(argc, argv)  the command-line run (the options of cpuvh_run:
 *                          --hydro --config <dir> --output <dir>)
 *   run_hydro_no_cli()     config "rhic-conf", output "cpu-vh_output" under
 *                          the working directory, initial state from the
 *                          vectors below when initialConditionType is 13
 *                          (initial_energy_density) or 14 (all sixteen)
 *
 * Vectors are x-fastest over the interior (i + nx (j + ny k)); e and the
 * pi/Pi components in GeV/fm^3 as the reference takes them (divided by
 * hbar c there and here, e plus the reference's 1e-3 fm^-4 floor).  Type 14
 * normalises u^tau = sqrt(1 + ux^2 + uy^2 + tau0^2 un^2) by default; with
 * reference_ut set it stores 1 + ux^2 + uy^2 + un^2, as the reference does
 * (InitialConditions.cpp:1681), for runs that must follow the reference.
 * Both methods return 0; a configuration the reference stops on ends the
 * process with the reference's message, as there.  The EoS table and bulk
 * viscosity come from rhic-conf/gpu.properties (the reference's
 * run_hydro_no_cli never loads its table, HydroWrapper.cpp:64-136). */
#pragma once
#include <vector>

namespace cpuvh {

// The sixteen initial vectors of types 13 and 14.
struct InitialTmunu {
    std::vector<double> e, ut, ux, uy, un;
    std::vector<double> pitt, pitx, pity, pitn, pixx, pixy, pixn, piyy, piyn, pinn;
    std::vector<double> Pi;
    bool reference_ut = false;  // type 14: u^tau = 1 + ux^2 + uy^2 + un^2, the reference's value
};

// The command-line run of cpuvh_run (argv as its main receives it) and the
// fixed-directory run, with vectors for IC types 13 and 14 (may be NULL).
int run_main(int argc, char **argv, const InitialTmunu *vectors);
int run_no_cli(const InitialTmunu *vectors);

}  // namespace cpuvh

class HYDRO {
  public:
    HYDRO() = default;
    ~HYDRO() = default;
    std::vector<double> initial_energy_density;
    std::vector<double> initial_ut;
    std::vector<double> initial_ux;
    std::vector<double> initial_uy;
    std::vector<double> initial_un;
    std::vector<double> initial_pitt;
    std::vector<double> initial_pitx;
    std::vector<double> initial_pity;
    std::vector<double> initial_pitn;
    std::vector<double> initial_pixx;
    std::vector<double> initial_pixy;
    std::vector<double> initial_pixn;
    std::vector<double> initial_piyy;
    std::vector<double> initial_piyn;
    std::vector<double> initial_pinn;
    std::vector<double> initial_Pi;
    // type 14: false (default) normalises u^tau; true stores the reference's
    // 1 + ux^2 + uy^2 + un^2 (InitialConditions.cpp:1681)
    bool reference_ut = false;
    int run_hydro(int argc, char **argv);
    int run_hydro_no_cli();
    void initialize_from_vector(std::vector<double> e);
    void initialize_from_vectors(std::vector<double> e, std::vector<double> ut, std::vector<double> ux,
                                 std::vector<double> uy, std::vector<double> un, std::vector<double> pitt,
                                 std::vector<double> pitx, std::vector<double> pity, std::vector<double> pitn,
                                 std::vector<double> pixx, std::vector<double> pixy, std::vector<double> pixn,
                                 std::vector<double> piyy, std::vector<double> piyn, std::vector<double> pinn,
                                 std::vector<double> Pi);

  private:
    cpuvh::InitialTmunu vectors() const;
};

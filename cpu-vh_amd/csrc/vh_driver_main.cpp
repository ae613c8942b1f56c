// cpuvh_run: the command-line driver (cpu-vh_amd/csrc/vh_driver.cpp).
#include "../../include/cpuvh_hydro.hpp"

int main(int argc, char **argv) { return cpuvh::run_main(argc, argv, nullptr); }

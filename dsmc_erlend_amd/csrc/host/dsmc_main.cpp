// dsmc_main.cpp - the dsmc_gpu executable (the reference's src/main.cpp on one GPU).
#include "dsmc_host.hpp"

int main(int argc, char** argv) { return dsmc::host::RunSimulation(argc, argv); }

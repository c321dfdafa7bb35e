// MC3D executable: `MC3D inputfile outputfile [options]` (cpp/3d/MC3D.cpp:34-94).
#include "vmc_cli.hpp"

int main(int argc, char **argv) { return vmc_cli::cli_main<3>(argc, argv); }

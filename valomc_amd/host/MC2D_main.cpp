// MC2D executable: `MC2D inputfile outputfile [-s] [options]` (cpp/2d/MC2D.cpp:38-122).
#include "vmc_cli.hpp"

int main(int argc, char **argv) { return vmc_cli::cli_main<2>(argc, argv); }

// MC2D drop-in: triangular engine with the reference class's public API
// (cpp/2d/MC2D.hpp:59-180); transport runs on the GPU through valomc.h.
#ifndef VMC_HOST_MC2D_HPP
#define VMC_HOST_MC2D_HPP
#include "MCEngine.hpp"
class MC2D : public MCEngine<2> {};
#endif

// MC3D drop-in: tetrahedral engine with the reference class's public API
// (cpp/3d/MC3D.hpp:51-173); transport runs on the GPU through valomc.h.
#ifndef VMC_HOST_MC3D_HPP
#define VMC_HOST_MC3D_HPP
#include "MCEngine.hpp"
class MC3D : public MCEngine<3> {};
#endif

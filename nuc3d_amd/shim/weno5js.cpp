// GPU definitions of nuc3d::weno5js (declared in the reference's weno5js.h;
// the reference's definitions are weno5js.cpp:6-624).
#include "weno5js.h"

#include "nuc3d_shim.hpp"

nuc3d::weno5js::weno5js() : ss(1.0e-6), p(2.0) {}

nuc3d::weno5js::~weno5js() {}

void nuc3d::weno5js::interpolationInner(const Field &fieldIN, const int upwind, Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_WENO5JS, NUC3D_PART_INNER, fieldIN, nullptr, upwind, fieldOUT, tilesize, 0, "weno5js");
}

void nuc3d::weno5js::interpolationBoundaryL(const Field &fieldIN, const Field &boundaryL, const int upwind,
                                            Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_WENO5JS, NUC3D_PART_LEFT, fieldIN, &boundaryL, upwind, fieldOUT, tilesize, 0,
                            "weno5js");
}

void nuc3d::weno5js::interpolationBoundaryR(const Field &fieldIN, const Field &boundaryR, const int upwind,
                                            Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_WENO5JS, NUC3D_PART_RIGHT, fieldIN, &boundaryR, upwind, fieldOUT, tilesize, 0,
                            "weno5js");
}

// GPU definitions of nuc3d::weno5z (declared in the reference's weno5z.hpp;
// the reference's definitions are weno5z.cpp:5-640, including the beta2 of
// its + inner faces, weno5z.cpp:80, which the inner part reproduces).
#include "weno5z.hpp"

#include "nuc3d_shim.hpp"

nuc3d::weno5z::weno5z() : ss(1.0e-6), p(2.0) {}

nuc3d::weno5z::~weno5z() {}

void nuc3d::weno5z::interpolationInner(const Field &fieldIN, const int upwind, Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_WENO5Z, NUC3D_PART_INNER, fieldIN, nullptr, upwind, fieldOUT, tilesize, 1, "weno5z");
}

void nuc3d::weno5z::interpolationBoundaryL(const Field &fieldIN, const Field &boundaryL, const int upwind,
                                           Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_WENO5Z, NUC3D_PART_LEFT, fieldIN, &boundaryL, upwind, fieldOUT, tilesize, 1,
                            "weno5z");
}

void nuc3d::weno5z::interpolationBoundaryR(const Field &fieldIN, const Field &boundaryR, const int upwind,
                                           Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_WENO5Z, NUC3D_PART_RIGHT, fieldIN, &boundaryR, upwind, fieldOUT, tilesize, 1,
                            "weno5z");
}

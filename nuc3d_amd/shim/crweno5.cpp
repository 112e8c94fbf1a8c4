// GPU definitions of nuc3d::crweno5 (declared in the reference's crweno5.hpp;
// the reference's definitions are crweno5.cpp:12-836).  The inner part solves
// one tridiagonal system per line on the GPU with no line-length limit (the
// reference's stack arrays hold 256 rows, crweno5.cpp:69-72).
#include "crweno5.hpp"

#include "nuc3d_shim.hpp"

nuc3d::crweno5::crweno5() : ss(1.0e-6), p(2.0) {}

nuc3d::crweno5::~crweno5() {}

void nuc3d::crweno5::interpolationInner(const Field &fieldIN, const int upwind, Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_CRWENO5, NUC3D_PART_INNER, fieldIN, nullptr, upwind, fieldOUT, tilesize, 0,
                            "crweno5");
}

void nuc3d::crweno5::interpolationBoundaryL(const Field &fieldIN, const Field &boundaryL, const int upwind,
                                            Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_CRWENO5, NUC3D_PART_LEFT, fieldIN, &boundaryL, upwind, fieldOUT, tilesize, 0,
                            "crweno5");
}

void nuc3d::crweno5::interpolationBoundaryR(const Field &fieldIN, const Field &boundaryR, const int upwind,
                                            Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_CRWENO5, NUC3D_PART_RIGHT, fieldIN, &boundaryR, upwind, fieldOUT, tilesize, 0,
                            "crweno5");
}

// GPU definitions of nuc3d::centraldifference6th (declared in the reference's
// centraldifference6th.h; the reference's definitions are
// centraldifference6th.cpp:16-183).
#include "centraldifference6th.h"

#include "nuc3d_shim.hpp"

nuc3d::centraldifference6th::centraldifference6th() {}

nuc3d::centraldifference6th::~centraldifference6th() {}

void nuc3d::centraldifference6th::differentialInner(const Field &fieldIN, Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::differentiate(NUC3D_CD6, NUC3D_PART_INNER, fieldIN, nullptr, fieldOUT, tilesize,
                              "centraldifference6th");
}

void nuc3d::centraldifference6th::differentialBoundaryL(const Field &fieldIN, const Field &boundaryL, Field &fieldOUT,
                                                        const int tilesize)
{
    nuc3d_shim::differentiate(NUC3D_CD6, NUC3D_PART_LEFT, fieldIN, &boundaryL, fieldOUT, tilesize,
                              "centraldifference6th");
}

void nuc3d::centraldifference6th::differentialBoundaryR(const Field &fieldIN, const Field &boundaryR, Field &fieldOUT,
                                                        const int tilesize)
{
    nuc3d_shim::differentiate(NUC3D_CD6, NUC3D_PART_RIGHT, fieldIN, &boundaryR, fieldOUT, tilesize,
                              "centraldifference6th");
}

// GPU definitions of nuc3d::centraldifference8th (declared in the reference's
// centraldifference8th.hpp; the reference's definitions are
// centraldifference8th.cpp:22-192).  The eighth-order stencil reads four
// ghost layers, so a ghost width below 4 is an error here (the reference reads
// past its buffers then).
#include "centraldifference8th.hpp"

#include "nuc3d_shim.hpp"

nuc3d::centraldifference8th::centraldifference8th() {}

nuc3d::centraldifference8th::~centraldifference8th() {}

void nuc3d::centraldifference8th::differentialInner(const Field &fieldIN, Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::differentiate(NUC3D_CD8, NUC3D_PART_INNER, fieldIN, nullptr, fieldOUT, tilesize,
                              "centraldifference8th");
}

void nuc3d::centraldifference8th::differentialBoundaryL(const Field &fieldIN, const Field &boundaryL, Field &fieldOUT,
                                                        const int tilesize)
{
    nuc3d_shim::differentiate(NUC3D_CD8, NUC3D_PART_LEFT, fieldIN, &boundaryL, fieldOUT, tilesize,
                              "centraldifference8th");
}

void nuc3d::centraldifference8th::differentialBoundaryR(const Field &fieldIN, const Field &boundaryR, Field &fieldOUT,
                                                        const int tilesize)
{
    nuc3d_shim::differentiate(NUC3D_CD8, NUC3D_PART_RIGHT, fieldIN, &boundaryR, fieldOUT, tilesize,
                              "centraldifference8th");
}

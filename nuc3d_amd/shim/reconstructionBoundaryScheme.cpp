// GPU definitions of nuc3d::ReconstructionboundaryScheme, the physical-face
// reconstruction closure (declared in the reference's
// reconstructionBoundaryScheme.hpp; the reference's definitions are
// reconstructionBoundaryScheme.cpp:13-474).
#include "reconstructionBoundaryScheme.hpp"

#include "nuc3d_shim.hpp"

nuc3d::ReconstructionboundaryScheme::ReconstructionboundaryScheme() {}

nuc3d::ReconstructionboundaryScheme::~ReconstructionboundaryScheme() {}

// the reference's inner part of the closure writes nothing
// (reconstructionBoundaryScheme.cpp:23-29)
void nuc3d::ReconstructionboundaryScheme::interpolationInner(const Field &, const int, Field &, const int) {}

void nuc3d::ReconstructionboundaryScheme::interpolationBoundaryL(const Field &fieldIN, const Field &boundaryL,
                                                                 const int upwind, Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_RECON_BOUNDARY, NUC3D_PART_LEFT, fieldIN, &boundaryL, upwind, fieldOUT, tilesize, 0,
                            "ReconstructionboundaryScheme");
}

void nuc3d::ReconstructionboundaryScheme::interpolationBoundaryR(const Field &fieldIN, const Field &boundaryR,
                                                                 const int upwind, Field &fieldOUT, const int tilesize)
{
    nuc3d_shim::interpolate(NUC3D_RECON_BOUNDARY, NUC3D_PART_RIGHT, fieldIN, &boundaryR, upwind, fieldOUT, tilesize, 0,
                            "ReconstructionboundaryScheme");
}

#define NUC3D_UNUSED(name) \
    void nuc3d::ReconstructionboundaryScheme::name(double *, double *) { nuc3d_shim::unused_member("ReconstructionboundaryScheme"); }
NUC3D_UNUSED(firstOrderP)
NUC3D_UNUSED(secondOrderP)
NUC3D_UNUSED(thirdOrderP)
NUC3D_UNUSED(fifthOrderP)
NUC3D_UNUSED(firstOrderN)
NUC3D_UNUSED(secondOrderN)
NUC3D_UNUSED(thirdOrderN)
NUC3D_UNUSED(fifthOrderN)
#undef NUC3D_UNUSED

// GPU definitions of nuc3d::boundaryScheme, the physical-face derivative
// closure (declared in the reference's boundaryScheme.hpp; the reference's
// definitions are boundaryScheme.cpp:17-126).
#include "boundaryScheme.hpp"

#include "nuc3d_shim.hpp"

nuc3d::boundaryScheme::boundaryScheme() {}

nuc3d::boundaryScheme::~boundaryScheme() {}

// the reference's inner part of the closure writes nothing (boundaryScheme.cpp:23-28)
void nuc3d::boundaryScheme::differentialInner(const Field &, Field &, const int) {}

void nuc3d::boundaryScheme::differentialBoundaryL(const Field &fieldIN, const Field &boundaryL, Field &fieldOUT,
                                                  const int tilesize)
{
    nuc3d_shim::differentiate(NUC3D_DIFF_BOUNDARY, NUC3D_PART_LEFT, fieldIN, &boundaryL, fieldOUT, tilesize,
                              "boundaryScheme");
}

void nuc3d::boundaryScheme::differentialBoundaryR(const Field &fieldIN, const Field &boundaryR, Field &fieldOUT,
                                                  const int tilesize)
{
    nuc3d_shim::differentiate(NUC3D_DIFF_BOUNDARY, NUC3D_PART_RIGHT, fieldIN, &boundaryR, fieldOUT, tilesize,
                              "boundaryScheme");
}

void nuc3d::boundaryScheme::firstOrderL(double *, double *, int &) { nuc3d_shim::unused_member("boundaryScheme"); }
void nuc3d::boundaryScheme::firstOrderR(double *, double *, int &) { nuc3d_shim::unused_member("boundaryScheme"); }
void nuc3d::boundaryScheme::secondOrder(double *, double *, int &) { nuc3d_shim::unused_member("boundaryScheme"); }
void nuc3d::boundaryScheme::fourthOrder(double *, double *, int &) { nuc3d_shim::unused_member("boundaryScheme"); }
void nuc3d::boundaryScheme::sixthOrder(double *, double *, int &) { nuc3d_shim::unused_member("boundaryScheme"); }

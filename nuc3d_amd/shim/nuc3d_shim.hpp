// Shared helpers of the drop-in operator definitions in this directory.
//
// Each .cpp here defines the members of one of the reference's operator
// classes -- declared by the reference's own header (weno5js.h, weno5z.hpp,
// crweno5.hpp, centraldifference6th.h, centraldifference8th.hpp,
// boundaryScheme.hpp, reconstructionBoundaryScheme.hpp, TVD-RK3.h) -- over the
// operator-level C ABI (include/nuc3d_ops.h), so the reference's
// fieldOperator3d, EulerData3D / NaiverStokesData3d and block driver run on
// the GPU operators unchanged.  Build them in place of the reference's
// translation units of the same names, with the reference's src/ and this
// repository's include/ on the include path, and link libnuc3d_hip.so.
#ifndef NUC3D_SHIM_HPP
#define NUC3D_SHIM_HPP

#include <cstdlib>
#include <iostream>

#include "field.h"
#include "nuc3d_hip.h"
#include "nuc3d_ops.h"

namespace nuc3d_shim {

// The reference reports an operator error on std::cout and exits
// (e.g. weno5js.cpp:205-206); so do these definitions.
inline void check(int status, const char *who)
{
    if (status != NUC3D_OK) {
        std::cout << who << " error: " << nuc3d_last_error() << std::endl;
        std::exit(-1);
    }
}

inline void check_shape(bool ok, const char *who, const char *what)
{
    if (!ok) {
        std::cout << who << " error: " << what << std::endl;
        std::exit(-1);
    }
}

// interpolationInner / BoundaryL / BoundaryR of a reconstruction scheme
inline void interpolate(int scheme, int part, const nuc3d::Field &in, const nuc3d::Field *ghost, int upwind,
                        nuc3d::Field &out, int w, int zquirk, const char *who)
{
    const int nx = in.getSizeX(), ny = in.getSizeY(), nz = in.getSizeZ();
    check_shape(out.getSizeX() == nx + 1 && out.getSizeY() == ny && out.getSizeZ() == nz, who,
                "output Field must have nx+1 faces per line");
    if (ghost)
        check_shape(ghost->getSizeX() == w && ghost->getSizeY() == ny && ghost->getSizeZ() == nz, who,
                    "ghost Field must hold w cells per line");
    check(nuc3d_op_interpolate(scheme, part, upwind, w, zquirk, nx, ny, nz, in.getDataPtr(),
                               ghost ? ghost->getDataPtr() : nullptr, out.getDataPtr(), nullptr),
          who);
}

// differentialInner / BoundaryL / BoundaryR of a derivative scheme
inline void differentiate(int scheme, int part, const nuc3d::Field &in, const nuc3d::Field *ghost,
                          nuc3d::Field &out, int w, const char *who)
{
    const int nx = in.getSizeX(), ny = in.getSizeY(), nz = in.getSizeZ();
    check_shape(out.getSizeX() == nx && out.getSizeY() == ny && out.getSizeZ() == nz, who,
                "output Field must have the input's shape");
    if (ghost && scheme != NUC3D_DIFF_BOUNDARY)
        check_shape(ghost->getSizeX() == w && ghost->getSizeY() == ny && ghost->getSizeZ() == nz, who,
                    "ghost Field must hold w cells per line");
    check(nuc3d_op_differentiate(scheme, part, w, nx, ny, nz, in.getDataPtr(), ghost ? ghost->getDataPtr() : nullptr,
                                 out.getDataPtr(), nullptr),
          who);
}

// members the reference's class declarations name but the GPU definitions
// never call (targets of member-pointer tables initialised in the headers)
[[noreturn]] inline void unused_member(const char *who)
{
    std::cout << who << " error: CPU stencil member called on the GPU operator" << std::endl;
    std::exit(-1);
}

}  // namespace nuc3d_shim

#endif

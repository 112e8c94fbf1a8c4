// ref_metrics.cpp -- runs the reference's own block::initial (block.cpp:31-98)
// on a mesh file and dumps what it computed: cell centres (getXYZ_center,
// block.cpp:147-235, tiled 5-point Lagrange), the nine coordinate
// derivatives, J and the nine metric terms (getJacobians, block.cpp:237-325)
// and the initial conserved state (initialQ, block.cpp:523-567).  TEST
// INFRASTRUCTURE ONLY: compiled with the reference's unmodified sources by
// oracle/Makefile (target "refmetrics") into oracle/_ref/.
//
// usage, in a case directory holding mesh/mesh_<rank>.x, inp/PhysModel.in,
// inp/Method.in, inp/IOController.in, bc_topo/BC_Topo_<rank>.in and flowData/:
//   ref_metrics        -> metrics_<rank>.bin (and block::initial's own
//                         flowData/GEO__id_<rank>.dat, block.cpp:87)
//   ref_metrics output -> the same, then block::Output (block.cpp:620-654):
//                         flowData/NUC3d_step_<step>_id_<rank>.dat (Tecplot)
//                         and the binary restart file
// metrics_<rank>.bin: int32 nx,ny,nz, then float64 J[N], met[9N] (xi_x..zeta_z),
//                     dxyz[9N] (x_xi,x_eta,x_zeta,y_xi..z_zeta), centre[3N], Q[5N]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "mpi.h"
#include "IOcontroller.h"
#include "MPICommunicator.h"
#include "block.h"
#include "boundaryConditions.hpp"
#include "euler3d.h"
#include "fieldOperator.h"
#include "physicsModel.h"

using namespace nuc3d;

namespace {

void dump(FILE *fp, const Field &f) { std::fwrite(f.getDataPtr(), sizeof(double), f.getSize(), fp); }

// reads the protected results of block::initial
struct MetricBlock : public block {
    void write(int rank)
    {
        char name[64];
        std::snprintf(name, sizeof name, "metrics_%d.bin", rank);
        FILE *fp = std::fopen(name, "wb");
        if (!fp) MPI_Abort(MPI_COMM_WORLD, 3);
        int dims[3] = {nx, ny, nz};
        std::fwrite(dims, sizeof(int), 3, fp);
        dump(fp, myFluxes->getJac());
        for (auto *v : {&myFluxes->getXi_xyz(), &myFluxes->getEta_xyz(), &myFluxes->getZeta_xyz()})
            for (auto &f : *v) dump(fp, f);
        for (auto *v : {&myFluxes->getDx(), &myFluxes->getDy(), &myFluxes->getDz()})
            for (auto &f : *v) dump(fp, f);
        for (auto &f : xyz_center) dump(fp, f);
        for (auto &f : myPDE.getQ()) dump(fp, f);
        std::fclose(fp);
    }
};

}  // namespace

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    {
        MPIComunicator3d_nonblocking mpi;  // its destructor calls MPI_Finalize
        physicsModel phys;
        fieldOperator3d op;
        boundaryCondition bc;
        IOController io;
        MetricBlock blk;
        blk.initial(op, phys, mpi, bc, io);
        blk.write(mpi.getMyId());
        if (argc > 1 && std::string(argv[1]) == "output") blk.Output(op, phys, mpi, bc, io);
    }
    return 0;
}

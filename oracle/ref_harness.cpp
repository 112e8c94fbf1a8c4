// ref_harness.cpp -- drives the reference's own classes on one block per MPI
// rank and dumps what the hot path produced.  TEST INFRASTRUCTURE ONLY: it is
// compiled together with the reference's unmodified sources (see
// oracle/Makefile, target "ref") into oracle/_ref/, and is used to pin the C
// restatement in nuc3d_oracle.c and to write golden fixtures.
//
// The call sequence is the reference's own: block::initialData
// (block.cpp:479-520) for the buffers, boundaryCondition::initialBC
// (boundaryConditions.cpp:25-66) for the topology, then per RK stage
// EulerData3D::solve / NaiverStokesData3d::solve followed by
// PDEData3d::solve (block.cpp:569-594).
//
// usage (run with the case directory as the working directory, which holds
// inp/PhysModel.in, inp/Method.in, bc_topo/BC_Topo_<rank>.in, in_<rank>.bin):
//   ref_harness rhs                 one RHS evaluation -> out_<rank>.bin
//   ref_harness steps <n> <cfl>     n TVD-RK3 steps    -> out_<rank>.bin
//   ref_harness post                postproc::solveQ   -> enstrophy.dat (+ out_<rank>.bin)
//                                   (needs inp/PostProc.in with SimutaneousData 1)
//
// in_<rank>.bin : int32 nx,ny,nz, then float64 J[N], met[9N], dxyz[9N], Q[5N]
//                 (met: xi_x..zeta_z; dxyz: x_xi,x_eta,x_zeta,y_xi..z_zeta;
//                  Q = conserved/J, i-fastest)
// out_<rank>.bin: float64 RHS[5N], Q[5N], W[5N] (rho,u,v,w,p), W0[3N] (T,e,c),
//                 dt_local, dt_global, maxEigen[3], nsteps, dt[nsteps],
//                 wall[nsteps] (MPI_Wtime around each step, max over ranks),
//                 res[nsteps] (PDEData3d::getRes after each step)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mpi.h"
#include "NaiverStokes3d.h"
#include "PDEData3d.hpp"
#include "boundaryConditions.hpp"
#include "bufferData.hpp"
#include "euler3d.h"
#include "fieldOperator.h"
#include "physicsModel.h"
#include "IOcontroller.h"
#include "postproc.hpp"

using namespace nuc3d;

static void load(Field &f, const double *src)
{
    std::memcpy(f.getDataPtr(), src, sizeof(double) * f.getSize());
}

static void dump(FILE *fp, const Field &f)
{
    std::fwrite(f.getDataPtr(), sizeof(double), f.getSize(), fp);
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    if (argc < 2) {
        std::fprintf(stderr, "usage: ref_harness rhs | steps <n> <cfl>\n");
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    std::string mode = argv[1];
    int nsteps = (mode == "steps" && argc > 2) ? std::atoi(argv[2]) : 0;
    double cfl = (mode == "steps" && argc > 3) ? std::atof(argv[3]) : 0.5;
    {
        MPIComunicator3d_nonblocking mpi;  // its destructor calls MPI_Finalize
        physicsModel phys;
        fieldOperator3d op;
        boundaryCondition bc;
        int rank = mpi.getMyId();

        char name[64];
        std::snprintf(name, sizeof name, "in_%d.bin", rank);
        FILE *fp = std::fopen(name, "rb");
        if (!fp) {
            std::fprintf(stderr, "missing %s\n", name);
            MPI_Abort(MPI_COMM_WORLD, 3);
        }
        int dims[3];
        if (std::fread(dims, sizeof(int), 3, fp) != 3) MPI_Abort(MPI_COMM_WORLD, 4);
        const int nx = dims[0], ny = dims[1], nz = dims[2];
        const long N = (long)nx * ny * nz;
        std::vector<double> buf(24 * N);
        if (std::fread(buf.data(), sizeof(double), 24 * N, fp) != (size_t)(24 * N)) MPI_Abort(MPI_COMM_WORLD, 5);
        std::fclose(fp);

        const int neqs = phys.getEqNum();
        PDEData3d pde;
        pde.initPDEData3d(nx, ny, nz, neqs);
        std::shared_ptr<EulerData3D> flux;
        if (phys.getMyModelName() == "NavierStokes3d")
            flux = std::make_shared<NaiverStokesData3d>(nx, ny, nz, neqs);
        else
            flux = std::make_shared<EulerData3D>(nx, ny, nz, neqs);

        int bfsize = std::max(op.bufferSize_xi, op.bufferSize_eta);
        bfsize = std::max(bfsize, op.bufferSize_zeta);
        VectorBuffer bufs;
        for (int n = 0; n < neqs; n++) bufs.push_back(bufferData(nx, ny, nz, bfsize));
        bc.initialBC(bufs, mpi);

        load(flux->jacobian, &buf[0]);
        for (int c = 0; c < 3; c++) {
            load(flux->xi_xyz[c], &buf[(1 + c) * N]);
            load(flux->eta_xyz[c], &buf[(4 + c) * N]);
            load(flux->zeta_xyz[c], &buf[(7 + c) * N]);
            load(flux->dx[c], &buf[(10 + c) * N]);
            load(flux->dy[c], &buf[(13 + c) * N]);
            load(flux->dz[c], &buf[(16 + c) * N]);
        }
        for (int c = 0; c < 5; c++) load(pde.getQ()[c], &buf[(19 + c) * N]);

        std::vector<double> dts, walls, res;
        if (mode == "rhs") {
            flux->solve(pde, op, bufs, phys, mpi, bc);
        } else if (mode == "post") {
            // block::Post (block.cpp:600-616): getPrim, then postproc::solvePost,
            // which with "SimutaneousData 1" in inp/PostProc.in runs solveQ and
            // appends step, time, enstrophy and kinetic energy to enstrophy.dat
            // on rank 0 (postproc.cpp:269-296, 725-848)
            VectorField prim, acoust, xyz;
            for (int c = 0; c < 5; c++) prim.push_back(Field(nx, ny, nz));
            for (int c = 0; c < 3; c++) acoust.push_back(Field(nx, ny, nz));
            phys.getPrim(flux->jacobian, pde.getQ(), prim, acoust);
            IOController io;
            postproc post(nx, ny, nz);
            post.solvePost(prim, acoust, xyz, flux->xi_xyz, flux->eta_xyz, flux->zeta_xyz, phys, op, bufs, mpi, bc,
                           io, 0, 0.0);
        } else {
            for (int s = 0; s < nsteps; s++) {
                MPI_Barrier(MPI_COMM_WORLD);
                double t0 = MPI_Wtime();
                for (int stage = 0; stage < op.getSteps(); stage++) {
                    flux->solve(pde, op, bufs, phys, mpi, bc);
                    pde.solve(op, cfl, stage);
                }
                double w = MPI_Wtime() - t0, wmax = 0.0;
                MPI_Allreduce(&w, &wmax, 1, MPI_DOUBLE, MPI_MAX, MPI_COMM_WORLD);
                dts.push_back(pde.getDt());
                walls.push_back(wmax);
                res.push_back(pde.getRes());
            }
        }

        std::snprintf(name, sizeof name, "out_%d.bin", rank);
        fp = std::fopen(name, "wb");
        for (int c = 0; c < 5; c++) dump(fp, pde.getRHS()[c]);
        for (int c = 0; c < 5; c++) dump(fp, pde.getQ()[c]);
        for (int c = 0; c < 5; c++) dump(fp, flux->W_Euler[c]);
        for (int c = 0; c < 3; c++) dump(fp, flux->W0_Euler[c]);
        double tail[6] = {flux->dt, pde.getDt(), flux->Flux_xi.maxEigen, flux->Flux_eta.maxEigen,
                          flux->Flux_zeta.maxEigen, (double)dts.size()};
        std::fwrite(tail, sizeof(double), 6, fp);
        if (!dts.empty()) std::fwrite(dts.data(), sizeof(double), dts.size(), fp);
        if (!walls.empty()) std::fwrite(walls.data(), sizeof(double), walls.size(), fp);
        if (!res.empty()) std::fwrite(res.data(), sizeof(double), res.size(), fp);
        std::fclose(fp);
    }
    return 0;
}

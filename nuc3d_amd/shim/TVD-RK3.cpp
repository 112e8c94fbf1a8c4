// GPU definitions of nuc3d::tvdrk3rd (declared in the reference's TVD-RK3.h;
// the reference's definitions are TVD-RK3.cpp:5-137).
#include "TVD-RK3.h"

#include <vector>

#include "nuc3d_shim.hpp"

nuc3d::tvdrk3rd::tvdrk3rd() : integration(3) {}

nuc3d::tvdrk3rd::~tvdrk3rd() {}

void nuc3d::tvdrk3rd::integrationAll(const VectorField &RHS, const VectorField &u_n, VectorField &u_i, double dt,
                                     int nstep)
{
    // stages other than 0, 1, 2 leave u_i untouched (TVD-RK3.cpp:17-30)
    if (nstep < 0 || nstep > 2) return;
    const int nf = (int)u_i.size();
    nuc3d_shim::check_shape(RHS.size() >= u_i.size() && u_n.size() >= u_i.size(), "tvdrk3rd",
                            "RHS and u_n need as many fields as u_i");
    if (nf == 0) return;
    const long long n = (long long)u_n.begin()->getSize();
    std::vector<const double *> r(nf), un(nf);
    std::vector<double *> ui(nf);
    for (int f = 0; f < nf; f++) {
        nuc3d_shim::check_shape(RHS[f].getSize() == n && u_n[f].getSize() == n && u_i[f].getSize() == n, "tvdrk3rd",
                                "fields of different sizes");
        r[f] = RHS[f].getDataPtr();
        un[f] = u_n[f].getDataPtr();
        ui[f] = u_i[f].getDataPtr();
    }
    nuc3d_shim::check(nuc3d_op_rk3(nstep, dt, nf, n, r.data(), un.data(), ui.data(), nullptr), "tvdrk3rd");
}

// the CPU stage loops the header declares are not used by this definition
void nuc3d::tvdrk3rd::rk1st(const VectorField &, const VectorField &, VectorField &, double)
{
    nuc3d_shim::unused_member("tvdrk3rd");
}
void nuc3d::tvdrk3rd::rk2nd(const VectorField &, const VectorField &, VectorField &, double)
{
    nuc3d_shim::unused_member("tvdrk3rd");
}
void nuc3d::tvdrk3rd::rk3rd(const VectorField &, const VectorField &, VectorField &, double)
{
    nuc3d_shim::unused_member("tvdrk3rd");
}

#pragma once

// Real scalar with a double-well potential:
//   phi'' = Lap(phi) - lambda * phi * (phi^2 - 1).
// The Laplacian order (1, 2 or 3 = 2nd/4th/6th order) and lambda come from
// the inputs file (project.laplacian_order, project.lambda).

#include <sledgehamr.h>

namespace RealScalar {

SLEDGEHAMR_ADD_SCALARS(Phi)
SLEDGEHAMR_ADD_CONJUGATE_MOMENTA(dPhi)

AMREX_GPU_DEVICE AMREX_FORCE_INLINE
void Rhs(const amrex::Array4<double>& rhs,
         const amrex::Array4<const double>& state,
         const int i, const int j, const int k, const int lev,
         const double time, const double dt, const double dx,
         const double* params) {
    const int order = static_cast<int>(params[0]);
    const double lambda = params[1];
    const double Phi = state(i, j, k, Scalar::Phi);

    double laplacian_Phi = 0;
    switch (order) {
        case 1:
            laplacian_Phi = sledgehamr::utils::Laplacian<1>(
                    state, i, j, k, Scalar::Phi, dx*dx);
            break;
        case 2:
            laplacian_Phi = sledgehamr::utils::Laplacian<2>(
                    state, i, j, k, Scalar::Phi, dx*dx);
            break;
        case 3:
            laplacian_Phi = sledgehamr::utils::Laplacian<3>(
                    state, i, j, k, Scalar::Phi, dx*dx);
            break;
    }

    rhs(i, j, k, Scalar::Phi)  = state(i, j, k, Scalar::dPhi);
    rhs(i, j, k, Scalar::dPhi) = laplacian_Phi - lambda*Phi*(Phi*Phi - 1.);
}

SLEDGEHAMR_FINISH_SETUP

class RealScalar : public sledgehamr::Sledgehamr {
  public:
    SLEDGEHAMR_INITIALIZE_PROJECT(RealScalar)

    void Init() override {
        amrex::ParmParse pp("project");
        pp.query("laplacian_order", laplacian_order);
        pp.query("lambda", lambda);
    }

    void SetParamsRhs(std::vector<double>& params, const double time,
                      const int lev) override {
        params = {static_cast<double>(laplacian_order), lambda};
    }

  private:
    int laplacian_order = 1;
    double lambda = 1.;
};

}; // namespace RealScalar

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

// Output functions for spectra and projections.
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
double ScaledPhi(const amrex::Array4<const double>& state, const int i,
                 const int j, const int k, const int lev, const double time,
                 const double dt, const double dx, const double* params) {
    return params[0] * state(i, j, k, Scalar::Phi);
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
double dPhi2(const amrex::Array4<const double>& state, const int i,
             const int j, const int k, const int lev, const double time,
             const double dt, const double dx, const double* params) {
    const double dPhi = state(i, j, k, Scalar::dPhi);
    return dPhi * dPhi;
}

AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
double GradientEnergy(const amrex::Array4<const double>& state, const int i,
                      const int j, const int k, const int lev,
                      const double time, const double dt, const double dx,
                      const double* params) {
    const double gx = sledgehamr::utils::Gradient<2>(state, i, j, k, Scalar::Phi, dx, 'x');
    const double gy = sledgehamr::utils::Gradient<2>(state, i, j, k, Scalar::Phi, dx, 'y');
    const double gz = sledgehamr::utils::Gradient<2>(state, i, j, k, Scalar::Phi, dx, 'z');
    return 0.5 * (gx * gx + gy * gy + gz * gz);
}

SLEDGEHAMR_OUTPUT_FUNCTIONS(ScaledPhi, dPhi2, GradientEnergy)

SLEDGEHAMR_FINISH_SETUP

class RealScalar : public sledgehamr::Sledgehamr {
  public:
    SLEDGEHAMR_INITIALIZE_PROJECT(RealScalar)

    void Init() override {
        amrex::ParmParse pp("project");
        pp.query("laplacian_order", laplacian_order);
        pp.query("lambda", lambda);
        pp.query("output_scale", output_scale);
    }

    void SetParamsSpectra(std::vector<double>& params,
                          const double time) override {
        params = {output_scale};
    }

    void SetParamsProjections(std::vector<double>& params,
                              const double time) override {
        params = {output_scale};
    }

    void SetParamsRhs(std::vector<double>& params, const double time,
                      const int lev) override {
        params = {static_cast<double>(laplacian_order), lambda};
    }

  private:
    int laplacian_order = 1;
    double lambda = 1.;
    double output_scale = 1.;
};

}; // namespace RealScalar

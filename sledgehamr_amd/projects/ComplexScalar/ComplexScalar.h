#pragma once

// Complex scalar psi = Psi1 + i Psi2 in a radiation-dominated background,
// V = (lambda/4)(|psi|^2 - 1)^2 with conformal time eta = time:
//   Pi' = -(2/eta) Pi + Lap(Psi) - lambda * Psi * (|psi|^2 - 1).
// 4th-order Laplacian (Laplacian<2>); lambda from project.lambda.

#include <sledgehamr.h>

namespace ComplexScalar {

SLEDGEHAMR_ADD_SCALARS(Psi1, Psi2)
SLEDGEHAMR_ADD_CONJUGATE_MOMENTA(Pi1, Pi2)

AMREX_GPU_DEVICE AMREX_FORCE_INLINE
void Rhs(const amrex::Array4<double>& rhs,
         const amrex::Array4<const double>& state,
         const int i, const int j, const int k, const int lev,
         const double time, const double dt, const double dx,
         const double* params) {
    const double lambda = params[0];
    const double damping = params[1];  // 2/eta at this stage time
    double Psi1 = state(i, j, k, Scalar::Psi1);
    double Psi2 = state(i, j, k, Scalar::Psi2);
    double Pi1  = state(i, j, k, Scalar::Pi1);
    double Pi2  = state(i, j, k, Scalar::Pi2);

    constexpr int order = 2;
    double laplacian_Psi1 = sledgehamr::utils::Laplacian<order>(
            state, i, j, k, Scalar::Psi1, dx*dx);
    double laplacian_Psi2 = sledgehamr::utils::Laplacian<order>(
            state, i, j, k, Scalar::Psi2, dx*dx);

    double potential = lambda*(Psi1*Psi1 + Psi2*Psi2 - 1.);

    rhs(i, j, k, Scalar::Psi1) =  Pi1;
    rhs(i, j, k, Scalar::Psi2) =  Pi2;
    rhs(i, j, k, Scalar::Pi1)  = -Pi1*damping + laplacian_Psi1 - Psi1*potential;
    rhs(i, j, k, Scalar::Pi2)  = -Pi2*damping + laplacian_Psi2 - Psi2*potential;
}

SLEDGEHAMR_FINISH_SETUP

class ComplexScalar : public sledgehamr::Sledgehamr {
  public:
    SLEDGEHAMR_INITIALIZE_PROJECT(ComplexScalar)

    void Init() override {
        amrex::ParmParse pp("project");
        pp.query("lambda", lambda);
    }

    void SetParamsRhs(std::vector<double>& params, const double time,
                      const int lev) override {
        // Conformal time eta = time; the damping 2/eta is evaluated once per
        // stage here instead of once per cell.
        params = {lambda, 2. / time};
    }

  private:
    double lambda = 1.;
};

}; // namespace ComplexScalar

#pragma once

// Complex scalar psi = Psi1 + i Psi2 in a radiation-dominated background,
// V = (lambda/4)(|psi|^2 - 1)^2 with conformal time eta = time:
//   Pi' = -(2/eta) Pi + Lap(Psi) - lambda * Psi * (|psi|^2 - 1).
// 4th-order Laplacian (Laplacian<2>); lambda from project.lambda.  The file
// layout follows the reference's projects/MinimalExample/MinimalExample.h
// (the project API it is written against); the physics is this workload's.
//
// Refinement criteria, both off by default:
//   project.tag_strings = 1   tags every cell whose plaquettes in the +x/+y,
//                             +x/+z and +y/+z planes are pierced by a string
//                             (nonzero winding of psi around the plaquette);
//   project.te_modifier = 1   compares te * dt for the momenta and
//                             te / max|Psi_n| (level 0, set per tagging) for
//                             the fields against the thresholds, instead of
//                             te itself.

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

/** Signed crossing of the positive real axis by the straight segment a -> b
 *  in the complex plane: +1 upward, -1 downward, 0 if the segment meets the
 *  axis at Re <= 0 or not at all.  Half-open in Im, so a corner on the axis
 *  is counted once.  The crossing abscissa (ar bi - br ai) / (bi - ai) is
 *  positive iff the two products compare as below (no division). */
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
int PositiveAxisCrossing(double ar, double ai, double br, double bi) {
    if (ai < 0. && bi >= 0.) return ar * bi > ai * br ? 1 : 0;
    if (ai >= 0. && bi < 0.) return ar * bi < ai * br ? -1 : 0;
    return 0;
}

/** Winding number of psi around the plaquette spanned at (i, j, k) by the
 *  unit steps a and b (corners (i,j,k), +a, +a+b, +b): the signed count of
 *  crossings of a ray from the origin by the closed polygon through the
 *  corner values. */
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
int PlaquetteWinding(const amrex::Array4<const double>& state,
                     const int i, const int j, const int k,
                     const int ax, const int ay, const int az,
                     const int bx, const int by, const int bz) {
    const int ci[5] = {i, i + ax, i + ax + bx, i + bx, i};
    const int cj[5] = {j, j + ay, j + ay + by, j + by, j};
    const int ck[5] = {k, k + az, k + az + bz, k + bz, k};
    int w = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e)
        w += PositiveAxisCrossing(
            state(ci[e], cj[e], ck[e], Scalar::Psi1),
            state(ci[e], cj[e], ck[e], Scalar::Psi2),
            state(ci[e + 1], cj[e + 1], ck[e + 1], Scalar::Psi1),
            state(ci[e + 1], cj[e + 1], ck[e + 1], Scalar::Psi2));
    return w;
}

/** User criterion: a string pierces one of the cell's three plaquettes in
 *  the positive index directions.  params[0] != 0 switches it on. */
template<> AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
bool TagCellForRefinement<true>(const amrex::Array4<const double>& state,
        const int i, const int j, const int k, const int lev, const double time,
        const double dt, const double dx, const double* params) {
    if (params[0] == 0.) return false;
    return PlaquetteWinding(state, i, j, k, 1, 0, 0, 0, 1, 0) != 0 ||
           PlaquetteWinding(state, i, j, k, 1, 0, 0, 0, 0, 1) != 0 ||
           PlaquetteWinding(state, i, j, k, 0, 1, 0, 0, 0, 1) != 0;
}

/** Truncation-error modifiers (params[0] != 0 switches them on): fields
 *  relative to their level-0 maximum params[1 + n], momenta times the
 *  level's dt. */
template<> AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
double TruncationModifier<Scalar::Psi1>(const amrex::Array4<const double>& state,
        const int i, const int j, const int k, const int lev, const double time,
        const double dt, const double dx, const double truncation_error,
        const double* params) {
    const double norm = params[1 + Scalar::Psi1];
    return params[0] != 0. && norm > 0. ? truncation_error / norm : truncation_error;
}
template<> AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
double TruncationModifier<Scalar::Psi2>(const amrex::Array4<const double>& state,
        const int i, const int j, const int k, const int lev, const double time,
        const double dt, const double dx, const double truncation_error,
        const double* params) {
    const double norm = params[1 + Scalar::Psi2];
    return params[0] != 0. && norm > 0. ? truncation_error / norm : truncation_error;
}
template<> AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
double TruncationModifier<Scalar::Pi1>(const amrex::Array4<const double>& state,
        const int i, const int j, const int k, const int lev, const double time,
        const double dt, const double dx, const double truncation_error,
        const double* params) {
    return params[0] != 0. ? truncation_error * dt : truncation_error;
}
template<> AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE
double TruncationModifier<Scalar::Pi2>(const amrex::Array4<const double>& state,
        const int i, const int j, const int k, const int lev, const double time,
        const double dt, const double dx, const double truncation_error,
        const double* params) {
    return params[0] != 0. ? truncation_error * dt : truncation_error;
}

SLEDGEHAMR_FINISH_SETUP

class ComplexScalar : public sledgehamr::Sledgehamr {
  public:
    SLEDGEHAMR_INITIALIZE_PROJECT(ComplexScalar)

    void Init() override {
        amrex::ParmParse pp("project");
        pp.query("lambda", lambda);
        pp.query("tag_strings", tag_strings);
        pp.query("te_modifier", te_modifier);
    }

    void SetParamsRhs(std::vector<double>& params, const double time,
                      const int lev) override {
        // Conformal time eta = time; the damping 2/eta is evaluated once per
        // stage here instead of once per cell.
        params = {lambda, 2. / time};
    }

    void SetParamsTagCellForRefinement(std::vector<double>& params,
                                       const double time,
                                       const int lev) override {
        params = {tag_strings ? 1. : 0.};
    }

    void SetParamsTruncationModifier(std::vector<double>& params,
                                     const double time,
                                     const int lev) override {
        // The field norms are taken afresh at every tagging.
        params = {te_modifier ? 1. : 0.};
        if (te_modifier) {
            const std::vector<double> m = LevelNorm0(0);
            params.insert(params.end(), m.begin(), m.end());
        }
        params.resize(1 + Scalar::NScalars, 0.);
    }

  private:
    double lambda = 1.;
    int tag_strings = 0;
    int te_modifier = 0;
};

}; // namespace ComplexScalar

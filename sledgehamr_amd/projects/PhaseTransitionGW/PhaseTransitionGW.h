#pragma once

// First-order phase transition with gravitational waves.
//   V = phi^2/2 - delta phi^3/3 + phi^4/4  ->  Phi'' = Lap(Phi) - V'(Phi)
//   u_ij'' = Lap(u_ij) + d_i Phi d_j Phi
// 6th-order Laplacian and gradient (Laplacian<3>, Gradient<3>).  Two GW
// spectrum modifiers: |k| u_ij, and du_ij plus a translated copy.

#include <sledgehamr.h>

namespace PhaseTransitionGW {

SLEDGEHAMR_ADD_SCALARS(Phi)
SLEDGEHAMR_ADD_CONJUGATE_MOMENTA(dPhi)

AMREX_GPU_DEVICE AMREX_FORCE_INLINE
void Rhs(const amrex::Array4<double>& rhs,
         const amrex::Array4<const double>& state,
         const int i, const int j, const int k, const int lev,
         const double time, const double dt, const double dx,
         const double* params) {
    const double delta = params[0];
    double Phi = state(i, j, k, Scalar::Phi);

    constexpr int order = 3;
    double laplacian_Phi = sledgehamr::utils::Laplacian<order>(
        state, i, j, k, Scalar::Phi, dx * dx);
    double dV = Phi - delta * Phi * Phi + Phi * Phi * Phi;

    rhs(i, j, k, Scalar::Phi) = state(i, j, k, Scalar::dPhi);
    rhs(i, j, k, Scalar::dPhi) = laplacian_Phi - dV;
}

template <>
AMREX_GPU_DEVICE AMREX_FORCE_INLINE void GravitationalWavesRhs<true>(
    const amrex::Array4<double> &rhs, const amrex::Array4<const double> &state,
    const int i, const int j, const int k, const int lev, const double time,
    const double dt, const double dx, const double *params) {
    double dx2 = dx * dx;
    constexpr int order = 3;
    double laplacian_u_xx =
        sledgehamr::utils::Laplacian<order>(state, i, j, k, Gw::u_xx, dx2);
    double laplacian_u_yy =
        sledgehamr::utils::Laplacian<order>(state, i, j, k, Gw::u_yy, dx2);
    double laplacian_u_zz =
        sledgehamr::utils::Laplacian<order>(state, i, j, k, Gw::u_zz, dx2);
    double laplacian_u_xy =
        sledgehamr::utils::Laplacian<order>(state, i, j, k, Gw::u_xy, dx2);
    double laplacian_u_xz =
        sledgehamr::utils::Laplacian<order>(state, i, j, k, Gw::u_xz, dx2);
    double laplacian_u_yz =
        sledgehamr::utils::Laplacian<order>(state, i, j, k, Gw::u_yz, dx2);

    double gx = sledgehamr::utils::Gradient<order>(state, i, j, k, Scalar::Phi, dx, 'x');
    double gy = sledgehamr::utils::Gradient<order>(state, i, j, k, Scalar::Phi, dx, 'y');
    double gz = sledgehamr::utils::Gradient<order>(state, i, j, k, Scalar::Phi, dx, 'z');

    rhs(i, j, k, Gw::u_xx) = state(i, j, k, Gw::du_xx);
    rhs(i, j, k, Gw::u_yy) = state(i, j, k, Gw::du_yy);
    rhs(i, j, k, Gw::u_zz) = state(i, j, k, Gw::du_zz);
    rhs(i, j, k, Gw::u_xy) = state(i, j, k, Gw::du_xy);
    rhs(i, j, k, Gw::u_xz) = state(i, j, k, Gw::du_xz);
    rhs(i, j, k, Gw::u_yz) = state(i, j, k, Gw::du_yz);
    rhs(i, j, k, Gw::du_xx) = laplacian_u_xx + gx * gx;
    rhs(i, j, k, Gw::du_yy) = laplacian_u_yy + gy * gy;
    rhs(i, j, k, Gw::du_zz) = laplacian_u_zz + gz * gz;
    rhs(i, j, k, Gw::du_xy) = laplacian_u_xy + gx * gy;
    rhs(i, j, k, Gw::du_xz) = laplacian_u_xz + gx * gz;
    rhs(i, j, k, Gw::du_yz) = laplacian_u_yz + gy * gz;
}

SLEDGEHAMR_FINISH_SETUP

/** Spectrum of |k| u_ij instead of du_ij: the six u components, each mode
 *  scaled by its wave number |n| dk (n the signed integer wave vector). */
struct SpectrumUTimesK {
    static void SelectComponents(int components[6]) {
        const int c[6] = {Gw::u_xx, Gw::u_xy, Gw::u_xz, Gw::u_yy, Gw::u_yz, Gw::u_zz};
        for (int n = 0; n < 6; ++n) components[n] = c[n];
    }
    AMREX_GPU_HOST_DEVICE static void FourierSpaceModification(
        double re[6], double im[6], int kx, int ky, int kz, double dk, int dimN,
        const double* params) {
        const double k = std::sqrt((double)(kx * kx + ky * ky + kz * kz)) * dk;
        for (int n = 0; n < 6; ++n) {
            re[n] *= k;
            im[n] *= k;
        }
    }
};

/** Spectrum of du_ij(x) + du_ij(x + d) for the distance d = (params[0],
 *  params[1], params[2]) (same length unit as sim.L): with the forward
 *  transform's exp(-i k.x), each mode times 1 + exp(i k.d). */
struct SpectrumPlusTranslatedCopy {
    static void SelectComponents(int components[6]) {
        const int c[6] = {Gw::du_xx, Gw::du_xy, Gw::du_xz, Gw::du_yy, Gw::du_yz,
                          Gw::du_zz};
        for (int n = 0; n < 6; ++n) components[n] = c[n];
    }
    AMREX_GPU_HOST_DEVICE static void FourierSpaceModification(
        double re[6], double im[6], int kx, int ky, int kz, double dk, int dimN,
        const double* params) {
        const double ph = (kx * params[0] + ky * params[1] + kz * params[2]) * dk;
        const double c = std::cos(ph), s = std::sin(ph);
        for (int n = 0; n < 6; ++n) {
            const double r = re[n], q = im[n];
            re[n] = r + (c * r - s * q);
            im[n] = q + (s * r + c * q);
        }
    }
};

SLEDGEHAMR_GW_SPECTRUM_MODIFIERS(SpectrumUTimesK, SpectrumPlusTranslatedCopy)

class PhaseTransitionGW : public sledgehamr::Sledgehamr {
  public:
    SLEDGEHAMR_INITIALIZE_PROJECT(PhaseTransitionGW)

    void Init() override {
        amrex::ParmParse pp("project");
        pp.query("delta", delta);
    }

    void SetParamsRhs(std::vector<double>& params, const double time,
                      const int lev) override {
        params = {delta};
    }

  private:
    double delta = 1.5;
};

}; // namespace PhaseTransitionGW

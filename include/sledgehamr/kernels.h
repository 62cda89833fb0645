// Cell kernels of the drop-in API.
//
// Reference interface: sledgehamr::kernels::AverageDownWithTruncationError
// (source/kernels.h:18-42) and KreissOligerDissipation<2,3>
// (source/kernels.h:56-106).  Same signatures and semantics; the fused MI355X
// kernels use shm::te_cell() below, which performs the identical FP64
// operation sequence so tag masks computed from it are bit-exact.
#pragma once

#include <type_traits>

#include "amrex_compat.h"

namespace shm {

/** Sum of the 8 fine children of coarse cell (i,j,k) in the fixed order
 *  (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0) (1,0,1) (0,1,1) (1,1,1).  The order
 *  is part of the result (FP64 addition is not associative), so every caller
 *  that must agree bit-for-bit uses this one routine. */
template <class Get>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double child_sum(const Get &f, int ii,
                                                          int jj, int kk) {
#pragma clang fp contract(off)
    double c = f(ii, jj, kk);
    c += f(ii + 1, jj, kk);
    c += f(ii, jj + 1, kk);
    c += f(ii, jj, kk + 1);
    c += f(ii + 1, jj + 1, kk);
    c += f(ii + 1, jj, kk + 1);
    c += f(ii, jj + 1, kk + 1);
    c += f(ii + 1, jj + 1, kk + 1);
    return c;
}

/** Truncation error of one coarse value against the mean of its children,
 *  returning the new (averaged) coarse value through `avg`. */
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double te_cell(double crse,
                                                        double child_sum_,
                                                        double *avg) {
#pragma clang fp contract(off)
    const double mean = 0.125 * child_sum_;
    *avg = mean;
    return std::fabs(crse - mean);
}

}  // namespace shm

namespace sledgehamr {
namespace kernels {

/** Average 2x2x2 fine cells onto coarse cell (i,j,k) for ncomp components,
 *  storing |crse - <fine>| into te at the even fine index (2i,2j,2k). */
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE void
AverageDownWithTruncationError(int i, int j, int k, const int ncomp,
                               amrex::Array4<double> const &crse,
                               amrex::Array4<double const> const &fine,
                               amrex::Array4<double> const &te) {
    const int ii = 2 * i, jj = 2 * j, kk = 2 * k;
    for (int n = 0; n < ncomp; ++n) {
        auto f = [&](int a, int b, int c) { return fine(a, b, c, n); };
        double avg;
        te(ii, jj, kk, n) =
            shm::te_cell(crse(i, j, k, n), shm::child_sum(f, ii, jj, kk), &avg);
        crse(i, j, k, n) = avg;
    }
}

/** Kreiss-Oliger dissipation term of order 2 (4th difference) or 3 (6th
 *  difference) for component c, scaled by strength/dx. */
template <int>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double
KreissOligerDissipation(const amrex::Array4<double const> &state, const int i,
                        const int j, const int k, const int c, const double dx,
                        const double dissipation_strength);

template <>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double KreissOligerDissipation<2>(
    const amrex::Array4<double const> &state, const int i, const int j,
    const int k, const int c, const double dx,
    const double dissipation_strength) {
    // Binomial 4th difference (1, -4, 6, -4, 1) along each axis.
    auto d4 = [&](int ax, int ay, int az) {
        return state(i + 2 * ax, j + 2 * ay, k + 2 * az, c) -
               4. * state(i + ax, j + ay, k + az, c) +
               state(i - 2 * ax, j - 2 * ay, k - 2 * az, c) -
               4. * state(i - ax, j - ay, k - az, c) +
               6. * state(i, j, k, c);
    };
    const double s = d4(1, 0, 0) + d4(0, 1, 0) + d4(0, 0, 1);
    return -dissipation_strength * s / 16. / dx;
}

template <>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double KreissOligerDissipation<3>(
    const amrex::Array4<double const> &state, const int i, const int j,
    const int k, const int c, const double dx,
    const double dissipation_strength) {
    // Binomial 6th difference (1, -6, 15, -20, 15, -6, 1) along each axis.
    auto d6 = [&](int ax, int ay, int az) {
        return state(i + 3 * ax, j + 3 * ay, k + 3 * az, c) -
               6. * state(i + 2 * ax, j + 2 * ay, k + 2 * az, c) +
               15. * state(i + ax, j + ay, k + az, c) +
               state(i - 3 * ax, j - 3 * ay, k - 3 * az, c) -
               6. * state(i - 2 * ax, j - 2 * ay, k - 2 * az, c) +
               15. * state(i - ax, j - ay, k - az, c) -
               20. * state(i, j, k, c);
    };
    const double s = d6(1, 0, 0) + d6(0, 1, 0) + d6(0, 0, 1);
    return dissipation_strength * s / 64. / dx;
}

}  // namespace kernels
}  // namespace sledgehamr

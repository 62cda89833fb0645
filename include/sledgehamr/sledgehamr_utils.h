// Finite-difference helpers used by project cell functions.
//
// Reference interface: sledgehamr::utils::Laplacian<p> and Gradient<p>
// (source/utils/sledgehamr_utils.h:50-188).  Template index p = order/2:
// p = 1, 2, 3 give the 2nd, 4th and 6th order central stencils, p = 0 is
// identically zero.  Signatures are unchanged so project headers compile as
// they are; the bodies are written as coefficient-ring sums.
#pragma once

#include "amrex_compat.h"

namespace sledgehamr {
namespace utils {

/** constexpr for-loop over an integer range, as used by the tag helpers. */
template <auto Start, auto End, auto Inc, class F>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE constexpr void constexpr_for(F &&f) {
    if constexpr (Start < End) {
        f(std::integral_constant<decltype(Start), Start>());
        constexpr_for<Start + Inc, End, Inc>(f);
    }
}

namespace detail {

/** Central second-derivative weights: lap = (w0 f0 + sum_m w_m (sum over
 *  6 neighbours at distance m)) / (den dx^2).  Rows are indexed by p. */
struct LapWeights {
    double w[4];
    double den;
};
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE constexpr LapWeights lap_weights(int p) {
    return p == 1   ? LapWeights{{-6., 1., 0., 0.}, 1.}
           : p == 2 ? LapWeights{{-90., 16., -1., 0.}, 12.}
           : p == 3 ? LapWeights{{-1470., 270., -27., 2.}, 180.}
                    : LapWeights{{0., 0., 0., 0.}, 1.};
}

/** Central first-derivative weights: grad = sum_m g_m (f(+m) - f(-m)) /
 *  (den dx). */
struct GradWeights {
    double g[4];
    double den;
};
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE constexpr GradWeights grad_weights(int p) {
    return p == 1   ? GradWeights{{0., 1., 0., 0.}, 2.}
           : p == 2 ? GradWeights{{0., 8., -1., 0.}, 12.}
           : p == 3 ? GradWeights{{0., 45., -9., 1.}, 60.}
                    : GradWeights{{0., 0., 0., 0.}, 1.};
}

}  // namespace detail

/** Ring sums on an FP32 tile (fused stage kernels): x/y neighbours from the
 *  staged LDS plane, z neighbours from the register queue, accumulated in
 *  FP64 (sums of FP32 values are then exact, so only the stored values
 *  carry rounding).  Returns false when the request is not the tile's own
 *  cell of a stencil slot, in which case the caller takes the generic path. */
template <int P>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE bool tile_lap_f32(
    amrex::Array4<amrex::Real const> const &state, int i, int j, int k, int c,
    double &sum) {
    if (state.mode != amrex::Array4<const double>::kTile) return false;
    const shm::TileCtx &tc = *static_cast<const shm::TileCtx *>(state.ext);
    if (!tc.f32 || i != tc.ci || j != tc.cj || k != tc.ck) return false;
    const int s = tc.slot[c];
    if (s < 0) return false;
    constexpr detail::LapWeights W = detail::lap_weights(P);
    const float *L = static_cast<const float *>(tc.lds) + s * tc.lds_cstride;
    // z neighbours: the register queue, or the LDS window's other planes.
    const float *Z = static_cast<const float *>(tc.zq);
    auto zv = [&](int m) -> double {
        return tc.zplane ? (double)L[tc.zplane[tc.R + m]] : (double)Z[tc.zq_at(s, m)];
    };
    double acc = W.w[0] * zv(0);
    constexpr_for<1, P + 1, 1>([&](auto m) {
        const double ring = ((double)L[m] + (double)L[-m]) +
                            ((double)L[m * tc.lds_pitch] + (double)L[-m * tc.lds_pitch]) +
                            (zv(m) + zv(-m));
        acc += W.w[m] * ring;
    });
    sum = acc;
    return true;
}

/** Laplacian of component c at cell (i,j,k) with squared spacing dx2. */
template <int P>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double
Laplacian(amrex::Array4<amrex::Real const> const &state, const int i,
          const int j, const int k, const int c, const double dx2) {
    static_assert(P >= 0 && P <= 3, "Laplacian order index must be 0..3");
    if constexpr (P == 0) {
        return 0.;
    } else {
        constexpr detail::LapWeights W = detail::lap_weights(P);
        const double scale = 1. / (W.den * dx2);  // loop invariant
        double fs;
        if (tile_lap_f32<P>(state, i, j, k, c, fs)) return fs * scale;
        double acc = W.w[0] * state(i, j, k, c);
        constexpr_for<1, P + 1, 1>([&](auto m) {
            const double ring = state(i + m, j, k, c) + state(i - m, j, k, c) +
                                state(i, j + m, k, c) + state(i, j - m, k, c) +
                                state(i, j, k + m, c) + state(i, j, k - m, c);
            acc += W.w[m] * ring;
        });
        return acc * scale;
    }
}

/** Central first derivative of component c along axis 'x', 'y' or 'z'. */
template <int P>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double
Gradient(amrex::Array4<amrex::Real const> const &state, const int i,
         const int j, const int k, const int c, const double dx,
         const char axis) {
    static_assert(P >= 0 && P <= 3, "Gradient order index must be 0..3");
    if constexpr (P == 0) {
        return 0.;
    } else {
        constexpr detail::GradWeights W = detail::grad_weights(P);
        const int ax = axis == 'x' ? 1 : 0, ay = axis == 'y' ? 1 : 0,
                  az = axis == 'z' ? 1 : 0;
        if (ax + ay + az == 0) return 0.;
        const double scale = 1. / (W.den * dx);  // loop invariant
        if (state.mode == amrex::Array4<const double>::kTile) {
            const shm::TileCtx &tc = *static_cast<const shm::TileCtx *>(state.ext);
            const int s = tc.slot[c];
            if (tc.f32 && s >= 0 && i == tc.ci && j == tc.cj && k == tc.ck) {
                // Differences of FP32 values are exact in FP64.
                double acc = 0.;
                if (az && tc.zplane) {
                    const float *L =
                        static_cast<const float *>(tc.lds) + s * tc.lds_cstride;
                    constexpr_for<1, P + 1, 1>([&](auto m) {
                        acc += W.g[m] * ((double)L[tc.zplane[tc.R + m]] -
                                         (double)L[tc.zplane[tc.R - m]]);
                    });
                } else if (az) {
                    const float *Z = static_cast<const float *>(tc.zq);
                    constexpr_for<1, P + 1, 1>([&](auto m) {
                        acc += W.g[m] * ((double)Z[tc.zq_at(s, m)] -
                                         (double)Z[tc.zq_at(s, -m)]);
                    });
                } else {
                    const float *L =
                        static_cast<const float *>(tc.lds) + s * tc.lds_cstride;
                    const int st = ax ? 1 : tc.lds_pitch;
                    constexpr_for<1, P + 1, 1>([&](auto m) {
                        acc += W.g[m] * ((double)L[m * st] - (double)L[-m * st]);
                    });
                }
                return acc * scale;
            }
        }
        double acc = 0.;
        constexpr_for<1, P + 1, 1>([&](auto m) {
            acc += W.g[m] * (state(i + m * ax, j + m * ay, k + m * az, c) -
                             state(i - m * ax, j - m * ay, k - m * az, c));
        });
        return acc * scale;
    }
}

/** Whether a value is a power of two (input validation helper). */
inline bool IsPowerOfTwo(int val) { return val > 0 && (val & (val - 1)) == 0; }

}  // namespace utils
}  // namespace sledgehamr

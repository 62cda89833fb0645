// Project API macros.
//
// Reference interface: source/macros.h.  The macro names, the entities they
// declare in the project namespace (enums _Scalar/_Momentum, namespace
// Scalar, enum Gw, l_scalar_fields, the TruncationModifier /
// TagCellForRefinement / GravitationalWavesRhs / GravitationalWavesBackreaction
// default templates, TruncationErrorTagCpu/Gpu) and their meaning are the
// same, so project headers written for the reference compile unchanged.
// Boost.Preprocessor is replaced by a local variadic FOR_EACH.  The cell
// loops that the reference generated inside the project class
// (macros.h:301-686) are replaced by the MI355X kernel table that
// SLEDGEHAMR_INITIALIZE_PROJECT instantiates for the project.
#pragma once

#include <vector>

#include "amrex_compat.h"
#include "kernels.h"
#include "scalars.h"
#include "sledgehamr_utils.h"

#define SHM_PP_CAT(a, b) SHM_PP_CAT_I(a, b)
#define SHM_PP_CAT_I(a, b) a##b
#define SHM_PP_FOR_EACH(M, ...) \
    SHM_PP_CAT(SHM_PP_FE_, SHM_PP_NARGS(__VA_ARGS__))(M, __VA_ARGS__)
#define SHM_PP_NARGS(...) SHM_PP_NARGS_I(__VA_ARGS__, 32,31,30,29,28,27,26,25,24,23,22,21,20,19,18,17,16,15,14,13,12,11,10,9,8,7,6,5,4,3,2,1)
#define SHM_PP_NARGS_I(_1,_2,_3,_4,_5,_6,_7,_8,_9,_10,_11,_12,_13,_14,_15,_16,_17,_18,_19,_20,_21,_22,_23,_24,_25,_26,_27,_28,_29,_30,_31,_32, N, ...) N
#define SHM_PP_FE_1(M, a) M(a)
#define SHM_PP_FE_2(M, a, ...) M(a) SHM_PP_FE_1(M, __VA_ARGS__)
#define SHM_PP_FE_3(M, a, ...) M(a) SHM_PP_FE_2(M, __VA_ARGS__)
#define SHM_PP_FE_4(M, a, ...) M(a) SHM_PP_FE_3(M, __VA_ARGS__)
#define SHM_PP_FE_5(M, a, ...) M(a) SHM_PP_FE_4(M, __VA_ARGS__)
#define SHM_PP_FE_6(M, a, ...) M(a) SHM_PP_FE_5(M, __VA_ARGS__)
#define SHM_PP_FE_7(M, a, ...) M(a) SHM_PP_FE_6(M, __VA_ARGS__)
#define SHM_PP_FE_8(M, a, ...) M(a) SHM_PP_FE_7(M, __VA_ARGS__)
#define SHM_PP_FE_9(M, a, ...) M(a) SHM_PP_FE_8(M, __VA_ARGS__)
#define SHM_PP_FE_10(M, a, ...) M(a) SHM_PP_FE_9(M, __VA_ARGS__)
#define SHM_PP_FE_11(M, a, ...) M(a) SHM_PP_FE_10(M, __VA_ARGS__)
#define SHM_PP_FE_12(M, a, ...) M(a) SHM_PP_FE_11(M, __VA_ARGS__)
#define SHM_PP_FE_13(M, a, ...) M(a) SHM_PP_FE_12(M, __VA_ARGS__)
#define SHM_PP_FE_14(M, a, ...) M(a) SHM_PP_FE_13(M, __VA_ARGS__)
#define SHM_PP_FE_15(M, a, ...) M(a) SHM_PP_FE_14(M, __VA_ARGS__)
#define SHM_PP_FE_16(M, a, ...) M(a) SHM_PP_FE_15(M, __VA_ARGS__)
#define SHM_PP_FE_17(M, a, ...) M(a) SHM_PP_FE_16(M, __VA_ARGS__)
#define SHM_PP_FE_18(M, a, ...) M(a) SHM_PP_FE_17(M, __VA_ARGS__)
#define SHM_PP_FE_19(M, a, ...) M(a) SHM_PP_FE_18(M, __VA_ARGS__)
#define SHM_PP_FE_20(M, a, ...) M(a) SHM_PP_FE_19(M, __VA_ARGS__)
#define SHM_PP_FE_21(M, a, ...) M(a) SHM_PP_FE_20(M, __VA_ARGS__)
#define SHM_PP_FE_22(M, a, ...) M(a) SHM_PP_FE_21(M, __VA_ARGS__)
#define SHM_PP_FE_23(M, a, ...) M(a) SHM_PP_FE_22(M, __VA_ARGS__)
#define SHM_PP_FE_24(M, a, ...) M(a) SHM_PP_FE_23(M, __VA_ARGS__)
#define SHM_PP_FE_25(M, a, ...) M(a) SHM_PP_FE_24(M, __VA_ARGS__)
#define SHM_PP_FE_26(M, a, ...) M(a) SHM_PP_FE_25(M, __VA_ARGS__)
#define SHM_PP_FE_27(M, a, ...) M(a) SHM_PP_FE_26(M, __VA_ARGS__)
#define SHM_PP_FE_28(M, a, ...) M(a) SHM_PP_FE_27(M, __VA_ARGS__)
#define SHM_PP_FE_29(M, a, ...) M(a) SHM_PP_FE_28(M, __VA_ARGS__)
#define SHM_PP_FE_30(M, a, ...) M(a) SHM_PP_FE_29(M, __VA_ARGS__)
#define SHM_PP_FE_31(M, a, ...) M(a) SHM_PP_FE_30(M, __VA_ARGS__)
#define SHM_PP_FE_32(M, a, ...) M(a) SHM_PP_FE_31(M, __VA_ARGS__)

#define SLEDGEHAMR_DO_PRAGMA(x) _Pragma(#x)

/** One scalar field: static registration plus Scalar::<name> id. */
#define SLEDGEHAMR_ADD_SCALAR(field)                                           \
    static sledgehamr::ScalarField SHM_PP_CAT(_s_, field) = {                  \
        #field, l_scalar_fields, false};                                       \
    namespace Scalar {                                                         \
    constexpr int field = _Scalar::field;                                      \
    }

/** One conjugate momentum: ids follow all scalar fields. */
#define SLEDGEHAMR_ADD_MOMENTUM(field)                                         \
    static sledgehamr::ScalarField SHM_PP_CAT(_s_, field) = {                  \
        #field, l_scalar_fields, true};                                        \
    namespace Scalar {                                                         \
    constexpr int field = _Momentum::field + _Scalar::NScalarFields;           \
    }

#define SHM_ENUM_ENTRY(x) x,

#define SLEDGEHAMR_SCALAR_ENUM(...)                                            \
    enum _Scalar { SHM_PP_FOR_EACH(SHM_ENUM_ENTRY, __VA_ARGS__) NScalarFields };

#define SLEDGEHAMR_MOMENTUM_ENUM(...)                                          \
    enum _Momentum {                                                           \
        SHM_PP_FOR_EACH(SHM_ENUM_ENTRY, __VA_ARGS__) NMomentumFields           \
    };

/** Gravitational-wave tensor components follow the project's scalars. */
#define SLEDGEHAMR_GW_ENUM                                                     \
    enum Gw {                                                                  \
        u_xx = Scalar::NScalars,                                               \
        u_yy,                                                                  \
        u_zz,                                                                  \
        u_xy,                                                                  \
        u_xz,                                                                  \
        u_yz,                                                                  \
        du_xx,                                                                 \
        du_yy,                                                                 \
        du_zz,                                                                 \
        du_xy,                                                                 \
        du_xz,                                                                 \
        du_yz,                                                                 \
        NGwScalars                                                             \
    };

/** Default truncation-error modifier f(te) = te (specialisable per comp). */
#define SLEDGEHAMR_TRUNCATION_MODIFIER                                         \
    template <int>                                                             \
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double TruncationModifier(        \
        const amrex::Array4<const double> &state, const int i, const int j,    \
        const int k, const int lev, const double time, const double dt,        \
        const double dx, const double truncation_error,                        \
        const double *params) {                                                \
        return truncation_error;                                               \
    }

/** Default user tagging criterion: off. */
#define SLEDGEHAMR_TAG_CELL_FOR_REFINEMENT                                     \
    template <bool>                                                            \
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE bool TagCellForRefinement(        \
        const amrex::Array4<const double> &state, const int i, const int j,    \
        const int k, const int lev, const double time, const double dt,        \
        const double dx, const double *params) {                               \
        return false;                                                          \
    }

/** Default GW source: none.  Device-only, like the projects' specialisations
 *  (projects/FirstOrderPhaseTransition/kernels_rhs.h:56-57). */
#define SLEDGEHAMR_GRAVITATIONAL_WAVES_RHS                                     \
    template <bool>                                                            \
    AMREX_GPU_DEVICE AMREX_FORCE_INLINE void GravitationalWavesRhs(            \
        const amrex::Array4<double> &rhs,                                      \
        const amrex::Array4<const double> &state, const int i, const int j,    \
        const int k, const int lev, const double time, const double dt,        \
        const double dx, const double *params) {}

/** Default GW backreaction onto the scalars: none. */
#define SLEDGEHAMR_GRAVITATIONAL_WAVES_BACKREACTION                            \
    template <bool>                                                            \
    AMREX_GPU_DEVICE AMREX_FORCE_INLINE void                                   \
    GravitationalWavesBackreaction(                                            \
        const amrex::Array4<double> &rhs,                                      \
        const amrex::Array4<const double> &state, const int i, const int j,    \
        const int k, const int lev, const double time, const double dt,        \
        const double dx, const double *params_scalars,                         \
        const double *params_gw) {}

#define SLEDGEHAMR_ADD_SCALARS(...)                                            \
    SLEDGEHAMR_SCALAR_ENUM(__VA_ARGS__)                                        \
    static std::vector<sledgehamr::ScalarField *> l_scalar_fields;             \
    SHM_PP_FOR_EACH(SLEDGEHAMR_ADD_SCALAR, __VA_ARGS__)

#define SLEDGEHAMR_ADD_CONJUGATE_MOMENTA(...)                                  \
    SLEDGEHAMR_MOMENTUM_ENUM(__VA_ARGS__)                                      \
    SHM_PP_FOR_EACH(SLEDGEHAMR_ADD_MOMENTUM, __VA_ARGS__)                      \
    namespace Scalar {                                                         \
    constexpr int NScalars =                                                   \
        _Scalar::NScalarFields + _Momentum::NMomentumFields;                   \
    }                                                                          \
    SLEDGEHAMR_GW_ENUM                                                         \
    SLEDGEHAMR_TRUNCATION_MODIFIER                                             \
    SLEDGEHAMR_TAG_CELL_FOR_REFINEMENT                                         \
    SLEDGEHAMR_GRAVITATIONAL_WAVES_RHS                                         \
    SLEDGEHAMR_GRAVITATIONAL_WAVES_BACKREACTION

/** TE criterion of one cell over NScalars components, counting tags per
 *  component (host loops).  Only all-even cells carry a truncation error. */
#define SLEDGEHAMR_TRUNCATION_ERROR_TAG_CPU                                    \
    template <int NScalars>                                                    \
    AMREX_FORCE_INLINE bool TruncationErrorTagCpu(                             \
        const amrex::Array4<const double> &state,                              \
        const amrex::Array4<const double> &te, const int i, const int j,       \
        const int k, const int lev, const double time, const double dt,        \
        const double dx, std::vector<double> &te_crit, long *ntags_trunc,      \
        const std::vector<double> &params) {                                   \
        if ((i & 1) || (j & 1) || (k & 1)) return false;                       \
        bool res = false;                                                      \
        sledgehamr::utils::constexpr_for<0, NScalars, 1>([&](auto n) {         \
            const double m = TruncationModifier<n>(state, i, j, k, lev, time,  \
                                                   dt, dx, te(i, j, k, n),     \
                                                   params.data());             \
            if (m >= te_crit[n]) {                                             \
                res = true;                                                    \
                ntags_trunc[n] += 8;                                           \
            }                                                                  \
        });                                                                    \
        return res;                                                            \
    }

/** Same criterion for device code (no counters). */
#define SLEDGEHAMR_TRUNCATION_ERROR_TAG_GPU                                    \
    template <int NScalars>                                                    \
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE bool TruncationErrorTagGpu(       \
        const amrex::Array4<double const> &state,                              \
        const amrex::Array4<double const> &te, const int i, const int j,       \
        const int k, const int lev, const double time, const double dt,        \
        const double dx, const double *te_crit, const double *params) {        \
        if ((i & 1) || (j & 1) || (k & 1)) return false;                       \
        bool res = false;                                                      \
        sledgehamr::utils::constexpr_for<0, NScalars, 1>([&](auto n) {         \
            if (TruncationModifier<n>(state, i, j, k, lev, time, dt, dx,       \
                                      te(i, j, k, n), params) >= te_crit[n])   \
                res = true;                                                    \
        });                                                                    \
        return res;                                                            \
    }

#define SLEDGEHAMR_FINISH_SETUP                                                \
    SLEDGEHAMR_TRUNCATION_ERROR_TAG_CPU                                        \
    SLEDGEHAMR_TRUNCATION_ERROR_TAG_GPU

/** Output cell functions of a project, for spectra and projections (what
 *  io_module->spectra / io_module->projections register in the reference,
 *  output_types/spectrum.cpp:12-113, output_types/projection.cpp:11-118).
 *  Each is double f(state, i, j, k, lev, time, dt, dx, params) with params
 *  as const double * (SetParamsSpectra / SetParamsProjections), declared
 *  AMREX_GPU_HOST_DEVICE so that it runs on the device.  A function's id is
 *  its position in the list.  Place after the functions, before the project
 *  class. */
#define SHM_OUT_NAME(f) #f,
#define SHM_OUT_CASE(f) \
    if (id == c++) return f(s, i, j, k, lev, t, dt, dx, p);
#define SLEDGEHAMR_OUTPUT_FUNCTIONS(...)                                       \
    struct ShmOutputFns {                                                      \
        static constexpr int count = SHM_PP_NARGS(__VA_ARGS__);                \
        static const char *name(int id) {                                      \
            static const char *const n[] = {                                   \
                SHM_PP_FOR_EACH(SHM_OUT_NAME, __VA_ARGS__)};                   \
            return id >= 0 && id < count ? n[id] : nullptr;                    \
        }                                                                      \
        AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE static double                 \
        eval(int id, const amrex::Array4<const double> &s, int i, int j,      \
             int k,                                                            \
             int lev, double t, double dt, double dx, const double *p) {       \
            int c = 0;                                                         \
            SHM_PP_FOR_EACH(SHM_OUT_CASE, __VA_ARGS__)                         \
            return 0.;                                                         \
        }                                                                      \
    };                                                                         \
    inline ShmOutputFns shm_select_output(int) { return {}; }

/** Projects without SLEDGEHAMR_OUTPUT_FUNCTIONS have no output functions. */
namespace shm {
struct NoOutputFns {
    static constexpr int count = 0;
    static const char *name(int) { return nullptr; }
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE static double
    eval(int, const amrex::Array4<const double> &, int, int, int, int, double,
         double, double, const double *) {
        return 0.;
    }
};
}  // namespace shm
inline shm::NoOutputFns shm_select_output(...) { return {}; }

/** Gravitational-wave spectrum modifiers of a project (the reference's
 *  GravitationalWavesSpectrumModifier, gravitational_waves.h:79-92).  Each
 *  listed struct M provides
 *    static void SelectComponents(int components[6]);   // host
 *    AMREX_GPU_HOST_DEVICE static void FourierSpaceModification(
 *        double re[6], double im[6], int kx, int ky, int kz, double dk,
 *        int dimN, const double *params);
 *  SelectComponents names the six tensor components transformed, in the
 *  order xx, xy, xz, yy, yz, zz (default: du_xx, du_xy, du_xz, du_yy, du_yz,
 *  du_zz).  FourierSpaceModification is the reference's
 *  FourierSpaceModifications applied to one mode of the x-halved R2C
 *  spectrum, where the reference loops over the spectral MultiFabs: the six
 *  amplitudes of the mode, its signed wave numbers (k = a >= dimN/2 ? a -
 *  dimN : a per axis, as the reference's modifiers form them), dk = 2 pi /
 *  L of the (zero-padded) transform, and per-call parameters (the
 *  reference's modifier constructor arguments).  A modifier's id is its
 *  position in the list.  Place after the structs, before the project
 *  class. */
#define SHM_GWM_NAME(m) #m,
#define SHM_GWM_SELECT(m) \
    if (id == c++) {      \
        m::SelectComponents(comps); \
        return true;      \
    }
#define SHM_GWM_MODIFY(m) \
    if (id == c++) {      \
        m::FourierSpaceModification(re, im, kx, ky, kz, dk, dimN, p); \
        return;           \
    }
#define SLEDGEHAMR_GW_SPECTRUM_MODIFIERS(...)                                  \
    struct ShmGwModifiers {                                                    \
        static constexpr int count = SHM_PP_NARGS(__VA_ARGS__);                \
        static const char *name(int id) {                                      \
            static const char *const n[] = {                                   \
                SHM_PP_FOR_EACH(SHM_GWM_NAME, __VA_ARGS__)};                   \
            return id >= 0 && id < count ? n[id] : nullptr;                    \
        }                                                                      \
        static bool select(int id, int comps[6]) {                             \
            int c = 0;                                                         \
            SHM_PP_FOR_EACH(SHM_GWM_SELECT, __VA_ARGS__)                       \
            return false;                                                      \
        }                                                                      \
        AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE static void                   \
        modify(int id, double re[6], double im[6], int kx, int ky, int kz,     \
               double dk, int dimN, const double *p) {                         \
            int c = 0;                                                         \
            SHM_PP_FOR_EACH(SHM_GWM_MODIFY, __VA_ARGS__)                       \
        }                                                                      \
    };                                                                         \
    inline ShmGwModifiers shm_select_gw_modifiers(int) { return {}; }

/** Projects without SLEDGEHAMR_GW_SPECTRUM_MODIFIERS have only the default
 *  (du_ij, unmodified). */
namespace shm {
struct NoGwModifiers {
    static constexpr int count = 0;
    static const char *name(int) { return nullptr; }
    static bool select(int, int *) { return false; }
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE static void modify(int, double *, double *,
                                                                int, int, int, double,
                                                                int, const double *) {}
};
}  // namespace shm
inline shm::NoGwModifiers shm_select_gw_modifiers(...) { return {}; }

/** Binds the project's device functions to the MI355X kernel table.
 *  ShmOps is the compile-time description of the project that the fused
 *  stage, tagging and probe kernels are instantiated with. */
#define SLEDGEHAMR_PRJ_OPS                                                     \
    struct ShmOps {                                                            \
        static constexpr int kNScalars = Scalar::NScalars;                     \
        static constexpr int kNFields = _Scalar::NScalarFields;                \
        static constexpr int kNGw = Gw::NGwScalars;                            \
        using Output = decltype(shm_select_output(0));                         \
        using GwModifiers = decltype(shm_select_gw_modifiers(0));              \
        AMREX_GPU_DEVICE AMREX_FORCE_INLINE static void                        \
        rhs(const amrex::Array4<double> &r,                                    \
            const amrex::Array4<const double> &s, int i, int j, int k,         \
            int lev, double t, double dt, double dx, const double *p) {        \
            Rhs(r, s, i, j, k, lev, t, dt, dx, p);                             \
        }                                                                      \
        AMREX_GPU_DEVICE AMREX_FORCE_INLINE static void                        \
        gw_rhs(const amrex::Array4<double> &r,                                 \
               const amrex::Array4<const double> &s, int i, int j, int k,      \
               int lev, double t, double dt, double dx, const double *p) {     \
            GravitationalWavesRhs<true>(r, s, i, j, k, lev, t, dt, dx, p);     \
        }                                                                      \
        AMREX_GPU_DEVICE AMREX_FORCE_INLINE static void                        \
        gw_backreaction(const amrex::Array4<double> &r,                        \
                        const amrex::Array4<const double> &s, int i, int j,    \
                        int k, int lev, double t, double dt, double dx,        \
                        const double *ps, const double *pg) {                  \
            GravitationalWavesBackreaction<true>(r, s, i, j, k, lev, t, dt,    \
                                                 dx, ps, pg);                  \
        }                                                                      \
        AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE static bool                   \
        tag_cell(const amrex::Array4<const double> &s, int i, int j, int k,    \
                 int lev, double t, double dt, double dx, const double *p) {   \
            return TagCellForRefinement<true>(s, i, j, k, lev, t, dt, dx, p);  \
        }                                                                      \
        template <int N>                                                       \
        AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE static double                 \
        te_modifier(const amrex::Array4<const double> &s, int i, int j, int k, \
                    int lev, double t, double dt, double dx, double te,        \
                    const double *p) {                                         \
            return TruncationModifier<N>(s, i, j, k, lev, t, dt, dx, te, p);   \
        }                                                                      \
    };

#define SLEDGEHAMR_PRJ_CONSTRUCTOR(prj)                                        \
    prj() {                                                                    \
        scalar_fields = l_scalar_fields;                                       \
        project_name = #prj;                                                   \
    }

/** Adds the generated members to the project class. */
#define SLEDGEHAMR_INITIALIZE_PROJECT(prj)                                     \
    SLEDGEHAMR_PRJ_OPS                                                         \
    SLEDGEHAMR_PRJ_CONSTRUCTOR(prj)                                            \
    const shm::KernelTable &ShmKernels() const override {                      \
        return shm::project_kernel_table<ShmOps>();                            \
    }

// ORACLE (test infrastructure only; never linked into the product).
//
// Minimal host-only restatement of the sledgehamr project API so that project
// headers compile with g++ for the FP64 CPU oracle.  Follows
// source/macros.h:39-249 (registration macros and default templates),
// source/utils/sledgehamr_utils.h:50-188 (Laplacian/Gradient, same expression
// order) and source/kernels.h:56-106 (Kreiss-Oliger).
#pragma once

#include <cmath>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <functional>
#include <vector>

#define AMREX_GPU_DEVICE
#define AMREX_GPU_HOST
#define AMREX_GPU_HOST_DEVICE
#define AMREX_FORCE_INLINE inline
#define AMREX_PRAGMA_SIMD

namespace amrex {
using Real = double;
using Long = long long;
struct Dim3 { int x, y, z; };

/** Strided view: x fastest, then y, z, component. */
template <class T>
struct Array4 {
    T *p = nullptr;
    Long jstride = 0, kstride = 0, nstride = 0;
    Dim3 begin{0, 0, 0};
    T &operator()(int i, int j, int k, int n = 0) const {
        return p[(i - begin.x) + (j - begin.y) * jstride +
                 (k - begin.z) * kstride + n * nstride];
    }
};

class ParmParse {
  public:
    explicit ParmParse(const std::string &prefix = "") : pre_(prefix) {}
    static std::map<std::string, std::vector<std::string>> &table() {
        static std::map<std::string, std::vector<std::string>> t;
        return t;
    }
    template <class T>
    bool query(const char *name, T &v) const {
        auto it = table().find(pre_.empty() ? name : pre_ + "." + name);
        if (it == table().end() || it->second.empty()) return false;
        std::istringstream s(it->second[0]);
        if constexpr (std::is_same_v<T, bool>) {
            const std::string &x = it->second[0];
            v = (x == "1" || x == "true" || x == "True");
        } else {
            s >> v;
        }
        return true;
    }
    template <class T>
    void get(const char *name, T &v) const {
        if (!query(name, v)) throw std::runtime_error(std::string("missing ") + name);
    }
    template <class T>
    bool queryarr(const char *name, std::vector<T> &v) const {
        auto it = table().find(pre_.empty() ? name : pre_ + "." + name);
        if (it == table().end()) return false;
        v.clear();
        for (auto &x : it->second) {
            std::istringstream s(x);
            T t;
            s >> t;
            v.push_back(t);
        }
        return true;
    }
    template <class T>
    void getarr(const char *name, std::vector<T> &v) const {
        if (!queryarr(name, v)) throw std::runtime_error(std::string("missing ") + name);
    }

  private:
    std::string pre_;
};
}  // namespace amrex

namespace sledgehamr {

class ScalarField {
  public:
    ScalarField(std::string s, std::vector<ScalarField *> &v, bool mom)
        : name(std::move(s)), id((int)v.size()), is_conjugate_momentum(mom) {
        v.push_back(this);
    }
    operator int() const { return id; }
    std::string name;
    const int id;
    bool is_conjugate_momentum;
};

class Sledgehamr {
  public:
    virtual ~Sledgehamr() = default;
    virtual void Init() {}
    virtual void SetParamsRhs(std::vector<double> &, const double, const int) {}
    virtual void SetParamsGravitationalWaveRhs(std::vector<double> &, const double,
                                               const int) {}
    virtual void SetParamsTagCellForRefinement(std::vector<double> &, const double,
                                               const int) {}
    virtual void SetParamsTruncationModifier(std::vector<double> &, const double,
                                             const int) {}
    virtual void SetParamsSpectra(std::vector<double> &, const double) {}
    virtual void SetParamsProjections(std::vector<double> &, const double) {}
    virtual bool CreateLevelIf(const int, const double) { return true; }
    virtual void BeforeTimestep(const double) {}
    /** grid_new[lev].norm0 over every component (set by the oracle). */
    std::vector<double> LevelNorm0(const int lev) const {
        return level_norm0_ ? level_norm0_(lev) : std::vector<double>();
    }
    std::vector<ScalarField *> scalar_fields;
    std::function<std::vector<double>(int)> level_norm0_;
};

namespace utils {

template <auto Start, auto End, auto Inc, class F>
constexpr void constexpr_for(F &&f) {
    if constexpr (Start < End) {
        f(std::integral_constant<decltype(Start), Start>());
        constexpr_for<Start + Inc, End, Inc>(f);
    }
}

// Laplacian<p>: sledgehamr_utils.h:50-94, same grouping of terms.
template <int>
double Laplacian(amrex::Array4<amrex::Real const> const &s, int i, int j, int k,
                 int c, double dx2);
template <>
inline double Laplacian<0>(amrex::Array4<amrex::Real const> const &, int, int, int,
                           int, double) {
    return 0;
}
template <>
inline double Laplacian<1>(amrex::Array4<amrex::Real const> const &s, int i, int j,
                           int k, int c, double dx2) {
    return (s(i + 1, j, k, c) + s(i - 1, j, k, c) + s(i, j + 1, k, c) +
            s(i, j - 1, k, c) + s(i, j, k + 1, c) + s(i, j, k - 1, c) -
            6. * s(i, j, k, c)) / dx2;
}
template <>
inline double Laplacian<2>(amrex::Array4<amrex::Real const> const &s, int i, int j,
                           int k, int c, double dx2) {
    return (-(s(i + 2, j, k, c) + s(i - 2, j, k, c) + s(i, j + 2, k, c) +
              s(i, j - 2, k, c) + s(i, j, k + 2, c) + s(i, j, k - 2, c)) +
            16. * (s(i + 1, j, k, c) + s(i - 1, j, k, c) + s(i, j + 1, k, c) +
                   s(i, j - 1, k, c) + s(i, j, k + 1, c) + s(i, j, k - 1, c)) -
            90. * s(i, j, k, c)) / (12. * dx2);
}
template <>
inline double Laplacian<3>(amrex::Array4<amrex::Real const> const &s, int i, int j,
                           int k, int c, double dx2) {
    return (2. * (s(i + 3, j, k, c) + s(i - 3, j, k, c) + s(i, j + 3, k, c) +
                  s(i, j - 3, k, c) + s(i, j, k + 3, c) + s(i, j, k - 3, c)) -
            27. * (s(i + 2, j, k, c) + s(i - 2, j, k, c) + s(i, j + 2, k, c) +
                   s(i, j - 2, k, c) + s(i, j, k + 2, c) + s(i, j, k - 2, c)) +
            270. * (s(i + 1, j, k, c) + s(i - 1, j, k, c) + s(i, j + 1, k, c) +
                    s(i, j - 1, k, c) + s(i, j, k + 1, c) + s(i, j, k - 1, c)) -
            1470. * s(i, j, k, c)) / (180. * dx2);
}

// Gradient<p>: sledgehamr_utils.h:106-188.
template <int P>
inline double Gradient(amrex::Array4<amrex::Real const> const &s, int i, int j,
                       int k, int c, double dx, char axis) {
    const int a = axis == 'x', b = axis == 'y', d = axis == 'z';
    auto f = [&](int m) { return s(i + m * a, j + m * b, k + m * d, c); };
    if (P == 1) return (f(1) - f(-1)) / (2. * dx);
    if (P == 2) return (-f(2) + 8. * f(1) + f(-2) - 8. * f(-1)) / (12. * dx);
    if (P == 3)
        return (+f(3) - f(-3) - 9. * f(2) + 9. * f(-2) + 45. * f(1) - 45. * f(-1)) /
               (60. * dx);
    return 0;
}

}  // namespace utils

namespace kernels {
// KreissOligerDissipation<2,3>: kernels.h:56-106.
template <int>
double KreissOligerDissipation(const amrex::Array4<double const> &, int, int, int,
                               int, double, double);
template <>
inline double KreissOligerDissipation<2>(const amrex::Array4<double const> &s, int i,
                                         int j, int k, int c, double dx, double st) {
    double dx4 = s(i + 2, j, k, c) - 4. * s(i + 1, j, k, c) + s(i - 2, j, k, c) -
                 4. * s(i - 1, j, k, c) + 6. * s(i, j, k, c);
    double dy4 = s(i, j + 2, k, c) - 4. * s(i, j + 1, k, c) + s(i, j - 2, k, c) -
                 4. * s(i, j - 1, k, c) + 6. * s(i, j, k, c);
    double dz4 = s(i, j, k + 2, c) - 4. * s(i, j, k + 1, c) + s(i, j, k - 2, c) -
                 4. * s(i, j, k - 1, c) + 6. * s(i, j, k, c);
    return -st * (dx4 + dy4 + dz4) / 16. / dx;
}
template <>
inline double KreissOligerDissipation<3>(const amrex::Array4<double const> &s, int i,
                                         int j, int k, int c, double dx, double st) {
    auto d6 = [&](int a, int b, int d) {
        return s(i + 3 * a, j + 3 * b, k + 3 * d, c) -
               6. * s(i + 2 * a, j + 2 * b, k + 2 * d, c) +
               15. * s(i + a, j + b, k + d, c) + s(i - 3 * a, j - 3 * b, k - 3 * d, c) -
               6. * s(i - 2 * a, j - 2 * b, k - 2 * d, c) +
               15. * s(i - a, j - b, k - d, c) - 20. * s(i, j, k, c);
    };
    return st * (d6(1, 0, 0) + d6(0, 1, 0) + d6(0, 0, 1)) / 64. / dx;
}
}  // namespace kernels
}  // namespace sledgehamr

#define ORC_CAT(a, b) ORC_CAT_I(a, b)
#define ORC_CAT_I(a, b) a##b
#define ORC_NARGS(...) ORC_NARGS_I(__VA_ARGS__, 8, 7, 6, 5, 4, 3, 2, 1)
#define ORC_NARGS_I(_1, _2, _3, _4, _5, _6, _7, _8, N, ...) N
#define ORC_FE(M, ...) ORC_CAT(ORC_FE_, ORC_NARGS(__VA_ARGS__))(M, __VA_ARGS__)
#define ORC_FE_1(M, a) M(a)
#define ORC_FE_2(M, a, ...) M(a) ORC_FE_1(M, __VA_ARGS__)
#define ORC_FE_3(M, a, ...) M(a) ORC_FE_2(M, __VA_ARGS__)
#define ORC_FE_4(M, a, ...) M(a) ORC_FE_3(M, __VA_ARGS__)
#define ORC_FE_5(M, a, ...) M(a) ORC_FE_4(M, __VA_ARGS__)
#define ORC_FE_6(M, a, ...) M(a) ORC_FE_5(M, __VA_ARGS__)
#define ORC_FE_7(M, a, ...) M(a) ORC_FE_6(M, __VA_ARGS__)
#define ORC_FE_8(M, a, ...) M(a) ORC_FE_7(M, __VA_ARGS__)
#define ORC_ENUM(x) x,
#define ORC_ADD_S(f)                                                           \
    static sledgehamr::ScalarField ORC_CAT(_s_, f) = {#f, l_scalar_fields, false}; \
    namespace Scalar { constexpr int f = _Scalar::f; }
#define ORC_ADD_M(f)                                                           \
    static sledgehamr::ScalarField ORC_CAT(_s_, f) = {#f, l_scalar_fields, true}; \
    namespace Scalar { constexpr int f = _Momentum::f + _Scalar::NScalarFields; }

#define SLEDGEHAMR_ADD_SCALARS(...)                                            \
    enum _Scalar { ORC_FE(ORC_ENUM, __VA_ARGS__) NScalarFields };              \
    static std::vector<sledgehamr::ScalarField *> l_scalar_fields;             \
    ORC_FE(ORC_ADD_S, __VA_ARGS__)

#define SLEDGEHAMR_ADD_CONJUGATE_MOMENTA(...)                                  \
    enum _Momentum { ORC_FE(ORC_ENUM, __VA_ARGS__) NMomentumFields };          \
    ORC_FE(ORC_ADD_M, __VA_ARGS__)                                             \
    namespace Scalar {                                                         \
    constexpr int NScalars = _Scalar::NScalarFields + _Momentum::NMomentumFields; \
    }                                                                          \
    enum Gw { u_xx = Scalar::NScalars, u_yy, u_zz, u_xy, u_xz, u_yz,           \
              du_xx, du_yy, du_zz, du_xy, du_xz, du_yz, NGwScalars };          \
    template <int>                                                             \
    inline double TruncationModifier(const amrex::Array4<const double> &, int, \
                                     int, int, int, double, double, double,    \
                                     double te, const double *) {              \
        return te;                                                             \
    }                                                                          \
    template <bool>                                                            \
    inline bool TagCellForRefinement(const amrex::Array4<const double> &, int, \
                                     int, int, int, double, double, double,    \
                                     const double *) {                         \
        return false;                                                          \
    }                                                                          \
    template <bool>                                                            \
    inline void GravitationalWavesRhs(const amrex::Array4<double> &,           \
                                      const amrex::Array4<const double> &,     \
                                      int, int, int, int, double, double,      \
                                      double, const double *) {}               \
    template <bool>                                                            \
    inline void GravitationalWavesBackreaction(                                \
        const amrex::Array4<double> &, const amrex::Array4<const double> &,    \
        int, int, int, int, double, double, double, const double *,            \
        const double *) {}

#define SLEDGEHAMR_FINISH_SETUP

/** Output cell functions for spectra and projections (the product header's
 *  SLEDGEHAMR_OUTPUT_FUNCTIONS); a function's id is its list position. */
#define ORC_OUT_CASE(f) \
    if (id == c++) return f(s, i, j, k, lev, t, dt, dx, p);
#define SLEDGEHAMR_OUTPUT_FUNCTIONS(...)                                       \
    struct OrcOutputFns {                                                      \
        static constexpr int count = ORC_NARGS(__VA_ARGS__);                   \
        static double eval(int id, const amrex::Array4<const double> &s,       \
                           int i, int j, int k, int lev, double t, double dt,  \
                           double dx, const double *p) {                       \
            int c = 0;                                                         \
            ORC_FE(ORC_OUT_CASE, __VA_ARGS__)                                  \
            return 0.;                                                         \
        }                                                                      \
    };                                                                         \
    inline OrcOutputFns orc_select_output(int) { return {}; }

namespace orc {
struct NoOutputFns {
    static constexpr int count = 0;
    static double eval(int, const amrex::Array4<const double> &, int, int, int,
                       int, double, double, double, const double *) {
        return 0.;
    }
};
}  // namespace orc
inline orc::NoOutputFns orc_select_output(...) { return {}; }

/** GW spectrum modifiers (the product header's
 *  SLEDGEHAMR_GW_SPECTRUM_MODIFIERS): SelectComponents and the per-mode
 *  FourierSpaceModification of each listed struct, by list position. */
#define ORC_GWM_SELECT(m) \
    if (id == c++) {      \
        m::SelectComponents(comps); \
        return true;      \
    }
#define ORC_GWM_MODIFY(m) \
    if (id == c++) {      \
        m::FourierSpaceModification(re, im, kx, ky, kz, dk, dimN, p); \
        return;           \
    }
#define SLEDGEHAMR_GW_SPECTRUM_MODIFIERS(...)                                  \
    struct OrcGwModifiers {                                                    \
        static constexpr int count = ORC_NARGS(__VA_ARGS__);                   \
        static bool select(int id, int comps[6]) {                             \
            int c = 0;                                                         \
            ORC_FE(ORC_GWM_SELECT, __VA_ARGS__)                                \
            return false;                                                      \
        }                                                                      \
        static void modify(int id, double re[6], double im[6], int kx, int ky, \
                           int kz, double dk, int dimN, const double *p) {     \
            int c = 0;                                                         \
            ORC_FE(ORC_GWM_MODIFY, __VA_ARGS__)                                \
        }                                                                      \
    };                                                                         \
    inline OrcGwModifiers orc_select_gw_modifiers(int) { return {}; }

namespace orc {
struct NoGwModifiers {
    static constexpr int count = 0;
    static bool select(int, int *) { return false; }
    static void modify(int, double *, double *, int, int, int, double, int,
                       const double *) {}
};
}  // namespace orc
inline orc::NoGwModifiers orc_select_gw_modifiers(...) { return {}; }

/** The oracle reaches the project's cell functions through OrcOps. */
#define SLEDGEHAMR_INITIALIZE_PROJECT(prj)                                     \
    struct OrcOps {                                                            \
        static constexpr int kNScalars = Scalar::NScalars;                     \
        static constexpr int kNFields = _Scalar::NScalarFields;                \
        static constexpr int kNGw = Gw::NGwScalars;                            \
        using Output = decltype(orc_select_output(0));                         \
        using GwModifiers = decltype(orc_select_gw_modifiers(0));              \
        static void rhs(const amrex::Array4<double> &r,                        \
                        const amrex::Array4<const double> &s, int i, int j,    \
                        int k, int lev, double t, double dt, double dx,        \
                        const double *p) {                                     \
            Rhs(r, s, i, j, k, lev, t, dt, dx, p);                             \
        }                                                                      \
        static void gw_rhs(const amrex::Array4<double> &r,                     \
                           const amrex::Array4<const double> &s, int i, int j, \
                           int k, int lev, double t, double dt, double dx,     \
                           const double *p) {                                  \
            GravitationalWavesRhs<true>(r, s, i, j, k, lev, t, dt, dx, p);     \
        }                                                                      \
        static void gw_back(const amrex::Array4<double> &r,                    \
                            const amrex::Array4<const double> &s, int i,       \
                            int j, int k, int lev, double t, double dt,        \
                            double dx, const double *p, const double *q) {     \
            GravitationalWavesBackreaction<true>(r, s, i, j, k, lev, t, dt,    \
                                                 dx, p, q);                    \
        }                                                                      \
        static bool tag(const amrex::Array4<const double> &s, int i, int j,    \
                        int k, int lev, double t, double dt, double dx,        \
                        const double *p) {                                     \
            return TagCellForRefinement<true>(s, i, j, k, lev, t, dt, dx, p);  \
        }                                                                      \
        template <int N>                                                       \
        static double te_mod(const amrex::Array4<const double> &s, int i,      \
                             int j, int k, int lev, double t, double dt,       \
                             double dx, double te, const double *p) {          \
            return TruncationModifier<N>(s, i, j, k, lev, t, dt, dx, te, p);   \
        }                                                                      \
    };                                                                         \
    prj() { scalar_fields = l_scalar_fields; }

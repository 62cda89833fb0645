// ORACLE -- test infrastructure only.  Imported, linked or run solely by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, always as
// the checker or the timed CPU baseline, never as the product path.
//
// FP64 CPU restatement of sledgehamr's per-level time-step path in the
// reference's own (unfused) structure: a per-stage Rhs pass over 1024x8x8
// tiles, separate Copy/Saxpy/LinComb passes, a ghost fill per stage.
//
// Functions and the reference code they follow:
//   fill_rhs / fill_add_rhs   source/macros.h:301-485 (FillRhs, FillAddRhs)
//   rkn                      source/integrators/rkn.cpp:25-120, tableaux :124-205
//   lsssprk3                 source/integrators/lsssprk3.cpp:14-31
//   leapfrog                 source/integrators/leapfrog.cpp:13-57
//   advance                  source/integrators/integrator.cpp:11-31
//   fill_boundary            level_synchronizer.cpp:94-214 (periodic int_dir,
//                            same-level box-to-box)
//   te_avgdown               source/kernels.h:18-42,
//                            level_synchronizer.cpp:240-357
//   tag_with_truncation      source/macros.h:489-566, sledgehamr.cpp:280-348
//   create_shadow            sledgehamr.cpp:234-251
//   gw_spectrum              gravitational_waves.cpp:52-322 (CPU branch),
//                            gravitational_waves.h:101-133 (Lambda)
//   k tables                 pySledgehamr/Sledgehamr.py:7-47 (CreateUniqueVals,
//                            IndexToK), type 3 from utils Gradient<3>
//
// Parity pins (tests/): Laplacian/Gradient exact on polynomials and on
// plane-wave symbols; integrator orders of convergence and exact harmonic
// solutions; SSPRK3 equal to u0 + dt/6 (k1 + k2 + 4 k3); hand-built TE/tag
// blocks; brute-force DFT and Lambda identities; spectra_ks.hdf5 tables.
#pragma once

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <array>
#include <sstream>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "shim/sledgehamr.h"

namespace orc {

struct Box {
    int lo[3], hi[3];
    int n(int d) const { return hi[d] - lo[d] + 1; }
    long long vol() const { return (long long)n(0) * n(1) * n(2); }
};

/** One box of a MultiFab: valid box grown by g, x fastest, comp slowest. */
struct Fab {
    Box box;
    int g = 0, nc = 0;
    long long jst = 0, kst = 0, nst = 0;
    std::vector<double> d;
    void define(const Box &b, int ghost, int ncomp) {
        box = b;
        g = ghost;
        nc = ncomp;
        jst = b.n(0) + 2 * g;
        kst = jst * (b.n(1) + 2 * g);
        nst = kst * (b.n(2) + 2 * g);
        d.assign((size_t)(nst * nc), 0.);
    }
    double &at(int i, int j, int k, int n) {
        return d[(size_t)((i - box.lo[0] + g) + (j - box.lo[1] + g) * jst +
                          (k - box.lo[2] + g) * kst + n * nst)];
    }
    amrex::Array4<double> arr() {
        amrex::Array4<double> a;
        a.p = d.data();
        a.jstride = jst;
        a.kstride = kst;
        a.nstride = nst;
        a.begin = {box.lo[0] - g, box.lo[1] - g, box.lo[2] - g};
        return a;
    }
    amrex::Array4<const double> carr() const {
        amrex::Array4<const double> a;
        a.p = d.data();
        a.jstride = jst;
        a.kstride = kst;
        a.nstride = nst;
        a.begin = {box.lo[0] - g, box.lo[1] - g, box.lo[2] - g};
        return a;
    }
};

struct MultiFab {
    std::vector<Fab> fabs;
    int nc = 0, g = 0;
    double t = -DBL_MAX;
    int istep = 0;
    void define(const std::vector<Box> &boxes, int ncomp, int ghost) {
        fabs.assign(boxes.size(), Fab());
        for (size_t b = 0; b < boxes.size(); ++b) fabs[b].define(boxes[b], ghost, ncomp);
        nc = ncomp;
        g = ghost;
    }
    bool defined() const { return !fabs.empty(); }
};

/** MultiFab::Copy(dst, src, scomp, dcomp, ncomp, nghost). */
inline void Copy(MultiFab &dst, MultiFab &src, int sc, int dc, int nc, int ng) {
    for (size_t b = 0; b < dst.fabs.size(); ++b) {
        Fab &D = dst.fabs[b], &S = src.fabs[b];
#pragma omp parallel for schedule(static)
        for (int k = D.box.lo[2] - ng; k <= D.box.hi[2] + ng; ++k)
            for (int n = 0; n < nc; ++n)
                for (int j = D.box.lo[1] - ng; j <= D.box.hi[1] + ng; ++j)
                    for (int i = D.box.lo[0] - ng; i <= D.box.hi[0] + ng; ++i)
                        D.at(i, j, k, dc + n) = S.at(i, j, k, sc + n);
    }
}

/** MultiFab::Saxpy(dst, a, src, scomp, dcomp, ncomp, nghost): dst += a*src. */
inline void Saxpy(MultiFab &dst, double a, MultiFab &src, int sc, int dc, int nc,
                  int ng) {
    for (size_t b = 0; b < dst.fabs.size(); ++b) {
        Fab &D = dst.fabs[b], &S = src.fabs[b];
#pragma omp parallel for schedule(static)
        for (int k = D.box.lo[2] - ng; k <= D.box.hi[2] + ng; ++k)
            for (int n = 0; n < nc; ++n)
                for (int j = D.box.lo[1] - ng; j <= D.box.hi[1] + ng; ++j)
                    for (int i = D.box.lo[0] - ng; i <= D.box.hi[0] + ng; ++i)
                        D.at(i, j, k, dc + n) += a * S.at(i, j, k, sc + n);
    }
}

/** MultiFab::LinComb(dst, a, x, xc, b, y, yc, dc, nc, ng): dst = a x + b y. */
inline void LinComb(MultiFab &dst, double a, MultiFab &x, int xc, double bb,
                    MultiFab &y, int yc, int dc, int nc, int ng) {
    for (size_t b = 0; b < dst.fabs.size(); ++b) {
        Fab &D = dst.fabs[b], &X = x.fabs[b], &Y = y.fabs[b];
#pragma omp parallel for schedule(static)
        for (int k = D.box.lo[2] - ng; k <= D.box.hi[2] + ng; ++k)
            for (int n = 0; n < nc; ++n)
                for (int j = D.box.lo[1] - ng; j <= D.box.hi[1] + ng; ++j)
                    for (int i = D.box.lo[0] - ng; i <= D.box.hi[0] + ng; ++i)
                        D.at(i, j, k, dc + n) =
                            a * X.at(i, j, k, xc + n) + bb * Y.at(i, j, k, yc + n);
    }
}

/** Ghost fill of boxes that are each their own periodic domain (an
 *  ensemble of independent realizations): images wrap inside the box. */
inline void fill_boundary_independent(MultiFab &m) {
    if (m.g == 0) return;
    for (size_t b = 0; b < m.fabs.size(); ++b) {
        Fab &D = m.fabs[b];
        const int g = D.g;
        auto wrap = [](int v, int lo, int n) { return lo + (((v - lo) % n) + n) % n; };
#pragma omp parallel for schedule(static)
        for (int k = D.box.lo[2] - g; k <= D.box.hi[2] + g; ++k)
            for (int j = D.box.lo[1] - g; j <= D.box.hi[1] + g; ++j)
                for (int i = D.box.lo[0] - g; i <= D.box.hi[0] + g; ++i) {
                    const int si = wrap(i, D.box.lo[0], D.box.n(0));
                    const int sj = wrap(j, D.box.lo[1], D.box.n(1));
                    const int sk = wrap(k, D.box.lo[2], D.box.n(2));
                    if (si == i && sj == j && sk == k) continue;
                    for (int n = 0; n < m.nc; ++n) D.at(i, j, k, n) = D.at(si, sj, sk, n);
                }
    }
}

/** Same-level ghost fill with periodic wrap (FillBoundary, int_dir BCs).
 *  Ghost cells whose periodic image lies in no box keep their values. */
inline void fill_boundary(MultiFab &m, int N) {
    if (m.g == 0) return;
    auto wrap = [N](int v) { return ((v % N) + N) % N; };
    for (size_t b = 0; b < m.fabs.size(); ++b) {
        Fab &D = m.fabs[b];
        const int g = D.g;
#pragma omp parallel for schedule(static)
        for (int k = D.box.lo[2] - g; k <= D.box.hi[2] + g; ++k)
            for (int j = D.box.lo[1] - g; j <= D.box.hi[1] + g; ++j)
                for (int i = D.box.lo[0] - g; i <= D.box.hi[0] + g; ++i) {
                    if (i >= D.box.lo[0] && i <= D.box.hi[0] && j >= D.box.lo[1] &&
                        j <= D.box.hi[1] && k >= D.box.lo[2] && k <= D.box.hi[2])
                        continue;
                    const int wi = wrap(i), wj = wrap(j), wk = wrap(k);
                    for (size_t s = 0; s < m.fabs.size(); ++s) {
                        const Box &sb = m.fabs[s].box;
                        // Shift of the image that lands in box s.
                        int si = wi, sj = wj, sk = wk;
                        if (!(si >= sb.lo[0] && si <= sb.hi[0] && sj >= sb.lo[1] &&
                              sj <= sb.hi[1] && sk >= sb.lo[2] && sk <= sb.hi[2]))
                            continue;
                        for (int n = 0; n < m.nc; ++n)
                            D.at(i, j, k, n) = m.fabs[s].at(si, sj, sk, n);
                        break;
                    }
                }
    }
}

/** Mean of the 8 children (amrex::average_down) of every coarse cell that
 *  lies under a fine box; coarse boxes are the fine boxes coarsened by 2. */
inline void average_down(MultiFab &fine, MultiFab &crse) {
    for (size_t b = 0; b < fine.fabs.size(); ++b) {
        Fab &F = fine.fabs[b];
        Fab &Cf = crse.fabs[b];
#pragma omp parallel for schedule(static)
        for (int k = F.box.lo[2] / 2; k <= (F.box.hi[2] + 1) / 2 - 1; ++k)
            for (int j = F.box.lo[1] / 2; j <= (F.box.hi[1] + 1) / 2 - 1; ++j)
                for (int i = F.box.lo[0] / 2; i <= (F.box.hi[0] + 1) / 2 - 1; ++i)
                    for (int n = 0; n < fine.nc; ++n) {
                        double c = 0;
                        for (int dk = 0; dk < 2; ++dk)
                            for (int dj = 0; dj < 2; ++dj)
                                for (int di = 0; di < 2; ++di)
                                    c += F.at(2 * i + di, 2 * j + dj, 2 * k + dk, n);
                        Cf.at(i, j, k, n) = c * 0.125;
                    }
    }
}

/** AverageDownTo (level_synchronizer.cpp:219-233): the mean of the 8
 *  children of every coarse cell under a fine box, written into whichever
 *  coarse box holds that cell. */
inline void average_down_to(MultiFab &fine, MultiFab &crse) {
    for (size_t b = 0; b < fine.fabs.size(); ++b) {
        Fab &F = fine.fabs[b];
#pragma omp parallel for schedule(static)
        for (int k = F.box.lo[2] / 2; k <= (F.box.hi[2] + 1) / 2 - 1; ++k)
            for (int j = F.box.lo[1] / 2; j <= (F.box.hi[1] + 1) / 2 - 1; ++j)
                for (int i = F.box.lo[0] / 2; i <= (F.box.hi[0] + 1) / 2 - 1; ++i) {
                    Fab *Cf = nullptr;
                    for (auto &c : crse.fabs)
                        if (i >= c.box.lo[0] && i <= c.box.hi[0] && j >= c.box.lo[1] &&
                            j <= c.box.hi[1] && k >= c.box.lo[2] && k <= c.box.hi[2]) {
                            Cf = &c;
                            break;
                        }
                    if (!Cf) continue;  // not properly nested: nothing to update
                    for (int n = 0; n < fine.nc; ++n) {
                        double c = 0;
                        for (int dk = 0; dk < 2; ++dk)
                            for (int dj = 0; dj < 2; ++dj)
                                for (int di = 0; di < 2; ++di)
                                    c += F.at(2 * i + di, 2 * j + dj, 2 * k + dk, n);
                        Cf->at(i, j, k, n) = c * 0.125;
                    }
                }
    }
}

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline double rng_uniform(uint64_t seed, uint64_t stream, uint64_t idx) {
    const uint64_t h =
        splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + idx);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

/** Fourier symbol of Gradient<p> at index a of an N-point axis
 *  (Sledgehamr.py:18-27; p = 3 from the Gradient<3> weights). */
inline double index_to_k(int a, int N, int type) {
    const int nt = a > N / 2 - 1 ? a - N : a;
    const double th = 2. * M_PI / N * nt;
    if (type == 1) return std::sin(th);
    if (type == 2) return (8. * std::sin(th) - std::sin(2. * th)) / 6.;
    if (type == 3)
        return (45. * std::sin(th) - 9. * std::sin(2. * th) + std::sin(3. * th)) / 30.;
    return 0.;
}

/** Sorted unique i^2 + j^2 + k^2 over fftfreq(N)*N (Sledgehamr.py:7-12). */
inline std::vector<long long> unique_k2(int N) {
    std::set<long long> sq;
    for (int a = 0; a < N; ++a) {
        const long long f = a < (N + 1) / 2 ? a : a - N;
        sq.insert(f * f);
    }
    std::set<long long> out;
    for (long long x : sq)
        for (long long y : sq)
            for (long long z : sq) out.insert(x + y + z);
    return std::vector<long long>(out.begin(), out.end());
}

/** Unnormalised forward DFT of a real N^3 array (x fastest) keeping
 *  x-frequencies 0..N/2 (the R2C layout, x halved); separable O(N^4). */
inline void r2c(const std::vector<double> &in, int N,
                std::vector<std::complex<double>> &out) {
    const int NH = N / 2 + 1;
    std::vector<std::complex<double>> tw(N);
    for (int m = 0; m < N; ++m)
        tw[m] = std::polar(1.0, -2. * M_PI * (double)m / (double)N);
    std::vector<std::complex<double>> a((size_t)NH * N * N);
    // x axis
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < (long long)N * N; ++r)
        for (int kx = 0; kx < NH; ++kx) {
            std::complex<double> s = 0;
            for (int x = 0; x < N; ++x)
                s += in[(size_t)(r * N + x)] * tw[(size_t)((long long)kx * x % N)];
            a[(size_t)(r * NH + kx)] = s;
        }
    // y axis
    std::vector<std::complex<double>> b(a.size());
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; ++z)
        for (int ky = 0; ky < N; ++ky)
            for (int kx = 0; kx < NH; ++kx) {
                std::complex<double> s = 0;
                for (int y = 0; y < N; ++y)
                    s += a[(size_t)(((long long)z * N + y) * NH + kx)] *
                         tw[(size_t)((long long)ky * y % N)];
                b[(size_t)(((long long)z * N + ky) * NH + kx)] = s;
            }
    // z axis
    out.assign(a.size(), 0.);
#pragma omp parallel for schedule(static)
    for (int kz = 0; kz < N; ++kz)
        for (int ky = 0; ky < N; ++ky)
            for (int kx = 0; kx < NH; ++kx) {
                std::complex<double> s = 0;
                for (int z = 0; z < N; ++z)
                    s += b[(size_t)(((long long)z * N + ky) * NH + kx)] *
                         tw[(size_t)((long long)kz * z % N)];
                out[(size_t)(((long long)kz * N + ky) * NH + kx)] = s;
            }
}

/** gw_GetProjection / gw_GetLambda (gravitational_waves.h:101-133). */
inline double gw_projection(int i, int j, const double abc[3]) {
    const double norm = abc[0] * abc[0] + abc[1] * abc[1] + abc[2] * abc[2];
    const int l = std::min(i, j), m = std::max(i, j);
    return (double)(l == m) - abc[l] * abc[m] / norm;
}
inline double gw_lambda(int i, int j, int l, int m, const int abc[3],
                        const double *itk) {
    if (abc[0] == 0 && abc[1] == 0 && abc[2] == 0) return 0.;
    const double d[3] = {itk[abc[0]], itk[abc[1]], itk[abc[2]]};
    return gw_projection(i, l, d) * gw_projection(j, m, d) -
           gw_projection(i, j, d) * gw_projection(l, m, d) / 2.;
}

// ---------------------------------------------------------------------------

struct Config {
    int N0 = 0, max_level = 0, nghost = 0;
    double L = 0, cfl = 0, t_start = 0, t_end = 0;
    bool gw = false;
    int integrator = 10;
    std::vector<double> nodes, bbar, b;
    std::vector<std::vector<double>> tableau;
    std::vector<double> te_crit, ko;
    bool ko_on = false;
    int ko_order = 0;
    bool shadow_hierarchy = false;
    bool independent = false;  // level-0 boxes are separate periodic domains
    double regrid_dt = DBL_MAX;
    // Local regrid (local_regrid.h:97-122; blocking factor per level, the
    // last value repeated).
    std::vector<int> bf;
    int n_error_buf = 1, max_local_regrids = 10;
    double vt_strong = 1.05, vt_weak = 1.1;
    bool local_regrid = false;  // shm.local_regrid: regrid at scheduled times
    bool keep_layout = false;   // shm.keep_layout_on_global_regrid
    int interp_type = 2;  // amr.interpolation_type: 0 piecewise constant,
                          // 1 limited linear, 2 quadratic, 4 quartic
};

template <class Proj>
class Sim {
    using Ops = typename Proj::OrcOps;

  public:
    Proj prj;
    Config cfg;
    int nc = 0;
    std::vector<int> dimN;
    std::vector<double> dx, dt;
    std::vector<MultiFab> grid_new, grid_old;
    MultiFab shadow, shadow_tmp;
    MultiFab rhs_out;  // FillRhs target of the C interface, per instance
    std::vector<double> rhs_te;  // last truncation errors (per coarse cell)

    explicit Sim(const std::string &inputs) {
        auto &T = amrex::ParmParse::table();
        T.clear();
        std::istringstream in(inputs);
        std::string line;
        while (std::getline(in, line)) {
            const size_t h = line.find('#');
            if (h != std::string::npos) line = line.substr(0, h);
            const size_t e = line.find('=');
            if (e == std::string::npos) continue;
            std::string key = line.substr(0, e);
            key.erase(0, key.find_first_not_of(" \t"));
            key.erase(key.find_last_not_of(" \t\r") + 1);
            std::istringstream vs(line.substr(e + 1));
            std::vector<std::string> v;
            std::string tok;
            while (vs >> tok) v.push_back(tok);
            T[key] = v;
        }
        amrex::ParmParse pp("");
        pp.get("amr.coarse_level_grid_size", cfg.N0);
        pp.query("amr.max_refinement_levels", cfg.max_level);
        pp.query("amr.nghost", cfg.nghost);
        pp.get("sim.L", cfg.L);
        pp.get("sim.cfl", cfg.cfl);
        pp.get("sim.t_start", cfg.t_start);
        pp.get("sim.t_end", cfg.t_end);
        pp.query("sim.gravitational_waves", cfg.gw);
        pp.query("shm.independent_boxes", cfg.independent);
        pp.query("amr.regrid_dt", cfg.regrid_dt);
        pp.get("integrator.type", cfg.integrator);
        nc = cfg.gw ? Ops::kNGw : Ops::kNScalars;
        std::vector<std::string> names;
        for (auto *f : prj.scalar_fields) names.push_back(f->name);
        for (const char *n : {"u_xx", "u_yy", "u_zz", "u_xy", "u_xz", "u_yz", "du_xx",
                              "du_yy", "du_zz", "du_xy", "du_xz", "du_yz"})
            names.push_back(n);
        double ted = DBL_MAX;
        pp.query("amr.te_crit", ted);
        cfg.te_crit.assign(nc, ted);
        for (int n = 0; n < nc; ++n) {
            pp.query(("amr.te_crit_" + names[n]).c_str(), cfg.te_crit[n]);
            if (cfg.te_crit[n] != DBL_MAX) cfg.shadow_hierarchy = true;
        }
        double kod = 0;
        pp.query("sim.dissipation_strength", kod);
        cfg.ko.assign(nc, kod);
        for (int n = 0; n < nc; ++n) {
            pp.query(("sim.dissipation_strength_" + names[n]).c_str(), cfg.ko[n]);
            if (cfg.ko[n] > 0) cfg.ko_on = true;
        }
        if (cfg.ko_on) {
            cfg.ko_order = cfg.nghost;
            pp.query("sim.dissipation_order", cfg.ko_order);
        }
        // LevelSynchronizer::LevelSynchronizer (level_synchronizer.cpp:27-47).
        pp.query("amr.interpolation_type", cfg.interp_type);
        if (cfg.interp_type != 0 && cfg.interp_type != 1 && cfg.interp_type != 2 &&
            cfg.interp_type != 4)
            throw std::runtime_error("amr.interpolation_type must be 0, 1, 2 or 4");
        switch (cfg.integrator) {
            case 20: {
                amrex::ParmParse pr("integrator.rkn");
                pr.getarr("weights_bar_b", cfg.bbar);
                pr.getarr("weights_b", cfg.b);
                pr.getarr("nodes", cfg.nodes);
                std::vector<double> bt;
                pr.getarr("tableau", bt);
                int k = 0;
                for (size_t i = 0; i < cfg.nodes.size(); ++i) {
                    std::vector<double> row;
                    for (size_t j = 0; j <= i; ++j) row.push_back(bt[k++]);
                    cfg.tableau.push_back(row);
                }
                break;
            }
            case 21:
                cfg.nodes = {0.0, 0.5, 1.0};
                cfg.tableau = {{0.0}, {1. / 8., 0.0}, {0.0, 0.5, 0.0}};
                cfg.bbar = {1. / 6., 1. / 3., 0.0};
                cfg.b = {1. / 6., 4. / 6., 1. / 6.};
                break;
            case 22:
                cfg.nodes = {0.0, 1. / 5., 2. / 3., 1.0};
                cfg.tableau = {{0.0},
                               {1. / 50., 0.0},
                               {-1. / 27., 7. / 27., 0.0},
                               {3. / 10., -2. / 35., 9. / 35., 0.0}};
                cfg.bbar = {14. / 336., 100. / 336., 54. / 336., 0.0};
                cfg.b = {14. / 336., 125. / 336., 162. / 336., 35. / 336.};
                break;
        }
        {
            // Checked here when given, else where the local regrid uses it.
            std::vector<int> b{8};
            const bool given = pp.queryarr("amr.blocking_factor", b);
            for (int lev = 0; lev <= cfg.max_level; ++lev) {
                const int v = b[std::min<size_t>(lev, b.size() - 1)];
                if (given) check_bf(v, cfg.N0 << lev);
                cfg.bf.push_back(v);
            }
        }
        pp.query("amr.n_error_buf", cfg.n_error_buf);
        pp.query("amr.max_local_regrids", cfg.max_local_regrids);
        pp.query("amr.volume_threshold_strong", cfg.vt_strong);
        pp.query("amr.volume_threshold_weak", cfg.vt_weak);
        pp.query("shm.local_regrid", cfg.local_regrid);
        pp.query("shm.keep_layout_on_global_regrid", cfg.keep_layout);
        for (int lev = 0; lev <= cfg.max_level; ++lev) {
            dimN.push_back(cfg.N0 * (1 << lev));
            dx.push_back(cfg.L / (double)dimN[lev]);
            dt.push_back(dx[lev] * cfg.cfl);
        }
        grid_new.resize(cfg.max_level + 1);
        grid_old.resize(cfg.max_level + 1);
        lr.no_local.assign(cfg.max_level + 1, 0);
        lr.do_global.assign(cfg.max_level + 1, 0);
        // MultiFab::norm0 of grid_new[lev] over every component (valid cells).
        prj.level_norm0_ = [this](int lev) {
            std::vector<double> m(nc, 0.);
            for (auto &f : grid_new[lev].fabs)
                for (int n = 0; n < nc; ++n)
                    for (int k = f.box.lo[2]; k <= f.box.hi[2]; ++k)
                        for (int j = f.box.lo[1]; j <= f.box.hi[1]; ++j)
                            for (int i = f.box.lo[0]; i <= f.box.hi[0]; ++i)
                                m[n] = std::max(m[n], std::fabs(f.at(i, j, k, n)));
            return m;
        };
        prj.Init();
        int mgs = cfg.N0;
        pp.query("amr.max_grid_size", mgs);
        std::vector<Box> boxes;
        for (int z = 0; z < cfg.N0; z += mgs)
            for (int y = 0; y < cfg.N0; y += mgs)
                for (int x = 0; x < cfg.N0; x += mgs)
                    boxes.push_back(Box{{x, y, z}, {x + mgs - 1, y + mgs - 1, z + mgs - 1}});
        set_level(0, boxes);
    }

    void set_level(int lev, const std::vector<Box> &boxes) {
        grid_new[lev].define(boxes, nc, cfg.nghost);
        grid_old[lev].define(boxes, nc, cfg.nghost);
        grid_new[lev].t = cfg.t_start;
        grid_old[lev].t = -DBL_MAX;
    }

    int finest() const {
        int f = 0;
        for (int l = 0; l <= cfg.max_level; ++l)
            if (grid_new[l].defined()) f = l;
        return f;
    }

    std::vector<double> params_rhs(double t, int lev) {
        std::vector<double> p;
        prj.SetParamsRhs(p, t, lev);
        p.resize(std::max<size_t>(p.size(), 64), 0.);
        return p;
    }
    std::vector<double> params_gw(double t, int lev) {
        std::vector<double> p;
        prj.SetParamsGravitationalWaveRhs(p, t, lev);
        p.resize(std::max<size_t>(p.size(), 64), 0.);
        return p;
    }

    /** FillRhs / FillAddRhs (macros.h:301-485): per-cell Rhs, GW source and
     *  backreaction, Kreiss-Oliger; with add: rhs = f + weight*rhs_old. */
    void fill_rhs(MultiFab &rhs, MultiFab &state, double time, int lev, double dtl,
                  double dxl, bool add, double weight) {
        const std::vector<double> p = params_rhs(time, lev);
        const std::vector<double> pg = params_gw(time, lev);
        const int ncomp = nc;
        for (size_t b = 0; b < state.fabs.size(); ++b) {
            Fab &S = state.fabs[b];
            Fab &Rf = rhs.fabs[b];
            const amrex::Array4<const double> sa = S.carr();
            const amrex::Array4<double> ra = Rf.arr();
            // 1024 x 8 x 8 tiles (MFIter tiling on CPU).
            const int tz = 8, ty = 8, tx = 1024;
            std::vector<std::array<int, 3>> tiles;
            for (int k0 = S.box.lo[2]; k0 <= S.box.hi[2]; k0 += tz)
                for (int j0 = S.box.lo[1]; j0 <= S.box.hi[1]; j0 += ty)
                    for (int i0 = S.box.lo[0]; i0 <= S.box.hi[0]; i0 += tx)
                        tiles.push_back({i0, j0, k0});
#pragma omp parallel for schedule(dynamic)
            for (size_t t = 0; t < tiles.size(); ++t) {
                const auto [i0, j0, k0] = tiles[t];
                for (int k = k0; k <= std::min(S.box.hi[2], k0 + tz - 1); ++k)
                    for (int j = j0; j <= std::min(S.box.hi[1], j0 + ty - 1); ++j)
                        for (int i = i0; i <= std::min(S.box.hi[0], i0 + tx - 1); ++i) {
                            double tmp[64];
                            if (add)
                                for (int n = 0; n < ncomp; ++n) tmp[n] = ra(i, j, k, n);
                            Ops::rhs(ra, sa, i, j, k, lev, time, dtl, dxl, p.data());
                            if (cfg.gw) {
                                Ops::gw_rhs(ra, sa, i, j, k, lev, time, dtl, dxl,
                                            pg.data());
                                Ops::gw_back(ra, sa, i, j, k, lev, time, dtl, dxl,
                                             p.data(), pg.data());
                            }
                            if (cfg.ko_on) {
                                for (int n = 0; n < ncomp; ++n)
                                    ra(i, j, k, n) +=
                                        cfg.ko_order == 2
                                            ? sledgehamr::kernels::KreissOligerDissipation<2>(
                                                  sa, i, j, k, n, dxl, cfg.ko[n])
                                            : sledgehamr::kernels::KreissOligerDissipation<3>(
                                                  sa, i, j, k, n, dxl, cfg.ko[n]);
                            }
                            if (add)
                                for (int n = 0; n < ncomp; ++n)
                                    ra(i, j, k, n) += weight * tmp[n];
                        }
            }
        }
    }

    /** FillPatch / FillIntermediatePatch (level_synchronizer.cpp:92-209) at
     *  `time`: on a fine level every ghost cell first takes level lev-1
     *  interpolated in time and space, then the same-level fill overwrites
     *  the ghost cells that boxes of the level cover (FillPatchTwoLevels). */
    void fill_patch(MultiFab &m, int lev, double time) {
        if (cfg.independent && lev == 0) {
            fill_boundary_independent(m);
            return;
        }
        if (lev > 0) fill_coarse_ghosts(m, lev, time);
        fill_boundary(m, lev < 0 ? cfg.N0 / 2 : dimN[lev]);
    }

    /** LevelSynchronizer::FillCoarsePatch (level_synchronizer.cpp:56-88):
     *  every cell of grid_new[lev], ghosts included, interpolated from level
     *  lev-1 at the level's time (InterpFromCoarseLevel). */
    void fill_coarse_patch(int lev) {
        if (lev < 1) throw std::runtime_error("FillCoarsePatch needs a level above 0");
        fill_coarse_ghosts(grid_new[lev], lev, grid_new[lev].t, true, true);
    }

    /** Weights of the level-lev states at `time`: GetLevelData picks the new
     *  or the old state within 1e-3 of their time difference, else both
     *  (level_synchronizer.cpp:613-632), combined linearly in time. */
    void time_weights(int lev, double time, double &wo, double &wn) const {
        const MultiFab &New = grid_new[lev], &Old = grid_old[lev];
        const double teps = std::fabs(New.t - Old.t) * 1.e-3;
        if (time > New.t - teps && time < New.t + teps) {
            wo = 0.;
            wn = 1.;
        } else if (time > Old.t - teps && time < Old.t + teps) {
            wo = 1.;
            wn = 0.;
        } else {
            wo = (New.t - time) / (New.t - Old.t);
            wn = (time - Old.t) / (New.t - Old.t);
        }
    }

    /** Coarse part of FillPatchTwoLevels for every ghost cell of level lev.
     *  Spatial interpolation, ratio 2, at the fine cell centre x, y, z =
     *  +-1/4 coarse cells from the parent's centre:
     *    type 0: the parent value;
     *    type 1: c + sx x + sy y + sz z with monotonized-central slopes
     *            (sl = c - c[-1], sr = c[+1] - c, sc = (sl + sr) / 2;
     *            s = sign(sc) min(|sc|, 2|sl|, 2|sr|) if sl sr > 0, else 0):
     *            conservative, exact for linear data, no new extrema;
     *    type 2: c + sx x + sy y + sz z + sxx x^2 + syy y^2 + szz z^2
     *            + sxy x y + sxz x z + syz y z, with central differences
     *            sx = (c[+1] - c[-1]) / 2, sxx = (c[+1] - 2 c + c[-1]) / 2,
     *            sxy = (c[+1,+1] - c[-1,+1] - c[+1,-1] + c[-1,-1]) / 4:
     *            exact for every polynomial of degree <= 2;
     *    type 4: the tensor product of the conservative quartic weights
     *            (3, -22, 128, 22, -3) / 128 over coarse offsets -2..2 for
     *            an upper child, mirrored for a lower one: the child's
     *            average of the quartic with the 5 coarse averages.
     *  Coarse values are the time-weighted states of level lev-1 at the
     *  periodic image of each coarse cell; a coarse cell that no box of
     *  level lev-1 holds is an error (the level is not properly nested). */
    void fill_coarse_ghosts(MultiFab &m, int lev, double time, bool all = false,
                            bool first_only = false) {
        MultiFab &New = grid_new[lev - 1], &Old = grid_old[lev - 1];
        double wo, wn;
        time_weights(lev - 1, time, wo, wn);
        if (first_only && wo != 0. && wn != 0.) {
            // FillCoarsePatch hands InterpFromCoarseLevel only the first state
            // of GetLevelData (level_synchronizer.cpp:64, 83): the old one.
            wo = 1.;
            wn = 0.;
        }
        const int Nc = dimN[lev - 1];
        auto wrap = [Nc](int v) { return ((v % Nc) + Nc) % Nc; };
        // 19 points: centre, 6 faces, 12 edges (i, j, k offsets).
        static const int P[19][3] = {
            {0, 0, 0},  {-1, 0, 0}, {1, 0, 0},   {0, -1, 0},  {0, 1, 0},
            {0, 0, -1}, {0, 0, 1},  {1, 1, 0},   {-1, 1, 0},  {1, -1, 0},
            {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1},  {1, 0, -1},  {-1, 0, -1},
            {0, 1, 1},  {0, -1, 1}, {0, 1, -1},  {0, -1, -1}};
        const int np = cfg.interp_type == 0 ? 1 : cfg.interp_type == 1 ? 7 : 19;
        if (cfg.interp_type == 4) {
            fill_coarse_ghosts_quartic(m, lev, wo, wn, all);
            return;
        }
        bool nested = true;
        for (Fab &F : m.fabs) {
            const int g = F.g;
#pragma omp parallel for schedule(dynamic)
            for (int k = F.box.lo[2] - g; k <= F.box.hi[2] + g; ++k)
                for (int j = F.box.lo[1] - g; j <= F.box.hi[1] + g; ++j)
                    for (int i = F.box.lo[0] - g; i <= F.box.hi[0] + g; ++i) {
                        if (!all && i >= F.box.lo[0] && i <= F.box.hi[0] &&
                            j >= F.box.lo[1] && j <= F.box.hi[1] && k >= F.box.lo[2] &&
                            k <= F.box.hi[2])
                            continue;
                        const int ic = i >> 1, jc = j >> 1, kc = k >> 1;
                        int fb[19], ci[19], cj[19], ck[19];
                        bool ok = true;
                        for (int p = 0; p < np; ++p) {
                            ci[p] = wrap(ic + P[p][0]);
                            cj[p] = wrap(jc + P[p][1]);
                            ck[p] = wrap(kc + P[p][2]);
                            fb[p] = -1;
                            for (size_t b = 0; b < New.fabs.size() && fb[p] < 0; ++b) {
                                const Box &B = New.fabs[b].box;
                                if (ci[p] >= B.lo[0] && ci[p] <= B.hi[0] && cj[p] >= B.lo[1] &&
                                    cj[p] <= B.hi[1] && ck[p] >= B.lo[2] && ck[p] <= B.hi[2])
                                    fb[p] = (int)b;
                            }
                            if (fb[p] < 0) ok = false;
                        }
                        if (!ok) {
#pragma omp atomic write
                            nested = false;
                            continue;
                        }
                        for (int n = 0; n < m.nc; ++n) {
                            double c[19];
                            for (int p = 0; p < np; ++p) {
                                const double vn = New.fabs[fb[p]].at(ci[p], cj[p], ck[p], n);
                                if (wo == 0.)
                                    c[p] = vn;
                                else {
                                    const double vo = Old.fabs[fb[p]].at(ci[p], cj[p], ck[p], n);
                                    c[p] = wn == 0. ? vo : wo * vo + wn * vn;
                                }
                            }
                            if (np == 1) {
                                F.at(i, j, k, n) = c[0];
                                continue;
                            }
                            const double x = (i & 1) ? 0.25 : -0.25;
                            const double y = (j & 1) ? 0.25 : -0.25;
                            const double z = (k & 1) ? 0.25 : -0.25;
                            if (np == 7) {
                                auto mc = [](double cm, double c0, double cp) {
                                    const double sl = c0 - cm, sr = cp - c0;
                                    const double sc = 0.5 * (sl + sr);
                                    if (!(sl * sr > 0.)) return 0.;
                                    const double a = std::min(std::fabs(sc),
                                                              std::min(2. * std::fabs(sl),
                                                                       2. * std::fabs(sr)));
                                    return sc > 0. ? a : -a;
                                };
                                F.at(i, j, k, n) = c[0] + x * mc(c[1], c[0], c[2]) +
                                                   y * mc(c[3], c[0], c[4]) +
                                                   z * mc(c[5], c[0], c[6]);
                                continue;
                            }
                            const double sx = 0.5 * (c[2] - c[1]), sy = 0.5 * (c[4] - c[3]),
                                         sz = 0.5 * (c[6] - c[5]);
                            const double sxx = 0.5 * (c[2] - 2. * c[0] + c[1]);
                            const double syy = 0.5 * (c[4] - 2. * c[0] + c[3]);
                            const double szz = 0.5 * (c[6] - 2. * c[0] + c[5]);
                            const double sxy = 0.25 * (c[7] - c[8] - c[9] + c[10]);
                            const double sxz = 0.25 * (c[11] - c[12] - c[13] + c[14]);
                            const double syz = 0.25 * (c[15] - c[16] - c[17] + c[18]);
                            F.at(i, j, k, n) = c[0] + x * sx + y * sy + z * sz +
                                               x * x * sxx + y * y * syy + z * z * szz +
                                               x * y * sxy + x * z * sxz + y * z * syz;
                        }
                    }
        }
        if (!nested)
            throw std::runtime_error("level " + std::to_string(lev) +
                                     " is not properly nested: a ghost cell needs coarse "
                                     "cells that no box of level " +
                                     std::to_string(lev - 1) + " holds");
    }

    /** Type 4 of fill_coarse_ghosts: sum over coarse offsets a, b, c in
     *  -2..2 (c outermost, a innermost) of wz[c] wy[b] wx[a] C. */
    void fill_coarse_ghosts_quartic(MultiFab &m, int lev, double wo, double wn,
                                    bool all) {
        MultiFab &New = grid_new[lev - 1], &Old = grid_old[lev - 1];
        const int Nc = dimN[lev - 1];
        auto wrap = [Nc](int v) { return ((v % Nc) + Nc) % Nc; };
        static const double WU[5] = {3. / 128., -22. / 128., 1., 22. / 128., -3. / 128.};
        bool nested = true;
        for (Fab &F : m.fabs) {
            const int g = F.g;
#pragma omp parallel for schedule(dynamic)
            for (int k = F.box.lo[2] - g; k <= F.box.hi[2] + g; ++k)
                for (int j = F.box.lo[1] - g; j <= F.box.hi[1] + g; ++j)
                    for (int i = F.box.lo[0] - g; i <= F.box.hi[0] + g; ++i) {
                        if (!all && i >= F.box.lo[0] && i <= F.box.hi[0] &&
                            j >= F.box.lo[1] && j <= F.box.hi[1] && k >= F.box.lo[2] &&
                            k <= F.box.hi[2])
                            continue;
                        const int ic = i >> 1, jc = j >> 1, kc = k >> 1;
                        double wx[5], wy[5], wz[5];
                        for (int q = 0; q < 5; ++q) {
                            wx[q] = (i & 1) ? WU[q] : WU[4 - q];
                            wy[q] = (j & 1) ? WU[q] : WU[4 - q];
                            wz[q] = (k & 1) ? WU[q] : WU[4 - q];
                        }
                        for (int n = 0; n < m.nc; ++n) {
                            double acc = 0.;
                            bool ok = true;
                            for (int cz = 0; cz < 5 && ok; ++cz)
                                for (int by = 0; by < 5 && ok; ++by)
                                    for (int ax = 0; ax < 5 && ok; ++ax) {
                                        const int ci = wrap(ic + ax - 2), cj = wrap(jc + by - 2),
                                                  ck = wrap(kc + cz - 2);
                                        int fb = -1;
                                        for (size_t b = 0; b < New.fabs.size() && fb < 0; ++b) {
                                            const Box &B = New.fabs[b].box;
                                            if (ci >= B.lo[0] && ci <= B.hi[0] && cj >= B.lo[1] &&
                                                cj <= B.hi[1] && ck >= B.lo[2] && ck <= B.hi[2])
                                                fb = (int)b;
                                        }
                                        if (fb < 0) {
                                            ok = false;
                                            break;
                                        }
                                        const double vn = New.fabs[fb].at(ci, cj, ck, n);
                                        double v;
                                        if (wo == 0.)
                                            v = vn;
                                        else {
                                            const double vo = Old.fabs[fb].at(ci, cj, ck, n);
                                            v = wn == 0. ? vo : wo * vo + wn * vn;
                                        }
                                        acc += wz[cz] * wy[by] * wx[ax] * v;
                                    }
                            if (!ok) {
#pragma omp atomic write
                                nested = false;
                                continue;
                            }
                            F.at(i, j, k, n) = acc;
                        }
                    }
        }
        if (!nested)
            throw std::runtime_error("level " + std::to_string(lev) +
                                     " is not properly nested: a ghost cell needs coarse "
                                     "cells that no box of level " +
                                     std::to_string(lev - 1) + " holds");
    }

    /** IntegratorLsssprk3::Integrate (lsssprk3.cpp:14-31). */
    void lsssprk3(MultiFab &mo, MultiFab &mn, int lev, double dtl, double dxl) {
        const double t0 = mo.t, t1 = t0 + dtl;
        MultiFab k1;
        k1.define(boxes_of(mo), nc, cfg.nghost);
        fill_rhs(k1, mo, t0, lev, dtl, dxl, false, 0.);
        LinComb(mn, 1, mo, 0, dtl, k1, 0, 0, nc, 0);
        fill_patch(mn, lev, t1);
        fill_rhs(k1, mn, t1, lev, dtl, dxl, true, 1.);
        LinComb(mn, 1, mo, 0, dtl / 4., k1, 0, 0, nc, 0);
        fill_patch(mn, lev, t0 + dtl / 2.);
        fill_rhs(k1, mn, t0 + dtl / 2., lev, dtl, dxl, true, 0.25);
        LinComb(mn, 1, mo, 0, dtl * 2. / 3., k1, 0, 0, nc, 0);
        fill_patch(mn, lev, t1);
    }

    /** IntegratorRkn::Integrate (rkn.cpp:25-120). */
    void rkn(MultiFab &mo, MultiFab &mn, int lev, double dtl, double dxl) {
        const double t = mo.t;
        const int ng = cfg.nghost;
        const int N = nc, Ngrav = cfg.gw ? 12 : 0, Nf = N - Ngrav;
        const int uN = Nf / 2, uN0 = 0, uN1 = uN;
        const int gN = Ngrav / 2, gN0 = Nf, gN1 = gN0 + gN;
        const int s = (int)cfg.nodes.size();
        std::vector<MultiFab> F(s);
        for (auto &f : F) f.define(boxes_of(mo), N, ng);
        for (int i = 0; i < s; ++i) {
            const double st = t + dtl * cfg.nodes[i];
            Copy(mn, mo, 0, 0, N, ng);
            Saxpy(mn, dtl * cfg.nodes[i], mo, uN1, uN0, uN, ng);
            if (Ngrav > 0) Saxpy(mn, dtl * cfg.nodes[i], mo, gN1, gN0, gN, ng);
            for (int j = 0; j < i; ++j) {
                Saxpy(mn, dtl * dtl * cfg.tableau[i][j], F[j], uN1, uN0, uN, ng);
                if (Ngrav > 0)
                    Saxpy(mn, dtl * dtl * cfg.tableau[i][j], F[j], gN1, gN0, gN, ng);
            }
            fill_patch(mn, lev, st);
            fill_rhs(F[i], mn, st, lev, dtl, dxl, false, 0.);
        }
        Copy(mn, mo, 0, 0, N, ng);
        Saxpy(mn, dtl, mn, uN1, uN0, uN, ng);
        if (Ngrav > 0) Saxpy(mn, dtl, mn, gN1, gN0, gN, ng);
        for (int i = 0; i < s; ++i) {
            Saxpy(mn, dtl * dtl * cfg.bbar[i], F[i], uN1, uN0, uN, ng);
            Saxpy(mn, dtl * cfg.b[i], F[i], uN1, uN1, uN, ng);
            if (Ngrav == 0) continue;
            Saxpy(mn, dtl * dtl * cfg.bbar[i], F[i], gN1, gN0, gN, ng);
            Saxpy(mn, dtl * cfg.b[i], F[i], gN1, gN1, gN, ng);
        }
        fill_patch(mn, lev, t + dtl);
    }

    /** IntegratorLeapfrog::Integrate (leapfrog.cpp:13-57). */
    void leapfrog(MultiFab &mo, MultiFab &mn, int lev, double dtl, double dxl) {
        const double t0 = mo.t, t1 = t0 + dtl;
        const int N = nc, Ngrav = cfg.gw ? 12 : 0, Nf = N - Ngrav;
        const int uN = Nf / 2, uN0 = 0, uN1 = uN;
        const int gN = Ngrav / 2, gN0 = Nf, gN1 = gN0 + gN;
        MultiFab a, vh;
        a.define(boxes_of(mo), N, cfg.nghost);
        vh.define(boxes_of(mo), N, cfg.nghost);
        fill_rhs(a, mo, t0, lev, dtl, dxl, false, 0.);
        LinComb(vh, 1, mo, 0, dtl / 2., a, 0, 0, N, 0);
        LinComb(mn, 1, mo, uN0, dtl, vh, uN1, uN0, uN, 0);
        if (Ngrav > 0) LinComb(mn, 1, mo, gN0, dtl, vh, gN1, gN0, gN, 0);
        fill_patch(mn, lev, t1);
        fill_rhs(a, mn, t1, lev, dtl, dxl, false, 0.);
        LinComb(mn, 1, vh, uN1, dtl / 2., a, uN1, uN1, uN, 0);
        if (Ngrav > 0) LinComb(mn, 1, vh, gN1, dtl / 2., a, gN1, gN1, gN, 0);
        fill_patch(mn, lev, t1);
    }

    static std::vector<Box> boxes_of(const MultiFab &m) {
        std::vector<Box> b;
        for (auto &f : m.fabs) b.push_back(f.box);
        return b;
    }

    void integrate(MultiFab &mo, MultiFab &mn, int lev, double dtl, double dxl) {
        if (cfg.integrator == 10)
            lsssprk3(mo, mn, lev, dtl, dxl);
        else if (cfg.integrator == 11)
            leapfrog(mo, mn, lev, dtl, dxl);
        else
            rkn(mo, mn, lev, dtl, dxl);
    }

    /** Integrator::Advance (integrator.cpp:11-31). */
    void advance(int lev) {
        if (lev >= 0) std::swap(grid_old[lev], grid_new[lev]);
        const double dtl = lev < 0 ? dt[0] * 2. : dt[lev];
        const double dxl = lev < 0 ? dx[0] * 2. : dx[lev];
        MultiFab &mo = lev < 0 ? shadow_tmp : grid_old[lev];
        MultiFab &mn = lev < 0 ? shadow : grid_new[lev];
        if (!mn.defined()) mn.define(boxes_of(mo), nc, cfg.nghost);
        fill_patch(mo, lev, mo.t);
        integrate(mo, mn, lev, dtl, dxl);
        mn.t = mo.t + dtl;
        mn.istep = mo.istep + 1;
    }

    /** Sledgehamr::CreateShadowLevel (sledgehamr.cpp:234-251). */
    void create_shadow() {
        std::vector<Box> cb;
        for (auto &f : grid_old[0].fabs) {
            Box b;
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = f.box.lo[a] / 2;
                b.hi[a] = (f.box.hi[a] + 1) / 2 - 1;
            }
            cb.push_back(b);
        }
        shadow.define(cb, nc, cfg.nghost);
        shadow_tmp.define(cb, nc, cfg.nghost);
        shadow_tmp.t = grid_old[0].t;
        average_down(grid_old[0], shadow_tmp);
        advance(-1);
    }

    /** AverageDownWithTruncationError over every coarse cell under the fine
     *  boxes (kernels.h:18-42; level_synchronizer.cpp:262-299), crse being
     *  the shadow level (lev 0) or level lev-1.  te is stored per fine box at
     *  even fine indices (here: per coarse cell, ncomp planes). */
    void te_avgdown(int lev, std::vector<double> &te, bool avgdown) {
        MultiFab &fine = grid_new[lev];
        const long long ncoarse = [&] {
            long long s = 0;
            for (auto &f : fine.fabs) s += f.box.vol() / 8;
            return s;
        }();
        te.assign((size_t)ncoarse * nc, 0.);
        long long off = 0;
        for (size_t b = 0; b < fine.fabs.size(); ++b) {
            Fab &F = fine.fabs[b];
            const int cnx = F.box.n(0) / 2, cny = F.box.n(1) / 2, cnz = F.box.n(2) / 2;
            for (int k = 0; k < cnz; ++k)
                for (int j = 0; j < cny; ++j)
                    for (int i = 0; i < cnx; ++i) {
                        const int ci = F.box.lo[0] / 2 + i, cj = F.box.lo[1] / 2 + j,
                                  ck = F.box.lo[2] / 2 + k;
                        const int ii = 2 * ci, jj = 2 * cj, kk = 2 * ck;
                        double *crse_cell[64];
                        for (int n = 0; n < nc; ++n) crse_cell[n] = coarse_ptr(lev, ci, cj, ck, n);
                        for (int n = 0; n < nc; ++n) {
                            const double volfrac = 1.0 / 8.0;
                            double c = F.at(ii, jj, kk, n);
                            c += F.at(ii + 1, jj, kk, n);
                            c += F.at(ii, jj + 1, kk, n);
                            c += F.at(ii, jj, kk + 1, n);
                            c += F.at(ii + 1, jj + 1, kk, n);
                            c += F.at(ii + 1, jj, kk + 1, n);
                            c += F.at(ii, jj + 1, kk + 1, n);
                            c += F.at(ii + 1, jj + 1, kk + 1, n);
                            const double avg = volfrac * c;
                            te[(size_t)(n * ncoarse + off + i + (long long)cnx * (j + (long long)cny * k))] =
                                std::fabs(*crse_cell[n] - avg);
                            if (avgdown) *crse_cell[n] = avg;
                        }
                    }
            off += (long long)cnx * cny * cnz;
        }
    }

    double *coarse_ptr(int lev, int ci, int cj, int ck, int n) {
        MultiFab &C = lev == 0 ? shadow : grid_new[lev - 1];
        for (auto &f : C.fabs)
            if (ci >= f.box.lo[0] && ci <= f.box.hi[0] && cj >= f.box.lo[1] &&
                cj <= f.box.hi[1] && ck >= f.box.lo[2] && ck <= f.box.hi[2])
                return &f.at(ci, cj, ck, n);
        throw std::runtime_error("coarse cell not covered");
    }

    /** TagWithTruncationCpu (macros.h:489-566) executed literally: per tile
     *  and cell in k, j, i order, CLEAR own cell, user tag, TE block SET.
     *  mask holds one byte per fine valid cell (box-major, x fastest).
     *  User criteria may read neighbours: the ghost cells are filled at the
     *  level's time first (the reference reads whatever its ghosts hold from
     *  the last fill; here that is defined as FillPatch at the tag time). */
    void tag(int lev, const std::vector<double> &te, bool with_te,
             std::vector<uint8_t> &mask, std::vector<long long> &counts) {
        MultiFab &st = grid_new[lev];
        const double time = st.t;
        fill_patch(st, lev, time);
        std::vector<double> ptag, pmod;
        prj.SetParamsTagCellForRefinement(ptag, time, lev);
        if (cfg.shadow_hierarchy) prj.SetParamsTruncationModifier(pmod, time, lev);
        ptag.resize(std::max<size_t>(ptag.size(), 1));
        pmod.resize(std::max<size_t>(pmod.size(), 1));
        long long total = 0, user = 0;
        std::vector<long long> trunc(nc, 0);
        long long nvalid = 0;
        for (auto &f : st.fabs) nvalid += f.box.vol();
        mask.assign((size_t)nvalid, 0);
        long long moff = 0, coff = 0;
        long long ncoarse = nvalid / 8;
        for (size_t b = 0; b < st.fabs.size(); ++b) {
            Fab &F = st.fabs[b];
            const amrex::Array4<const double> sa = F.carr();
            const Box &bx = F.box;
            const int nx = bx.n(0), ny = bx.n(1);
            auto mk = [&](int i, int j, int k) -> uint8_t & {
                return mask[(size_t)(moff + (i - bx.lo[0]) +
                                     (long long)nx * ((j - bx.lo[1]) +
                                                      (long long)ny * (k - bx.lo[2])))];
            };
            const int cnx = nx / 2, cny = ny / 2;
            for (int k0 = bx.lo[2]; k0 <= bx.hi[2]; k0 += 8)
                for (int j0 = bx.lo[1]; j0 <= bx.hi[1]; j0 += 8)
                    for (int k = k0; k <= std::min(bx.hi[2], k0 + 7); ++k)
                        for (int j = j0; j <= std::min(bx.hi[1], j0 + 7); ++j)
                            for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
                                mk(i, j, k) = 0;  // CLEAR
                                const bool res = Ops::tag(sa, i, j, k, lev, time, dt[lev],
                                                          dx[lev], ptag.data());
                                if (res) {
                                    mk(i, j, k) = 1;
                                    ++user;
                                    ++total;
                                }
                                if (!with_te) continue;
                                bool te_res = false;
                                if (!((i % 2 != 0) || (j % 2 != 0) || (k % 2 != 0))) {
                                    const long long cidx =
                                        coff + (i - bx.lo[0]) / 2 +
                                        (long long)cnx * ((j - bx.lo[1]) / 2 +
                                                          (long long)cny * ((k - bx.lo[2]) / 2));
                                    sledgehamr::utils::constexpr_for<0, 64, 1>([&](auto n) {
                                        if constexpr (n < Ops::kNGw) {
                                            if (n >= nc) return;
                                            const double m = Ops::template te_mod<n>(
                                                sa, i, j, k, lev, time, dt[lev], dx[lev],
                                                te[(size_t)(n * ncoarse + cidx)], pmod.data());
                                            if (m >= cfg.te_crit[n]) {
                                                te_res = true;
                                                trunc[n] += 8;
                                            }
                                        }
                                    });
                                }
                                if (te_res) {
                                    for (int d = 0; d < 8; ++d)
                                        mk(i + (d & 1), j + ((d >> 1) & 1), k + (d >> 2)) = 1;
                                    total += 8 - (int)res;
                                }
                            }
            moff += bx.vol();
            coff += bx.vol() / 8;
        }
        counts.assign(2 + nc, 0);
        counts[0] = total;
        counts[1] = user;
        for (int n = 0; n < nc; ++n) counts[2 + n] = trunc[n];
    }

    /** GravitationalWaves::ComputeSpectrum, CPU branch.  mode 0 unbinned
     *  (lower_bound over unique |k|^2), 1 rounded |k|, 2 log bins (new). */
    void gw_spectrum(int mode, int nbins, int zp, int proj, std::vector<double> &spec,
                     std::vector<double> &kv, int modifier = -1,
                     const std::vector<double> &mparams = {}) {
        const int N = cfg.N0 * zp;
        const int NS = Ops::kNScalars;
        // GravitationalWavesSpectrumModifier::SelectComponents
        // (gravitational_waves.h:79-92); default du_xx, du_xy, du_xz, du_yy,
        // du_yz, du_zz.
        int comps[6] = {NS + 6, NS + 9, NS + 10, NS + 7, NS + 11, NS + 8};
        using Mods = typename Ops::GwModifiers;
        if (modifier >= 0 && !Mods::select(modifier, comps))
            throw std::runtime_error("no such GW spectrum modifier");
        const double dk = 2. * M_PI / (cfg.L * zp);
        std::vector<double> mp = mparams;
        mp.resize(std::max<size_t>(mp.size(), 1), 0.);
        const int mat[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        std::vector<std::vector<std::complex<double>>> u(6);
        for (int c = 0; c < 6; ++c) {
            std::vector<double> dense((size_t)N * N * N, 0.);
            for (auto &f : grid_new[0].fabs)
                for (int k = f.box.lo[2]; k <= f.box.hi[2]; ++k)
                    for (int j = f.box.lo[1]; j <= f.box.hi[1]; ++j)
                        for (int i = f.box.lo[0]; i <= f.box.hi[0]; ++i)
                            dense[(size_t)(i + (long long)N * (j + (long long)N * k))] =
                                f.at(i, j, k, comps[c]);
            r2c(dense, N, u[c]);
        }
        std::vector<double> itk(N);
        for (int a = 0; a < N; ++a) itk[a] = index_to_k(a, N, proj);
        std::vector<long long> ks;
        kv.clear();
        if (mode == 0) {
            ks = unique_k2(N);
            for (long long v : ks) kv.push_back((double)v);
        } else if (mode == 1) {
            for (int k = 0; k <= std::sqrt(3.) / 2. * (double)N + .5; ++k) {
                ks.push_back((long long)k * k);
                kv.push_back((double)k * k);
            }
        } else {
            const double lk = std::log(std::sqrt(3.) * N / 2.);
            for (int i = 0; i < nbins; ++i) kv.push_back(std::exp(lk * i / nbins));
        }
        const int kmax = mode == 2 ? nbins : (int)ks.size();
        spec.assign(kmax, 0.);
        const int NH = N / 2 + 1;
        for (int c = 0; c < N; ++c)
            for (int b = 0; b < N; ++b)
                for (int a = 0; a < NH; ++a) {
                    const double multpl = (a == 0 || a == N / 2) ? 1. : 2.;
                    const int abc[3] = {a, b, c};
                    const int li = a >= N / 2 ? a - N : a;
                    const int lj = b >= N / 2 ? b - N : b;
                    const int lk = c >= N / 2 ? c - N : c;
                    const unsigned sq = li * li + lj * lj + lk * lk;
                    long index;
                    if (mode == 0) {
                        index = std::lower_bound(ks.begin(), ks.end(), (long long)sq) - ks.begin();
                    } else if (mode == 1) {
                        index = (long)(std::sqrt((double)sq) + .5);
                    } else {
                        if (sq == 0) continue;
                        index = (long)std::floor((double)nbins * 0.5 * std::log((double)sq) /
                                                 std::log(std::sqrt(3.) * N / 2.));
                        index = std::max(0L, std::min((long)nbins - 1, index));
                    }
                    const size_t e = (size_t)(((long long)c * N + b) * NH + a);
                    double re[6], im[6];
                    for (int q = 0; q < 6; ++q) {
                        re[q] = u[q][e].real();
                        im[q] = u[q][e].imag();
                    }
                    // FourierSpaceModifications (gravitational_waves.cpp:96),
                    // mode by mode.
                    if (modifier >= 0) Mods::modify(modifier, re, im, li, lj, lk, dk, N, mp.data());
                    double running = 0;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            for (int l = 0; l < 3; ++l)
                                for (int m = 0; m < 3; ++m)
                                    running += gw_lambda(i, j, l, m, abc, itk.data()) * multpl *
                                               (re[mat[i][j]] * re[mat[l][m]] +
                                                im[mat[i][j]] * im[mat[l][m]]);
                    if (index >= 0 && index < kmax) spec[index] += running;
                }
        const double n6 = std::pow((double)N, 6);
        for (double &v : spec) v /= n6;
    }

    // -----------------------------------------------------------------------
    // Output functions (the project's SLEDGEHAMR_OUTPUT_FUNCTIONS).  Ghost
    // cells are filled at the level's time first (FillPatch), as they are
    // at output time after a step.
    // -----------------------------------------------------------------------

    /** Spectrum::Compute (output_types/spectrum.cpp:12-113) on level 0:
     *  spec[b] = sum over modes of the x-halved R2C transform with
     *  |k|^2 = ks[b] of mult pre |F|^2, mult = 1 on the x = 0 and x = N/2
     *  planes else 2, pre = t / (dk N^6), dk = 2 pi / L (:49-51); ks the
     *  sorted unique |k|^2 (lower_bound, :80-84). */
    void spectrum(int fn, std::vector<double> &spec, std::vector<long long> &ks) {
        using Out = typename Ops::Output;
        if (fn < 0 || fn >= Out::count) throw std::runtime_error("output function id");
        MultiFab &S = grid_new[0];
        const int N = dimN[0];
        const double time = S.t;
        fill_patch(S, 0, time);
        std::vector<double> p;
        prj.SetParamsSpectra(p, time);
        p.resize(std::max<size_t>(p.size(), 64), 0.);
        std::vector<double> dense((size_t)N * N * N, 0.);
        for (Fab &f : S.fabs) {
            const amrex::Array4<const double> sa = f.carr();
            for (int k = f.box.lo[2]; k <= f.box.hi[2]; ++k)
                for (int j = f.box.lo[1]; j <= f.box.hi[1]; ++j)
                    for (int i = f.box.lo[0]; i <= f.box.hi[0]; ++i)
                        dense[(size_t)(i + (long long)N * (j + (long long)N * k))] =
                            Out::eval(fn, sa, i, j, k, 0, time, dt[0], dx[0], p.data());
        }
        std::vector<std::complex<double>> F;
        r2c(dense, N, F);
        ks = unique_k2(N);
        spec.assign(ks.size(), 0.);
        const double pre = std::pow(1. / N, 6) * time / (2. * M_PI / cfg.L);
        const int NH = N / 2 + 1;
        for (int c = 0; c < N; ++c)
            for (int b = 0; b < N; ++b)
                for (int a = 0; a < NH; ++a) {
                    const double multpl = (a == 0 || a == N / 2) ? 1. : 2.;
                    const int li = a >= N / 2 ? a - N : a;
                    const int lj = b >= N / 2 ? b - N : b;
                    const int lk = c >= N / 2 ? c - N : c;
                    const long long sq = (long long)li * li + lj * lj + lk * lk;
                    const size_t index =
                        std::lower_bound(ks.begin(), ks.end(), sq) - ks.begin();
                    spec[index] += multpl * pre *
                                   std::norm(F[(size_t)(((long long)c * N + b) * NH + a)]);
                }
    }

    /** Projection::Compute (output_types/projection.cpp:11-118) along z onto
     *  the finest level up to max_level: every cell of level lev that the
     *  next level does not hold (no finer box contains (2i, 2j, 2k)) adds
     *  ratio * value (mode 0) or maximises (mode 1, from 0) the ratio^2
     *  pixels (i ratio + a) D + (j ratio + b), counting each contribution. */
    void projection(int fn, int mode, int max_level, std::vector<double> &data,
                    std::vector<int> &n, int &D) {
        using Out = typename Ops::Output;
        if (fn < 0 || fn >= Out::count) throw std::runtime_error("output function id");
        const int ml = std::min(max_level < 0 ? INT32_MAX : max_level, finest());
        D = dimN[ml];
        data.assign((size_t)D * D, 0.);
        n.assign((size_t)D * D, 0);
        std::vector<double> p;
        prj.SetParamsProjections(p, grid_new[0].t);
        p.resize(std::max<size_t>(p.size(), 64), 0.);
        for (int lev = 0; lev <= ml; ++lev) {
            MultiFab &S = grid_new[lev];
            const double time = S.t;
            fill_patch(S, lev, time);
            const int ratio = D / dimN[lev];
            for (Fab &f : S.fabs) {
                const amrex::Array4<const double> sa = f.carr();
                for (int k = f.box.lo[2]; k <= f.box.hi[2]; ++k)
                    for (int j = f.box.lo[1]; j <= f.box.hi[1]; ++j)
                        for (int i = f.box.lo[0]; i <= f.box.hi[0]; ++i) {
                            bool covered = false;
                            if (lev < ml)
                                for (const Fab &g : grid_new[lev + 1].fabs) {
                                    const Box &B = g.box;
                                    if (2 * i >= B.lo[0] && 2 * i <= B.hi[0] &&
                                        2 * j >= B.lo[1] && 2 * j <= B.hi[1] &&
                                        2 * k >= B.lo[2] && 2 * k <= B.hi[2])
                                        covered = true;
                                }
                            if (covered) continue;
                            const double val =
                                Out::eval(fn, sa, i, j, k, lev, time, dt[lev], dx[lev], p.data());
                            for (int ia = 0; ia < ratio; ++ia)
                                for (int ja = 0; ja < ratio; ++ja) {
                                    const size_t ind =
                                        (size_t)(i * ratio + ia) * D + (j * ratio + ja);
                                    if (mode == 0)
                                        data[ind] += ratio * val;
                                    else
                                        data[ind] = std::max(data[ind], val);
                                    n[ind] += 1;
                                }
                        }
            }
        }
    }

    // -----------------------------------------------------------------------
    // Sub-cycled stepping on a static layout: TimeStepper::Advance with
    // ScheduleRegrid, SynchronizeLevels and DoRegridIfScheduled
    // (time_stepper.cpp:90-171, 221-311, 347-397) and RegridScheduler
    // (regrid_scheduler.cpp:9-74).  A scheduled regrid computes truncation
    // errors and tags as the reference does; the box layout stays.
    // -----------------------------------------------------------------------
    struct Sched { int lowest; double t; };
    std::vector<Sched> sched;
    std::vector<double> last_regrid;
    // Order of the level operations of evolve(): 'S' a regrid scheduled by
    // lev for time t, 'H' shadow level created, 'A' lev advanced to t, 'D'
    // lev+1 averaged onto lev, 'T' truncation errors of lev (with its
    // average-down) and tags, 'R' regrid from lev at t.
    struct Event { char kind; int lev; double t; };
    std::vector<Event> events;

    int find_sched(double t) const {
        for (size_t i = 0; i < sched.size(); ++i)
            if (sched[i].t == t) return (int)i;
        return -1;
    }
    bool do_regrid(int lev, double t) const {
        const int id = find_sched(t);
        return id >= 0 && lev == sched[id].lowest;
    }
    bool need_te(int lev, double t) const {
        const int id = find_sched(t);
        return id >= 0 && lev >= sched[id].lowest;
    }
    double regrid_dt_of(int lev) const { return cfg.regrid_dt / std::pow(2., lev); }

    void schedule_regrid(int lev) {
        const double time = grid_new[lev].t;
        const int istep = grid_new[lev].istep;
        if (do_regrid(lev, time + dt[lev])) return;
        if (grid_old[lev].t == time) return;
        if (lev >= cfg.max_level) return;
        if (istep % 2 == 0 && lev > 0) return;
        const double next = lev > 0 ? time + 3. * dt[lev] : time + 2. * dt[lev];
        if (lev + 1 > 0 && !prj.CreateLevelIf(lev + 1, next)) return;
        if (next <= last_regrid[lev] + regrid_dt_of(lev) && !lr.do_global[lev]) return;
        if (lev == 0 && grid_old[lev].t == -DBL_MAX) return;
        const int id = find_sched(time + dt[lev]);
        if (id < 0)
            sched.push_back({lev, time + dt[lev]});
        else
            sched[id].lowest = std::min(lev, sched[id].lowest);
        events.push_back({'S', lev, time + dt[lev]});
        if (lev == 0) {
            create_shadow();
            events.push_back({'H', -1, shadow.t});
        }
    }

    void synchronize_levels(int lev) {
        const bool te = need_te(lev, grid_new[lev].t);
        const int finest = this->finest();
        if (lev < finest && !(cfg.shadow_hierarchy && te)) {
            average_down_to(grid_new[lev + 1], grid_new[lev]);
            events.push_back({'D', lev, grid_new[lev].t});
        }
        if (lev >= 1 - (cfg.shadow_hierarchy ? 1 : 0) && te) {
            events.push_back({'T', lev, grid_new[lev].t});
            std::vector<double> t;
            te_avgdown(lev, t, true);
            std::vector<uint8_t> m;
            std::vector<long long> c;
            tag(lev, t, true, m, c);
            last_tags[lev] = c;
            last_mask[lev] = {grid_new[lev].t, m};
            if (lev == 0) {
                shadow = MultiFab();
                shadow_tmp = MultiFab();
            }
        }
    }

    // ---------------------------------------------------------------------
    // Local regrid: local_regrid/local_regrid.cpp, unique_layout.cpp.
    // Blocks are blocking-factor cubes of a level, (i, j, k) block indices.
    // ---------------------------------------------------------------------
    using Block = std::array<int, 3>;
    using BlockSet = std::set<Block>;
    struct LocalRegridState {
        int nregrids = 0, veto_level = -1;
        std::vector<long long> last_numPts;
        std::vector<char> no_local, do_global;
        std::vector<double> min_distance, latest;
        long long global_wanted = 0, attempts = 0;
    } lr;
    // Tag mask of the last tagging per level (time, mask box-major x fastest).
    std::map<int, std::pair<double, std::vector<uint8_t>>> last_mask;

    long long num_pts(const std::vector<Box> &bl) const {
        long long n = 0;
        for (const Box &b : bl) n += b.vol();
        return n;
    }
    std::vector<Box> boxes_of(int lev) const {
        std::vector<Box> v;
        for (auto &f : grid_new[lev].fabs) v.push_back(f.box);
        return v;
    }
    static bool covers(const MultiFab &m, int i, int j, int k) {
        for (auto &f : m.fabs)
            if (i >= f.box.lo[0] && i <= f.box.hi[0] && j >= f.box.lo[1] &&
                j <= f.box.hi[1] && k >= f.box.lo[2] && k <= f.box.hi[2])
                return true;
        return false;
    }

    /** LocalRegrid::WrapIndices (local_regrid.cpp:474-490): entry c in 1..N
     *  is cell (c - 1/2) bf, inside block c-1; entries 0 and N+1 are cells of
     *  the periodic images of the last and the first block. */
    static void check_bf(int v, int n) {
        if (v < 2 || (v & (v - 1)) || n % v)
            throw std::runtime_error("amr.blocking_factor must be a power of 2 dividing "
                                     "the level size");
    }
    std::vector<int> wrapped_index(int lev) const {
        check_bf(cfg.bf[lev], dimN[lev]);
        const int bf = cfg.bf[lev], N = dimN[lev] / bf;
        std::vector<int> w(N + 2);
        w[0] = dimN[lev] - bf / 2;
        w[N + 1] = bf / 2;
        for (int i = 1; i <= N; ++i) w[i] = (int)(((double)i - 0.5) * (double)bf);
        return w;
    }

    /** LocalRegrid::DetermineNewBoxArray, TagAndMeasure and CheckBorders
     *  (local_regrid.cpp:511-680) over the tag mask of level lev (box-major,
     *  x fastest), boxes in order and cells in k, j, i order.  Per box, a
     *  border map over the 1-based wrapped block coordinates c0-1 .. c1+1 of
     *  level lev+1 (c = 2 lo / bf + 1) marks blocks whose wrapped-index cell
     *  no box of lev+1 holds.  A tagged cell ci (fine index f = 2 ci, own
     *  block cf = f / bf + 1) measures, for each of the 27 blocks cf + d that
     *  is still marked, d^2 = sum over axes of 0 if lo < f < hi, else
     *  min((f - lo)^2, (f - hi)^2), lo = bf (cf + d - 1), hi = bf (cf + d) - 1;
     *  below (n_error_buf + 1)^2 the block (wrapped) is added and unmarked
     *  for the rest of this box.  Returns the smallest measured d^2, INT_MAX
     *  when nothing was measured. */
    int select_blocks(int lev, const std::vector<uint8_t> &mask, BlockSet &add) const {
        const double threshold = (double)((cfg.n_error_buf + 1) * (cfg.n_error_buf + 1));
        const int ibff = cfg.bf[lev + 1];
        const double bff = (double)ibff;
        const std::vector<int> wi = wrapped_index(lev + 1);
        const MultiFab &fine = grid_new[lev + 1];
        int closest = INT_MAX;
        long long moff = 0;
        for (const Fab &F : grid_new[lev].fabs) {
            const Box &bx = F.box;
            int c0[3], xs[3];
            for (int a = 0; a < 3; ++a) {
                c0[a] = (int)((double)(bx.lo[a] * 2) / bff) + 1;
                xs[a] = (int)((double)(bx.hi[a] * 2) / bff) + 1 - c0[a] + 1;
            }
            const size_t sx = (size_t)xs[0] + 2, sy = (size_t)xs[1] + 2;
            std::vector<char> border(sx * sy * (size_t)(xs[2] + 2));
            auto B = [&](int i, int j, int k) -> char & {
                return border[(size_t)i + sx * ((size_t)j + sy * (size_t)k)];
            };
            int remaining = 0;
            for (int k = -1; k <= xs[2]; ++k)
                for (int j = -1; j <= xs[1]; ++j)
                    for (int i = -1; i <= xs[0]; ++i) {
                        B(i + 1, j + 1, k + 1) =
                            !covers(fine, wi[c0[0] + i], wi[c0[1] + j], wi[c0[2] + k]);
                        remaining += B(i + 1, j + 1, k + 1);
                    }
            const int nx = bx.n(0), ny = bx.n(1);
            for (int k = bx.lo[2]; k <= bx.hi[2] && remaining > 0; ++k)
                for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
                    for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
                        if (!mask[(size_t)(moff + (i - bx.lo[0]) +
                                           (long long)nx * ((j - bx.lo[1]) +
                                                            (long long)ny * (k - bx.lo[2])))])
                            continue;
                        const int fi[3] = {2 * i, 2 * j, 2 * k};
                        int cfi[3];
                        for (int a = 0; a < 3; ++a) cfi[a] = (int)((double)fi[a] / bff) + 1;
                        for (int kk = -1; kk <= 1; ++kk)
                            for (int jj = -1; jj <= 1; ++jj)
                                for (int ii = -1; ii <= 1; ++ii) {
                                    const int d[3] = {ii, jj, kk};
                                    int ref[3];
                                    for (int a = 0; a < 3; ++a) ref[a] = cfi[a] - c0[a] + d[a] + 1;
                                    char &b = B(ref[0], ref[1], ref[2]);
                                    if (!b) continue;
                                    int d_sq = 0;
                                    for (int a = 0; a < 3; ++a) {
                                        const int lo = ibff * (cfi[a] + d[a] - 1);
                                        const int hi = ibff * (cfi[a] + d[a]) - 1;
                                        d_sq += (lo < fi[a] && fi[a] < hi)
                                                    ? 0
                                                    : std::min((fi[a] - lo) * (fi[a] - lo),
                                                               (fi[a] - hi) * (fi[a] - hi));
                                    }
                                    closest = std::min(closest, d_sq);
                                    if (d_sq < threshold) {
                                        b = 0;
                                        Block nb;
                                        for (int a = 0; a < 3; ++a)
                                            nb[a] = (int)((double)wi[cfi[a] + d[a]] / bff - 0.5);
                                        add.insert(nb);
                                    }
                                }
                    }
            moff += bx.vol();
        }
        return closest;
    }

    /** UniqueLayout::BoxList (unique_layout.cpp:133-170): per block row
     *  (i, j), ascending, one box per run of consecutive k. */
    static std::vector<Box> block_boxes(const BlockSet &s, int bf) {
        std::vector<Box> out;
        for (auto it = s.begin(); it != s.end();) {
            const Block b0 = *it;
            int km = b0[2];
            ++it;
            while (it != s.end() && (*it)[0] == b0[0] && (*it)[1] == b0[1] &&
                   (*it)[2] == km + 1) {
                km = (*it)[2];
                ++it;
            }
            out.push_back(Box{{b0[0] * bf, b0[1] * bf, b0[2] * bf},
                              {(b0[0] + 1) * bf - 1, (b0[1] + 1) * bf - 1, (km + 1) * bf - 1}});
        }
        return out;
    }

    /** Joins runs into larger boxes (the reference leaves this to AMReX's
     *  BoxList::simplify, local_regrid.cpp:67-76): boxes with the same x and
     *  z ranges that touch along y merge, then boxes with the same y and z
     *  ranges that touch along x; the result is sorted by (z, y, x) of lo. */
    static std::vector<Box> merge_boxes(std::vector<Box> v) {
        auto merge = [&](int a, int o1, int o2) {
            std::sort(v.begin(), v.end(), [&](const Box &p, const Box &q) {
                return std::make_tuple(p.lo[o1], p.hi[o1], p.lo[o2], p.hi[o2], p.lo[a]) <
                       std::make_tuple(q.lo[o1], q.hi[o1], q.lo[o2], q.hi[o2], q.lo[a]);
            });
            std::vector<Box> out;
            for (const Box &b : v) {
                if (!out.empty()) {
                    Box &p = out.back();
                    if (p.lo[o1] == b.lo[o1] && p.hi[o1] == b.hi[o1] && p.lo[o2] == b.lo[o2] &&
                        p.hi[o2] == b.hi[o2] && p.hi[a] + 1 == b.lo[a]) {
                        p.hi[a] = b.hi[a];
                        continue;
                    }
                }
                out.push_back(b);
            }
            v = out;
        };
        merge(1, 0, 2);
        merge(0, 1, 2);
        std::sort(v.begin(), v.end(), [](const Box &p, const Box &q) {
            return std::make_tuple(p.lo[2], p.lo[1], p.lo[0]) <
                   std::make_tuple(q.lo[2], q.lo[1], q.lo[0]);
        });
        return v;
    }

    /** LocalRegrid::FixNesting (local_regrid.cpp:687-729): the runs added to
     *  level l, grown by nghost + 4 cells, coarsened by 2 (floor), cut at
     *  the periodic images of the domain (WrapBoxArray, :735-757); for each
     *  piece, wrapped block coordinates lo / bf + 1 .. hi / bf + 2 of level
     *  l-1 whose cell no box of l-1 holds are added to l-1. */
    void fix_nesting(int l, std::vector<BlockSet> &add) const {
        const std::vector<Box> runs = block_boxes(add[l], cfg.bf[l]);
        const int g = cfg.nghost + 4, Nc = dimN[l - 1];
        auto floordiv2 = [](int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); };
        std::vector<Box> pieces;
        for (Box b : runs) {
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = floordiv2(b.lo[a] - g);
                b.hi[a] = floordiv2(b.hi[a] + g);
            }
            for (int i = -1; i <= 1; ++i)
                for (int j = -1; j <= 1; ++j)
                    for (int k = -1; k <= 1; ++k) {
                        const int sh[3] = {i * Nc, j * Nc, k * Nc};
                        Box p;
                        bool empty = false;
                        for (int a = 0; a < 3; ++a) {
                            p.lo[a] = std::max(b.lo[a], sh[a]) - sh[a];
                            p.hi[a] = std::min(b.hi[a], sh[a] + Nc - 1) - sh[a];
                            if (p.hi[a] < p.lo[a]) empty = true;
                        }
                        if (!empty) pieces.push_back(p);
                    }
        }
        const double bfc = (double)cfg.bf[l - 1];
        const std::vector<int> wi = wrapped_index(l - 1);
        for (const Box &p : pieces) {
            int cl[3], ch[3];
            for (int a = 0; a < 3; ++a) {
                cl[a] = (int)((double)p.lo[a] / bfc + 1.);
                ch[a] = (int)((double)p.hi[a] / bfc + 2.);
            }
            for (int cx = cl[0]; cx <= ch[0]; ++cx)
                for (int cy = cl[1]; cy <= ch[1]; ++cy)
                    for (int cz = cl[2]; cz <= ch[2]; ++cz)
                        if (!covers(grid_new[l - 1], wi[cx], wi[cy], wi[cz]))
                            add[l - 1].insert(Block{(int)((double)wi[cx] / bfc - 0.5),
                                                    (int)((double)wi[cy] / bfc - 0.5),
                                                    (int)((double)wi[cz] / bfc - 0.5)});
        }
    }

    /** Valid cells of `boxes` as FillPatch (level_synchronizer.cpp:92-209)
     *  fills them at `time` from level lev's data `src` and level lev-1:
     *  coarse data interpolated in time and space, then src where it holds
     *  the cell. */
    MultiFab fill_patch_boxes(int lev, const std::vector<Box> &boxes, const MultiFab &src,
                              double time) {
        MultiFab m;
        m.define(boxes, nc, cfg.nghost);
        m.t = time;
        fill_coarse_ghosts(m, lev, time, true, false);
        for (Fab &D : m.fabs)
            for (const Fab &S : src.fabs) {
                int lo[3], hi[3];
                bool empty = false;
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::max(D.box.lo[a], S.box.lo[a]);
                    hi[a] = std::min(D.box.hi[a], S.box.hi[a]);
                    if (hi[a] < lo[a]) empty = true;
                }
                if (empty) continue;
                for (int n = 0; n < nc; ++n)
                    for (int k = lo[2]; k <= hi[2]; ++k)
                        for (int j = lo[1]; j <= hi[1]; ++j)
                            for (int i = lo[0]; i <= hi[0]; ++i)
                                D.at(i, j, k, n) = const_cast<Fab &>(S).at(i, j, k, n);
            }
        return m;
    }

    /** LocalRegrid::AddBoxes (local_regrid.cpp:763-828): the level keeps its
     *  boxes and data and gains `nb` after them, filled by FillPatch at the
     *  new and at the old time (the old state from the new one when the
     *  level has no old time yet). */
    void add_boxes(int lev, const std::vector<Box> &nb) {
        if (lev < 1) throw std::runtime_error("boxes are added to levels above 0");
        if (nb.empty()) return;
        MultiFab &New = grid_new[lev], &Old = grid_old[lev];
        const MultiFab fn = fill_patch_boxes(lev, nb, New, New.t);
        const MultiFab fo =
            Old.t == -DBL_MAX ? fn : fill_patch_boxes(lev, nb, Old, Old.t);
        for (size_t b = 0; b < nb.size(); ++b) {
            New.fabs.push_back(fn.fabs[b]);
            Old.fabs.push_back(fo.fabs[b]);
        }
    }

    /** Sledgehamr::RemakeLevel (sledgehamr.cpp:173-190): level lev on new
     *  boxes, its state FillPatch-ed at the level's time from the level's
     *  data and level lev-1.  The old state, which the reference leaves
     *  undefined, is a copy of the new one. */
    void remake_level(int lev, const std::vector<Box> &boxes) {
        if (lev < 1) throw std::runtime_error("levels above 0 are remade");
        MultiFab &New = grid_new[lev], &Old = grid_old[lev];
        MultiFab fn = fill_patch_boxes(lev, boxes, New, New.t);
        fn.istep = New.istep;
        MultiFab fo = fn;
        fo.t = Old.t;
        fo.istep = Old.istep;
        New = fn;
        Old = fo;
        did_global_regrid(lev - 1);
    }

    /** LocalRegrid::DidGlobalRegrid (local_regrid.cpp:34-46). */
    void did_global_regrid(int lev) {
        std::fill(lr.no_local.begin(), lr.no_local.end(), 0);
        std::fill(lr.do_global.begin(), lr.do_global.end(), 0);
        lr.nregrids = 0;
        for (int l = lev + 1; l < (int)lr.last_numPts.size(); ++l)
            lr.last_numPts[l] = num_pts(boxes_of(l));
    }

    /** Tags ErrorEst would hand DetermineNewBoxArray for level l: none at
     *  t_start with a shadow hierarchy or when the project declines level
     *  l+1 (sledgehamr.cpp:207-230), else the mask of the level's tagging at
     *  its current time. */
    std::vector<uint8_t> regrid_mask(int l) {
        const double t = grid_new[l].t;
        long long n = num_pts(boxes_of(l));
        if ((t == cfg.t_start && cfg.shadow_hierarchy) || !prj.CreateLevelIf(l + 1, t))
            return std::vector<uint8_t>((size_t)n, 0);
        auto it = last_mask.find(l);
        if (it == last_mask.end() || it->second.first != t || (long long)it->second.second.size() != n)
            throw std::runtime_error("no tags of level " + std::to_string(l) + " at its time");
        return it->second.second;
    }

    /** LocalRegrid::DoAttemptRegrid (local_regrid.cpp:407-440) with its
     *  Prechecks (:106-145), DetermineAllBoxArrays (:163-172), FixAllNesting
     *  (:176-182), CheckForVeto / CheckThresholds / ComputeLatestPossible-
     *  RegridTime (:214-322) and DealWithVeto (:331-400).  Returns 0 when a
     *  global regrid is wanted, 1 after adding boxes (possibly none), 2 when
     *  the regrid is delayed. */
    int local_regrid(int lev) {
        ++lr.attempts;
        const int finest = this->finest();
        if (lr.nregrids++ >= cfg.max_local_regrids) return 0;
        if (cfg.vt_weak <= 1.) {
            lr.veto_level = lev - 1;
            return 0;
        }
        lr.veto_level = -1;
        if (lr.do_global[lev]) return 0;
        if (lev == finest) return 0;
        while ((int)lr.last_numPts.size() <= finest)
            lr.last_numPts.push_back(num_pts(boxes_of((int)lr.last_numPts.size())));
        std::vector<BlockSet> add(finest + 1);
        lr.min_distance.assign(finest + 1, -1.);
        for (int l = lev; l < finest && !lr.no_local[l]; ++l) {
            const int d = select_blocks(l, regrid_mask(l), add[l + 1]);
            lr.min_distance[l + 1] = d == INT_MAX ? -1. : std::sqrt((double)d);
        }
        for (int l = finest; l > 1; --l) fix_nesting(l, add);
        std::vector<std::vector<Box>> ba(finest + 1);
        for (int l = 1; l <= finest; ++l) ba[l] = merge_boxes(block_boxes(add[l], cfg.bf[l]));
        bool veto = false;
        lr.latest.assign(finest + 1, -1.);
        for (int l = 1; l <= finest; ++l) {
            if (!veto) {
                const double Nb = (double)num_pts(ba[l]);
                const double Nc = (double)num_pts(boxes_of(l));
                double Nr = (double)lr.last_numPts[l];
                if (Nr == 0) Nr = Nc;
                const double fV = (Nb + Nc) / Nr;
                veto = fV > cfg.vt_weak;
                if (fV > cfg.vt_strong && lr.veto_level == -1) lr.veto_level = l - 1;
            }
            if (l > lev) {
                double delay = DBL_MAX;
                if (lr.min_distance[l] >= 0)
                    delay = lr.min_distance[l] / (double)cfg.n_error_buf * regrid_dt_of(l);
                lr.latest[l] = grid_new[l].t + delay;
            }
        }
        if (veto) {
            const int vl = lr.veto_level;
            if (vl >= lev || vl <= 0) return 0;
            lr.do_global[vl] = 1;
            double nsteps = grid_new[vl].istep % 2 == 0 ? 1 : 2;
            if (cfg.shadow_hierarchy) nsteps = 0;
            const double target = grid_new[vl].t + nsteps * dt[vl];
            bool wait = true;
            for (int l = lev + 1; l <= finest; ++l)
                if (lr.latest[l] < target) {
                    wait = false;
                    break;
                }
            if (wait) {
                for (int l = lev; l <= finest; ++l) lr.no_local[l] = 1;
                return 2;
            }
        }
        for (int l = 1; l <= finest; ++l)
            if (!ba[l].empty()) add_boxes(l, ba[l]);
        return 1;
    }

    void do_regrid_if_scheduled(int lev) {
        const double time = grid_new[lev].t;
        if (!do_regrid(lev, time)) return;
        events.push_back({'R', lev, time});
        const bool global = cfg.local_regrid && local_regrid(lev) == 0;
        for (int l = lev; l <= finest(); ++l) last_regrid[l] = time;
        const int id = find_sched(time);
        if (id >= 0) sched.erase(sched.begin() + id);
        if (global) {
            // A global regrid (AMReX clustering) is not restated: stop, or
            // keep the layout with the counters as they are when asked to.
            ++lr.global_wanted;
            if (!cfg.keep_layout)
                throw std::runtime_error("a global regrid is wanted");
        }
    }

    void step(int lev) {
        if (cfg.shadow_hierarchy) schedule_regrid(lev);
        if (lev == 0 && !shadow.defined()) prj.BeforeTimestep(grid_new[lev].t);
        advance(lev);
        events.push_back({'A', lev, grid_new[lev].t});
        if (lev != finest()) {
            step(lev + 1);
            step(lev + 1);
        }
        synchronize_levels(lev);
        do_regrid_if_scheduled(lev);
        if (lev == 0)
            for (int l = 1; l <= finest(); ++l) grid_new[l].t = grid_new[0].t;
    }

    void evolve(int ncoarse) {
        if (last_regrid.empty()) last_regrid.assign(cfg.max_level + 1, cfg.t_start);
        for (int n = 0; n < ncoarse; ++n) step(0);
    }

    std::map<int, std::vector<long long>> last_tags;  // counters of the last TE per level

    /** Seeded bubbles (kind 3: over vacuum, nucleated by p[7]; kind 4: added
     *  to the current field, nucleated in (p[6], p[7]]) and closed string
     *  loops (kind 5), as the device generator defines them: objects drawn
     *  from streams 100+ / 200+ in level-0 cell units, periodic minimum
     *  image, bubble wall phi_in (1 - tanh((r - R0)/w))/2 combined by max,
     *  loops psi = prod tanh(d/w) exp(i atan2(h, rho - R)). */
    void init_objects(int lev, int kind, uint64_t seed, const double *p, int np) {
        MultiFab &m = grid_new[lev];
        const double N0 = (double)cfg.N0;
        const double cell = N0 / (double)dimN[lev];
        struct Obj { double v[8]; };
        std::vector<Obj> obj;
        const int n = np > 0 ? (int)p[0] : 0;
        for (int i = 0; i < n; ++i) {
            Obj o{};
            if (kind == 3 || kind == 4) {
                o.v[0] = N0 * rng_uniform(seed, 100, i);
                o.v[1] = N0 * rng_uniform(seed, 101, i);
                o.v[2] = N0 * rng_uniform(seed, 102, i);
                o.v[3] = p[4] + (p[5] - p[4]) * rng_uniform(seed, 103, i);
            } else {
                const double u = 2. * rng_uniform(seed, 204, i) - 1.;
                const double ph = 2. * M_PI * rng_uniform(seed, 205, i);
                const double st = std::sqrt(std::max(0., 1. - u * u));
                o.v[0] = N0 * rng_uniform(seed, 200, i);
                o.v[1] = N0 * rng_uniform(seed, 201, i);
                o.v[2] = N0 * rng_uniform(seed, 202, i);
                o.v[3] = p[1] + (p[2] - p[1]) * rng_uniform(seed, 203, i);
                o.v[4] = st * std::cos(ph);
                o.v[5] = st * std::sin(ph);
                o.v[6] = u;
            }
            obj.push_back(o);
        }
        auto mi = [N0](double d) { return d - N0 * std::floor(d / N0 + 0.5); };
        for (auto &f : m.fabs)
#pragma omp parallel for schedule(static)
            for (int k = f.box.lo[2]; k <= f.box.hi[2]; ++k)
                for (int j = f.box.lo[1]; j <= f.box.hi[1]; ++j)
                    for (int i = f.box.lo[0]; i <= f.box.hi[0]; ++i) {
                        const double x = (i + 0.5) * cell, y = (j + 0.5) * cell,
                                     z = (k + 0.5) * cell;
                        if (kind == 3 || kind == 4) {
                            double phi = kind == 3 ? 0. : f.at(i, j, k, 0);
                            for (const Obj &o : obj) {
                                const bool now = kind == 3 ? o.v[3] <= p[7]
                                                           : (o.v[3] > p[6] && o.v[3] <= p[7]);
                                if (!now) continue;
                                const double dx = mi(x - o.v[0]), dy = mi(y - o.v[1]),
                                             dz = mi(z - o.v[2]);
                                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                                phi = std::max(phi, p[1] * 0.5 * (1. - std::tanh((r - p[2]) / p[3])));
                            }
                            if (kind == 3)
                                for (int c = 0; c < nc; ++c) f.at(i, j, k, c) = 0.;
                            f.at(i, j, k, 0) = phi;
                        } else {
                            double mod = 1., th = 0.;
                            for (const Obj &o : obj) {
                                const double dx = mi(x - o.v[0]), dy = mi(y - o.v[1]),
                                             dz = mi(z - o.v[2]);
                                const double h = dx * o.v[4] + dy * o.v[5] + dz * o.v[6];
                                const double px = dx - h * o.v[4], py = dy - h * o.v[5],
                                             pz = dz - h * o.v[6];
                                const double sr = std::sqrt(px * px + py * py + pz * pz) - o.v[3];
                                mod *= std::tanh(std::sqrt(sr * sr + h * h) / p[3]);
                                th += std::atan2(h, sr);
                            }
                            for (int c = 0; c < nc; ++c) f.at(i, j, k, c) = 0.;
                            f.at(i, j, k, 0) = mod * std::cos(th);
                            f.at(i, j, k, 1) = mod * std::sin(th);
                        }
                    }
    }

    /** Seeded complex random phase, identical to the device generator. */
    void init_complex_random_phase(int lev, uint64_t seed, double sigma) {
        MultiFab &m = grid_new[lev];
        const int N = dimN[lev];
        for (auto &f : m.fabs)
#pragma omp parallel for schedule(static)
            for (int k = f.box.lo[2]; k <= f.box.hi[2]; ++k)
                for (int j = f.box.lo[1]; j <= f.box.hi[1]; ++j)
                    for (int i = f.box.lo[0]; i <= f.box.hi[0]; ++i) {
                        const uint64_t g = (uint64_t)i + (uint64_t)N * ((uint64_t)j + (uint64_t)N * (uint64_t)k);
                        const double u1 = rng_uniform(seed, 1, g);
                        const double u2 = rng_uniform(seed, 2, g);
                        const double u3 = rng_uniform(seed, 3, g);
                        const double theta = 2. * M_PI * u1;
                        const double gg = std::sqrt(-2. * std::log(1. - u2)) * std::cos(2. * M_PI * u3);
                        const double mm = 1. + sigma * gg;
                        for (int n = 0; n < nc; ++n) f.at(i, j, k, n) = 0.;
                        f.at(i, j, k, 0) = mm * std::cos(theta);
                        f.at(i, j, k, 1) = mm * std::sin(theta);
                    }
    }
};

}  // namespace orc

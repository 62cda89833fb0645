// ORACLE -- test infrastructure only (see oracle.hpp).
// C entry points of the FP64 oracle for one project, mirroring the product's
// C ABI so tests drive both with the same call sequence.
#pragma once

#include <cstring>
#include <string>

#include "oracle.hpp"

#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

#define ORC_DEFINE_API(Proj)                                                   \
    using OrcSim = orc::Sim<Proj>;                                             \
    static thread_local std::string orc_err;                                   \
    template <class F>                                                         \
    static int orc_guard(F &&f) {                                              \
        try {                                                                  \
            f();                                                               \
            return 0;                                                          \
        } catch (const std::exception &e) {                                    \
            orc_err = e.what();                                                \
            return 1;                                                          \
        }                                                                      \
    }                                                                          \
    static orc::MultiFab &orc_level(OrcSim *s, int lev);                       \
    static long long orc_cells(OrcSim *s, int lev) {                          \
        long long n = 0;                                                       \
        for (auto &f : orc_level(s, lev).fabs) n += f.box.vol();              \
        return n;                                                              \
    }                                                                          \
    template <class F>                                                         \
    static void orc_valid(orc::MultiFab &m, F &&f) {                           \
        long long o = 0;                                                       \
        for (auto &fb : m.fabs)                                                \
            for (int k = fb.box.lo[2]; k <= fb.box.hi[2]; ++k)                 \
                for (int j = fb.box.lo[1]; j <= fb.box.hi[1]; ++j)             \
                    for (int i = fb.box.lo[0]; i <= fb.box.hi[0]; ++i)         \
                        f(fb, i, j, k, o++);                                   \
    }                                                                          \
    static orc::MultiFab &orc_level(OrcSim *s, int lev) {                      \
        if (lev >= 0) return s->grid_new[lev];                                 \
        if (!s->shadow.defined()) {                                            \
            std::vector<orc::Box> cb;                                          \
            for (auto &f : s->grid_new[0].fabs) {                              \
                orc::Box b;                                                    \
                for (int a = 0; a < 3; ++a) {                                  \
                    b.lo[a] = f.box.lo[a] / 2;                                 \
                    b.hi[a] = (f.box.hi[a] + 1) / 2 - 1;                       \
                }                                                              \
                cb.push_back(b);                                               \
            }                                                                  \
            s->shadow.define(cb, s->nc, s->cfg.nghost);                        \
            s->shadow.t = s->grid_new[0].t;                                    \
        }                                                                      \
        return s->shadow;                                                      \
    }                                                                          \
    ORC_EXPORT const char *orc_last_error() { return orc_err.c_str(); }        \
    ORC_EXPORT void *orc_create(const char *inputs) {                          \
        try {                                                                  \
            return new OrcSim(inputs);                                         \
        } catch (const std::exception &e) {                                    \
            orc_err = e.what();                                                \
            return nullptr;                                                    \
        }                                                                      \
    }                                                                          \
    ORC_EXPORT void orc_destroy(void *h) { delete static_cast<OrcSim *>(h); }  \
    ORC_EXPORT int orc_ncomp(void *h) { return static_cast<OrcSim *>(h)->nc; } \
    ORC_EXPORT long long orc_ncells(void *h, int lev) {                        \
        return orc_cells(static_cast<OrcSim *>(h), lev);                       \
    }                                                                          \
    ORC_EXPORT int orc_set_boxes(void *h, int lev, const int *b, int nb) {    \
        return orc_guard([&] {                                                 \
            std::vector<orc::Box> v(nb);                                       \
            for (int i = 0; i < nb; ++i)                                       \
                for (int a = 0; a < 3; ++a) {                                  \
                    v[i].lo[a] = b[6 * i + a];                                 \
                    v[i].hi[a] = b[6 * i + 3 + a];                             \
                }                                                              \
            static_cast<OrcSim *>(h)->set_level(lev, v);                       \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_set_time(void *h, int lev, double t, int istep) {       \
        auto *s = static_cast<OrcSim *>(h);                                    \
        orc_level(s, lev).t = t;                                               \
        orc_level(s, lev).istep = istep;                                       \
        return 0;                                                              \
    }                                                                          \
    ORC_EXPORT double orc_time(void *h, int lev) {                             \
        return orc_level(static_cast<OrcSim *>(h), lev).t;                     \
    }                                                                          \
    ORC_EXPORT double orc_index_to_k(int a, int N, int type) {                 \
        return orc::index_to_k(a, N, type);                                    \
    }                                                                          \
    ORC_EXPORT int orc_upload(void *h, int lev, int comp, const double *d) {   \
        return orc_guard([&] {                                                 \
            orc_valid(orc_level(static_cast<OrcSim *>(h), lev),                \
                      [&](orc::Fab &f, int i, int j, int k, long long o) {     \
                          f.at(i, j, k, comp) = d[o];                          \
                      });                                                      \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_upload_old(void *h, int lev, int comp, const double *d, \
                                  double t) {                                  \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            orc_valid(s->grid_old[lev],                                        \
                      [&](orc::Fab &f, int i, int j, int k, long long o) {     \
                          f.at(i, j, k, comp) = d[o];                          \
                      });                                                      \
            s->grid_old[lev].t = t;                                            \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_fill_coarse_patch(void *h, int lev) {                   \
        return orc_guard(                                                      \
            [&] { static_cast<OrcSim *>(h)->fill_coarse_patch(lev); });        \
    }                                                                          \
    ORC_EXPORT int orc_fill_patch(void *h, int lev, double time) {             \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            s->fill_patch(s->grid_new[lev], lev, time);                        \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_download_fab(void *h, int lev, int box, int comp,       \
                                    double *d) {                               \
        return orc_guard([&] {                                                 \
            orc::Fab &f = static_cast<OrcSim *>(h)->grid_new[lev].fabs.at(box);\
            std::memcpy(d, f.d.data() + comp * f.nst, sizeof(double) * f.nst); \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_spectrum(void *h, int fn, double *bins, long long *ks,  \
                                int *n) {                                      \
        return orc_guard([&] {                                                 \
            std::vector<double> b;                                             \
            std::vector<long long> k;                                          \
            static_cast<OrcSim *>(h)->spectrum(fn, b, k);                      \
            if (bins) {                                                        \
                std::memcpy(bins, b.data(), sizeof(double) * b.size());        \
                std::memcpy(ks, k.data(), sizeof(long long) * k.size());       \
            }                                                                  \
            *n = (int)b.size();                                                \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_projection(void *h, int fn, int mode, int ml, double *d,\
                                  int *c, int *D) {                            \
        return orc_guard([&] {                                                 \
            std::vector<double> dv;                                            \
            std::vector<int> cv;                                               \
            static_cast<OrcSim *>(h)->projection(fn, mode, ml, dv, cv, *D);    \
            if (d) {                                                           \
                std::memcpy(d, dv.data(), sizeof(double) * dv.size());         \
                std::memcpy(c, cv.data(), sizeof(int) * cv.size());            \
            }                                                                  \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_download(void *h, int lev, int comp, double *d,         \
                                int old) {                                     \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            orc_valid(old ? s->grid_old[lev] : orc_level(s, lev),              \
                      [&](orc::Fab &f, int i, int j, int k, long long o) {     \
                          d[o] = f.at(i, j, k, comp);                          \
                      });                                                      \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_fill_rhs(void *h, int lev, double time, double w,       \
                                int add) {                                     \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            if (!add || !s->rhs_out.defined())                                 \
                s->rhs_out.define(OrcSim::boxes_of(s->grid_new[lev]), s->nc,   \
                                  s->cfg.nghost);                              \
            s->fill_patch(s->grid_new[lev], lev, s->grid_new[lev].t);          \
            s->fill_rhs(s->rhs_out, s->grid_new[lev], time, lev, s->dt[lev],   \
                        s->dx[lev], add != 0, w);                              \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_download_rhs(void *h, int comp, double *d) {            \
        return orc_guard([&] {                                                 \
            orc_valid(static_cast<OrcSim *>(h)->rhs_out,                       \
                      [&](orc::Fab &f, int i, int j, int k,                    \
                          long long o) { d[o] = f.at(i, j, k, comp); });       \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_advance(void *h, int lev) {                             \
        return orc_guard([&] { static_cast<OrcSim *>(h)->advance(lev); });    \
    }                                                                          \
    ORC_EXPORT int orc_evolve(void *h, int ncoarse) {                          \
        return orc_guard([&] { static_cast<OrcSim *>(h)->evolve(ncoarse); });  \
    }                                                                          \
    ORC_EXPORT int orc_last_tags(void *h, int lev, long long *counts, int n) { \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            auto it = s->last_tags.find(lev);                                  \
            if (it == s->last_tags.end())                                      \
                throw std::runtime_error("no truncation errors at this level"); \
            for (int i = 0; i < n && i < (int)it->second.size(); ++i)          \
                counts[i] = it->second[i];                                     \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_events(void *h, char *kind, int *lev, double *t,       \
                              int maxn, int *n) {                              \
        return orc_guard([&] {                                                 \
            const auto &ev = static_cast<OrcSim *>(h)->events;                 \
            *n = (int)ev.size();                                               \
            for (int i = 0; i < (int)ev.size() && i < maxn; ++i) {             \
                kind[i] = ev[i].kind;                                          \
                lev[i] = ev[i].lev;                                            \
                t[i] = ev[i].t;                                                \
            }                                                                  \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_create_shadow(void *h) {                                \
        return orc_guard([&] { static_cast<OrcSim *>(h)->create_shadow(); });  \
    }                                                                          \
    ORC_EXPORT double orc_shadow_time(void *h) {                               \
        return static_cast<OrcSim *>(h)->shadow.t;                             \
    }                                                                          \
    ORC_EXPORT int orc_te_tag(void *h, int lev, int flags, unsigned char *mk,  \
                              long long *counts, double *te) {                 \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            std::vector<double> t;                                             \
            if (flags & 1) s->te_avgdown(lev, t, (flags & 2) != 0);            \
            std::vector<uint8_t> m;                                            \
            std::vector<long long> c;                                          \
            s->tag(lev, t, (flags & 1) != 0, m, c);                            \
            if (mk) std::memcpy(mk, m.data(), m.size());                       \
            if (counts) std::memcpy(counts, c.data(), c.size() * 8);           \
            if (te && !t.empty()) std::memcpy(te, t.data(), t.size() * 8);     \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_boxes(void *h, int lev, int *b, int maxbox, int *nbox) { \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            const auto &f = s->grid_new[lev].fabs;                             \
            *nbox = (int)f.size();                                             \
            if (b && (int)f.size() <= maxbox)                                  \
                for (size_t i = 0; i < f.size(); ++i)                          \
                    for (int a = 0; a < 3; ++a) {                              \
                        b[6 * i + a] = f[i].box.lo[a];                         \
                        b[6 * i + 3 + a] = f[i].box.hi[a];                     \
                    }                                                          \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_regrid_select(void *h, int lev, const unsigned char *mask, \
                                     int *blocks, int maxblock, int *nblock,   \
                                     double *dist) {                           \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            std::vector<uint8_t> m;                                            \
            if (mask)                                                          \
                m.assign(mask, mask + orc_cells(s, lev));                      \
            else                                                               \
                m = s->regrid_mask(lev);                                       \
            typename OrcSim::BlockSet add;                                     \
            const int d = s->select_blocks(lev, m, add);                       \
            *dist = d == INT_MAX ? -1. : std::sqrt((double)d);                 \
            *nblock = (int)add.size();                                         \
            int i = 0;                                                         \
            if (blocks && (int)add.size() <= maxblock)                         \
                for (const auto &b : add) {                                    \
                    for (int a = 0; a < 3; ++a) blocks[3 * i + a] = b[a];      \
                    ++i;                                                       \
                }                                                              \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_fix_nesting(void *h, int l, const int *blocks, int nb,  \
                                   int *out, int maxout, int *nout) {          \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            std::vector<typename OrcSim::BlockSet> add(l + 1);                 \
            for (int i = 0; i < nb; ++i)                                       \
                add[l].insert({blocks[3 * i], blocks[3 * i + 1], blocks[3 * i + 2]}); \
            s->fix_nesting(l, add);                                            \
            *nout = (int)add[l - 1].size();                                    \
            int i = 0;                                                         \
            if (out && (int)add[l - 1].size() <= maxout)                       \
                for (const auto &b : add[l - 1]) {                             \
                    for (int a = 0; a < 3; ++a) out[3 * i + a] = b[a];         \
                    ++i;                                                       \
                }                                                              \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_block_layout(const int *blocks, int nb, int bf, int *out, \
                                    int maxout, int *nout) {                   \
        return orc_guard([&] {                                                 \
            typename OrcSim::BlockSet set;                                     \
            for (int i = 0; i < nb; ++i)                                       \
                set.insert({blocks[3 * i], blocks[3 * i + 1], blocks[3 * i + 2]}); \
            const auto v = OrcSim::merge_boxes(OrcSim::block_boxes(set, bf));  \
            *nout = (int)v.size();                                             \
            if (out && (int)v.size() <= maxout)                                \
                for (size_t i = 0; i < v.size(); ++i)                          \
                    for (int a = 0; a < 3; ++a) {                              \
                        out[6 * i + a] = v[i].lo[a];                           \
                        out[6 * i + 3 + a] = v[i].hi[a];                       \
                    }                                                          \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_add_boxes(void *h, int lev, const int *b, int nb) {     \
        return orc_guard([&] {                                                 \
            std::vector<orc::Box> v(nb);                                       \
            for (int i = 0; i < nb; ++i)                                       \
                for (int a = 0; a < 3; ++a) {                                  \
                    v[i].lo[a] = b[6 * i + a];                                 \
                    v[i].hi[a] = b[6 * i + 3 + a];                             \
                }                                                              \
            static_cast<OrcSim *>(h)->add_boxes(lev, v);                       \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_remake_level(void *h, int lev, const int *b, int nb) {  \
        return orc_guard([&] {                                                 \
            std::vector<orc::Box> v(nb);                                       \
            for (int i = 0; i < nb; ++i)                                       \
                for (int a = 0; a < 3; ++a) {                                  \
                    v[i].lo[a] = b[6 * i + a];                                 \
                    v[i].hi[a] = b[6 * i + 3 + a];                             \
                }                                                              \
            static_cast<OrcSim *>(h)->remake_level(lev, v);                    \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_local_regrid(void *h, int lev, int *outcome) {          \
        return orc_guard([&] { *outcome = static_cast<OrcSim *>(h)->local_regrid(lev); }); \
    }                                                                          \
    ORC_EXPORT int orc_regrid_stats(void *h, long long *st, double *md, int nlev) { \
        return orc_guard([&] {                                                 \
            auto &r = static_cast<OrcSim *>(h)->lr;                            \
            st[0] = r.attempts;                                                \
            st[1] = r.global_wanted;                                           \
            st[2] = r.nregrids;                                                \
            st[3] = r.veto_level;                                              \
            for (int l = 0; l < nlev; ++l)                                     \
                md[l] = l < (int)r.min_distance.size() ? r.min_distance[l] : -1.; \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_set_tags(void *h, int lev, const unsigned char *mask) { \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            s->last_mask[lev] = {s->grid_new[lev].t,                           \
                                 std::vector<uint8_t>(mask, mask + orc_cells(s, lev))}; \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_gw_spectrum(void *h, int mode, int nbins, int zp,       \
                                   int proj, double *bins, double *k,          \
                                   int *nout) {                                \
        return orc_guard([&] {                                                 \
            std::vector<double> b, kv;                                         \
            static_cast<OrcSim *>(h)->gw_spectrum(mode, nbins, zp, proj, b, kv); \
            if (bins && *nout >= (int)b.size()) {                              \
                std::memcpy(bins, b.data(), b.size() * 8);                     \
                std::memcpy(k, kv.data(), kv.size() * 8);                      \
            }                                                                  \
            *nout = (int)b.size();                                             \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_gw_spectrum_mod(void *h, int mode, int nbins, int zp,   \
                                       int proj, int modifier,                 \
                                       const double *mp, int nmp,              \
                                       double *bins, double *k, int *nout) {   \
        return orc_guard([&] {                                                 \
            std::vector<double> b, kv, pv(mp, mp + (mp ? nmp : 0));            \
            static_cast<OrcSim *>(h)->gw_spectrum(mode, nbins, zp, proj, b, kv, \
                                                  modifier, pv);               \
            if (bins && *nout >= (int)b.size()) {                              \
                std::memcpy(bins, b.data(), b.size() * 8);                     \
                std::memcpy(k, kv.data(), kv.size() * 8);                      \
            }                                                                  \
            *nout = (int)b.size();                                             \
        });                                                                    \
    }                                                                          \
    ORC_EXPORT int orc_init_state(void *h, int lev, int kind,                  \
                                  unsigned long long seed, const double *p,    \
                                  int np) {                                    \
        return orc_guard([&] {                                                 \
            auto *s = static_cast<OrcSim *>(h);                                \
            if (kind == 1)                                                     \
                s->init_complex_random_phase(lev, seed, np > 0 ? p[0] : 0.);   \
            else if (kind >= 3 && kind <= 5)                                   \
                s->init_objects(lev, kind, seed, p, np);                       \
            else                                                               \
                throw std::runtime_error("unknown initial-state kind");        \
        });                                                                    \
    }                                                                          \
    /* Spectral building blocks, exposed for their own identity tests. */     \
    ORC_EXPORT void orc_r2c(int N, const double *in, double *out) {            \
        std::vector<double> v(in, in + (size_t)N * N * N);                     \
        std::vector<std::complex<double>> c;                                   \
        orc::r2c(v, N, c);                                                     \
        std::memcpy(out, c.data(), c.size() * sizeof(c[0]));                   \
    }                                                                          \
    ORC_EXPORT double orc_gw_lambda(int i, int j, int l, int m, int a, int b,  \
                                    int c, int N, int type) {                  \
        std::vector<double> itk(N);                                            \
        for (int q = 0; q < N; ++q) itk[q] = orc::index_to_k(q, N, type);      \
        const int abc[3] = {a, b, c};                                          \
        return orc::gw_lambda(i, j, l, m, abc, itk.data());                    \
    }                                                                          \
    ORC_EXPORT int orc_unique_k2(int N, long long *out, int maxn) {            \
        const std::vector<long long> v = orc::unique_k2(N);                    \
        if (out && maxn >= (int)v.size())                                      \
            std::memcpy(out, v.data(), v.size() * sizeof(long long));          \
        return (int)v.size();                                                  \
    }

// extern "C" entry points (include/shm.h) over shm::Runtime.
#include "../../include/shm.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "shm_hdf5.hpp"
#include "shm_runtime.hpp"

extern "C" sledgehamr::Sledgehamr *shm_project_factory();

struct shm_ctx {
    std::unique_ptr<sledgehamr::Sledgehamr> sim;
    std::unique_ptr<shm::Runtime> rt;
    std::string err;
};

namespace {

template <class F>
int guard(shm_ctx *ctx, F &&f) {
    if (!ctx) return SHM_ERR_ARG;
    try {
        f();
        ctx->err.clear();
        return SHM_OK;
    } catch (const shm::Error &e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return SHM_ERR_ARG;
    } catch (...) {
        ctx->err = "unknown error";
        return SHM_ERR_INTERNAL;
    }
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) int shm_create(const shm_config *cfg,
                                                      shm_ctx **out) {
    if (!cfg || !out || !cfg->inputs) return SHM_ERR_ARG;
    *out = nullptr;
    auto ctx = std::make_unique<shm_ctx>();
    const int rc = guard(ctx.get(), [&] {
        if (hipSetDevice(cfg->device) != hipSuccess)
            throw shm::Error(SHM_ERR_HIP, "hipSetDevice failed");
        ctx->sim.reset(shm_project_factory());
        ctx->rt = std::make_unique<shm::Runtime>(ctx->sim.get(),
                                                 ctx->sim->ShmKernels());
        ctx->rt->init_from_inputs(cfg->inputs, cfg->precision);
    });
    if (rc != SHM_OK) {
        // Keep the context so the caller can read the error, then destroy.
        ctx->rt.reset();
        *out = ctx.release();
        return rc;
    }
    *out = ctx.release();
    return SHM_OK;
}

__attribute__((visibility("default"))) int shm_destroy(shm_ctx *ctx) {
    if (!ctx) return SHM_ERR_ARG;
    try {
        ctx->rt.reset();
        ctx->sim.reset();
    } catch (...) {
    }
    delete ctx;
    return SHM_OK;
}

__attribute__((visibility("default"))) const char *shm_last_error(
    const shm_ctx *ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

__attribute__((visibility("default"))) const char *shm_project_name(
    const shm_ctx *ctx) {
    return (ctx && ctx->sim) ? ctx->sim->project_name.c_str() : "";
}

#define SHM_NEED_RT                                                           \
    if (!ctx->rt) throw shm::Error(SHM_ERR_STATE, "context not initialised")

__attribute__((visibility("default"))) int shm_get_info(shm_ctx *ctx,
                                                        shm_info_t *info) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!info) throw shm::Error(SHM_ERR_ARG, "null info");
        sledgehamr::Sledgehamr &S = *ctx->sim;
        shm::Runtime &R = *ctx->rt;
        const shm::KernelTable &kt = S.ShmKernels();
        std::memset(info, 0, sizeof(*info));
        info->ncomp = R.ncomp();
        info->nscalars = kt.nscalars;
        info->nfields = kt.nfields;
        info->nghost = S.nghost;
        info->stencil_radius = R.radius();
        info->precision = R.precision();
        info->max_level = S.GetMaxLevel();
        info->finest_level = S.GetFinestLevel();
        info->coarse_level_grid_size = S.GetDimN(0);
        info->with_gravitational_waves = S.with_gravitational_waves ? 1 : 0;
        info->L = S.GetL();
        info->cfl = S.GetDt(0) / S.GetDx(0);
        amrex::ParmParse pp("");
        int it = 0;
        pp.query("integrator.type", it);
        info->integrator_type = it;
        double ts = 0, te = 0;
        pp.query("sim.t_start", ts);
        pp.query("sim.t_end", te);
        info->t_start = ts;
        info->t_end = te;
        double crit = DBL_MAX;
        pp.query("amr.te_crit", crit);
        info->shadow_hierarchy = crit != DBL_MAX;
        info->graph_replays = ctx->rt->graph_replays;
        info->gw_groups = ctx->rt->gw_groups() ? 1 : 0;
    });
}

__attribute__((visibility("default"))) int shm_set_box_table(shm_ctx *ctx,
                                                             int lev,
                                                             const int *boxes,
                                                             int nbox) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!boxes || nbox <= 0) throw shm::Error(SHM_ERR_ARG, "empty box table");
        std::vector<shm::Box3> b(nbox);
        for (int i = 0; i < nbox; ++i)
            for (int a = 0; a < 3; ++a) {
                b[i].lo[a] = boxes[6 * i + a];
                b[i].hi[a] = boxes[6 * i + 3 + a];
            }
        ctx->rt->set_level(lev, b);
    });
}

__attribute__((visibility("default"))) int shm_get_box_table(
    shm_ctx *ctx, int lev, int *boxes, int maxbox, int *nbox) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        const shm::LevelState &ls = ctx->rt->level(lev);
        if (nbox) *nbox = (int)ls.L.boxes.size();
        if (boxes)
            for (int i = 0; i < (int)ls.L.boxes.size() && i < maxbox; ++i)
                for (int a = 0; a < 3; ++a) {
                    boxes[6 * i + a] = ls.L.boxes[i].lo[a];
                    boxes[6 * i + 3 + a] = ls.L.boxes[i].hi[a];
                }
    });
}

__attribute__((visibility("default"))) int shm_level_info(shm_ctx *ctx, int lev,
                                                          long long *ncells,
                                                          double *t,
                                                          int *istep) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        const shm::LevelState &ls = ctx->rt->level(lev);
        if (ncells) *ncells = ls.L.nvalid();
        if (t) *t = ls.t;
        if (istep) *istep = ls.istep;
    });
}

__attribute__((visibility("default"))) int shm_set_level_time(shm_ctx *ctx,
                                                              int lev, double t,
                                                              int istep) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        shm::LevelState &ls = ctx->rt->level(lev);
        ls.t = t;
        ls.istep = istep;
    });
}

__attribute__((visibility("default"))) int shm_upload(shm_ctx *ctx, int lev,
                                                      int comp,
                                                      const double *data) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!data) throw shm::Error(SHM_ERR_ARG, "null data");
        ctx->rt->upload(lev, comp, data);
    });
}

__attribute__((visibility("default"))) int shm_download(shm_ctx *ctx, int lev,
                                                        int comp, double *data,
                                                        int old_state) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!data) throw shm::Error(SHM_ERR_ARG, "null data");
        ctx->rt->download(lev, comp, data, old_state != 0);
    });
}

__attribute__((visibility("default"))) int shm_init_state(
    shm_ctx *ctx, int lev, int kind, unsigned long long seed, const double *p,
    int np) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->init_state(lev, kind, seed, p, p ? np : 0);
    });
}

__attribute__((visibility("default"))) int shm_fill_rhs(shm_ctx *ctx, int lev,
                                                        double time,
                                                        double weight, int add) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->fill_rhs(lev, time, weight, add != 0);
    });
}

__attribute__((visibility("default"))) int shm_download_rhs(shm_ctx *ctx,
                                                            int lev, int comp,
                                                            double *data) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!data) throw shm::Error(SHM_ERR_ARG, "null data");
        ctx->rt->download_rhs(lev, comp, data);
    });
}

__attribute__((visibility("default"))) int shm_advance_level(shm_ctx *ctx,
                                                             int lev) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->advance(lev);
    });
}

__attribute__((visibility("default"))) int shm_evolve(shm_ctx *ctx,
                                                      int ncoarse) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->step(ncoarse);
    });
}

__attribute__((visibility("default"))) int shm_create_shadow(shm_ctx *ctx) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->create_shadow_level();
    });
}

__attribute__((visibility("default"))) int shm_te_tag(shm_ctx *ctx, int lev,
                                                      int flags,
                                                      unsigned char *mask,
                                                      long long *counts,
                                                      double *te) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->te_tag(lev, flags, mask, counts, te);
    });
}

__attribute__((visibility("default"))) int shm_event_log(shm_ctx *ctx, char *kinds,
                                                         int *levs, double *times,
                                                         int maxn, int *n) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!n) throw shm::Error(SHM_ERR_ARG, "null n");
        const auto &ev = ctx->rt->events_;
        *n = (int)ev.size();
        if (kinds && levs && times)
            for (int i = 0; i < (int)ev.size() && i < maxn; ++i) {
                kinds[i] = ev[i].kind;
                levs[i] = ev[i].lev;
                times[i] = ev[i].t;
            }
    });
}

__attribute__((visibility("default"))) int shm_hdf5_version(int *version3) {
    if (!version3) return SHM_ERR_ARG;
    try {
        const shm::h5::Api &a = shm::h5::api();
        for (int i = 0; i < 3; ++i) version3[i] = (int)a.version[i];
        return SHM_OK;
    } catch (const shm::Error &e) {
        return e.code;
    } catch (...) {
        return SHM_ERR_INTERNAL;
    }
}

__attribute__((visibility("default"))) int shm_read_initial_state(shm_ctx *ctx, int lev,
                                                                  const char *path) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->read_initial_state(lev, path ? path : "");
    });
}

__attribute__((visibility("default"))) int shm_write_checkpoint(
    shm_ctx *ctx, const char *folder, const int *next_id, const double *last_time_written,
    int noutput, int idx_checkpoints) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!folder) throw shm::Error(SHM_ERR_ARG, "null folder");
        if (noutput < 0 || (noutput > 0 && (!next_id || !last_time_written)))
            throw shm::Error(SHM_ERR_ARG, "output counters missing");
        std::vector<int> nid(next_id, next_id + noutput);
        std::vector<double> lt(last_time_written, last_time_written + noutput);
        ctx->rt->write_checkpoint(folder, nid, lt, idx_checkpoints);
    });
}

__attribute__((visibility("default"))) int shm_read_checkpoint(
    shm_ctx *ctx, const char *folder, int *next_id, double *last_time_written, int noutput,
    int idx_checkpoints) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!folder) throw shm::Error(SHM_ERR_ARG, "null folder");
        const bool want = next_id != nullptr;
        if (want && noutput > 0 && !last_time_written)
            throw shm::Error(SHM_ERR_ARG, "null last_time_written");
        std::vector<int> nid;
        std::vector<double> lt;
        ctx->rt->read_checkpoint(folder, want ? noutput : -1, idx_checkpoints, nid, lt);
        if (want)
            for (int i = 0; i < noutput; ++i) {
                next_id[i] = nid[i];
                last_time_written[i] = lt[i];
            }
    });
}

__attribute__((visibility("default"))) int shm_delete_checkpoint(shm_ctx *ctx,
                                                                 const char *folder) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!folder) throw shm::Error(SHM_ERR_ARG, "null folder");
        ctx->rt->delete_checkpoint(folder);
    });
}

__attribute__((visibility("default"))) int shm_write_gw_spectrum(shm_ctx *ctx,
                                                                 const char *prefix,
                                                                 int modifier,
                                                                 const double *params,
                                                                 int nparams) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!prefix) throw shm::Error(SHM_ERR_ARG, "null prefix");
        if (nparams < 0 || (nparams > 0 && !params))
            throw shm::Error(SHM_ERR_ARG, "modifier parameters missing");
        std::vector<double> mp(params, params + nparams);
        ctx->rt->write_gw_spectrum_file(prefix, modifier, mp);
    });
}

__attribute__((visibility("default"))) int shm_write_spectra(shm_ctx *ctx, const char *prefix,
                                                             const int *fns, int nfn) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!prefix) throw shm::Error(SHM_ERR_ARG, "null prefix");
        if (nfn < 0 || (nfn > 0 && !fns)) throw shm::Error(SHM_ERR_ARG, "null fns");
        ctx->rt->write_spectra_file(prefix, std::vector<int>(fns, fns + nfn));
    });
}

__attribute__((visibility("default"))) int shm_write_projections(shm_ctx *ctx,
                                                                 const char *prefix,
                                                                 const int *fns,
                                                                 const int *modes, int nfn) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!prefix) throw shm::Error(SHM_ERR_ARG, "null prefix");
        if (nfn < 0 || (nfn > 0 && !fns)) throw shm::Error(SHM_ERR_ARG, "null fns");
        std::vector<int> m;
        if (modes) m.assign(modes, modes + nfn);
        ctx->rt->write_projections_file(prefix, std::vector<int>(fns, fns + nfn), m);
    });
}

__attribute__((visibility("default"))) int shm_te_tag_noalloc(shm_ctx *ctx, int lev,
                                                              int flags,
                                                              unsigned char *mask,
                                                              long long *counts) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->te_tag_noalloc(lev, flags, mask, counts);
    });
}

__attribute__((visibility("default"))) int shm_gw_spectrum(
    shm_ctx *ctx, int mode, int nbins, int zero_padding, int projection_type,
    double *bins, double *kvals, int *nout) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!nout) throw shm::Error(SHM_ERR_ARG, "null nout");
        std::vector<double> b, k;
        ctx->rt->gw_spectrum(mode, nbins, zero_padding, projection_type, b, k, -1, {},
                             bins == nullptr);
        if (bins) {
            if (*nout < (int)b.size())
                throw shm::Error(SHM_ERR_ARG, "output arrays too small");
            std::memcpy(bins, b.data(), b.size() * sizeof(double));
            if (kvals) std::memcpy(kvals, k.data(), k.size() * sizeof(double));
        }
        *nout = (int)b.size();
    });
}

__attribute__((visibility("default"))) int shm_gw_spectrum_mod(
    shm_ctx *ctx, int mode, int nbins, int zero_padding, int projection_type,
    int modifier, const double *params, int nparams, double *bins, double *kvals,
    int *nout) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!nout) throw shm::Error(SHM_ERR_ARG, "null nout");
        if (nparams < 0 || (nparams > 0 && !params))
            throw shm::Error(SHM_ERR_ARG, "bad modifier parameters");
        std::vector<double> p(params, params + nparams), b, k;
        ctx->rt->gw_spectrum(mode, nbins, zero_padding, projection_type, b, k, modifier, p,
                             bins == nullptr);
        if (bins) {
            if (*nout < (int)b.size())
                throw shm::Error(SHM_ERR_ARG, "output arrays too small");
            std::memcpy(bins, b.data(), b.size() * sizeof(double));
            if (kvals) std::memcpy(kvals, k.data(), k.size() * sizeof(double));
        }
        *nout = (int)b.size();
    });
}

__attribute__((visibility("default"))) int shm_gw_modifiers(shm_ctx *ctx, int *count) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!count) throw shm::Error(SHM_ERR_ARG, "null count");
        *count = ctx->rt->kernels().ngwmod;
    });
}

__attribute__((visibility("default"))) int shm_gw_modifier_name(shm_ctx *ctx, int id,
                                                                char *buf, int len) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        const shm::KernelTable &kt = ctx->rt->kernels();
        if (id < 0 || id >= kt.ngwmod || !kt.gwmod_name)
            throw shm::Error(SHM_ERR_ARG, "no such GW spectrum modifier");
        if (!buf || len <= 0) throw shm::Error(SHM_ERR_ARG, "bad buffer");
        std::snprintf(buf, (size_t)len, "%s", kt.gwmod_name(id));
    });
}

__attribute__((visibility("default"))) int shm_fill_coarse_patch(shm_ctx *ctx,
                                                                int lev) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->fill_coarse_patch(lev);
    });
}

static std::vector<shm::Box3> box_list(const int *boxes, int nbox) {
    if (nbox < 0 || (nbox > 0 && !boxes)) throw shm::Error(SHM_ERR_ARG, "bad box list");
    std::vector<shm::Box3> b(nbox);
    for (int i = 0; i < nbox; ++i)
        for (int a = 0; a < 3; ++a) {
            b[i].lo[a] = boxes[6 * i + a];
            b[i].hi[a] = boxes[6 * i + 3 + a];
        }
    return b;
}

__attribute__((visibility("default"))) int shm_regrid_select(
    shm_ctx *ctx, int lev, const unsigned char *mask, int *blocks, int maxblock,
    int *nblock, double *min_distance) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        shm::Runtime::BlockSet s;
        int d = INT_MAX;
        ctx->rt->regrid_select(lev, mask, s, d);
        if (nblock) *nblock = (int)s.size();
        if (min_distance) *min_distance = d == INT_MAX ? -1. : std::sqrt((double)d);
        if (blocks) {
            if ((int)s.size() > maxblock)
                throw shm::Error(SHM_ERR_ARG, "block buffer too small");
            int i = 0;
            for (const auto &b : s) {
                for (int a = 0; a < 3; ++a) blocks[3 * i + a] = b[a];
                ++i;
            }
        }
    });
}

__attribute__((visibility("default"))) int shm_add_boxes(shm_ctx *ctx, int lev,
                                                         const int *boxes, int nbox) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->add_boxes(lev, box_list(boxes, nbox));
    });
}

__attribute__((visibility("default"))) int shm_remake_level(shm_ctx *ctx, int lev,
                                                            const int *boxes, int nbox) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->remake_level(lev, box_list(boxes, nbox));
    });
}

__attribute__((visibility("default"))) int shm_local_regrid(shm_ctx *ctx, int lev,
                                                            int *outcome) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        const int r = ctx->rt->local_regrid(lev);
        if (outcome) *outcome = r;
    });
}

__attribute__((visibility("default"))) int shm_regrid_stats(shm_ctx *ctx,
                                                            long long *stats,
                                                            double *min_distance,
                                                            int nlev) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        const shm::RegridState &r = ctx->rt->regrid_state();
        if (stats) {
            stats[0] = r.attempts;
            stats[1] = r.global_wanted;
            stats[2] = r.nregrids;
            stats[3] = r.veto_level;
        }
        if (min_distance)
            for (int l = 0; l < nlev; ++l)
                min_distance[l] = l < (int)r.min_distance.size() ? r.min_distance[l] : -1.;
    });
}

__attribute__((visibility("default"))) int shm_upload_array(shm_ctx *ctx, int lev,
                                                           int comp,
                                                           const double *data,
                                                           int layout) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!data) throw shm::Error(SHM_ERR_ARG, "null data");
        ctx->rt->upload_array(lev, comp, data, layout);
    });
}

__attribute__((visibility("default"))) int shm_output_functions(shm_ctx *ctx,
                                                               int *count) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!count) throw shm::Error(SHM_ERR_ARG, "null count");
        *count = ctx->rt->output_count();
    });
}

__attribute__((visibility("default"))) int shm_output_function_name(
    shm_ctx *ctx, int id, char *buf, int len) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        const char *n = ctx->rt->output_name(id);
        if (!n) throw shm::Error(SHM_ERR_ARG, "output function id out of range");
        if (!buf || len <= (int)std::strlen(n))
            throw shm::Error(SHM_ERR_ARG, "name buffer too small");
        std::strcpy(buf, n);
    });
}

__attribute__((visibility("default"))) int shm_spectrum(shm_ctx *ctx, int fn,
                                                       double *bins,
                                                       long long *k_sq, int *n) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!n) throw shm::Error(SHM_ERR_ARG, "null n");
        std::vector<double> b;
        std::vector<long long> k;
        ctx->rt->spectrum(fn, b, k, bins == nullptr);
        if (bins) {
            if (*n < (int)b.size()) throw shm::Error(SHM_ERR_ARG, "output arrays too small");
            std::memcpy(bins, b.data(), b.size() * sizeof(double));
            if (k_sq) std::memcpy(k_sq, k.data(), k.size() * sizeof(long long));
        }
        *n = (int)b.size();
    });
}

__attribute__((visibility("default"))) int shm_projection(
    shm_ctx *ctx, int fn, int mode, int max_level, double *data, int *counts,
    int *dimN) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!dimN) throw shm::Error(SHM_ERR_ARG, "null dimN");
        std::vector<double> d;
        std::vector<int> c;
        int D = 0;
        ctx->rt->projection(fn, mode, max_level, d, c, D, data == nullptr);
        if (data) {
            if (*dimN < D) throw shm::Error(SHM_ERR_ARG, "output arrays too small");
            std::memcpy(data, d.data(), d.size() * sizeof(double));
            if (counts) std::memcpy(counts, c.data(), c.size() * sizeof(int));
        }
        *dimN = D;
    });
}

__attribute__((visibility("default"))) int shm_sync(shm_ctx *ctx) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->sync();
    });
}

__attribute__((visibility("default"))) int shm_stream_probe(
    shm_ctx *ctx, unsigned long long bytes, int reps, double *read_gbs,
    double *write_gbs, double *copy_gbs) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (bytes < (1u << 20) || reps < 1)
            throw shm::Error(SHM_ERR_ARG, "stream probe needs >= 1 MiB and reps >= 1");
        ctx->rt->stream_probe(bytes, reps, read_gbs, write_gbs, copy_gbs);
    });
}

__attribute__((visibility("default"))) int shm_level_tag_counts(shm_ctx *ctx, int lev,
                                                                long long *counts) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!counts) throw shm::Error(SHM_ERR_ARG, "null counts");
        const std::vector<long long> &c = ctx->rt->last_tags(lev);
        std::copy(c.begin(), c.end(), counts);
    });
}

__attribute__((visibility("default"))) int shm_autotune(shm_ctx *ctx, int lev) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->autotune(lev);
    });
}

__attribute__((visibility("default"))) int shm_set_launch(shm_ctx *ctx,
                                                          const shm_launch_t *cfg) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!cfg) throw shm::Error(SHM_ERR_ARG, "null launch shape");
        if (cfg->variant >= shm::kNVar || cfg->variant < -1)
            throw shm::Error(SHM_ERR_ARG, "bad tile variant");
        if (cfg->tile_z < 0) throw shm::Error(SHM_ERR_ARG, "bad tile_z");
        auto flag = [](int v, const char *what) {
            if (v < -1 || v > 1)
                throw shm::Error(SHM_ERR_ARG, std::string("bad ") + what);
            return v;
        };
        shm::LaunchCfg c;
        c.variant = cfg->variant;
        c.tz = cfg->tile_z;
        c.xcd_split = flag(cfg->xcd_split, "xcd_split");
        c.persistent = flag(cfg->persistent, "persistent");
        c.zc_minor = flag(cfg->zc_minor, "zc_minor");
        c.fused = flag(cfg->fused, "fused");
        if (cfg->fused_variant >= shm::kNVar || cfg->fused_variant < -1)
            throw shm::Error(SHM_ERR_ARG, "bad fused tile variant");
        c.fused_variant = cfg->fused_variant;
        ctx->rt->launch = c;
        ctx->rt->bump_epoch();
    });
}

__attribute__((visibility("default"))) int shm_get_launch(shm_ctx *ctx, int lev,
                                                          shm_launch_t *cfg) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        if (!cfg) throw shm::Error(SHM_ERR_ARG, "null launch shape");
        if (lev < -1 || lev >= ctx->rt->nlevels())
            throw shm::Error(SHM_ERR_ARG, "bad level");
        const shm::LaunchCfg c = ctx->rt->resolved_launch(lev);
        cfg->variant = c.variant;
        cfg->tile_z = c.tz;
        cfg->xcd_split = c.xcd_split;
        cfg->persistent = c.persistent;
        cfg->zc_minor = c.zc_minor;
        cfg->fused = c.fused && ctx->rt->fused_possible(ctx->rt->level(lev)) ? 1 : 0;
        cfg->fused_variant = c.fused_variant;
    });
}

__attribute__((visibility("default"))) int shm_set_graphs(shm_ctx *ctx, int on) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->sync();
        ctx->rt->graphs = on != 0;
    });
}

__attribute__((visibility("default"))) int shm_set_timing(shm_ctx *ctx, int on) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->collect_timing();
        ctx->rt->timing_on = on != 0;
        ctx->rt->timing = shm::Timing();
    });
}

__attribute__((visibility("default"))) int shm_get_timing(
    shm_ctx *ctx, double *stage_ms, long long *stage_launches,
    long long *stage_cells, double *halo_ms, long long *halo_launches,
    double *fused_ms, long long *fused_launches) {
    return guard(ctx, [&] {
        SHM_NEED_RT;
        ctx->rt->collect_timing();
        const shm::Timing &t = ctx->rt->timing;
        if (stage_ms) *stage_ms = t.stage_ms;
        if (stage_launches) *stage_launches = t.stage_launches;
        if (stage_cells) *stage_cells = t.stage_cells;
        if (halo_ms) *halo_ms = t.halo_ms;
        if (halo_launches) *halo_launches = t.halo_launches;
        if (fused_ms) *fused_ms = t.fused_ms;
        if (fused_launches) *fused_launches = t.fused_launches;
    });
}

}  // extern "C"

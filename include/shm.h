/* C ABI of the MI355X sledgehamr time-step library (libshm_<Project>.so).
 *
 * One shared library is built per project: it holds that project's cell
 * functions (Rhs, GravitationalWavesRhs, TagCellForRefinement,
 * TruncationModifier) compiled into the gfx950 kernels.  Host code talks to
 * it only through the functions below: plain pointers, sizes and status
 * codes; no C++ or torch types.
 *
 * Conventions
 *   - Every function returns SHM_OK (0) or a positive status code; the text
 *     of the last error of a context is available from shm_last_error().
 *   - Field data are exchanged as FP64 arrays of the valid cells of a level,
 *     box after box in box-table order, x fastest within a box (the Array4
 *     order of the reference).  Boxes are inclusive index boxes
 *     {lo_x, lo_y, lo_z, hi_x, hi_y, hi_z}.
 *   - Calls on one context are stream-ordered on the device; functions that
 *     return data synchronise.  One context per host thread.
 */
#ifndef SHM_H_
#define SHM_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shm_ctx shm_ctx;

enum shm_status {
    SHM_OK = 0,
    SHM_ERR_ARG = 1,          /* invalid argument or inputs-file value */
    SHM_ERR_HIP = 2,          /* HIP / hipFFT runtime failure */
    SHM_ERR_STATE = 3,        /* call not valid in the current state */
    SHM_ERR_UNSUPPORTED = 4,  /* configuration not built / not supported */
    SHM_ERR_INTERNAL = 5,
    SHM_ERR_IO = 6            /* a file could not be read or written */
};

enum shm_precision {
    SHM_PREC_INPUTS = -1,     /* take shm.precision from the inputs file */
    SHM_PREC_FP64 = 0,        /* FP64 storage and arithmetic */
    SHM_PREC_FP32C = 1,       /* FP32 storage + FP32 remainder (compensated) */
    SHM_PREC_FP32C16 = 2,     /* FP32 storage + bf16 remainder (compensated) */
    SHM_PREC_FP32CS = 3       /* stored as FP32C16; stage states and stencils
                                 formed in FP64 from value + remainder
                                 (compensated stencils) */
};

enum shm_te_flags {
    SHM_TE_COMPUTE = 1,       /* compute truncation errors */
    SHM_TE_AVGDOWN = 2,       /* replace coarse data by the child mean */
    SHM_TE_TAGS = 4           /* write the tag mask and counters */
};

typedef struct shm_config {
    const char *inputs;       /* inputs-file text (ParmParse syntax) */
    int precision;            /* enum shm_precision */
    int device;               /* HIP device ordinal */
} shm_config;

typedef struct shm_info_t {
    int ncomp;                /* components incl. GW tensor when enabled */
    int nscalars;             /* project fields + momenta */
    int nfields;              /* project fields */
    int nghost;
    int stencil_radius;       /* reach of the cell functions (probed) */
    int precision;
    int max_level;
    int finest_level;
    int coarse_level_grid_size;
    int integrator_type;
    int with_gravitational_waves;
    int shadow_hierarchy;
    double L, cfl, t_start, t_end;
    long long graph_replays;  /* level steps run from captured graphs */
    int gw_groups;            /* GW stages run as three component groups
                                 (shm.gw_split and the dependency probe) */
} shm_info_t;

/* Sledgehamr::Sledgehamr + InitSledgehamr (sledgehamr.cpp:15-77): build the
 * project, parse the inputs file, create level 0 (one periodic box, or
 * chopped by amr.max_grid_size), probe the cell functions. */
int shm_create(const shm_config *cfg, shm_ctx **out);
int shm_destroy(shm_ctx *ctx);
const char *shm_last_error(const shm_ctx *ctx);
const char *shm_project_name(const shm_ctx *ctx);
int shm_get_info(shm_ctx *ctx, shm_info_t *info);

/* Define / replace the boxes of a level (AmrCore::MakeNewLevel* for a
 * static layout; boxes must be properly nested and even-aligned above 0). */
int shm_set_box_table(shm_ctx *ctx, int lev, const int *boxes, int nbox);
int shm_get_box_table(shm_ctx *ctx, int lev, int *boxes, int maxbox, int *nbox);
int shm_level_info(shm_ctx *ctx, int lev, long long *ncells, double *t,
                   int *istep);
int shm_set_level_time(shm_ctx *ctx, int lev, double t, int istep);

/* State transfer for grid_new of a level (or grid_old with old_state=1). */
int shm_upload(shm_ctx *ctx, int lev, int comp, const double *data);
int shm_download(shm_ctx *ctx, int lev, int comp, double *data, int old_state);

/* Seeded device-side initial states (kind 1: complex random phase with
 * modulus 1 + p[0] N(0,1); kind 2: zero). */
int shm_init_state(shm_ctx *ctx, int lev, int kind, unsigned long long seed,
                   const double *p, int np);

/* Sledgehamr::FillRhs / FillAddRhs (sledgehamr.h:71-88, macros.h:301-485):
 * rhs = f(grid_new) or rhs = f(grid_new) + weight * rhs. */
int shm_fill_rhs(shm_ctx *ctx, int lev, double time, double weight, int add);
int shm_download_rhs(shm_ctx *ctx, int lev, int comp, double *data);

/* Integrator::Advance(lev) (integrator.cpp:11-31) with the configured
 * integrator (integrator.type 10, 11, 20, 21, 22). */
int shm_advance_level(shm_ctx *ctx, int lev);

/* TimeStepper::Advance(0) repeated ncoarse times (time_stepper.cpp:90-133).
 * With shm.local_regrid = 1, a scheduled regrid whose local regrid asks for
 * a global one (LocalRegrid::DoAttemptRegrid returning false) ends the call
 * with SHM_ERR_UNSUPPORTED right after that regrid: the caller runs its
 * global regrid (AmrCore::regrid) and hands the new box lists to
 * shm_remake_level, then continues.  With shm.keep_layout_on_global_regrid
 * = 1 the layout stays instead and the evolve goes on. */
int shm_evolve(shm_ctx *ctx, int ncoarse);

/* Counters of the most recent truncation-error tagging that shm_evolve ran
 * on lev at a scheduled regrid (Sledgehamr::ErrorEst's ntags_total,
 * ntags_user, ntags_trunc[ncomp], sledgehamr.cpp:280-348); the layout
 * itself stays static.  counts holds 2 + ncomp values. */
int shm_level_tag_counts(shm_ctx *ctx, int lev, long long *counts);

/* The order of the level operations shm_evolve ran so far (the sub-cycled
 * schedule of TimeStepper::Advance, time_stepper.cpp:90-171, 221-311):
 * kinds[i] is 'S' (a regrid scheduled by levs[i] for times[i]), 'H' (shadow
 * level created), 'A' (levs[i] advanced to times[i]), 'D' (levs[i]+1
 * averaged onto levs[i]), 'T' (truncation errors and tags of levs[i]) or
 * 'R' (regrid from levs[i]).  No counterpart in the reference (its log
 * messages); for tests and tracing.  Arrays hold maxn entries; *n receives
 * the number of events (call with NULL arrays to query it). */
int shm_event_log(shm_ctx *ctx, char *kinds, int *levs, double *times, int maxn,
                  int *n);

/* Sledgehamr::CreateShadowLevel (sledgehamr.cpp:234-251). */
int shm_create_shadow(shm_ctx *ctx);

/* LevelSynchronizer::ComputeTruncationErrors (level_synchronizer.cpp:240-357)
 * fused with Sledgehamr::ErrorEst/TagWithTruncationCpu (sledgehamr.cpp:280-348,
 * macros.h:489-566).  mask: one byte per valid cell of lev (may be NULL);
 * counts: [ntags_total, ntags_user, ntags_trunc[ncomp]] (may be NULL);
 * te: truncation errors per coarse cell, ncomp planes (may be NULL). */
int shm_te_tag(shm_ctx *ctx, int lev, int flags, unsigned char *mask,
               long long *counts, double *te);

/* shm_te_tag without device memory allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc/hipFree: counters, parameter records, the
 * coarse/fine fill record and the tag mask live in a fixed device scratch
 * array of 32 MiB per GPU, held until the process ends, so the level may
 * hold at most 33 550 336 valid cells (larger levels: SHM_ERR_ARG), and at
 * most 62 components and 64 tagging parameters each.  The level must be
 * set up by an earlier allocating call (its shadow level or coarsened copy
 * and coarse/fine fill plans; otherwise SHM_ERR_STATE).  Outputs as
 * shm_te_tag without te; the level's kept mask used by shm_regrid_select
 * is not updated.  Every call on a GPU shares that GPU's scratch array and
 * leaves its contents behind: calls that overlap on one GPU (two contexts
 * used from two host threads) race. */
int shm_te_tag_noalloc(shm_ctx *ctx, int lev, int flags, unsigned char *mask,
                       long long *counts);

/* GravitationalWaves::ComputeSpectrum (gravitational_waves.cpp:52-322).
 * mode 0: unbinned (unique |k|^2), 1: rounded |k|, 2: nbins log bins.
 * bins/kvals must hold *nout values (call with bins=NULL to query *nout;
 * that call computes nothing). */
int shm_gw_spectrum(shm_ctx *ctx, int mode, int nbins, int zero_padding,
                    int projection_type, double *bins, double *kvals,
                    int *nout);

/* GravitationalWaves::ComputeSpectrum(file, modifier) with a project
 * spectrum modifier (GravitationalWavesSpectrumModifier, gravitational_waves.h:
 * 79-92, called at gravitational_waves.cpp:69 and :96): modifier is the
 * position of the struct in the project's SLEDGEHAMR_GW_SPECTRUM_MODIFIERS
 * list (-1: the default du_ij, unmodified, as shm_gw_spectrum); params /
 * nparams are its per-call parameters (the reference modifier's constructor
 * arguments).  Other arguments as shm_gw_spectrum. */
int shm_gw_spectrum_mod(shm_ctx *ctx, int mode, int nbins, int zero_padding,
                        int projection_type, int modifier, const double *params,
                        int nparams, double *bins, double *kvals, int *nout);
/* The project's GW spectrum modifiers: their number, and the name of id. */
int shm_gw_modifiers(shm_ctx *ctx, int *count);
int shm_gw_modifier_name(shm_ctx *ctx, int id, char *buf, int len);

/* LevelSynchronizer::FillCoarsePatch (level_synchronizer.cpp:56-88): every
 * cell of level lev (lev >= 1), ghosts included, interpolated from level
 * lev-1 at the level's time (amr.interpolation_type); between the coarse
 * level's two times the old state is used, as the reference does. */
int shm_fill_coarse_patch(shm_ctx *ctx, int lev);

/* LocalRegrid::DetermineNewBoxArray with TagAndMeasure / CheckBorders
 * (local_regrid/local_regrid.cpp:511-680) for tagging level lev, on the
 * device: the blocks (amr.blocking_factor cubes) of level lev+1 that tagged
 * cells within amr.n_error_buf of a coarse/fine boundary ask for, and the
 * shortest distance of a tagged cell to such a boundary (-1: none measured).
 * mask: one byte per valid cell of lev (box-major, x fastest), or NULL for
 * the mask of the level's last tagging at its time.  blocks receives 3 ints
 * (i, j, k) per block in (i, j, k) order; *nblock their number. */
int shm_regrid_select(shm_ctx *ctx, int lev, const unsigned char *mask,
                      int *blocks, int maxblock, int *nblock,
                      double *min_distance);

/* LocalRegrid::AddBoxes (local_regrid.cpp:763-828): level lev keeps its
 * boxes and data and gains nbox boxes after them (lo[3], hi[3] each, no
 * overlap), filled as FillPatch does at the level's new and old times. */
int shm_add_boxes(shm_ctx *ctx, int lev, const int *boxes, int nbox);

/* Sledgehamr::RemakeLevel (sledgehamr.cpp:173-190) on a new box list, state
 * FillPatch-ed at the level's time from its previous data and level lev-1;
 * the local regrid's counters reset as after a global regrid
 * (LocalRegrid::DidGlobalRegrid, local_regrid.cpp:34-46). */
int shm_remake_level(shm_ctx *ctx, int lev, const int *boxes, int nbox);

/* LocalRegrid::DoAttemptRegrid (local_regrid.cpp:407-440) from tagging level
 * lev, using the masks of each level's last tagging: new blocks, nesting
 * fixes, volume thresholds and vetoes, then AddBoxes.  *outcome: 0 a global
 * regrid is wanted (not done here), 1 boxes added (possibly none), 2 the
 * regrid is delayed.  shm_evolve calls it at scheduled regrids when the
 * input shm.local_regrid = 1. */
int shm_local_regrid(shm_ctx *ctx, int lev, int *outcome);

/* Counters of the local regrid: stats[4] = attempts, global regrids wanted,
 * consecutive local regrids (nregrids), veto level; min_distance[nlev]: the
 * shortest tagged-cell distance per level of the last attempt (-1: none). */
int shm_regrid_stats(shm_ctx *ctx, long long *stats, double *min_distance,
                     int nlev);

/* Initial-state arrays in the reference's orders (the HDF5 reader's data):
 * layout 1: the whole level, ind = i N^2 + j N + k (FillLevel::FromArray,
 * fill_level.cpp:205-223); layout 2: one chunk per box in box order,
 * box-relative ind = i ly lz + j lz + k (FillLevel::FromArrayChunks,
 * fill_level.cpp:225-251); layout 0: as shm_upload. */
int shm_upload_array(shm_ctx *ctx, int lev, int comp, const double *data,
                     int layout);

/* Output cell functions the project declares (SLEDGEHAMR_OUTPUT_FUNCTIONS;
 * the device form of io_module->spectra / io_module->projections).  *count
 * receives their number; shm_output_function_name copies the name of id
 * into buf (len bytes, NUL-terminated). */
int shm_output_functions(shm_ctx *ctx, int *count);
int shm_output_function_name(shm_ctx *ctx, int id, char *buf, int len);

/* Spectrum::Compute (output_types/spectrum.cpp:12-113) of output function fn
 * on level 0: bins[b] = sum over modes with |k|^2 = k_sq[b] of
 * mult t |F(k)|^2 / (dk N^6), dk = 2 pi / L.  bins / k_sq must hold *n
 * values (call with bins = NULL to query *n; that call computes nothing). */
int shm_spectrum(shm_ctx *ctx, int fn, double *bins, long long *k_sq, int *n);

/* Projection::Compute (output_types/projection.cpp:11-118) of output function
 * fn along z onto the finest level up to max_level (< 0: all levels).
 * mode 0: sum of ratio * value, 1: maximum (from 0).  data / counts hold
 * (*dimN)^2 values, x index slow (call with data = NULL to query *dimN; that
 * call computes nothing). */
int shm_projection(shm_ctx *ctx, int fn, int mode, int max_level, double *data,
                   int *counts, int *dimN);

int shm_sync(shm_ctx *ctx);

/* Achievable HBM bandwidth: plain float4 read, write and copy kernels over
 * buffers of `bytes` bytes, mean of `reps` launches (GB/s; copy counts read
 * plus write).  No counterpart in the reference: the measured-copy reference
 * printed beside the HBM peak. */
int shm_stream_probe(shm_ctx *ctx, unsigned long long bytes, int reps,
                     double *read_gbs, double *write_gbs, double *copy_gbs);

/* Launch shape of the fused stage kernels.  No counterpart in the reference
 * (AMReX picks its own launch bounds): the tile shape and work order of the
 * persistent tile-table launch.  A field of -1 (tile_z: 0) is not set: the
 * tuned value of the level, else the built-in default, applies. */
typedef struct shm_launch_t {
    int variant;      /* tile shape: 0..2 = 64 x {4, 8, 16} cells per z plane,
                         one row per thread; 3 = 32 x 8 (two rows per
                         wavefront); 4 = 64 x 8, two rows per thread */
    int tile_z;       /* z-chunk length of a tile (0: automatic) */
    int xcd_split;    /* 1: each XCD walks one contiguous range of tiles */
    int persistent;   /* 1: grid = resident capacity; 0: one tile per workgroup */
    int zc_minor;     /* tile order: 0 = z chunk outermost, 1 = between y and x */
    int fused;        /* SSPRK3, or RKN with a first node 0, on levels of
                         periodic boxes: stages one and two in one launch */
    int fused_variant;/* tile shape of the fused launch (0..2 as variant) */
} shm_launch_t;

/* Times the stage kernels of a level over every launch shape not fixed by
 * shm_set_launch and keeps the fastest for that level. */
int shm_autotune(shm_ctx *ctx, int lev);
/* Fixes the launch shape of every level: each call replaces the whole
 * shape, and fields at -1 (tile_z 0) are left to the defaults and the
 * autotuner. */
int shm_set_launch(shm_ctx *ctx, const shm_launch_t *cfg);
/* The launch shape a level's stage kernels use now. */
int shm_get_launch(shm_ctx *ctx, int lev, shm_launch_t *cfg);

/* Level steps as captured HIP graphs (also shm.graphs = 1 in the inputs
 * file).  No counterpart in the reference: the stage launches and ghost
 * fills of one Integrator::Integrate are captured per (level, state slot)
 * on the first step after an eager one and replayed with fresh stage
 * records afterwards; layout or launch-shape changes recapture.  Steps run
 * eagerly while kernel timing is on. */
int shm_set_graphs(shm_ctx *ctx, int on);

/* Kernel timing with HIP events on the library's stream.  stage_* count every
 * stage launch (a fused stage-one-and-two launch updates its cells twice);
 * fused_* count the fused launches alone.  Any pointer may be NULL. */
int shm_set_timing(shm_ctx *ctx, int on);
int shm_get_timing(shm_ctx *ctx, double *stage_ms, long long *stage_launches,
                   long long *stage_cells, double *halo_ms,
                   long long *halo_launches, double *fused_ms,
                   long long *fused_launches);

/* ----------------------------------------------------------------------
 * HDF5 files.  libhdf5 (1.10 or newer) is loaded at run time: the file named
 * by $SHM_HDF5_LIB, else libhdf5.so on the loader's path, else common install
 * paths; when none loads these return SHM_ERR_UNSUPPORTED.  Missing files,
 * datasets of the wrong size and failed writes return SHM_ERR_IO.
 * ---------------------------------------------------------------------- */

/* Version of the loaded libhdf5 (major, minor, release).  Needs no context;
 * returns SHM_ERR_UNSUPPORTED when the library cannot be loaded. */
int shm_hdf5_version(int *version3);

/* FillLevel::FromInitialStateFile / FromHdf5File (fill_level.cpp:17-190) for
 * level lev, which must be defined.  path NULL or "": input.initial_state.
 * A folder holding Meta.hdf5 restarts from that checkpoint (as
 * shm_read_checkpoint without output counters).  Otherwise, per component
 * <name> (the project's scalars, then u_xx .. du_yz): the file
 * input.initial_state_<name>, else path; no file: zeros.  In that file,
 * datasets <name>_<b> / data_<b> are per-box chunks (b = position in the box
 * table; box-relative ind = i ly lz + j lz + k), else <name> / data is the
 * whole level (ind = i N^2 + j N + k), else the component is set to zero.
 * A folder path without Meta.hdf5 holds chunk files <name>_<b>.hdf5 with
 * dataset <name> or data (the reference reads these and then leaves the
 * level unchanged, fill_level.cpp:93-99; here they are applied).
 * input.upsample > 1 returns SHM_ERR_UNSUPPORTED. */
int shm_read_initial_state(shm_ctx *ctx, int lev, const char *path);

/* Checkpoint::Write (output_types/checkpoint.cpp:18-98) into folder:
 * Meta.hdf5 with the reference's datasets (Header = t, 1, finest level, N,
 * nghost, ncomp, noutput, idx_checkpoints; blocking_factors; isteps; next_id
 * and last_time_written of the noutput output types, entry idx_checkpoints
 * counted as written), and per level Level_<lev>/Cell.hdf5 (in place of
 * AMReX's MultiFab files): boxes (lo, hi), level (t, istep, N) and one FP64
 * dataset per component, the valid cells box after box, x fastest.  next_id
 * and last_time_written may be NULL when noutput is 0. */
int shm_write_checkpoint(shm_ctx *ctx, const char *folder, const int *next_id,
                         const double *last_time_written, int noutput,
                         int idx_checkpoints);

/* Checkpoint::Read + UpdateLevels + UpdateOutputModules (checkpoint.cpp:
 * 122-187, 195-275): defines levels 0..min(finest, max level) with the
 * checkpoint's boxes and data, every level at the checkpoint's time with its
 * own step count; a blocking factor smaller than the inputs' marks a global
 * regrid as wanted.  The component count must match (else SHM_ERR_ARG), as
 * must noutput / idx_checkpoints when next_id is not NULL; next_id and
 * last_time_written then receive the stored counters. */
int shm_read_checkpoint(shm_ctx *ctx, const char *folder, int *next_id,
                        double *last_time_written, int noutput,
                        int idx_checkpoints);

/* Checkpoint::Delete (checkpoint.cpp:277-308): removes the files of a
 * checkpoint written by shm_write_checkpoint, then its folders if empty. */
int shm_delete_checkpoint(shm_ctx *ctx, const char *folder);

/* IOModule::WriteGravitationalWaveSpectrum (io_module.cpp:334-353) into
 * prefix/spectra.hdf5: Header = t, N zp, number of bins, L zp, zp, unbinned;
 * k (|k|^2 per bin, or with output.gw_spectra.log_bins = n > 0 the lower
 * |k| edge of n log bins, plus a dataset log_bins); Spectrum.  Settings from
 * output.gw_spectra.{projection_type, zero_padding_factor, unbinned}.
 * modifier / params as shm_gw_spectrum_mod. */
int shm_write_gw_spectrum(shm_ctx *ctx, const char *prefix, int modifier,
                          const double *params, int nparams);

/* IOModule::WriteSpectra (io_module.cpp:307-327, spectrum.cpp:101-112) into
 * prefix/spectra.hdf5: Header = t, N, number of bins; k_sq; one dataset per
 * output function fns[0..nfn), named after it. */
int shm_write_spectra(shm_ctx *ctx, const char *prefix, const int *fns, int nfn);

/* IOModule::WriteProjections (io_module.cpp:280-301, projection.cpp:76-86)
 * into prefix/projections.hdf5: Header = t, N of the projected level;
 * <name>_data and <name>_n per output function, with modes[i] (NULL: 0) and
 * the level limit output.projections.max_level. */
int shm_write_projections(shm_ctx *ctx, const char *prefix, const int *fns,
                          const int *modes, int nfn);

#ifdef __cplusplus
}
#endif

#endif /* SHM_H_ */

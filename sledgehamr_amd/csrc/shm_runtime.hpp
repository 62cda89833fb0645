// Host runtime: owns device memory, box/tile tables, copy plans and the
// stage sequences of the integrators for one project instance.
#pragma once

#include <hipfft/hipfft.h>
#include <sledgehamr.h>

#include <array>
#include <cfloat>
#include <map>
#include <set>
#include <tuple>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "shm_layout.hpp"
#include "shm_stage.hpp"
#include "shm_table.hpp"

namespace shm {

/** The per-GPU device scratch array (shm_te_tag_noalloc, level norms):
 *  fixed regions at its head, then the tag mask of the allocation-free
 *  tagging. */
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kScratchWords = 64;              // words per head region
constexpr size_t kScratchCountsOff = 0;        // tag counters (u64)
constexpr size_t kScratchCritOff = 512;        // te_crit (double)
constexpr size_t kScratchTagPrmOff = 1024;     // TagCellForRefinement params
constexpr size_t kScratchModPrmOff = 1536;     // TruncationModifier params
constexpr size_t kScratchNormOff = 2048;       // level-norm maxima (u64 bits)
constexpr size_t kScratchRecOff = 2560;        // coarse/fine fill record
constexpr size_t kScratchMaskOff = 4096;       // tag mask, one byte per cell
constexpr long long kScratchMaxCells = (long long)(kScratchBytes - kScratchMaskOff);

struct Box3 {
    int lo[3], hi[3];  // inclusive, global index space of the level
    int n(int d) const { return hi[d] - lo[d] + 1; }
    long long vol() const { return (long long)n(0) * n(1) * n(2); }
};

/** Owning device allocation. */
class DevMem {
  public:
    DevMem() = default;
    ~DevMem() { reset(); }
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    DevMem(DevMem &&o) noexcept : p_(o.p_), bytes_(o.bytes_) {
        o.p_ = nullptr;
        o.bytes_ = 0;
    }
    DevMem &operator=(DevMem &&o) noexcept {
        if (this != &o) {
            reset();
            p_ = o.p_;
            bytes_ = o.bytes_;
            o.p_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    void alloc(size_t bytes);  // zero-filled
    void reset();
    void *get() const { return p_; }
    template <class T>
    T *as() const { return static_cast<T *>(p_); }
    size_t bytes() const { return bytes_; }
    explicit operator bool() const { return p_ != nullptr; }

  private:
    void *p_ = nullptr;
    size_t bytes_ = 0;
};

/** Box layout of one level (or of a coarsened copy of one). */
struct Layout {
    int N = 0;          // domain cells per axis at this level
    int g = 0;          // ghost width
    bool periodic = true;
    bool independent = false;  // every box is its own periodic domain
    std::vector<Box3> boxes;
    std::vector<BoxDesc> hbox;
    DevMem dbox;
    long long cstride = 0;  // elements per component plane
    std::vector<long long> vprefix;  // valid cells per box (prefix, nbox+1)
    DevMem dvprefix;
    std::vector<long long> cprefix;  // coarse (half) cells per box (prefix)
    DevMem dcprefix;
    std::vector<HaloSeg> halo;       // ghosts read by stage kernels
    DevMem dhalo;
    long long halo_total = 0;
    std::vector<HaloSeg> halo_full;  // every ghost cell
    DevMem dhalo_full;
    long long halo_full_total = 0;
    std::map<std::tuple<int, int, int, int>, std::pair<std::shared_ptr<DevMem>, int>> tiles;

    void build(const std::vector<Box3> &b, int ghost, int domain_n,
               bool is_periodic, bool is_independent = false);
    const TileDesc *get_tiles(int ty, int tz, int zc_minor, int *ntiles, int tx = 64);
    long long nvalid() const { return vprefix.empty() ? 0 : vprefix.back(); }
};

/** Device state of one level: the old/new state pair and stage scratch. */
struct LevelState {
    Layout L;
    int ncomp = 0;
    DevMem U[2], Ulo[2];
    int cur = 0;               // index of grid_new in U
    bool ghosts_ok[2] = {false, false};
    DevMem K[2];               // SSPRK3 k accumulators (k1, k2)
    DevMem F[5];               // RKN forces
    DevMem Y2;                 // RKN second-stage field states (rkn_state)
    DevMem Y3;                 // RKN third-stage field states (three stages)
    DevMem VH;                 // leapfrog half-step momenta (FP64)
    DevMem RHS;                // FillRhs target
    double t = -DBL_MAX;       // time of grid_new
    double t_old = -DBL_MAX;   // time of grid_old
    int istep = 0;
    // Coarse/fine ghost fill plan (levels above 0): the ghost-shell slabs of
    // the boxes, and per box the boxes of the coarser level that its coarse
    // interpolation stencil may touch.
    std::vector<CfPSeg> cf_segs;
    DevMem dcf_segs, dcf_start, dcf_cand;
    long long cf_total = 0;
    // Workgroup tiles of the segments and their largest coarse brick.
    DevMem dcf_tiles, dcf_full_tiles;
    int cf_ntile = 0, cf_full_ntile = 0, cf_lds = 0, cf_full_lds = 0;
    // The ghost shell a level step reads: the stencil radius instead of all
    // ghost layers.
    std::vector<CfPSeg> cf_step;
    DevMem dcf_step, dcf_step_tiles;
    long long cf_step_total = 0;
    int cf_step_ntile = 0, cf_step_lds = 0;
    // Every cell of each ghosted box (FillCoarsePatch).
    std::vector<CfPSeg> cf_full;
    DevMem dcf_full;
    long long cf_full_total = 0;
    bool cf_built = false;
    // Mask of the last tagging (one byte per valid cell) and its time.
    DevMem tags;
    double tags_t = -DBL_MAX;
    // This level's boxes coarsened onto level lev-1 (average-down and
    // truncation errors), with copy plans from and to level lev-1; rebuilt
    // when the epoch changes.
    std::unique_ptr<LevelState> mirror;
    DevMem mplan_in, mplan_out;
    int mnseg_in = 0, mnseg_out = 0;
    long long mtot_in = 0, mtot_out = 0;
    long mirror_epoch = -1;
    bool defined() const { return !L.boxes.empty(); }
};

/** Counters and per-level values of the local regrid (LocalRegrid's
 *  members, local_regrid.h:93-150). */
struct RegridState {
    int nregrids = 0, veto_level = -1;
    std::vector<long long> last_numPts;
    std::vector<char> no_local, do_global;
    std::vector<double> min_distance, latest;
    long long attempts = 0, global_wanted = 0;
};

/** Launch shape of the stage kernels.  A field of -1 (tz: 0) is not set:
 *  the level's tuned value, else the default, applies. */
struct LaunchCfg {
    int variant = -1;     // tile shape index into kVariants
    int tz = 0;           // z-chunk length of a tile (0: automatic)
    int xcd_split = -1;   // each XCD walks one contiguous range of tiles
    int persistent = -1;  // grid = resident capacity (1) or one tile per workgroup (0)
    int zc_minor = -1;    // tile order: z-chunk outermost (0) or innermost (1)
    int fused = -1;       // SSPRK3 stages one and two in one launch (0/1)
    int fused_variant = -1;  // tile shape of that launch
};

/** A captured level step (swap excluded): the stage launches and ghost
 *  fills of Integrator::Integrate for one (level, old-state slot), with its
 *  own pinned / device parameter records that the host refreshes before
 *  every replay. */
struct StepGraph {
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;   // recorded after the last replay
    double *host = nullptr;      // pinned parameter records (capacity nrec)
    DevMem dev;                  // device copies read by the kernels
    int nrec = 0;
    // What fills record i before a replay: kind 0 stage parameters, kind 1
    // coarse/fine time weights (2: first state only); at entry tidx of the
    // step's time table.
    std::vector<std::pair<int, int>> recs;
    long epoch = -1;             // layout / launch-shape generation captured
};

struct Timing {
    double stage_ms = 0, halo_ms = 0;
    long long stage_launches = 0, halo_launches = 0;
    long long stage_cells = 0;  // valid cells updated by stage launches
    // Launches of the fused stage-one-and-two kernel (also counted above).
    double fused_ms = 0;
    long long fused_launches = 0;
};

class Runtime {
  public:
    Runtime(sledgehamr::Sledgehamr *sim, const KernelTable &kt);
    ~Runtime();

    // Setup.
    void init_from_inputs(const std::string &inputs, int precision_override);
    void set_level(int lev, const std::vector<Box3> &boxes);

    // Data movement (valid cells, box-major, x fastest).
    void upload(int lev, int comp, const double *host);
    /** Upload from the reference's array orders: 1 the whole level with
     *  ind = i N^2 + j N + k (FillLevel::FromArray), 2 one chunk per box,
     *  box-relative ind = i ly lz + j lz + k (FillLevel::FromArrayChunks),
     *  chunks in box order. */
    void upload_array(int lev, int comp, const double *host, int layout);
    void download(int lev, int comp, double *host, bool old_state = false);
    void download_rhs(int lev, int comp, double *host);
    void init_state(int lev, int kind, uint64_t seed, const double *p, int np);

    // Reference operations.
    void fill_rhs(int lev, double time, double weight, bool add);
    /** LevelSynchronizer::FillCoarsePatch: every cell of grid_new[lev],
     *  ghosts included, interpolated from level lev-1 at the level's time. */
    void fill_coarse_patch(int lev);
    void advance(int lev);               // Integrator::Advance(lev)
    void create_shadow_level();          // Sledgehamr::CreateShadowLevel
    /** Maximum |value| per component over the valid cells of the current
     *  state of level lev (Sledgehamr::LevelNorm0). */
    std::vector<double> level_norm0(int lev);
    void te_tag(int lev, int flags, uint8_t *mask, long long *counts,
                double *te_host);
    /** te_tag without device allocation: counters, parameters, the fill
     *  record and the mask in the GPU's scratch array; the level's kept
     *  mask (for the local regrid) is left as it is. */
    void te_tag_noalloc(int lev, int flags, uint8_t *mask, long long *counts);
    /** GravitationalWaves::ComputeSpectrum with the project's spectrum
     *  modifier `modifier` (-1: default components, unmodified). */
    void gw_spectrum(int mode, int nbins, int zero_padding, int proj_type,
                     std::vector<double> &bins, std::vector<double> &kvals,
                     int modifier = -1, const std::vector<double> &mparams = {},
                     bool count_only = false);
    void step(int ncoarse);              // TimeStepper::Advance(0)
    /** Spectrum::Compute of output function fn on level 0: bins per unique
     *  |k|^2 (k_sq). */
    void spectrum(int fn, std::vector<double> &bins, std::vector<long long> &k_sq,
                  bool count_only = false);
    /** Projection::Compute of output function fn along z onto the finest
     *  level up to max_level (< 0: all levels); D x D pixels, x index slow. */
    void projection(int fn, int mode, int max_level, std::vector<double> &data,
                    std::vector<int> &count, int &D, bool count_only = false);
    int output_count() const { return kt_.nout; }
    const KernelTable &kernels() const { return kt_; }
    const char *output_name(int id) const {
        return kt_.out_name ? kt_.out_name(id) : nullptr;
    }
    /** Tag counters of the most recent scheduled truncation-error tagging
     *  of lev: [total, user, trunc[ncomp]]. */
    const std::vector<long long> &last_tags(int lev) const;

    // Local regrid (local_regrid/local_regrid.cpp).
    using Block = std::array<int, 3>;
    using BlockSet = std::set<Block>;
    /** LocalRegrid::DetermineNewBoxArray of tagging level lev from `mask`
     *  (host bytes, one per valid cell; null: the level's last tagging at
     *  its time): the blocks of level lev+1 to add, in (i, j, k) order, and
     *  the smallest measured squared distance (INT_MAX: none measured). */
    void regrid_select(int lev, const uint8_t *mask, BlockSet &blocks, int &min_dsq);
    /** LocalRegrid::AddBoxes: the level keeps its boxes and data and gains
     *  `boxes` after them, FillPatch-ed at the new and the old time. */
    void add_boxes(int lev, const std::vector<Box3> &boxes);
    /** Sledgehamr::RemakeLevel on `boxes`, then LocalRegrid::DidGlobalRegrid. */
    void remake_level(int lev, const std::vector<Box3> &boxes);
    /** LocalRegrid::DoAttemptRegrid from tagging level lev: 0 a global
     *  regrid is wanted, 1 boxes added (possibly none), 2 delayed. */
    int local_regrid(int lev);
    /** LocalRegrid::FixNesting of level l: blocks of level l-1 that the
     *  blocks added to level l need. */
    void fix_nesting(int l, std::vector<BlockSet> &add) const;
    /** UniqueLayout::BoxList: one box per run of consecutive k blocks. */
    static std::vector<Box3> block_boxes(const BlockSet &s, int bf);
    /** Runs joined along y, then x, where their other extents agree; sorted
     *  by (z, y, x) of their low corner. */
    static std::vector<Box3> merge_boxes(std::vector<Box3> v);
    const RegridState &regrid_state() const { return lr_; }
    int blocking_factor(int lev) const { return bf_[lev]; }

    void sync();
    void autotune(int lev);
    /** Plain read / write / copy bandwidth over buffers of `bytes` (GB/s). */
    void stream_probe(size_t bytes, int reps, double *read_gbs,
                      double *write_gbs, double *copy_gbs);
    void collect_timing();

    // State queries.
    int ncomp() const { return ncomp_; }
    int nfields() const { return kt_.nfields; }
    int radius() const { return R_; }
    int precision() const { return (int)prec_; }
    /** GW level stages run as three component groups. */
    bool gw_groups() const { return gw_groups_; }
    const std::vector<unsigned long long> &gw_dependencies() const { return gw_dep_; }
    LevelState &level(int lev);
    bool has_level(int lev) const {
        return lev >= 0 && lev < (int)levels_.size() && levels_[lev].defined();
    }
    int nlevels() const { return (int)levels_.size(); }
    sledgehamr::Sledgehamr *sim() { return sim_; }

    Timing timing;
    bool timing_on = false;
    LaunchCfg launch;   // user overrides of the launch shape
    bool graphs = false;   // replay captured level steps (shm.graphs)
    long long graph_replays = 0;
    /** Invalidates captured steps (layout or launch shape changed). */
    void bump_epoch() { ++epoch_; }
    /** Launch shape used for level lev (overrides, tuned values, defaults). */
    LaunchCfg resolved_launch(int lev) const;
    /** Fused SSPRK3 stages one and two possible at this level (every box
     *  periodic on itself, kernel built for the precision and radius). */
    bool fused_possible(const LevelState &ls) const;
    const FusedEntry *fused_row() const;  // fused12 or fusedrkn, by integrator

  private:
    struct EventPair {
        hipEvent_t a, b;
        int kind;  // 0 stage kernel, 1 halo kernel, 2 fused stages one and two
    };
    EventPair *timed_pair(int kind);
    std::vector<EventPair> ev_pool_;
    size_t ev_used_ = 0;
    void define_shadow();
    void fill_params(double *blk, double time, double dt, double dx, int lev);
    void ensure_arrays(LevelState &ls, int integrator);
    size_t esize() const { return prec_ == Prec::F64 ? 8 : 4; }
    size_t losize() const { return sprec() == Prec::F32C16 ? 2 : 4; }
    bool comp() const { return prec_ != Prec::F64; }
    /** Storage precision: the auxiliary kernels see only the stored arrays. */
    Prec sprec() const { return storage_prec(prec_); }
    /** Stencils read hi + lo (FP64 tiles of stored pairs): the remainders
     *  of the states a step reads must be valid at ghost cells too. */
    bool tile_lo() const { return prec_ == Prec::F32CS; }
    // Element size of the stored RKN stage states: the tile type.
    size_t ysize() const { return prec_ == Prec::F64 || tile_lo() ? 8 : 4; }
    /** Remainder array of state slot `which` when stencils read it. */
    void *tile_lo_of(LevelState &ls, int which) const {
        return tile_lo() ? ls.Ulo[which].get() : nullptr;
    }
    void launch_stage(int pol, LevelState &ls, StageArgs &a, int lev);
    /** Fused SSPRK3 stages one and two possible at this level (every box
     *  periodic on itself, kernel built for the precision and radius). */
    void launch_fused(LevelState &ls, const double *prm1, const double *prm2,
                      double dt, int lev);
    /** Same-level ghosts of state slot `which`: the leading words, and the
     *  remainders when the stencils read them. */
    void halo_state(LevelState &ls, int which, const std::vector<int> &comps);
    void halo(LevelState &ls, void *arr, const std::vector<int> &comps,
              bool full = false, int elem_bytes = 0);
    /** Ghost cells of grid_new of lev with their remainders, for functions
     *  that read the full-precision state at neighbours. */
    void fill_output_ghosts(int lev);
    void check_output_fn(int fn) const;
    /** FillPatch of state `which` of level lev at `time`: coarse/fine
     *  ghosts (levels above 0), then the same-level copy plan.  for_step:
     *  only what a level step reads on levels above 0 (step_ghost_comps,
     *  leading words). */
    void fill_patch(LevelState &ls, int which, int lev, double time, bool for_step = false);
    /** Level lev has ghost cells filled from level lev-1. */
    bool has_cf(int lev) const { return lev >= 1 && has_level(lev - 1); }
    void build_cf(int lev);
    /** Arguments of a coarse/fine fill of level lev at `time` (layout,
     *  coarse states and their time weights); the caller adds the target.
     *  rec_dev: device memory that takes the fill record instead of the
     *  record ring (the allocation-free tagging). */
    CfArgs cf_args(int lev, double time, bool first_state_only = false, bool step = false,
                   double *rec_dev = nullptr);
    /** Coarse/fine record of level lev at `time` (see CfArgs::rec). */
    void cf_record(double *rec, int lev, double time, bool first_state_only) const;
    /** A parameter record for the device: the next slot of the ring, or,
     *  while a step is captured, the graph's next record (filled again by
     *  the host before every replay from (kind, time index)). */
    const double *push_record(const std::vector<double> &blk, int kind, double time);
    std::vector<double> step_times_;  // stage times and t0 + dt of the step
    void cf_launch(CfArgs &a, int mode);
    /** kCfState fill of the components `comps` (empty: all) of a state of
     *  level lev; lo null: the leading words only. */
    void cf_fill_state(int lev, double time, void *hi, void *lo,
                       const std::vector<int> &comps = {}, bool step = false,
                       double *rec_dev = nullptr);
    /** Components whose coarse/fine ghost cells a level step reads: the
     *  stencil components; every component with dissipation, or for RKN,
     *  whose stage states at ghost cells compose phi0 + c pi0 + ... */
    std::vector<int> step_ghost_comps() const;
    /** SSPRK3 step of a level with coarse/fine ghost cells: the step-start
     *  state and both stage k arrays are filled at once, before stage one. */
    bool cf_merged(int lev) const { return has_cf(lev) && integrator_ == 10; }
    /** That one fill: u0 (leading words) at t0, k1 and k2 for t0 + dt and
     *  t0 + dt/2, in the ghost layers the stencil reads. */
    void cf_fill_ssprk(LevelState &ls, int lev, int old, double t0, double dt);
    /** Every cell of state `which` of level lev, ghosts included, from level
     *  lev-1 at `time` (first_state_only: FillCoarsePatch's choice). */
    void fill_coarse_full(int lev, int which, double time, bool first_state_only);
    /** Level lev on `boxes`: every cell from level lev-1, then the valid
     *  data of the level's previous layout; state grid_old too when `both`
     *  (else a copy of grid_new). */
    void relayout(int lev, const std::vector<Box3> &boxes, bool both);
    /** Blocks of level lev whose wrapped-index cell a box holds. */
    std::vector<uint8_t> block_cover(int lev) const;
    void did_global_regrid(int lev);
    void check_bf(int lev) const;
    /** Valid cells of component comp_id of state `which`: from the host
     *  array (to_state) or into it, in one kernel and one copy. */
    void valid_io(LevelState &ls, int comp_id, int which, double *host, bool to_state);
    /** The coarsened copy of level lev (>= 1) on level lev-1, built once
     *  per epoch with its copy plans. */
    LevelState &coarse_mirror(int lev);
    /** Copy of ncomp components along a plan between two layouts, for the
     *  state (es = element size) and, compensated, its remainder. */
    void copy_state(void *dst, void *dstlo, const Layout &dl, const void *src,
                    const void *srclo, const Layout &sl, const DevMem &plan, int nseg,
                    long long total);
    void check_new_boxes(int lev, const std::vector<Box3> &boxes,
                         const std::vector<Box3> &keep) const;
    void integrate(LevelState &ls, int lev, double dt, double dx,
                   double t_old);
    /** Times at which the integrator evaluates its stages. */
    std::vector<double> stage_times(double t0, double dt) const;
    bool replay_step(LevelState &ls, int lev, double dt, double dx, double t0);
    void release_graphs();
    // Parameter source while a step is being captured: stage i reads the
    // graph's device record i, copied from its pinned record inside the graph.
    StepGraph *capture_ = nullptr;
    std::map<std::pair<int, int>, StepGraph> graphs_;  // (level, old slot)
    std::map<int, long> warm_;   // level -> epoch of its last eager step
    long epoch_ = 0;
    void probe();
    bool check_gw_groups();
    std::vector<int> stencil_comps() const;
    // (field, momentum) pairs of the RKN / leapfrog integrators: the
    // project's fields, then the six GW tensor fields (rkn.cpp:32-44).
    int pair_field(int p) const {
        return p < kt_.nfields ? p : kt_.nscalars + (p - kt_.nfields);
    }
    int pair_mom(int p) const {
        return p < kt_.nfields ? kt_.nfields + p : kt_.nscalars + 6 + (p - kt_.nfields);
    }
    std::vector<int> all_comps() const;
    bool te_pass_fused(int lev) const;
    void avg_down(LevelState &fine, int fine_which, LevelState &crse,
                  int crse_which);
    void scheduled_step(int lev);
    void synchronize_levels(int lev);

    sledgehamr::Sledgehamr *sim_;
    const KernelTable &kt_;
    Prec prec_ = Prec::F32C;
    int ncomp_ = 0;
    int R_ = 1;
    bool gw_ = false, ko_ = false;
    int interp_type_ = 2;  // amr.interpolation_type (0 or 2)
    bool independent_ = false;  // level-0 boxes are separate periodic domains
    std::map<int, std::vector<long long>> last_tags_;
    int integrator_ = 10;
    std::vector<double> nodes_, bbar_, b_;
    std::vector<std::vector<double>> tableau_;
    std::vector<LevelState> levels_;
    LevelState shadow_;
    hipStream_t stream_ = nullptr;
    int ncu_ = 256;
    DevMem prm_dev_;
    double *prm_host_ = nullptr;
    std::vector<hipEvent_t> prm_ev_;
    int prm_slot_ = 0;
    DevMem ko_dev_;
    std::vector<LaunchCfg> tuned_;  // per level + 1 (index 0: shadow level)
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    // regrid/TE schedule of the time stepper
    std::vector<std::pair<int, double>> schedule_;
    std::vector<double> regrid_dt_, last_regrid_;
    double te_lowest_time_ = -DBL_MAX;
    // Local regrid inputs (amr.blocking_factor per level, amr.n_error_buf,
    // amr.max_local_regrids, amr.volume_threshold_strong / _weak) and state.
    std::vector<int> bf_;
    int n_error_buf_ = 1, max_local_regrids_ = 10;
    double vt_strong_ = 1.05, vt_weak_ = 1.1;
    bool local_regrid_on_ = false;  // shm.local_regrid
    bool keep_layout_on_global_ = false;  // shm.keep_layout_on_global_regrid
    int rkn_state_ = 1;             // shm.rkn_state
    // GW stages as three component groups (shm.gw_split and the dependency
    // probe); the probe's masks per output.
    bool gw_groups_ = false;
    std::vector<unsigned long long> gw_dep_;
    // Tagging's inputs and outputs, kept between calls (each call ends with
    // a stream synchronisation) so that tagging allocates nothing per call.
    DevMem te_crit_, te_ptag_, te_pmod_, te_mask_, te_counts_, te_out_;
    unsigned char *scratch_ = nullptr;  // this GPU's scratch array

  public:
    /** Order of the level operations of shm_evolve: 'S' a regrid scheduled
     *  by lev for time t, 'H' shadow level created, 'A' lev advanced to t,
     *  'D' lev+1 averaged onto lev, 'T' truncation errors and tags of lev,
     *  'R' regrid from lev at t. */
    struct Event {
        char kind;
        int lev;
        double t;
    };
    std::vector<Event> events_;

  private:
    /** GW spectrum work space, kept between calls: the D2Z plan and buffers
     *  for one padded size, the k tables for one (size, projection type,
     *  binning). */
    struct GwSpecCache {
        int N = 0, proj = 0, mode = -1, nbins = 0;
        hipfftHandle plan = 0;
        bool have_plan = false;
        DevMem dense, out[6], ditk, dsq, dbins, dparams;
        std::vector<double> kvals;
        int nbin_eff = 0;
    } gwc_;
    void release_gw_cache();
    /** Scalar spectrum work space (plan, buffers, unique |k|^2 table) for
     *  one level size, kept between calls. */
    struct OutSpecCache {
        int N = 0;
        hipfftHandle plan = 0;
        bool have_plan = false;
        DevMem dense, F, dsq, dbins, dparams;
        std::vector<long long> k_sq;
    } osc_;
    void release_out_cache();
    // Component names: the project's scalars, then the GW tensor.
    std::vector<std::string> names_;

  public:
    // HDF5 files (shm_io.cpp).
    /** FillLevel::FromInitialStateFile (fill_level.cpp:17-30): a checkpoint
     *  folder restarts every level, any other path is read as initial-state
     *  data for level lev. */
    void read_initial_state(int lev, const std::string &path);
    /** Checkpoint::Write (output_types/checkpoint.cpp:18-98). */
    void write_checkpoint(const std::string &folder, const std::vector<int> &next_id,
                          const std::vector<double> &last_time, int idx_checkpoints);
    /** Checkpoint::Read + UpdateLevels + UpdateOutputModules
     *  (checkpoint.cpp:122-187, 205-275). */
    void read_checkpoint(const std::string &folder, int noutput, int idx_checkpoints,
                         std::vector<int> &next_id, std::vector<double> &last_time);
    /** Checkpoint::Delete (checkpoint.cpp:277-308). */
    void delete_checkpoint(const std::string &folder);
    /** IOModule::WriteGravitationalWaveSpectrum (io_module.cpp:334-353). */
    void write_gw_spectrum_file(const std::string &prefix, int modifier,
                                const std::vector<double> &mparams);
    /** IOModule::WriteSpectra (io_module.cpp:307-327). */
    void write_spectra_file(const std::string &prefix, const std::vector<int> &fns);
    /** IOModule::WriteProjections (io_module.cpp:280-301). */
    void write_projections_file(const std::string &prefix, const std::vector<int> &fns,
                                const std::vector<int> &modes);
    const std::vector<std::string> &comp_names() const { return names_; }

  private:
    void read_hdf5_level(int lev, const std::string &path);
    /** Device buffers one tagging writes: context buffers (te_tag) or
     *  regions of the scratch array (te_tag_noalloc). */
    struct TagBuffers {
        uint8_t *mask = nullptr;
        unsigned long long *counts = nullptr;
        double *crit = nullptr;
        double *te = nullptr;
        DevMem *ptag = nullptr, *pmod = nullptr;  // context parameter buffers
        double *ptag_dev = nullptr, *pmod_dev = nullptr;  // or fixed regions
        double *cf_rec = nullptr;  // coarse/fine record target (null: ring)
        bool keep_tags = false;    // the mask is the level's kept mask
    };
    void te_tag_run(int lev, int flags, TagBuffers &tb, uint8_t *mask,
                    long long *counts, double *te_host);
    /** RKN steps of level lev whose first stage also writes the second
     *  stage's field states (first node 0, three or more stages, no
     *  coarse/fine ghosts). */
    bool rkn_y(int lev) const;
    RegridState lr_;
};

/** Last error text for the C ABI (per context). */
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

}  // namespace shm

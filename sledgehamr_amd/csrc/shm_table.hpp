// Kernel table: the per-project set of instantiated MI355X kernels, seen by
// the generic runtime as plain function pointers.
#pragma once

#include <hip/hip_runtime.h>

#include "shm_layout.hpp"

namespace shm {

struct StageArgs;
struct FusedArgs;
struct TeTagArgs;
struct ProbeOut;
struct BinArgs;

/** One instantiated stage program per kind: the integrator stage is a
 *  compile-time property of the kernel, so its store path has no branches. */
enum StageKind : int {
    kSsprk1 = 0,   // k1 = f(u0)
    kSsprk2,       // k2 = f(u0 + dt k1) + k1
    kSsprk3,       // k3 = f(u0 + dt/4 k2) + k2/4; u = u0 + 2dt/3 k3
    kFillRhs,      // rhs = f(u)
    kFillAddRhs,   // rhs = f(u) + w rhs
    kRkn0,         // RKN non-final stage with 0..3 earlier stages
    kRkn1,
    kRkn2,
    kRkn3,
    kRknF0,        // RKN final stage with 0..4 earlier stages
    kRknF1,
    kRknF2,
    kRknF3,
    kRknF4,
    kLeap1,        // leapfrog kick + drift
    kLeap2,        // leapfrog second kick
    kRkn0Y,        // RKN first stage at node 0 that also writes the second
                   // stage's field states
    kRkn1Y,        // RKN second stage reading them
    kRkn1YY,       // the same in a three-stage step, also writing the third
                   // stage's field states
    kRknF2Y,       // final stage of a three-stage step reading them
    kNKind
};

/** Tile shape variants: NW wavefronts per workgroup, RY rows per thread, TX
 *  cells in x (tile = TX x (NW * 64 / TX) * RY cells in x/y).  The 32-wide
 *  tile serves boxes narrower than 64 cells.  The autotuner picks one per
 *  level. */
struct TileVariant {
    int nw, ry, tx;
};
constexpr int kNVar = 5;
constexpr TileVariant kVariants[kNVar] = {
    {4, 1, 64}, {8, 1, 64}, {16, 1, 64}, {4, 1, 32}, {4, 2, 64}};
constexpr int kDefaultVariant = 1;

using StageFn = hipError_t (*)(const StageArgs &, int grid, hipStream_t);
using FusedFn = hipError_t (*)(const FusedArgs &, int grid, hipStream_t);
using OccFn = int (*)();
using TeTagFn = hipError_t (*)(const TeTagArgs &, int grid, hipStream_t);
using ProbeFn = hipError_t (*)(ProbeOut *, const double *, const double *,
                               double, double, double, hipStream_t);
using OutFn = hipError_t (*)(const OutArgs &, int grid, hipStream_t);
using GwBinFn = hipError_t (*)(const BinArgs &, int grid, hipStream_t);
/** Dependency probe of the cell functions: dep[n] = bit mask of the
 *  components output n reads (by perturbation). */
using GwDepFn = hipError_t (*)(unsigned long long *dep, const double *params,
                               const double *gw_params, double time, double dt,
                               double dx, hipStream_t);

struct StageEntry {
    StageFn fn = nullptr;
    OccFn occupancy = nullptr;  // resident workgroups per CU
    int threads = 0;
    int ty = 0;
    int tx = 64;
};

/** Fused first two stages of SSPRK3 or RKN (tile 64 x nw). */
struct FusedEntry {
    FusedFn fn = nullptr;
    int (*occupancy)() = nullptr;
    int threads = 0;
    int ty = 0;
};

struct KernelTable {
    int nscalars = 0, nfields = 0, ngw = 0;
    bool built_gw = false, built_ko = false;
    int r_mask = 0;
    // [kind][precision][R][gw][ko][variant]; every precision (shm::Prec).
    StageEntry stage[kNKind][4][4][2][2][kNVar];
    FusedEntry fused12[4][4][2][kNVar];   // [precision][R][gw][variant]
    FusedEntry fusedrkn[4][4][2][kNVar];  // the same for RKN stages one and two
    // Split stages of GW levels without dissipation: three group launches
    // per stage ([kind][precision][R][group][variant]), the groups' output
    // and staged component masks, and the probe that validates the split.
    StageEntry stage_grp[kNKind][4][4][3][kNVar];
    unsigned long long grp_out[3] = {}, grp_staged[3] = {};
    GwDepFn gwdep = nullptr;
    // Auxiliary kernels per storage precision (F64, F32C, F32C16).
    TeTagFn te_tag[3][2] = {};  // [storage precision][gw]
    ProbeFn probe[2] = {};      // [gw]
    // Output functions for spectra and projections (SLEDGEHAMR_OUTPUT_FUNCTIONS).
    int nout = 0;
    const char *(*out_name)(int) = nullptr;
    OutFn output[3] = {};       // [storage precision]
    // GW spectrum binning with the project's spectrum modifiers
    // (SLEDGEHAMR_GW_SPECTRUM_MODIFIERS).
    GwBinFn gw_bin = nullptr;
    int ngwmod = 0;
    const char *(*gwmod_name)(int) = nullptr;
    bool (*gwmod_select)(int, int *) = nullptr;
};

}  // namespace shm

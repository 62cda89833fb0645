// Kernel table: the per-project set of instantiated MI355X kernels, seen by
// the generic runtime as plain function pointers.
#pragma once

#include <hip/hip_runtime.h>

#include "shm_layout.hpp"

namespace shm {

struct StageArgs;
struct TeTagArgs;
struct ProbeOut;

enum PolId : int { kPolSsprk = 0, kPolRkn = 1, kPolLeapfrog = 2, kNPol = 3 };

/** Tile shape variants: NW wavefronts per workgroup, RY rows per thread
 *  (tile = 64 x NW*RY cells in x/y).  The autotuner picks one per level. */
struct TileVariant {
    int nw, ry;
};
constexpr int kNVar = 3;
constexpr TileVariant kVariants[kNVar] = {{4, 1}, {4, 2}, {4, 4}};

using StageFn = hipError_t (*)(const StageArgs &, int grid, hipStream_t);
using OccFn = int (*)();
using TeTagFn = hipError_t (*)(const TeTagArgs &, int grid, hipStream_t);
using ProbeFn = hipError_t (*)(ProbeOut *, const double *, const double *,
                               double, double, double, hipStream_t);

struct StageEntry {
    StageFn fn = nullptr;
    OccFn occupancy = nullptr;  // resident workgroups per CU
    int threads = 0;
    int ty = 0;
};

struct KernelTable {
    int nscalars = 0, nfields = 0, ngw = 0;
    bool built_gw = false, built_ko = false;
    int r_mask = 0;
    // [policy][precision][R][gw][ko][variant]
    StageEntry stage[kNPol][2][4][2][2][kNVar];
    TeTagFn te_tag[2][2] = {};  // [precision][gw]
    ProbeFn probe[2] = {};      // [gw]
};

}  // namespace shm

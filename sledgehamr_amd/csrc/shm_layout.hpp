// Plain-old-data descriptors shared by host runtime and device kernels.
//
// HBM layout of a level ("component planes"): every array of a level holds
// ncomp planes of `cstride` elements.  Inside a plane each box owns a padded
// brick: x fastest, then y, then z (the Array4 order of the reference), with
// g ghost layers on every side.  The first valid cell (lo) of a box sits at
// element `base`, which is 32-element aligned, and row/plane pitches are
// multiples of 32 elements, so every 64-wide x row of valid cells starts on a
// 128 B (FP32) / 256 B (FP64) boundary.
#pragma once

#include <cstdint>

namespace shm {

struct BoxDesc {
    int lo[3];        // global index of the first valid cell
    int n[3];         // valid extents
    long long base;   // element offset of cell lo inside a component plane
    long long org;    // element offset of cell lo - (g, g, g): every cell of
                      // the ghosted brick is org + an unsigned 32-bit offset
    int jstride;      // elements per row
    int kstride;      // elements per xy-plane
    int wrap;         // bit a: the box spans the periodic domain along axis
                      // a, so stage kernels wrap that index instead of
                      // reading ghost cells
    int pad_;
};

/** One work unit of a fused stage kernel: an x/y tile of one box marched
 *  along z over [z0, z1).  Coordinates are box-local (0 = first valid). */
struct TileDesc {
    int box;
    int x0, y0;
    int z0, z1;
};

/** Ghost-cell copy segment: destination cells [lo, lo+n) (global index, in
 *  dst_box's ghost region) take the values of src_box at index - shift,
 *  where shift is a multiple of the domain size per axis (periodic wrap). */
struct HaloSeg {
    int dst_box, src_box;
    int lo[3];
    int n[3];
    int shift[3];
    long long start;  // prefix sum of segment volumes
};

/** Coarse/fine pairing for truncation errors: one coarse box per fine box
 *  (the coarse box is the fine box coarsened by 2). */
struct TePair {
    int fine_box;
    int crse_box;
};

/** Coarse/fine ghost fill: what a ghost cell receives (see cf_fill_kernel). */
enum CfMode { kCfState = 0, kCfK = 1, kCfForce = 2, kCfSsprk = 3 };

/** Coarse parent cells of a region of fine cells [flo, flo + fn) of fine
 *  box dst_box: parent p covers fine cells 2p and 2p + 1 along each axis. */
struct CfPSeg {
    int dst_box;
    int plo[3], pn[3];
    int flo[3], fn[3];
    long long start;  // prefix sum of parent cells
};

/** A workgroup's share of a CfPSeg: parents plo .. plo + pn - 1. */
struct CfTile {
    int seg;
    int plo[3], pn[3];
    int pad_;
};

struct CfArgs {
    const BoxDesc *fboxes;
    const BoxDesc *cboxes;
    const CfPSeg *segs;        // parents of ghost-shell slabs (or whole boxes)
    const CfTile *tiles;       // one per workgroup
    int ntile;
    int lds_cells;             // largest coarse brick of a tile (doubles)
    long long total;           // parent cells
    const int *cand_start;     // per fine box: range in cand
    const int *cand;           // coarse boxes the box's coarse stencil touches
    int Nc;                    // coarse domain cells per axis
    long long fcstride, ccstride;
    const void *cs[2], *cslo[2];  // the coarse level's two state slots
    // Record read on the device (so captured steps take fresh values):
    // [0] weight of the old state, [1] of the new one, [2] old state used,
    // [3] new state used, [4] slot of the new state.
    const double *rec;
    // kCfSsprk: records of t0 + dt and t0 + dt/2 (rec: t0); the state's
    // leading words go to dst, k1 and k2 to kd[0] and kd[1], with c = cf[0]
    // and cf[1] (dt, dt/4).
    const double *rec3[2];
    void *kd[2];
    int ncomp;
    int comp[32];              // interpolated component (fine = coarse index)
    void *dst, *dstlo;         // state (kCfState; dstlo null: leading words
                               // only), k (kCfK), F_m (kCfForce)
    const void *u0;            // fine u0 / phi0, pi0 (kCfK, kCfForce)
    const void *u0lo;          // their remainders (FP64 tiles of stored pairs)
    double c0;                 // kCfK: c; kCfForce: coefficient of pi0
    int mom[32];               // kCfForce: momentum component of each pair
    int pair[32];              // kCfForce: pair index (force component)
    int m;                     // kCfForce: index of the force written
    const void *f[4];          // kCfForce: F_0 .. F_{m-1}
    double cf[5];              // kCfForce: coefficients of F_0 .. F_m
};

/** Evaluation of a project output function over the valid cells of a level
 *  (spectra: into a dense N^3 array; projections: summed or maximised along
 *  z onto the D x D grid of the finest projected level). */
struct OutArgs {
    const BoxDesc *boxes;
    const long long *vprefix;  // valid cells per box, prefix (nbox + 1)
    int nbox;
    long long total;
    long long cstride;
    int ncomp;
    const void *arr, *arrlo;   // state (with its remainder)
    int fn;                    // output function id
    const double *params;
    int lev;
    double time, dt, dx;
    int target;                // 0 dense, 1 projection
    double *dense;             // target 0: N^3, x fastest
    int N;
    const int *cov_start;      // target 1, per box: range of cov
    const int *cov;            // finer-level boxes, lo[3] hi[3] each
    int ratio, D, mode;        // mode 0 sum of ratio * value, 1 maximum
    double *proj;              // D * D, x index slow
    int *nproj;
};

/** A box of a tagging level for the local regrid's block selection: its
 *  valid cells, the offset of its first mask byte, and its border map, the
 *  1-based wrapped block coordinates c0 - 1 .. c0 + xs of the finer level
 *  (x fastest), from entry boff on. */
struct RegridBox {
    int lo[3], n[3];
    int c0[3], xs[3];
    long long voff;
    long long boff;
};

struct RegridArgs {
    const RegridBox *boxes;
    int nbox;
    long long nvalid;
    const uint8_t *mask;    // one byte per valid cell, box-major, x fastest
    const uint8_t *border;  // 1: the block holds no finer cell at its centre
    int *first;             // per border entry: first cell adding it (box-local)
    uint8_t *added;         // Np^3 blocks of the finer level, x fastest
    int *min_dsq;
    int bf, Np, thr;        // blocking factor, blocks per axis, (n_error_buf+1)^2
};

/** Valid cells of one component of a level and a dense FP64 array of them
 *  (box-major, x fastest). */
struct ValidIoArgs {
    const BoxDesc *boxes;
    const long long *vprefix;  // nbox + 1 prefix sums of valid cells
    int nbox;
    long long total;
    void *hi, *lo;             // state arrays (lo null without compensation)
    long long comp_off;        // component * plane stride
    double *dense;
};

}  // namespace shm

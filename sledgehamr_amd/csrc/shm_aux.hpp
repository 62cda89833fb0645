// Truncation-error / tagging, gravitational-wave spectrum and initial-state
// kernels for gfx950.
#pragma once

#include <sledgehamr.h>

#include <climits>
#include <cstdint>

#include "shm_layout.hpp"
#include "shm_stage.hpp"

namespace shm {

// ---------------------------------------------------------------------------
// Counter-based random numbers shared by device generators and the oracle:
// splitmix64 of (seed, stream, cell index) -> uniform in [0, 1).
// ---------------------------------------------------------------------------

__host__ __device__ AMREX_FORCE_INLINE uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__host__ __device__ AMREX_FORCE_INLINE double rng_uniform(uint64_t seed,
                                                          uint64_t stream,
                                                          uint64_t idx) {
    const uint64_t h = splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + idx);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Fused average-down + truncation error + tagging.
// ---------------------------------------------------------------------------

struct TeTagArgs {
    const BoxDesc *fboxes;   // fine boxes
    const BoxDesc *cboxes;   // coarse box holding each fine box's parent cells
    long long fcstride, ccstride;
    const void *fine;        // fine state (hi part for FP32C)
    const void *finelo;      // fine remainder (FP32C)
    void *crse;              // coarse state, averaged in place when requested
    void *crselo;            // coarse remainder (FP32C)
    const long long *cprefix;  // coarse cells per fine box, prefix (nbox+1)
    const long long *mprefix;  // fine valid cells per fine box, prefix
    int nbox;
    long long total;
    int ncomp;
    double *te_out;          // optional: te per coarse cell, ncomp planes of total
    uint8_t *mask;           // fine-cell tags (0/1), box-major x-fastest
    unsigned long long *counts;  // [total, user, trunc[ncomp]]
    const double *te_crit;
    const double *params_tag;
    const double *params_mod;
    int lev;
    double time, dt, dx;
    int flags;
};

enum TeTagFlags : int {
    kTeCompute = 1,   // compute truncation errors (else user tags only)
    kTeAvgDown = 2,   // write the child mean into the coarse state
    kTeTags = 4,      // write the tag mask and counters
};

/** Plain view of a box of a stored state; with a remainder the view reads
 *  hi + lo in FP64. */
template <Prec P>
__device__ AMREX_FORCE_INLINE amrex::Array4<const double> plain_view(
    const void *arr, const void *arrlo, const BoxDesc &b, long long cstride,
    int ncomp) {
    using T = PrecTraits<P>;
    amrex::Array4<const double> v;
    if constexpr (T::comp) {
        v.mode = std::is_same_v<typename T::L, float>
                     ? amrex::Array4<const double>::kPlain32C
                     : amrex::Array4<const double>::kPlain32C16;
        v.pf = static_cast<const float *>(arr) + b.base;
        v.plo = static_cast<const typename T::L *>(arrlo) + b.base;
    } else {
        v.mode = amrex::Array4<const double>::kPlain64;
        v.p = static_cast<const double *>(arr) + b.base;
    }
    v.jstride = b.jstride;
    v.kstride = b.kstride;
    v.nstride = cstride;
    v.begin = {b.lo[0], b.lo[1], b.lo[2]};
    v.end = {b.lo[0] + b.n[0], b.lo[1] + b.n[1], b.lo[2] + b.n[2]};
    v.ncomp = ncomp;
    return v;
}

__device__ AMREX_FORCE_INLINE unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

/** One thread per coarse cell.  Reproduces the effective CPU semantics of
 *  the reference tag loop (macros.h:489-566): a fine cell is tagged if its
 *  user criterion holds, or if it is the all-even child of its parent and
 *  some component passes TruncationModifier<n>(te) >= te_crit[n] (FP64).
 *  Counters follow macros.h:509-527 (8 per passing component, 8 - user for
 *  the block total). */
template <class Ops, Prec P, bool GW>
__global__ void __launch_bounds__(256) te_tag_kernel(const TeTagArgs a) {
    using S = typename PrecTraits<P>::S;
    constexpr int NC = GW ? Ops::kNGw : Ops::kNScalars;
    const long long gstride = (long long)gridDim.x * blockDim.x;
    unsigned long long c_total = 0, c_user = 0, c_trunc[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) c_trunc[n] = 0;

    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < a.total; t += gstride) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.cprefix[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int b = lo;
        const BoxDesc fb = a.fboxes[b];
        const BoxDesc cb = a.cboxes[b];
        const int cnx = fb.n[0] >> 1, cny = fb.n[1] >> 1;
        long long r = t - a.cprefix[b];
        const int cx = (int)(r % cnx);
        r /= cnx;
        const int cy = (int)(r % cny);
        const int cz = (int)(r / cny);
        const int ii = fb.lo[0] + 2 * cx, jj = fb.lo[1] + 2 * cy,
                  kk = fb.lo[2] + 2 * cz;
        const int ci = (fb.lo[0] >> 1) + cx, cj = (fb.lo[1] >> 1) + cy,
                  ck = (fb.lo[2] >> 1) + cz;
        const amrex::Array4<const double> fv =
            plain_view<P>(a.fine, a.finelo, fb, a.fcstride, NC);

        bool te_pass = false;
        if (a.flags & kTeCompute) {
            const long long ce = cb.base + (ci - cb.lo[0]) +
                                 (long long)(cj - cb.lo[1]) * cb.jstride +
                                 (long long)(ck - cb.lo[2]) * cb.kstride;
            sledgehamr::utils::constexpr_for<0, NC, 1>([&](auto n) {
                auto f = [&](int x, int y, int z) { return fv(x, y, z, n); };
                const double sum = child_sum(f, ii, jj, kk);
                const long long cen = (long long)n * a.ccstride + ce;
                const double crse =
                    load_state_at<PrecTraits<P>>(a.crse, a.crselo, cen);
                double avg;
                const double te = te_cell(crse, sum, &avg);
                if (a.flags & kTeAvgDown) store_state_at<PrecTraits<P>>(a.crse, a.crselo, cen, avg);
                if (a.te_out) a.te_out[(long long)n * a.total + t] = te;
                const double m = Ops::template te_modifier<n>(
                    fv, ii, jj, kk, a.lev, a.time, a.dt, a.dx, te, a.params_mod);
                if (m >= a.te_crit[n]) {
                    te_pass = true;
                    c_trunc[n] += 8;
                }
            });
        }
        if (a.flags & kTeTags) {
            const int fnx = fb.n[0], fny = fb.n[1];
            uint8_t *mb = a.mask + a.mprefix[b];
            int user_even = 0;
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                const int dx_ = d & 1, dy_ = (d >> 1) & 1, dz_ = d >> 2;
                const bool u = Ops::tag_cell(fv, ii + dx_, jj + dy_, kk + dz_,
                                             a.lev, a.time, a.dt, a.dx,
                                             a.params_tag);
                if (d == 0) user_even = u;
                c_user += u;
                c_total += u;
                const bool tag = u || (d == 0 && te_pass);
                mb[(2 * cx + dx_) + (long long)(2 * cy + dy_) * fnx +
                   (long long)(2 * cz + dz_) * fnx * fny] = tag ? 1 : 0;
            }
            if (te_pass) c_total += 8 - user_even;
        }
    }
    if (a.flags & kTeTags) {
        const unsigned long long s_total = wave_sum(c_total);
        const unsigned long long s_user = wave_sum(c_user);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.counts[0], s_total);
            atomicAdd(&a.counts[1], s_user);
        }
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            const unsigned long long s = wave_sum(c_trunc[n]);
            if ((threadIdx.x & 63) == 0) atomicAdd(&a.counts[2 + n], s);
        }
    }
}

// ---------------------------------------------------------------------------
// Box-to-box copy between two layouts (coarse data onto a coarsened-fine
// layout and back, shadow levels).  Segments index dst boxes; src cells are
// dst cells minus `shift`.
// ---------------------------------------------------------------------------

struct CopyArgs {
    void *dst;
    const void *src;
    long long dcstride, scstride;
    const BoxDesc *dboxes;
    const BoxDesc *sboxes;
    const HaloSeg *segs;
    int nseg;
    long long total;
    int ncomp;
};

template <class T>
__global__ void __launch_bounds__(256) copy_kernel(const CopyArgs a) {
    T *dst = static_cast<T *>(a.dst);
    const T *src = static_cast<const T *>(a.src);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < a.total; t += stride) {
        int lo = 0, hi = a.nseg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.segs[mid].start <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const HaloSeg &sg = a.segs[lo];
        long long r = t - sg.start;
        const int x = (int)(r % sg.n[0]);
        r /= sg.n[0];
        const int y = (int)(r % sg.n[1]);
        const int z = (int)(r / sg.n[1]);
        const int gi = sg.lo[0] + x, gj = sg.lo[1] + y, gk = sg.lo[2] + z;
        const BoxDesc &db = a.dboxes[sg.dst_box];
        const BoxDesc &sb = a.sboxes[sg.src_box];
        const long long de = db.base + (gi - db.lo[0]) +
                             (long long)(gj - db.lo[1]) * db.jstride +
                             (long long)(gk - db.lo[2]) * db.kstride;
        const long long se = sb.base + (gi - sg.shift[0] - sb.lo[0]) +
                             (long long)(gj - sg.shift[1] - sb.lo[1]) * sb.jstride +
                             (long long)(gk - sg.shift[2] - sb.lo[2]) * sb.kstride;
        for (int c = 0; c < a.ncomp; ++c)
            dst[(long long)c * a.dcstride + de] = src[(long long)c * a.scstride + se];
    }
}

/** 8-cell mean of a fine level onto a coarser layout (amrex::average_down,
 *  used to build the shadow level and to synchronise levels): one thread
 *  per coarse cell of each fine box, FP64 sums in child_sum order. */
struct AvgArgs {
    const BoxDesc *fboxes;
    const BoxDesc *cboxes;
    long long fcstride, ccstride;
    const void *fine;
    const void *finelo;
    void *crse;
    void *crselo;
    const long long *cprefix;
    int nbox;
    long long total;
    int ncomp;
};

template <Prec P>
__global__ void __launch_bounds__(256) avgdown_kernel(const AvgArgs a) {
    using S = typename PrecTraits<P>::S;
    const long long gstride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < a.total; t += gstride) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.cprefix[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BoxDesc fb = a.fboxes[lo];
        const BoxDesc cb = a.cboxes[lo];
        const int cnx = fb.n[0] >> 1, cny = fb.n[1] >> 1;
        long long r = t - a.cprefix[lo];
        const int cx = (int)(r % cnx);
        r /= cnx;
        const int cy = (int)(r % cny);
        const int cz = (int)(r / cny);
        const int ci = (fb.lo[0] >> 1) + cx, cj = (fb.lo[1] >> 1) + cy,
                  ck = (fb.lo[2] >> 1) + cz;
        const long long ce = cb.base + (ci - cb.lo[0]) +
                             (long long)(cj - cb.lo[1]) * cb.jstride +
                             (long long)(ck - cb.lo[2]) * cb.kstride;
        const long long fe = fb.base + 2 * cx + (long long)(2 * cy) * fb.jstride +
                             (long long)(2 * cz) * fb.kstride;
        for (int n = 0; n < a.ncomp; ++n) {
            const long long off = (long long)n * a.fcstride + fe;
            auto f = [&](int x, int y, int z) {
                const long long e = off + x + (long long)y * fb.jstride +
                                    (long long)z * fb.kstride;
                return load_state_at<PrecTraits<P>>(a.fine, a.finelo, e);
            };
            const double sum = child_sum(f, 0, 0, 0);
            store_state_at<PrecTraits<P>>(a.crse, a.crselo,
                                          (long long)n * a.ccstride + ce, 0.125 * sum);
        }
    }
}

// ---------------------------------------------------------------------------
// Coarse/fine ghost fill (FillPatchTwoLevels, level_synchronizer.cpp:92-209).
// Every ghost cell of every fine box takes the coarse level interpolated in
// time and space; the same-level copy plan, run afterwards, overwrites the
// ghost cells other fine boxes cover.  Spatial interpolation, ratio 2, at
// the fine cell centre x, y, z = +-1/4 coarse cells from the parent centre:
//   0 piecewise constant: the parent value;
//   1 limited linear: c + sx x + sy y + sz z with monotonized-central
//              slopes (conservative, exact for linear data);
//   2 quadratic: c + sx x + sy y + sz z + sxx x^2 + syy y^2 + szz z^2
//              + sxy x y + sxz x z + syz y z from central differences of
//              the 19-point coarse neighbourhood (exact for polynomials of
//              degree <= 2);
//   4 quartic: tensor product of the conservative child weights
//              (3, -22, 128, 22, -3) / 128 over the 5^3 coarse neighbourhood.
// What a ghost cell receives depends on how the stage kernels form the
// state they read there:
//   kCfState  the state itself (value and remainder);
//   kCfK      k of a composed SSPRK3 stage state u0 + c k: (v - u0) / c;
//   kCfForce  the last force F_m of a composed RKN stage state
//             phi0 + c0 pi0 + sum_{j<=m} c_j F_j:
//             (v - phi0 - c0 pi0 - sum_{j<m} c_j F_j) / c_m.
// ---------------------------------------------------------------------------

__device__ AMREX_FORCE_INLINE int cf_wrap(int v, int n) {
    // Coarse indices lie within a period of the domain: no division.
    while (v < 0) v += n;
    while (v >= n) v -= n;
    return v;
}

/** The 3D indices (x fastest, nx x ny columns) of first, first + step,
 *  first + 2 step, ... without a division per step. */
struct Walk3 {
    int x, y, z, dx, dy, dz, nx, ny;
    __device__ Walk3(int first, int step, int nx_, int ny_) : nx(nx_), ny(ny_) {
        x = first % nx;
        y = (first / nx) % ny;
        z = first / (nx * ny);
        dx = step % nx;
        dy = (step / nx) % ny;
        dz = step / (nx * ny);
    }
    __device__ void next() {
        x += dx;
        int c = x >= nx;
        if (c) x -= nx;
        y += dy + c;
        c = y >= ny;
        if (c) y -= ny;
        z += dz + c;
    }
};

template <Prec P, int MODE, int INTERP>
__global__ void __launch_bounds__(256) cf_fill_kernel(const CfArgs a) {
    using T = PrecTraits<P>;
    using S = typename T::S;
    // Stage states are composed in the tile type V from u0's tile value (the
    // leading word for FP32 tiles, hi + lo for FP64 tiles) and coefficients
    // in V: the ghost values below compose to the interpolated state the
    // same way.
    using V = typename T::V;
    auto u0_at = [&](long long e) -> double {
        return (double)ld_tile<T>(a.u0, a.u0lo, e, 0u);
    };
    // Time-weighted coarse values of the tile's brick: its parents grown by
    // the stencil reach, x fastest.
    extern __shared__ double brick[];
    // Points of the type 0 / 1 / 2 stencils: centre, 6 faces, then the 12
    // edges as (+,+) (-,+) (+,-) (-,-) pairs.
    constexpr int NPT = INTERP == 0 ? 1 : INTERP == 1 ? 7 : 19;
    constexpr int R = INTERP == 0 ? 0 : INTERP == 4 ? 2 : 1;
    constexpr int PO[19][3] = {
        {0, 0, 0},   {-1, 0, 0}, {1, 0, 0},  {0, -1, 0}, {0, 1, 0},
        {0, 0, -1},  {0, 0, 1},  {1, 1, 0},  {-1, 1, 0}, {1, -1, 0},
        {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
        {0, 1, 1},   {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
    const CfTile tl = a.tiles[blockIdx.x];
    const CfPSeg &sg = a.segs[tl.seg];
    const int fb = sg.dst_box;
    const BoxDesc &db = a.fboxes[fb];
    const int bxn = tl.pn[0] + 2 * R, byn = tl.pn[1] + 2 * R;
    const int nb = bxn * byn * (tl.pn[2] + 2 * R);
    const int np = tl.pn[0] * tl.pn[1] * tl.pn[2];
    const int cs0 = a.cand_start[fb], ncand = a.cand_start[fb + 1] - cs0;
    // Coarse states and their time weights (kCfSsprk: three times, one
    // brick each).
    constexpr int NTIME = MODE == kCfSsprk ? 3 : 1;
    double twO[NTIME], twN[NTIME];
    bool tuO[NTIME], tuN[NTIME];
#pragma unroll
    for (int ti = 0; ti < NTIME; ++ti) {
        const double *rc = ti == 0 ? a.rec : a.rec3[ti - 1];
        twO[ti] = rc[0];
        twN[ti] = rc[1];
        tuO[ti] = rc[2] != 0.;
        tuN[ti] = rc[3] != 0.;
    }
    const double wo = a.rec[0], wn = a.rec[1];
    const bool use_old = a.rec[2] != 0., use_new = a.rec[3] != 0.;
    const int ns = a.rec[4] != 0. ? 1 : 0;
    const void *cnew = a.cs[ns], *cnewlo = a.cslo[ns];
    const void *cold = a.cs[ns ^ 1], *coldlo = a.cslo[ns ^ 1];
    auto cval = [&](int comp, long long ce) -> double {
        const long long e = (long long)comp * a.ccstride + ce;
        double v = 0.;
        if (use_new) v = load_state_at<T>(cnew, cnewlo, e);
        if (use_old) {
            const double vo = load_state_at<T>(cold, coldlo, e);
            v = use_new ? wo * vo + wn * v : vo;
        }
        return v;
    };
    // What the ghost cell at fine element fe receives for component slot n.
    auto put = [&](int n, long long fe, double v) {
        const int comp = a.comp[n];
        if constexpr (MODE == kCfState) {
            const long long e = (long long)comp * a.fcstride + fe;
            // Without a remainder array (dstlo null) only the leading word.
            if constexpr (T::comp) {
                if (a.dstlo)
                    store_state_at<T>(a.dst, a.dstlo, e, v);
                else
                    st<float>(a.dst, e, 0u, (float)v);
            } else {
                store_state_at<T>(a.dst, a.dstlo, e, v);
            }
        } else if constexpr (MODE == kCfK) {
            const long long e = (long long)comp * a.fcstride + fe;
            const double u0 = u0_at(e);
            st<S>(a.dst, e, 0u, (S)((v - u0) / (double)(V)a.c0));
        } else {
            double yv = u0_at((long long)comp * a.fcstride + fe) +
                        (double)(V)a.c0 * u0_at((long long)a.mom[n] * a.fcstride + fe);
            const long long pe = (long long)a.pair[n] * a.fcstride + fe;
            for (int j = 0; j < a.m; ++j)
                yv += (double)(V)a.cf[j] * (double)ld<S>(a.f[j], pe, 0u);
            st<S>(a.dst, pe, 0u, (S)((v - yv) / (double)(V)a.cf[a.m]));
        }
    };
    for (int n = 0; n < a.ncomp; ++n) {
        const int comp = a.comp[n];
        __syncthreads();  // the previous component's brick has been read
        // Brick cells, coalesced along x: the periodic image of each coarse
        // cell in the candidate boxes of the fine box (every one is held:
        // nesting is checked when the plan is built).
        int hint = 0;
        Walk3 w(threadIdx.x, blockDim.x, bxn, byn);
        for (int q = threadIdx.x; q < nb; q += blockDim.x, w.next()) {
            const int ci = cf_wrap(tl.plo[0] - R + w.x, a.Nc);
            const int cj = cf_wrap(tl.plo[1] - R + w.y, a.Nc);
            const int ck = cf_wrap(tl.plo[2] - R + w.z, a.Nc);
            long long e = -1;
            for (int k = 0; k < ncand && e < 0; ++k) {
                const int qq = hint + k < ncand ? hint + k : hint + k - ncand;
                const BoxDesc &cb = a.cboxes[a.cand[cs0 + qq]];
                if (ci >= cb.lo[0] && ci < cb.lo[0] + cb.n[0] && cj >= cb.lo[1] &&
                    cj < cb.lo[1] + cb.n[1] && ck >= cb.lo[2] && ck < cb.lo[2] + cb.n[2]) {
                    e = cb.base + (ci - cb.lo[0]) + (long long)(cj - cb.lo[1]) * cb.jstride +
                        (long long)(ck - cb.lo[2]) * cb.kstride;
                    hint = qq;
                }
            }
            if constexpr (MODE == kCfSsprk) {
                // Both coarse states once, weighted per time as cval does.
                double vn = 0., vo = 0.;
                const long long ce = (long long)comp * a.ccstride + (e >= 0 ? e : 0);
                if (tuN[0] || tuN[1] || tuN[2]) vn = load_state_at<T>(cnew, cnewlo, ce);
                if (tuO[0] || tuO[1] || tuO[2]) vo = load_state_at<T>(cold, coldlo, ce);
#pragma unroll
                for (int ti = 0; ti < NTIME; ++ti) {
                    double v = 0.;
                    if (tuN[ti]) v = vn;
                    if (tuO[ti]) v = tuN[ti] ? twO[ti] * vo + twN[ti] * v : vo;
                    brick[ti * nb + q] = e >= 0 ? v : __builtin_nan("");
                }
            } else {
                brick[q] = e >= 0 ? cval(comp, e) : __builtin_nan("");
            }
        }
        __syncthreads();
        Walk3 pw(threadIdx.x, blockDim.x, tl.pn[0], tl.pn[1]);
        for (int p = threadIdx.x; p < np; p += blockDim.x, pw.next()) {
            const int px = pw.x, py = pw.y, pz = pw.z;
            const int pc[3] = {tl.plo[0] + px, tl.plo[1] + py, tl.plo[2] + pz};
            // Children 2 p + c of the parent inside the segment, per axis.
            int c0[3], c1[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                c0[d] = max(0, sg.flo[d] - 2 * pc[d]);
                c1[d] = min(1, sg.flo[d] + sg.fn[d] - 1 - 2 * pc[d]);
            }
            const int b0 = (px + R) + bxn * ((py + R) + byn * (pz + R));
            auto C = [&](int di, int dj, int dk) -> double {
                return brick[b0 + di + bxn * (dj + byn * dk)];
            };
            auto fine_elem = [&](int cx, int cy, int cz) -> long long {
                return db.base + (2 * pc[0] + cx - db.lo[0]) +
                       (long long)(2 * pc[1] + cy - db.lo[1]) * db.jstride +
                       (long long)(2 * pc[2] + cz - db.lo[2]) * db.kstride;
            };
            if constexpr (MODE == kCfSsprk) {
                // The interpolated value of child (x, y, z) from brick bb,
                // by the same expressions as the per-parent forms below.
                auto value = [&](const double *bb, int x, int y, int z) -> double {
                    auto Cb = [&](int di, int dj, int dk) -> double {
                        return bb[b0 + di + bxn * (dj + byn * dk)];
                    };
                    if constexpr (INTERP == 4) {
                        auto wu = [](int q) -> double {
                            return q == 0 ? 3. / 128. : q == 1 ? -22. / 128. : q == 2 ? 1.
                                                      : q == 3 ? 22. / 128. : -3. / 128.;
                        };
                        double acc = 0.;
#pragma unroll 1
                        for (int cz = 0; cz < 5; ++cz)
#pragma unroll 1
                            for (int by = 0; by < 5; ++by)
#pragma unroll
                                for (int ax = 0; ax < 5; ++ax) {
                                    const double wz = z ? wu(cz) : wu(4 - cz);
                                    const double wy = y ? wu(by) : wu(4 - by);
                                    const double wx = x ? wu(ax) : wu(4 - ax);
                                    acc += wz * wy * wx * Cb(ax - 2, by - 2, cz - 2);
                                }
                        return acc;
                    } else {
                        double c[NPT];
#pragma unroll
                        for (int q = 0; q < NPT; ++q) c[q] = Cb(PO[q][0], PO[q][1], PO[q][2]);
                        const double xo = x ? 0.25 : -0.25, yo = y ? 0.25 : -0.25,
                                     zo = z ? 0.25 : -0.25;
                        if constexpr (INTERP == 2) {
                            const double sx = 0.5 * (c[2] - c[1]), sy = 0.5 * (c[4] - c[3]),
                                         sz = 0.5 * (c[6] - c[5]);
                            const double sxx = 0.5 * (c[2] - 2. * c[0] + c[1]);
                            const double syy = 0.5 * (c[4] - 2. * c[0] + c[3]);
                            const double szz = 0.5 * (c[6] - 2. * c[0] + c[5]);
                            const double sxy = 0.25 * (c[7] - c[8] - c[9] + c[10]);
                            const double sxz = 0.25 * (c[11] - c[12] - c[13] + c[14]);
                            const double syz = 0.25 * (c[15] - c[16] - c[17] + c[18]);
                            return c[0] + xo * sx + yo * sy + zo * sz + xo * xo * sxx +
                                   yo * yo * syy + zo * zo * szz + xo * yo * sxy +
                                   xo * zo * sxz + yo * zo * syz;
                        } else if constexpr (INTERP == 1) {
                            auto mc = [](double cm, double c0v, double cp) {
                                const double sl = c0v - cm, sr = cp - c0v;
                                const double sc = 0.5 * (sl + sr);
                                if (!(sl * sr > 0.)) return 0.;
                                const double m =
                                    fmin(fabs(sc), fmin(2. * fabs(sl), 2. * fabs(sr)));
                                return sc > 0. ? m : -m;
                            };
                            return c[0] + xo * mc(c[1], c[0], c[2]) +
                                   yo * mc(c[3], c[0], c[4]) + zo * mc(c[5], c[0], c[6]);
                        } else {
                            return c[0];
                        }
                    }
                };
                // u0 at t0 (its leading word; with FP64 tiles of stored
                // pairs hi and lo), then k1 and k2 such that u0 + dt k1 and
                // u0 + dt/4 k2 are the states at t0 + dt and t0 + dt/2, as
                // the separate state and k fills write them.
                for (int z = c0[2]; z <= c1[2]; ++z)
                    for (int y = c0[1]; y <= c1[1]; ++y)
                        for (int x = c0[0]; x <= c1[0]; ++x) {
                            const long long e = (long long)comp * a.fcstride + fine_elem(x, y, z);
                            const double u0v = value(brick, x, y, z);
                            double u0t;  // u0's tile value as the stage kernels read it
                            if constexpr (T::comp && T::exact) {
                                store_state_at<T>(a.dst, a.dstlo, e, u0v);
                                u0t = pair_value<T>(u0v);
                            } else {
                                const S u0s = (S)u0v;
                                st<S>(a.dst, e, 0u, u0s);
                                u0t = (double)u0s;
                            }
                            st<S>(a.kd[0], e, 0u,
                                  (S)((value(brick + nb, x, y, z) - u0t) /
                                      (double)(V)a.cf[0]));
                            st<S>(a.kd[1], e, 0u,
                                  (S)((value(brick + 2 * nb, x, y, z) - u0t) /
                                      (double)(V)a.cf[1]));
                        }
            } else if constexpr (INTERP == 4) {
                // Tensor product of the conservative quartic child weights;
                // an upper child (odd fine index) takes WU, a lower one its
                // mirror.
                // WU = (3, -22, 128, 22, -3) / 128 as selects (no indexed
                // array, which would live in scratch memory).
                auto wu = [](int q) -> double {
                    return q == 0 ? 3. / 128. : q == 1 ? -22. / 128. : q == 2 ? 1.
                                              : q == 3 ? 22. / 128. : -3. / 128.;
                };
                double acc[2][2][2] = {{{0., 0.}, {0., 0.}}, {{0., 0.}, {0., 0.}}};
#pragma unroll 1
                for (int cz = 0; cz < 5; ++cz)
#pragma unroll 1
                    for (int by = 0; by < 5; ++by)
#pragma unroll
                        for (int ax = 0; ax < 5; ++ax) {
                            const double v = C(ax - 2, by - 2, cz - 2);
                            const double wz[2] = {wu(4 - cz), wu(cz)};
                            const double wy[2] = {wu(4 - by), wu(by)};
                            const double wx[2] = {wu(4 - ax), wu(ax)};
#pragma unroll
                            for (int z = 0; z < 2; ++z)
#pragma unroll
                                for (int y = 0; y < 2; ++y)
#pragma unroll
                                    for (int x = 0; x < 2; ++x)
                                        acc[z][y][x] += wz[z] * wy[y] * wx[x] * v;
                        }
                // Compile-time child indices keep acc in registers.
#pragma unroll
                for (int z = 0; z < 2; ++z)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
#pragma unroll
                        for (int x = 0; x < 2; ++x)
                            if (z >= c0[2] && z <= c1[2] && y >= c0[1] && y <= c1[1] &&
                                x >= c0[0] && x <= c1[0])
                                put(n, fine_elem(x, y, z), acc[z][y][x]);
            } else {
                double c[NPT];
#pragma unroll
                for (int q = 0; q < NPT; ++q) c[q] = C(PO[q][0], PO[q][1], PO[q][2]);
                // Slopes of the parent, shared by its children at x, y, z =
                // +-1/4 coarse cells.
                double sx = 0., sy = 0., sz = 0., sxx = 0., syy = 0., szz = 0., sxy = 0.,
                       sxz = 0., syz = 0.;
                if constexpr (INTERP == 2) {
                    sx = 0.5 * (c[2] - c[1]);
                    sy = 0.5 * (c[4] - c[3]);
                    sz = 0.5 * (c[6] - c[5]);
                    sxx = 0.5 * (c[2] - 2. * c[0] + c[1]);
                    syy = 0.5 * (c[4] - 2. * c[0] + c[3]);
                    szz = 0.5 * (c[6] - 2. * c[0] + c[5]);
                    sxy = 0.25 * (c[7] - c[8] - c[9] + c[10]);
                    sxz = 0.25 * (c[11] - c[12] - c[13] + c[14]);
                    syz = 0.25 * (c[15] - c[16] - c[17] + c[18]);
                } else if constexpr (INTERP == 1) {
                    // Monotonized-central slopes per direction.
                    auto mc = [](double cm, double c0v, double cp) {
                        const double sl = c0v - cm, sr = cp - c0v;
                        const double sc = 0.5 * (sl + sr);
                        if (!(sl * sr > 0.)) return 0.;
                        const double m = fmin(fabs(sc), fmin(2. * fabs(sl), 2. * fabs(sr)));
                        return sc > 0. ? m : -m;
                    };
                    sx = mc(c[1], c[0], c[2]);
                    sy = mc(c[3], c[0], c[4]);
                    sz = mc(c[5], c[0], c[6]);
                }
                for (int z = c0[2]; z <= c1[2]; ++z)
                    for (int y = c0[1]; y <= c1[1]; ++y)
                        for (int x = c0[0]; x <= c1[0]; ++x) {
                            const double xo = x ? 0.25 : -0.25, yo = y ? 0.25 : -0.25,
                                         zo = z ? 0.25 : -0.25;
                            double v;
                            if constexpr (INTERP == 2)
                                v = c[0] + xo * sx + yo * sy + zo * sz + xo * xo * sxx +
                                    yo * yo * syy + zo * zo * szz + xo * yo * sxy +
                                    xo * zo * sxz + yo * zo * syz;
                            else if constexpr (INTERP == 1)
                                v = c[0] + xo * sx + yo * sy + zo * sz;
                            else
                                v = c[0];
                            put(n, fine_elem(x, y, z), v);
                        }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Gravitational-wave spectrum: pack -> hipFFT D2Z -> projected binning.
// ---------------------------------------------------------------------------

/** Copy one component of one box into a dense N^3 FP64 array (x fastest). */
struct PackArgs {
    const void *arr;
    const void *arrlo;
    long long cstride;
    BoxDesc box;
    int comp;
    double *out;
    int N;
};

template <Prec P>
__global__ void __launch_bounds__(256) pack_kernel(const PackArgs a) {
    using S = typename PrecTraits<P>::S;
    const long long nb = (long long)a.box.n[0] * a.box.n[1] * a.box.n[2];
    const long long gstride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nb;
         t += gstride) {
        const int x = (int)(t % a.box.n[0]);
        const int y = (int)((t / a.box.n[0]) % a.box.n[1]);
        const int z = (int)(t / ((long long)a.box.n[0] * a.box.n[1]));
        const long long e = (long long)a.comp * a.cstride + a.box.base + x +
                            (long long)y * a.box.jstride +
                            (long long)z * a.box.kstride;
        const long long o = (long long)(a.box.lo[0] + x) +
                            (long long)a.N * ((a.box.lo[1] + y) +
                                              (long long)a.N * (a.box.lo[2] + z));
        a.out[o] = load_state_at<PrecTraits<P>>(a.arr, a.arrlo, e);
    }
}

struct BinArgs {
    const double2 *u[6];     // R2C outputs [z][y][x], x halved (N/2+1)
    int N;
    const double *index_to_k;
    int mode;                // 0 unbinned (lookup), 1 round(|n|), 2 log bins
    const int *sq_to_bin;    // mode 0 table, size 3*(N/2)^2 + 1
    int nbins;
    double log_kmax;         // mode 2: ln(sqrt(3) N / 2)
    double *bins;            // FP64 accumulators
    int modifier;            // project spectrum modifier id, -1: none
    const double *mparams;   // its parameters
    double dk;               // 2 pi / L of the (zero-padded) transform
};

/** sum_{ijlm} Lambda_{ij,lm}(k) A_ij A_lm = Tr(APAP) - Tr(PA)^2 / 2 for a
 *  symmetric A and the transverse projector P = 1 - k k^T / |k|^2; equal to
 *  the 81-term contraction of gw_GetLambda (gravitational_waves.h:101-133). */
__device__ AMREX_FORCE_INLINE double lambda_contract(const double P[3][3],
                                                     const double A[3][3]) {
    double B[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            B[i][j] = P[i][0] * A[0][j] + P[i][1] * A[1][j] + P[i][2] * A[2][j];
    double tr = B[0][0] + B[1][1] + B[2][2];
    double tr2 = 0.;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) tr2 += B[i][j] * B[j][i];
    return tr2 - 0.5 * tr * tr;
}

template <class Mods, int LDSBINS>
__global__ void __launch_bounds__(256) gw_bin_kernel(const BinArgs a) {
    __shared__ double sb[LDSBINS > 0 ? LDSBINS : 1];
    const bool use_lds = LDSBINS > 0 && a.nbins <= LDSBINS;
    if (use_lds)
        for (int i = threadIdx.x; i < a.nbins; i += blockDim.x) sb[i] = 0.;
    __syncthreads();
    const int N = a.N, NH = N / 2 + 1;
    const long long total = (long long)NH * N * N;
    const long long gstride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < total; t += gstride) {
        const int ax = (int)(t % NH);
        const int by = (int)((t / NH) % N);
        const int cz = (int)(t / ((long long)NH * N));
        if (ax == 0 && by == 0 && cz == 0) continue;  // Lambda(0) = 0
        const int li = ax >= N / 2 ? ax - N : ax;
        const int lj = by >= N / 2 ? by - N : by;
        const int lk = cz >= N / 2 ? cz - N : cz;
        const unsigned sq = (unsigned)(li * li + lj * lj + lk * lk);
        int bin;
        if (a.mode == 0) {
            bin = a.sq_to_bin[sq];
        } else if (a.mode == 1) {
            bin = (int)(sqrt((double)sq) + 0.5);
        } else {
            bin = (int)floor((double)a.nbins * 0.5 * log((double)sq) / a.log_kmax);
            bin = bin < 0 ? 0 : (bin >= a.nbins ? a.nbins - 1 : bin);
        }
        if (bin < 0 || bin >= a.nbins) continue;
        const double kv[3] = {a.index_to_k[ax], a.index_to_k[by],
                              a.index_to_k[cz]};
        const double nrm = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double Pm[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Pm[i][j] = (i == j ? 1. : 0.) - kv[i] * kv[j] / nrm;
        double2 v[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) v[c] = a.u[c][t];
        if constexpr (Mods::count > 0) {
            // The project's FourierSpaceModification of this mode.
            if (a.modifier >= 0) {
                double re[6], im[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    re[c] = v[c].x;
                    im[c] = v[c].y;
                }
                Mods::modify(a.modifier, re, im, li, lj, lk, a.dk, N, a.mparams);
#pragma unroll
                for (int c = 0; c < 6; ++c) v[c] = make_double2(re[c], im[c]);
            }
        }
        // Component order xx, xy, xz, yy, yz, zz.
        const double Ar[3][3] = {{v[0].x, v[1].x, v[2].x},
                                 {v[1].x, v[3].x, v[4].x},
                                 {v[2].x, v[4].x, v[5].x}};
        const double Ai[3][3] = {{v[0].y, v[1].y, v[2].y},
                                 {v[1].y, v[3].y, v[4].y},
                                 {v[2].y, v[4].y, v[5].y}};
        const double mult = (ax == 0 || ax == N / 2) ? 1. : 2.;
        const double val =
            mult * (lambda_contract(Pm, Ar) + lambda_contract(Pm, Ai));
        if (use_lds)
            atomicAdd(&sb[bin], val);
        else
            atomicAdd(&a.bins[bin], val);
    }
    if (use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < a.nbins; i += blockDim.x)
            if (sb[i] != 0.) atomicAdd(&a.bins[i], sb[i]);
    }
}

// ---------------------------------------------------------------------------
// Output functions: spectra (output_types/spectrum.cpp:12-113) and
// projections (output_types/projection.cpp:11-118).
// ---------------------------------------------------------------------------

/** max with a CAS loop (values of either sign). */
__device__ AMREX_FORCE_INLINE void atomic_max_double(double *p, double v) {
    unsigned long long *q = reinterpret_cast<unsigned long long *>(p);
    unsigned long long old = *q;
    while (__longlong_as_double((long long)old) < v) {
        const unsigned long long prev =
            atomicCAS(q, old, (unsigned long long)__double_as_longlong(v));
        if (prev == old) break;
        old = prev;
    }
}

/** One thread per valid cell of a level: the output function on a full-
 *  precision view of the state.  Projection: a cell contributes unless the
 *  finer level holds fine cell (2i, 2j, 2k); it adds ratio * value to (or
 *  maximises) the ratio x ratio pixels (i ratio + a) D + (j ratio + b). */
template <class Ops, Prec P>
__global__ void __launch_bounds__(256) output_kernel(const OutArgs a) {
    using Out = typename Ops::Output;
    const long long gstride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < a.total; t += gstride) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.vprefix[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BoxDesc &b = a.boxes[lo];
        long long r = t - a.vprefix[lo];
        const int x = (int)(r % b.n[0]);
        r /= b.n[0];
        const int y = (int)(r % b.n[1]);
        const int z = (int)(r / b.n[1]);
        const int i = b.lo[0] + x, j = b.lo[1] + y, k = b.lo[2] + z;
        if (a.target == 1) {
            bool covered = false;
            for (int q = a.cov_start[lo]; q < a.cov_start[lo + 1] && !covered; ++q) {
                const int *f = a.cov + 6 * q;
                covered = 2 * i >= f[0] && 2 * i <= f[3] && 2 * j >= f[1] &&
                          2 * j <= f[4] && 2 * k >= f[2] && 2 * k <= f[5];
            }
            if (covered) continue;
        }
        const amrex::Array4<const double> v =
            plain_view<P>(a.arr, a.arrlo, b, a.cstride, a.ncomp);
        const double val =
            Out::eval(a.fn, v, i, j, k, a.lev, a.time, a.dt, a.dx, a.params);
        if (a.target == 0) {
            a.dense[(long long)i + (long long)a.N * (j + (long long)a.N * k)] = val;
        } else {
            for (int ia = 0; ia < a.ratio; ++ia)
                for (int ja = 0; ja < a.ratio; ++ja) {
                    const long long ind =
                        (long long)(i * a.ratio + ia) * a.D + (j * a.ratio + ja);
                    if (a.mode == 0)
                        atomicAdd(&a.proj[ind], a.ratio * val);
                    else
                        atomic_max_double(&a.proj[ind], val);
                    atomicAdd(&a.nproj[ind], 1);
                }
        }
    }
}

struct SpecBinArgs {
    const double2 *F;      // R2C output [z][y][x], x halved (N/2 + 1)
    int N;
    const int *sq_to_bin;  // unique |k|^2 index, size 3 (N/2)^2 + 1
    int nbins;
    double *bins;
};

/** sum over modes of mult |F|^2 per unique |k|^2, mult = 1 on the x = 0 and
 *  x = N/2 planes of the halved layout, else 2 (spectrum.cpp:73-90). */
template <int LDSBINS>
__global__ void __launch_bounds__(256) scalar_bin_kernel(const SpecBinArgs a) {
    __shared__ double sb[LDSBINS];
    const bool use_lds = a.nbins <= LDSBINS;
    if (use_lds)
        for (int i = threadIdx.x; i < a.nbins; i += blockDim.x) sb[i] = 0.;
    __syncthreads();
    const int N = a.N, NH = N / 2 + 1;
    const long long total = (long long)NH * N * N;
    const long long gstride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < total; t += gstride) {
        const int ax = (int)(t % NH);
        const int by = (int)((t / NH) % N);
        const int cz = (int)(t / ((long long)NH * N));
        const int li = ax >= N / 2 ? ax - N : ax;
        const int lj = by >= N / 2 ? by - N : by;
        const int lk = cz >= N / 2 ? cz - N : cz;
        const int bin = a.sq_to_bin[li * li + lj * lj + lk * lk];
        if (bin < 0 || bin >= a.nbins) continue;
        const double2 f = a.F[t];
        const double val = ((ax == 0 || ax == N / 2) ? 1. : 2.) * (f.x * f.x + f.y * f.y);
        if (use_lds)
            atomicAdd(&sb[bin], val);
        else
            atomicAdd(&a.bins[bin], val);
    }
    if (use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < a.nbins; i += blockDim.x)
            if (sb[i] != 0.) atomicAdd(&a.bins[i], sb[i]);
    }
}

// ---------------------------------------------------------------------------
// Seeded initial states (identical formulas in the oracle).
// ---------------------------------------------------------------------------

struct InitArgs {
    void *arr;
    void *arrlo;
    long long cstride;
    const BoxDesc *boxes;
    const long long *vprefix;  // valid cells per box, prefix
    int nbox;
    long long total;
    int N;                     // domain cells per axis at this level
    int kind;
    uint64_t seed;
    int nfields;
    int ncomp;
    double p[8];
    double cell;               // level-0 cells per cell of this level
    double N0;                 // level-0 cells per axis (periodic length)
    const double *obj;         // object table (bubbles / loops), 8 per object
    int nobj;
};

enum InitKind : int {
    kInitComplexRandomPhase = 1,  // psi = (1 + p0 N(0,1)) exp(i 2 pi U), pi = 0
    kInitZero = 2,
    kInitBubbles = 3,        // vacuum plus the bubbles nucleated by p[7]
    kInitAddBubbles = 4,     // inject the bubbles nucleated in (p[6], p[7]]
    kInitStringLoops = 5,    // psi = prod tanh(d/w) exp(i theta), unit winding
};

/** Minimum periodic image of a displacement along one axis. */
__host__ __device__ AMREX_FORCE_INLINE double min_image(double d, double L) {
    return d - L * floor(d / L + 0.5);
}

/** Bubble wall: phi_in (1 - tanh((r - R0) / w)) / 2 at distance r (all in
 *  level-0 cells).  Object row: centre x, y, z, nucleation time. */
__host__ __device__ AMREX_FORCE_INLINE double bubble_profile(
    const double *o, double x, double y, double z, double L, double phi_in,
    double R0, double w) {
    const double dx = min_image(x - o[0], L), dy = min_image(y - o[1], L),
                 dz = min_image(z - o[2], L);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    return phi_in * 0.5 * (1. - tanh((r - R0) / w));
}

/** One closed string loop: object row centre x, y, z, radius, unit normal.
 *  Returns the modulus factor tanh(d / w) (d = distance to the string) and
 *  the winding angle atan2(h, rho - R) of the point around the string. */
__host__ __device__ AMREX_FORCE_INLINE void loop_factor(
    const double *o, double x, double y, double z, double L, double w,
    double &mod, double &theta) {
    const double dx = min_image(x - o[0], L), dy = min_image(y - o[1], L),
                 dz = min_image(z - o[2], L);
    const double h = dx * o[4] + dy * o[5] + dz * o[6];
    const double px = dx - h * o[4], py = dy - h * o[5], pz = dz - h * o[6];
    const double s = sqrt(px * px + py * py + pz * pz) - o[3];
    mod = tanh(sqrt(s * s + h * h) / w);
    theta = atan2(h, s);
}

template <Prec P>
__global__ void __launch_bounds__(256) init_kernel(const InitArgs a) {
    using S = typename PrecTraits<P>::S;
    const long long gstride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
         t < a.total; t += gstride) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.vprefix[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BoxDesc b = a.boxes[lo];
        long long r = t - a.vprefix[lo];
        const int x = (int)(r % b.n[0]);
        r /= b.n[0];
        const int y = (int)(r % b.n[1]);
        const int z = (int)(r / b.n[1]);
        const long long e =
            b.base + x + (long long)y * b.jstride + (long long)z * b.kstride;
        const uint64_t gidx = (uint64_t)(b.lo[0] + x) +
                              (uint64_t)a.N * ((uint64_t)(b.lo[1] + y) +
                                               (uint64_t)a.N * (uint64_t)(b.lo[2] + z));
        // Cell centre in level-0 cell units.
        const double px = (b.lo[0] + x + 0.5) * a.cell, py = (b.lo[1] + y + 0.5) * a.cell,
                     pz = (b.lo[2] + z + 0.5) * a.cell;
        if (a.kind == kInitAddBubbles) {
            // Injection: the first field only, read-modify-write.
            double phi = load_state_at<PrecTraits<P>>(a.arr, a.arrlo, e);
            for (int o = 0; o < a.nobj; ++o) {
                const double *ob = a.obj + 8 * o;
                if (ob[3] > a.p[6] && ob[3] <= a.p[7])
                    phi = fmax(phi, bubble_profile(ob, px, py, pz, a.N0, a.p[1],
                                                   a.p[2], a.p[3]));
            }
            store_state_at<PrecTraits<P>>(a.arr, a.arrlo, e, phi);
            continue;
        }
        double v[32];
        for (int n = 0; n < a.ncomp && n < 32; ++n) v[n] = 0.;
        if (a.kind == kInitBubbles) {
            double phi = 0.;
            for (int o = 0; o < a.nobj; ++o) {
                const double *ob = a.obj + 8 * o;
                if (ob[3] <= a.p[7])
                    phi = fmax(phi, bubble_profile(ob, px, py, pz, a.N0, a.p[1],
                                                   a.p[2], a.p[3]));
            }
            v[0] = phi;
        } else if (a.kind == kInitStringLoops) {
            double mod = 1., th = 0.;
            for (int o = 0; o < a.nobj; ++o) {
                double m, t;
                loop_factor(a.obj + 8 * o, px, py, pz, a.N0, a.p[3], m, t);
                mod *= m;
                th += t;
            }
            v[0] = mod * cos(th);
            v[1] = mod * sin(th);
        }
        if (a.kind == kInitComplexRandomPhase) {
            const double u1 = rng_uniform(a.seed, 1, gidx);
            const double u2 = rng_uniform(a.seed, 2, gidx);
            const double u3 = rng_uniform(a.seed, 3, gidx);
            const double theta = 2. * M_PI * u1;
            const double g = sqrt(-2. * log(1. - u2)) * cos(2. * M_PI * u3);
            const double m = 1. + a.p[0] * g;
            v[0] = m * cos(theta);
            v[1] = m * sin(theta);
        }
        for (int n = 0; n < a.ncomp && n < 32; ++n)
            store_state_at<PrecTraits<P>>(a.arr, a.arrlo, (long long)n * a.cstride + e, v[n]);
    }
}

// ---------------------------------------------------------------------------
// Local regrid: blocks of the finer level a tagged cell asks for.
// ---------------------------------------------------------------------------

/** LocalRegrid::CheckBorders (local_regrid.cpp:563-620) for every tagged
 *  cell.  A cell at fine index f = 2 ci, in 1-based block cf = f / bf + 1,
 *  measures for each of the 27 blocks cf + d marked in its box's border map
 *  d^2 = sum over axes of 0 inside (lo < f < hi), else min((f - lo)^2,
 *  (f - hi)^2), lo = bf (cf + d - 1), hi = bf (cf + d) - 1.
 *  Pass 0: below thr the block (wrapped) is added, and the box-local index
 *  of the first such cell in k, j, i order is kept per border entry.
 *  Pass 1: the reference scans a box's cells in that order and unmarks a
 *  block once added, so a cell measures a block iff its index is at most
 *  that first index; the smallest measured d^2 is kept. */
template <int PASS>
__global__ void __launch_bounds__(256) regrid_select_kernel(const RegridArgs a) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < a.nvalid;
         idx += stride) {
        if (!a.mask[idx]) continue;
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.boxes[mid].voff <= idx) lo = mid;
            else hi = mid - 1;
        }
        const RegridBox &B = a.boxes[lo];
        const int local = (int)(idx - B.voff);
        const int r = local / B.n[0];
        const int ci[3] = {B.lo[0] + local % B.n[0], B.lo[1] + r % B.n[1],
                           B.lo[2] + r / B.n[1]};
        int fi[3], cf[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            fi[d] = 2 * ci[d];
            cf[d] = fi[d] / a.bf + 1;
        }
        const long long sx = B.xs[0] + 2, sy = B.xs[1] + 2;
        int best = INT_MAX;
        for (int kk = -1; kk <= 1; ++kk)
            for (int jj = -1; jj <= 1; ++jj)
                for (int ii = -1; ii <= 1; ++ii) {
                    const int off[3] = {ii, jj, kk};
                    int ref[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) ref[d] = cf[d] - B.c0[d] + off[d] + 1;
                    const long long e = B.boff + ref[0] + sx * (ref[1] + sy * ref[2]);
                    if (!a.border[e]) continue;
                    int d_sq = 0;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const int blo = a.bf * (cf[d] + off[d] - 1);
                        const int bhi = a.bf * (cf[d] + off[d]) - 1;
                        const int dl = fi[d] - blo, dh = fi[d] - bhi;
                        d_sq += (blo < fi[d] && fi[d] < bhi) ? 0 : min(dl * dl, dh * dh);
                    }
                    if (PASS == 0) {
                        if (d_sq < a.thr) {
                            atomicMin(&a.first[e], local);
                            int b[3];
#pragma unroll
                            for (int d = 0; d < 3; ++d)
                                b[d] = (cf[d] + off[d] - 1 + a.Np) % a.Np;
                            a.added[b[0] + (long long)a.Np * (b[1] + (long long)a.Np * b[2])] = 1;
                        }
                    } else if (local <= a.first[e]) {
                        best = min(best, d_sq);
                    }
                }
        if (PASS == 1 && best != INT_MAX) atomicMin(a.min_dsq, best);
    }
}

// ---------------------------------------------------------------------------
// Uploads and downloads: valid cells <-> one dense FP64 array in one pass.
// ---------------------------------------------------------------------------

/** TO_STATE: dense -> state (leading word and remainder as the stage kernels
 *  store them); else state -> dense (leading word plus remainder). */
template <Prec P, bool TO_STATE>
__global__ void __launch_bounds__(256) valid_io_kernel(const ValidIoArgs a) {
    using T = PrecTraits<P>;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < a.total;
         t += stride) {
        int lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.vprefix[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BoxDesc &b = a.boxes[lo];
        const long long r = t - a.vprefix[lo];
        const long long row = r / b.n[0];
        const long long e = a.comp_off + b.base + (r - row * b.n[0]) +
                            (row % b.n[1]) * b.jstride + (row / b.n[1]) * b.kstride;
        if constexpr (TO_STATE)
            store_state_at<T>(a.hi, a.lo, e, a.dense[t]);
        else
            a.dense[t] = load_state_at<T>(a.hi, a.lo, e);
    }
}

/** Maximum |value| per component over the valid cells of a level (the
 *  reference's MultiFab::norm0 as projects call it on grid_new).  Non-negative
 *  doubles order like their bit patterns, so each wave folds its maximum
 *  and one lane takes an atomic maximum of the bits per component. */
struct Norm0Args {
    const BoxDesc *boxes;
    const long long *vprefix;
    int nbox;
    long long total;
    long long cstride;
    int ncomp;
    const void *hi, *lo;
    unsigned long long *out;  // ncomp bit patterns, zero before the launch
};

template <Prec P>
__global__ void __launch_bounds__(256) norm0_kernel(const Norm0Args a) {
    using T = PrecTraits<P>;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (int c = 0; c < a.ncomp; ++c) {
        double m = 0.;
        for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < a.total;
             t += stride) {
            int lo = 0, hi = a.nbox - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (a.vprefix[mid] <= t)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const BoxDesc &b = a.boxes[lo];
            const long long r = t - a.vprefix[lo];
            const long long row = r / b.n[0];
            const long long e = (long long)c * a.cstride + b.base + (r - row * b.n[0]) +
                                (row % b.n[1]) * b.jstride + (row / b.n[1]) * b.kstride;
            m = fmax(m, fabs(load_state_at<T>(a.hi, a.lo, e)));
        }
        unsigned long long bits = (unsigned long long)__double_as_longlong(m);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long v = __shfl_xor(bits, o, 64);
            bits = v > bits ? v : bits;
        }
        if ((threadIdx.x & 63) == 0) atomicMax(&a.out[c], bits);
    }
}

}  // namespace shm

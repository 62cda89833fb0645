// Minimal AMReX vocabulary for sledgehamr project headers compiled by hipcc
// for gfx950.  Only what project cell functions touch is provided:
// attribute macros, Real, Dim3, Box, Array4, lbound/ubound, Print and a
// ParmParse reader for the inputs file.
//
// Array4<const double> and Array4<double> carry a *mode*.  In plain mode they
// are the usual strided views (x fastest, then y, z, component).  Inside the
// fused stage kernels they become views onto the LDS-staged tile (centre plane
// in LDS, z-neighbours in registers) or onto a register sink for the Rhs
// output.  Every accessor is force-inlined, so after inlining the mode tests
// and offsets fold to constants and a project's Rhs() compiles to LDS/register
// accesses without a line of the project changing.
//
// Reference interface: amrex::Array4 as used by source/macros.h:301-485 and
// projects/*/kernels_rhs.h (state(i,j,k,comp) reads, rhs(i,j,k,comp) writes).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#define AMREX_GPU_DEVICE __device__
#define AMREX_GPU_HOST __host__
#define AMREX_GPU_HOST_DEVICE __host__ __device__
#define AMREX_FORCE_INLINE __attribute__((always_inline)) inline
#define AMREX_INLINE inline
#define AMREX_RESTRICT __restrict__
#define AMREX_PRAGMA_SIMD
#define AMREX_SPACEDIM 3

namespace shm {
struct TileCtx;
struct ProbeCtx;
}  // namespace shm

namespace amrex {

using Real = double;
using Long = long long;

struct Dim3 {
    int x, y, z;
};

struct IntVect {
    int v[3] = {0, 0, 0};
    AMREX_GPU_HOST_DEVICE IntVect() {}
    AMREX_GPU_HOST_DEVICE explicit IntVect(int a) : v{a, a, a} {}
    AMREX_GPU_HOST_DEVICE IntVect(int a, int b, int c) : v{a, b, c} {}
    AMREX_GPU_HOST_DEVICE int operator[](int i) const { return v[i]; }
    AMREX_GPU_HOST_DEVICE int &operator[](int i) { return v[i]; }
};

/** Cell-centred index box, inclusive bounds (AMReX convention). */
struct Box {
    IntVect lo, hi;
    AMREX_GPU_HOST_DEVICE Box() {}
    AMREX_GPU_HOST_DEVICE Box(IntVect l, IntVect h) : lo(l), hi(h) {}
    AMREX_GPU_HOST_DEVICE IntVect smallEnd() const { return lo; }
    AMREX_GPU_HOST_DEVICE IntVect bigEnd() const { return hi; }
    AMREX_GPU_HOST_DEVICE int length(int d) const { return hi[d] - lo[d] + 1; }
    AMREX_GPU_HOST_DEVICE Long numPts() const {
        return (Long)length(0) * length(1) * length(2);
    }
};

AMREX_GPU_HOST_DEVICE inline Dim3 lbound(const Box &b) {
    return {b.lo[0], b.lo[1], b.lo[2]};
}
AMREX_GPU_HOST_DEVICE inline Dim3 ubound(const Box &b) {
    return {b.hi[0], b.hi[1], b.hi[2]};
}

/** Generic plain strided view (used for char tag arrays and host code). */
template <class T>
struct Array4 {
    T *AMREX_RESTRICT p = nullptr;
    Long jstride = 0, kstride = 0, nstride = 0;
    Dim3 begin{0, 0, 0}, end{0, 0, 0};  // end is exclusive
    int ncomp = 1;

    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE T &operator()(int i, int j, int k,
                                                           int n = 0) const {
        return p[(i - begin.x) + (j - begin.y) * jstride +
                 (k - begin.z) * kstride + n * nstride];
    }
    AMREX_GPU_HOST_DEVICE int nComp() const { return ncomp; }
    AMREX_GPU_HOST_DEVICE bool contains(int i, int j, int k) const {
        return i >= begin.x && i < end.x && j >= begin.y && j < end.y &&
               k >= begin.z && k < end.z;
    }
};

/** Read-only state view handed to Rhs(), TagCellForRefinement() and
 *  TruncationModifier<n>().  Storage may be FP64 or FP32 (values are widened
 *  to double on read), or the LDS/register tile of a fused stage kernel. */
template <>
struct Array4<const double> {
    // kPlain32C / kPlain32C16: FP32 value plus an FP32 / bf16 remainder,
    // read back as their FP64 sum.
    enum Mode : int {
        kPlain64 = 0,
        kPlain32 = 1,
        kTile = 2,
        kProbe = 3,
        kPlain32C = 4,
        kPlain32C16 = 5
    };

    const double *p = nullptr;
    const float *pf = nullptr;
    const void *plo = nullptr;  // remainder plane (kPlain32C, kPlain32C16)
    Long jstride = 0, kstride = 0, nstride = 0;
    Dim3 begin{0, 0, 0}, end{0, 0, 0};
    int ncomp = 1;
    int mode = kPlain64;
    const void *ext = nullptr;  // shm::TileCtx or shm::ProbeCtx

    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double operator()(int i, int j,
                                                               int k,
                                                               int n = 0) const;
    AMREX_GPU_HOST_DEVICE int nComp() const { return ncomp; }
    AMREX_GPU_HOST_DEVICE bool contains(int i, int j, int k) const {
        return i >= begin.x && i < end.x && j >= begin.y && j < end.y &&
               k >= begin.z && k < end.z;
    }
};

/** Writable view handed to Rhs() / GravitationalWavesRhs<true>() as `rhs`.
 *  In sink mode every component of the current cell lives in a register
 *  array owned by the stage kernel; (i,j,k) must be the cell being updated. */
template <>
struct Array4<double> {
    enum Mode : int { kPlain = 0, kSink = 1 };

    double *p = nullptr;
    Long jstride = 0, kstride = 0, nstride = 0;
    Dim3 begin{0, 0, 0}, end{0, 0, 0};
    int ncomp = 1;
    int mode = kPlain;
    double *sink = nullptr;

    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double &operator()(int i, int j,
                                                                int k,
                                                                int n = 0) const {
        if (mode == kSink) return sink[n];
        return p[(i - begin.x) + (j - begin.y) * jstride +
                 (k - begin.z) * kstride + n * nstride];
    }
    AMREX_GPU_HOST_DEVICE int nComp() const { return ncomp; }
};

/** amrex::Print() replacement: writes to stdout on the host. */
class Print {
  public:
    Print() = default;
    ~Print() { std::cout << ss_.str() << std::flush; }
    template <class T>
    Print &operator<<(const T &v) {
        ss_ << v;
        return *this;
    }
    Print &operator<<(std::ostream &(*f)(std::ostream &)) {
        ss_ << f;
        return *this;
    }

  private:
    std::ostringstream ss_;
};

template <class T>
using Vector = std::vector<T>;

/** Inputs-file reader with ParmParse's query/get semantics.  The table is a
 *  process-wide key -> tokens map filled by ParmParse::ParseText(). */
class ParmParse {
  public:
    explicit ParmParse(const std::string &prefix = "") : prefix_(prefix) {}

    static void ParseText(const std::string &text);
    static void Clear();
    static bool Contains(const std::string &full_key);

    template <class T>
    bool query(const char *name, T &val) const {
        const std::vector<std::string> *tok = find(name);
        if (!tok || tok->empty()) return false;
        val = convert<T>((*tok)[0]);
        return true;
    }
    template <class T>
    void get(const char *name, T &val) const {
        if (!query(name, val))
            throw std::runtime_error("ParmParse: missing required key '" +
                                     full(name) + "'");
    }
    template <class T>
    bool queryarr(const char *name, std::vector<T> &val) const {
        const std::vector<std::string> *tok = find(name);
        if (!tok) return false;
        val.clear();
        for (const auto &s : *tok) val.push_back(convert<T>(s));
        return true;
    }
    template <class T>
    void getarr(const char *name, std::vector<T> &val) const {
        if (!queryarr(name, val))
            throw std::runtime_error("ParmParse: missing required key '" +
                                     full(name) + "'");
    }
    bool contains(const char *name) const { return find(name) != nullptr; }

  private:
    std::string full(const char *name) const {
        return prefix_.empty() ? std::string(name) : prefix_ + "." + name;
    }
    const std::vector<std::string> *find(const char *name) const;
    template <class T>
    static T convert(const std::string &s);

    std::string prefix_;
};

template <>
inline int ParmParse::convert<int>(const std::string &s) {
    return std::stoi(s);
}
template <>
inline long ParmParse::convert<long>(const std::string &s) {
    return std::stol(s);
}
template <>
inline double ParmParse::convert<double>(const std::string &s) {
    return std::stod(s);
}
template <>
inline std::string ParmParse::convert<std::string>(const std::string &s) {
    return s;
}
template <>
inline bool ParmParse::convert<bool>(const std::string &s) {
    if (s == "true" || s == "True" || s == "TRUE" || s == "1") return true;
    if (s == "false" || s == "False" || s == "FALSE" || s == "0") return false;
    throw std::runtime_error("ParmParse: cannot parse bool '" + s + "'");
}

inline void Abort(const std::string &msg) { throw std::runtime_error(msg); }
inline void Error(const std::string &msg) { throw std::runtime_error(msg); }

}  // namespace amrex

namespace shm {

/** Tile view context: the fused stage kernel fills one per cell it updates.
 *  `lds` points at the cell's own position (stencil slot 0) in the staged
 *  centre plane, `zq` at the cell's register queue of z-neighbours
 *  (zq_cstride entries per stencil slot; plane ck + dk at entry
 *  dk + R + zq_phase, taken modulo zq_len, so a queue that rotates instead
 *  of shifting passes its phase), `ctr` at the widened centre values of
 *  every component.  Components whose slot is negative are point-wise: only
 *  the centre may be read. */
struct TileCtx {
    int ci, cj, ck;
    int R;
    int f32;
    int lds_pitch;
    int lds_cstride;
    int zq_cstride;
    int zq_phase = 0;
    int zq_len = 0x7fffffff;
    const void *lds;
    const void *zq;
    const double *ctr;
    const signed char *slot;
    // LDS-window mode (zq unused): element offset, relative to the centre
    // plane, of plane ck + dk for dk = -R..R.
    const int *zplane = nullptr;

    /** Queue entry of plane ck + dk, |dk| <= R, in stencil slot s. */
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE int zq_at(int s, int dk) const {
        int q = dk + R + zq_phase;
        if (q >= zq_len) q -= zq_len;
        return s * zq_cstride + q;
    }

    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double get(int i, int j, int k,
                                                        int n) const {
        const int di = i - ci, dj = j - cj, dk = k - ck;
        if (di == 0 && dj == 0 && dk == 0) return ctr[n];
        const int s = slot[n];
        if (s < 0) return __builtin_nan("");
        if (dk != 0 && zplane) {
            const int l = s * lds_cstride + zplane[dk + R] + dj * lds_pitch + di;
            return f32 ? (double)static_cast<const float *>(lds)[l]
                       : static_cast<const double *>(lds)[l];
        }
        if (dk != 0) {
            if (di != 0 || dj != 0) return __builtin_nan("");
            const int q = zq_at(s, dk);
            return f32 ? (double)static_cast<const float *>(zq)[q]
                       : static_cast<const double *>(zq)[q];
        }
        const int l = s * lds_cstride + dj * lds_pitch + di;
        return f32 ? (double)static_cast<const float *>(lds)[l]
                   : static_cast<const double *>(lds)[l];
    }
};

/** Probe context: records the reach of every component read by a cell
 *  function (run once per context to size halos and validate the tile
 *  assumptions).  Reads return a smooth positive value; with `hashed` set
 *  they return values that vary with component and offset (in [0.75,
 *  1.25)), plus pert[n] times a second such value, so that perturbing one
 *  component shows which outputs read it. */
struct ProbeCtx {
    static constexpr int kMaxComp = 64;
    int ci, cj, ck;
    int ncomp;
    mutable int reach[kMaxComp];     // max |offset| along any axis
    mutable int offaxis[kMaxComp];   // read with more than one nonzero offset
    mutable int zreach[kMaxComp];    // max |dk|
    int hashed = 0;
    unsigned seed = 0;
    double pert[kMaxComp];

    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE static double unit(unsigned h) {
        h ^= h >> 16;
        h *= 0x7feb352du;
        h ^= h >> 15;
        h *= 0x846ca68bu;
        h ^= h >> 16;
        return (double)(h >> 8) * (1.0 / 16777216.0);
    }

    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double get(int i, int j, int k,
                                                        int n) const {
        const int di = i - ci, dj = j - cj, dk = k - ck;
        if (n >= 0 && n < kMaxComp) {
            int ai = di < 0 ? -di : di, aj = dj < 0 ? -dj : dj,
                ak = dk < 0 ? -dk : dk;
            int m = ai > aj ? ai : aj;
            m = m > ak ? m : ak;
            if (m > reach[n]) reach[n] = m;
            if (ak > zreach[n]) zreach[n] = ak;
            int nz = (di != 0) + (dj != 0) + (dk != 0);
            if (nz > 1 && (dk != 0)) offaxis[n] = 1;
        }
        if (hashed && n >= 0 && n < kMaxComp) {
            const unsigned key = seed ^ ((unsigned)n * 0x9e3779b1u) ^
                                 ((unsigned)(di + 16) * 0x85ebca77u) ^
                                 ((unsigned)(dj + 16) * 0xc2b2ae3du) ^
                                 ((unsigned)(dk + 16) * 0x27d4eb2fu);
            return 0.75 + 0.5 * unit(key) + pert[n] * (0.5 + unit(key ^ 0x5bd1e995u));
        }
        return 1.0 + 0.001 * (double)(n + 1);
    }
};

}  // namespace shm

namespace amrex {
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE double
Array4<const double>::operator()(int i, int j, int k, int n) const {
    if (mode == kTile)
        return static_cast<const shm::TileCtx *>(ext)->get(i, j, k, n);
    if (mode == kProbe)
        return static_cast<const shm::ProbeCtx *>(ext)->get(i, j, k, n);
    const Long off = (i - begin.x) + (j - begin.y) * jstride +
                     (k - begin.z) * kstride + n * nstride;
    if (mode == kPlain32C)
        return (double)pf[off] + (double)static_cast<const float *>(plo)[off];
    if (mode == kPlain32C16) {
        const unsigned int bits =
            (unsigned int)static_cast<const unsigned short *>(plo)[off] << 16;
        return (double)pf[off] + (double)__builtin_bit_cast(float, bits);
    }
    return mode == kPlain32 ? (double)pf[off] : p[off];
}
}  // namespace amrex

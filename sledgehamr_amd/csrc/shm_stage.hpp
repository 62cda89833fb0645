// Fused stage kernels for gfx950.
//
// One kernel template evaluates a project's Rhs() and applies an integrator
// stage update in registers, so each stored array is read once and written at
// most once per stage.  Work is a persistent loop over a device tile table
// that spans every box of a level: a tile is 64 cells in x (one wavefront per
// row, fully coalesced; 32 and two rows per wavefront for narrow boxes) by
// TY rows in y, marched along z.
//
// The z-march keeps, per thread and stencil component, a register queue of
// the 2R+1 values of its column; the centre plane with its x/y ghost ring is
// staged in LDS (double buffered, one barrier per plane).  The project's
// Rhs() reads the tile through amrex::Array4<const double> in tile mode and
// writes into a register sink; point-wise components (momenta) are read at
// the centre only.  Plane z+R+1, its halo ring, the point-wise values and the
// store-time inputs of plane z+1 are prefetched one plane ahead.
//
// Reference behaviour reproduced: the Rhs cell loop with the optional GW
// source/backreaction and Kreiss-Oliger term (source/macros.h:301-485), and
// the stage combinations of IntegratorLsssprk3 (lsssprk3.cpp:14-31),
// IntegratorRkn (rkn.cpp:25-186) and IntegratorLeapfrog (leapfrog.cpp:13-57).
#pragma once

#include <sledgehamr.h>

#include <cstdint>
#include <type_traits>

#include "shm_layout.hpp"

namespace shm {

/** Storage precision of a level's state.
 *   F64    : FP64 state and stage arrays.
 *   F32C   : FP32 state + FP32 remainder (u = hi + lo, ~48 significant bits).
 *   F32C16 : FP32 state + bf16 remainder (~32 significant bits).
 *   F32CS  : stored as F32C16; stage states and stencil tiles are formed in
 *            FP64 from hi + lo (compensated stencils).
 *  Stage arrays (k, RKN forces) are FP32 in every FP32 mode.
 *
 *  S is the storage type of the state and the stage arrays, L that of the
 *  remainder, V the type of the stage states the stencils read (tile
 *  values).  `exact`: a tile value carries the whole stored value (hi + lo),
 *  so the stage updates need no separate remainder at the centre. */
enum class Prec : int { F64 = 0, F32C = 1, F32C16 = 2, F32CS = 3 };
constexpr int kNPrec = 4;

template <Prec P>
struct PrecTraits;
template <>
struct PrecTraits<Prec::F64> {
    using S = double;
    using L = double;  // unused
    using V = double;
    static constexpr bool comp = false;
    static constexpr bool exact = true;
};
template <>
struct PrecTraits<Prec::F32C> {
    using S = float;
    using L = float;
    using V = float;
    static constexpr bool comp = true;
    static constexpr bool exact = false;
};
template <>
struct PrecTraits<Prec::F32C16> {
    using S = float;
    using L = uint16_t;  // bf16 bit pattern
    using V = float;
    static constexpr bool comp = true;
    static constexpr bool exact = false;
};
template <>
struct PrecTraits<Prec::F32CS> {
    using S = float;
    using L = uint16_t;  // bf16 bit pattern
    using V = double;
    static constexpr bool comp = true;
    static constexpr bool exact = true;
};
/** Storage twin of a precision: the auxiliary kernels (tagging, outputs,
 *  transfers, ghost copies) see only the stored arrays. */
constexpr Prec storage_prec(Prec p) { return p == Prec::F32CS ? Prec::F32C16 : p; }

__device__ AMREX_FORCE_INLINE uint16_t f32_to_bf16_rne(float f) {
    const uint32_t b = __float_as_uint(f);
    return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ AMREX_FORCE_INLINE float bf16_to_f32(uint16_t h) {
    return __uint_as_float((uint32_t)h << 16);
}

/** Per-stage parameter record in device memory (written by the host before
 *  the stage runs, so a captured graph replays with fresh values). */
constexpr int kPrmTime = 0, kPrmDt = 1, kPrmDx = 2;
constexpr int kPrmUser = 8;    // SetParamsRhs() values
constexpr int kMaxParams = 64;
constexpr int kPrmGw = kPrmUser + kMaxParams;  // SetParamsGravitationalWaveRhs()
constexpr int kPrmBlock = kPrmGw + kMaxParams;

/** Register copy of a stage record's user (and, with GW, GW) parameters,
 *  taken at kernel entry.  The projects index them with constants, so the
 *  copies of unused entries are removed and the rest stay in registers. */
template <bool GW>
struct ParamRegs {
    double user[kMaxParams];
    double gw[GW ? kMaxParams : 1];
    __device__ AMREX_FORCE_INLINE explicit ParamRegs(const double *prm) {
#pragma unroll
        for (int i = 0; i < kMaxParams; ++i) user[i] = prm[kPrmUser + i];
        if constexpr (GW) {
#pragma unroll
            for (int i = 0; i < kMaxParams; ++i) gw[i] = prm[kPrmGw + i];
        }
    }
    __device__ AMREX_FORCE_INLINE const double *gw_ptr(const double *prm) const {
        if constexpr (GW)
            return gw;
        else
            return prm + kPrmGw;
    }
};

/** Kernel argument block shared by all stage policies. */
struct StageArgs {
    const BoxDesc *boxes;
    const TileDesc *tiles;
    int ntiles;
    int lev;
    long long cstride;
    const double *prm;          // stage record: time, dt, dx, params, gw params
    const double *ko_strength;
    int ko_order;
    int flags;
    int nprev;
    int xcd_split;
    int ghost;          // ghost width g of the level's bricks
    const void *x;      // stage-state source
    const void *xlo;    // its remainder (read by FP64 tiles of stored pairs)
    const void *u0;     // old state (hi part for FP32C)
    const void *u0lo;   // old state remainder (FP32C)
    const void *kin;    // SSPRK3 k in / FillAddRhs previous rhs
    void *kout;         // SSPRK3 k out / FillRhs out
    void *y;            // SSPRK3 stage state out
    void *y2;           // RKN third-stage field states (PolRkn1YY, PolRknF2Y)
    void *unew;         // new state (hi)
    void *unewlo;       // new state remainder
    const void *f[5];   // RKN forces of earlier stages
    void *fout;         // RKN force of this stage
    void *vh;           // leapfrog half-step momenta (FP64)
    double c[20];       // stage coefficients
};

// Component bookkeeping is evaluated at compile time: forced inlining lets
// the unrolled loops fold every index, so the tile queues and the per-cell
// arrays stay in registers (an out-of-line call would leave them in scratch
// memory).
#define SHM_CT __host__ __device__ __attribute__((always_inline))

/** Which components need neighbour access (stencil) in the fused kernel.
 *  Default: the project's fields plus the six u_ij when GW are on; with
 *  Kreiss-Oliger dissipation every component is differenced.
 *
 *  G >= 0 selects one group of a split stage (GW levels without
 *  dissipation, where the fourteen components of one pass need more
 *  registers than a wavefront can hold at useful occupancy): each group is
 *  a launch that computes and stores only its own outputs and stages only
 *  the components they read.
 *    G = 0: the project's scalars (Rhs and the GW backreaction);
 *    G = 1: the pairs (u_xx, du_xx), (u_yy, du_yy), (u_zz, du_zz);
 *    G = 2: the pairs (u_xy, du_xy), (u_xz, du_xz), (u_yz, du_yz);
 *  groups 1 and 2 also stage the scalars (the GW source reads them).  A
 *  component a group does not stage reads as NaN; the runtime splits only
 *  when a dependency probe shows that every output reads its own group's
 *  components alone (Runtime::check_gw_groups). */
template <class Ops, bool GW, bool KO, int G = -1>
struct CompSet {
    static_assert(G < 0 || (GW && !KO && G < 3), "groups split GW levels only");
    static constexpr int NC = GW ? Ops::kNGw : Ops::kNScalars;
    static constexpr int NF = Ops::kNFields;
    static constexpr int NSC = Ops::kNScalars;
    static constexpr int kGroups = 3;
    /** Outputs of this group. */
    SHM_CT static constexpr bool out(int n) {
        if (G < 0) return true;
        if (G == 0) return n < NSC;
        if (n < NSC) return false;
        return G == 1 ? (n - NSC) % 6 < 3 : (n - NSC) % 6 >= 3;
    }
    /** Components this group loads. */
    SHM_CT static constexpr bool staged(int n) { return G < 0 || out(n) || n < NSC; }
    SHM_CT static constexpr bool base_stencil(int n) {
        return KO || n < NF ||
               (GW && n >= Ops::kNScalars && n < Ops::kNScalars + 6);
    }
    SHM_CT static constexpr bool stencil(int n) { return base_stencil(n) && staged(n); }
    SHM_CT static constexpr bool point(int n) { return !base_stencil(n) && staged(n); }
    SHM_CT static constexpr int count() {
        int c = 0;
        for (int n = 0; n < NC; ++n) c += stencil(n) ? 1 : 0;
        return c;
    }
    SHM_CT static constexpr int count_points() {
        int c = 0;
        for (int n = 0; n < NC; ++n) c += point(n) ? 1 : 0;
        return c;
    }
    static constexpr int NS = count();
    static constexpr int NP = count_points();
    SHM_CT static constexpr int slot(int n) {
        if (!stencil(n)) return -1;
        int c = 0;
        for (int m = 0; m < n; ++m) c += stencil(m) ? 1 : 0;
        return c;
    }
    SHM_CT static constexpr int comp_of_slot(int s) {
        int c = 0;
        for (int n = 0; n < NC; ++n)
            if (stencil(n)) {
                if (c == s) return n;
                ++c;
            }
        return -1;
    }
    SHM_CT static constexpr int comp_of_point(int p) {
        int c = 0;
        for (int n = 0; n < NC; ++n)
            if (point(n)) {
                if (c == p) return n;
                ++c;
            }
        return -1;
    }
    // (field, momentum) pairs as the RKN/leapfrog integrators pair them
    // (rkn.cpp:32-44): first and second halves of the scalars, then of the
    // twelve GW components.  Pair indices address the force arrays.
    static constexpr int NPAIR = NF + (GW ? 6 : 0);
    SHM_CT static constexpr int pair_field(int p) {
        return p < NF ? p : Ops::kNScalars + (p - NF);
    }
    SHM_CT static constexpr int pair_mom(int p) {
        return p < NF ? NF + p : Ops::kNScalars + 6 + (p - NF);
    }
    SHM_CT static constexpr int pair_of_field(int n) {
        return n < NF ? n
                      : ((GW && n >= Ops::kNScalars && n < Ops::kNScalars + 6)
                             ? NF + (n - Ops::kNScalars)
                             : -1);
    }
    // The outputs a stage writes: components (SSPRK3, fills) and pairs
    // (RKN, leapfrog) of this group, in index order.
    SHM_CT static constexpr int count_out() {
        int c = 0;
        for (int n = 0; n < NC; ++n) c += out(n) ? 1 : 0;
        return c;
    }
    SHM_CT static constexpr int count_out_pairs() {
        int c = 0;
        for (int p = 0; p < NPAIR; ++p) c += out(pair_field(p)) ? 1 : 0;
        return c;
    }
    static constexpr int NO = count_out();
    static constexpr int NOP = count_out_pairs();
    SHM_CT static constexpr int out_comp(int i) {
        int c = 0;
        for (int n = 0; n < NC; ++n)
            if (out(n)) {
                if (c == i) return n;
                ++c;
            }
        return -1;
    }
    SHM_CT static constexpr int out_pair(int i) {
        int c = 0;
        for (int p = 0; p < NPAIR; ++p)
            if (out(pair_field(p))) {
                if (c == i) return p;
                ++c;
            }
        return -1;
    }
    /** Bit masks of out() and staged() (host-side group checks). */
    static constexpr unsigned long long out_mask() {
        unsigned long long m = 0;
        for (int n = 0; n < NC; ++n)
            if (out(n)) m |= 1ull << n;
        return m;
    }
    static constexpr unsigned long long staged_mask() {
        unsigned long long m = 0;
        for (int n = 0; n < NC; ++n)
            if (staged(n)) m |= 1ull << n;
        return m;
    }
};

// ---------------------------------------------------------------------------
// Stage policies: how the stage state is formed and what a cell writes.
//
// Addressing: every array of a level shares one layout, so a cell is named
// by (org, off): org = uniform element offset of the box brick's corner
// (SGPRs), off = unsigned 32-bit offset inside the brick (one VGPR shared
// by all arrays and components).  Component n of array X lives at
// X + org + n * cstride + off, which maps onto global_load/store with a
// scalar base and a 32-bit vector offset.
// ---------------------------------------------------------------------------

// The byte offset is formed in 32 bits so the address is (uniform 64-bit
// base) + zext(32-bit offset): the scalar-base form of global_load/store.
template <class T>
__device__ AMREX_FORCE_INLINE T ld(const void *p, long long base, unsigned off) {
    const char *b = reinterpret_cast<const char *>(static_cast<const T *>(p) + base);
    return *reinterpret_cast<const T *>(b + (unsigned)(off * (unsigned)sizeof(T)));
}
// Read once, at the centre only (remainders): a non-temporal load keeps it
// from displacing the stencil halos in L2.
template <class T>
__device__ AMREX_FORCE_INLINE T ld_nt(const void *p, long long base, unsigned off) {
    const char *b = reinterpret_cast<const char *>(static_cast<const T *>(p) + base);
    return __builtin_nontemporal_load(
        reinterpret_cast<const T *>(b + (unsigned)(off * (unsigned)sizeof(T))));
}
// Stage outputs are written once and not read again by the same kernel:
// non-temporal stores keep them from displacing the stencil halos in L2.
template <class T>
__device__ AMREX_FORCE_INLINE void st(void *p, long long base, unsigned off, T v) {
    char *b = reinterpret_cast<char *>(static_cast<T *>(p) + base);
    __builtin_nontemporal_store(v, reinterpret_cast<T *>(b + (unsigned)(off * (unsigned)sizeof(T))));
}

/** a * b + c with one rounding in the storage type (stage states). */
template <class S>
__device__ AMREX_FORCE_INLINE S fma_st(S a, S b, S c) {
    if constexpr (std::is_same_v<S, float>)
        return __builtin_fmaf(a, b, c);
    else
        return __builtin_fma(a, b, c);
}

/** Split a full-precision value into the stored pair (or keep it FP64):
 *  hi = fl32(u), lo = u - hi rounded to the remainder type. */
template <class T>
__device__ AMREX_FORCE_INLINE void store_state(void *hi, void *lo, long long base,
                                               unsigned off, double u) {
    using S = typename T::S;
    if constexpr (T::comp) {
        const float h = (float)u;
        st<float>(hi, base, off, h);
        const float r = (float)(u - (double)h);
        if constexpr (std::is_same_v<typename T::L, float>)
            st<float>(lo, base, off, r);
        else
            st<uint16_t>(lo, base, off, f32_to_bf16_rne(r));
    } else {
        st<S>(hi, base, off, (S)u);
    }
}
/** The value a stored pair of u reads back as (hi + lo after rounding). */
template <class T>
__device__ AMREX_FORCE_INLINE double pair_value(double u) {
    if constexpr (!T::comp) {
        return (double)(typename T::S)u;
    } else {
        const float h = (float)u;
        const float r = (float)(u - (double)h);
        if constexpr (std::is_same_v<typename T::L, float>)
            return (double)h + (double)r;
        else
            return (double)h + (double)bf16_to_f32(f32_to_bf16_rne(r));
    }
}
/** Memory sources of the load helpers and of the policies' load hooks.
 *
 *  MemLoad reads at once.  MemFetch issues the same loads and keeps their
 *  raw words, in call order, without using them; MemReplay hands those
 *  words to a second call of the same hook (same arguments), which then
 *  forms its values.  The stage kernel fetches the next plane before a
 *  plane's right-hand sides and replays after them, so the arithmetic that
 *  forms a stage state from several arrays does not wait for the loads
 *  before the plane's work.  One hook call reads at most kRawWords 32-bit
 *  words: the largest is a final RKN stage of 5 stages in FP64 (phi0, pi0
 *  and four forces, 12 words). */
constexpr int kRawWords = 12;
/** An empty instruction that reads and rewrites x in its VGPRs: the value
 *  is complete at this point of the program, and the compiler neither
 *  moves its computation later nor merges it with later work. */
template <class X>
__device__ AMREX_FORCE_INLINE void pin_reg(X &x) {
    asm volatile("" : "+v"(x));
}
template <int K>
struct RawWords {
    uint32_t w[K];
};
struct MemLoad {
    template <class S>
    __device__ AMREX_FORCE_INLINE S ld(const void *p, long long b, unsigned off) {
        return ::shm::ld<S>(p, b, off);
    }
    template <class S>
    __device__ AMREX_FORCE_INLINE S ld_nt(const void *p, long long b, unsigned off) {
        return ::shm::ld_nt<S>(p, b, off);
    }
};
template <int K>
struct MemFetch {
    RawWords<K> &r;
    int i = 0;
    __device__ AMREX_FORCE_INLINE explicit MemFetch(RawWords<K> &raw) : r(raw) {}
    // Words beyond K are dropped (and replayed as 0): a hook that read
    // more than K words would give wrong values, never an access outside
    // the buffer.  The indices are constants once the hooks are inlined,
    // so the checks fold away.
    __device__ AMREX_FORCE_INLINE void word(uint32_t x) {
        if (i < K) r.w[i] = x;
        ++i;
    }
    template <class S>
    __device__ AMREX_FORCE_INLINE void put(S v) {
        if constexpr (sizeof(S) == 8) {
            const unsigned long long u = (unsigned long long)__double_as_longlong((double)v);
            word((uint32_t)u);
            word((uint32_t)(u >> 32));
        } else if constexpr (std::is_same_v<S, float>) {
            word(__float_as_uint(v));
        } else {
            word((uint32_t)v);
        }
    }
    template <class S>
    __device__ AMREX_FORCE_INLINE S ld(const void *p, long long b, unsigned off) {
        put<S>(::shm::ld<S>(p, b, off));
        return S();
    }
    template <class S>
    __device__ AMREX_FORCE_INLINE S ld_nt(const void *p, long long b, unsigned off) {
        put<S>(::shm::ld_nt<S>(p, b, off));
        return S();
    }
};
template <int K>
struct MemReplay {
    const RawWords<K> &r;
    int i = 0;
    __device__ AMREX_FORCE_INLINE explicit MemReplay(const RawWords<K> &raw) : r(raw) {}
    __device__ AMREX_FORCE_INLINE uint32_t word() {
        const uint32_t x = i < K ? r.w[i] : 0u;
        ++i;
        return x;
    }
    template <class S>
    __device__ AMREX_FORCE_INLINE S get() {
        if constexpr (sizeof(S) == 8) {
            const unsigned long long lo = word();
            const unsigned long long hi = word();
            return (S)__longlong_as_double((long long)(lo | (hi << 32)));
        } else if constexpr (std::is_same_v<S, float>) {
            return __uint_as_float(word());
        } else {
            return (S)word();
        }
    }
    template <class S>
    __device__ AMREX_FORCE_INLINE S ld(const void *, long long, unsigned) {
        return get<S>();
    }
    template <class S>
    __device__ AMREX_FORCE_INLINE S ld_nt(const void *, long long, unsigned) {
        return get<S>();
    }
};

/** Remainder of the stored pair as FP64 (0 without compensation). */
template <class T, class M>
__device__ AMREX_FORCE_INLINE double load_lo(M &m, const void *lo, long long base,
                                             unsigned off) {
    if constexpr (!T::comp)
        return 0.;
    else if constexpr (std::is_same_v<typename T::L, float>)
        return (double)m.template ld_nt<float>(lo, base, off);
    else
        return (double)bf16_to_f32(m.template ld_nt<uint16_t>(lo, base, off));
}
template <class T>
__device__ AMREX_FORCE_INLINE double load_lo(const void *lo, long long base,
                                             unsigned off) {
    MemLoad m;
    return load_lo<T>(m, lo, base, off);
}
template <class T, class M>
__device__ AMREX_FORCE_INLINE double load_state(M &m, const void *hi, const void *lo,
                                                long long base, unsigned off) {
    return (double)m.template ld<typename T::S>(hi, base, off) + load_lo<T>(m, lo, base, off);
}
template <class T>
__device__ AMREX_FORCE_INLINE double load_state(const void *hi, const void *lo,
                                                long long base, unsigned off) {
    MemLoad m;
    return load_state<T>(m, hi, lo, base, off);
}
/** Tile value (type V) of a stored state: the stored word for FP32 and FP64
 *  tiles, hi + lo for the FP64 tiles of a stored pair.  NT: the value is
 *  read at the centre only (non-temporal loads). */
template <class T, bool NT = false, class M>
__device__ AMREX_FORCE_INLINE typename T::V ld_tile(M &m, const void *hi, const void *lo,
                                                    long long base, unsigned off) {
    using S = typename T::S;
    if constexpr (T::comp && T::exact) {
        const double h = NT ? (double)m.template ld_nt<S>(hi, base, off)
                            : (double)m.template ld<S>(hi, base, off);
        if constexpr (std::is_same_v<typename T::L, float>)
            return h + (double)(NT ? m.template ld_nt<float>(lo, base, off)
                                   : m.template ld<float>(lo, base, off));
        else
            return h + (double)bf16_to_f32(NT ? m.template ld_nt<uint16_t>(lo, base, off)
                                               : m.template ld<uint16_t>(lo, base, off));
    } else {
        return (typename T::V)(NT ? m.template ld_nt<S>(hi, base, off)
                                  : m.template ld<S>(hi, base, off));
    }
}
template <class T, bool NT = false>
__device__ AMREX_FORCE_INLINE typename T::V ld_tile(const void *hi, const void *lo,
                                                    long long base, unsigned off) {
    MemLoad m;
    return ld_tile<T, NT>(m, hi, lo, base, off);
}
/** Same, addressed by a 64-bit element index (auxiliary kernels). */
template <class T>
__device__ AMREX_FORCE_INLINE void store_state_at(void *hi, void *lo,
                                                  long long e, double u) {
    store_state<T>(hi, lo, e, 0u, u);
}
template <class T>
__device__ AMREX_FORCE_INLINE double load_state_at(const void *hi,
                                                   const void *lo, long long e) {
    return load_state<T>(hi, lo, e, 0u);
}

/** Low-storage SSPRK3 (lsssprk3.cpp:14-31) and FillRhs / FillAddRhs.
 *
 *  The stage states y1 = u0 + dt k1 and y2 = u0 + dt/4 k2 are not stored:
 *  stages 2 and 3 rebuild them at every stencil point from u0 and the k
 *  accumulator (X = u0 + c2 * k_in), which removes one write and one read
 *  per component and stage.  k_in and k_out are different arrays (k is read
 *  at neighbouring cells while it is being rewritten).
 *    stage 1: k1 = f(u0)                        -> k_out
 *    stage 2: k2 = f(u0 + dt k1) + k1           -> k_out
 *    stage 3: k3 = f(u0 + dt/4 k2) + k2/4;  u = u0 + 2dt/3 k3  (compensated)
 *    fill   : rhs = f(X)  /  rhs = f(X) + w rhs
 *  Coefficients: c0 weight of k_in, c1 update coefficient, c2 stage-state
 *  coefficient. */
enum SsprkKind : int { kSsStage1 = 1, kSsStage2 = 2, kSsStage3 = 3, kSsFill = 4, kSsFillAdd = 5 };

// Policy hooks used by the stage kernel:
//   load_st(a, org, n, off)          stage-state value of component n
//   load_keep(a, org, n, off, keep)  the same, and the nkeep() raw values
//                                    the store reads at the centre cell
//   load_ext(a, org, off, e)         next() further store-time inputs
//   store(a, org, off, rhs, ctr, e, kc)  combine and write; kc[n] holds the
//                                    kept raw values of component n
template <int K>
struct PolSsprk {
    static constexpr bool composed = (K == kSsStage2 || K == kSsStage3);
    // The next plane's stage states are formed before this plane's stores,
    // by every plane (after the last one from dummy loads), so that no wait
    // for the loads drifts past the stores.  Measured faster for the SSPRK3
    // stages; the RKN stages (less work per plane) do better when the
    // formation is left where the compiler places it, after the stores.
    static constexpr bool kFormEarly = true;
    template <class T>
    using Ext = typename T::S;
    // Kept per component: stage 2: k_in; stage 3: k_in, u0.
    template <class CS, class T>
    static constexpr int nkeep() {
        return K == kSsStage2 ? 1 : K == kSsStage3 ? 2 : 0;
    }
    // Loaded at store time per component: stage 3 remainder of u0 (unless
    // the kept u0 is the whole stored value); fill-add previous rhs.
    template <class CS, class T>
    static constexpr int next() {
        return CS::NO * ((K == kSsStage3 && T::comp && !T::exact) || K == kSsFillAdd ? 1 : 0);
    }
    /** Stage state of component n: u0 + c2 k_in in the tile type (composed
     *  stages), else the source state. */
    template <class CS, class T, bool NT, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_x(M &m, const StageArgs &a,
                                                              long long org, int n,
                                                              unsigned off,
                                                              typename T::V *keep) {
        using S = typename T::S;
        using V = typename T::V;
        const long long b = org + n * a.cstride;
        if constexpr (composed) {
            const V k = (V)(NT ? m.template ld_nt<S>(a.kin, b, off)
                               : m.template ld<S>(a.kin, b, off));
            const V u = ld_tile<T, NT>(m, a.u0, a.u0lo, b, off);
            if (keep) {
                keep[0] = k;
                if constexpr (K == kSsStage3) keep[1] = u;
            }
            return fma_st<V>((V)a.c[2], k, u);
        } else {
            return ld_tile<T, NT>(m, a.x, a.xlo, b, off);
        }
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_st(M &m, const StageArgs &a,
                                                               long long org, int n,
                                                               unsigned off) {
        return load_x<CS, T, false>(m, a, org, n, off, nullptr);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_keep(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        return load_x<CS, T, false>(m, a, org, n, off, keep);
    }
    /** load_keep for a point component: read at the centre only, so
     *  non-temporal (the stencil components' halos keep L2). */
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_pt(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        return load_x<CS, T, true>(m, a, org, n, off, keep);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static void load_ext(M &m, const StageArgs &a,
                                                       long long org, unsigned off,
                                                       typename T::S *e) {
        using S = typename T::S;
        int i = 0;
#pragma unroll
        for (int o = 0; o < CS::NO; ++o) {
            const int n = CS::out_comp(o);
            const long long b = org + n * a.cstride;
            if constexpr (K == kSsStage3 && T::comp && !T::exact)
                e[i++] = (S)load_lo<T>(m, a.u0lo, b, off);
            if constexpr (K == kSsFillAdd) e[i++] = m.template ld<S>(a.kin, b, off);
        }
    }
    template <class CS, class T, int NK>
    __device__ AMREX_FORCE_INLINE static void store(const StageArgs &a,
                                                    long long org, unsigned off,
                                                    const double *rhs,
                                                    const double *ctr,
                                                    const typename T::S *e,
                                                    const typename T::V (*kc)[NK]) {
        using S = typename T::S;
        int i = 0;
#pragma unroll
        for (int o = 0; o < CS::NO; ++o) {
            const int n = CS::out_comp(o);
            const long long b = org + n * a.cstride;
            double k = rhs[n];
            if constexpr (composed) k += a.c[0] * (double)kc[n][0];
            if constexpr (K == kSsFillAdd) k += a.c[0] * (double)e[i++];
            if constexpr (K == kSsStage3) {
                double u0 = (double)kc[n][1];
                if constexpr (T::comp && !T::exact) u0 += (double)e[i++];
                store_state<T>(a.unew, a.unewlo, b, off, u0 + a.c[1] * k);
            } else {
                st<S>(a.kout, b, off, (S)k);
            }
        }
    }
};

/** Runge-Kutta-Nystroem stage with NPREV earlier stages (rkn.cpp:52-119).
 *  Stage state of a field: y = phi0 + c0*pi0 + sum_j c[1+j] F_j; momenta
 *  stay at pi0.  Non-final stages write F (the momentum right-hand sides);
 *  the final stage writes phi = phi0 + c6*pi0 + sum_j c[7+j] F_j and
 *  pi = pi0 + sum_j c[12+j] F_j, with this stage's F (weights c[11], c[16])
 *  taken from registers.  The final stage keeps phi0, pi0 and the F_j that
 *  formed a field's stage state for its own update. */
template <int NPREV, bool FINAL, bool NODE0 = false>
struct PolRkn {
    static constexpr bool kFormEarly = false;
    static_assert(NPREV >= 0 && NPREV <= 4, "RKN supports up to 5 stages");
    template <class T>
    using Ext = typename T::S;
    template <class CS, class T>
    static constexpr int nkeep() {
        return FINAL ? 2 + NPREV : 0;
    }
    // Final stage, per pair: the remainders of phi0 and pi0 (unless the
    // kept values are the whole stored values).
    template <class CS, class T>
    static constexpr int next() {
        return FINAL && T::comp && !T::exact ? CS::NOP * 2 : 0;
    }

    /** Stage state in the tile type (FP32 tiles hold FP32 values, so their
     *  sums are formed in FP32; FP64 tiles of stored pairs start from
     *  hi + lo). */
    template <class CS, class T, bool NT, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_y(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        using S = typename T::S;
        using V = typename T::V;
        const int p = CS::pair_of_field(n);
        const V phi0 = ld_tile<T, NT>(m, a.u0, a.u0lo, org + n * a.cstride, off);
        if (p < 0) return phi0;
        // A first stage at node 0 (RKN4, RKN5) has the state phi0: pi0 is
        // then read at the centre only.  NODE0: known when compiled, so the
        // loads do not depend on a branch.
        if constexpr (!FINAL && NPREV == 0) {
            if constexpr (NODE0)
                return phi0;
            else if (a.c[0] == 0.)
                return phi0;
        }
        const V pi0 = ld_tile<T, NT>(m, a.u0, a.u0lo, org + CS::pair_mom(p) * a.cstride, off);
        V y = phi0;
        y += (V)a.c[0] * pi0;
        if constexpr (FINAL) {
            keep[0] = phi0;
            keep[1] = pi0;
        }
#pragma unroll
        for (int j = 0; j < NPREV; ++j) {
            const V fj = (V)(NT ? m.template ld_nt<S>(a.f[j], org + p * a.cstride, off)
                                : m.template ld<S>(a.f[j], org + p * a.cstride, off));
            if constexpr (FINAL) keep[2 + j] = fj;
            y += (V)a.c[1 + j] * fj;
        }
        return y;
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_keep(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        return load_y<CS, T, false>(m, a, org, n, off, keep);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_st(M &m, const StageArgs &a,
                                                               long long org, int n,
                                                               unsigned off) {
        typename T::V keep[FINAL ? 2 + NPREV : 1];
        return load_y<CS, T, false>(m, a, org, n, off, keep);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_pt(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        return load_y<CS, T, false>(m, a, org, n, off, keep);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static void load_ext(M &m, const StageArgs &a,
                                                       long long org, unsigned off,
                                                       typename T::S *e) {
        using S = typename T::S;
        if constexpr (FINAL && T::comp && !T::exact) {
            int i = 0;
#pragma unroll
            for (int o = 0; o < CS::NOP; ++o) {
                const int p = CS::out_pair(o);
                e[i++] = (S)load_lo<T>(m, a.u0lo, org + CS::pair_field(p) * a.cstride, off);
                e[i++] = (S)load_lo<T>(m, a.u0lo, org + CS::pair_mom(p) * a.cstride, off);
            }
        }
    }
    template <class CS, class T, int NK>
    __device__ AMREX_FORCE_INLINE static void store(const StageArgs &a,
                                                    long long org, unsigned off,
                                                    const double *rhs,
                                                    const double *ctr,
                                                    const typename T::S *e,
                                                    const typename T::V (*kc)[NK]) {
        using S = typename T::S;
        int i = 0;
#pragma unroll
        for (int o = 0; o < CS::NOP; ++o) {
            const int p = CS::out_pair(o);
            const int nf = CS::pair_field(p), nm = CS::pair_mom(p);
            const double fi = rhs[nm];
            if constexpr (!FINAL) {
                st<S>(a.fout, org + p * a.cstride, off, (S)fi);
            } else {
                const long long bf = org + nf * a.cstride, bm = org + nm * a.cstride;
                double phi = (double)kc[nf][0];
                double pi0 = (double)kc[nf][1];
                if constexpr (T::comp && !T::exact) {
                    phi += (double)e[i++];
                    pi0 += (double)e[i++];
                }
                phi += a.c[6] * pi0;
                double pi = pi0;
#pragma unroll
                for (int j = 0; j < NPREV; ++j) {
                    const double fj = (double)kc[nf][2 + j];
                    phi += a.c[7 + j] * fj;
                    pi += a.c[12 + j] * fj;
                }
                phi += a.c[11] * fi;
                pi += a.c[16] * fi;
                store_state<T>(a.unew, a.unewlo, bf, off, phi);
                store_state<T>(a.unew, a.unewlo, bm, off, pi);
            }
        }
    }
};

/** First RKN stage at node 0 (its state is u0) that also writes, per pair,
 *  the second stage's field state y1 = phi0 + c[18] pi0 + c[19] F0 into a.y,
 *  formed in the tile type exactly as PolRkn<1, false> forms it from the
 *  stored phi0, pi0 and F0.  The second stage (PolRkn1Y) then reads y1 at
 *  its stencil points instead of composing it from three arrays there.
 *  y1 is stored in the tile type (FP64 for the FP64 tiles of a stored
 *  pair), so the stored value is the composed one. */
struct PolRkn0Y : PolRkn<0, false, true> {
    template <class CS, class T, int NK>
    __device__ AMREX_FORCE_INLINE static void store(const StageArgs &a,
                                                    long long org, unsigned off,
                                                    const double *rhs,
                                                    const double *ctr,
                                                    const typename T::S *e,
                                                    const typename T::V (*kc)[NK]) {
        using S = typename T::S;
        using Y = typename T::V;
        PolRkn<0, false>::template store<CS, T, NK>(a, org, off, rhs, ctr, e, kc);
#pragma unroll
        for (int o = 0; o < CS::NOP; ++o) {
            const int p = CS::out_pair(o);
            const S f0 = (S)rhs[CS::pair_mom(p)];
            Y y = (Y)ctr[CS::pair_field(p)];
            y += (Y)a.c[18] * (Y)ctr[CS::pair_mom(p)];
            y += (Y)a.c[19] * (Y)f0;
            st<Y>(a.y, org + p * a.cstride, off, y);
        }
    }
};

/** Second RKN stage reading the field states PolRkn0Y wrote; momenta stay
 *  at pi0.  Writes F1 as PolRkn<1, false> does. */
struct PolRkn1Y : PolRkn<1, false> {
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_st(M &m, const StageArgs &a,
                                                               long long org, int n,
                                                               unsigned off) {
        using V = typename T::V;
        const int p = CS::pair_of_field(n);
        if (p < 0) return ld_tile<T>(m, a.u0, a.u0lo, org + n * a.cstride, off);
        return m.template ld<V>(a.y, org + p * a.cstride, off);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_keep(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *) {
        return load_st<CS, T>(m, a, org, n, off);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_pt(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        return load_keep<CS, T>(m, a, org, n, off, keep);
    }
};

/** Second stage of a three-stage RKN step (RKN4) that reads the field
 *  states PolRkn0Y wrote and also writes, per pair, the final stage's field
 *  state y2 = phi0 + c[12] pi0 + c[13] F0 + c[14] F1, formed in the tile
 *  type exactly as PolRkn<2, true> forms it from the stored phi0, pi0, F0
 *  and the F1 written here (same terms, order and rounding).  The final
 *  stage (PolRknF2Y) then reads y2 at its stencil points instead of
 *  composing it from four arrays there.  y2 is stored in the tile type. */
struct PolRkn1YY : PolRkn1Y {
    // phi0 and F0 of each output pair at the centre are store-time inputs,
    // prefetched with the plane (loaded inside store() they would be
    // waited for there), in the tile type.
    template <class T>
    using Ext = typename T::V;
    template <class CS, class T>
    static constexpr int next() {
        return CS::NOP * 2;
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static void load_ext(M &m, const StageArgs &a,
                                                       long long org, unsigned off,
                                                       typename T::V *e) {
        using S = typename T::S;
        using V = typename T::V;
        int i = 0;
#pragma unroll
        for (int o = 0; o < CS::NOP; ++o) {
            const int p = CS::out_pair(o);
            e[i++] = ld_tile<T>(m, a.u0, a.u0lo, org + CS::pair_field(p) * a.cstride, off);
            e[i++] = (V)m.template ld<S>(a.f[0], org + p * a.cstride, off);
        }
    }
    template <class CS, class T, int NK>
    __device__ AMREX_FORCE_INLINE static void store(const StageArgs &a,
                                                    long long org, unsigned off,
                                                    const double *rhs,
                                                    const double *ctr,
                                                    const typename T::V *e,
                                                    const typename T::V (*kc)[NK]) {
        using S = typename T::S;
        using Y = typename T::V;
        PolRkn<1, false>::template store<CS, T, NK>(a, org, off, rhs, ctr,
                                                    (const S *)nullptr, kc);
        int i = 0;
#pragma unroll
        for (int o = 0; o < CS::NOP; ++o) {
            const int p = CS::out_pair(o);
            const S f1 = (S)rhs[CS::pair_mom(p)];
            const Y phi0 = e[i++];
            const Y f0 = e[i++];
            Y y = phi0;
            y += (Y)a.c[12] * (Y)ctr[CS::pair_mom(p)];
            y += (Y)a.c[13] * f0;
            y += (Y)a.c[14] * (Y)f1;
            st<Y>(a.y2, org + p * a.cstride, off, y);
        }
    }
};
/** Final stage of a three-stage RKN step reading the field states
 *  PolRkn1YY wrote at its stencil points; phi0, pi0, F0 and F1 are read at
 *  the centre for the update, as PolRkn<2, true> keeps them. */
struct PolRknF2Y : PolRkn<2, true> {
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_st(M &m, const StageArgs &a,
                                                               long long org, int n,
                                                               unsigned off) {
        using V = typename T::V;
        const int p = CS::pair_of_field(n);
        if (p < 0) return ld_tile<T>(m, a.u0, a.u0lo, org + n * a.cstride, off);
        return m.template ld<V>(a.y2, org + p * a.cstride, off);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_keep(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        using S = typename T::S;
        using V = typename T::V;
        const int p = CS::pair_of_field(n);
        if (p < 0) return ld_tile<T>(m, a.u0, a.u0lo, org + n * a.cstride, off);
        // Fields a component group stages without updating (the scalars in
        // the GW tensor groups) need no centre values.
        if (!CS::out(n)) return m.template ld<V>(a.y2, org + p * a.cstride, off);
        keep[0] = ld_tile<T>(m, a.u0, a.u0lo, org + n * a.cstride, off);
        keep[1] = ld_tile<T>(m, a.u0, a.u0lo, org + CS::pair_mom(p) * a.cstride, off);
        keep[2] = (V)m.template ld<S>(a.f[0], org + p * a.cstride, off);
        keep[3] = (V)m.template ld<S>(a.f[1], org + p * a.cstride, off);
        return m.template ld<V>(a.y2, org + p * a.cstride, off);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_pt(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        return load_keep<CS, T>(m, a, org, n, off, keep);
    }
};

/** Leapfrog kick-drift-kick (leapfrog.cpp:41-56).
 *  Kick 1 (X = u0): vh = pi0 + c0*a(pi) (stored FP64), phi1 = phi0 + c1*vh.
 *  Kick 2 (X = u_new with phi1 and the momenta u_new held before the step,
 *  exactly as the reference's second Rhs sees them): pi1 = vh + c0*a(pi). */
template <bool KICK2>
struct PolLeapfrog {
    static constexpr bool kFormEarly = false;
    template <class T>
    using Ext = double;
    template <class CS, class T>
    static constexpr int nkeep() { return 0; }
    template <class CS, class T>
    static constexpr int next() { return CS::NOP * (KICK2 ? 1 : 2); }

    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_st(M &m, const StageArgs &a,
                                                               long long org, int n,
                                                               unsigned off) {
        return ld_tile<T>(m, a.x, a.xlo, org + n * a.cstride, off);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_keep(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *) {
        return load_st<CS, T>(m, a, org, n, off);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static typename T::V load_pt(
        M &m, const StageArgs &a, long long org, int n, unsigned off, typename T::V *keep) {
        return load_keep<CS, T>(m, a, org, n, off, keep);
    }
    template <class CS, class T, class M>
    __device__ AMREX_FORCE_INLINE static void load_ext(M &m, const StageArgs &a,
                                                       long long org, unsigned off,
                                                       double *e) {
        int i = 0;
#pragma unroll
        for (int o = 0; o < CS::NOP; ++o) {
            const int p = CS::out_pair(o);
            if constexpr (!KICK2) {
                e[i++] = load_state<T>(m, a.u0, a.u0lo, org + CS::pair_mom(p) * a.cstride, off);
                e[i++] = load_state<T>(m, a.u0, a.u0lo, org + CS::pair_field(p) * a.cstride, off);
            } else {
                e[i++] = m.template ld<double>(a.vh, org + p * a.cstride, off);
            }
        }
    }
    template <class CS, class T, int NK>
    __device__ AMREX_FORCE_INLINE static void store(const StageArgs &a,
                                                    long long org, unsigned off,
                                                    const double *rhs,
                                                    const double *ctr,
                                                    const double *e,
                                                    const typename T::V (*)[NK]) {
        int i = 0;
#pragma unroll
        for (int o = 0; o < CS::NOP; ++o) {
            const int p = CS::out_pair(o);
            const int nf = CS::pair_field(p), nm = CS::pair_mom(p);
            const long long bp = org + p * a.cstride;
            const long long bf = org + nf * a.cstride, bm = org + nm * a.cstride;
            if constexpr (!KICK2) {
                const double pi0 = e[i++];
                const double phi0 = e[i++];
                const double vh = pi0 + a.c[0] * rhs[nm];
                st<double>(a.vh, bp, off, vh);
                store_state<T>(a.unew, a.unewlo, bf, off, phi0 + a.c[1] * vh);
            } else {
                const double vh = e[i++];
                store_state<T>(a.unew, a.unewlo, bm, off, vh + a.c[0] * rhs[nm]);
            }
        }
    }
};

// ---------------------------------------------------------------------------
// The fused stage kernel.
// ---------------------------------------------------------------------------

/** Halo-ring cell h of a (TX + 2R) x (TY + 2R) plane: rows above, rows
 *  below, then the R columns left and right of each interior row.  Returns
 *  tile-local (x, y), both >= -R. */
template <int R, int TY, int TX>
__device__ AMREX_FORCE_INLINE void ring_xy(int h, int &x, int &y) {
    constexpr int PX = TX + 2 * R;
    if (h < R * PX) {
        y = -R + h / PX;
        x = -R + h % PX;
    } else if (h < 2 * R * PX) {
        h -= R * PX;
        y = TY + h / PX;
        x = -R + h % PX;
    } else {
        h -= 2 * R * PX;
        y = h / (2 * R);
        const int c = h % (2 * R);
        x = c < R ? -R + c : TX + (c - R);
    }
}

template <class Ops, class Pol, Prec P, int R, bool GW, bool KO, int NW, int RY, int TX,
          int G = -1>
__global__ void __launch_bounds__(64 * NW)
    stage_kernel(const StageArgs a) {
    using T = PrecTraits<P>;
    using V = typename T::V;  // tile values
    using CS = CompSet<Ops, GW, KO, G>;  // G >= 0: one group of a split stage
    constexpr int NC = CS::NC;
    constexpr int NS = CS::NS;
    constexpr int NP = CS::NP;
    // TX cells of a row per thread group (64: one wavefront per row; 32:
    // two rows per wavefront, for narrow boxes); ROWS rows per workgroup.
    static_assert(TX == 64 || TX == 32, "tile width 64 or 32");
    constexpr int ROWS = NW * 64 / TX;
    constexpr int TY = ROWS * RY;
    constexpr int PX = TX + 2 * R, PY = TY + 2 * R;
    constexpr int Q = 2 * R + 1;
    constexpr int NT = 64 * NW;
    constexpr int H = PX * PY - TX * TY;
    constexpr int HPT = (H + NT - 1) / NT;
    static_assert(R >= 1 && R <= 3, "stencil radius must be 1..3");

    __shared__ V lds[2][NS][PY][PX];

    signed char slot[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) slot[n] = (signed char)CS::slot(n);

    const int tx = threadIdx.x & (TX - 1);
    const int w = threadIdx.x / TX;
    const int g = a.ghost;
    // Halo-ring cells of this thread (tile-local) and their LDS slots.
    int hx[HPT], hy[HPT], hl[HPT];
#pragma unroll
    for (int m = 0; m < HPT; ++m) {
        const int h = threadIdx.x + m * NT;
        hx[m] = 0;
        hy[m] = 0;
        if (h < H) ring_xy<R, TY, TX>(h, hx[m], hy[m]);
        hl[m] = (hy[m] + R) * PX + (hx[m] + R);
    }
    const double time = a.prm[kPrmTime], dt = a.prm[kPrmDt], dx = a.prm[kPrmDx];
    // The parameters are read once, here: a read inside the plane loop would
    // make the right-hand side wait for every load in flight before it,
    // the next plane's prefetch included.
    const ParamRegs<GW> pr(a.prm);
    const double *params = pr.user;
    const double *gw_params = pr.gw_ptr(a.prm);
    double ko_s[KO ? NC : 1];  // dissipation strengths, read once as well
    if constexpr (KO) {
#pragma unroll
        for (int n = 0; n < NC; ++n) ko_s[n] = a.ko_strength[n];
    }

    // Persistent walk over the tile table.  With xcd_split the tiles are cut
    // into 8 contiguous ranges, one per XCD (workgroups are dealt to XCDs
    // round-robin), so tiles sharing halos run behind the same L2.
    int t, tstep, tend;
    if (a.xcd_split) {
        const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3;
        const int per = (a.ntiles + 7) >> 3;
        t = xcd * per + l;
        tend = min(a.ntiles, (xcd + 1) * per);
        tstep = gridDim.x >> 3;
    } else {
        t = blockIdx.x;
        tend = a.ntiles;
        tstep = gridDim.x;
    }

    unsigned step = 0;
    for (; t < tend; t += tstep) {
        const TileDesc td = a.tiles[t];
        const BoxDesc bd = a.boxes[td.box];
        const int nx = bd.n[0], ny = bd.n[1], nz = bd.n[2];
        const unsigned js = bd.jstride, ks = bd.kstride;
        const long long org = bd.org;
        const int wrap = bd.wrap;
        // Along axes the box spans periodically, indices within the stencil
        // reach of the box are wrapped onto the box itself (no ghost reads).
        auto wx = [&](int v) {
            return (wrap & 1) ? (v < 0 ? v + nx : (v >= nx ? v - nx : v)) : v;
        };
        auto wy = [&](int v) {
            return (wrap & 2) ? (v < 0 ? v + ny : (v >= ny ? v - ny : v)) : v;
        };
        auto wz = [&](int v) {
            return (wrap & 4) ? (v < 0 ? v + nz : (v >= nz ? v - nz : v)) : v;
        };
        const int lx = td.x0 + tx;
        const bool ldx = lx < nx + R;
        const bool cpx = lx < nx;

        // Brick offsets of this thread's cells and halo-ring cells at z = -g.
        unsigned coff[RY];
        bool ldy[RY], cpy[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int ly = td.y0 + w + r * ROWS;
            ldy[r] = ly < ny + R;
            cpy[r] = ly < ny;
            coff[r] = (unsigned)(wx(lx) + g) + (unsigned)(wy(ly) + g) * js;
        }
        // Loads are unconditional: a cell no valid cell reads (beyond the
        // ghost reach of a partial tile) loads the brick corner instead, and
        // its value is never used.
        unsigned hoff[HPT];
#pragma unroll
        for (int m = 0; m < HPT; ++m) {
            const int h = threadIdx.x + m * NT;
            const int gx = td.x0 + hx[m], gy = td.y0 + hy[m];
            const bool ok = h < H && gx < nx + R && gy < ny + R;
            hoff[m] = ok ? (unsigned)(wx(gx) + g) + (unsigned)(wy(gy) + g) * js : 0u;
        }
        unsigned qoff[RY];  // queue loads: own column, clamped like the ring
#pragma unroll
        for (int r = 0; r < RY; ++r) qoff[r] = (ldx && ldy[r]) ? coff[r] : 0u;
        auto zoff = [&](int z) { return (unsigned)(wz(z) + g) * ks; };

        using E = typename Pol::template Ext<T>;
        constexpr int NE = Pol::template next<CS, T>();
        constexpr int NK = Pol::template nkeep<CS, T>();
        constexpr int NKA = NK > 0 ? NK : 1;
        // Raw values a stencil component's store reads at the centre (k_in,
        // u0, earlier RKN forces) were loaded R+1 planes earlier for the
        // queue head; when the registers allow they ride along in a delay
        // line instead of being read again at store time, when they have
        // left L2.  Otherwise they are reloaded with the point values.
        constexpr bool DELAY =
            NK > 0 && NS * RY * (R + 2) * NK * (int)(sizeof(V) / 4) <= 64;
        V q[NS][RY][Q];
        V nq[NS][RY];
        V hv[NS][HPT];
        V pv[NP > 0 ? NP : 1][RY];
        E ev[RY][NE > 0 ? NE : 1];  // store-time inputs of the next plane
        V kq[NS][RY][DELAY ? R + 1 : 1][NKA];  // kept values, planes z..z+R
        V nk[NS][RY][NKA];  // kept values of the queue head / of plane z+1
        V pk[NP > 0 ? NP : 1][RY][NKA];  // kept values of point components

        // Prologue: planes z0-R .. z0+R-1 into q[..][1..2R].
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int r = 0; r < RY; ++r) {
#pragma unroll
                for (int m = 1; m < Q; ++m) {
                    MemLoad ml;
                    const unsigned o = qoff[r] + zoff(td.z0 - R + m - 1);
                    if (DELAY && m > R)
                        q[s][r][m] = Pol::template load_keep<CS, T>(
                            ml, a, org, CS::comp_of_slot(s), o, kq[s][r][DELAY ? m - R : 0]);
                    else
                        q[s][r][m] = Pol::template load_st<CS, T>(
                            ml, a, org, CS::comp_of_slot(s), o);
                    if constexpr (Pol::kFormEarly) {
                        pin_reg(q[s][r][m]);
                        if (DELAY && m > R) {
#pragma unroll
                            for (int j = 0; j < NK; ++j) pin_reg(kq[s][r][DELAY ? m - R : 0][j]);
                        }
                    }
                }
            }

        // The next plane's loads (queue head zq, halo ring and point values
        // of plane zp) in two steps: fetch() issues them and keeps the raw
        // words, form() makes the stage states, kept values and store-time
        // inputs from those words.  A plane fetches its successor before its
        // right-hand sides and forms it after them, so the loads arrive
        // while the plane is computed; the plane's stores follow form(), so
        // its wait covers nothing issued after the loads.
        constexpr int KE = 2 * (NE > 0 ? NE : 1);  // ext words: <= 2 per value
        RawWords<kRawWords> rq[NS][RY], rh[NS][HPT], rp[NP > 0 ? NP : 1][RY];
        RawWords<kRawWords> rk[NS][RY];
        RawWords<KE> re[RY];
        // live = false (after the last plane): the same loads, all from the
        // brick's first cell, and the formed values are not used.  Every
        // plane then issues and waits for the same loads, so no wait for
        // them is left to a later point of the program, where it would also
        // wait for stores issued in between.
        auto plane_loads = [&](auto tag, int zq, int zp, bool live) {
            constexpr bool FORM = decltype(tag)::value;
            using MQ = std::conditional_t<FORM, MemReplay<kRawWords>, MemFetch<kRawWords>>;
            using ME = std::conditional_t<FORM, MemReplay<KE>, MemFetch<KE>>;
            const unsigned oq = live ? zoff(zq) : 0u, op = live ? zoff(zp) : 0u;
            unsigned qo[RY], ho[HPT];
#pragma unroll
            for (int r = 0; r < RY; ++r) qo[r] = live ? qoff[r] : 0u;
#pragma unroll
            for (int m = 0; m < HPT; ++m) ho[m] = live ? hoff[m] : 0u;
            V dk[NKA];
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int r = 0; r < RY; ++r) {
                    MQ mq(rq[s][r]);
                    V y;
                    if constexpr (DELAY)
                        y = Pol::template load_keep<CS, T>(mq, a, org, CS::comp_of_slot(s),
                                                           qo[r] + oq, FORM ? nk[s][r] : dk);
                    else
                        y = Pol::template load_st<CS, T>(mq, a, org, CS::comp_of_slot(s),
                                                         qo[r] + oq);
                    if constexpr (FORM) nq[s][r] = y;
                }
#pragma unroll
            for (int m = 0; m < HPT; ++m)
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    MQ mh(rh[s][m]);
                    const V y = Pol::template load_st<CS, T>(mh, a, org, CS::comp_of_slot(s),
                                                             ho[m] + op);
                    if constexpr (FORM) hv[s][m] = y;
                }
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int r = 0; r < RY; ++r) {
                    MQ mp(rp[p][r]);
                    const V y = Pol::template load_pt<CS, T>(mp, a, org, CS::comp_of_point(p),
                                                             qo[r] + op, FORM ? pk[p][r] : dk);
                    if constexpr (FORM) pv[p][r] = y;
                }
            if constexpr (NK > 0 && !DELAY) {
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int r = 0; r < RY; ++r) {
                        MQ mk(rk[s][r]);
                        (void)Pol::template load_keep<CS, T>(mk, a, org, CS::comp_of_slot(s),
                                                             qo[r] + op, FORM ? nk[s][r] : dk);
                    }
            }
            if constexpr (NE > 0) {
                E de[NE];
#pragma unroll
                for (int r = 0; r < RY; ++r) {
                    ME me(re[r]);
                    Pol::template load_ext<CS, T>(me, a, org, qo[r] + op, FORM ? ev[r] : de);
                }
            }
            if constexpr (FORM && Pol::kFormEarly) {
                // The formed values are complete here: the compiler would
                // otherwise sink their arithmetic past the plane's stores
                // (or into the next plane), and the waits for their loads
                // would then also wait for those stores.
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int r = 0; r < RY; ++r) {
                        pin_reg(nq[s][r]);
#pragma unroll
                        for (int j = 0; j < NK; ++j) pin_reg(nk[s][r][j]);
                    }
#pragma unroll
                for (int m = 0; m < HPT; ++m)
#pragma unroll
                    for (int s = 0; s < NS; ++s) pin_reg(hv[s][m]);
#pragma unroll
                for (int p = 0; p < NP; ++p)
#pragma unroll
                    for (int r = 0; r < RY; ++r) {
                        pin_reg(pv[p][r]);
#pragma unroll
                        for (int j = 0; j < NK; ++j) pin_reg(pk[p][r][j]);
                    }
#pragma unroll
                for (int r = 0; r < RY; ++r)
#pragma unroll
                    for (int i = 0; i < NE; ++i) pin_reg(ev[r][i]);
            }
        };
        using Fetch = std::false_type;
        using Form = std::true_type;
        plane_loads(Fetch{}, td.z0 + R, td.z0, true);
        plane_loads(Form{}, td.z0 + R, td.z0, true);

        for (int z = td.z0; z < td.z1; ++z, ++step) {
            const int buf = step & 1;
            // Advance the register queues by one plane.
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int r = 0; r < RY; ++r) {
#pragma unroll
                    for (int m = 0; m < Q - 1; ++m) q[s][r][m] = q[s][r][m + 1];
                    q[s][r][Q - 1] = nq[s][r];
                    if constexpr (NK > 0) {
                        constexpr int D = DELAY ? R : 0;
#pragma unroll
                        for (int m = 0; m < D; ++m)
#pragma unroll
                            for (int j = 0; j < NK; ++j)
                                kq[s][r][m][j] = kq[s][r][m + 1][j];
#pragma unroll
                        for (int j = 0; j < NK; ++j) kq[s][r][D][j] = nk[s][r][j];
                    }
                }
            // Stage plane z: own centres and the halo ring.
#pragma unroll
            for (int s = 0; s < NS; ++s) {
#pragma unroll
                for (int r = 0; r < RY; ++r)
                    lds[buf][s][w + r * ROWS + R][tx + R] = q[s][r][R];
#pragma unroll
                for (int m = 0; m < HPT; ++m)
                    if (m < HPT - 1 || threadIdx.x + m * NT < H)
                        (&lds[buf][s][0][0])[hl[m]] = hv[s][m];
            }
            V pc[NP > 0 ? NP : 1][RY];
            V pkc[NP > 0 ? NP : 1][RY][NKA];
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int r = 0; r < RY; ++r) {
                    pc[p][r] = pv[p][r];
#pragma unroll
                    for (int j = 0; j < NK; ++j) pkc[p][r][j] = pk[p][r][j];
                }
            E ec[RY][NE > 0 ? NE : 1];
#pragma unroll
            for (int r = 0; r < RY; ++r)
#pragma unroll
                for (int i = 0; i < NE; ++i) ec[r][i] = ev[r][i];

            const bool more = z + 1 < td.z1;
            if (Pol::kFormEarly || more) plane_loads(Fetch{}, z + 1 + R, z + 1, more);
            __syncthreads();

            const unsigned zc = zoff(z);
            double rhs[RY][NC];
            double ctr[RY][NC];
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                if (!(cpx && cpy[r])) continue;
                const int gi = bd.lo[0] + lx, gj = bd.lo[1] + td.y0 + w + r * ROWS,
                          gk = bd.lo[2] + z;
                if constexpr (G >= 0) {
                    // Components of other groups: not loaded, read as NaN.
#pragma unroll
                    for (int n = 0; n < NC; ++n) ctr[r][n] = __builtin_nan("");
                }
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    ctr[r][CS::comp_of_slot(s)] = (double)q[s][r][R];
#pragma unroll
                for (int p = 0; p < NP; ++p)
                    ctr[r][CS::comp_of_point(p)] = (double)pc[p][r];

                TileCtx tc;
                tc.ci = gi;
                tc.cj = gj;
                tc.ck = gk;
                tc.R = R;
                tc.f32 = std::is_same_v<V, float> ? 1 : 0;
                tc.lds_pitch = PX;
                tc.lds_cstride = PX * PY;
                tc.zq_cstride = RY * Q;
                tc.lds = &lds[buf][0][w + r * ROWS + R][tx + R];
                tc.zq = &q[0][r][0];
                tc.ctr = ctr[r];
                tc.slot = slot;
                amrex::Array4<const double> sv;
                sv.mode = amrex::Array4<const double>::kTile;
                sv.ext = &tc;
                sv.ncomp = NC;

#pragma unroll
                for (int n = 0; n < NC; ++n) rhs[r][n] = 0.;
                amrex::Array4<double> rv;
                rv.mode = amrex::Array4<double>::kSink;
                rv.sink = rhs[r];
                rv.ncomp = NC;

                Ops::rhs(rv, sv, gi, gj, gk, a.lev, time, dt, dx, params);
                if constexpr (GW) {
                    Ops::gw_rhs(rv, sv, gi, gj, gk, a.lev, time, dt, dx,
                                gw_params);
                    Ops::gw_backreaction(rv, sv, gi, gj, gk, a.lev, time, dt,
                                         dx, params, gw_params);
                }
                if constexpr (KO) {
                    if (a.ko_order == 2) {
#pragma unroll
                        for (int n = 0; n < NC; ++n)
                            rhs[r][n] += sledgehamr::kernels::KreissOligerDissipation<2>(
                                sv, gi, gj, gk, n, dx, ko_s[n]);
                    } else if (a.ko_order == 3) {
#pragma unroll
                        for (int n = 0; n < NC; ++n)
                            rhs[r][n] += sledgehamr::kernels::KreissOligerDissipation<3>(
                                sv, gi, gj, gk, n, dx, ko_s[n]);
                    }
                }
            }

            if (Pol::kFormEarly || more) plane_loads(Form{}, z + 1 + R, z + 1, more);

#pragma unroll
            for (int r = 0; r < RY; ++r) {
                if (!(cpx && cpy[r])) continue;
                V kc[NC][NKA];
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int j = 0; j < NK; ++j)
                        kc[CS::comp_of_slot(s)][j] = kq[s][r][0][j];
#pragma unroll
                for (int p = 0; p < NP; ++p)
#pragma unroll
                    for (int j = 0; j < NK; ++j)
                        kc[CS::comp_of_point(p)][j] = pkc[p][r][j];
                Pol::template store<CS, T, NKA>(a, org, coff[r] + zc, rhs[r], ctr[r],
                                                ec[r], kc);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Ghost-cell fill from a copy plan (periodic wrap and box-to-box).
// ---------------------------------------------------------------------------

struct HaloArgs {
    void *arr;
    long long cstride;
    const BoxDesc *boxes;
    const HaloSeg *segs;
    int nseg;
    long long total;
    int ncomp;
    int comps[32];
};

template <class T>
__global__ void __launch_bounds__(256) halo_kernel(const HaloArgs a) {
    T *arr = static_cast<T *>(a.arr);
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
        const BoxDesc &db = a.boxes[sg.dst_box];
        const BoxDesc &sb = a.boxes[sg.src_box];
        const long long de = db.base + (gi - db.lo[0]) +
                             (long long)(gj - db.lo[1]) * db.jstride +
                             (long long)(gk - db.lo[2]) * db.kstride;
        const long long se = sb.base + (gi - sg.shift[0] - sb.lo[0]) +
                             (long long)(gj - sg.shift[1] - sb.lo[1]) * sb.jstride +
                             (long long)(gk - sg.shift[2] - sb.lo[2]) * sb.kstride;
        for (int c = 0; c < a.ncomp; ++c) {
            const long long cb = (long long)a.comps[c] * a.cstride;
            arr[cb + de] = arr[cb + se];
        }
    }
}

// ---------------------------------------------------------------------------
// Probe: run the project's cell functions once with a recording view.
// ---------------------------------------------------------------------------

struct ProbeOut {
    int ncomp;
    int reach[ProbeCtx::kMaxComp];
    int zreach[ProbeCtx::kMaxComp];
    int offaxis[ProbeCtx::kMaxComp];
    int tag_reach[ProbeCtx::kMaxComp];
};

/** Dependencies of the GW cell functions (Rhs, GravitationalWavesRhs and
 *  the backreaction, summed per output as a stage sums them): dep[n] gets
 *  bit m when perturbing component m (at every offset) changes output n, on
 *  two sets of offset-dependent inputs.  An output that is not a number is
 *  taken to read everything. */
template <class Ops>
__global__ void gwdep_kernel(unsigned long long *dep, const double *params,
                             const double *gw_params, double time, double dt,
                             double dx) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    constexpr int NC = Ops::kNGw;
    static_assert(NC <= 64, "dependency masks hold 64 components");
    ProbeCtx pc;
    pc.ci = 7;
    pc.cj = 9;
    pc.ck = 11;
    pc.ncomp = NC;
    pc.hashed = 1;
    for (int n = 0; n < ProbeCtx::kMaxComp; ++n) {
        pc.reach[n] = 0;
        pc.zreach[n] = 0;
        pc.offaxis[n] = 0;
        pc.pert[n] = 0.;
    }
    amrex::Array4<const double> sv;
    sv.mode = amrex::Array4<const double>::kProbe;
    sv.ext = &pc;
    sv.ncomp = NC;
    auto eval = [&](double *out) {
        for (int n = 0; n < NC; ++n) out[n] = 0.;
        amrex::Array4<double> rv;
        rv.mode = amrex::Array4<double>::kSink;
        rv.sink = out;
        rv.ncomp = NC;
        Ops::rhs(rv, sv, 7, 9, 11, 0, time, dt, dx, params);
        Ops::gw_rhs(rv, sv, 7, 9, 11, 0, time, dt, dx, gw_params);
        Ops::gw_backreaction(rv, sv, 7, 9, 11, 0, time, dt, dx, params, gw_params);
    };
    unsigned long long d[NC];
    for (int n = 0; n < NC; ++n) d[n] = 0;
    for (int trial = 0; trial < 2; ++trial) {
        pc.seed = 0x2545f491u * (unsigned)(trial + 1);
        double base[NC], o[NC];
        eval(base);
        for (int m = 0; m < NC; ++m) {
            pc.pert[m] = 1e-3;
            eval(o);
            pc.pert[m] = 0.;
            for (int n = 0; n < NC; ++n)
                if (!(o[n] == base[n]) || !(base[n] == base[n])) d[n] |= 1ull << m;
        }
    }
    for (int n = 0; n < NC; ++n) dep[n] = d[n];
}

template <class Ops, bool GW>
__global__ void probe_kernel(ProbeOut *out, const double *params,
                             const double *gw_params, double time, double dt,
                             double dx) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    constexpr int NC = GW ? Ops::kNGw : Ops::kNScalars;
    ProbeCtx pc;
    pc.ci = 7;
    pc.cj = 9;
    pc.ck = 11;
    pc.ncomp = NC;
    for (int n = 0; n < ProbeCtx::kMaxComp; ++n) {
        pc.reach[n] = 0;
        pc.zreach[n] = 0;
        pc.offaxis[n] = 0;
    }
    amrex::Array4<const double> sv;
    sv.mode = amrex::Array4<const double>::kProbe;
    sv.ext = &pc;
    sv.ncomp = NC;
    double rhs[NC];
    for (int n = 0; n < NC; ++n) rhs[n] = 0.;
    amrex::Array4<double> rv;
    rv.mode = amrex::Array4<double>::kSink;
    rv.sink = rhs;
    Ops::rhs(rv, sv, 7, 9, 11, 0, time, dt, dx, params);
    if constexpr (GW) {
        Ops::gw_rhs(rv, sv, 7, 9, 11, 0, time, dt, dx, gw_params);
        Ops::gw_backreaction(rv, sv, 7, 9, 11, 0, time, dt, dx, params,
                             gw_params);
    }
    out->ncomp = NC;
    for (int n = 0; n < ProbeCtx::kMaxComp; ++n) {
        out->reach[n] = pc.reach[n];
        out->zreach[n] = pc.zreach[n];
        out->offaxis[n] = pc.offaxis[n];
    }
}

}  // namespace shm

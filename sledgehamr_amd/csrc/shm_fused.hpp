// Fused first and second stages of the low-storage SSPRK3 step
// (lsssprk3.cpp:20-25): k2 = f(u0 + dt k1) + k1 with k1 = f(u0), in one pass
// over the level, so k1 is neither written nor read back.
//
// The second stage needs u1 = u0 + dt k1 within R cells of each tile cell,
// so the first stage is evaluated on the tile plus a ring of R cells (the
// ring values are what the neighbouring tiles compute for themselves, so
// the result equals the two-launch path bit for bit).  Per tile and plane:
//   * u0 of the stencil components over the tile plus 2R cells sits in a
//     circular LDS window of 2R+2 planes; u0 of the point-wise components
//     stays in the registers of the thread that loaded it;
//   * stage one at plane zz writes u1 of the stencil components into a
//     second window (tile plus R, R+2 planes) for the x and y neighbours,
//     and keeps u1 of the tile's own cells in a register queue of 2R+2
//     planes (their z neighbours) and k1 and the point-wise u1 in delay
//     lines of as many planes (the queues rotate: the plane loop is
//     unrolled by their length);
//   * stage two at plane zz - R - 1 reads plane zz - R - 1 of the u1 window,
//     the queue and the delay line, and writes k2.
// Stage two trails stage one by R+1 planes, so it reads only planes of
// earlier passes, and each window holds one plane more than its readers
// use: one barrier per plane suffices (a wave that runs ahead into the next
// plane writes slots no slower wave still reads).  Planes are loaded one
// plane ahead.  Applies to boxes that wrap periodically along every axis (a
// single box spanning the domain, or an ensemble of independent boxes),
// where the extra reach needs no ghost cells.
//
// With RKN set, the same pass runs the first two stages of a Runge-Kutta-
// Nystroem step whose first node is 0 (rkn.cpp:52-119; RKN4 and RKN5): the
// first stage's state is u0, its forces F0 (the momenta's right-hand sides)
// are written for the tile's own cells, and the second stage's state of a
// field is phi0 + c0 pi0 + c1 F0 (momenta stay at pi0), formed in the storage
// type as the separate second stage forms it; that stage writes F1.
#pragma once

#include "shm_stage.hpp"

namespace shm {

struct FusedArgs {
    const BoxDesc *boxes;
    const TileDesc *tiles;
    int ntiles;
    int lev;
    const double *prm1;  // stage record of stage one (t0)
    const double *prm2;  // stage record of stage two (t0 + dt)
    int xcd_split;
    int ghost;
    long long cstride;
    const void *u0;
    const void *u0lo;    // remainder of u0 (FP64 tiles of stored pairs)
    void *kout;          // SSPRK3: k2; RKN: F1
    double dt;           // u1 = u0 + dt k1
    void *f0out;         // RKN: F0
    double c0, c1;       // RKN: second-stage state phi0 + c0 pi0 + c1 F0
};

/** Static shape of a fused tile: 64 x NW cells, radius R. */
template <class Ops, Prec P, int R, bool GW, bool KO, int NW>
struct FusedShape {
    using S = typename PrecTraits<P>::S;  // storage (k1, F0)
    using V = typename PrecTraits<P>::V;  // tile values (u0, u1)
    using CS = CompSet<Ops, GW, KO>;
    static constexpr int NC = CS::NC, NS = CS::NS, NP = CS::NP;
    static constexpr int TX = 64, TY = NW, NT = 64 * NW, W = 2 * R + 1, WS = W + 1;
    static constexpr int E1X = TX + 2 * R, E1Y = TY + 2 * R, E1 = E1X * E1Y;
    static constexpr int E0X = TX + 4 * R, E0Y = TY + 4 * R, E0 = E0X * E0Y;
    static constexpr int L0 = (E0 + NT - 1) / NT;  // window points per thread
    static constexpr int L1 = (E1 + NT - 1) / NT;  // stage-one points per thread
    static constexpr int WS1 = R + 2;  // u1 window planes
    static constexpr long lds_bytes = (long)sizeof(V) * ((long)NS * WS * E0 + (long)NS * WS1 * E1);
    // Instantiated when the LDS fits one workgroup per CU and k1 and the
    // point-wise u1 over the R+2 planes stage two spans stay within 56
    // registers (their queues are padded to 2R+2 planes to rotate).
    static constexpr bool fits =
        !KO && lds_bytes <= 160 * 1024 && L1 <= 3 &&
        (NC * (int)(sizeof(S) / 4) + NP * (int)(sizeof(V) / 4)) * (R + 2) <= 56;
};

/** f(integral_constant<int, I>) for each I of the sequence, in order. */
template <int... I, class Fn>
__device__ __attribute__((always_inline)) inline void unroll_phases(
    std::integer_sequence<int, I...>, Fn &&f) {
    (f(std::integral_constant<int, I>{}), ...);
}

template <class Ops, Prec P, int R, bool GW, bool KO, int NW, bool RKN>
__global__ void __launch_bounds__(64 * NW) fused12_kernel(const FusedArgs a) {
    using F = FusedShape<Ops, P, R, GW, KO, NW>;
    using T = PrecTraits<P>;
    using S = typename F::S;
    using V = typename F::V;
    using CS = typename F::CS;
    constexpr int NC = F::NC, NS = F::NS, NP = F::NP, NPA = NP > 0 ? NP : 1;
    constexpr int W = F::W, WS = F::WS, WS1 = F::WS1, NT = F::NT;
    constexpr int E0X = F::E0X, E0 = F::E0, E1X = F::E1X, E1 = F::E1;
    constexpr int L0 = F::L0, L1 = F::L1;
    static_assert(F::fits, "fused tile exceeds LDS or register budget");

    __shared__ V w0[NS][WS][E0];  // u0, stencil components, tile + 2R
    __shared__ V w1[NS][WS1][E1];  // u1, stencil components, tile + R

    signed char slot[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n) slot[n] = (signed char)CS::slot(n);

    const int tid = threadIdx.x;
    const int tx = tid & 63, w = tid >> 6;
    const int g = a.ghost;
    const double t1 = a.prm1[kPrmTime], t2 = a.prm2[kPrmTime];
    const double dt = a.prm1[kPrmDt], dx = a.prm1[kPrmDx];
    // Parameters in registers, read once (a read in the plane loop would
    // wait for the next plane's prefetch).
    const ParamRegs<GW> pr1(a.prm1), pr2(a.prm2);
    const double *par1 = pr1.user, *par2 = pr2.user;
    const double *gpar1 = pr1.gw_ptr(a.prm1), *gpar2 = pr2.gw_ptr(a.prm2);
    const V cdt = (V)a.dt;

    // Stage-one points of this thread: its own cell (x = tx, y = w), then a
    // ring cell of the tile + R region.
    int e1x[L1], e1y[L1];
    bool e1ok[L1];
#pragma unroll
    for (int m = 0; m < L1; ++m) {
        e1x[m] = 0;
        e1y[m] = 0;
        e1ok[m] = m == 0 || tid + NT * m < E1;
        if (m == 0) {
            e1x[m] = tx;
            e1y[m] = w;
        } else if (e1ok[m]) {
            ring_xy<R, NW, 64>(tid + NT * (m - 1), e1x[m], e1y[m]);
        }
    }

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

    for (; t < tend; t += tstep) {
        const TileDesc td = a.tiles[t];
        const BoxDesc bd = a.boxes[td.box];
        const int nx = bd.n[0], ny = bd.n[1], nz = bd.n[2];
        const unsigned js = bd.jstride, ks = bd.kstride;
        const long long org = bd.org;
        auto wrapi = [](int v, int n) { return ((v % n) + n) % n; };
        auto zoff = [&](int z) { return (unsigned)(wrapi(z, nz) + g) * ks; };
        auto col = [&](int x, int y) {
            return (unsigned)(wrapi(td.x0 + x, nx) + g) +
                   (unsigned)(wrapi(td.y0 + y, ny) + g) * js;
        };
        // Window points (tile + 2R) of this thread.
        unsigned c0[L0];
        int l0[L0];
        bool ok0[L0];
#pragma unroll
        for (int m = 0; m < L0; ++m) {
            const int e = tid + NT * m;
            ok0[m] = e < E0;
            const int ee = ok0[m] ? e : 0;
            l0[m] = ee;
            c0[m] = col(ee % E0X - 2 * R, ee / E0X - 2 * R);
        }
        unsigned c1[L1];
        int l1[L1];
#pragma unroll
        for (int m = 0; m < L1; ++m) {
            c1[m] = col(e1x[m], e1y[m]);
            l1[m] = (e1y[m] + R) * E1X + (e1x[m] + R);
        }
        const bool own = td.x0 + tx < nx && td.y0 + w < ny;
        const unsigned cown = (unsigned)(td.x0 + tx + g) + (unsigned)(td.y0 + w + g) * js;

        // Plane bookkeeping without divisions in the z loop: wrapped plane
        // indices advance by one per plane.  Window slots are those of
        // consecutive planes modulo WS (the lowest loaded plane z0 - 2R has
        // slot 0), so with WS equal to the unroll length they are constants
        // of the plane's phase; the u1 window (WS1 slots) keeps a counter.
        auto next = [](int v, int n) { return v + 1 == n ? 0 : v + 1; };
        int zwin = wrapi(td.z0 - 2 * R, nz);  // next window plane to load
        int zpt = wrapi(td.z0 - R, nz);       // next point plane to load
        int zst = wrapi(td.z0, nz);           // next plane to store

        // Next plane's loads (registers).  The point components alternate
        // between two buffers: plane zz's are read while plane zz+1's load.
        V v0[NS][L0];
        V vp[2][NPA][L1];
        auto load_window = [&]() {
            const unsigned zo = (unsigned)(zwin + g) * ks;
            zwin = next(zwin, nz);
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int m = 0; m < L0; ++m)
                    v0[s][m] = ld_tile<T>(a.u0, a.u0lo, org + CS::comp_of_slot(s) * a.cstride,
                                          c0[m] + zo);
        };
        auto load_point = [&](V (&dst)[NPA][L1]) {
            const unsigned zo = (unsigned)(zpt + g) * ks;
            zpt = next(zpt, nz);
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int m = 0; m < L1; ++m)
                    dst[p][m] = ld_tile<T>(a.u0, a.u0lo, org + CS::comp_of_point(p) * a.cstride,
                                           c1[m] + zo);
        };
        auto store_window = [&](const int sl) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int m = 0; m < L0; ++m)
                    if (ok0[m]) w0[s][sl][l0[m]] = v0[s][m];
        };

        // Prologue: planes z0-2R .. z0-1 of the window; prefetch plane z0
        // of the window and plane z0-R of the point components.
        for (int z = 0; z < 2 * R; ++z) {
            load_window();
            store_window(z);
        }
        load_window();
        load_point(vp[0]);
        int s1w = 0;       // u1 window slot of plane zz
        int swin = 2 * R;  // window slot of plane zz + R

        // Register queues of QL = 2R+2 planes.  They rotate instead of
        // shifting: the plane loop is unrolled by QL, and the plane of phase
        // P (its distance from plane z0 - R, modulo QL) keeps logical entry j
        // (j = QL-1 the newest) in physical entry (j + P) mod QL, so a plane
        // moves no register.
        constexpr int QL = 2 * R + 2;
        static_assert(QL == WS, "window slots are constants of the phase");
        S dk1[QL][NC];   // k1 of the own cell, planes zz-2R-1 .. zz (stage two reads zz-R-1)
        V du1[QL][NPA];  // u1 of the point components, same planes
        V uq[NS][QL];    // u1 of the own cell, stencil components, same planes

        // A plane's global stores are issued at the top of the next plane,
        // after its window stores have waited for the prefetched loads and
        // before the next prefetch: a store between a load and its use
        // would make that use wait for the store too (the wait counts
        // loads and stores in issue order).
        constexpr int NPR = CS::NPAIR > 0 ? CS::NPAIR : 1;
        constexpr int NKS = RKN ? NPR : NC;
        S pend_k[NKS];       // stage two's k2 (SSPRK3) or F1 (RKN)
        S pend_f[NPR];       // RKN: stage one's F0
        bool has_k = false, has_f = false;
        unsigned o_k = 0, o_f = 0;
        // Output bases per component, uniform across the workgroup: a store
        // adds only the cell's 32-bit offset.
        S *kb[NKS];
        S *fb[NPR];
#pragma unroll
        for (int n = 0; n < NKS; ++n) kb[n] = static_cast<S *>(a.kout) + org + n * a.cstride;
#pragma unroll
        for (int p = 0; p < NPR; ++p)
            fb[p] = RKN ? static_cast<S *>(a.f0out) + org + p * a.cstride : nullptr;
        auto flush = [&]() {
            if (has_k) {
#pragma unroll
                for (int n = 0; n < NKS; ++n) st<S>(kb[n], 0, o_k, pend_k[n]);
            }
            if constexpr (RKN) {
                if (has_f) {
#pragma unroll
                    for (int p = 0; p < CS::NPAIR; ++p) st<S>(fb[p], 0, o_f, pend_f[p]);
                }
            }
            has_k = false;
            has_f = false;
        };

        auto plane = [&](auto phase, const int zz) __attribute__((always_inline)) {
            constexpr int PH = decltype(phase)::value;
            auto ix = [](int j) { return (j + PH) % QL; };  // logical -> physical
            // Stage one runs on plane zz (window planes zz-R .. zz+R) up to
            // the last plane stage two needs; stage two on plane zz-R-1.
            const bool one = zz < td.z1 + R;
            const V (&vpc)[NPA][L1] = vp[PH & 1];  // point components of plane zz
            // Plane zz + R enters window slot (zz + R - (z0 - 2R)) mod WS
            // (a counter: as a constant of the phase its address would stay
            // live across the unrolled loop).
            if (one) {
                store_window(swin);
                swin = swin + 1 == WS ? 0 : swin + 1;
            }
            // Wait for this plane's point components here, before the next
            // plane's loads: a wait issued after those would wait for them
            // too (the counter drains in issue order).  On the last plane
            // the buffer holds an older plane, long arrived.
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int m = 0; m < L1; ++m) pin_reg(vp[PH & 1][p][m]);
            flush();
            if (zz + 1 < td.z1 + R) {
                load_window();
                load_point(vp[(PH + 1) & 1]);
            }
            __syncthreads();

            // ---- stage one at plane zz over the tile + R ----
            // Plane zz - R + d is in window slot (PH + d) mod WS.
            constexpr int sz = (PH + R) % WS;
            if (one) {
                int zpl0[W];
#pragma unroll
                for (int d = 0; d < W; ++d) zpl0[d] = ((PH + d) % WS - sz) * E0;
#pragma unroll
                for (int m = 0; m < L1; ++m) {
                    if (!e1ok[m]) continue;
                    const int lc = (e1y[m] + 2 * R) * E0X + (e1x[m] + 2 * R);
                    const int gi = bd.lo[0] + td.x0 + e1x[m], gj = bd.lo[1] + td.y0 + e1y[m],
                              gk = bd.lo[2] + zz;
                    V u0v[NC];
                    double ctr[NC];
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        u0v[CS::comp_of_slot(s)] = w0[s][sz][lc];
                        ctr[CS::comp_of_slot(s)] = (double)w0[s][sz][lc];
                    }
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        u0v[CS::comp_of_point(p)] = vpc[p][m];
                        ctr[CS::comp_of_point(p)] = (double)vpc[p][m];
                    }
                    TileCtx tc;
                    tc.ci = gi;
                    tc.cj = gj;
                    tc.ck = gk;
                    tc.R = R;
                    tc.f32 = std::is_same_v<V, float> ? 1 : 0;
                    tc.lds_pitch = E0X;
                    tc.lds_cstride = WS * E0;
                    tc.zq_cstride = 0;
                    tc.lds = &w0[0][sz][lc];
                    tc.zq = nullptr;
                    tc.zplane = zpl0;
                    tc.ctr = ctr;
                    tc.slot = slot;
                    amrex::Array4<const double> sv;
                    sv.mode = amrex::Array4<const double>::kTile;
                    sv.ext = &tc;
                    sv.ncomp = NC;
                    double rhs[NC];
#pragma unroll
                    for (int n = 0; n < NC; ++n) rhs[n] = 0.;
                    amrex::Array4<double> rv;
                    rv.mode = amrex::Array4<double>::kSink;
                    rv.sink = rhs;
                    rv.ncomp = NC;
                    Ops::rhs(rv, sv, gi, gj, gk, a.lev, t1, dt, dx, par1);
                    if constexpr (GW) {
                        Ops::gw_rhs(rv, sv, gi, gj, gk, a.lev, t1, dt, dx, gpar1);
                        Ops::gw_backreaction(rv, sv, gi, gj, gk, a.lev, t1, dt, dx, par1, gpar1);
                    }
                    if constexpr (RKN) {
                        // F0 per pair in the storage type, as the separate first
                        // stage stores it.
                        S f0[NPR];
#pragma unroll
                        for (int p = 0; p < CS::NPAIR; ++p) f0[p] = (S)rhs[CS::pair_mom(p)];
                        auto state2 = [&](int n) -> V {
                            const int p = CS::pair_of_field(n);
                            V y = u0v[n];
                            if (p < 0) return y;
                            y += (V)a.c0 * u0v[CS::pair_mom(p)];
                            y += (V)a.c1 * (V)f0[p];
                            return y;
                        };
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            const V y = state2(CS::comp_of_slot(s));
                            w1[s][s1w][l1[m]] = y;
                            if (m == 0) uq[s][ix(QL - 1)] = y;
                        }
                        if (m == 0) {
#pragma unroll
                            for (int p = 0; p < NP; ++p) du1[ix(QL - 1)][p] = state2(CS::comp_of_point(p));
                            has_f = own && zz >= td.z0 && zz < td.z1;
                            o_f = cown + (unsigned)(zz + g) * ks;
#pragma unroll
                            for (int p = 0; p < CS::NPAIR; ++p) pend_f[p] = f0[p];
                        }
                    } else {
                        // k1 in the storage type, as the two-launch path stores
                        // it; u1 = u0 + dt k1 in the storage type.
                        S k1[NC];
#pragma unroll
                        for (int n = 0; n < NC; ++n) k1[n] = (S)rhs[n];
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            const int n = CS::comp_of_slot(s);
                            const V y = fma_st<V>(cdt, (V)k1[n], u0v[n]);
                            w1[s][s1w][l1[m]] = y;
                            if (m == 0) uq[s][ix(QL - 1)] = y;
                        }
                        if (m == 0) {
#pragma unroll
                            for (int n = 0; n < NC; ++n) dk1[ix(QL - 1)][n] = k1[n];
#pragma unroll
                            for (int p = 0; p < NP; ++p) {
                                const int n = CS::comp_of_point(p);
                                du1[ix(QL - 1)][p] = fma_st<V>(cdt, (V)k1[n], u0v[n]);
                            }
                        }
                    }
                }
            }

            // ---- stage two at plane zc = zz - R - 1 on the tile ----
            const int zc = zz - R - 1;
            const int sc = next(s1w, WS1);  // u1 window slot of plane zc
            s1w = sc;
            if (zc < td.z0) return;
            const int lc = l1[0];
            const int gi = bd.lo[0] + td.x0 + tx, gj = bd.lo[1] + td.y0 + w,
                      gk = bd.lo[2] + zc;
            double ctr[NC];
#pragma unroll
            for (int s = 0; s < NS; ++s) ctr[CS::comp_of_slot(s)] = (double)uq[s][ix(R)];
#pragma unroll
            for (int p = 0; p < NP; ++p) ctr[CS::comp_of_point(p)] = (double)du1[ix(R)][p];
            TileCtx tc;
            tc.ci = gi;
            tc.cj = gj;
            tc.ck = gk;
            tc.R = R;
            tc.f32 = std::is_same_v<V, float> ? 1 : 0;
            tc.lds_pitch = E1X;
            tc.lds_cstride = WS1 * E1;
            tc.zq_cstride = QL;
            tc.zq_phase = PH;
            tc.zq_len = QL;
            tc.lds = &w1[0][sc][lc];
            tc.zq = &uq[0][0];
            tc.zplane = nullptr;
            tc.ctr = ctr;
            tc.slot = slot;
            amrex::Array4<const double> sv;
            sv.mode = amrex::Array4<const double>::kTile;
            sv.ext = &tc;
            sv.ncomp = NC;
            double rhs[NC];
#pragma unroll
            for (int n = 0; n < NC; ++n) rhs[n] = 0.;
            amrex::Array4<double> rv;
            rv.mode = amrex::Array4<double>::kSink;
            rv.sink = rhs;
            rv.ncomp = NC;
            Ops::rhs(rv, sv, gi, gj, gk, a.lev, t2, dt, dx, par2);
            if constexpr (GW) {
                Ops::gw_rhs(rv, sv, gi, gj, gk, a.lev, t2, dt, dx, gpar2);
                Ops::gw_backreaction(rv, sv, gi, gj, gk, a.lev, t2, dt, dx, par2, gpar2);
            }
            const unsigned zso = (unsigned)(zst + g) * ks;
            zst = next(zst, nz);
            has_k = own;
            o_k = cown + zso;
            if constexpr (RKN) {
#pragma unroll
                for (int p = 0; p < CS::NPAIR; ++p) pend_k[p] = (S)rhs[CS::pair_mom(p)];
            } else {
#pragma unroll
                for (int n = 0; n < NC; ++n) pend_k[n] = (S)(rhs[n] + (double)dk1[ix(R)][n]);
            }
        };
        for (int zz = td.z0 - R; zz <= td.z1 + R; zz += QL)
            unroll_phases(std::make_integer_sequence<int, QL>{}, [&](auto phase) {
                if (zz + phase <= td.z1 + R) plane(phase, zz + phase);
            });
        flush();
        __syncthreads();  // the next tile's prologue rewrites the windows
    }
}

}  // namespace shm

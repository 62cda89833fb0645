// Instantiation of the kernel table for one project.
//
// The generated translation units of a project include the project header,
// define SHM_PROJECT_CLASS and include this file:
//   * one TU per (precision, stencil radius) with SHM_TU_PREC / SHM_TU_R set
//     instantiates the fused stage kernels of that pair;
//   * the main TU (neither macro set) instantiates the tagging and probe
//     kernels and registers the project class.
// Build-time selection (defaults in brackets):
//   SHM_R_MASK    bit r set -> stencil radius r instantiated    [0xE: 1,2,3]
//   SHM_GW_MODES  bit 0: without GW tensor, bit 1: with GW      [1]
//   SHM_WITH_KO   Kreiss-Oliger variants (default tile only)    [0]
#pragma once

#include "shm_aux.hpp"
#include "shm_fused.hpp"
#include "shm_stage.hpp"
#include "shm_table.hpp"

#ifndef SHM_R_MASK
#define SHM_R_MASK 0xE
#endif
#ifndef SHM_GW_MODES
#define SHM_GW_MODES 1
#endif
#ifndef SHM_WITH_KO
#define SHM_WITH_KO 0
#endif

namespace shm {

void fill_table_aux(KernelTable &kt);
void fill_table_p0_r1(KernelTable &kt);
void fill_table_p0_r2(KernelTable &kt);
void fill_table_p0_r3(KernelTable &kt);
void fill_table_p1_r1(KernelTable &kt);
void fill_table_p1_r2(KernelTable &kt);
void fill_table_p1_r3(KernelTable &kt);
void fill_table_p2_r1(KernelTable &kt);
void fill_table_p2_r2(KernelTable &kt);
void fill_table_p2_r3(KernelTable &kt);
void fill_table_p3_r1(KernelTable &kt);
void fill_table_p3_r2(KernelTable &kt);
void fill_table_p3_r3(KernelTable &kt);

template <class Ops>
const KernelTable &project_kernel_table() {
    static const KernelTable table = [] {
        KernelTable kt;
        kt.nscalars = Ops::kNScalars;
        kt.nfields = Ops::kNFields;
        kt.ngw = Ops::kNGw;
        kt.built_gw = (SHM_GW_MODES & 2) != 0;
        kt.built_ko = SHM_WITH_KO != 0;
        kt.r_mask = SHM_R_MASK;
#if SHM_R_MASK & 2
        fill_table_p0_r1(kt);
        fill_table_p1_r1(kt);
        fill_table_p2_r1(kt);
        fill_table_p3_r1(kt);
#endif
#if SHM_R_MASK & 4
        fill_table_p0_r2(kt);
        fill_table_p1_r2(kt);
        fill_table_p2_r2(kt);
        fill_table_p3_r2(kt);
#endif
#if SHM_R_MASK & 8
        fill_table_p0_r3(kt);
        fill_table_p1_r3(kt);
        fill_table_p2_r3(kt);
        fill_table_p3_r3(kt);
#endif
        fill_table_aux(kt);
        return kt;
    }();
    return table;
}

template <int K>
struct KindPol;
template <> struct KindPol<kSsprk1> { using T = PolSsprk<kSsStage1>; };
template <> struct KindPol<kSsprk2> { using T = PolSsprk<kSsStage2>; };
template <> struct KindPol<kSsprk3> { using T = PolSsprk<kSsStage3>; };
template <> struct KindPol<kFillRhs> { using T = PolSsprk<kSsFill>; };
template <> struct KindPol<kFillAddRhs> { using T = PolSsprk<kSsFillAdd>; };
template <> struct KindPol<kRkn0> { using T = PolRkn<0, false>; };
template <> struct KindPol<kRkn1> { using T = PolRkn<1, false>; };
template <> struct KindPol<kRkn2> { using T = PolRkn<2, false>; };
template <> struct KindPol<kRkn3> { using T = PolRkn<3, false>; };
template <> struct KindPol<kRknF0> { using T = PolRkn<0, true>; };
template <> struct KindPol<kRknF1> { using T = PolRkn<1, true>; };
template <> struct KindPol<kRknF2> { using T = PolRkn<2, true>; };
template <> struct KindPol<kRknF3> { using T = PolRkn<3, true>; };
template <> struct KindPol<kRknF4> { using T = PolRkn<4, true>; };
template <> struct KindPol<kRkn0Y> { using T = PolRkn0Y; };
template <> struct KindPol<kRkn1Y> { using T = PolRkn1Y; };
template <> struct KindPol<kRkn1YY> { using T = PolRkn1YY; };
template <> struct KindPol<kRknF2Y> { using T = PolRknF2Y; };
template <> struct KindPol<kLeap1> { using T = PolLeapfrog<false>; };
template <> struct KindPol<kLeap2> { using T = PolLeapfrog<true>; };

template <class Ops, int K, Prec P, int R, bool GW, bool KO, int V>
hipError_t launch_stage(const StageArgs &a, int grid, hipStream_t s) {
    constexpr TileVariant v = kVariants[V];
    hipLaunchKernelGGL((stage_kernel<Ops, typename KindPol<K>::T, P, R, GW, KO,
                                     v.nw, v.ry, v.tx>),
                       dim3(grid), dim3(64 * v.nw), 0, s, a);
    return hipGetLastError();
}

/** One group of a split GW stage (CompSet group G). */
template <class Ops, int K, Prec P, int R, int V, int G>
hipError_t launch_stage_grp(const StageArgs &a, int grid, hipStream_t s) {
    constexpr TileVariant v = kVariants[V];
    hipLaunchKernelGGL((stage_kernel<Ops, typename KindPol<K>::T, P, R, true, false,
                                     v.nw, v.ry, v.tx, G>),
                       dim3(grid), dim3(64 * v.nw), 0, s, a);
    return hipGetLastError();
}

template <class Ops, int K, Prec P, int R, int V, int G>
int occupancy_stage_grp() {
    constexpr TileVariant v = kVariants[V];
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb,
            stage_kernel<Ops, typename KindPol<K>::T, P, R, true, false, v.nw, v.ry,
                         v.tx, G>,
            64 * v.nw, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <class Ops, int K, Prec P, int R, bool GW, bool KO, int V>
int occupancy_stage() {
    constexpr TileVariant v = kVariants[V];
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb,
            stage_kernel<Ops, typename KindPol<K>::T, P, R, GW, KO, v.nw, v.ry, v.tx>,
            64 * v.nw, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <class Ops, int K, Prec P, int R, bool GW, bool KO>
void fill_variants(KernelTable &kt) {
    auto put = [&](auto vc) {
        constexpr int V = decltype(vc)::value;
        using TV = typename PrecTraits<P>::V;  // tile values in LDS
        constexpr int TX = kVariants[V].tx;
        constexpr int TY = kVariants[V].nw * 64 / TX * kVariants[V].ry;
        constexpr long lds = 2L * CompSet<Ops, GW, KO>::NS * (TY + 2 * R) *
                             (TX + 2 * R) * (long)sizeof(TV);
        // Keep at least two workgroups per CU (160 KB LDS); dissipation
        // variants only in the default tile shape.
        if constexpr (lds <= 80 * 1024 && (!KO || V == kDefaultVariant)) {
            StageEntry &e = kt.stage[K][(int)P][R][GW ? 1 : 0][KO ? 1 : 0][V];
            e.fn = &launch_stage<Ops, K, P, R, GW, KO, V>;
            e.occupancy = &occupancy_stage<Ops, K, P, R, GW, KO, V>;
            e.threads = 64 * kVariants[V].nw;
            e.ty = TY;
            e.tx = TX;
        }
    };
    put(std::integral_constant<int, 0>{});
    put(std::integral_constant<int, 1>{});
    put(std::integral_constant<int, 2>{});
    put(std::integral_constant<int, 3>{});
    put(std::integral_constant<int, 4>{});
}

/** Split-stage groups of a GW kind: tile shapes 64 x 4, 64 x 8, 32 x 8 and
 *  64 x 8 with two rows per thread (the 64 x 16 tile is left out to bound
 *  the build). */
template <class Ops, int K, Prec P, int R>
void fill_group_variants(KernelTable &kt) {
    auto put = [&](auto vc, auto gc) {
        constexpr int V = decltype(vc)::value;
        constexpr int G = decltype(gc)::value;
        using TV = typename PrecTraits<P>::V;
        constexpr int TX = kVariants[V].tx;
        constexpr int TY = kVariants[V].nw * 64 / TX * kVariants[V].ry;
        constexpr long lds = 2L * CompSet<Ops, true, false, G>::NS * (TY + 2 * R) *
                             (TX + 2 * R) * (long)sizeof(TV);
        if constexpr (lds <= 80 * 1024) {
            StageEntry &e = kt.stage_grp[K][(int)P][R][G][V];
            e.fn = &launch_stage_grp<Ops, K, P, R, V, G>;
            e.occupancy = &occupancy_stage_grp<Ops, K, P, R, V, G>;
            e.threads = 64 * kVariants[V].nw;
            e.ty = TY;
            e.tx = TX;
        }
    };
    auto per_group = [&](auto vc) {
        put(vc, std::integral_constant<int, 0>{});
        put(vc, std::integral_constant<int, 1>{});
        put(vc, std::integral_constant<int, 2>{});
    };
    per_group(std::integral_constant<int, 0>{});
    per_group(std::integral_constant<int, 1>{});
    per_group(std::integral_constant<int, 3>{});
    per_group(std::integral_constant<int, 4>{});
}

/** The kinds a split GW step uses: SSPRK3 and RKN stages (fills, leapfrog
 *  and dissipation run unsplit). */
template <class Ops, Prec P, int R>
void fill_group_kinds(KernelTable &kt) {
    fill_group_variants<Ops, kSsprk1, P, R>(kt);
    fill_group_variants<Ops, kSsprk2, P, R>(kt);
    fill_group_variants<Ops, kSsprk3, P, R>(kt);
    fill_group_variants<Ops, kRkn0, P, R>(kt);
    fill_group_variants<Ops, kRkn1, P, R>(kt);
    fill_group_variants<Ops, kRkn2, P, R>(kt);
    fill_group_variants<Ops, kRkn3, P, R>(kt);
    fill_group_variants<Ops, kRknF0, P, R>(kt);
    fill_group_variants<Ops, kRknF1, P, R>(kt);
    fill_group_variants<Ops, kRknF2, P, R>(kt);
    fill_group_variants<Ops, kRknF3, P, R>(kt);
    fill_group_variants<Ops, kRknF4, P, R>(kt);
    fill_group_variants<Ops, kRkn0Y, P, R>(kt);
    fill_group_variants<Ops, kRkn1Y, P, R>(kt);
    fill_group_variants<Ops, kRkn1YY, P, R>(kt);
    fill_group_variants<Ops, kRknF2Y, P, R>(kt);
}

template <class Ops, Prec P, int R, bool GW, bool KO>
void fill_kinds(KernelTable &kt) {
    fill_variants<Ops, kSsprk1, P, R, GW, KO>(kt);
    fill_variants<Ops, kSsprk2, P, R, GW, KO>(kt);
    fill_variants<Ops, kSsprk3, P, R, GW, KO>(kt);
    fill_variants<Ops, kFillRhs, P, R, GW, KO>(kt);
    fill_variants<Ops, kFillAddRhs, P, R, GW, KO>(kt);
    fill_variants<Ops, kRkn0, P, R, GW, KO>(kt);
    fill_variants<Ops, kRkn1, P, R, GW, KO>(kt);
    fill_variants<Ops, kRkn2, P, R, GW, KO>(kt);
    fill_variants<Ops, kRkn3, P, R, GW, KO>(kt);
    fill_variants<Ops, kRknF0, P, R, GW, KO>(kt);
    fill_variants<Ops, kRknF1, P, R, GW, KO>(kt);
    fill_variants<Ops, kRknF2, P, R, GW, KO>(kt);
    fill_variants<Ops, kRknF3, P, R, GW, KO>(kt);
    fill_variants<Ops, kRknF4, P, R, GW, KO>(kt);
    // Stored RKN stage states (in the tile type).
    fill_variants<Ops, kRkn0Y, P, R, GW, KO>(kt);
    fill_variants<Ops, kRkn1Y, P, R, GW, KO>(kt);
    fill_variants<Ops, kRkn1YY, P, R, GW, KO>(kt);
    fill_variants<Ops, kRknF2Y, P, R, GW, KO>(kt);
    // Leapfrog writes u_new momenta in place while the stage reads them
    // point-wise; with dissipation the momenta are differenced, so that
    // combination is not instantiated.
    if constexpr (!KO) {
        fill_variants<Ops, kLeap1, P, R, GW, KO>(kt);
        fill_variants<Ops, kLeap2, P, R, GW, KO>(kt);
    }
}

template <class Ops, Prec P, int R, bool GW, int V, bool RKN>
hipError_t launch_fused(const FusedArgs &a, int grid, hipStream_t s) {
    constexpr int NW = kVariants[V].nw;
    hipLaunchKernelGGL((fused12_kernel<Ops, P, R, GW, false, NW, RKN>), dim3(grid),
                       dim3(64 * NW), 0, s, a);
    return hipGetLastError();
}

template <class Ops, Prec P, int R, bool GW, int V, bool RKN>
int occupancy_fused() {
    constexpr int NW = kVariants[V].nw;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, fused12_kernel<Ops, P, R, GW, false, NW, RKN>, 64 * NW, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

/** Fused first two stages: SSPRK3 (fused12) and RKN with a first node 0
 *  (fusedrkn), on 64-wide tiles whose windows fit. */
template <class Ops, Prec P, int R, bool GW>
void fill_fused(KernelTable &kt) {
    auto put = [&](auto vc) {
        constexpr int V = decltype(vc)::value;
        if constexpr (kVariants[V].ry == 1 && kVariants[V].tx == 64 &&
                      FusedShape<Ops, P, R, GW, false, kVariants[V].nw>::fits) {
            FusedEntry &e = kt.fused12[(int)P][R][GW ? 1 : 0][V];
            e.fn = &launch_fused<Ops, P, R, GW, V, false>;
            e.occupancy = &occupancy_fused<Ops, P, R, GW, V, false>;
            e.threads = 64 * kVariants[V].nw;
            e.ty = kVariants[V].nw;
            FusedEntry &r = kt.fusedrkn[(int)P][R][GW ? 1 : 0][V];
            r.fn = &launch_fused<Ops, P, R, GW, V, true>;
            r.occupancy = &occupancy_fused<Ops, P, R, GW, V, true>;
            r.threads = e.threads;
            r.ty = e.ty;
        }
    };
    put(std::integral_constant<int, 0>{});
    put(std::integral_constant<int, 1>{});
    put(std::integral_constant<int, 2>{});
    put(std::integral_constant<int, 3>{});
    put(std::integral_constant<int, 4>{});
}

template <class Ops, Prec P, int R>
void fill_radius(KernelTable &kt) {
    if constexpr (SHM_GW_MODES & 1) {
        fill_kinds<Ops, P, R, false, false>(kt);
        fill_fused<Ops, P, R, false>(kt);
        if constexpr (SHM_WITH_KO) fill_kinds<Ops, P, R, false, true>(kt);
    }
    if constexpr (SHM_GW_MODES & 2) {
        fill_kinds<Ops, P, R, true, false>(kt);
        fill_group_kinds<Ops, P, R>(kt);
        if constexpr (SHM_WITH_KO) fill_kinds<Ops, P, R, true, true>(kt);
    }
}

template <class Ops, Prec P, bool GW>
hipError_t launch_te_tag(const TeTagArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((te_tag_kernel<Ops, P, GW>), dim3(grid), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

template <class Ops, Prec P>
hipError_t launch_output(const OutArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((output_kernel<Ops, P>), dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <class Ops>
hipError_t launch_gw_bin(const BinArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((gw_bin_kernel<typename Ops::GwModifiers, 2048>), dim3(grid),
                       dim3(256), 0, s, a);
    return hipGetLastError();
}

template <class Ops>
hipError_t launch_gwdep(unsigned long long *dep, const double *p, const double *pg,
                        double t, double dt, double dx, hipStream_t s) {
    hipLaunchKernelGGL((gwdep_kernel<Ops>), dim3(1), dim3(64), 0, s, dep, p, pg, t, dt, dx);
    return hipGetLastError();
}

template <class Ops, bool GW>
hipError_t launch_probe(ProbeOut *out, const double *p, const double *pg,
                        double t, double dt, double dx, hipStream_t s) {
    hipLaunchKernelGGL((probe_kernel<Ops, GW>), dim3(1), dim3(64), 0, s, out, p,
                       pg, t, dt, dx);
    return hipGetLastError();
}

}  // namespace shm

#ifdef SHM_PROJECT_CLASS
using ShmProjectOps = SHM_PROJECT_CLASS::ShmOps;

#if defined(SHM_TU_PREC) && defined(SHM_TU_R)
namespace shm {
void SHM_PP_CAT(SHM_PP_CAT(fill_table_p, SHM_TU_PREC),
                SHM_PP_CAT(_r, SHM_TU_R))(KernelTable &kt) {
    fill_radius<ShmProjectOps, (Prec)SHM_TU_PREC, SHM_TU_R>(kt);
}
}  // namespace shm
#else
namespace shm {
void fill_table_aux(KernelTable &kt) {
    using Out = typename ShmProjectOps::Output;
    kt.nout = Out::count;
    kt.out_name = &Out::name;
    if constexpr (Out::count > 0) {
        kt.output[0] = &launch_output<ShmProjectOps, Prec::F64>;
        kt.output[1] = &launch_output<ShmProjectOps, Prec::F32C>;
        kt.output[2] = &launch_output<ShmProjectOps, Prec::F32C16>;
    }
    if constexpr (SHM_GW_MODES & 1) {
        kt.te_tag[0][0] = &launch_te_tag<ShmProjectOps, Prec::F64, false>;
        kt.te_tag[1][0] = &launch_te_tag<ShmProjectOps, Prec::F32C, false>;
        kt.te_tag[2][0] = &launch_te_tag<ShmProjectOps, Prec::F32C16, false>;
        kt.probe[0] = &launch_probe<ShmProjectOps, false>;
    }
    if constexpr (SHM_GW_MODES & 2) {
        kt.te_tag[0][1] = &launch_te_tag<ShmProjectOps, Prec::F64, true>;
        kt.te_tag[1][1] = &launch_te_tag<ShmProjectOps, Prec::F32C, true>;
        kt.te_tag[2][1] = &launch_te_tag<ShmProjectOps, Prec::F32C16, true>;
        kt.probe[1] = &launch_probe<ShmProjectOps, true>;
        kt.gwdep = &launch_gwdep<ShmProjectOps>;
        kt.grp_out[0] = CompSet<ShmProjectOps, true, false, 0>::out_mask();
        kt.grp_out[1] = CompSet<ShmProjectOps, true, false, 1>::out_mask();
        kt.grp_out[2] = CompSet<ShmProjectOps, true, false, 2>::out_mask();
        kt.grp_staged[0] = CompSet<ShmProjectOps, true, false, 0>::staged_mask();
        kt.grp_staged[1] = CompSet<ShmProjectOps, true, false, 1>::staged_mask();
        kt.grp_staged[2] = CompSet<ShmProjectOps, true, false, 2>::staged_mask();
        using Mods = typename ShmProjectOps::GwModifiers;
        kt.gw_bin = &launch_gw_bin<ShmProjectOps>;
        kt.ngwmod = Mods::count;
        kt.gwmod_name = &Mods::name;
        kt.gwmod_select = &Mods::select;
    }
}
}  // namespace shm

/** Registers the project class of this library: shm_create() builds it. */
extern "C" __attribute__((visibility("default"))) sledgehamr::Sledgehamr *
shm_project_factory() {
    return new SHM_PROJECT_CLASS();
}
#endif
#endif

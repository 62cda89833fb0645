// Instantiation of the kernel table for one project.  Included (after the
// project header) by the generated per-project translation unit.
//
// Build-time selection (defaults in brackets):
//   SHM_R_MASK  bit r set -> stencil radius r instantiated      [0xE: 1,2,3]
//   SHM_WITH_GW instantiate the 12-component GW variants        [0]
//   SHM_WITH_KO instantiate Kreiss-Oliger variants              [0]
#pragma once

#include "shm_aux.hpp"
#include "shm_stage.hpp"
#include "shm_table.hpp"

#ifndef SHM_R_MASK
#define SHM_R_MASK 0xE
#endif
#ifndef SHM_WITH_GW
#define SHM_WITH_GW 0
#endif
#ifndef SHM_WITH_KO
#define SHM_WITH_KO 0
#endif

namespace shm {

template <class Pol>
struct PolOf;
template <>
struct PolOf<PolSsprk> {
    static constexpr int id = kPolSsprk;
};
template <>
struct PolOf<PolRkn> {
    static constexpr int id = kPolRkn;
};
template <>
struct PolOf<PolLeapfrog> {
    static constexpr int id = kPolLeapfrog;
};

template <class Ops, class Pol, Prec P, int R, bool GW, bool KO, int V>
hipError_t launch_stage(const StageArgs &a, int grid, hipStream_t s) {
    constexpr TileVariant v = kVariants[V];
    hipLaunchKernelGGL((stage_kernel<Ops, Pol, P, R, GW, KO, v.nw, v.ry>),
                       dim3(grid), dim3(64 * v.nw), 0, s, a);
    return hipGetLastError();
}

template <class Ops, class Pol, Prec P, int R, bool GW, bool KO, int V>
int occupancy_stage() {
    constexpr TileVariant v = kVariants[V];
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, stage_kernel<Ops, Pol, P, R, GW, KO, v.nw, v.ry>, 64 * v.nw,
            0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <class Ops, Prec P, bool GW>
hipError_t launch_te_tag(const TeTagArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((te_tag_kernel<Ops, P, GW>), dim3(grid), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

template <class Ops, bool GW>
hipError_t launch_probe(ProbeOut *out, const double *p, const double *pg,
                        double t, double dt, double dx, hipStream_t s) {
    hipLaunchKernelGGL((probe_kernel<Ops, GW>), dim3(1), dim3(64), 0, s, out, p,
                       pg, t, dt, dx);
    return hipGetLastError();
}

template <class Ops, class Pol, Prec P, int R, bool GW, bool KO>
void fill_variants(KernelTable &kt) {
    auto put = [&](auto vc) {
        constexpr int V = decltype(vc)::value;
        using S = typename PrecTraits<P>::S;
        constexpr int TY = kVariants[V].nw * kVariants[V].ry;
        constexpr long lds = 2L * CompSet<Ops, GW, KO>::NS * (TY + 2 * R) *
                             (64 + 2 * R) * (long)sizeof(S);
        // Keep at least two workgroups per CU (160 KB LDS).
        if constexpr (lds <= 80 * 1024) {
            StageEntry &e =
                kt.stage[PolOf<Pol>::id][(int)P][R][GW ? 1 : 0][KO ? 1 : 0][V];
            e.fn = &launch_stage<Ops, Pol, P, R, GW, KO, V>;
            e.occupancy = &occupancy_stage<Ops, Pol, P, R, GW, KO, V>;
            e.threads = 64 * kVariants[V].nw;
            e.ty = kVariants[V].nw * kVariants[V].ry;
        }
    };
    put(std::integral_constant<int, 0>{});
    put(std::integral_constant<int, 1>{});
    put(std::integral_constant<int, 2>{});
}

template <class Ops, Prec P, int R, bool GW, bool KO>
void fill_policies(KernelTable &kt) {
    fill_variants<Ops, PolSsprk, P, R, GW, KO>(kt);
    fill_variants<Ops, PolRkn, P, R, GW, KO>(kt);
    // Leapfrog writes u_new momenta in place while the stage reads them
    // point-wise; with dissipation the momenta are differenced, so that
    // combination is not instantiated.
    if constexpr (!KO) fill_variants<Ops, PolLeapfrog, P, R, GW, KO>(kt);
}

template <class Ops, Prec P, int R>
void fill_radius(KernelTable &kt) {
    if constexpr ((SHM_R_MASK >> R) & 1) {
        fill_policies<Ops, P, R, false, false>(kt);
        if constexpr (SHM_WITH_KO) fill_policies<Ops, P, R, false, true>(kt);
        if constexpr (SHM_WITH_GW) {
            fill_policies<Ops, P, R, true, false>(kt);
            if constexpr (SHM_WITH_KO) fill_policies<Ops, P, R, true, true>(kt);
        }
    }
}

template <class Ops>
const KernelTable &project_kernel_table() {
    static const KernelTable table = [] {
        KernelTable kt;
        kt.nscalars = Ops::kNScalars;
        kt.nfields = Ops::kNFields;
        kt.ngw = Ops::kNGw;
        kt.built_gw = SHM_WITH_GW != 0;
        kt.built_ko = SHM_WITH_KO != 0;
        kt.r_mask = SHM_R_MASK;
        fill_radius<Ops, Prec::F64, 1>(kt);
        fill_radius<Ops, Prec::F64, 2>(kt);
        fill_radius<Ops, Prec::F64, 3>(kt);
        fill_radius<Ops, Prec::F32C, 1>(kt);
        fill_radius<Ops, Prec::F32C, 2>(kt);
        fill_radius<Ops, Prec::F32C, 3>(kt);
        kt.te_tag[0][0] = &launch_te_tag<Ops, Prec::F64, false>;
        kt.te_tag[1][0] = &launch_te_tag<Ops, Prec::F32C, false>;
        kt.probe[0] = &launch_probe<Ops, false>;
        if constexpr (SHM_WITH_GW) {
            kt.te_tag[0][1] = &launch_te_tag<Ops, Prec::F64, true>;
            kt.te_tag[1][1] = &launch_te_tag<Ops, Prec::F32C, true>;
            kt.probe[1] = &launch_probe<Ops, true>;
        }
        return kt;
    }();
    return table;
}

}  // namespace shm

namespace sledgehamr {
class Sledgehamr;
}

/** Registers the project class of this library: shm_create() builds it. */
#define SHM_REGISTER_PROJECT(cls)                                              \
    extern "C" __attribute__((visibility("default")))                         \
    sledgehamr::Sledgehamr *shm_project_factory() {                            \
        return new cls();                                                      \
    }

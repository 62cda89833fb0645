// Host runtime for the MI355X time-step path.  See shm_runtime.hpp.
//
// Reference behaviour followed (file:line in the sledgehamr sources):
//   Integrator::Advance                 integrators/integrator.cpp:11-31
//   IntegratorLsssprk3::Integrate       integrators/lsssprk3.cpp:14-31
//   IntegratorRkn::Integrate, tableaux  integrators/rkn.cpp:25-205
//   IntegratorLeapfrog::Integrate       integrators/leapfrog.cpp:13-57
//   FillPatch / FillIntermediatePatch   level_synchronizer.cpp:94-214
//   ComputeTruncationErrors             level_synchronizer.cpp:240-357
//   CreateShadowLevel                   sledgehamr.cpp:234-251
//   ParseInput / ParseInputScalars      sledgehamr.cpp:389-548
//   per-level dx/dt                     sledgehamr.cpp:27-31
#include "shm_runtime.hpp"

#include <hipfft/hipfft.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <sstream>

#include "shm_aux.hpp"
#include "shm_fused.hpp"

// ---------------------------------------------------------------------------
// ParmParse table.
// ---------------------------------------------------------------------------

namespace {
std::map<std::string, std::vector<std::string>> &pp_table() {
    static std::map<std::string, std::vector<std::string>> t;
    return t;
}
}  // namespace

namespace amrex {

void ParmParse::ParseText(const std::string &text) {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = line.substr(0, eq);
        key.erase(0, key.find_first_not_of(" \t"));
        key.erase(key.find_last_not_of(" \t\r") + 1);
        if (key.empty()) continue;
        std::istringstream vs(line.substr(eq + 1));
        std::vector<std::string> vals;
        std::string tok;
        while (vs >> tok) {
            if (tok.size() >= 2 && tok.front() == '"' && tok.back() == '"')
                tok = tok.substr(1, tok.size() - 2);
            vals.push_back(tok);
        }
        pp_table()[key] = vals;
    }
}

void ParmParse::Clear() { pp_table().clear(); }

bool ParmParse::Contains(const std::string &full_key) {
    return pp_table().count(full_key) != 0;
}

const std::vector<std::string> *ParmParse::find(const char *name) const {
    auto it = pp_table().find(full(name));
    return it == pp_table().end() ? nullptr : &it->second;
}

}  // namespace amrex

namespace shm {

// ---------------------------------------------------------------------------
// Stream probe: plain read / write / copy kernels over large buffers, the
// achievable-bandwidth reference printed beside the HBM peak.
// ---------------------------------------------------------------------------

typedef float sf4 __attribute__((ext_vector_type(4)));

// Bandwidth probes: one float4 per thread and one 4 KB chunk per workgroup,
// dispatched in address order (the pattern that reaches the highest
// sustained rates on this part; grid-stride loops run slower).
__global__ void __launch_bounds__(256) stream_read_kernel(const sf4 *a, long long n,
                                                          float *out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const sf4 v = a[i];
    const float s = v.x + v.y + v.z + v.w;
    if (s == 12345.f) out[threadIdx.x] = s;  // keeps the load alive
}

__global__ void __launch_bounds__(256) stream_write_kernel(sf4 *a, long long n,
                                                           float v) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = sf4{v, v, v, v};
}

__global__ void __launch_bounds__(256) stream_copy_kernel(const sf4 *a, sf4 *b,
                                                          long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}


#define SHM_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            throw Error(2, std::string("HIP error ") + hipGetErrorString(e_) + \
                               " at " #call);                                  \
    } while (0)

// ---------------------------------------------------------------------------
// Device memory and layouts.
// ---------------------------------------------------------------------------

void DevMem::alloc(size_t bytes) {
    reset();
    if (bytes == 0) return;
    SHM_HIP(hipMalloc(&p_, bytes));
    SHM_HIP(hipMemset(p_, 0, bytes));
    bytes_ = bytes;
}

void DevMem::reset() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
}

/** Copies v into d, growing d only when it is too small. */
template <class T>
static void upload_into(DevMem &d, const std::vector<T> &v) {
    const size_t b = std::max<size_t>(1, v.size()) * sizeof(T);
    if (d.bytes() < b) d.alloc(b);
    if (!v.empty())
        SHM_HIP(hipMemcpy(d.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

template <class T>
static void upload_vec(DevMem &d, const std::vector<T> &v) {
    d.alloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty())
        SHM_HIP(hipMemcpy(d.get(), v.data(), v.size() * sizeof(T),
                          hipMemcpyHostToDevice));
}

static long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

void Layout::build(const std::vector<Box3> &b, int ghost, int domain_n,
                   bool is_periodic, bool is_independent) {
    boxes = b;
    g = ghost;
    N = domain_n;
    periodic = is_periodic;
    independent = is_independent;
    hbox.clear();
    vprefix.assign(1, 0);
    cprefix.assign(1, 0);
    long long end = 0;
    for (const Box3 &bx : boxes) {
        for (int d = 0; d < 3; ++d)
            if (bx.n(d) <= 0)
                throw Error(1, "box with non-positive extent");
        BoxDesc bd;
        for (int d = 0; d < 3; ++d) {
            bd.lo[d] = bx.lo[d];
            bd.n[d] = bx.n(d);
        }
        // Rows of valid cells start on a 64-element boundary (one 128-byte
        // line holds a tile row of a bf16 remainder, two of an FP32 array)
        // for boxes at least one tile wide, on 32 elements otherwise.
        const int al = bx.n(0) >= 64 ? 64 : 32;
        bd.jstride = (int)round_up(bx.n(0) + 2 * g, al);
        bd.kstride = bd.jstride * (bx.n(1) + 2 * g);
        const long long start = round_up(end + g, al) - g;
        bd.base = start + (long long)g * bd.kstride + (long long)g * bd.jstride + g;
        bd.org = start;
        bd.wrap = 0;
        bd.pad_ = 0;
        for (int a = 0; a < 3; ++a)
            if (independent || (periodic && bx.n(a) == N)) bd.wrap |= 1 << a;
        end = start + (long long)bd.kstride * (bx.n(2) + 2 * g);
        hbox.push_back(bd);
        vprefix.push_back(vprefix.back() + bx.vol());
        cprefix.push_back(cprefix.back() + (long long)(bx.n(0) / 2) *
                                               (bx.n(1) / 2) * (bx.n(2) / 2));
    }
    cstride = round_up(end + 64, 64);
    upload_vec(dbox, hbox);
    upload_vec(dvprefix, vprefix);
    upload_vec(dcprefix, cprefix);

    // Ghost copy plan: intersect each box's ghost shell with every
    // (periodically shifted) valid box, itself included for shift != 0.
    // Independent boxes are each periodic on themselves.
    halo_full.clear();
    halo_full_total = 0;
    const int smax = (periodic || independent) ? 1 : 0;
    for (int d = 0; d < (int)boxes.size(); ++d) {
        const Box3 &db = boxes[d];
        int glo[3], ghi[3];
        for (int a = 0; a < 3; ++a) {
            glo[a] = db.lo[a] - g;
            ghi[a] = db.hi[a] + g;
        }
        if (g == 0) continue;
        for (int s = 0; s < (int)boxes.size(); ++s) {
            if (independent && s != d) continue;
            const Box3 &sb = boxes[s];
            for (int sz = -smax; sz <= smax; ++sz)
                for (int sy = -smax; sy <= smax; ++sy)
                    for (int sx = -smax; sx <= smax; ++sx) {
                        if (s == d && sx == 0 && sy == 0 && sz == 0) continue;
                        const int per[3] = {independent ? db.n(0) : N,
                                            independent ? db.n(1) : N,
                                            independent ? db.n(2) : N};
                        const int sh[3] = {sx * per[0], sy * per[1], sz * per[2]};
                        HaloSeg seg;
                        bool empty = false;
                        for (int a = 0; a < 3; ++a) {
                            const int lo = std::max(glo[a], sb.lo[a] + sh[a]);
                            const int hi = std::min(ghi[a], sb.hi[a] + sh[a]);
                            if (hi < lo) {
                                empty = true;
                                break;
                            }
                            seg.lo[a] = lo;
                            seg.n[a] = hi - lo + 1;
                            seg.shift[a] = sh[a];
                        }
                        if (empty) continue;
                        seg.dst_box = d;
                        seg.src_box = s;
                        seg.start = halo_full_total;
                        halo_full_total += (long long)seg.n[0] * seg.n[1] * seg.n[2];
                        halo_full.push_back(seg);
                    }
        }
    }
    // Stage kernels wrap boxes that span the domain along all axes, so their
    // ghosts are left out of the per-stage plan.
    halo.clear();
    halo_total = 0;
    for (HaloSeg seg : halo_full) {
        if (hbox[seg.dst_box].wrap == 7) continue;
        seg.start = halo_total;
        halo_total += (long long)seg.n[0] * seg.n[1] * seg.n[2];
        halo.push_back(seg);
    }
    upload_vec(dhalo_full, halo_full);
    upload_vec(dhalo, halo);
    tiles.clear();
}

const TileDesc *Layout::get_tiles(int ty, int tz, int zc_minor, int *ntiles, int tx) {
    auto key = std::make_tuple(ty, tz, zc_minor, tx);
    auto it = tiles.find(key);
    if (it == tiles.end()) {
        std::vector<TileDesc> tv;
        for (int b = 0; b < (int)boxes.size(); ++b) {
            const Box3 &bx = boxes[b];
            const int nzc = (bx.n(2) + tz - 1) / tz;
            const int ny = (bx.n(1) + ty - 1) / ty, nx = (bx.n(0) + tx - 1) / tx;
            auto push = [&](int x, int y, int zc) {
                TileDesc t;
                t.box = b;
                t.x0 = x * tx;
                t.y0 = y * ty;
                t.z0 = zc * tz;
                t.z1 = std::min(t.z0 + tz, bx.n(2));
                tv.push_back(t);
            };
            // x fastest; the z chunk outermost or between y and x.
            if (zc_minor) {
                for (int y = 0; y < ny; ++y)
                    for (int zc = 0; zc < nzc; ++zc)
                        for (int x = 0; x < nx; ++x) push(x, y, zc);
            } else {
                for (int zc = 0; zc < nzc; ++zc)
                    for (int y = 0; y < ny; ++y)
                        for (int x = 0; x < nx; ++x) push(x, y, zc);
            }
        }
        auto mem = std::make_shared<DevMem>();
        upload_vec(*mem, tv);
        it = tiles.emplace(key, std::make_pair(mem, (int)tv.size())).first;
    }
    *ntiles = it->second.second;
    return it->second.first->as<TileDesc>();
}

/** Copy plan from `src` boxes onto `dst` boxes (valid regions, no shift). */
static std::vector<HaloSeg> copy_plan(const Layout &dst, const Layout &src,
                                      long long *total) {
    std::vector<HaloSeg> segs;
    *total = 0;
    for (int d = 0; d < (int)dst.boxes.size(); ++d)
        for (int s = 0; s < (int)src.boxes.size(); ++s) {
            HaloSeg seg;
            bool empty = false;
            for (int a = 0; a < 3; ++a) {
                const int lo = std::max(dst.boxes[d].lo[a], src.boxes[s].lo[a]);
                const int hi = std::min(dst.boxes[d].hi[a], src.boxes[s].hi[a]);
                if (hi < lo) {
                    empty = true;
                    break;
                }
                seg.lo[a] = lo;
                seg.n[a] = hi - lo + 1;
                seg.shift[a] = 0;
            }
            if (empty) continue;
            seg.dst_box = d;
            seg.src_box = s;
            seg.start = *total;
            *total += (long long)seg.n[0] * seg.n[1] * seg.n[2];
            segs.push_back(seg);
        }
    return segs;
}

/** Box-to-box copy of ncomp components of element size es (8, 4 or 2). */
static void launch_copy(const CopyArgs &c, size_t es, int ncu, hipStream_t s) {
    if (c.total == 0) return;
    const int grid = (int)std::min<long long>((c.total + 255) / 256, 8LL * ncu);
    if (es == 8)
        hipLaunchKernelGGL(copy_kernel<double>, dim3(grid), dim3(256), 0, s, c);
    else if (es == 4)
        hipLaunchKernelGGL(copy_kernel<float>, dim3(grid), dim3(256), 0, s, c);
    else
        hipLaunchKernelGGL(copy_kernel<uint16_t>, dim3(grid), dim3(256), 0, s, c);
    SHM_HIP(hipGetLastError());
}

static std::vector<Box3> coarsen(const std::vector<Box3> &b) {
    std::vector<Box3> c;
    for (const Box3 &x : b) {
        Box3 y;
        for (int a = 0; a < 3; ++a) {
            if ((x.lo[a] & 1) || (x.n(a) & 1))
                throw Error(1, "box not coarsenable by 2 (odd corner or size)");
            y.lo[a] = x.lo[a] >> 1;
            y.hi[a] = ((x.hi[a] + 1) >> 1) - 1;
        }
        c.push_back(y);
    }
    return c;
}

// ---------------------------------------------------------------------------
// Device scratch array: the working memory of the allocation-free tagging
// entry point (shm_te_tag_noalloc) and of the level norms.  One array per
// GPU for the whole process, shared by every context of this library on
// that GPU; its device address is looked up once per device.
// ---------------------------------------------------------------------------

static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static_assert(kScratchRecOff + sizeof(double) * kPrmBlock <= kScratchMaskOff,
              "coarse/fine record overlaps the mask region");

static void *scratch_on_current_device() {
    static std::mutex mu;
    static std::map<int, void *> ptrs;
    int dev = 0;
    SHM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) return it->second;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptrs[dev] = p;
    return p;
}

/** Zeroes n 64-bit words (counters that are summed atomically). */
__global__ void __launch_bounds__(64) clear_words_kernel(unsigned long long *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Runtime.
// ---------------------------------------------------------------------------

constexpr int kPrmSlots = 1024;

Runtime::Runtime(sledgehamr::Sledgehamr *sim, const KernelTable &kt)
    : sim_(sim), kt_(kt) {
    int dev = 0;
    SHM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SHM_HIP(hipGetDeviceProperties(&prop, dev));
    ncu_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    SHM_HIP(hipStreamCreate(&stream_));  // blocking: ordered after null-stream memsets
    prm_dev_.alloc(sizeof(double) * kPrmBlock * kPrmSlots);
    SHM_HIP(hipHostMalloc((void **)&prm_host_,
                          sizeof(double) * kPrmBlock * kPrmSlots,
                          hipHostMallocDefault));
    std::memset(prm_host_, 0, sizeof(double) * kPrmBlock * kPrmSlots);
    prm_ev_.assign(kPrmSlots, nullptr);
    SHM_HIP(hipEventCreate(&ev0_));
    SHM_HIP(hipEventCreate(&ev1_));
    scratch_ = static_cast<unsigned char *>(scratch_on_current_device());
    sim_->level_norm0_ = [this](int lev) { return level_norm0(lev); };
}

std::vector<double> Runtime::level_norm0(int lev) {
    LevelState &ls = level(lev);
    std::vector<double> out(ncomp_, 0.);
    const long long total = ls.L.nvalid();
    if (total == 0 || !ls.U[ls.cur]) return out;
    if (ncomp_ > kScratchWords) throw Error(4, "too many components for the level norm");
    // The maxima live in the scratch array: a project's host hooks may call
    // this from the allocation-free tagging.
    unsigned long long *dmax =
        reinterpret_cast<unsigned long long *>(scratch_ + kScratchNormOff);
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, stream_, dmax, ncomp_);
    SHM_HIP(hipGetLastError());
    Norm0Args a{};
    a.boxes = ls.L.dbox.as<BoxDesc>();
    a.vprefix = ls.L.dvprefix.as<long long>();
    a.nbox = (int)ls.L.boxes.size();
    a.total = total;
    a.cstride = ls.L.cstride;
    a.ncomp = ncomp_;
    a.hi = ls.U[ls.cur].get();
    a.lo = comp() ? ls.Ulo[ls.cur].get() : nullptr;
    a.out = dmax;
    const int grid = (int)std::min<long long>((total + 255) / 256, 8LL * ncu_);
    switch (sprec()) {
        case Prec::F64:
            hipLaunchKernelGGL(norm0_kernel<Prec::F64>, dim3(grid), dim3(256), 0, stream_, a);
            break;
        case Prec::F32C:
            hipLaunchKernelGGL(norm0_kernel<Prec::F32C>, dim3(grid), dim3(256), 0, stream_, a);
            break;
        default:
            hipLaunchKernelGGL(norm0_kernel<Prec::F32C16>, dim3(grid), dim3(256), 0, stream_,
                               a);
            break;
    }
    SHM_HIP(hipGetLastError());
    std::vector<unsigned long long> bits(ncomp_);
    SHM_HIP(hipMemcpyAsync(bits.data(), dmax, bits.size() * 8, hipMemcpyDeviceToHost,
                           stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
    for (int c = 0; c < ncomp_; ++c) std::memcpy(&out[c], &bits[c], 8);
    return out;
}

Runtime::~Runtime() {
    sim_->level_norm0_ = nullptr;
    if (stream_) (void)hipStreamSynchronize(stream_);
    release_graphs();
    release_gw_cache();
    release_out_cache();
    for (EventPair &p : ev_pool_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : prm_ev_)
        if (e) (void)hipEventDestroy(e);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (prm_host_) (void)hipHostFree(prm_host_);
    levels_.clear();
    if (stream_) (void)hipStreamDestroy(stream_);
}

LevelState &Runtime::level(int lev) {
    if (lev == -1) {
        define_shadow();
        return shadow_;
    }
    if (!has_level(lev))
        throw Error(1, "level " + std::to_string(lev) + " is not defined");
    return levels_[lev];
}

std::vector<int> Runtime::all_comps() const {
    std::vector<int> c(ncomp_);
    for (int n = 0; n < ncomp_; ++n) c[n] = n;
    return c;
}

std::vector<int> Runtime::stencil_comps() const {
    std::vector<int> c;
    for (int n = 0; n < ncomp_; ++n) {
        const bool s = ko_ || n < kt_.nfields ||
                       (gw_ && n >= kt_.nscalars && n < kt_.nscalars + 6);
        if (s) c.push_back(n);
    }
    return c;
}

void Runtime::init_from_inputs(const std::string &inputs,
                               int precision_override) {
    amrex::ParmParse::Clear();
    amrex::ParmParse::ParseText(inputs);
    amrex::ParmParse pp("");
    sledgehamr::Sledgehamr &S = *sim_;

    std::string pname;
    if (pp.query("project.name", pname) && pname != S.project_name)
        throw Error(1, "inputs select project '" + pname +
                           "' but this library holds '" + S.project_name + "'");

    pp.query("amr.nghost", S.nghost);
    pp.get("amr.coarse_level_grid_size", S.coarse_level_grid_size);
    S.max_level = 0;
    pp.query("amr.max_refinement_levels", S.max_level);
    // The reference requires a power of two (sledgehamr.cpp:438-444).
    // shm.allow_any_grid_size lifts that for any even size (the shadow
    // level halves the coarse level), e.g. a 384^3 level.
    bool any_size = false;
    pp.query("shm.allow_any_grid_size", any_size);
    if (any_size) {
        if (S.coarse_level_grid_size < 2 || (S.coarse_level_grid_size & 1))
            throw Error(1, "amr.coarse_level_grid_size needs to be even");
    } else if (!sledgehamr::utils::IsPowerOfTwo(S.coarse_level_grid_size)) {
        throw Error(1, "amr.coarse_level_grid_size needs to be a power of 2!");
    }
    pp.get("sim.t_start", S.t_start);
    pp.get("sim.t_end", S.t_end);
    pp.get("sim.L", S.L);
    pp.get("sim.cfl", S.cfl);
    pp.query("sim.gravitational_waves", S.with_gravitational_waves);
    pp.query("amr.tagging_on_gpu", S.tagging_on_gpu);
    S.dimN.clear();
    S.dx.clear();
    S.dt.clear();
    for (int lev = 0; lev <= S.max_level; ++lev) {
        S.dimN.push_back(S.coarse_level_grid_size * (1 << lev));
        S.dx.push_back(S.L / (double)S.dimN[lev]);
        S.dt.push_back(S.dx[lev] * S.cfl);
    }

    gw_ = S.with_gravitational_waves;
    if (gw_ && !kt_.built_gw)
        throw Error(4, "gravitational waves requested but this library was "
                       "built without GW kernels");
    ncomp_ = gw_ ? kt_.ngw : kt_.nscalars;
    if ((int)S.scalar_fields.size() != kt_.nscalars)
        throw Error(3, "scalar field registry does not match the kernel table");
    std::vector<std::string> &names = names_;
    names.clear();
    for (auto *f : S.scalar_fields) names.push_back(f->name);
    if (gw_)
        for (const char *n : {"u_xx", "u_yy", "u_zz", "u_xy", "u_xz", "u_yz",
                              "du_xx", "du_yy", "du_zz", "du_xy", "du_xz",
                              "du_yz"})
            names.push_back(n);

    // Truncation-error thresholds and dissipation (ParseInputScalars).
    S.te_crit.assign(ncomp_, DBL_MAX);
    double te_default = DBL_MAX;
    pp.query("amr.te_crit", te_default);
    if (!(te_default > 0)) throw Error(1, "amr.te_crit needs to be > 0!");
    S.shadow_hierarchy = false;
    for (int n = 0; n < ncomp_; ++n) {
        S.te_crit[n] = te_default;
        pp.query(("amr.te_crit_" + names[n]).c_str(), S.te_crit[n]);
        if (!(S.te_crit[n] > 0))
            throw Error(1, "amr.te_crit_" + names[n] + " needs to be > 0!");
        if (S.te_crit[n] != DBL_MAX) S.shadow_hierarchy = true;
    }
    S.dissipation_strength.assign(ncomp_, 0.);
    double ko_default = 0;
    pp.query("sim.dissipation_strength", ko_default);
    S.with_dissipation = false;
    for (int n = 0; n < ncomp_; ++n) {
        S.dissipation_strength[n] = ko_default;
        pp.query(("sim.dissipation_strength_" + names[n]).c_str(),
                 S.dissipation_strength[n]);
        if (S.dissipation_strength[n] < 0)
            throw Error(1, "dissipation strength needs to be >= 0!");
        if (S.dissipation_strength[n] > 0) S.with_dissipation = true;
    }
    ko_ = S.with_dissipation;
    // Coarse/fine interpolation (LevelSynchronizer, level_synchronizer.cpp:27-47).
    pp.query("amr.interpolation_type", interp_type_);
    if (interp_type_ != 0 && interp_type_ != 1 && interp_type_ != 2 && interp_type_ != 4)
        throw Error(1, "amr.interpolation_type must be 0, 1, 2 or 4");
    if (ko_) {
        S.dissipation_order = S.nghost;
        pp.query("sim.dissipation_order", S.dissipation_order);
        if (S.dissipation_order != 2 && S.dissipation_order != 3)
            throw Error(1, "Currently only sim.dissipation_order = 2 or 3 supported!");
        if (!kt_.built_ko)
            throw Error(4, "dissipation requested but this library was built "
                           "without Kreiss-Oliger kernels");
    }
    ko_dev_.alloc(sizeof(double) * std::max(1, ncomp_));
    SHM_HIP(hipMemcpy(ko_dev_.get(), S.dissipation_strength.data(),
                      sizeof(double) * ncomp_, hipMemcpyHostToDevice));

    // Integrator (TimeStepper::SetIntegrator, rkn.cpp:188-271).
    pp.get("integrator.type", integrator_);
    switch (integrator_) {
        case 10:
        case 11:
            break;
        case 20: {
            amrex::ParmParse pr("integrator.rkn");
            pr.getarr("weights_bar_b", bbar_);
            pr.getarr("weights_b", b_);
            pr.getarr("nodes", nodes_);
            std::vector<double> bt;
            pr.getarr("tableau", bt);
            const int s = (int)nodes_.size();
            if ((int)bbar_.size() != s || (int)b_.size() != s)
                throw Error(1, "integrator.weights should be the same length "
                               "as integrator.nodes");
            if ((int)bt.size() != s * (s + 1) / 2)
                throw Error(1, "integrator.tableau incorrect length - should "
                               "include the Butcher Tableau diagonal.");
            tableau_.clear();
            int k = 0;
            for (int i = 0; i < s; ++i) {
                std::vector<double> row;
                for (int j = 0; j <= i; ++j) row.push_back(bt[k++]);
                if (row.back() != 0.0)
                    throw Error(1, "RKN integrator currently only supports "
                                   "explicit Butcher tableaus.");
                tableau_.push_back(row);
            }
            break;
        }
        case 21:
            nodes_ = {0.0, 0.5, 1.0};
            tableau_ = {{0.0}, {1. / 8., 0.0}, {0.0, 0.5, 0.0}};
            bbar_ = {1. / 6., 1. / 3., 0.0};
            b_ = {1. / 6., 4. / 6., 1. / 6.};
            break;
        case 22:
            nodes_ = {0.0, 1. / 5., 2. / 3., 1.0};
            tableau_ = {{0.0},
                        {1. / 50., 0.0},
                        {-1. / 27., 7. / 27., 0.0},
                        {3. / 10., -2. / 35., 9. / 35., 0.0}};
            bbar_ = {14. / 336., 100. / 336., 54. / 336., 0.0};
            b_ = {14. / 336., 125. / 336., 162. / 336., 35. / 336.};
            break;
        default:
            throw Error(4, "#error: Unknown integration type: " +
                               std::to_string(integrator_));
    }
    if (integrator_ >= 20 && (nodes_.empty() || nodes_.size() > 5))
        throw Error(4, "RKN tableaus with 1..5 stages are supported");
    if ((integrator_ >= 20 || integrator_ == 11) &&
        2 * kt_.nfields != kt_.nscalars)
        throw Error(4, "RKN/leapfrog need one conjugate momentum per field");
    if (integrator_ == 11 && ko_)
        throw Error(4, "leapfrog with Kreiss-Oliger dissipation is not supported");

    std::string prec;
    prec_ = Prec::F32C16;
    if (pp.query("shm.precision", prec)) {
        if (prec == "fp64" || prec == "FP64")
            prec_ = Prec::F64;
        else if (prec == "fp32c" || prec == "FP32C")
            prec_ = Prec::F32C;
        else if (prec == "fp32c16" || prec == "FP32C16")
            prec_ = Prec::F32C16;
        else if (prec == "fp32cs" || prec == "FP32CS")
            prec_ = Prec::F32CS;
        else
            throw Error(1, "shm.precision must be fp64, fp32c, fp32c16 or fp32cs");
    }
    if (precision_override == 0) prec_ = Prec::F64;
    if (precision_override == 1) prec_ = Prec::F32C;
    if (precision_override == 2) prec_ = Prec::F32C16;
    if (precision_override == 3) prec_ = Prec::F32CS;
    pp.query("shm.tile_variant", launch.variant);
    pp.query("shm.tile_z", launch.tz);
    pp.query("shm.xcd_split", launch.xcd_split);
    pp.query("shm.persistent", launch.persistent);
    // Ensemble of independent realizations: every level-0 box is its own
    // periodic domain (boxes may share index ranges).
    pp.query("shm.independent_boxes", independent_);
    if (independent_ && S.max_level > 0)
        throw Error(1, "shm.independent_boxes needs a single level");
    pp.query("shm.tile_order_zc_minor", launch.zc_minor);
    pp.query("shm.graphs", graphs);
    pp.query("shm.fused", launch.fused);
    pp.query("shm.rkn_state", rkn_state_);
    int gw_split = 1;
    pp.query("shm.gw_split", gw_split);
    if (launch.variant >= kNVar) throw Error(1, "shm.tile_variant out of range");

    // Project initialisation, then the reach of its cell functions.
    sim_->Init();
    probe();
    gw_groups_ = gw_split != 0 && check_gw_groups();
    if (R_ > S.nghost)
        throw Error(1, "cell functions reach " + std::to_string(R_) +
                           " cells but amr.nghost = " + std::to_string(S.nghost));

    // Level 0: the periodic domain, optionally chopped (amr.max_grid_size).
    const int N0 = S.coarse_level_grid_size;
    int mgs = N0;
    pp.query("amr.max_grid_size", mgs);
    if (mgs <= 0 || N0 % mgs != 0) throw Error(1, "amr.max_grid_size must divide the grid");
    std::vector<Box3> boxes;
    for (int z = 0; z < N0; z += mgs)
        for (int y = 0; y < N0; y += mgs)
            for (int x = 0; x < N0; x += mgs)
                boxes.push_back(Box3{{x, y, z}, {x + mgs - 1, y + mgs - 1, z + mgs - 1}});
    levels_.clear();
    levels_.resize(S.max_level + 1);
    set_level(0, boxes);
    S.finest_level = 0;
    tuned_.assign(S.max_level + 2, LaunchCfg());
    // Local regrid (local_regrid.cpp:449-457, local_regrid.h:93-122); the
    // blocking factor per level, the last value repeated.
    {
        // Checked here when given, else where the local regrid uses it.
        std::vector<int> b{8};
        const bool given = pp.queryarr("amr.blocking_factor", b);
        if (b.empty()) throw Error(1, "amr.blocking_factor needs a value");
        bf_.clear();
        for (int lev = 0; lev <= S.max_level; ++lev) {
            bf_.push_back(b[std::min<size_t>(lev, b.size() - 1)]);
            if (given) check_bf(lev);
        }
    }
    pp.query("amr.n_error_buf", n_error_buf_);
    pp.query("amr.max_local_regrids", max_local_regrids_);
    pp.query("amr.volume_threshold_strong", vt_strong_);
    pp.query("amr.volume_threshold_weak", vt_weak_);
    if (n_error_buf_ < 0) throw Error(1, "amr.n_error_buf must be >= 0");
    {
        int on = 0;
        pp.query("shm.local_regrid", on);
        local_regrid_on_ = on != 0;
        int keep = 0;
        pp.query("shm.keep_layout_on_global_regrid", keep);
        keep_layout_on_global_ = keep != 0;
    }
    lr_ = RegridState();
    lr_.no_local.assign(S.max_level + 1, 0);
    lr_.do_global.assign(S.max_level + 1, 0);
    regrid_dt_.clear();
    last_regrid_.clear();
    double reg_dt = DBL_MAX;
    pp.query("amr.regrid_dt", reg_dt);
    for (int lev = 0; lev <= S.max_level; ++lev) {
        regrid_dt_.push_back(reg_dt / std::pow(2., lev));
        last_regrid_.push_back(S.t_start);
    }
}

/** Whether the stages of a GW level may run as three component groups
 *  (CompSet<..., G>): every output must read only the components its group
 *  stages, by the dependency probe of the project's cell functions. */
bool Runtime::check_gw_groups() {
    if (!gw_ || ko_ || !kt_.gwdep) return false;
    DevMem dep;
    dep.alloc(sizeof(unsigned long long) * 64);
    std::vector<double> blk(kPrmBlock, 0.);
    fill_params(blk.data(), sim_->t_start, sim_->dt[0], sim_->dx[0], 0);
    DevMem d;
    upload_vec(d, blk);
    SHM_HIP(kt_.gwdep(dep.as<unsigned long long>(), d.as<double>() + kPrmUser,
                      d.as<double>() + kPrmGw, sim_->t_start, sim_->dt[0], sim_->dx[0],
                      stream_));
    std::vector<unsigned long long> h(64, 0);
    SHM_HIP(hipMemcpyAsync(h.data(), dep.get(), sizeof(unsigned long long) * 64,
                           hipMemcpyDeviceToHost, stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
    gw_dep_.assign(h.begin(), h.begin() + ncomp_);
    for (int g = 0; g < 3; ++g)
        for (int n = 0; n < ncomp_; ++n)
            if (((kt_.grp_out[g] >> n) & 1) && (h[n] & ~kt_.grp_staged[g])) return false;
    return true;
}

void Runtime::probe() {
    const int gwi = gw_ ? 1 : 0;
    if (!kt_.probe[gwi]) throw Error(4, "probe kernel missing");
    DevMem out;
    out.alloc(sizeof(ProbeOut));
    std::vector<double> blk(kPrmBlock, 0.);
    fill_params(blk.data(), sim_->t_start, sim_->dt[0], sim_->dx[0], 0);
    DevMem d;
    upload_vec(d, blk);
    SHM_HIP(kt_.probe[gwi](out.as<ProbeOut>(), d.as<double>() + kPrmUser,
                           d.as<double>() + kPrmGw, sim_->t_start, sim_->dt[0],
                           sim_->dx[0], stream_));
    ProbeOut po;
    SHM_HIP(hipMemcpyAsync(&po, out.get(), sizeof(po), hipMemcpyDeviceToHost,
                           stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
    int R = 1;
    const std::vector<int> sc = stencil_comps();
    for (int n = 0; n < ncomp_; ++n) {
        const bool is_st = std::find(sc.begin(), sc.end(), n) != sc.end();
        if (po.offaxis[n])
            throw Error(4, "cell functions read component " + std::to_string(n) +
                               " at a point off the z axis and off the xy "
                               "plane; the tile view supports star stencils "
                               "and in-plane diagonals");
        if (!is_st && po.reach[n] > 0)
            throw Error(4, "cell functions difference component " +
                               std::to_string(n) +
                               " which is stored point-wise (momenta are "
                               "read at the centre only)");
        R = std::max(R, po.reach[n]);
    }
    if (ko_) R = std::max(R, sim_->dissipation_order);
    if (R > 3) throw Error(4, "stencil radius above 3 is not supported");
    if (!((kt_.r_mask >> R) & 1))
        throw Error(4, "stencil radius " + std::to_string(R) +
                           " not built into this library");
    R_ = R;
}

void Runtime::set_level(int lev, const std::vector<Box3> &boxes) {
    bump_epoch();
    if (lev < 0 || lev > sim_->max_level) throw Error(1, "level out of range");
    if ((int)levels_.size() <= lev) levels_.resize(lev + 1);
    LevelState &ls = levels_[lev];
    ls = LevelState();
    const int N = sim_->dimN[lev];
    for (const Box3 &b : boxes)
        for (int a = 0; a < 3; ++a)
            if (b.lo[a] < 0 || b.hi[a] >= N || b.hi[a] < b.lo[a])
                throw Error(1, "box outside the level domain");
    const bool indep = independent_ && lev == 0;
    ls.L.build(boxes, sim_->nghost, N, /*periodic=*/!indep, indep);
    ls.ncomp = ncomp_;
    ensure_arrays(ls, integrator_);
    ls.t = sim_->t_start;
    ls.t_old = -DBL_MAX;
    ls.istep = 0;
    if (lev + 1 < (int)levels_.size()) levels_[lev + 1].cf_built = false;
    if (lev > sim_->finest_level) sim_->finest_level = lev;
}

void Runtime::ensure_arrays(LevelState &ls, int integrator) {
    const size_t plane = (size_t)ls.L.cstride;
    const size_t nc = ls.ncomp;
    const size_t es = esize();
    const int npair = kt_.nfields + (gw_ ? 6 : 0);
    for (int b = 0; b < 2; ++b) {
        if (!ls.U[b]) ls.U[b].alloc(plane * nc * es);
        if (comp() && !ls.Ulo[b]) ls.Ulo[b].alloc(plane * nc * losize());
    }
    if (integrator == 10) {
        if (!ls.K[0]) ls.K[0].alloc(plane * nc * es);
        if (!ls.K[1]) ls.K[1].alloc(plane * nc * es);
    } else if (integrator >= 20) {
        for (size_t i = 0; i + 1 < nodes_.size(); ++i)
            if (!ls.F[i]) ls.F[i].alloc(plane * npair * es);
        // Stored stage states are in the tile type (FP64 for FP64 tiles).
        if (nodes_.size() >= 3 && nodes_[0] == 0. && !ls.Y2)
            ls.Y2.alloc(plane * npair * ysize());
        if (gw_ && nodes_.size() == 3 && nodes_[0] == 0. && !ls.Y3)
            ls.Y3.alloc(plane * npair * ysize());
    } else if (integrator == 11) {
        if (!ls.VH) ls.VH.alloc(plane * npair * 8);
    }
}

void Runtime::fill_params(double *blk, double time, double dt, double dx,
                          int lev) {
    std::fill(blk, blk + kPrmBlock, 0.);
    blk[kPrmTime] = time;
    blk[kPrmDt] = dt;
    blk[kPrmDx] = dx;
    std::vector<double> p;
    sim_->SetParamsRhs(p, time, lev);
    if ((int)p.size() > kMaxParams) throw Error(4, "too many Rhs parameters");
    std::copy(p.begin(), p.end(), blk + kPrmUser);
    if (gw_) {
        std::vector<double> pg;
        sim_->SetParamsGravitationalWaveRhs(pg, time, lev);
        if ((int)pg.size() > kMaxParams) throw Error(4, "too many GW parameters");
        std::copy(pg.begin(), pg.end(), blk + kPrmGw);
    }
}

// ---------------------------------------------------------------------------
// Launch helpers.
// ---------------------------------------------------------------------------

LaunchCfg Runtime::resolved_launch(int lev) const {
    const LaunchCfg t = lev + 1 >= 0 && lev + 1 < (int)tuned_.size() ? tuned_[lev + 1]
                                                                     : LaunchCfg();
    auto pick = [](int user, int tuned, int dflt) {
        return user >= 0 ? user : (tuned >= 0 ? tuned : dflt);
    };
    LaunchCfg c;
    c.variant = pick(launch.variant, t.variant, kDefaultVariant);
    c.tz = launch.tz > 0 ? launch.tz : t.tz;
    c.xcd_split = pick(launch.xcd_split, t.xcd_split, 1);
    c.persistent = pick(launch.persistent, t.persistent, 1);
    c.zc_minor = pick(launch.zc_minor, t.zc_minor, 0);
    c.fused = pick(launch.fused, t.fused, 0);
    int dv = -1;
    const FusedEntry *row = fused_row();
    for (int v = kNVar - 1; v >= 0 && dv < 0; --v)
        if (row[v].fn) dv = v;
    c.fused_variant = pick(launch.fused_variant, t.fused_variant, dv);
    return c;
}

bool Runtime::rkn_y(int lev) const {
    // Stored stage states (in the tile type, ysize()); not across
    // coarse/fine ghost cells, whose fills write the composed parts.
    return rkn_state_ != 0 && integrator_ >= 20 && nodes_.size() >= 3 && nodes_[0] == 0. &&
           !has_cf(lev);
}

const FusedEntry *Runtime::fused_row() const {
    return integrator_ >= 20 ? kt_.fusedrkn[(int)prec_][R_][gw_ ? 1 : 0]
                             : kt_.fused12[(int)prec_][R_][gw_ ? 1 : 0];
}

/** The first two stages run in one pass: SSPRK3, or RKN whose first node is
 *  0 and which has a third stage, without dissipation, on a level of boxes
 *  that wrap on themselves. */
bool Runtime::fused_possible(const LevelState &ls) const {
    if (ko_) return false;
    if (integrator_ >= 20) {
        if (nodes_.size() < 3 || nodes_[0] != 0.) return false;
    } else if (integrator_ != 10) {
        return false;
    }
    const FusedEntry *row = fused_row();
    bool any = false;
    for (int v = 0; v < kNVar; ++v) any = any || row[v].fn != nullptr;
    if (!any) return false;
    for (const BoxDesc &b : ls.L.hbox)
        if (b.wrap != 7) return false;
    return !ls.L.hbox.empty();
}

void Runtime::launch_fused(LevelState &ls, const double *prm1, const double *prm2,
                           double dt, int lev) {
    const LaunchCfg lc = resolved_launch(lev);
    int v = lc.fused_variant;
    const FusedEntry *row = fused_row();
    if (v < 0 || v >= kNVar || !row[v].fn) {
        v = -1;
        for (int d = kNVar - 1; d >= 0 && v < 0; --d)
            if (row[d].fn) v = d;
    }
    if (v < 0) throw Error(4, "fused stage kernel not built for this configuration");
    const FusedEntry &e = row[v];
    int maxn = 1;
    for (const Box3 &b : ls.L.boxes) maxn = std::max(maxn, b.n(2));
    // Whole z columns: every chunk boundary costs 2R extra stage-one planes.
    const int z = lc.tz > 0 ? lc.tz : maxn;
    int ntiles = 0;
    FusedArgs a{};
    a.tiles = ls.L.get_tiles(e.ty, z, 0, &ntiles);
    a.boxes = ls.L.dbox.as<BoxDesc>();
    a.ntiles = ntiles;
    a.lev = lev;
    a.prm1 = prm1;
    a.prm2 = prm2;
    a.ghost = ls.L.g;
    a.cstride = ls.L.cstride;
    a.u0 = ls.U[ls.cur ^ 1].get();
    a.u0lo = tile_lo_of(ls, ls.cur ^ 1);
    a.kout = ls.K[1].get();
    a.dt = dt;
    if (integrator_ >= 20) {
        // RKN stages one and two: F0, F1 and the second stage's state.
        a.f0out = ls.F[0].get();
        a.kout = ls.F[1].get();
        a.c0 = dt * nodes_[1];
        a.c1 = dt * dt * tableau_[1][0];
    }
    if (ntiles == 0) return;
    const int occ = e.occupancy();
    int grid = lc.persistent ? (int)std::min<long long>(ntiles, (long long)occ * ncu_)
                             : ntiles;
    a.xcd_split = lc.xcd_split ? 1 : 0;
    if (a.xcd_split) grid = (int)round_up(grid, 8);
    EventPair *ep = timing_on ? timed_pair(2) : nullptr;
    if (ep) SHM_HIP(hipEventRecord(ep->a, stream_));
    SHM_HIP(e.fn(a, grid, stream_));
    if (ep) {
        SHM_HIP(hipEventRecord(ep->b, stream_));
        timing.stage_cells += 2 * ls.L.nvalid();  // two stages
    }
}

void Runtime::launch_stage(int kind, LevelState &ls, StageArgs &a, int lev) {
    const int pi = (int)prec_;
    const LaunchCfg lc = resolved_launch(lev);
    int v = lc.variant;
    const auto &row = kt_.stage[kind][pi][R_][gw_ ? 1 : 0][ko_ ? 1 : 0];
    // Variants over the LDS budget (and non-default dissipation variants)
    // are not instantiated: take the nearest built tile shape.
    if (!row[v].fn) {
        int best = -1;
        for (int d = 1; d < kNVar && best < 0; ++d) {
            if (v - d >= 0 && row[v - d].fn) best = v - d;
            else if (v + d < kNVar && row[v + d].fn) best = v + d;
        }
        if (best >= 0) v = best;
    }
    const StageEntry &e = row[v];
    if (!e.fn) throw Error(4, "stage kernel not built for this configuration");
    // GW levels: three component-group launches when the split applies and
    // the groups are built for this kind and tile shape.
    const StageEntry *grp[3] = {nullptr, nullptr, nullptr};
    if (gw_groups_) {
        const auto &gr = kt_.stage_grp[kind][pi][R_];
        if (gr[0][v].fn && gr[1][v].fn && gr[2][v].fn)
            for (int g = 0; g < 3; ++g) grp[g] = &gr[g][v];
    }
    int z = lc.tz;
    if (z <= 0) {
        // Enough tiles to fill every CU several times, long z runs otherwise.
        int maxn = 1;
        long long xy = 0;
        for (const Box3 &b : ls.L.boxes) {
            maxn = std::max(maxn, b.n(2));
            xy += (long long)((b.n(0) + e.tx - 1) / e.tx) * ((b.n(1) + e.ty - 1) / e.ty);
        }
        z = maxn;
        while (z > 16 && xy * ((maxn + z - 1) / z) < 8LL * ncu_) z = (z + 1) / 2;
    }
    int ntiles = 0;
    a.tiles = ls.L.get_tiles(e.ty, z, lc.zc_minor, &ntiles, e.tx);
    a.boxes = ls.L.dbox.as<BoxDesc>();
    a.ntiles = ntiles;
    a.cstride = ls.L.cstride;
    a.lev = lev;
    a.ko_strength = ko_dev_.as<double>();
    a.ko_order = sim_->dissipation_order;
    a.ghost = ls.L.g;
    if (ntiles == 0) return;
    a.xcd_split = lc.xcd_split ? 1 : 0;
    auto grid_of = [&](const StageEntry &x) {
        const int occ = x.occupancy();
        int grid = lc.persistent ? (int)std::min<long long>(ntiles, (long long)occ * ncu_)
                                 : ntiles;
        if (a.xcd_split) grid = (int)round_up(grid, 8);
        return grid;
    };
    EventPair *ep = timing_on ? timed_pair(0) : nullptr;
    if (ep) SHM_HIP(hipEventRecord(ep->a, stream_));
    if (grp[0]) {
        for (int g = 0; g < 3; ++g) SHM_HIP(grp[g]->fn(a, grid_of(*grp[g]), stream_));
    } else {
        SHM_HIP(e.fn(a, grid_of(e), stream_));
    }
    if (ep) {
        SHM_HIP(hipEventRecord(ep->b, stream_));
        timing.stage_cells += ls.L.nvalid();
    }
}

void Runtime::halo(LevelState &ls, void *arr, const std::vector<int> &comps,
                   bool full, int elem_bytes) {
    const int eb = elem_bytes > 0 ? elem_bytes : (int)esize();
    const long long total = full ? ls.L.halo_full_total : ls.L.halo_total;
    if (total == 0 || comps.empty()) return;
    HaloArgs h;
    h.arr = arr;
    h.cstride = ls.L.cstride;
    h.boxes = ls.L.dbox.as<BoxDesc>();
    h.segs = (full ? ls.L.dhalo_full : ls.L.dhalo).as<HaloSeg>();
    h.nseg = (int)(full ? ls.L.halo_full : ls.L.halo).size();
    h.total = total;
    for (size_t i = 0; i < comps.size(); i += 32) {
        h.ncomp = (int)std::min<size_t>(32, comps.size() - i);
        for (int c = 0; c < h.ncomp; ++c) h.comps[c] = comps[i + c];
        const int grid = (int)std::min<long long>((h.total + 255) / 256, 8LL * ncu_);
        EventPair *ep = timing_on ? timed_pair(1) : nullptr;
        if (ep) SHM_HIP(hipEventRecord(ep->a, stream_));
        if (eb == 8)
            hipLaunchKernelGGL(halo_kernel<double>, dim3(grid), dim3(256), 0,
                               stream_, h);
        else if (eb == 4)
            hipLaunchKernelGGL(halo_kernel<float>, dim3(grid), dim3(256), 0,
                               stream_, h);
        else
            hipLaunchKernelGGL(halo_kernel<uint16_t>, dim3(grid), dim3(256), 0,
                               stream_, h);
        SHM_HIP(hipGetLastError());
        if (ep) SHM_HIP(hipEventRecord(ep->b, stream_));
    }
}

void Runtime::halo_state(LevelState &ls, int which, const std::vector<int> &comps) {
    halo(ls, ls.U[which].get(), comps);
    if (tile_lo()) halo(ls, ls.Ulo[which].get(), comps, false, (int)losize());
}

void Runtime::fill_patch(LevelState &ls, int which, int lev, double time, bool for_step) {
    if (has_cf(lev)) {
        // The coarser level may have changed (its step, average-down).
        if (for_step) {
            // A step reads the leading words of step_ghost_comps() at ghost
            // cells; the remainders and other components stay as they are.
            const std::vector<int> comps = step_ghost_comps();
            cf_fill_state(lev, time, ls.U[which].get(), tile_lo_of(ls, which), comps, true);
            halo_state(ls, which, comps);
            ls.ghosts_ok[which] = false;
            return;
        }
        cf_fill_state(lev, time, ls.U[which].get(),
                      comp() ? ls.Ulo[which].get() : nullptr);
    } else if (ls.ghosts_ok[which]) {
        return;
    }
    halo_state(ls, which, all_comps());
    ls.ghosts_ok[which] = true;
}

void Runtime::cf_fill_ssprk(LevelState &ls, int lev, int old, double t0, double dt) {
    CfArgs a = cf_args(lev, t0, false, true);
    std::vector<double> rec(kPrmBlock, 0.);
    const double tt[2] = {t0 + dt, t0 + dt / 2.};
    for (int q = 0; q < 2; ++q) {
        cf_record(rec.data(), lev, tt[q], false);
        a.rec3[q] = push_record(rec, 1, tt[q]);
    }
    a.dst = ls.U[old].get();
    a.dstlo = tile_lo_of(ls, old);
    a.kd[0] = ls.K[0].get();
    a.kd[1] = ls.K[1].get();
    a.cf[0] = dt;
    a.cf[1] = dt / 4.;
    const std::vector<int> sc = stencil_comps();
    for (size_t c0 = 0; c0 < sc.size(); c0 += 32) {
        a.ncomp = (int)std::min<size_t>(32, sc.size() - c0);
        for (int q = 0; q < a.ncomp; ++q) a.comp[q] = sc[c0 + q];
        cf_launch(a, kCfSsprk);
    }
}

std::vector<int> Runtime::step_ghost_comps() const {
    if (ko_ || integrator_ >= 20) return all_comps();
    return stencil_comps();
}

// ---------------------------------------------------------------------------
// Coarse/fine ghost fill (FillPatchTwoLevels, level_synchronizer.cpp:92-209).
// ---------------------------------------------------------------------------

static int floordiv2(int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

/** Ghost-shell slabs of every box of level lev, and per box the boxes of
 *  level lev-1 its coarse stencil may touch.  Every coarse cell that stencil
 *  can reach must lie in a box of level lev-1 (proper nesting), else the
 *  ghost cells would have no coarse data. */
void Runtime::build_cf(int lev) {
    LevelState &ls = levels_[lev];
    const LevelState &cs = levels_[lev - 1];
    const int g = ls.L.g, Nc = cs.L.N;
    const int s = interp_type_ == 0 ? 0 : interp_type_ == 4 ? 2 : 1;  // stencil reach
    ls.cf_segs.clear();
    ls.cf_total = 0;
    ls.cf_full.clear();
    ls.cf_full_total = 0;
    ls.cf_step.clear();
    ls.cf_step_total = 0;
    const int gs = std::min(g, R_);  // ghost layers a stencil reads
    std::vector<int> start(1, 0), cand;
    // The parent cells of fine cells [flo, flo + fn) of box b.
    auto parents = [](std::vector<CfPSeg> &v, long long &tot, int b, const int *flo,
                      const int *fn) {
        CfPSeg p{};
        p.dst_box = b;
        long long n = 1;
        for (int a = 0; a < 3; ++a) {
            p.flo[a] = flo[a];
            p.fn[a] = fn[a];
            p.plo[a] = floordiv2(flo[a]);
            p.pn[a] = floordiv2(flo[a] + fn[a] - 1) - p.plo[a] + 1;
            n *= p.pn[a];
        }
        p.start = tot;
        tot += n;
        v.push_back(p);
    };
    for (int b = 0; b < (int)ls.L.boxes.size(); ++b) {
        const Box3 &bx = ls.L.boxes[b];
        const int nx = bx.n(0), ny = bx.n(1), nz = bx.n(2);
        const int slab[6][6] = {
            {bx.lo[0] - g, bx.lo[1] - g, bx.lo[2] - g, nx + 2 * g, ny + 2 * g, g},
            {bx.lo[0] - g, bx.lo[1] - g, bx.hi[2] + 1, nx + 2 * g, ny + 2 * g, g},
            {bx.lo[0] - g, bx.lo[1] - g, bx.lo[2], nx + 2 * g, g, nz},
            {bx.lo[0] - g, bx.hi[1] + 1, bx.lo[2], nx + 2 * g, g, nz},
            {bx.lo[0] - g, bx.lo[1], bx.lo[2], g, ny, nz},
            {bx.hi[0] + 1, bx.lo[1], bx.lo[2], g, ny, nz}};
        for (int q = 0; g > 0 && q < 6; ++q)
            parents(ls.cf_segs, ls.cf_total, b, slab[q], slab[q] + 3);
        const int sslab[6][6] = {
            {bx.lo[0] - gs, bx.lo[1] - gs, bx.lo[2] - gs, nx + 2 * gs, ny + 2 * gs, gs},
            {bx.lo[0] - gs, bx.lo[1] - gs, bx.hi[2] + 1, nx + 2 * gs, ny + 2 * gs, gs},
            {bx.lo[0] - gs, bx.lo[1] - gs, bx.lo[2], nx + 2 * gs, gs, nz},
            {bx.lo[0] - gs, bx.hi[1] + 1, bx.lo[2], nx + 2 * gs, gs, nz},
            {bx.lo[0] - gs, bx.lo[1], bx.lo[2], gs, ny, nz},
            {bx.hi[0] + 1, bx.lo[1], bx.lo[2], gs, ny, nz}};
        for (int q = 0; gs > 0 && q < 6; ++q)
            parents(ls.cf_step, ls.cf_step_total, b, sslab[q], sslab[q] + 3);
        {
            const int flo[3] = {bx.lo[0] - g, bx.lo[1] - g, bx.lo[2] - g};
            const int fn[3] = {nx + 2 * g, ny + 2 * g, nz + 2 * g};
            parents(ls.cf_full, ls.cf_full_total, b, flo, fn);
        }
        // Coarse cells the interpolation of this box's ghost cells reads.
        int glo[3], ghi[3];
        long long need = 1;
        for (int a = 0; a < 3; ++a) {
            glo[a] = floordiv2(bx.lo[a] - g) - s;
            ghi[a] = floordiv2(bx.hi[a] + g) + s;
            need *= ghi[a] - glo[a] + 1;
        }
        long long have = 0;
        for (int c = 0; c < (int)cs.L.boxes.size(); ++c) {
            const Box3 &cb = cs.L.boxes[c];
            bool hit = false;
            for (int sz = -1; sz <= 1; ++sz)
                for (int sy = -1; sy <= 1; ++sy)
                    for (int sx = -1; sx <= 1; ++sx) {
                        const int sh[3] = {sx * Nc, sy * Nc, sz * Nc};
                        long long v = 1;
                        for (int a = 0; a < 3 && v > 0; ++a) {
                            const int lo = std::max(glo[a], cb.lo[a] + sh[a]);
                            const int hi = std::min(ghi[a], cb.hi[a] + sh[a]);
                            v *= std::max(0, hi - lo + 1);
                        }
                        if (v > 0) {
                            have += v;
                            hit = true;
                        }
                    }
            if (hit) cand.push_back(c);
        }
        if (have != need)
            throw Error(1, "level " + std::to_string(lev) +
                               " is not properly nested: ghost cells of box " +
                               std::to_string(b) + " need coarse cells that no box of level " +
                               std::to_string(lev - 1) + " holds");
        start.push_back((int)cand.size());
    }
    upload_vec(ls.dcf_segs, ls.cf_segs);
    upload_vec(ls.dcf_full, ls.cf_full);
    upload_vec(ls.dcf_step, ls.cf_step);
    // Workgroup tiles: at most 32 x 8 x 4 parents (16 x 8 x 4 for the
    // quartic's reach 2), so the coarse brick of a tile, its parents grown
    // by the reach, takes at most 16 KB of LDS as doubles and several
    // workgroups share a CU.
    auto tile = [s](const std::vector<CfPSeg> &v, DevMem &d, int &n, int &lds) {
        const int lim[3] = {s > 1 ? 16 : 32, 8, 4};
        std::vector<CfTile> t;
        lds = 0;
        for (int q = 0; q < (int)v.size(); ++q)
            for (int z = 0; z < v[q].pn[2]; z += lim[2])
                for (int y = 0; y < v[q].pn[1]; y += lim[1])
                    for (int x = 0; x < v[q].pn[0]; x += lim[0]) {
                        CfTile c{};
                        c.seg = q;
                        const int o[3] = {x, y, z};
                        int cells = 1;
                        for (int a = 0; a < 3; ++a) {
                            c.plo[a] = v[q].plo[a] + o[a];
                            c.pn[a] = std::min(lim[a], v[q].pn[a] - o[a]);
                            cells *= c.pn[a] + 2 * s;
                        }
                        lds = std::max(lds, cells);
                        t.push_back(c);
                    }
        if (t.size() > (size_t)INT_MAX) throw Error(1, "too many coarse/fine tiles");
        n = (int)t.size();
        upload_vec(d, t);
    };
    tile(ls.cf_segs, ls.dcf_tiles, ls.cf_ntile, ls.cf_lds);
    tile(ls.cf_full, ls.dcf_full_tiles, ls.cf_full_ntile, ls.cf_full_lds);
    tile(ls.cf_step, ls.dcf_step_tiles, ls.cf_step_ntile, ls.cf_step_lds);
    upload_vec(ls.dcf_start, start);
    if (cand.empty()) cand.push_back(0);
    upload_vec(ls.dcf_cand, cand);
    ls.cf_built = true;
}

void Runtime::cf_record(double *rec, int lev, double time, bool first_state_only) const {
    const LevelState &cs = levels_[lev - 1];
    std::fill(rec, rec + 8, 0.);
    rec[4] = cs.cur;
    // GetLevelData (level_synchronizer.cpp:613-632): the new or the old
    // state within 1e-3 of their time difference, else both, linear in time.
    const double teps = std::fabs(cs.t - cs.t_old) * 1.e-3;
    if (time > cs.t - teps && time < cs.t + teps) {
        rec[3] = 1.;
    } else if (time > cs.t_old - teps && time < cs.t_old + teps) {
        rec[2] = 1.;
    } else if (first_state_only) {
        // InterpFromCoarseLevel in FillCoarsePatch receives only the first
        // state GetLevelData returns (level_synchronizer.cpp:64, 83).
        rec[2] = 1.;
    } else {
        rec[2] = rec[3] = 1.;
        rec[0] = (cs.t - time) / (cs.t - cs.t_old);
        rec[1] = (time - cs.t_old) / (cs.t - cs.t_old);
    }
}

static const double *push_params(double *host_ring, std::vector<hipEvent_t> &evs,
                                 int &slot, double *dev_ring, hipStream_t s,
                                 const std::vector<double> &blk);

const double *Runtime::push_record(const std::vector<double> &blk, int kind, double time) {
    if (capture_) {
        // A copy node from the graph's pinned record into its device record;
        // the host fills the pinned record before every replay.
        const int i = (int)capture_->recs.size();
        if (i >= capture_->nrec) throw Error(5, "too many parameter records in a captured step");
        int tidx = -1;
        for (size_t q = 0; q < step_times_.size() && tidx < 0; ++q)
            if (step_times_[q] == time) tidx = (int)q;
        if (tidx < 0) throw Error(5, "record time outside the captured step's time table");
        capture_->recs.push_back({kind, tidx});
        double *h = capture_->host + (size_t)i * kPrmBlock;
        double *d = capture_->dev.as<double>() + (size_t)i * kPrmBlock;
        std::copy(blk.begin(), blk.end(), h);
        SHM_HIP(hipMemcpyAsync(d, h, sizeof(double) * kPrmBlock, hipMemcpyHostToDevice,
                               stream_));
        return d;
    }
    return push_params(prm_host_, prm_ev_, prm_slot_, prm_dev_.as<double>(), stream_, blk);
}

CfArgs Runtime::cf_args(int lev, double time, bool first_state_only, bool step,
                        double *rec_dev) {
    LevelState &ls = levels_[lev];
    if (!ls.cf_built) build_cf(lev);
    LevelState &cs = levels_[lev - 1];
    CfArgs a{};
    a.fboxes = ls.L.dbox.as<BoxDesc>();
    a.cboxes = cs.L.dbox.as<BoxDesc>();
    a.segs = (step ? ls.dcf_step : ls.dcf_segs).as<CfPSeg>();
    a.tiles = (step ? ls.dcf_step_tiles : ls.dcf_tiles).as<CfTile>();
    a.ntile = step ? ls.cf_step_ntile : ls.cf_ntile;
    a.lds_cells = step ? ls.cf_step_lds : ls.cf_lds;
    a.total = step ? ls.cf_step_total : ls.cf_total;
    a.cand_start = ls.dcf_start.as<int>();
    a.cand = ls.dcf_cand.as<int>();
    a.Nc = cs.L.N;
    a.fcstride = ls.L.cstride;
    a.ccstride = cs.L.cstride;
    for (int q = 0; q < 2; ++q) {
        a.cs[q] = cs.U[q].get();
        a.cslo[q] = comp() ? cs.Ulo[q].get() : nullptr;
    }
    std::vector<double> rec(kPrmBlock, 0.);
    cf_record(rec.data(), lev, time, first_state_only);
    if (rec_dev) {
        // Synchronous: the record lives on this stack frame (the null
        // stream orders it with the context's blocking stream).
        SHM_HIP(hipMemcpy(rec_dev, rec.data(), sizeof(double) * kPrmBlock,
                          hipMemcpyHostToDevice));
        a.rec = rec_dev;
    } else {
        a.rec = push_record(rec, first_state_only ? 2 : 1, time);
    }
    return a;
}

template <Prec P, int M>
static void cf_dispatch(const CfArgs &a, int interp, int grid, hipStream_t s) {
    const size_t lds = sizeof(double) * (size_t)a.lds_cells * (M == kCfSsprk ? 3 : 1);
    switch (interp) {
        case 0:
            hipLaunchKernelGGL((cf_fill_kernel<P, M, 0>), dim3(grid), dim3(256), lds, s, a);
            break;
        case 1:
            hipLaunchKernelGGL((cf_fill_kernel<P, M, 1>), dim3(grid), dim3(256), lds, s, a);
            break;
        case 2:
            hipLaunchKernelGGL((cf_fill_kernel<P, M, 2>), dim3(grid), dim3(256), lds, s, a);
            break;
        default:
            hipLaunchKernelGGL((cf_fill_kernel<P, M, 4>), dim3(grid), dim3(256), lds, s, a);
            break;
    }
}

template <Prec P>
static void cf_dispatch_mode(const CfArgs &a, int mode, int interp, int grid,
                             hipStream_t s) {
    if (mode == kCfState)
        cf_dispatch<P, kCfState>(a, interp, grid, s);
    else if (mode == kCfK)
        cf_dispatch<P, kCfK>(a, interp, grid, s);
    else if (mode == kCfSsprk)
        cf_dispatch<P, kCfSsprk>(a, interp, grid, s);
    else
        cf_dispatch<P, kCfForce>(a, interp, grid, s);
}

void Runtime::cf_launch(CfArgs &a, int mode) {
    if (a.total == 0 || a.ncomp == 0) return;
    const int grid = a.ntile;  // one workgroup per tile
    if (a.lds_cells * sizeof(double) * (mode == kCfSsprk ? 3 : 1) > 65536)
        throw Error(5, "coarse/fine brick exceeds LDS");
    const int ip = interp_type_;
    EventPair *ep = timing_on ? timed_pair(1) : nullptr;
    if (ep) SHM_HIP(hipEventRecord(ep->a, stream_));
    switch (prec_) {
        case Prec::F64: cf_dispatch_mode<Prec::F64>(a, mode, ip, grid, stream_); break;
        case Prec::F32C: cf_dispatch_mode<Prec::F32C>(a, mode, ip, grid, stream_); break;
        case Prec::F32C16: cf_dispatch_mode<Prec::F32C16>(a, mode, ip, grid, stream_); break;
        case Prec::F32CS: cf_dispatch_mode<Prec::F32CS>(a, mode, ip, grid, stream_); break;
    }
    SHM_HIP(hipGetLastError());
    if (ep) SHM_HIP(hipEventRecord(ep->b, stream_));
}

void Runtime::fill_coarse_patch(int lev) {
    if (!has_cf(lev)) throw Error(1, "FillCoarsePatch needs a defined level above 0");
    LevelState &ls = level(lev);
    // Between the coarse level's two times the old state alone, not a time
    // interpolation (the first state GetLevelData returns).
    fill_coarse_full(lev, ls.cur, ls.t, true);
}

void Runtime::fill_coarse_full(int lev, int which, double time, bool first_state_only) {
    LevelState &ls = level(lev);
    CfArgs a = cf_args(lev, time, first_state_only);
    a.segs = ls.dcf_full.as<CfPSeg>();
    a.tiles = ls.dcf_full_tiles.as<CfTile>();
    a.ntile = ls.cf_full_ntile;
    a.lds_cells = ls.cf_full_lds;
    a.total = ls.cf_full_total;
    a.dst = ls.U[which].get();
    a.dstlo = comp() ? ls.Ulo[which].get() : nullptr;
    for (int c0 = 0; c0 < ncomp_; c0 += 32) {
        a.ncomp = std::min(32, ncomp_ - c0);
        for (int n = 0; n < a.ncomp; ++n) a.comp[n] = c0 + n;
        cf_launch(a, kCfState);
    }
    ls.ghosts_ok[which] = false;
}

void Runtime::cf_fill_state(int lev, double time, void *hi, void *lo,
                            const std::vector<int> &comps, bool step, double *rec_dev) {
    const std::vector<int> cs = comps.empty() ? all_comps() : comps;
    CfArgs a = cf_args(lev, time, false, step, rec_dev);
    a.dst = hi;
    a.dstlo = lo;
    for (size_t c0 = 0; c0 < cs.size(); c0 += 32) {
        a.ncomp = (int)std::min<size_t>(32, cs.size() - c0);
        for (int n = 0; n < a.ncomp; ++n) a.comp[n] = cs[c0 + n];
        cf_launch(a, kCfState);
    }
}

// ---------------------------------------------------------------------------
// Integrators.
// ---------------------------------------------------------------------------

/** Stage record ring: the host writes a pinned slot, the stream copies it to
 *  the matching device slot; a slot is rewritten only after its copy ran. */
static const double *push_params(double *host_ring,
                                 std::vector<hipEvent_t> &evs, int &slot,
                                 double *dev_ring, hipStream_t s,
                                 const std::vector<double> &blk) {
    const int i = slot;
    slot = (slot + 1) % kPrmSlots;
    if (evs[i]) SHM_HIP(hipEventSynchronize(evs[i]));
    else SHM_HIP(hipEventCreateWithFlags(&evs[i], hipEventDisableTiming));
    double *h = host_ring + (size_t)i * kPrmBlock;
    std::copy(blk.begin(), blk.end(), h);
    double *d = dev_ring + (size_t)i * kPrmBlock;
    SHM_HIP(hipMemcpyAsync(d, h, sizeof(double) * kPrmBlock,
                           hipMemcpyHostToDevice, s));
    SHM_HIP(hipEventRecord(evs[i], s));
    return d;
}

void Runtime::integrate(LevelState &ls, int lev, double dt, double dx,
                        double t0) {
    const int old = ls.cur ^ 1, nw = ls.cur;
    std::vector<double> blk(kPrmBlock);
    const std::vector<double> st = stage_times(t0, dt);
    step_times_ = st;
    step_times_.push_back(t0 + dt);
    auto prm = [&](double time) -> const double * {
        fill_params(blk.data(), time, dt, dx, lev);
        return push_record(blk, 0, time);
    };
    const int npair = kt_.nfields + (gw_ ? 6 : 0);
    std::vector<int> pairs(npair);
    for (int p = 0; p < npair; ++p) pairs[p] = p;
    void *ulo_old = comp() ? ls.Ulo[old].get() : nullptr;
    void *ulo_new = comp() ? ls.Ulo[nw].get() : nullptr;
    ls.ghosts_ok[nw] = false;
    // Ghost cells at coarse/fine boundaries take level lev-1 at each stage
    // time (FillIntermediatePatch).  Stage states are composed from u0 and
    // the stage arrays, so the fill writes the k or force ghost values that
    // compose to the interpolated state.
    const bool cf = has_cf(lev);
    if (cf && integrator_ >= 20) {
        if (nodes_[0] != 0.)
            throw Error(4, "coarse/fine ghost fill needs an RKN tableau with nodes[0] = 0");
        for (size_t i = 1; i < nodes_.size(); ++i)
            if (tableau_[i][i - 1] == 0.)
                throw Error(4, "coarse/fine ghost fill needs a nonzero RKN tableau "
                               "sub-diagonal");
        if (ko_)
            throw Error(4, "dissipation with RKN on levels with coarse/fine ghost "
                           "cells is not supported (stage momenta at ghost cells)");
    }
    auto cf_fill_k = [&](void *k, double time, double c) {
        CfArgs f = cf_args(lev, time, false, true);
        f.dst = k;
        f.u0 = ls.U[old].get();
        f.u0lo = tile_lo_of(ls, old);
        f.c0 = c;
        const std::vector<int> sc = stencil_comps();
        for (size_t c0 = 0; c0 < sc.size(); c0 += 32) {
            f.ncomp = (int)std::min<size_t>(32, sc.size() - c0);
            for (int q = 0; q < f.ncomp; ++q) f.comp[q] = sc[c0 + q];
            cf_launch(f, kCfK);
        }
    };

    if (integrator_ == 10) {
        // Low-storage SSPRK3 with the stage states rebuilt from u0 and k:
        // k1 -> K[0], k2 -> K[1] (k is read at neighbours, so not in place).
        if (resolved_launch(lev).fused && fused_possible(ls)) {
            // Stages one and two in one pass: k1 stays on chip.
            const double *p1 = prm(st[0]);
            const double *p2 = prm(st[1]);
            launch_fused(ls, p1, p2, dt, lev);
        } else {
            StageArgs a{};
            a.prm = prm(st[0]);
            a.x = ls.U[old].get();
            a.xlo = ulo_old;
            a.kout = ls.K[0].get();
            launch_stage(kSsprk1, ls, a, lev);
            // Coarse/fine ghosts of k1 and k2 were written with the step's
            // start state (cf_fill_ssprk); stage kernels write valid cells.
            if (cf && !cf_merged(lev)) cf_fill_k(ls.K[0].get(), st[1], dt);
            halo(ls, ls.K[0].get(), stencil_comps());

            StageArgs b{};
            b.prm = prm(st[1]);
            b.u0 = ls.U[old].get();
            b.u0lo = ulo_old;
            b.kin = ls.K[0].get();
            b.kout = ls.K[1].get();
            b.c[0] = 1.;
            b.c[2] = dt;
            launch_stage(kSsprk2, ls, b, lev);
            if (cf && !cf_merged(lev)) cf_fill_k(ls.K[1].get(), st[2], dt / 4.);
            halo(ls, ls.K[1].get(), stencil_comps());
        }

        StageArgs c{};
        c.prm = prm(st[2]);
        c.u0 = ls.U[old].get();
        c.u0lo = ulo_old;
        c.kin = ls.K[1].get();
        c.unew = ls.U[nw].get();
        c.unewlo = ulo_new;
        c.c[0] = 0.25;
        c.c[1] = dt * 2. / 3.;
        c.c[2] = dt / 4.;
        launch_stage(kSsprk3, ls, c, lev);
    } else if (integrator_ >= 20) {
        const int s = (int)nodes_.size();
        const bool ry = rkn_y(lev);
        int i0 = 0;
        if (resolved_launch(lev).fused && fused_possible(ls)) {
            // Stages one and two in one pass (boxes wrap on themselves, so
            // there are no coarse/fine ghosts); F0 and F1 as below.
            launch_fused(ls, prm(st[0]), prm(st[1]), dt, lev);
            halo(ls, ls.F[0].get(), pairs);
            halo(ls, ls.F[1].get(), pairs);
            i0 = 2;
        }
        // Three stages on GW levels: the second stage also writes the final
        // stage's field states, which saves the final stage composing them
        // from four arrays at every stencil point of seven stencil
        // components.  With one or two stencil components the composed
        // reads hit the caches and the extra array costs more than it saves
        // (64-box RealScalar ensemble: 1.52 -> 1.66 ms per step).
        const bool ry2 = ry && s == 3 && i0 == 0 && gw_;
        for (int i = i0; i < s; ++i) {
            StageArgs a{};
            a.prm = prm(st[i]);
            a.u0 = ls.U[old].get();
            a.u0lo = ulo_old;
            a.c[0] = dt * nodes_[i];
            a.nprev = i;
            for (int j = 0; j < i; ++j) {
                a.c[1 + j] = dt * dt * tableau_[i][j];
                a.f[j] = ls.F[j].get();
            }
            if (i < s - 1) {
                a.fout = ls.F[i].get();
                int kind = kRkn0 + i;
                if (ry && i == 0) {
                    // Also the second stage's field states, at its node.
                    kind = kRkn0Y;
                    a.y = ls.Y2.get();
                    a.c[18] = dt * nodes_[1];
                    a.c[19] = dt * dt * tableau_[1][0];
                } else if (ry && i == 1) {
                    kind = kRkn1Y;
                    a.y = ls.Y2.get();
                    if (ry2) {
                        kind = kRkn1YY;
                        a.y2 = ls.Y3.get();
                        a.c[12] = dt * nodes_[2];
                        a.c[13] = dt * dt * tableau_[2][0];
                        a.c[14] = dt * dt * tableau_[2][1];
                    }
                }
                launch_stage(kind, ls, a, lev);
                if (cf) {
                    // Stage i+1 reads phi0 + dt c_{i+1} pi0 + sum_j dt^2 a_{i+1,j} F_j.
                    CfArgs f = cf_args(lev, st[i + 1], false, true);
                    f.dst = ls.F[i].get();
                    f.u0 = ls.U[old].get();
                    f.u0lo = tile_lo_of(ls, old);
                    f.c0 = dt * nodes_[i + 1];
                    f.m = i;
                    for (int j = 0; j <= i; ++j) f.cf[j] = dt * dt * tableau_[i + 1][j];
                    for (int j = 0; j < i; ++j) f.f[j] = ls.F[j].get();
                    for (int p0 = 0; p0 < npair; p0 += 32) {
                        f.ncomp = std::min(32, npair - p0);
                        for (int q = 0; q < f.ncomp; ++q) {
                            f.pair[q] = p0 + q;
                            f.comp[q] = pair_field(p0 + q);
                            f.mom[q] = pair_mom(p0 + q);
                        }
                        cf_launch(f, kCfForce);
                    }
                }
                halo(ls, ls.F[i].get(), pairs);
                if (ry && i == 0) halo(ls, ls.Y2.get(), pairs, false, (int)ysize());
                if (ry2 && i == 1) halo(ls, ls.Y3.get(), pairs, false, (int)ysize());
            } else {
                a.c[6] = dt;
                for (int j = 0; j < i; ++j) {
                    a.c[7 + j] = dt * dt * bbar_[j];
                    a.c[12 + j] = dt * b_[j];
                }
                a.c[11] = dt * dt * bbar_[i];
                a.c[16] = dt * b_[i];
                a.unew = ls.U[nw].get();
                a.unewlo = ulo_new;
                int kind = kRknF0 + i;
                if (ry2) {
                    kind = kRknF2Y;
                    a.y2 = ls.Y3.get();
                }
                launch_stage(kind, ls, a, lev);
            }
        }
    } else if (integrator_ == 11) {
        StageArgs a{};
        a.prm = prm(st[0]);
        a.x = ls.U[old].get();
        a.xlo = ulo_old;
        a.u0 = ls.U[old].get();
        a.u0lo = ulo_old;
        a.vh = ls.VH.get();
        a.unew = ls.U[nw].get();
        a.unewlo = ulo_new;
        a.c[0] = dt / 2.;
        a.c[1] = dt;
        launch_stage(kLeap1, ls, a, lev);
        // The second kick reads the fields' leading words at ghost cells
        // (and their remainders with FP64 tiles).
        if (cf)
            cf_fill_state(lev, st[1], ls.U[nw].get(), tile_lo_of(ls, nw), stencil_comps(),
                          true);
        halo_state(ls, nw, stencil_comps());

        StageArgs b{};
        b.prm = prm(st[1]);
        b.x = ls.U[nw].get();
        b.xlo = ulo_new;
        b.vh = ls.VH.get();
        b.unew = ls.U[nw].get();
        b.unewlo = ulo_new;
        b.c[0] = dt / 2.;
        launch_stage(kLeap2, ls, b, lev);
    }
    // FillIntermediatePatch at the end of the step.  On levels with
    // coarse/fine ghost cells the next step fills them again, so the fill is
    // left to what reads them in between (tagging, outputs), with the same
    // coarse data and time.
    if (!cf) halo_state(ls, nw, all_comps());
    ls.ghosts_ok[nw] = !cf;
}

std::vector<double> Runtime::stage_times(double t0, double dt) const {
    if (integrator_ == 10) return {t0, t0 + dt, t0 + dt / 2.};
    if (integrator_ == 11) return {t0, t0 + dt};
    std::vector<double> t;
    for (double c : nodes_) t.push_back(t0 + dt * c);
    return t;
}

void Runtime::release_graphs() {
    for (auto &kv : graphs_) {
        StepGraph &g = kv.second;
        if (g.done) (void)hipEventSynchronize(g.done);
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.done) (void)hipEventDestroy(g.done);
        if (g.host) (void)hipHostFree(g.host);
    }
    graphs_.clear();
}

/** Runs Integrate for a level from a captured graph: captured on the first
 *  step after an eager one with the same layout and launch shape (so every
 *  lazily built table exists), replayed afterwards.  Returns false when the
 *  step has to run eagerly. */
bool Runtime::replay_step(LevelState &ls, int lev, double dt, double dx, double t0) {
    if (!graphs || timing_on || lev < 0) return false;
    auto w = warm_.find(lev);
    if (w == warm_.end() || w->second != epoch_) return false;
    const std::vector<double> st = stage_times(t0, dt);
    StepGraph &g = graphs_[{lev, ls.cur ^ 1}];
    if (g.epoch != epoch_) {
        if (g.exec) {
            if (g.done) SHM_HIP(hipEventSynchronize(g.done));
            SHM_HIP(hipGraphExecDestroy(g.exec));
            g.exec = nullptr;
        }
        if (!g.host) {
            // Stage parameters plus the coarse/fine fills of a step.
            g.nrec = 4 * (int)st.size() + 8;
            SHM_HIP(hipHostMalloc((void **)&g.host, sizeof(double) * kPrmBlock * g.nrec,
                                  hipHostMallocDefault));
            g.dev.alloc(sizeof(double) * kPrmBlock * g.nrec);
            SHM_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
        }
        hipGraph_t graph = nullptr;
        capture_ = &g;
        g.recs.clear();
        SHM_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            integrate(ls, lev, dt, dx, t0);
        } catch (...) {
            capture_ = nullptr;
            (void)hipStreamEndCapture(stream_, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        capture_ = nullptr;
        SHM_HIP(hipStreamEndCapture(stream_, &graph));
        const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        SHM_HIP(ie);
        g.epoch = epoch_;
    } else {
        ls.ghosts_ok[ls.cur] = !has_cf(lev);  // what integrate() records eagerly
    }
    // The previous replay of this graph has read its pinned records.
    SHM_HIP(hipEventSynchronize(g.done));
    std::vector<double> times = st;
    times.push_back(t0 + dt);
    for (size_t i = 0; i < g.recs.size(); ++i) {
        double *h = g.host + i * kPrmBlock;
        const double time = times[g.recs[i].second];
        if (g.recs[i].first == 0)
            fill_params(h, time, dt, dx, lev);
        else
            cf_record(h, lev, time, g.recs[i].first == 2);
    }
    SHM_HIP(hipGraphLaunch(g.exec, stream_));
    SHM_HIP(hipEventRecord(g.done, stream_));
    ++graph_replays;
    return true;
}

void Runtime::advance(int lev) {
    LevelState &ls = level(lev);
    ls.cur ^= 1;  // swap grid_old / grid_new
    const int old = ls.cur ^ 1;
    ls.t_old = ls.t;
    const double dt = sim_->dt[lev], dx = sim_->dx[lev];
    if (cf_merged(lev)) {
        cf_fill_ssprk(ls, lev, old, ls.t_old, dt);
        halo_state(ls, old, step_ghost_comps());
        ls.ghosts_ok[old] = false;
    } else {
        fill_patch(ls, old, lev, ls.t_old, /*for_step=*/true);
    }
    if (!replay_step(ls, lev, dt, dx, ls.t_old)) {
        integrate(ls, lev, dt, dx, ls.t_old);
        if (lev >= 0) warm_[lev] = epoch_;
    }
    ls.t = ls.t_old + dt;
    ls.istep += 1;
}

void Runtime::fill_rhs(int lev, double time, double weight, bool add) {
    LevelState &ls = level(lev);
    if (!ls.RHS) ls.RHS.alloc((size_t)ls.L.cstride * ls.ncomp * esize());
    fill_patch(ls, ls.cur, lev, ls.t);
    std::vector<double> blk(kPrmBlock);
    fill_params(blk.data(), time, sim_->dt[std::max(lev, 0)],
                sim_->dx[std::max(lev, 0)], lev);
    StageArgs a{};
    a.prm = push_params(prm_host_, prm_ev_, prm_slot_,
                        prm_dev_.as<double>(), stream_, blk);
    a.x = ls.U[ls.cur].get();
    a.xlo = comp() ? ls.Ulo[ls.cur].get() : nullptr;
    a.kin = ls.RHS.get();
    a.kout = ls.RHS.get();
    a.c[0] = weight;
    launch_stage(add ? kFillAddRhs : kFillRhs, ls, a, lev);
}

// ---------------------------------------------------------------------------
// Shadow level, truncation errors and tags.
// ---------------------------------------------------------------------------

void Runtime::avg_down(LevelState &fine, int fw, LevelState &crse, int cw) {
    AvgArgs a{};
    a.fboxes = fine.L.dbox.as<BoxDesc>();
    a.cboxes = crse.L.dbox.as<BoxDesc>();
    a.fcstride = fine.L.cstride;
    a.ccstride = crse.L.cstride;
    a.fine = fine.U[fw].get();
    a.finelo = comp() ? fine.Ulo[fw].get() : nullptr;
    a.crse = crse.U[cw].get();
    a.crselo = comp() ? crse.Ulo[cw].get() : nullptr;
    a.cprefix = fine.L.dcprefix.as<long long>();
    a.nbox = (int)fine.L.boxes.size();
    a.total = fine.L.cprefix.back();
    a.ncomp = ncomp_;
    const int grid = (int)std::max<long long>(1, std::min<long long>((a.total + 255) / 256, 8LL * ncu_));
    switch (sprec()) {
        case Prec::F64:
            hipLaunchKernelGGL(avgdown_kernel<Prec::F64>, dim3(grid), dim3(256),
                               0, stream_, a);
            break;
        case Prec::F32C:
            hipLaunchKernelGGL(avgdown_kernel<Prec::F32C>, dim3(grid), dim3(256),
                               0, stream_, a);
            break;
        case Prec::F32C16:
            hipLaunchKernelGGL(avgdown_kernel<Prec::F32C16>, dim3(grid),
                               dim3(256), 0, stream_, a);
            break;
    }
    SHM_HIP(hipGetLastError());
    crse.ghosts_ok[cw] = false;
}

void Runtime::define_shadow() {
    if (independent_)
        throw Error(4, "no shadow level for independent boxes");
    LevelState &l0 = levels_.at(0);
    if (!shadow_.defined() || shadow_.L.boxes.size() != l0.L.boxes.size()) {
        shadow_ = LevelState();
        shadow_.L.build(coarsen(l0.L.boxes), sim_->nghost,
                        sim_->coarse_level_grid_size / 2, true);
        shadow_.ncomp = ncomp_;
        ensure_arrays(shadow_, integrator_);
        shadow_.t = l0.t;
    }
}

void Runtime::create_shadow_level() {
    LevelState &l0 = level(0);
    const int old = l0.cur ^ 1;
    if (l0.t_old == -DBL_MAX)
        throw Error(3, "no valid grid_old on level 0 to build the shadow level");
    define_shadow();
    // shadow_level_tmp = average_down(grid_old[0]) at grid_old[0].t
    avg_down(l0, old, shadow_, 0);
    shadow_.t = l0.t_old;
    // Integrator::Advance(-1): one step of 2 dt on 2 dx from buffer 0 into 1.
    shadow_.cur = 1;
    shadow_.t_old = shadow_.t;
    fill_patch(shadow_, 0, -1, shadow_.t);
    const double dt = sim_->dt[0] * 2., dx = sim_->dx[0] * 2.;
    integrate(shadow_, -1, dt, dx, shadow_.t_old);
    shadow_.t = shadow_.t_old + dt;
}

LevelState &Runtime::coarse_mirror(int lev) {
    LevelState &fine = level(lev);
    if (!fine.mirror || fine.mirror_epoch != epoch_) {
        LevelState &cl = level(lev - 1);
        auto m = std::make_unique<LevelState>();
        m->L.build(coarsen(fine.L.boxes), 0, sim_->dimN[lev - 1], true);
        m->ncomp = ncomp_;
        m->U[0].alloc((size_t)m->L.cstride * ncomp_ * esize());
        if (comp()) m->Ulo[0].alloc((size_t)m->L.cstride * ncomp_ * losize());
        m->cur = 0;
        const std::vector<HaloSeg> in = copy_plan(m->L, cl.L, &fine.mtot_in);
        const std::vector<HaloSeg> out = copy_plan(cl.L, m->L, &fine.mtot_out);
        upload_vec(fine.mplan_in, in);
        upload_vec(fine.mplan_out, out);
        fine.mnseg_in = (int)in.size();
        fine.mnseg_out = (int)out.size();
        fine.mirror = std::move(m);
        fine.mirror_epoch = epoch_;
    }
    return *fine.mirror;
}

void Runtime::copy_state(void *dst, void *dstlo, const Layout &dl, const void *src,
                         const void *srclo, const Layout &sl, const DevMem &plan,
                         int nseg, long long total) {
    if (total == 0) return;
    CopyArgs c{};
    c.dst = dst;
    c.src = src;
    c.dcstride = dl.cstride;
    c.scstride = sl.cstride;
    c.dboxes = dl.dbox.as<BoxDesc>();
    c.sboxes = sl.dbox.as<BoxDesc>();
    c.segs = plan.as<HaloSeg>();
    c.nseg = nseg;
    c.total = total;
    c.ncomp = ncomp_;
    launch_copy(c, esize(), ncu_, stream_);
    if (comp()) {
        c.dst = dstlo;
        c.src = srclo;
        launch_copy(c, losize(), ncu_, stream_);
    }
}

void Runtime::te_tag(int lev, int flags, uint8_t *mask, long long *counts,
                     double *te_host) {
    LevelState &fine = level(lev);
    const sledgehamr::Sledgehamr &S = *sim_;
    // Device buffers of the context, kept between calls.
    TagBuffers tb;
    upload_into(te_crit_, S.te_crit);
    tb.crit = te_crit_.as<double>();
    // The mask stays with the level for the local regrid.
    const size_t nmask = (size_t)std::max<long long>(1, fine.L.nvalid());
    if (flags & kTeTags) {
        if (fine.tags.bytes() != nmask) fine.tags.alloc(nmask);
        tb.mask = fine.tags.as<uint8_t>();
        tb.keep_tags = true;
    } else {
        if (te_mask_.bytes() < nmask) te_mask_.alloc(nmask);
        tb.mask = te_mask_.as<uint8_t>();
    }
    if (te_counts_.bytes() < sizeof(unsigned long long) * (2 + ncomp_))
        te_counts_.alloc(sizeof(unsigned long long) * (2 + ncomp_));
    tb.counts = te_counts_.as<unsigned long long>();
    const long long total = fine.L.cprefix.back();
    if (te_host && (flags & kTeCompute)) {
        if (te_out_.bytes() < sizeof(double) * total * ncomp_)
            te_out_.alloc(sizeof(double) * total * ncomp_);
        tb.te = te_out_.as<double>();
    }
    tb.ptag = &te_ptag_;
    tb.pmod = &te_pmod_;
    te_tag_run(lev, flags, tb, mask, counts, te_host);
}

void Runtime::te_tag_noalloc(int lev, int flags, uint8_t *mask, long long *counts) {
    if (lev < 0 || lev >= (int)levels_.size() || !has_level(lev))
        throw Error(1, "no such level");
    LevelState &fine = levels_[lev];
    const sledgehamr::Sledgehamr &S = *sim_;
    if (2 + ncomp_ > kScratchWords || (int)S.te_crit.size() > kScratchWords)
        throw Error(4, "too many components for the allocation-free tagging");
    if (fine.L.nvalid() > kScratchMaxCells)
        throw Error(1, "level " + std::to_string(lev) + " holds " +
                           std::to_string(fine.L.nvalid()) +
                           " valid cells; the allocation-free tagging takes at most " +
                           std::to_string(kScratchMaxCells));
    // What the tagging would otherwise build on first use.
    if (flags & kTeCompute) {
        if (lev == 0 && !shadow_.defined())
            throw Error(3, "no shadow level for truncation errors on level 0");
        if (lev > 0 && (!fine.mirror || fine.mirror_epoch != epoch_))
            throw Error(3, "the coarsened copy of level " + std::to_string(lev) +
                               " is not set up: run shm_te_tag on the level first");
    }
    if (has_cf(lev) && !fine.cf_built)
        throw Error(3, "the coarse/fine fill of level " + std::to_string(lev) +
                           " is not set up: run shm_te_tag or a step on the level first");
    TagBuffers tb;
    tb.counts = reinterpret_cast<unsigned long long *>(scratch_ + kScratchCountsOff);
    tb.crit = reinterpret_cast<double *>(scratch_ + kScratchCritOff);
    tb.ptag_dev = reinterpret_cast<double *>(scratch_ + kScratchTagPrmOff);
    tb.pmod_dev = reinterpret_cast<double *>(scratch_ + kScratchModPrmOff);
    tb.cf_rec = reinterpret_cast<double *>(scratch_ + kScratchRecOff);
    tb.mask = scratch_ + kScratchMaskOff;
    if (!S.te_crit.empty())
        SHM_HIP(hipMemcpy(tb.crit, S.te_crit.data(), sizeof(double) * S.te_crit.size(),
                          hipMemcpyHostToDevice));
    te_tag_run(lev, flags, tb, mask, counts, nullptr);
}

void Runtime::te_tag_run(int lev, int flags, TagBuffers &tb, uint8_t *mask,
                         long long *counts, double *te_host) {
    LevelState &fine = level(lev);
    const double time = fine.t;
    LevelState *crse = nullptr;
    if (flags & kTeCompute) {
        if (lev == 0) {
            if (!shadow_.defined())
                throw Error(3, "no shadow level for truncation errors on level 0");
            if (std::fabs(shadow_.t - fine.t) >= std::fabs(shadow_.t) * 1e-8 &&
                shadow_.t != fine.t)
                throw Error(3, "Shadow level not sync'd with coarse level!");
            crse = &shadow_;
        } else {
            LevelState &cl = level(lev - 1);
            LevelState &m = coarse_mirror(lev);
            copy_state(m.U[0].get(), m.Ulo[0].get(), m.L, cl.U[cl.cur].get(),
                       cl.Ulo[cl.cur].get(), cl.L, fine.mplan_in, fine.mnseg_in,
                       fine.mtot_in);
            crse = &m;
        }
    }
    // User criteria may read neighbours of a cell: complete every ghost cell
    // of the fine state first (stage kernels wrap periodic boxes instead).
    // The level's end-of-step coarse/fine ghost fill (FillIntermediatePatch
    // at its time), then the same-level ghosts.
    if (has_cf(lev))
        cf_fill_state(lev, time, fine.U[fine.cur].get(),
                      comp() ? fine.Ulo[fine.cur].get() : nullptr, {}, false, tb.cf_rec);
    halo(fine, fine.U[fine.cur].get(), all_comps(), /*full=*/true);
    if (comp()) halo(fine, fine.Ulo[fine.cur].get(), all_comps(), true, (int)losize());
    fine.ghosts_ok[fine.cur] = false;
    const sledgehamr::Sledgehamr &S = *sim_;
    std::vector<double> ptag, pmod;
    sim_->SetParamsTagCellForRefinement(ptag, time, lev);
    if (S.shadow_hierarchy) sim_->SetParamsTruncationModifier(pmod, time, lev);
    ptag.resize(std::max<size_t>(ptag.size(), 1));
    pmod.resize(std::max<size_t>(pmod.size(), 1));
    const double *dptag, *dpmod;
    if (tb.ptag) {
        upload_into(*tb.ptag, ptag);
        upload_into(*tb.pmod, pmod);
        dptag = tb.ptag->as<double>();
        dpmod = tb.pmod->as<double>();
    } else {
        if ((int)ptag.size() > kScratchWords || (int)pmod.size() > kScratchWords)
            throw Error(4, "too many tagging parameters for the allocation-free tagging");
        SHM_HIP(hipMemcpy(tb.ptag_dev, ptag.data(), sizeof(double) * ptag.size(),
                          hipMemcpyHostToDevice));
        SHM_HIP(hipMemcpy(tb.pmod_dev, pmod.data(), sizeof(double) * pmod.size(),
                          hipMemcpyHostToDevice));
        dptag = tb.ptag_dev;
        dpmod = tb.pmod_dev;
    }
    if (tb.keep_tags) fine.tags_t = -DBL_MAX;
    // Every valid cell's mask byte is written; the counters are summed
    // atomically, so they start at zero.
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, stream_, tb.counts,
                       2 + ncomp_);
    SHM_HIP(hipGetLastError());
    const long long total = fine.L.cprefix.back();

    TeTagArgs a{};
    a.fboxes = fine.L.dbox.as<BoxDesc>();
    a.cboxes = crse ? crse->L.dbox.as<BoxDesc>() : fine.L.dbox.as<BoxDesc>();
    a.fcstride = fine.L.cstride;
    a.ccstride = crse ? crse->L.cstride : 0;
    a.fine = fine.U[fine.cur].get();
    a.finelo = comp() ? fine.Ulo[fine.cur].get() : nullptr;
    a.crse = crse ? crse->U[crse->cur].get() : nullptr;
    a.crselo = (crse && comp()) ? crse->Ulo[crse->cur].get() : nullptr;
    a.cprefix = fine.L.dcprefix.as<long long>();
    a.mprefix = fine.L.dvprefix.as<long long>();
    a.nbox = (int)fine.L.boxes.size();
    a.total = total;
    a.ncomp = ncomp_;
    a.te_out = tb.te;
    a.mask = tb.mask;
    a.counts = tb.counts;
    a.te_crit = tb.crit;
    a.params_tag = dptag;
    a.params_mod = dpmod;
    a.lev = lev;
    a.time = time;
    a.dt = sim_->dt[lev];
    a.dx = sim_->dx[lev];
    a.flags = flags;
    const TeTagFn fn = kt_.te_tag[(int)sprec()][gw_ ? 1 : 0];
    if (!fn) throw Error(4, "tagging kernel not built");
    if (total > 0) {
        const int grid = (int)std::min<long long>((total + 255) / 256, 8LL * ncu_);
        SHM_HIP(fn(a, grid, stream_));
    }
    if (crse && lev >= 1 && (flags & kTeAvgDown)) {
        LevelState &cl = level(lev - 1);
        copy_state(cl.U[cl.cur].get(), cl.Ulo[cl.cur].get(), cl.L, crse->U[0].get(),
                   crse->Ulo[0].get(), crse->L, fine.mplan_out, fine.mnseg_out,
                   fine.mtot_out);
        cl.ghosts_ok[cl.cur] = false;
    }
    if (tb.keep_tags && (flags & kTeTags)) fine.tags_t = time;
    if (mask && (flags & kTeTags))
        SHM_HIP(hipMemcpyAsync(mask, tb.mask, fine.L.nvalid(), hipMemcpyDeviceToHost,
                               stream_));
    unsigned long long hc[kScratchWords];
    SHM_HIP(hipMemcpyAsync(hc, tb.counts, sizeof(unsigned long long) * (2 + ncomp_),
                           hipMemcpyDeviceToHost, stream_));
    if (te_host && tb.te)
        SHM_HIP(hipMemcpyAsync(te_host, tb.te, sizeof(double) * total * ncomp_,
                               hipMemcpyDeviceToHost, stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
    if (counts)
        for (int i = 0; i < 2 + ncomp_; ++i) counts[i] = (long long)hc[i];
}

// ---------------------------------------------------------------------------
// Data movement and initial states.
// ---------------------------------------------------------------------------

void Runtime::upload_array(int lev, int comp_id, const double *host, int layout) {
    if (layout == 0) return upload(lev, comp_id, host);
    if (layout != 1 && layout != 2) throw Error(1, "array layout must be 0, 1 or 2");
    if (lev == -1) define_shadow();
    LevelState &ls = level(lev);
    const long long N = ls.L.N;
    std::vector<double> v((size_t)ls.L.nvalid());
    long long chunk = 0;
    for (size_t b = 0; b < ls.L.boxes.size(); ++b) {
        const Box3 &bx = ls.L.boxes[b];
        const long long nx = bx.n(0), ny = bx.n(1), nz = bx.n(2);
        double *dst = v.data() + ls.L.vprefix[b];
        for (long long z = 0; z < nz; ++z)
            for (long long y = 0; y < ny; ++y)
                for (long long x = 0; x < nx; ++x) {
                    const long long i = bx.lo[0] + x, j = bx.lo[1] + y, k = bx.lo[2] + z;
                    const long long ind = layout == 1 ? (i * N + j) * N + k
                                                      : chunk + (x * ny + y) * nz + z;
                    dst[(z * ny + y) * nx + x] = host[ind];
                }
        chunk += nx * ny * nz;
    }
    upload(lev, comp_id, v.data());
}

void Runtime::valid_io(LevelState &ls, int comp_id, int which, double *host, bool to_state) {
    if (comp_id < 0 || comp_id >= ncomp_) throw Error(1, "component out of range");
    const long long n = ls.L.nvalid();
    if (n == 0) return;
    DevMem dense;
    dense.alloc(sizeof(double) * (size_t)n);
    ValidIoArgs a{};
    a.boxes = ls.L.dbox.as<BoxDesc>();
    a.vprefix = ls.L.dvprefix.as<long long>();
    a.nbox = (int)ls.L.boxes.size();
    a.total = n;
    a.hi = ls.U[which].get();
    a.lo = comp() ? ls.Ulo[which].get() : nullptr;
    a.comp_off = (long long)comp_id * ls.L.cstride;
    a.dense = dense.as<double>();
    const int grid = (int)std::min<long long>((n + 255) / 256, 8LL * ncu_);
    if (to_state)
        SHM_HIP(hipMemcpy(dense.get(), host, sizeof(double) * (size_t)n,
                          hipMemcpyHostToDevice));
    switch (sprec()) {
        case Prec::F64:
            if (to_state)
                hipLaunchKernelGGL((valid_io_kernel<Prec::F64, true>), dim3(grid), dim3(256), 0,
                                   stream_, a);
            else
                hipLaunchKernelGGL((valid_io_kernel<Prec::F64, false>), dim3(grid), dim3(256),
                                   0, stream_, a);
            break;
        case Prec::F32C:
            if (to_state)
                hipLaunchKernelGGL((valid_io_kernel<Prec::F32C, true>), dim3(grid), dim3(256),
                                   0, stream_, a);
            else
                hipLaunchKernelGGL((valid_io_kernel<Prec::F32C, false>), dim3(grid),
                                   dim3(256), 0, stream_, a);
            break;
        case Prec::F32C16:
            if (to_state)
                hipLaunchKernelGGL((valid_io_kernel<Prec::F32C16, true>), dim3(grid),
                                   dim3(256), 0, stream_, a);
            else
                hipLaunchKernelGGL((valid_io_kernel<Prec::F32C16, false>), dim3(grid),
                                   dim3(256), 0, stream_, a);
            break;
    }
    SHM_HIP(hipGetLastError());
    if (!to_state)
        SHM_HIP(hipMemcpyAsync(host, dense.get(), sizeof(double) * (size_t)n,
                               hipMemcpyDeviceToHost, stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
}

void Runtime::upload(int lev, int comp_id, const double *host) {
    if (lev == -1) define_shadow();
    LevelState &ls = level(lev);
    // The stream may still read the state (steps queued before the upload).
    sync();
    valid_io(ls, comp_id, ls.cur, const_cast<double *>(host), true);
    ls.ghosts_ok[ls.cur] = false;
}

void Runtime::download(int lev, int comp_id, double *host, bool old_state) {
    LevelState &ls = level(lev);
    sync();
    valid_io(ls, comp_id, old_state ? (ls.cur ^ 1) : ls.cur, host, false);
}

void Runtime::download_rhs(int lev, int comp, double *host) {
    LevelState &ls = level(lev);
    if (!ls.RHS) throw Error(3, "no rhs has been computed on this level");
    sync();
    const size_t es = esize();
    for (size_t b = 0; b < ls.L.boxes.size(); ++b) {
        const Box3 &bx = ls.L.boxes[b];
        const BoxDesc &bd = ls.L.hbox[b];
        const int nx = bx.n(0), ny = bx.n(1), nz = bx.n(2);
        std::vector<char> buf((size_t)bx.vol() * es);
        for (int z = 0; z < nz; ++z) {
            const char *base = static_cast<const char *>(ls.RHS.get()) +
                               ((size_t)comp * ls.L.cstride + bd.base +
                                (size_t)z * bd.kstride) * es;
            SHM_HIP(hipMemcpy2D(buf.data() + (size_t)z * nx * ny * es,
                                (size_t)nx * es, base, (size_t)bd.jstride * es,
                                (size_t)nx * es, ny, hipMemcpyDeviceToHost));
        }
        double *dst = host + ls.L.vprefix[b];
        for (long long i = 0; i < bx.vol(); ++i)
            dst[i] = prec_ == Prec::F64
                         ? reinterpret_cast<const double *>(buf.data())[i]
                         : (double)reinterpret_cast<const float *>(buf.data())[i];
    }
}

void Runtime::init_state(int lev, int kind, uint64_t seed, const double *p,
                         int np) {
    LevelState &ls = level(lev);
    InitArgs a{};
    a.arr = ls.U[ls.cur].get();
    a.arrlo = comp() ? ls.Ulo[ls.cur].get() : nullptr;
    a.cstride = ls.L.cstride;
    a.boxes = ls.L.dbox.as<BoxDesc>();
    a.vprefix = ls.L.dvprefix.as<long long>();
    a.nbox = (int)ls.L.boxes.size();
    a.total = ls.L.nvalid();
    a.N = ls.L.N;
    a.kind = kind;
    a.seed = seed;
    a.nfields = kt_.nfields;
    a.ncomp = ncomp_;
    for (int i = 0; i < np && i < 8; ++i) a.p[i] = p[i];
    if (ncomp_ > 32) throw Error(4, "initial-state generator supports <= 32 components");
    const double N0 = (double)sim_->coarse_level_grid_size;
    a.N0 = N0;
    a.cell = N0 / (double)ls.L.N;
    // Object tables, 8 doubles per object, drawn from the counter-based
    // generator (streams 100+: bubbles, 200+: loops), positions in level-0
    // cells so every level sees the same objects.
    std::vector<double> obj;
    if (kind == kInitBubbles || kind == kInitAddBubbles) {
        // p: count, phi_in, R0, wall width w, nucleation times in
        // [t_lo, t_hi], injection window (t_from, t_to].
        if (np < 8 || p[0] < 0 || p[0] > 4096 || p[3] <= 0)
            throw Error(1, "bubbles take 8 parameters: n, phi_in, R0, w, t_lo, t_hi, t_from, t_to");
        for (int i = 0; i < (int)p[0]; ++i) {
            const double row[8] = {
                N0 * rng_uniform(seed, 100, i), N0 * rng_uniform(seed, 101, i),
                N0 * rng_uniform(seed, 102, i),
                p[4] + (p[5] - p[4]) * rng_uniform(seed, 103, i), 0., 0., 0., 0.};
            obj.insert(obj.end(), row, row + 8);
        }
    } else if (kind == kInitStringLoops) {
        // p: count, r_min, r_max (level-0 cells), core width w.
        if (np < 4 || p[0] < 0 || p[0] > 4096 || p[3] <= 0 || ncomp_ < 2)
            throw Error(1, "string loops take 4 parameters: n, r_min, r_max, w");
        for (int i = 0; i < (int)p[0]; ++i) {
            const double u = 2. * rng_uniform(seed, 204, i) - 1.;
            const double ph = 2. * M_PI * rng_uniform(seed, 205, i);
            const double st = std::sqrt(std::max(0., 1. - u * u));
            const double row[8] = {
                N0 * rng_uniform(seed, 200, i), N0 * rng_uniform(seed, 201, i),
                N0 * rng_uniform(seed, 202, i),
                p[1] + (p[2] - p[1]) * rng_uniform(seed, 203, i),
                st * std::cos(ph), st * std::sin(ph), u, 0.};
            obj.insert(obj.end(), row, row + 8);
        }
    } else if (kind != kInitComplexRandomPhase && kind != kInitZero) {
        throw Error(1, "unknown initial-state kind");
    }
    if (kind == kInitAddBubbles) {
        // Only the bubbles nucleated in (t_from, t_to] change the level
        // (the kernel tests the same window): pass those alone, and leave
        // the level untouched when there are none.
        std::vector<double> now;
        for (size_t o = 0; o < obj.size(); o += 8)
            if (obj[o + 3] > p[6] && obj[o + 3] <= p[7])
                now.insert(now.end(), obj.begin() + o, obj.begin() + o + 8);
        if (now.empty()) return;
        obj.swap(now);
    }
    DevMem dobj;
    if (!obj.empty()) upload_vec(dobj, obj);
    a.obj = obj.empty() ? nullptr : dobj.as<double>();
    a.nobj = (int)(obj.size() / 8);
    const int grid = (int)std::min<long long>((a.total + 255) / 256, 8LL * ncu_);
    switch (sprec()) {
        case Prec::F64:
            hipLaunchKernelGGL(init_kernel<Prec::F64>, dim3(grid), dim3(256), 0,
                               stream_, a);
            break;
        case Prec::F32C:
            hipLaunchKernelGGL(init_kernel<Prec::F32C>, dim3(grid), dim3(256), 0,
                               stream_, a);
            break;
        case Prec::F32C16:
            hipLaunchKernelGGL(init_kernel<Prec::F32C16>, dim3(grid), dim3(256),
                               0, stream_, a);
            break;
    }
    SHM_HIP(hipGetLastError());
    if (dobj) SHM_HIP(hipStreamSynchronize(stream_));  // table freed on return
    ls.ghosts_ok[ls.cur] = false;
}

void Runtime::sync() { SHM_HIP(hipStreamSynchronize(stream_)); }

void Runtime::stream_probe(size_t bytes, int reps, double *read_gbs,
                           double *write_gbs, double *copy_gbs) {
    bytes = bytes / 16 * 16;
    const long long n = (long long)(bytes / 16);
    DevMem A, B, O;
    A.alloc(bytes);
    B.alloc(bytes);
    O.alloc(256 * sizeof(float));
    const long long g64 = (n + 255) / 256;
    if (g64 > 0x7fffffffLL) throw Error(1, "stream probe buffer too large");
    const int grid = (int)g64;
    auto timed = [&](auto &&launch) {
        launch();
        SHM_HIP(hipEventRecord(ev0_, stream_));
        for (int r = 0; r < reps; ++r) launch();
        SHM_HIP(hipEventRecord(ev1_, stream_));
        SHM_HIP(hipEventSynchronize(ev1_));
        float ms = 0;
        SHM_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
        return ms / reps;
    };
    const double r_ms = timed([&] {
        hipLaunchKernelGGL(stream_read_kernel, dim3(grid), dim3(256), 0, stream_,
                           A.as<sf4>(), n, O.as<float>());
    });
    const double w_ms = timed([&] {
        hipLaunchKernelGGL(stream_write_kernel, dim3(grid), dim3(256), 0, stream_,
                           B.as<sf4>(), n, 1.f);
    });
    const double c_ms = timed([&] {
        hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, stream_,
                           A.as<const sf4>(), B.as<sf4>(), n);
    });
    SHM_HIP(hipGetLastError());
    if (read_gbs) *read_gbs = bytes / (r_ms * 1e-3) / 1e9;
    if (write_gbs) *write_gbs = bytes / (w_ms * 1e-3) / 1e9;
    if (copy_gbs) *copy_gbs = 2. * bytes / (c_ms * 1e-3) / 1e9;
}

/** Event pair for one timed launch, from a pool that grows on demand. */
Runtime::EventPair *Runtime::timed_pair(int kind) {
    if (ev_used_ == ev_pool_.size()) {
        EventPair p{};
        SHM_HIP(hipEventCreate(&p.a));
        SHM_HIP(hipEventCreate(&p.b));
        ev_pool_.push_back(p);
    }
    EventPair *p = &ev_pool_[ev_used_++];
    p->kind = kind;
    return p;
}

/** Resolve recorded launch timings into `timing` (after a synchronisation). */
void Runtime::collect_timing() {
    if (ev_used_ == 0) return;
    sync();
    for (size_t i = 0; i < ev_used_; ++i) {
        float ms = 0;
        SHM_HIP(hipEventElapsedTime(&ms, ev_pool_[i].a, ev_pool_[i].b));
        if (ev_pool_[i].kind != 1) {  // 0: stage, 2: fused stages one and two
            timing.stage_ms += ms;
            timing.stage_launches += 1;
            if (ev_pool_[i].kind == 2) {
                timing.fused_ms += ms;
                timing.fused_launches += 1;
            }
        } else {
            timing.halo_ms += ms;
            timing.halo_launches += 1;
        }
    }
    ev_used_ = 0;
}

// ---------------------------------------------------------------------------
// Autotuner: time the first stage of the configured integrator (it writes
// only stage scratch) for each tile variant and a few z-chunk lengths.
// ---------------------------------------------------------------------------

/** Times the stage kernels of a level over the launch shapes (tile variant,
 *  z-chunk length, persistent or one tile per workgroup, tile order, XCD
 *  split) and keeps the fastest for the level.  Only kernels that write
 *  stage scratch run (SSPRK3 stages 1 and 2, stage 3 into a scratch state,
 *  the first RKN stage), so the level's state is unchanged.  Fields of
 *  `launch` that are set are kept. */
void Runtime::autotune(int lev) {
    LevelState &ls = level(lev);
    const int old = ls.cur;
    fill_patch(ls, old, lev, ls.t);
    const double dt = sim_->dt[lev], dx = sim_->dx[lev];
    std::vector<double> blk(kPrmBlock);
    fill_params(blk.data(), ls.t, dt, dx, lev);
    const double *prm = push_params(prm_host_, prm_ev_, prm_slot_,
                                    prm_dev_.as<double>(), stream_, blk);
    std::vector<std::pair<int, StageArgs>> runs;
    StageArgs a{};
    a.prm = prm;
    if (integrator_ == 10) {
        StageArgs s1 = a;
        s1.x = ls.U[old].get();
        s1.xlo = comp() ? ls.Ulo[old].get() : nullptr;
        s1.kout = ls.K[0].get();
        runs.push_back({kSsprk1, s1});
        StageArgs s2 = a;
        s2.u0 = ls.U[old].get();
        s2.u0lo = comp() ? ls.Ulo[old].get() : nullptr;
        s2.kin = ls.K[0].get();
        s2.kout = ls.K[1].get();
        s2.c[0] = 1.;
        s2.c[2] = dt;
        runs.push_back({kSsprk2, s2});
    } else if (integrator_ >= 20) {
        // Every stage of a step: the stages differ in how many arrays they
        // read (the last one composes phi0, pi0 and every earlier force at
        // each stencil point), so the shape is chosen on their sum.  The
        // final stage is added below, with a scratch target.
        const int s = (int)nodes_.size();
        if (s < 2) return;
        for (int i = 0; i + 1 < s; ++i) {
            StageArgs r = a;
            r.u0 = ls.U[old].get();
            r.u0lo = comp() ? ls.Ulo[old].get() : nullptr;
            r.c[0] = dt * nodes_[i];
            r.nprev = i;
            for (int j = 0; j < i; ++j) {
                r.c[1 + j] = dt * dt * tableau_[i][j];
                r.f[j] = ls.F[j].get();
            }
            r.fout = ls.F[i].get();
            int kind = kRkn0 + i;
            if (rkn_y(lev) && i < 2) {
                kind = i == 0 ? kRkn0Y : kRkn1Y;
                r.y = ls.Y2.get();
                r.c[18] = dt * nodes_[1];
                r.c[19] = dt * dt * tableau_[1][0];
                if (i == 1 && s == 3 && gw_) {
                    kind = kRkn1YY;
                    r.y2 = ls.Y3.get();
                    r.c[12] = dt * nodes_[2];
                    r.c[13] = dt * dt * tableau_[2][0];
                    r.c[14] = dt * dt * tableau_[2][1];
                }
            }
            runs.push_back({kind, r});
        }
    } else {
        return;  // leapfrog kick 1 writes u_new fields; keep the default
    }
    int maxn = 1;
    for (const Box3 &b : ls.L.boxes) maxn = std::max(maxn, b.n(2));
    auto choices = [](int user, std::vector<int> all) {
        return user >= 0 ? std::vector<int>{user} : all;
    };
    std::vector<int> zs = {0};
    for (int z = 8; z < maxn; z *= 2) zs.push_back(z);
    zs.push_back(maxn);
    if (launch.tz > 0) zs = {launch.tz};
    std::vector<int> vs;
    for (int v = 0; v < kNVar; ++v)
        if (kt_.stage[runs[0].first][(int)prec_][R_][gw_ ? 1 : 0][ko_ ? 1 : 0][v].fn)
            vs.push_back(v);
    vs = choices(launch.variant, vs);
    const std::vector<int> ps = choices(launch.persistent, {1, 0});
    const std::vector<int> os = choices(launch.zc_minor, {0, 1});
    const std::vector<int> xs = choices(launch.xcd_split, {1, 0});

    // Final stages write the new state into a scratch target, so the
    // level's states are untouched.  SSPRK3: stage three is timed per shape
    // too; it is the stage kernel that remains when stages one and two run
    // fused.
    DevMem t3, t3lo;
    StageArgs s3{};
    const bool tune3 = integrator_ == 10;
    if (tune3 || integrator_ >= 20) {
        t3.alloc((size_t)ls.L.cstride * ls.ncomp * esize());
        if (comp()) t3lo.alloc((size_t)ls.L.cstride * ls.ncomp * losize());
    }
    if (integrator_ >= 20) {
        const int i = (int)nodes_.size() - 1;
        StageArgs r = a;
        r.u0 = ls.U[old].get();
        r.u0lo = comp() ? ls.Ulo[old].get() : nullptr;
        r.c[0] = dt * nodes_[i];
        r.nprev = i;
        for (int j = 0; j < i; ++j) {
            r.c[1 + j] = dt * dt * tableau_[i][j];
            r.f[j] = ls.F[j].get();
            r.c[7 + j] = dt * dt * bbar_[j];
            r.c[12 + j] = dt * b_[j];
        }
        r.c[6] = dt;
        r.c[11] = dt * dt * bbar_[i];
        r.c[16] = dt * b_[i];
        r.unew = t3.get();
        r.unewlo = comp() ? t3lo.get() : nullptr;
        int kind = kRknF0 + i;
        if (rkn_y(lev) && i == 2 && gw_) {
            kind = kRknF2Y;
            r.y2 = ls.Y3.get();
        }
        runs.push_back({kind, r});
    }
    if (tune3) {
        s3 = a;
        s3.u0 = ls.U[old].get();
        s3.u0lo = comp() ? ls.Ulo[old].get() : nullptr;
        s3.kin = ls.K[1].get();
        s3.unew = t3.get();
        s3.unewlo = comp() ? t3lo.get() : nullptr;
        s3.c[0] = 0.25;
        s3.c[1] = dt * 2. / 3.;
        s3.c[2] = dt / 4.;
    }
    auto time_ms = [&](auto &&fn, int reps = 3) {
        fn();  // warm-up
        SHM_HIP(hipEventRecord(ev0_, stream_));
        for (int k = 0; k < reps; ++k) fn();
        SHM_HIP(hipEventRecord(ev1_, stream_));
        SHM_HIP(hipEventSynchronize(ev1_));
        float ms = 0;
        SHM_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
        return (double)ms / reps;
    };

    const LaunchCfg saved = launch;
    const bool saved_t = timing_on;
    timing_on = false;
    double best = 1e30, best3 = 1e30;
    LaunchCfg bc, bc3;
    // Screening times every shape briefly; the few fastest are then timed
    // again at length, since among hundreds of short, noisy timings the
    // smallest is often a mediocre shape that was timed lucky.
    std::vector<std::pair<double, LaunchCfg>> seen, seen3;
    auto run_all = [&] {
        for (auto &r : runs) {
            StageArgs c = r.second;
            launch_stage(r.first, ls, c, lev);
        }
    };
    auto run3 = [&] {
        StageArgs c = s3;
        launch_stage(kSsprk3, ls, c, lev);
    };
    for (int v : vs)
        for (int z : zs)
            for (int p : ps)
                for (int o : os)
                    for (int x : xs) {
                        if (o == 1 && (z == 0 || z >= maxn)) continue;  // one chunk
                        launch.variant = v;
                        launch.tz = z;
                        launch.persistent = p;
                        launch.zc_minor = o;
                        launch.xcd_split = x;
                        double ms = time_ms(run_all);
                        if (tune3) {
                            const double m3 = time_ms(run3);
                            seen3.push_back({m3, launch});
                            ms += m3;
                        }
                        seen.push_back({ms, launch});
                    }
    auto by_time = [](const std::pair<double, LaunchCfg> &x,
                      const std::pair<double, LaunchCfg> &y) { return x.first < y.first; };
    constexpr int kFinalists = 4, kLongReps = 12;
    std::stable_sort(seen.begin(), seen.end(), by_time);
    std::stable_sort(seen3.begin(), seen3.end(), by_time);
    for (int i = 0; i < (int)seen.size() && i < kFinalists; ++i) {
        launch = seen[i].second;
        double ms = time_ms(run_all, kLongReps);
        if (tune3) ms += time_ms(run3, kLongReps);
        if (ms < best) {
            best = ms;
            bc = launch;
        }
    }
    for (int i = 0; i < (int)seen3.size() && i < kFinalists; ++i) {
        launch = seen3[i].second;
        const double m3 = time_ms(run3, kLongReps);
        if (m3 < best3) {
            best3 = m3;
            bc3 = launch;
        }
    }
    // On a level of periodic boxes: one fused launch for the first two
    // stages plus the remaining stages (SSPRK3: stage three at its best
    // shape; RKN: the later stages at the best unfused shape), against the
    // unfused step.  Each candidate is timed exactly as a step runs it (one
    // z chunk length for all its launches), with that chunk length or with
    // whole columns, which spare the fused pass its 2R extra planes a chunk.
    bc.fused = 0;
    bc.fused_variant = -1;
    if (fused_possible(ls) && saved.fused != 0 && (integrator_ == 10 || integrator_ >= 20)) {
        const FusedEntry *row = fused_row();
        const LaunchCfg base = integrator_ == 10 ? bc3 : bc;
        for (int v = 0; v < kNVar; ++v) {
            if (!row[v].fn) continue;
            if (saved.fused_variant >= 0 && v != saved.fused_variant) continue;
            for (int tzc : {base.tz, 0}) {
                launch = base;
                launch.tz = tzc;
                launch.fused = 1;
                launch.fused_variant = v;
                const double ms = time_ms(
                    [&] {
                        launch_fused(ls, prm, prm, dt, lev);
                        if (integrator_ == 10) {
                            run3();
                        } else {
                            for (size_t k = 2; k < runs.size(); ++k) {
                                StageArgs c = runs[k].second;
                                launch_stage(runs[k].first, ls, c, lev);
                            }
                        }
                    },
                    kLongReps);
                if (ms < best) {
                    best = ms;
                    bc = launch;
                }
                if (base.tz == 0) break;  // one chunk length to try
            }
        }
    }
    launch = saved;
    timing_on = saved_t;
    tuned_[lev + 1] = bc;
    bump_epoch();
}

// ---------------------------------------------------------------------------
// Gravitational-wave spectrum (GravitationalWaves::ComputeSpectrum).
// ---------------------------------------------------------------------------

static double index_to_k(int a, int N, int type) {
    const int nt = a >= N / 2 ? a - N : a;
    const double th = 2. * M_PI * nt / N;
    switch (type) {
        case 1:
            return std::sin(th);
        case 2:
            return (8. * std::sin(th) - std::sin(2. * th)) / 6.;
        case 3:
            return (45. * std::sin(th) - 9. * std::sin(2. * th) +
                    std::sin(3. * th)) / 30.;
    }
    return 0.;
}

void Runtime::release_gw_cache() {
    if (gwc_.have_plan) (void)hipfftDestroy(gwc_.plan);
    gwc_ = GwSpecCache();
}

/** GravitationalWaves::ComputeSpectrum (gravitational_waves.cpp:52-322):
 *  the six selected components of level 0, zero padded, through hipFFT D2Z,
 *  the project's Fourier-space modification per mode, and the Lambda
 *  projected sum per bin, divided by N^6.  The plan, the buffers and the k
 *  tables stay with the context for the next call of the same size. */
void Runtime::gw_spectrum(int mode, int nbins, int zp, int proj,
                          std::vector<double> &bins, std::vector<double> &kvals,
                          int modifier, const std::vector<double> &mparams,
                          bool count_only) {
    if (independent_)
        throw Error(4, "the GW spectrum needs one periodic domain, not independent boxes");
    if (!gw_) throw Error(3, "gravitational waves are not enabled");
    if (proj < 1 || proj > 3) throw Error(1, "projection_type must be 1, 2 or 3");
    if (zp < 1 || !sledgehamr::utils::IsPowerOfTwo(zp))
        throw Error(1, "zero padding must be a power of 2");
    if (mode < 0 || mode > 2) throw Error(1, "spectrum mode must be 0, 1 or 2");
    if (!kt_.gw_bin) throw Error(4, "GW spectrum kernel not built");
    LevelState &l0 = level(0);
    const int N = sim_->coarse_level_grid_size * zp;
    const int NH = N / 2 + 1;
    const int NS = kt_.nscalars;
    // GravitationalWavesSpectrumModifier::SelectComponents
    // (gravitational_waves.h:79-92): du_xx, du_xy, du_xz, du_yy, du_yz, du_zz
    // unless the project's modifier names others.
    int comps[6] = {NS + 6, NS + 9, NS + 10, NS + 7, NS + 11, NS + 8};
    if (modifier >= 0) {
        if (modifier >= kt_.ngwmod || !kt_.gwmod_select(modifier, comps))
            throw Error(1, "no such GW spectrum modifier");
        for (int c = 0; c < 6; ++c)
            if (comps[c] < 0 || comps[c] >= ncomp_)
                throw Error(1, "GW spectrum modifier selects a component out of range");
    }
    GwSpecCache &w = gwc_;
    if (w.N != N) {
        release_gw_cache();
        w.N = N;
        w.dense.alloc(sizeof(double) * (size_t)N * N * N);
        for (int c = 0; c < 6; ++c) w.out[c].alloc(sizeof(double2) * (size_t)NH * N * N);
        if (hipfftPlan3d(&w.plan, N, N, N, HIPFFT_D2Z) != HIPFFT_SUCCESS)
            throw Error(2, "hipfftPlan3d failed");
        w.have_plan = true;
    }
    if (hipfftSetStream(w.plan, stream_) != HIPFFT_SUCCESS)
        throw Error(2, "hipfftSetStream failed");
    if (w.proj != proj) {
        std::vector<double> itk(N);
        for (int a = 0; a < N; ++a) itk[a] = index_to_k(a, N, proj);
        upload_vec(w.ditk, itk);
        w.proj = proj;
    }
    if (w.mode != mode || (mode == 2 && w.nbins != nbins)) {
        w.kvals.clear();
        if (mode == 0) {
            // Unique |n|^2 over the signed frequencies: every |f| in 0..N/2
            // occurs on each axis, so mark f1^2 + f2^2 + f3^2 for
            // 0 <= f1 <= f2 <= f3 <= N/2.
            const long long h = N / 2, maxsq = 3 * h * h;
            std::vector<char> seen(maxsq + 1, 0);
            for (long long a = 0; a <= h; ++a)
                for (long long b = a; b <= h; ++b)
                    for (long long c = b; c <= h; ++c) seen[a * a + b * b + c * c] = 1;
            std::vector<int> sq2bin(maxsq + 1, -1);
            int i = 0;
            for (long long v = 0; v <= maxsq; ++v)
                if (seen[v]) {
                    sq2bin[v] = i++;
                    w.kvals.push_back((double)v);
                }
            upload_vec(w.dsq, sq2bin);
            w.nbin_eff = i;
        } else if (mode == 1) {
            int nb = 0;
            for (int k = 0; k <= std::sqrt(3.) / 2. * (double)N + .5; ++k) {
                w.kvals.push_back((double)k * k);
                ++nb;
            }
            w.nbin_eff = nb;
        } else {
            if (nbins <= 0) throw Error(1, "log binning needs nbins > 0");
            const double lk = std::log(std::sqrt(3.) * N / 2.);
            for (int i = 0; i < nbins; ++i) w.kvals.push_back(std::exp(lk * i / nbins));
            w.nbin_eff = nbins;
        }
        w.dbins.alloc(sizeof(double) * std::max(1, w.nbin_eff));
        w.mode = mode;
        w.nbins = nbins;
    }
    kvals = w.kvals;
    const int nb_eff = w.nbin_eff;
    if (count_only) {
        bins.assign(nb_eff, 0.);
        return;
    }

    // Zero padding: the level fills the corner [0, N/zp)^3 of the dense
    // array (its boxes tile it completely); the rest keeps the zeros of its
    // allocation.
    for (int c = 0; c < 6; ++c) {
        for (size_t b = 0; b < l0.L.boxes.size(); ++b) {
            PackArgs p{};
            p.arr = l0.U[l0.cur].get();
            p.arrlo = comp() ? l0.Ulo[l0.cur].get() : nullptr;
            p.cstride = l0.L.cstride;
            p.box = l0.L.hbox[b];
            p.comp = comps[c];
            p.out = w.dense.as<double>();
            p.N = N;
            const long long nb = l0.L.boxes[b].vol();
            const int grid = (int)std::min<long long>((nb + 255) / 256, 8LL * ncu_);
            switch (sprec()) {
                case Prec::F64:
                    hipLaunchKernelGGL(pack_kernel<Prec::F64>, dim3(grid),
                                       dim3(256), 0, stream_, p);
                    break;
                case Prec::F32C:
                    hipLaunchKernelGGL(pack_kernel<Prec::F32C>, dim3(grid),
                                       dim3(256), 0, stream_, p);
                    break;
                default:
                    hipLaunchKernelGGL(pack_kernel<Prec::F32C16>, dim3(grid),
                                       dim3(256), 0, stream_, p);
                    break;
            }
            SHM_HIP(hipGetLastError());
        }
        if (hipfftExecD2Z(w.plan, w.dense.as<hipfftDoubleReal>(),
                          w.out[c].as<hipfftDoubleComplex>()) != HIPFFT_SUCCESS)
            throw Error(2, "hipfftExecD2Z failed");
    }
    std::vector<double> mp = mparams;
    mp.resize(std::max<size_t>(mp.size(), 1), 0.);
    upload_into(w.dparams, mp);
    // The bins are summed with atomics: clear them with a kernel-free copy
    // of zeros, outside any captured sequence.
    upload_into(w.dbins, std::vector<double>(std::max(1, nb_eff), 0.));
    BinArgs ba{};
    for (int c = 0; c < 6; ++c) ba.u[c] = w.out[c].as<double2>();
    ba.N = N;
    ba.index_to_k = w.ditk.as<double>();
    ba.mode = mode;
    ba.sq_to_bin = mode == 0 ? w.dsq.as<int>() : nullptr;
    ba.nbins = nb_eff;
    ba.log_kmax = std::log(std::sqrt(3.) * N / 2.);
    ba.bins = w.dbins.as<double>();
    ba.modifier = modifier;
    ba.mparams = w.dparams.as<double>();
    ba.dk = 2. * M_PI / (sim_->L * zp);
    const long long total = (long long)NH * N * N;
    const int grid = (int)std::min<long long>((total + 255) / 256, 4LL * ncu_);
    SHM_HIP(kt_.gw_bin(ba, grid, stream_));
    bins.assign(nb_eff, 0.);
    SHM_HIP(hipMemcpyAsync(bins.data(), w.dbins.get(), sizeof(double) * nb_eff,
                           hipMemcpyDeviceToHost, stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
    const double n6 = std::pow((double)N, 6);
    for (double &v : bins) v /= n6;
}

// ---------------------------------------------------------------------------
// Output functions: spectra (output_types/spectrum.cpp:12-113) and
// projections (output_types/projection.cpp:11-118).
// ---------------------------------------------------------------------------

void Runtime::check_output_fn(int fn) const {
    if (kt_.nout == 0)
        throw Error(4, "the project declares no output functions "
                       "(SLEDGEHAMR_OUTPUT_FUNCTIONS)");
    if (fn < 0 || fn >= kt_.nout) throw Error(1, "output function id out of range");
    if (independent_)
        throw Error(4, "spectra and projections need one periodic domain, not "
                       "independent boxes");
}

void Runtime::fill_output_ghosts(int lev) {
    // FillPatch at the level's time, with every ghost cell written (boxes
    // spanning the domain included, whose stage kernels wrap instead) and
    // the remainders copied too.
    LevelState &ls = level(lev);
    void *lo = comp() ? ls.Ulo[ls.cur].get() : nullptr;
    if (has_cf(lev)) cf_fill_state(lev, ls.t, ls.U[ls.cur].get(), lo);
    halo(ls, ls.U[ls.cur].get(), all_comps(), true);
    if (comp()) halo(ls, lo, all_comps(), true, (int)losize());
}

/** Sorted unique |n|^2 over the signed frequencies of an N^3 grid (the
 *  k-table of ReadSpectrumKs, sledgehamr.cpp:554-601): every |f| in 0..N/2
 *  occurs on each axis, so mark f1^2 + f2^2 + f3^2 for 0 <= f1 <= f2 <= f3
 *  <= N/2. */
static std::vector<long long> unique_k2(int N) {
    const long long h = N / 2, maxsq = 3 * h * h;
    std::vector<char> seen(maxsq + 1, 0);
    for (long long a = 0; a <= h; ++a)
        for (long long b = a; b <= h; ++b)
            for (long long c = b; c <= h; ++c) seen[a * a + b * b + c * c] = 1;
    std::vector<long long> v;
    for (long long x = 0; x <= maxsq; ++x)
        if (seen[x]) v.push_back(x);
    return v;
}

void Runtime::spectrum(int fn, std::vector<double> &bins, std::vector<long long> &k_sq,
                       bool count_only) {
    check_output_fn(fn);
    LevelState &l0 = level(0);
    const int N = l0.L.N, NH = N / 2 + 1;
    OutSpecCache &w = osc_;
    if (w.N != N) {
        release_out_cache();
        w.N = N;
        w.k_sq = unique_k2(N);
        const long long maxsq = 3LL * (N / 2) * (N / 2);
        std::vector<int> sq2bin(maxsq + 1, -1);
        for (size_t b = 0; b < w.k_sq.size(); ++b) sq2bin[w.k_sq[b]] = (int)b;
        upload_vec(w.dsq, sq2bin);
        w.dense.alloc(sizeof(double) * (size_t)N * N * N);
        w.F.alloc(sizeof(double2) * (size_t)NH * N * N);
        w.dbins.alloc(sizeof(double) * w.k_sq.size());
        if (hipfftPlan3d(&w.plan, N, N, N, HIPFFT_D2Z) != HIPFFT_SUCCESS)
            throw Error(2, "hipfftPlan3d failed");
        w.have_plan = true;
    }
    k_sq = w.k_sq;
    const int nbins = (int)k_sq.size();
    if (count_only) {
        bins.assign(nbins, 0.);
        return;
    }
    fill_output_ghosts(0);
    std::vector<double> p;
    sim_->SetParamsSpectra(p, l0.t);
    p.resize(std::max<size_t>(p.size(), 1), 0.);
    upload_into(w.dparams, p);
    SHM_HIP(hipMemsetAsync(w.dense.get(), 0, w.dense.bytes(), stream_));
    SHM_HIP(hipMemsetAsync(w.dbins.get(), 0, w.dbins.bytes(), stream_));
    OutArgs a{};
    a.boxes = l0.L.dbox.as<BoxDesc>();
    a.vprefix = l0.L.dvprefix.as<long long>();
    a.nbox = (int)l0.L.boxes.size();
    a.total = l0.L.nvalid();
    a.cstride = l0.L.cstride;
    a.ncomp = ncomp_;
    a.arr = l0.U[l0.cur].get();
    a.arrlo = comp() ? l0.Ulo[l0.cur].get() : nullptr;
    a.fn = fn;
    a.params = w.dparams.as<double>();
    a.lev = 0;
    a.time = l0.t;
    a.dt = sim_->dt[0];
    a.dx = sim_->dx[0];
    a.target = 0;
    a.dense = w.dense.as<double>();
    a.N = N;
    const int grid = (int)std::min<long long>((a.total + 255) / 256, 8LL * ncu_);
    SHM_HIP(kt_.output[(int)sprec()](a, grid, stream_));
    if (hipfftSetStream(w.plan, stream_) != HIPFFT_SUCCESS)
        throw Error(2, "hipfftSetStream failed");
    if (hipfftExecD2Z(w.plan, w.dense.as<hipfftDoubleReal>(),
                      w.F.as<hipfftDoubleComplex>()) != HIPFFT_SUCCESS)
        throw Error(2, "hipfftExecD2Z failed");
    SpecBinArgs sb{};
    sb.F = w.F.as<double2>();
    sb.N = N;
    sb.sq_to_bin = w.dsq.as<int>();
    sb.nbins = nbins;
    sb.bins = w.dbins.as<double>();
    const long long modes = (long long)NH * N * N;
    const int bgrid = (int)std::min<long long>((modes + 255) / 256, 4LL * ncu_);
    hipLaunchKernelGGL(scalar_bin_kernel<2048>, dim3(bgrid), dim3(256), 0, stream_, sb);
    SHM_HIP(hipGetLastError());
    bins.assign(nbins, 0.);
    SHM_HIP(hipMemcpyAsync(bins.data(), w.dbins.get(), sizeof(double) * nbins,
                           hipMemcpyDeviceToHost, stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
    // pre = t / (dk N^6), dk = 2 pi / L (spectrum.cpp:49-51).
    const double pre = std::pow(1. / N, 6) * l0.t / (2. * M_PI / sim_->L);
    for (double &v : bins) v *= pre;
}

void Runtime::release_out_cache() {
    if (osc_.have_plan) (void)hipfftDestroy(osc_.plan);
    osc_ = OutSpecCache();
}

void Runtime::projection(int fn, int mode, int max_level, std::vector<double> &data,
                         std::vector<int> &count, int &D, bool count_only) {
    check_output_fn(fn);
    if (mode != 0 && mode != 1) throw Error(1, "projection mode must be 0 (sum) or 1 (max)");
    int ml = max_level < 0 ? INT32_MAX : max_level;
    int finest = 0;
    while (has_level(finest + 1)) ++finest;
    ml = std::min(ml, finest);
    D = sim_->dimN[ml];
    if (count_only) return;
    std::vector<double> p;
    sim_->SetParamsProjections(p, level(0).t);
    p.resize(std::max<size_t>(p.size(), 1), 0.);
    DevMem dp, dproj, dn;
    upload_vec(dp, p);
    dproj.alloc(sizeof(double) * (size_t)D * D);
    dn.alloc(sizeof(int) * (size_t)D * D);
    SHM_HIP(hipMemsetAsync(dproj.get(), 0, dproj.bytes(), stream_));
    SHM_HIP(hipMemsetAsync(dn.get(), 0, dn.bytes(), stream_));
    std::vector<DevMem> keep;
    for (int lev = 0; lev <= ml; ++lev) {
        LevelState &ls = level(lev);
        fill_output_ghosts(lev);
        // Per box, the finer boxes that may hold fine cell (2i, 2j, 2k).
        std::vector<int> start(1, 0), cov;
        for (const Box3 &b : ls.L.boxes) {
            if (lev < ml)
                for (const Box3 &f : level(lev + 1).L.boxes) {
                    bool hit = true;
                    for (int d = 0; d < 3; ++d)
                        hit = hit && 2 * b.lo[d] <= f.hi[d] && 2 * b.hi[d] >= f.lo[d];
                    if (hit)
                        cov.insert(cov.end(), {f.lo[0], f.lo[1], f.lo[2], f.hi[0], f.hi[1], f.hi[2]});
                }
            start.push_back((int)(cov.size() / 6));
        }
        if (cov.empty()) cov.assign(6, 0);
        keep.emplace_back();
        upload_vec(keep.back(), start);
        const int *dstart = keep.back().as<int>();
        keep.emplace_back();
        upload_vec(keep.back(), cov);
        OutArgs a{};
        a.boxes = ls.L.dbox.as<BoxDesc>();
        a.vprefix = ls.L.dvprefix.as<long long>();
        a.nbox = (int)ls.L.boxes.size();
        a.total = ls.L.nvalid();
        a.cstride = ls.L.cstride;
        a.ncomp = ncomp_;
        a.arr = ls.U[ls.cur].get();
        a.arrlo = comp() ? ls.Ulo[ls.cur].get() : nullptr;
        a.fn = fn;
        a.params = dp.as<double>();
        a.lev = lev;
        a.time = ls.t;
        a.dt = sim_->dt[lev];
        a.dx = sim_->dx[lev];
        a.target = 1;
        a.cov_start = dstart;
        a.cov = keep.back().as<int>();
        a.ratio = D / sim_->dimN[lev];
        a.D = D;
        a.mode = mode;
        a.proj = dproj.as<double>();
        a.nproj = dn.as<int>();
        const int grid = (int)std::min<long long>((a.total + 255) / 256, 8LL * ncu_);
        SHM_HIP(kt_.output[(int)sprec()](a, grid, stream_));
    }
    data.assign((size_t)D * D, 0.);
    count.assign((size_t)D * D, 0);
    SHM_HIP(hipMemcpyAsync(data.data(), dproj.get(), dproj.bytes(), hipMemcpyDeviceToHost, stream_));
    SHM_HIP(hipMemcpyAsync(count.data(), dn.get(), dn.bytes(), hipMemcpyDeviceToHost, stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
}

// ---------------------------------------------------------------------------
// Time stepper (TimeStepper::Advance, time_stepper.cpp:90-311) for a static
// box layout: the schedule decides when truncation errors are computed; the
// regrid itself is replaced by tagging (ErrorEst) on the current layout.
// ---------------------------------------------------------------------------

const std::vector<long long> &Runtime::last_tags(int lev) const {
    auto it = last_tags_.find(lev);
    if (it == last_tags_.end())
        throw Error(3, "no scheduled truncation-error tagging at this level yet");
    return it->second;
}

void Runtime::step(int ncoarse) {
    for (int n = 0; n < ncoarse; ++n) scheduled_step(0);
}

void Runtime::scheduled_step(int lev) {
    sledgehamr::Sledgehamr &S = *sim_;
    LevelState &ls = level(lev);
    const double time = ls.t;
    const int istep = ls.istep;
    if (S.shadow_hierarchy && lev < S.finest_level) {
        auto find = [&](double t) {
            for (size_t i = 0; i < schedule_.size(); ++i)
                if (schedule_[i].second == t) return (int)i;
            return -1;
        };
        const double tnext = time + S.dt[lev];
        int id = find(tnext);
        bool already = id >= 0 && schedule_[id].first == lev;
        bool ok = !already && ls.t_old != time && lev < S.max_level &&
                  !(istep % 2 == 0 && lev > 0);
        const double opp = lev > 0 ? time + 3. * S.dt[lev] : time + 2. * S.dt[lev];
        if (ok && !S.DoCreateLevelIf(lev + 1, opp)) ok = false;
        if (ok && opp <= last_regrid_[lev] + regrid_dt_[lev] && !lr_.do_global[lev])
            ok = false;
        if (ok && lev == 0 && ls.t_old == -DBL_MAX) ok = false;
        if (ok) {
            if (id < 0)
                schedule_.emplace_back(lev, tnext);
            else
                schedule_[id].first = std::min(schedule_[id].first, lev);
            events_.push_back({'S', lev, tnext});
            if (lev == 0) {
                create_shadow_level();
                events_.push_back({'H', -1, shadow_.t});
            }
        }
    }
    if (lev == 0 && !shadow_.defined()) S.BeforeTimestep(ls.t);
    advance(lev);
    events_.push_back({'A', lev, ls.t});
    if (lev < S.finest_level) {
        scheduled_step(lev + 1);
        scheduled_step(lev + 1);
    }
    synchronize_levels(lev);
    // DoRegridIfScheduled (time_stepper.cpp:287-296, DoRegrid :346-390): the
    // local regrid when shm.local_regrid is set.  Where the reference falls
    // back to a global regrid (AMReX clustering, not part of this library)
    // the evolve stops with a status, the regrid done with: the caller runs
    // its global regrid and remakes the levels (shm_remake_level).  With
    // shm.keep_layout_on_global_regrid = 1 the layout stays instead and the
    // local regrid's counters are left as they are (no regrid ran).
    for (size_t i = 0; i < schedule_.size(); ++i) {
        if (schedule_[i].second == ls.t && schedule_[i].first == lev) {
            events_.push_back({'R', lev, ls.t});
            const bool global = local_regrid_on_ && local_regrid(lev) == 0;
            for (int l = lev; l <= S.finest_level; ++l) last_regrid_[l] = ls.t;
            schedule_.erase(schedule_.begin() + i);
            if (global) {
                ++lr_.global_wanted;
                if (!keep_layout_on_global_)
                    throw Error(4, "a global regrid of level " + std::to_string(lev) +
                                       " is wanted at t = " + std::to_string(ls.t) +
                                       " (remake the levels with shm_remake_level, or set "
                                       "shm.keep_layout_on_global_regrid = 1)");
            }
            break;
        }
    }
    if (lev == 0)
        for (int l = 1; l <= S.finest_level; ++l)
            if (has_level(l)) levels_[l].t = levels_[0].t;
}

void Runtime::synchronize_levels(int lev) {
    sledgehamr::Sledgehamr &S = *sim_;
    LevelState &ls = level(lev);
    bool need_te = false;
    for (auto &s : schedule_)
        if (s.second == ls.t && lev >= s.first) need_te = true;
    if (lev < S.finest_level && has_level(lev + 1)) {
        if (!(S.shadow_hierarchy && need_te)) {
            LevelState &fine = level(lev + 1);
            // amrex::average_down onto the coarse boxes covering each fine box.
            LevelState &m = coarse_mirror(lev + 1);
            avg_down(fine, fine.cur, m, 0);
            copy_state(ls.U[ls.cur].get(), ls.Ulo[ls.cur].get(), ls.L, m.U[0].get(),
                       m.Ulo[0].get(), m.L, fine.mplan_out, fine.mnseg_out, fine.mtot_out);
            ls.ghosts_ok[ls.cur] = false;
            events_.push_back({'D', lev, ls.t});
        }
    }
    if (lev >= 1 - (S.shadow_hierarchy ? 1 : 0) && need_te) {
        events_.push_back({'T', lev, ls.t});
        std::vector<long long> counts(2 + ncomp_, 0);
        te_tag(lev, kTeCompute | kTeAvgDown | kTeTags, nullptr, counts.data(), nullptr);
        last_tags_[lev] = counts;
        if (lev == 0) shadow_ = LevelState();
    }
}

// ---------------------------------------------------------------------------
// Local regrid (local_regrid/local_regrid.cpp, unique_layout.cpp).  Blocks
// are blocking-factor cubes of a level, (i, j, k) block indices.
// ---------------------------------------------------------------------------

static int fdiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

void Runtime::check_bf(int lev) const {
    const int v = bf_[lev];
    if (v < 2 || (v & (v - 1)) || sim_->dimN[lev] % v)
        throw Error(1, "amr.blocking_factor must be a power of 2 dividing the level size");
}

std::vector<uint8_t> Runtime::block_cover(int lev) const {
    check_bf(lev);
    // LocalRegrid::WrapIndices (local_regrid.cpp:474-490) tests the cell bf/2
    // into each block: blocks q with lo <= q bf + bf/2 <= hi per axis.
    const int bf = bf_[lev], Np = sim_->dimN[lev] / bf;
    std::vector<uint8_t> cov((size_t)Np * Np * Np, 0);
    if (!has_level(lev)) return cov;
    for (const Box3 &b : levels_[lev].L.boxes) {
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = -fdiv(-(b.lo[a] - bf / 2), bf);
            hi[a] = fdiv(b.hi[a] - bf / 2, bf);
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x)
                    cov[(size_t)x + (size_t)Np * ((size_t)y + (size_t)Np * z)] = 1;
    }
    return cov;
}

void Runtime::regrid_select(int lev, const uint8_t *host_mask, BlockSet &blocks,
                            int &min_dsq) {
    blocks.clear();
    min_dsq = INT_MAX;
    if (lev < 0 || !has_level(lev) || !has_level(lev + 1))
        throw Error(1, "block selection needs the tagging level and the level above it");
    LevelState &ls = level(lev);
    const long long nv = ls.L.nvalid();
    DevMem dmask;
    const uint8_t *mask = nullptr;
    if (host_mask) {
        dmask.alloc((size_t)std::max<long long>(1, nv));
        SHM_HIP(hipMemcpy(dmask.get(), host_mask, (size_t)nv, hipMemcpyHostToDevice));
        mask = dmask.as<uint8_t>();
    } else {
        if (!ls.tags || ls.tags_t != ls.t)
            throw Error(3, "level " + std::to_string(lev) + " has no tags at its time");
        mask = ls.tags.as<uint8_t>();
    }
    const int bf = bf_[lev + 1], Np = sim_->dimN[lev + 1] / bf;
    const std::vector<uint8_t> cov = block_cover(lev + 1);
    std::vector<RegridBox> rb;
    std::vector<uint8_t> border;
    long long voff = 0;
    for (const Box3 &b : ls.L.boxes) {
        if (b.vol() > INT_MAX) throw Error(1, "box too large for the block selection");
        RegridBox r{};
        for (int a = 0; a < 3; ++a) {
            r.lo[a] = b.lo[a];
            r.n[a] = b.n(a);
            r.c0[a] = 2 * b.lo[a] / bf + 1;
            r.xs[a] = 2 * b.hi[a] / bf + 1 - r.c0[a] + 1;
        }
        r.voff = voff;
        r.boff = (long long)border.size();
        // Border map of the box (GetBoxCoarseFineBorders, local_regrid.cpp:
        // 500-530): wrapped blocks with no cell of the finer level.
        for (int k = -1; k <= r.xs[2]; ++k)
            for (int j = -1; j <= r.xs[1]; ++j)
                for (int i = -1; i <= r.xs[0]; ++i) {
                    const int q[3] = {(r.c0[0] + i - 1 + Np) % Np, (r.c0[1] + j - 1 + Np) % Np,
                                      (r.c0[2] + k - 1 + Np) % Np};
                    border.push_back(
                        cov[(size_t)q[0] + (size_t)Np * ((size_t)q[1] + (size_t)Np * q[2])] ? 0
                                                                                           : 1);
                }
        rb.push_back(r);
        voff += b.vol();
    }
    if (nv == 0) return;
    DevMem dboxes, dborder, dfirst, dadded, dmin;
    upload_vec(dboxes, rb);
    upload_vec(dborder, border);
    upload_vec(dfirst, std::vector<int>(border.size(), INT_MAX));
    upload_vec(dmin, std::vector<int>(1, INT_MAX));
    dadded.alloc((size_t)Np * Np * Np);
    RegridArgs a{};
    a.boxes = dboxes.as<RegridBox>();
    a.nbox = (int)rb.size();
    a.nvalid = nv;
    a.mask = mask;
    a.border = dborder.as<uint8_t>();
    a.first = dfirst.as<int>();
    a.added = dadded.as<uint8_t>();
    a.min_dsq = dmin.as<int>();
    a.bf = bf;
    a.Np = Np;
    a.thr = (n_error_buf_ + 1) * (n_error_buf_ + 1);
    const int grid = (int)std::min<long long>((nv + 255) / 256, 8LL * ncu_);
    hipLaunchKernelGGL(regrid_select_kernel<0>, dim3(grid), dim3(256), 0, stream_, a);
    SHM_HIP(hipGetLastError());
    hipLaunchKernelGGL(regrid_select_kernel<1>, dim3(grid), dim3(256), 0, stream_, a);
    SHM_HIP(hipGetLastError());
    std::vector<uint8_t> added((size_t)Np * Np * Np);
    SHM_HIP(hipMemcpyAsync(added.data(), dadded.get(), added.size(), hipMemcpyDeviceToHost,
                           stream_));
    SHM_HIP(hipMemcpyAsync(&min_dsq, dmin.get(), sizeof(int), hipMemcpyDeviceToHost, stream_));
    SHM_HIP(hipStreamSynchronize(stream_));
    for (int x = 0; x < Np; ++x)
        for (int y = 0; y < Np; ++y)
            for (int z = 0; z < Np; ++z)
                if (added[(size_t)x + (size_t)Np * ((size_t)y + (size_t)Np * z)])
                    blocks.insert(Block{x, y, z});
}

std::vector<Box3> Runtime::block_boxes(const BlockSet &s, int bf) {
    std::vector<Box3> out;
    for (auto it = s.begin(); it != s.end();) {
        const Block b0 = *it;
        int km = b0[2];
        for (++it; it != s.end() && (*it)[0] == b0[0] && (*it)[1] == b0[1] &&
                   (*it)[2] == km + 1;
             ++it)
            km = (*it)[2];
        out.push_back(Box3{{b0[0] * bf, b0[1] * bf, b0[2] * bf},
                           {(b0[0] + 1) * bf - 1, (b0[1] + 1) * bf - 1, (km + 1) * bf - 1}});
    }
    return out;
}

std::vector<Box3> Runtime::merge_boxes(std::vector<Box3> v) {
    // Join along axis a boxes whose extents along o1 and o2 agree.
    auto join = [&v](int a, int o1, int o2) {
        std::sort(v.begin(), v.end(), [&](const Box3 &p, const Box3 &q) {
            return std::make_tuple(p.lo[o1], p.hi[o1], p.lo[o2], p.hi[o2], p.lo[a]) <
                   std::make_tuple(q.lo[o1], q.hi[o1], q.lo[o2], q.hi[o2], q.lo[a]);
        });
        std::vector<Box3> out;
        for (const Box3 &b : v) {
            if (!out.empty()) {
                Box3 &p = out.back();
                if (p.lo[o1] == b.lo[o1] && p.hi[o1] == b.hi[o1] && p.lo[o2] == b.lo[o2] &&
                    p.hi[o2] == b.hi[o2] && p.hi[a] + 1 == b.lo[a]) {
                    p.hi[a] = b.hi[a];
                    continue;
                }
            }
            out.push_back(b);
        }
        v.swap(out);
    };
    join(1, 0, 2);
    join(0, 1, 2);
    std::sort(v.begin(), v.end(), [](const Box3 &p, const Box3 &q) {
        return std::make_tuple(p.lo[2], p.lo[1], p.lo[0]) <
               std::make_tuple(q.lo[2], q.lo[1], q.lo[0]);
    });
    return v;
}

void Runtime::fix_nesting(int l, std::vector<BlockSet> &add) const {
    // The runs added to level l grown by nghost + 4 and coarsened by 2
    // (floor), cut at the periodic images of the domain (WrapBoxArray,
    // local_regrid.cpp:735-757); per piece the wrapped block coordinates
    // lo / bf + 1 .. hi / bf + 2 of level l-1 (local_regrid.cpp:687-729).
    const std::vector<Box3> runs = block_boxes(add[l], bf_[l]);
    const int g = sim_->nghost + 4, Nc = sim_->dimN[l - 1];
    const int bfc = bf_[l - 1], Npc = Nc / bfc;
    const std::vector<uint8_t> cov = block_cover(l - 1);
    for (Box3 b : runs) {
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = floordiv2(b.lo[a] - g);
            b.hi[a] = floordiv2(b.hi[a] + g);
        }
        for (int i = -1; i <= 1; ++i)
            for (int j = -1; j <= 1; ++j)
                for (int k = -1; k <= 1; ++k) {
                    const int sh[3] = {i * Nc, j * Nc, k * Nc};
                    int cl[3], ch[3];
                    bool empty = false;
                    for (int a = 0; a < 3; ++a) {
                        const int lo = std::max(b.lo[a], sh[a]) - sh[a];
                        const int hi = std::min(b.hi[a], sh[a] + Nc - 1) - sh[a];
                        if (hi < lo) empty = true;
                        cl[a] = lo / bfc + 1;
                        ch[a] = hi / bfc + 2;
                    }
                    if (empty) continue;
                    for (int cx = cl[0]; cx <= ch[0]; ++cx)
                        for (int cy = cl[1]; cy <= ch[1]; ++cy)
                            for (int cz = cl[2]; cz <= ch[2]; ++cz) {
                                const Block q{(cx - 1) % Npc, (cy - 1) % Npc, (cz - 1) % Npc};
                                if (!cov[(size_t)q[0] +
                                         (size_t)Npc * ((size_t)q[1] + (size_t)Npc * q[2])])
                                    add[l - 1].insert(q);
                            }
                }
    }
}

void Runtime::check_new_boxes(int lev, const std::vector<Box3> &boxes,
                              const std::vector<Box3> &keep) const {
    const int N = sim_->dimN[lev];
    auto overlap = [](const Box3 &p, const Box3 &q) {
        for (int a = 0; a < 3; ++a)
            if (p.hi[a] < q.lo[a] || q.hi[a] < p.lo[a]) return false;
        return true;
    };
    for (size_t i = 0; i < boxes.size(); ++i) {
        const Box3 &b = boxes[i];
        for (int a = 0; a < 3; ++a) {
            if (b.lo[a] < 0 || b.hi[a] >= N || b.hi[a] < b.lo[a])
                throw Error(1, "box outside the level domain");
            if ((b.lo[a] | (b.hi[a] + 1)) & 1)
                throw Error(1, "boxes above level 0 need even corners");
        }
        for (size_t j = 0; j < i; ++j)
            if (overlap(b, boxes[j])) throw Error(1, "new boxes overlap");
        for (const Box3 &k : keep)
            if (overlap(b, k)) throw Error(1, "a new box overlaps a box of the level");
    }
}

void Runtime::relayout(int lev, const std::vector<Box3> &boxes, bool both) {
    LevelState prev = std::move(levels_[lev]);
    LevelState &ls = levels_[lev];
    ls = LevelState();
    bump_epoch();
    if (lev + 1 < (int)levels_.size()) levels_[lev + 1].cf_built = false;
    try {
        ls.L.build(boxes, sim_->nghost, sim_->dimN[lev], /*periodic=*/true);
        ls.ncomp = ncomp_;
        ensure_arrays(ls, integrator_);
        ls.t = prev.t;
        ls.t_old = prev.t_old;
        ls.istep = prev.istep;
        ls.cur = prev.cur;
        long long tot = 0;
        const std::vector<HaloSeg> plan = copy_plan(ls.L, prev.L, &tot);
        DevMem dplan;
        upload_vec(dplan, plan);
        // FillPatch: coarse data everywhere, then the level's own valid data.
        auto fill = [&](int s, double time) {
            fill_coarse_full(lev, s, time, false);
            if (tot == 0) return;
            CopyArgs c{};
            c.dst = ls.U[s].get();
            c.src = prev.U[s].get();
            c.dcstride = ls.L.cstride;
            c.scstride = prev.L.cstride;
            c.dboxes = ls.L.dbox.as<BoxDesc>();
            c.sboxes = prev.L.dbox.as<BoxDesc>();
            c.segs = dplan.as<HaloSeg>();
            c.nseg = (int)plan.size();
            c.total = tot;
            c.ncomp = ncomp_;
            launch_copy(c, esize(), ncu_, stream_);
            if (comp()) {
                c.dst = ls.Ulo[s].get();
                c.src = prev.Ulo[s].get();
                launch_copy(c, losize(), ncu_, stream_);
            }
        };
        fill(ls.cur, ls.t);
        const int od = ls.cur ^ 1;
        if (both && ls.t_old != -DBL_MAX) {
            fill(od, ls.t_old);
        } else {
            const size_t n = (size_t)ls.L.cstride * ncomp_;
            SHM_HIP(hipMemcpyAsync(ls.U[od].get(), ls.U[ls.cur].get(), n * esize(),
                                   hipMemcpyDeviceToDevice, stream_));
            if (comp())
                SHM_HIP(hipMemcpyAsync(ls.Ulo[od].get(), ls.Ulo[ls.cur].get(),
                                       n * losize(), hipMemcpyDeviceToDevice, stream_));
        }
        ls.ghosts_ok[0] = ls.ghosts_ok[1] = false;
        SHM_HIP(hipStreamSynchronize(stream_));
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        levels_[lev] = std::move(prev);
        levels_[lev].cf_built = false;
        bump_epoch();
        throw;
    }
}

void Runtime::add_boxes(int lev, const std::vector<Box3> &boxes) {
    if (lev < 1 || !has_level(lev) || !has_level(lev - 1))
        throw Error(1, "boxes are added to a defined level above 0");
    check_new_boxes(lev, boxes, level(lev).L.boxes);
    if (boxes.empty()) return;
    std::vector<Box3> all = level(lev).L.boxes;
    all.insert(all.end(), boxes.begin(), boxes.end());
    relayout(lev, all, true);
}

void Runtime::remake_level(int lev, const std::vector<Box3> &boxes) {
    if (lev < 1 || !has_level(lev) || !has_level(lev - 1))
        throw Error(1, "a defined level above 0 is remade");
    if (boxes.empty()) throw Error(1, "a remade level needs boxes");
    check_new_boxes(lev, boxes, {});
    relayout(lev, boxes, false);
    did_global_regrid(lev - 1);
}

void Runtime::did_global_regrid(int lev) {
    // LocalRegrid::DidGlobalRegrid (local_regrid.cpp:34-46).
    std::fill(lr_.no_local.begin(), lr_.no_local.end(), 0);
    std::fill(lr_.do_global.begin(), lr_.do_global.end(), 0);
    lr_.nregrids = 0;
    for (int l = lev + 1; l < (int)lr_.last_numPts.size(); ++l)
        lr_.last_numPts[l] = has_level(l) ? levels_[l].L.nvalid() : 0;
}

int Runtime::local_regrid(int lev) {
    sledgehamr::Sledgehamr &S = *sim_;
    if (lev < 0 || !has_level(lev)) throw Error(1, "no such level");
    ++lr_.attempts;
    int finest = 0;
    for (int l = 0; l < (int)levels_.size(); ++l)
        if (has_level(l)) finest = l;
    // Prechecks (local_regrid.cpp:106-145); no restarts here.
    if (lr_.nregrids++ >= max_local_regrids_) return 0;
    if (vt_weak_ <= 1.) {
        lr_.veto_level = lev - 1;
        return 0;
    }
    lr_.veto_level = -1;
    if (lr_.do_global[lev]) return 0;
    if (lev == finest) return 0;
    // InitializeLocalRegrid (:151-160).
    while ((int)lr_.last_numPts.size() <= finest) {
        const int l = (int)lr_.last_numPts.size();
        lr_.last_numPts.push_back(has_level(l) ? levels_[l].L.nvalid() : 0);
    }
    // DetermineAllBoxArrays (:163-172); ErrorEst leaves the tags clear at
    // t_start with a shadow hierarchy and where the project declines the
    // finer level (sledgehamr.cpp:207-230).
    std::vector<BlockSet> add(finest + 1);
    lr_.min_distance.assign(finest + 1, -1.);
    for (int l = lev; l < finest && !lr_.no_local[l]; ++l) {
        const double t = levels_[l].t;
        if ((t == S.t_start && S.shadow_hierarchy) || !S.DoCreateLevelIf(l + 1, t)) continue;
        int d = INT_MAX;
        BlockSet sel;
        regrid_select(l, nullptr, sel, d);
        add[l + 1].insert(sel.begin(), sel.end());
        lr_.min_distance[l + 1] = d == INT_MAX ? -1. : std::sqrt((double)d);
    }
    // FixAllNesting (:176-182), JoinAllBoxArrays (:186-193).
    for (int l = finest; l > 1; --l) fix_nesting(l, add);
    std::vector<std::vector<Box3>> ba(finest + 1);
    for (int l = 1; l <= finest; ++l) ba[l] = merge_boxes(block_boxes(add[l], bf_[l]));
    // CheckForVeto (:304-322): CheckThresholds until the first veto,
    // ComputeLatestPossibleRegridTime for every level above lev.
    bool veto = false;
    lr_.latest.assign(finest + 1, -1.);
    for (int l = 1; l <= finest; ++l) {
        if (!veto) {
            double nb = 0;
            for (const Box3 &b : ba[l]) nb += (double)b.vol();
            const double nc = (double)levels_[l].L.nvalid();
            double nr = (double)lr_.last_numPts[l];
            if (nr == 0) nr = nc;
            const double fv = (nb + nc) / nr;
            veto = fv > vt_weak_;
            if (fv > vt_strong_ && lr_.veto_level == -1) lr_.veto_level = l - 1;
        }
        if (l > lev) {
            double delay = DBL_MAX;
            if (lr_.min_distance[l] >= 0)
                delay = lr_.min_distance[l] / (double)n_error_buf_ * regrid_dt_[l];
            lr_.latest[l] = levels_[l].t + delay;
        }
    }
    // DealWithVeto (:331-400).
    if (veto) {
        const int vl = lr_.veto_level;
        if (vl >= lev || vl <= 0) return 0;
        lr_.do_global[vl] = 1;
        double nsteps = levels_[vl].istep % 2 == 0 ? 1 : 2;
        if (S.shadow_hierarchy) nsteps = 0;
        const double target = levels_[vl].t + nsteps * S.dt[vl];
        bool wait = true;
        for (int l = lev + 1; l <= finest; ++l)
            if (lr_.latest[l] < target) {
                wait = false;
                break;
            }
        if (wait) {
            for (int l = lev; l <= finest; ++l) lr_.no_local[l] = 1;
            return 2;
        }
    }
    // AddAllBoxes (:196-204).
    for (int l = 1; l <= finest; ++l)
        if (!ba[l].empty()) add_boxes(l, ba[l]);
    return 1;
}

}  // namespace shm

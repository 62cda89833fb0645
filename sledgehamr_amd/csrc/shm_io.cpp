// HDF5 files of the reference's I/O paths: the initial-state reader
// (fill_level.cpp:17-251), checkpoints (output_types/checkpoint.cpp:18-308)
// and the spectrum, projection and GW spectrum files (io_module.cpp:280-353,
// output_types/spectrum.cpp:101-112, output_types/projection.cpp:76-86,
// gravitational_waves.cpp:313-321).  Host code only: the data reaches and
// leaves the device through the context's upload / download paths.
#include "shm_hdf5.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <mutex>

#include "../../include/shm.h"
#include "shm_runtime.hpp"

namespace fs = std::filesystem;

namespace shm {
namespace h5 {

namespace {

template <class F>
void bind(void *so, const char *name, F &fn, std::string &missing) {
    fn = reinterpret_cast<F>(dlsym(so, name));
    if (!fn) missing += std::string(missing.empty() ? "" : ", ") + name;
}

bool bind_type(void *so, const char *name, hid_t &out, std::string &missing) {
    const hid_t *g = reinterpret_cast<const hid_t *>(dlsym(so, name));
    if (!g) {
        missing += std::string(missing.empty() ? "" : ", ") + name;
        return false;
    }
    out = *g;
    return true;
}

}  // namespace

const Api &api() {
    static std::once_flag once;
    static Api a{};
    static std::string err;
    std::call_once(once, [] {
        std::vector<std::string> cand;
        if (const char *e = std::getenv("SHM_HDF5_LIB")) cand.push_back(e);
        for (const char *n :
             {"libhdf5.so", "libhdf5.so.310", "libhdf5.so.200", "libhdf5.so.103",
              "libhdf5_serial.so", "libhdf5_serial.so.103", "/opt/conda/lib/libhdf5.so",
              "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
              "/usr/local/lib/libhdf5.so"})
            cand.push_back(n);
        void *so = nullptr;
        std::string tried;
        for (const std::string &c : cand) {
            so = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (so) break;
            tried += (tried.empty() ? "" : ", ") + c;
        }
        if (!so) {
            err = "libhdf5 could not be loaded (tried " + tried +
                  "; set SHM_HDF5_LIB to its path)";
            return;
        }
        std::string missing;
        bind(so, "H5open", a.open, missing);
        bind(so, "H5get_libversion", a.get_libversion, missing);
        bind(so, "H5Eset_auto2", a.eset_auto, missing);
        bind(so, "H5Fcreate", a.fcreate, missing);
        bind(so, "H5Fopen", a.fopen, missing);
        bind(so, "H5Fclose", a.fclose, missing);
        bind(so, "H5Lexists", a.lexists, missing);
        bind(so, "H5Dopen2", a.dopen, missing);
        bind(so, "H5Dcreate2", a.dcreate, missing);
        bind(so, "H5Dread", a.dread, missing);
        bind(so, "H5Dwrite", a.dwrite, missing);
        bind(so, "H5Dclose", a.dclose, missing);
        bind(so, "H5Dget_space", a.dget_space, missing);
        bind(so, "H5Sget_simple_extent_npoints", a.sget_npoints, missing);
        bind(so, "H5Screate_simple", a.screate_simple, missing);
        bind(so, "H5Sclose", a.sclose, missing);
        if (!missing.empty()) {
            err = "libhdf5 lacks " + missing;
            return;
        }
        if (a.get_libversion(&a.version[0], &a.version[1], &a.version[2]) < 0 ||
            a.version[0] < 1 || (a.version[0] == 1 && a.version[1] < 10)) {
            err = "libhdf5 1.10 or newer is needed (64-bit hid_t)";
            return;
        }
        // The predefined type ids are set by the library's initialisation.
        if (a.open() < 0) {
            err = "H5open failed";
            return;
        }
        bind_type(so, "H5T_NATIVE_DOUBLE_g", a.native_double, missing);
        bind_type(so, "H5T_NATIVE_INT_g", a.native_int, missing);
        bind_type(so, "H5T_NATIVE_LLONG_g", a.native_llong, missing);
        bind_type(so, "H5T_IEEE_F64LE_g", a.ieee_f64le, missing);
        bind_type(so, "H5T_STD_I32LE_g", a.std_i32le, missing);
        bind_type(so, "H5T_STD_I64LE_g", a.std_i64le, missing);
        if (!missing.empty()) {
            err = "libhdf5 lacks " + missing;
            return;
        }
        // Absent datasets are an expected outcome of the lookups here: keep
        // the library's error stack off stderr.
        a.eset_auto(kErrDefault, nullptr, nullptr);
    });
    if (!err.empty()) throw Error(SHM_ERR_UNSUPPORTED, err);
    return a;
}

namespace {

hid_t mem_type(const Api &a, Type t) {
    switch (t) {
        case Type::F64: return a.native_double;
        case Type::I32: return a.native_int;
        default: return a.native_llong;
    }
}

}  // namespace

File File::create(const std::string &path) {
    const Api &a = api();
    const hid_t id = a.fcreate(path.c_str(), kAccTrunc, kDefault, kDefault);
    if (id < 0) throw Error(SHM_ERR_IO, "Could not create file: " + path);
    return File(id, path);
}

File File::open(const std::string &path) {
    const Api &a = api();
    if (!fs::is_regular_file(path)) throw Error(SHM_ERR_IO, "Could not open file: " + path);
    const hid_t id = a.fopen(path.c_str(), kAccRdonly, kDefault);
    if (id < 0) throw Error(SHM_ERR_IO, "Could not open file: " + path);
    return File(id, path);
}

File::~File() {
    if (id_ >= 0) api().fclose(id_);
}

bool File::has(const std::string &name) const {
    return api().lexists(id_, name.c_str(), kDefault) > 0;
}

std::string File::find(const std::vector<std::string> &names) const {
    for (const std::string &n : names)
        if (has(n)) return n;
    return "";
}

long long File::size(const std::string &name) const {
    const Api &a = api();
    const hid_t d = a.dopen(id_, name.c_str(), kDefault);
    if (d < 0) throw Error(SHM_ERR_IO, "no dataset " + name + " in " + path_);
    const hid_t s = a.dget_space(d);
    const long long n = s < 0 ? -1 : (long long)a.sget_npoints(s);
    if (s >= 0) a.sclose(s);
    a.dclose(d);
    if (n < 0) throw Error(SHM_ERR_IO, "dataset " + name + " in " + path_ + " has no extent");
    return n;
}

void File::read(const std::string &name, Type t, void *data, long long n) const {
    const long long have = size(name);
    if (have != n)
        throw Error(SHM_ERR_IO, "dataset " + name + " in " + path_ + " holds " +
                                    std::to_string(have) + " values, " + std::to_string(n) +
                                    " are needed");
    const Api &a = api();
    const hid_t d = a.dopen(id_, name.c_str(), kDefault);
    if (d < 0) throw Error(SHM_ERR_IO, "no dataset " + name + " in " + path_);
    const herr_t rc = n > 0 ? a.dread(d, mem_type(a, t), kAll, kAll, kDefault, data) : 0;
    a.dclose(d);
    if (rc < 0) throw Error(SHM_ERR_IO, "reading " + name + " from " + path_ + " failed");
}

void File::write(const std::string &name, Type t, const void *data, long long n,
                 bool as_f64) const {
    const Api &a = api();
    const hsize_t dims[1] = {(hsize_t)n};
    const hid_t s = a.screate_simple(1, dims, nullptr);
    if (s < 0) throw Error(SHM_ERR_IO, "H5Screate_simple failed");
    const hid_t ft = (as_f64 || t == Type::F64) ? a.ieee_f64le
                     : t == Type::I32           ? a.std_i32le
                                                : a.std_i64le;
    const hid_t d = a.dcreate(id_, name.c_str(), ft, s, kDefault, kDefault, kDefault);
    a.sclose(s);
    if (d < 0) throw Error(SHM_ERR_IO, "creating dataset " + name + " in " + path_ + " failed");
    const herr_t rc = n > 0 ? a.dwrite(d, mem_type(a, t), kAll, kAll, kDefault, data) : 0;
    a.dclose(d);
    if (rc < 0) throw Error(SHM_ERR_IO, "writing " + name + " to " + path_ + " failed");
}

}  // namespace h5

using h5::File;
using h5::Type;

namespace {

std::string meta_name(const std::string &folder) { return folder + "/Meta.hdf5"; }
std::string level_dir(const std::string &folder, int lev) {
    return folder + "/Level_" + std::to_string(lev);
}
std::string level_file(const std::string &folder, int lev) {
    return level_dir(folder, lev) + "/Cell.hdf5";
}

int query_upsample() {
    amrex::ParmParse pp("input");
    int up = 1;
    pp.query("upsample", up);
    if (up < 1 || !sledgehamr::utils::IsPowerOfTwo(up))
        throw Error(SHM_ERR_ARG, "Upsample factor input.upsample is not a power of 2!");
    if (up > 1)
        throw Error(SHM_ERR_UNSUPPORTED,
                    "input.upsample > 1 (interpolated initial states) is not supported");
    return up;
}

}  // namespace

// ---------------------------------------------------------------------------
// Initial state (FillLevel, fill_level.cpp).
// ---------------------------------------------------------------------------

void Runtime::read_initial_state(int lev, const std::string &path_arg) {
    std::string path = path_arg;
    if (path.empty()) {
        amrex::ParmParse pp("");
        pp.query("input.initial_state", path);
    }
    // FromInitialStateFile (fill_level.cpp:17-30): a checkpoint restarts
    // the whole hierarchy.
    if (!path.empty() && fs::is_regular_file(meta_name(path))) {
        std::vector<int> nid;
        std::vector<double> lt;
        read_checkpoint(path, -1, -1, nid, lt);
        return;
    }
    read_hdf5_level(lev, path);
}

/** FillLevel::FromHdf5File (fill_level.cpp:57-190) for a level held by one
 *  process: the reference's rank-indexed chunks (one box per rank) become
 *  one chunk per box, indexed by the box's position in the level's box
 *  table. */
void Runtime::read_hdf5_level(int lev, const std::string &path) {
    LevelState &ls = level(lev);
    const auto &boxes = ls.L.boxes;
    const long long nvalid = ls.L.nvalid();
    const long long N = ls.L.N;
    sync();
    if (!path.empty() && fs::is_directory(path)) {
        // Directory of per-component chunk files <dir>/<name>_<box>.hdf5,
        // each with the dataset <name> or "data" (fill_level.cpp:65-100).
        // The reference reads the chunks and leaves the level unchanged
        // (its FromArrayChunks call is commented out, :93-99); here the
        // chunks are applied, in the box-relative order of FromArrayChunks.
        query_upsample();
        for (int f = 0; f < ncomp_; ++f) {
            const std::string &name = names_[f];
            std::vector<double> v((size_t)nvalid);
            long long off = 0;
            for (size_t b = 0; b < boxes.size(); ++b) {
                const std::string file = path + "/" + name + "_" + std::to_string(b) + ".hdf5";
                if (!fs::is_regular_file(file))
                    throw Error(SHM_ERR_IO, "Could not find initial state chunk " + file + "!");
                File h = File::open(file);
                const std::string ds = h.find({name, "data"});
                if (ds.empty())
                    throw Error(SHM_ERR_IO, "Could not find initial state chunk " + file + "!");
                h.read(ds, Type::F64, v.data() + off, boxes[b].vol());
                off += boxes[b].vol();
            }
            upload_array(lev, f, v.data(), 2);
        }
        return;
    }

    amrex::ParmParse pp("input");
    for (int f = 0; f < ncomp_; ++f) {
        const std::string &name = names_[f];
        std::string file;
        pp.query(("initial_state_" + name).c_str(), file);
        if (file.empty()) file = path;
        if (file.empty()) {  // FromConst(f, 0)
            std::vector<double> zero((size_t)nvalid, 0.);
            upload(lev, f, zero.data());
            continue;
        }
        File h = File::open(file);
        // Chunked file: <name>_<box> or data_<box> (fill_level.cpp:134-139).
        const std::string chunk = h.find({name + "_0", "data_0"});
        query_upsample();
        if (chunk.empty()) {
            // Whole level, ind = i N^2 + j N + k (FromArray, :205-223); a
            // missing dataset sets the component to zero (:158-171).
            const std::string ds = h.find({name, "data"});
            if (ds.empty()) {
                std::vector<double> zero((size_t)nvalid, 0.);
                upload(lev, f, zero.data());
                continue;
            }
            std::vector<double> v((size_t)(N * N * N));
            h.read(ds, Type::F64, v.data(), N * N * N);
            upload_array(lev, f, v.data(), 1);
            continue;
        }
        std::vector<double> v((size_t)nvalid);
        long long off = 0;
        for (size_t b = 0; b < boxes.size(); ++b) {
            const std::string ds =
                h.find({name + "_" + std::to_string(b), "data_" + std::to_string(b)});
            if (ds.empty())
                throw Error(SHM_ERR_IO, "Could not find initial state chunk " + name + "_" +
                                            std::to_string(b) + "!");
            h.read(ds, Type::F64, v.data() + off, boxes[b].vol());
            off += boxes[b].vol();
        }
        upload_array(lev, f, v.data(), 2);
    }
}

// ---------------------------------------------------------------------------
// Checkpoints (output_types/checkpoint.cpp).  Meta.hdf5 holds the
// reference's datasets unchanged.  The level data, which the reference
// writes in AMReX's own MultiFab format (VisMF, :91-95), is one HDF5 file
// per level here: Level_<lev>/Cell.hdf5 with "boxes" (lo, hi per box),
// "level" (t, istep, N) and one FP64 dataset per component named after it,
// the valid cells box after box, x fastest within a box.
// ---------------------------------------------------------------------------

void Runtime::write_checkpoint(const std::string &folder, const std::vector<int> &next_id_in,
                               const std::vector<double> &last_time_in,
                               int idx_checkpoints) {
    if (folder.empty()) throw Error(SHM_ERR_ARG, "empty checkpoint folder");
    if (next_id_in.size() != last_time_in.size())
        throw Error(SHM_ERR_ARG, "next_id and last_time_written differ in length");
    int finest = -1;
    for (int l = 0; l < nlevels(); ++l)
        if (has_level(l)) finest = l;
    if (finest < 0) throw Error(SHM_ERR_STATE, "no level to write");
    for (int l = 0; l <= finest; ++l)
        if (!has_level(l)) throw Error(SHM_ERR_STATE, "level " + std::to_string(l) + " undefined");
    sync();
    std::error_code ec;
    for (int lev = 0; lev <= finest; ++lev) {
        fs::create_directories(level_dir(folder, lev), ec);
        if (ec) throw Error(SHM_ERR_IO, "Could not create " + level_dir(folder, lev));
    }
    const int noutput = (int)next_id_in.size();
    {
        File h = File::create(meta_name(folder));
        const double header[8] = {level(0).t,          1.0,
                                  (double)finest,      (double)sim_->dimN[0],
                                  (double)sim_->nghost, (double)ncomp_,
                                  (double)noutput,     (double)idx_checkpoints};
        h.write("Header", Type::F64, header, 8);
        std::vector<int> bf(finest + 1), istep(finest + 1);
        for (int lev = 0; lev <= finest; ++lev) {
            bf[lev] = bf_[lev];
            istep[lev] = level(lev).istep;
        }
        h.write("blocking_factors", Type::I32, bf.data(), finest + 1, true);
        h.write("isteps", Type::I32, istep.data(), finest + 1, true);
        // The checkpoint counts itself as written (checkpoint.cpp:80-82).
        std::vector<int> next_id = next_id_in;
        std::vector<double> last_time = last_time_in;
        if (idx_checkpoints >= 0 && idx_checkpoints < noutput) {
            next_id[idx_checkpoints]++;
            last_time[idx_checkpoints] = level(0).t;
        }
        h.write("next_id", Type::I32, next_id.data(), noutput, true);
        h.write("last_time_written", Type::F64, last_time.data(), noutput);
    }
    for (int lev = 0; lev <= finest; ++lev) {
        LevelState &ls = level(lev);
        File h = File::create(level_file(folder, lev));
        std::vector<int> bx;
        for (const Box3 &b : ls.L.boxes)
            for (int a = 0; a < 6; ++a) bx.push_back(a < 3 ? b.lo[a] : b.hi[a - 3]);
        h.write("boxes", Type::I32, bx.data(), (long long)bx.size());
        const double info[3] = {ls.t, (double)ls.istep, (double)ls.L.N};
        h.write("level", Type::F64, info, 3);
        std::vector<double> v((size_t)ls.L.nvalid());
        for (int c = 0; c < ncomp_; ++c) {
            download(lev, c, v.data());
            h.write(names_[c], Type::F64, v.data(), (long long)v.size());
        }
    }
}

void Runtime::read_checkpoint(const std::string &folder, int noutput, int idx_checkpoints,
                              std::vector<int> &next_id, std::vector<double> &last_time) {
    double header[8];
    {
        if (!fs::is_regular_file(meta_name(folder)))
            throw Error(SHM_ERR_IO, "Could not find checkpoint header!");
        File h = File::open(meta_name(folder));
        if (!h.has("Header")) throw Error(SHM_ERR_IO, "Could not find checkpoint header!");
        h.read("Header", Type::F64, header, 8);
    }
    const double time = header[0];
    const int finest = std::min((int)header[2], sim_->max_level);
    if ((int)header[3] != sim_->dimN[0])
        throw Error(SHM_ERR_ARG, "checkpoint coarse level has N = " +
                                     std::to_string((int)header[3]) + ", the inputs " +
                                     std::to_string(sim_->dimN[0]));
    if ((int)header[5] != ncomp_)
        throw Error(SHM_ERR_ARG, "Number of scalar fields has changed!");
    for (int l = finest + 1; l < nlevels(); ++l)
        if (has_level(l))
            throw Error(SHM_ERR_STATE, "level " + std::to_string(l) +
                                           " exists above the checkpoint's finest level");
    // Output modules (UpdateOutputModules, checkpoint.cpp:195-229).
    File meta = File::open(meta_name(folder));
    if (noutput >= 0) {
        if (noutput != (int)header[6] || idx_checkpoints != (int)header[7])
            throw Error(SHM_ERR_ARG, "Number of output types changed!");
        next_id.assign(noutput, 0);
        last_time.assign(noutput, 0.);
        if (noutput > 0) {
            if (!meta.has("next_id")) throw Error(SHM_ERR_IO, "Could not find next_id!");
            if (!meta.has("last_time_written"))
                throw Error(SHM_ERR_IO, "Could not find last_time_written!");
            meta.read("next_id", Type::I32, next_id.data(), noutput);
            meta.read("last_time_written", Type::F64, last_time.data(), noutput);
        }
    }
    const int nlev_file = (int)header[2] + 1;
    std::vector<int> istep(nlev_file), bf(nlev_file);
    if (!meta.has("isteps")) throw Error(SHM_ERR_IO, "Could not find isteps!");
    if (!meta.has("blocking_factors")) throw Error(SHM_ERR_IO, "Could not find blocking_factors!");
    meta.read("isteps", Type::I32, istep.data(), nlev_file);
    meta.read("blocking_factors", Type::I32, bf.data(), nlev_file);

    sync();
    for (int lev = 0; lev <= finest; ++lev) {
        File h = File::open(level_file(folder, lev));
        const long long nb6 = h.size("boxes");
        if (nb6 % 6) throw Error(SHM_ERR_IO, "malformed box table in " + h.path());
        std::vector<int> bx((size_t)nb6);
        h.read("boxes", Type::I32, bx.data(), nb6);
        std::vector<Box3> boxes(nb6 / 6);
        for (size_t b = 0; b < boxes.size(); ++b)
            for (int a = 0; a < 3; ++a) {
                boxes[b].lo[a] = bx[6 * b + a];
                boxes[b].hi[a] = bx[6 * b + 3 + a];
            }
        set_level(lev, boxes);
        LevelState &ls = level(lev);
        std::vector<double> v((size_t)ls.L.nvalid());
        for (int c = 0; c < ncomp_; ++c) {
            if (!h.has(names_[c]))
                throw Error(SHM_ERR_IO, "no data for " + names_[c] + " in " + h.path());
            h.read(names_[c], Type::F64, v.data(), (long long)v.size());
            upload(lev, c, v.data());
        }
        // Every level restarts at the checkpoint's time (checkpoint.cpp:
        // 158-159) with its own step count (:243-274).
        ls.t = time;
        ls.istep = istep[lev];
        // A coarser blocking factor in the inputs than in the checkpoint
        // needs a global regrid first (:264-269).
        if (bf[lev] < bf_[lev]) {
            if ((int)lr_.do_global.size() < sim_->max_level + 1)
                lr_.do_global.resize(sim_->max_level + 1, 0);
            for (int l = 0; l <= lev; ++l) lr_.do_global[l] = 1;
        }
    }
    sim_->finest_level = finest;
}

void Runtime::delete_checkpoint(const std::string &folder) {
    double header[8];
    {
        if (!fs::is_regular_file(meta_name(folder)))
            throw Error(SHM_ERR_ARG, "Not a valid checkpoint: " + folder);
        File h = File::open(meta_name(folder));
        if (!h.has("Header")) throw Error(SHM_ERR_ARG, "Not a valid checkpoint: " + folder);
        h.read("Header", Type::F64, header, 8);
    }
    // Only the files a checkpoint consists of are removed; a folder is
    // removed once it is empty.
    std::error_code ec;
    fs::remove(meta_name(folder), ec);
    fs::remove(folder + "/BoxArrays", ec);
    for (int lev = 0; lev <= (int)header[2]; ++lev) {
        fs::remove(level_file(folder, lev), ec);
        fs::remove(level_dir(folder, lev), ec);
    }
    fs::remove(folder, ec);
}

// ---------------------------------------------------------------------------
// Spectrum, projection and GW spectrum files.
// ---------------------------------------------------------------------------

void Runtime::write_gw_spectrum_file(const std::string &prefix, int modifier,
                                     const std::vector<double> &mparams) {
    amrex::ParmParse pp("output.gw_spectra");
    int proj = 2, zp = 1, log_bins = 0;
    bool unbinned = true;
    pp.query("projection_type", proj);
    pp.query("zero_padding_factor", zp);
    pp.query("unbinned", unbinned);
    // Not in the reference: nbins logarithmic |k| bins instead.
    pp.query("log_bins", log_bins);
    const int mode = log_bins > 0 ? 2 : unbinned ? 0 : 1;
    std::vector<double> bins, kvals;
    gw_spectrum(mode, log_bins, zp, proj, bins, kvals, modifier, mparams);
    File h = File::create(prefix + "/spectra.hdf5");
    const double header[6] = {level(0).t,
                              (double)(sim_->dimN[0] * zp),
                              (double)bins.size(),
                              sim_->L * zp,
                              (double)zp,
                              mode == 0 ? 1.0 : 0.0};
    h.write("Header", Type::F64, header, 6);
    h.write("k", Type::F64, kvals.data(), (long long)kvals.size());
    h.write("Spectrum", Type::F64, bins.data(), (long long)bins.size());
    if (mode == 2) h.write("log_bins", Type::I32, &log_bins, 1, true);
}

void Runtime::write_spectra_file(const std::string &prefix, const std::vector<int> &fns) {
    if (fns.empty()) return;
    for (int fn : fns) check_output_fn(fn);
    File h = File::create(prefix + "/spectra.hdf5");
    for (size_t i = 0; i < fns.size(); ++i) {
        std::vector<double> bins;
        std::vector<long long> ksq;
        spectrum(fns[i], bins, ksq);
        if (i == 0) {
            const double header[3] = {level(0).t, (double)sim_->dimN[0], (double)bins.size()};
            h.write("Header", Type::F64, header, 3);
            h.write("k_sq", Type::I64, ksq.data(), (long long)ksq.size(), true);
        }
        h.write(output_name(fns[i]), Type::F64, bins.data(), (long long)bins.size());
    }
}

void Runtime::write_projections_file(const std::string &prefix, const std::vector<int> &fns,
                                     const std::vector<int> &modes) {
    if (fns.empty()) return;
    if (!modes.empty() && modes.size() != fns.size())
        throw Error(SHM_ERR_ARG, "one projection mode per output function");
    for (int fn : fns) check_output_fn(fn);
    amrex::ParmParse pp("output.projections");
    int mlevel = -1;
    pp.query("max_level", mlevel);
    File h = File::create(prefix + "/projections.hdf5");
    for (size_t i = 0; i < fns.size(); ++i) {
        std::vector<double> data;
        std::vector<int> count;
        int D = 0;
        projection(fns[i], modes.empty() ? 0 : modes[i], mlevel, data, count, D);
        if (i == 0) {
            const double header[2] = {level(0).t, (double)D};
            h.write("Header", Type::F64, header, 2);
        }
        const std::string id = output_name(fns[i]);
        h.write(id + "_data", Type::F64, data.data(), (long long)data.size());
        h.write(id + "_n", Type::I32, count.data(), (long long)count.size(), true);
    }
}

}  // namespace shm

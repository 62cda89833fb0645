// HDF5 access for the file formats of the reference's I/O paths.
//
// libhdf5 is opened at run time (dlopen), so the product library loads and
// runs without it; the file entry points then fail with a status that names
// the missing library.  The declarations below restate the HDF5 1.10+ C ABI
// that the reference's utils::hdf5 helpers use (utils/hdf5_utils.h:19-136):
// hid_t is 64-bit, H5P_DEFAULT / H5S_ALL / H5E_DEFAULT are 0, the file access
// flags are H5F_ACC_RDONLY = 0 and H5F_ACC_TRUNC = 2, and the predefined
// datatypes are the globals H5T_*_g that H5open() initialises.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace shm {
namespace h5 {

using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
using hssize_t = long long;

constexpr unsigned kAccRdonly = 0x0000u;
constexpr unsigned kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0;  // H5P_DEFAULT
constexpr hid_t kAll = 0;      // H5S_ALL
constexpr hid_t kErrDefault = 0;  // H5E_DEFAULT

/** Element types of the datasets this code reads and writes. */
enum class Type { F64, I32, I64 };

/** The loaded library: function pointers and the predefined type ids. */
struct Api {
    herr_t (*open)();
    herr_t (*get_libversion)(unsigned *, unsigned *, unsigned *);
    herr_t (*eset_auto)(hid_t, void *, void *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    htri_t (*lexists)(hid_t, const char *, hid_t);
    hid_t (*dopen)(hid_t, const char *, hid_t);
    hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dclose)(hid_t);
    hid_t (*dget_space)(hid_t);
    hssize_t (*sget_npoints)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    hid_t native_double, native_int, native_llong, ieee_f64le, std_i32le, std_i64le;
    unsigned version[3];
};

/** The library, loaded on first use: $SHM_HDF5_LIB if set, else
 *  libhdf5.so from the loader's search path, else the versioned names.
 *  Throws shm::Error(SHM_ERR_UNSUPPORTED) when none loads or the version is
 *  below 1.10 (32-bit hid_t). */
const Api &api();

/** One open HDF5 file (closed by the destructor). */
class File {
  public:
    /** H5Fcreate with H5F_ACC_TRUNC. */
    static File create(const std::string &path);
    /** H5Fopen read-only; throws when the file cannot be opened. */
    static File open(const std::string &path);
    File(File &&o) noexcept : id_(o.id_), path_(std::move(o.path_)) { o.id_ = -1; }
    File(const File &) = delete;
    File &operator=(const File &) = delete;
    ~File();

    /** H5Lexists(name) > 0. */
    bool has(const std::string &name) const;
    /** First of names that exists, or "" (utils::hdf5::FindDataset). */
    std::string find(const std::vector<std::string> &names) const;
    /** Number of elements of dataset name (throws if absent). */
    long long size(const std::string &name) const;
    /** Reads the whole dataset into data (n elements, converted to type
     *  t by the library); throws when the dataset holds a different count. */
    void read(const std::string &name, Type t, void *data, long long n) const;
    /** Creates a 1-d dataset of n elements: file type IEEE F64LE for F64,
     *  STD I32LE / I64LE for the integer types unless as_f64 (the
     *  reference's Write<int> stores its int arrays as F64LE,
     *  hdf5_utils.h:86-111). */
    void write(const std::string &name, Type t, const void *data, long long n,
               bool as_f64 = false) const;

    const std::string &path() const { return path_; }

  private:
    File(hid_t id, std::string path) : id_(id), path_(std::move(path)) {}
    hid_t id_;
    std::string path_;
};

}  // namespace h5
}  // namespace shm

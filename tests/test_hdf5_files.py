"""CPU checks of the HDF5 layer the file entry points use.

* The HDF5 C ABI that sledgehamr_amd/csrc/shm_hdf5.hpp restates (hid_t
  width, H5P_DEFAULT / H5S_ALL / H5E_DEFAULT, the H5F_ACC_* flags) equals
  the installed hdf5.h: a program that includes both is compiled and run.
* The product library loads libhdf5 by itself (shm_hdf5_version needs no
  GPU context).
* The tests' own ctypes writer and reader agree (round trip, type
  conversion), so the GPU tests can build files in the reference's layouts
  and read back what the library writes.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import h5util
import sledgehamr_amd as shm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDF5_INC = "/opt/conda/include"
HDF5_LIB = "/opt/conda/lib"

ABI_PROBE = r"""
#include <hdf5.h>
#include <cstdio>
#include <type_traits>
#include "shm_hdf5.hpp"
static_assert(std::is_same<hid_t, shm::h5::hid_t>::value, "hid_t");
static_assert(std::is_same<herr_t, shm::h5::herr_t>::value, "herr_t");
static_assert(std::is_same<htri_t, shm::h5::htri_t>::value, "htri_t");
static_assert(std::is_same<hsize_t, shm::h5::hsize_t>::value, "hsize_t");
static_assert(std::is_same<hssize_t, shm::h5::hssize_t>::value, "hssize_t");
static_assert(H5P_DEFAULT == shm::h5::kDefault, "H5P_DEFAULT");
static_assert(H5S_ALL == shm::h5::kAll, "H5S_ALL");
static_assert(H5E_DEFAULT == shm::h5::kErrDefault, "H5E_DEFAULT");
int main() {
    const bool ok = H5F_ACC_TRUNC == shm::h5::kAccTrunc &&
                    H5F_ACC_RDONLY == shm::h5::kAccRdonly;
    std::printf(ok ? "ABI OK\n" : "ABI MISMATCH\n");
    return ok ? 0 : 1;
}
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(HDF5_INC, "hdf5.h")),
                    reason="hdf5.h not installed")
def test_restated_abi_matches_hdf5_header(tmp_path):
    src = tmp_path / "abi.cpp"
    src.write_text(ABI_PROBE)
    exe = tmp_path / "abi"
    subprocess.run(["g++", "-std=c++17", "-I", HDF5_INC, "-I",
                    os.path.join(ROOT, "sledgehamr_amd", "csrc"), str(src), "-o", str(exe),
                    "-L", HDF5_LIB, "-lhdf5", f"-Wl,-rpath,{HDF5_LIB}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and "ABI OK" in out.stdout, out.stdout + out.stderr


def test_library_loads_libhdf5():
    if not os.path.exists(shm.library_path("RealScalar")):
        pytest.skip("library not built")
    v = shm.hdf5_version("RealScalar")
    assert v[0] == 1 and v[1] >= 10


@pytest.mark.skipif(not h5util.available(), reason="libhdf5 not found")
def test_h5util_round_trip(tmp_path):
    p = str(tmp_path / "x.hdf5")
    a = np.random.default_rng(0).standard_normal(1000)
    b = np.arange(-5, 7, dtype=np.int32)
    c = a.astype(np.float32)
    h5util.write(p, {"a": a, "b": b, "c": c})
    assert h5util.has(p, "a") and not h5util.has(p, "nope")
    assert np.array_equal(h5util.read(p, "a"), a)
    assert np.array_equal(h5util.read(p, "b", np.int32), b)
    assert np.array_equal(h5util.read(p, "b"), b.astype(np.float64))
    assert np.array_equal(h5util.read(p, "c"), c.astype(np.float64))

#!/bin/bash
# Regenerates spectra_ks.json from the reference's data/spectra_ks.hdf5
# (needs the reference checkout and the HDF5 C library; build in /tmp).
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${SLEDGEHAMR_REF:-/root/reference}
HDF5=${HDF5_ROOT:-/opt/conda}
gcc -O2 -std=c99 -I"$HDF5/include" "$HERE/extract_spectra_ks.c" -o /tmp/extract_spectra_ks \
    -L"$HDF5/lib" -lhdf5 -Wl,-rpath,"$HDF5/lib"
/tmp/extract_spectra_ks "$REF/data/spectra_ks.hdf5" "$HERE/spectra_ks.json"

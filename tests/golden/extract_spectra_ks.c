/* Writes tests/golden/spectra_ks.json from the reference's data file
 * data/spectra_ks.hdf5 (the k-bin tables the reference loads for the GW
 * spectrum, sledgehamr.cpp:554-615).
 *
 * Per N: <N>_nks (number of unique |k|^2), <N>_bins (the unique |k|^2,
 * int64), <N>_k1 / <N>_k2 (FP64 per-axis k of the order-2 / order-4
 * gradient symbols).  Full tables are kept for N = 32, 64, 128; for
 * N = 256, 512 only the count, sum and last value of <N>_bins.
 *
 * Build and run in a container holding the reference (run_extract.sh). */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>

static long long read_len(hid_t f, const char *name) {
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (d < 0) { fprintf(stderr, "missing dataset %s\n", name); exit(1); }
    hid_t s = H5Dget_space(d);
    hsize_t n = 0;
    H5Sget_simple_extent_dims(s, &n, NULL);
    H5Sclose(s);
    H5Dclose(d);
    return (long long)n;
}

static void *read_all(hid_t f, const char *name, hid_t type, size_t es, long long *n) {
    *n = read_len(f, name);
    void *buf = malloc((size_t)*n * es);
    hid_t d = H5Dopen2(f, name, H5P_DEFAULT);
    if (H5Dread(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) {
        fprintf(stderr, "read failed: %s\n", name);
        exit(1);
    }
    H5Dclose(d);
    return buf;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s spectra_ks.hdf5 out.json\n", argv[0]);
        return 2;
    }
    hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return 1;
    FILE *o = fopen(argv[2], "w");
    const int Ns[5] = {32, 64, 128, 256, 512};
    fprintf(o, "{\n  \"source\": \"data/spectra_ks.hdf5\",\n");
    for (int t = 0; t < 5; ++t) {
        const int N = Ns[t];
        char nm[64];
        long long n;
        snprintf(nm, sizeof nm, "/%d_nks", N);
        long long *nks = read_all(f, nm, H5T_NATIVE_LLONG, 8, &n);
        snprintf(nm, sizeof nm, "/%d_bins", N);
        long long *bins = read_all(f, nm, H5T_NATIVE_LLONG, 8, &n);
        fprintf(o, "  \"%d\": {\n    \"nks\": %lld,\n    \"nbins\": %lld,\n", N, nks[0], n);
        long long sum = 0;
        for (long long i = 0; i < n; ++i) sum += bins[i];
        fprintf(o, "    \"bins_sum\": %lld,\n    \"bins_last\": %lld", sum, bins[n - 1]);
        if (N <= 128) {
            fprintf(o, ",\n    \"bins\": [");
            for (long long i = 0; i < n; ++i) fprintf(o, "%s%lld", i ? ", " : "", bins[i]);
            fprintf(o, "]");
            for (int kk = 1; kk <= 2; ++kk) {
                long long m;
                snprintf(nm, sizeof nm, "/%d_k%d", N, kk);
                double *k = read_all(f, nm, H5T_NATIVE_DOUBLE, 8, &m);
                fprintf(o, ",\n    \"k%d\": [", kk);
                for (long long i = 0; i < m; ++i) fprintf(o, "%s%.17g", i ? ", " : "", k[i]);
                fprintf(o, "]");
                free(k);
            }
        }
        fprintf(o, "\n  }%s\n", t < 4 ? "," : "");
        free(nks);
        free(bins);
    }
    fprintf(o, "}\n");
    fclose(o);
    H5Fclose(f);
    return 0;
}

// Host harness for the sort-key header mhip_sfc.hpp (compiled with g++ by
// tests/test_hilbert_host.py): keys of n cells, exactly as the device forms them.
#include "../../meshoid_amd/csrc/mhip_sfc.hpp"

extern "C" void keys(const uint32_t *cells, long n, int dim, uint64_t *out) {
    for (long i = 0; i < n; i++) {
        const uint32_t *c = cells + i * 3;
        out[i] = mhip::hilbert_key(c[0], c[1], c[2], dim);
    }
}

extern "C" void cells_of(const float *v, long n, float o, float s, uint32_t *out) {
    for (long i = 0; i < n; i++) out[i] = mhip::key_cell(v[i], o, s);
}

extern "C" int sort_lo_bit(int dim) { return mhip::key_sort_lo_bit(dim); }

// Host harness for mhip_sortnet.hpp (compiled with g++ by tests/test_sortnet_host.py).
#include "../../meshoid_amd/csrc/mhip_sortnet.hpp"

extern "C" void sort32(float *v) {
    float a[32];
    for (int i = 0; i < 32; i++) a[i] = v[i];
    mhip::oddeven_sort<32>(a);
    for (int i = 0; i < 32; i++) v[i] = a[i];
}

extern "C" void merge33(float *A, const float *S) {
    float a[33], s[32];
    for (int i = 0; i < 33; i++) a[i] = A[i];
    for (int i = 0; i < 32; i++) s[i] = S[i];
    mhip::merge_into<32>(a, s);
    for (int i = 0; i < 33; i++) A[i] = a[i];
}

extern "C" int comparators32() {
    int c = 0;
    for (int p = 1; p < 32; p <<= 1)
        for (int k = p; k >= 1; k >>= 1)
            for (int j = k % p; j + k < 32; j += 2 * k)
                for (int i = 0; i < k; i++)
                    if (i + j + k < 32 && (i + j) / (2 * p) == (i + j + k) / (2 * p)) c++;
    return c;
}

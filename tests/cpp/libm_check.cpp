// Host build of the HIP library's transcendentals (xitils_amd/csrc/xt_libm.h,
// compiled here by g++ with -ffp-contract=off as the library is by hipcc):
// reads kind and float arguments from stdin-free argv files and writes the
// results, for tests/test_libm.py to compare with float64 and with the oracle.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../xitils_amd/csrc/xt_libm.h"

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    const int kind = atoi(argv[1]);
    FILE* f = fopen(argv[2], "rb");
    if (!f) return 3;
    std::vector<float> x;
    float v;
    while (fread(&v, 4, 1, f) == 1) x.push_back(v);
    fclose(f);
    const size_t n = kind == 3 ? x.size() / 2 : x.size();
    std::vector<float> out(n);
    for (size_t i = 0; i < n; ++i) {
        float s, c;
        switch (kind) {
            case 0: xt::xt_sincosf(x[i], &s, &c); out[i] = s; break;
            case 1: xt::xt_sincosf(x[i], &s, &c); out[i] = c; break;
            case 2: out[i] = xt::xt_pow5f(x[i]); break;
            case 3: out[i] = xt::xt_atan2f(x[2 * i], x[2 * i + 1]); break;
            default: out[i] = xt::xt_acosf(x[i]); break;
        }
    }
    FILE* g = fopen(argv[3], "wb");
    if (!g) return 4;
    fwrite(out.data(), 4, n, g);
    fclose(g);
    return 0;
}

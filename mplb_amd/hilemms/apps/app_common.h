// app_common.h -- command-line options and result dumps shared by the
// example drivers (Key=Value arguments, like the reference's Config=).
#ifndef APP_COMMON_H
#define APP_COMMON_H
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "mplb.h"

inline std::map<std::string, std::string> AppOptions(int argc, const char** argv) {
    std::map<std::string, std::string> o;
    for (int i = 1; i < argc; i++) {
        const std::string a{argv[i]};
        const size_t eq = a.find('=');
        if (eq != std::string::npos) o[a.substr(0, eq)] = a.substr(eq + 1);
    }
    return o;
}
inline double OptD(const std::map<std::string, std::string>& o, const char* k, double d) {
    auto it = o.find(k);
    return it == o.end() ? d : std::atof(it->second.c_str());
}

// counter-based splitmix64 density noise: rho0 = 1 + amp*(2u-1) with
// u = splitmix64(seed ^ linear index) as a 53-bit fraction
inline double DensityNoise(uint64_t index, uint64_t seed, double amp) {
    uint64_t z = (index ^ seed) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 1.0 + amp * (2.0 * u - 1.0);
}

// raw little-endian doubles: for every block, every macro variable in id
// order, the interior [k][j][i]
inline void DumpMacroVars(const std::string& path) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        ops_printf("Error! cannot write %s\n", path.c_str());
        return;
    }
    for (const auto& idBlock : g_Block()) {
        for (auto& idField : g_MacroVars()) {
            ops_dat d = idField.second.at(idBlock.first);
            const double* p = reinterpret_cast<const double*>(d->data.data());
            for (int k = 0; k < d->size[2]; k++)
                for (int j = 0; j < d->size[1]; j++)
                    for (int i = 0; i < d->size[0]; i++) {
                        const double v = p[d->offset(i, j, k)];
                        std::fwrite(&v, sizeof v, 1, fp);
                    }
        }
    }
    std::fclose(fp);
}
#endif

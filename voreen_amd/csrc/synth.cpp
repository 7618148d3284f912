// synth.cpp -- seeded synthetic volumes (declared in include/hipdvr_synth.h).
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/hipdvr.h"
#include "../../include/hipdvr_synth.h"

namespace {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
inline uint64_t hash64(uint64_t x) {
    uint64_t s = x;
    return splitmix64(s);
}

template <typename F>
void parallel_slices(uint32_t nz, F&& f) {
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (uint32_t z = t; z < nz; z += nt) f(z);
        });
    for (auto& x : th) x.join();
}

// Value noise on an integer lattice: hash -> [-1, 1), smoothstep-interpolated.
inline double lattice(uint64_t seed, int oct, int64_t x, int64_t y, int64_t z) {
    uint64_t h = seed ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(oct + 1));
    h = hash64(h ^ ((uint64_t)x * 0xD6E8FEB86659FD93ULL));
    h = hash64(h ^ ((uint64_t)y * 0xA0761D6478BD642FULL));
    h = hash64(h ^ ((uint64_t)z * 0xE7037ED1A0B428DBULL));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

double value_noise(uint64_t seed, int oct, double x, double y, double z) {
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    const double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    double c[2][2][2];
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) c[k][j][i] = lattice(seed, oct, ix + i, iy + j, iz + k);
    double r = 0.0;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j) {
            const double a = c[k][j][0] + (c[k][j][1] - c[k][j][0]) * tx;
            r += (k ? tz : 1.0 - tz) * (j ? ty : 1.0 - ty) * a;
        }
    return r;
}

}  // namespace

extern "C" int hipdvr_synth_blobs(uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz,
                                  int32_t dtype, void* out) {
    if (!out || nx == 0 || ny == 0 || nz == 0) return HIPDVR_ERR_INVALID;
    if (dtype != HIPDVR_U8 && dtype != HIPDVR_U16) return HIPDVR_ERR_INVALID;
    uint64_t s = seed;
    double c[8][3], sig[8], amp[8];
    for (int k = 0; k < 8; ++k) {
        for (int i = 0; i < 3; ++i) c[k][i] = 0.2 + 0.6 * uniform01(s);
        sig[k] = 0.05 + 0.10 * uniform01(s);
        amp[k] = 0.3 + 0.7 * uniform01(s);
    }
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> f(n);
    std::vector<float> zmax(nz, 0.f);
    parallel_slices(nz, [&](uint32_t z) {
        const double pz = (z + 0.5) / nz;
        float m = 0.f;
        for (uint32_t y = 0; y < ny; ++y) {
            const double py = (y + 0.5) / ny;
            for (uint32_t x = 0; x < nx; ++x) {
                const double px = (x + 0.5) / nx;
                double v = 0.0;
                for (int k = 0; k < 8; ++k) {
                    const double dx = px - c[k][0], dy = py - c[k][1], dz = pz - c[k][2];
                    v += amp[k] * std::exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * sig[k] * sig[k]));
                }
                const float fv = (float)v;
                f[((size_t)z * ny + y) * nx + x] = fv;
                m = std::max(m, fv);
            }
        }
        zmax[z] = m;
    });
    const double fmax = (double)*std::max_element(zmax.begin(), zmax.end());
    const double range = dtype == HIPDVR_U8 ? 255.0 : 65535.0;
    const int nscale = dtype == HIPDVR_U8 ? 1 : 257;
    parallel_slices(nz, [&](uint32_t z) {
        for (size_t i = (size_t)z * ny * nx; i < (size_t)(z + 1) * ny * nx; ++i) {
            const uint64_t h = hash64(seed ^ (0xD1B54A32D192ED03ULL * (i + 1)));
            const int noise = ((int)(h % 5) - 2) * nscale;
            double v = std::floor(range * (double)f[i] / fmax + 0.5) + noise;
            v = std::min(std::max(v, 0.0), range);
            if (dtype == HIPDVR_U8) ((uint8_t*)out)[i] = (uint8_t)v;
            else ((uint16_t*)out)[i] = (uint16_t)v;
        }
    });
    return HIPDVR_OK;
}

extern "C" int hipdvr_synth_shells(uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz,
                                   uint16_t* out) {
    if (!out || nx == 0 || ny == 0 || nz == 0) return HIPDVR_ERR_INVALID;
    const double ax = 0.62, ay = 0.56, az = 0.50;
    const double edges[5] = {1.0, 0.85, 0.70, 0.50, 0.30};
    const double dens[5] = {0.15, 0.45, 0.70, 0.25, 0.90};
    parallel_slices(nz, [&](uint32_t z) {
        const double qz = (z + 0.5) / nz - 0.5;
        for (uint32_t y = 0; y < ny; ++y) {
            const double qy = (y + 0.5) / ny - 0.5;
            for (uint32_t x = 0; x < nx; ++x) {
                const double qx = (x + 0.5) / nx - 0.5;
                const double r = std::sqrt((qx / ax) * (qx / ax) + (qy / ay) * (qy / ay) + (qz / az) * (qz / az));
                uint16_t v = 0;
                if (r < 1.0) {
                    double d = dens[0];
                    for (int k = 0; k < 5; ++k)
                        if (r < edges[k]) d = dens[k];
                    double fbm = 0.0, a = 1.0, fr = 8.0, norm = 0.0;
                    for (int o = 0; o < 4; ++o) {
                        fbm += a * value_noise(seed, o, (qx + 0.5) * fr, (qy + 0.5) * fr, (qz + 0.5) * fr);
                        norm += a;
                        a *= 0.5;
                        fr *= 2.0;
                    }
                    const double val = 65535.0 * (d + 0.05 * fbm / norm);
                    v = (uint16_t)std::min(std::max(std::floor(val + 0.5), 0.0), 65535.0);
                }
                out[((size_t)z * ny + y) * nx + x] = v;
            }
        }
    });
    return HIPDVR_OK;
}

// tf.cpp -- host-side transfer-function tables of libhipdvr.so.
//
// hipdvr_tf_from_keys: the RGBA8 lookup table of a 1D key transfer function,
//   TransFunc1D::createTextureData (src/core/datastructures/transfunc/1d/
//   transfunc1d.cpp:85-116) with TransFunc1DKeys::getMappingForValueUByte
//   (1d/1dkeys/transfunc1dkeys.cpp:119-154) and TransFuncBase::applyAlpha
//   (transfuncbase.cpp:177-210).  Key interpolation keeps the reference's
//   byte truncation (GLubyte += int * float).
// hipdvr_preint_build: PreIntegrationTable::computeTable, exact and integral
//   variants (1d/preintegrationtable.cpp:67-204).
// Compiled without FMA contraction so the float sequences are the reference's.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/hipdvr.h"

namespace {

inline int iround_tgt(float f) { return (int)std::floor(f + 0.5f); }
inline float smaxf(float a, float b) { return (a < b) ? b : a; }
inline float sminf(float a, float b) { return (b < a) ? b : a; }

void map_ubyte(const hipdvr_tf_key* keys, int n, float value, uint8_t out[4]) {
    if (n <= 0) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    value = (value < 0.f) ? 0.f : value;
    value = (value > 1.f) ? 1.f : value;
    int k = 0;
    while (k < n && value > keys[k].intensity) ++k;
    if (k == 0) {
        std::memcpy(out, keys[0].color_l, 4);
    } else if (k == n) {
        std::memcpy(out, keys[n - 1].color_r, 4);
    } else {
        const hipdvr_tf_key& L = keys[k - 1];
        const hipdvr_tf_key& R = keys[k];
        const float fraction = (value - L.intensity) / (R.intensity - L.intensity);
        for (int c = 0; c < 4; ++c) {
            const int diff = (int)R.color_l[c] - (int)L.color_r[c];
            const float s = (float)L.color_r[c] + (float)diff * fraction;
            out[c] = (uint8_t)s;
        }
    }
}

inline float gamma_of(float v, float gamma) { return gamma != 1.f ? std::pow(v, gamma) : v; }

bool keys_valid(const hipdvr_tf_key* keys, int32_t n) {
    if (n < 0 || (n > 0 && !keys)) return false;
    for (int i = 1; i < n; ++i)
        if (!(keys[i].intensity >= keys[i - 1].intensity)) return false;
    return true;
}

template <typename F>
void parallel_rows(int n, F&& f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if ((int)nt > n) nt = (unsigned)(n > 0 ? n : 1);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int i = (int)t; i < n; i += (int)nt) f(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" int hipdvr_tf_from_keys(const hipdvr_tf_key* keys, int32_t nkeys, int32_t width,
                                   float gamma, float thr_lo, float thr_hi, int32_t alpha_mode,
                                   uint8_t* out) {
    if (!out || width < 2 || !keys_valid(keys, nkeys)) return HIPDVR_ERR_INVALID;
    if (alpha_mode < HIPDVR_ALPHA_ZERO || alpha_mode > HIPDVR_ALPHA_ONE) return HIPDVR_ERR_INVALID;
    const int front_end = iround_tgt(thr_lo * (float)width);
    const int back_start = iround_tgt(thr_hi * (float)width);
    for (int x = 0; x < width; ++x) {
        uint8_t* o = out + 4 * (size_t)x;
        if (x < front_end || x > back_start) {
            o[0] = o[1] = o[2] = o[3] = 0;
        } else {
            const float v = (float)x / (float)(width - 1);
            map_ubyte(keys, nkeys, gamma_of(v, gamma), o);
        }
        if (alpha_mode == HIPDVR_ALPHA_ONE) {
            if (o[3]) o[3] = 255;
        } else if (alpha_mode == HIPDVR_ALPHA_ZERO) {
            o[3] = 0;
        }
    }
    return HIPDVR_OK;
}

namespace hipdvr {

// TF colour per pre-integration table index (preintegrationtable.cpp:74-96):
// index i maps intensity i/(n-1) through gamma, the keys and the threshold.
bool preint_index_colors(const hipdvr_tf_key* keys, int32_t nkeys, float gamma, float thr_lo, float thr_hi,
                         int32_t alpha_mode, uint32_t n_, float* tfb) {
    if (n_ < 2 || !keys_valid(keys, nkeys)) return false;
    const int n = (int)n_;
    for (size_t i = 0; i < (size_t)n * 4; ++i) tfb[i] = 0.f;
    const int front_end = iround_tgt(thr_lo * (float)n);
    const int back_start = iround_tgt(thr_hi * (float)n);
    for (int i = front_end < 0 ? 0 : front_end; i < back_start && i < n; ++i) {
        const float intensity = (float)i / (float)(n - 1);
        uint8_t c8[4];
        map_ubyte(keys, nkeys, gamma_of(intensity, gamma), c8);
        float* v = &tfb[4 * (size_t)i];
        for (int c = 0; c < 4; ++c) v[c] = (float)c8[c] / 255.f;
        if (alpha_mode == HIPDVR_ALPHA_ONE && v[3] != 0.f) v[3] = 1.f;
        else if (alpha_mode == HIPDVR_ALPHA_ZERO) v[3] = 0.f;
    }
    return true;
}

}  // namespace hipdvr

extern "C" int hipdvr_preint_build(const hipdvr_tf_key* keys, int32_t nkeys, float gamma,
                                   float thr_lo, float thr_hi, int32_t alpha_mode, uint32_t n_,
                                   float step, int32_t use_integral, float* table) {
    if (!table || n_ < 2 || n_ > 4096 || !keys_valid(keys, nkeys) || !(step > 0.f))
        return HIPDVR_ERR_INVALID;
    const int n = (int)n_;
    std::vector<float> tfb((size_t)n * 4, 0.f);
    hipdvr::preint_index_colors(keys, nkeys, gamma, thr_lo, thr_hi, alpha_mode, n_, tfb.data());
    const float e_full = step * 200.0f;
    if (!use_integral) {
        parallel_rows(n, [&](int sb) {
            for (int sf = 0; sf < n; ++sf) {
                float* dst = table + 4 * ((size_t)sb * n + sf);
                if (sb != sf) {
                    const float scale = 1.0f / (std::fabs((float)(sb - sf)) + 1);
                    const int incr = (sb < sf) ? -1 : 1;
                    float r = 0.f, g = 0.f, b = 0.f, a = 0.f;
                    // the float alpha against the double 0.95 (preintegrationtable.cpp:115):
                    // a composite alpha of exactly 0.95f composites one more sample
                    for (int s = sf; (incr == 1 ? s <= sb : s >= sb) && ((double)a < 0.95); s += incr) {
                        const float* cc = &tfb[4 * (size_t)s];
                        float ca = cc[3];
                        if (ca > 0.0f) {
                            ca = 1.f - std::pow(1.f - ca, step * 200.0f * scale);
                            const float w = (1.0f - a) * ca;
                            const float nr = r + w * cc[0];
                            const float ng = g + w * cc[1];
                            const float nb = b + w * cc[2];
                            a = a + (1.0f - a) * ca;
                            r = nr;
                            g = ng;
                            b = nb;
                        }
                    }
                    const float den = smaxf(a, 0.001f);
                    r /= den;
                    g /= den;
                    b /= den;
                    dst[0] = sminf(smaxf(r, 0.f), 1.f);
                    dst[1] = sminf(smaxf(g, 0.f), 1.f);
                    dst[2] = sminf(smaxf(b, 0.f), 1.f);
                    dst[3] = sminf(smaxf(a, 0.f), 1.f);
                } else {
                    const float* cc = &tfb[4 * (size_t)sf];
                    for (int c = 0; c < 4; ++c) dst[c] = sminf(smaxf(cc[c], 0.f), 1.f);
                    dst[3] = 1.f - std::pow(1.f - dst[3], e_full);
                }
            }
        });
    } else {
        std::vector<float> intf((size_t)n * 4, 0.f);
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        for (int i = 0; i < n; ++i) {
            const float* cc = &tfb[4 * (size_t)i];
            if (cc[3] > 0.0f) {
                acc[0] += cc[0] * cc[3];
                acc[1] += cc[1] * cc[3];
                acc[2] += cc[2] * cc[3];
                acc[3] += cc[3];
            }
            std::memcpy(&intf[4 * (size_t)i], acc, sizeof acc);
        }
        parallel_rows(n, [&](int sb) {
            for (int sf = 0; sf < n; ++sf) {
                const int smin = sb < sf ? sb : sf;
                const int smax = sb < sf ? sf : sb;
                float col[4];
                if (smax != smin) {
                    const float factor = step * 200.f / (float)(smax - smin);
                    for (int c = 0; c < 3; ++c)
                        col[c] = (intf[4 * (size_t)smax + c] - intf[4 * (size_t)smin + c]) * factor;
                    col[3] = 1.f - std::exp(-(intf[4 * (size_t)smax + 3] - intf[4 * (size_t)smin + 3]) * factor);
                    const float den = smaxf(col[3], 0.001f);
                    for (int c = 0; c < 3; ++c) col[c] /= den;
                } else {
                    std::memcpy(col, &tfb[4 * (size_t)smin], sizeof col);
                    col[3] = 1.f - std::pow(1.f - col[3], e_full);
                }
                float* dst = table + 4 * ((size_t)sb * n + sf);
                for (int c = 0; c < 4; ++c) dst[c] = sminf(smaxf(col[c], 0.f), 1.f);
            }
        });
    }
    return HIPDVR_OK;
}

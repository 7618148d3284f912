// The HIPRaycaster processor and the Voreen shim under AddressSanitizer and
// UndefinedBehaviorSanitizer, through the test driver's C entry points: the
// description, every property with valid and invalid values, volumes of both
// voxel types, TFs, entry/exit images, the step rules and process() (which,
// without a GPU, takes its error path).  Built by tests/test_sanitizers.py.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hipdvr.h"
#include "hipdvr_learned.h"

extern "C" {
void* hrc_create();
void hrc_destroy(void* h);
int hrc_describe(void* h, char* buf, int n);
int hrc_set_property(void* h, const char* id, const char* value);
int hrc_set_volume(void* h, const void* data, const uint32_t dims[3], int dtype, float rwm_scale, float rwm_offset,
                   const float spacing[3], const float offset[3]);
int hrc_set_tf(void* h, const hipdvr_tf_key* keys, int n, float dom_lo, float dom_hi);
int hrc_set_images(void* h, const float* entry, const float* exit_, int w, int hh);
int hrc_connect(void* h, int entry, int exit_);
int hrc_is_ready(void* h);
void hrc_set_interaction(int on);
float hrc_render_step(void* h);
float hrc_step(void* h);
int hrc_process(void* h, float* out, int w, int hh);
int hrc_last_error(void* h, char* buf, int n);
int hrc_stats(void* h, hipdvr_stats* st);
int hrc_learned_stats(void* h, hipdvr_learned_stats* st);
int hrc_invalidation(void* h);
}

int main() {
    void* h = hrc_create();
    std::vector<char> buf(1 << 16);
    const int n = hrc_describe(h, buf.data(), (int)buf.size());
    (void)hrc_describe(h, buf.data(), 5);   // truncated copy
    const char* props[][3] = {
        {"samplingRate", "4", "-3"}, {"classification", "pre-integrated-gpu", "nope"},
        {"compositing", "mip", "xx"}, {"shading", "phong", "bad"}, {"gradient", "sobel", "?"},
        {"textureFilterMode_", "nearest", "cubic"}, {"hipKernel", "slab", "fastest"},
        {"hipBrickSize", "32", "7"}, {"hipLearnedStepping", "true", "maybe"}, {"hipPsnrFloor", "50", "-1"},
        {"hipRemarchStride", "4", "0"}, {"hipLearnedTrainingTFs", "16", "999999"},
        {"interactionQuality", "0.5", "2"}, {"preIntegrationTableSize", "512", "3"}, {"noSuchProperty", "1", "2"}};
    for (auto& p : props) {
        (void)hrc_set_property(h, p[0], p[1]);
        (void)hrc_set_property(h, p[0], p[2]);
    }
    (void)hrc_set_property(h, "hipLearnedStepping", "false");
    (void)hrc_set_property(h, "classification", "transfer-function");
    for (int dtype = 0; dtype < 2; ++dtype) {
        const uint32_t dims[3] = {13, 7, 5};
        std::vector<uint16_t> vox(13 * 7 * 5);
        for (size_t i = 0; i < vox.size(); ++i) vox[i] = (uint16_t)(i * 37u);
        const float sp[3] = {1.f, 1.f, 2.f}, off[3] = {0.f, 0.f, 0.f};
        hrc_set_volume(h, vox.data(), dims, dtype, dtype ? 65535.f : 255.f, 0.f, sp, off);
        hipdvr_tf_key keys[2] = {{0.2f, 0, {1, 2, 3, 4}, {1, 2, 3, 4}}, {0.8f, 1, {9, 9, 9, 9}, {200, 100, 50, 25}}};
        hrc_set_tf(h, keys, 2, 0.f, 1.f);
        std::vector<float> en(4 * 9 * 6, 0.25f), ex(4 * 9 * 6, 0.75f);
        hrc_set_images(h, en.data(), ex.data(), 9, 6);
        hrc_connect(h, 1, 1);
        (void)hrc_is_ready(h);
        hrc_set_interaction(1);
        (void)hrc_step(h);
        (void)hrc_render_step(h);
        hrc_set_interaction(0);
        std::vector<float> out(4 * 9 * 6);
        (void)hrc_process(h, out.data(), 9, 6);
        (void)hrc_last_error(h, buf.data(), (int)buf.size());
        hipdvr_stats st;
        hipdvr_learned_stats ls;
        (void)hrc_stats(h, &st);
        (void)hrc_learned_stats(h, &ls);
        (void)hrc_invalidation(h);
    }
    hrc_destroy(h);
    std::printf("ok %d\n", n);
    return 0;
}

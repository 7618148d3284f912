// tgt_probe.cpp -- golden vectors from the reference's own tgt code
// (TEST INFRASTRUCTURE ONLY).
//
// Built by `make -C oracle ref` against the header-only tgt of the reference
// (voreen/ext/tgt, with the boost headers it bundles) into oracle/_ref/; its
// output is the fixture tests/golden/tgt_ray_samples.txt.  For seeded rays
// it evaluates, with tgt's own types and functions, what
// CPURaycaster::directRendering computes before a voxel is read
// (cpuraycaster.cpp:219-247): direction = last - first, tend =
// tgt::length(direction), direction = tgt::normalize(direction), the loop's
// t += step, sample = first + t * direction, the texture-to-voxel matrix of
// VolumeBase::getTextureToVoxelMatrix (volumebase.cpp:221-224: translation by
// -0.5 times scale by the dimensions) applied as mat4 * vec4, tgt::clamp to
// [0, dim - 1] and tgt::iround.  For a set of cameras and window sizes it
// prints the projection * view matrix tgt::Camera gives MeshEntryExitPoints
// (meshentryexitpoints.cpp:210-214: getProjectionMatrix(port size),
// getViewMatrix()).  Floats are printed as C99 hex literals.
#include <cstdint>
#include <cstdio>

#include "tgt/camera.h"
#include "tgt/matrix.h"
#include "tgt/vector.h"

namespace {

uint64_t g_state = 0x7467745F70726F62ULL;

uint64_t next_u64() {
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

float uniform01() { return (float)(next_u64() >> 40) * (1.0f / 16777216.0f); }

}  // namespace

int main() {
    const int rays = 96;
    const float rates[4] = {0.5f, 1.0f, 2.0f, 4.0f};
    for (int r = 0; r < rays; ++r) {
        tgt::svec3 dims(1 + next_u64() % 600, 1 + next_u64() % 600, 1 + next_u64() % 600);
        if (r % 16 == 5) dims = tgt::svec3(1, 1 + next_u64() % 7, 1);
        tgt::vec3 first(uniform01(), uniform01(), uniform01());
        tgt::vec3 last(uniform01(), uniform01(), uniform01());
        if (r % 16 == 3) last = first;   // degenerate ray
        if (r % 16 == 9) {               // along an axis, from a face
            first = tgt::vec3(0.0f, uniform01(), uniform01());
            last = tgt::vec3(1.0f, first.y, first.z);
        }
        const size_t m = dims.x > dims.y ? (dims.x > dims.z ? dims.x : dims.z) : (dims.y > dims.z ? dims.y : dims.z);
        const float step = 1.0f / ((float)m * rates[r % 4]);   // VolumeRaycaster::getSamplingStepSize

        // cpuraycaster.cpp:219-230
        float tend;
        tgt::vec3 direction = last - first;
        if (direction == tgt::vec3(0.0f) && last != tgt::vec3(0.0f) && first != tgt::vec3(0.0f)) {
            tend = step / 2.0f;
        } else {
            tend = tgt::length(direction);
            direction = tgt::normalize(direction);
        }
        const tgt::mat4 tex_to_voxel =
            tgt::mat4::createTranslation(tgt::vec3(-0.5f)) * tgt::mat4::createScale(tgt::vec3(dims));
        std::printf("ray %d dims %zu %zu %zu step %a first %a %a %a last %a %a %a tend %a dir %a %a %a\n", r, dims.x,
                    dims.y, dims.z, step, first.x, first.y, first.z, last.x, last.y, last.z, tend, direction.x,
                    direction.y, direction.z);
        // the loop's samples (no early termination): print a spread of indices
        float t = 0.0f;
        for (int k = 0; k < 255 * 255; ++k) {
            const bool print = k < 4 || k % 97 == 0 || !(t + step <= tend);
            if (print) {
                const tgt::vec3 sample = first + t * direction;
                const tgt::vec3 vp = tgt::clamp((tex_to_voxel * tgt::vec4(sample, 1.f)).xyz(), tgt::vec3::zero,
                                                tgt::vec3(dims) - tgt::vec3::one);
                const tgt::ivec3 iv = tgt::iround(vp);
                std::printf("  k %d t %a vp %a %a %a iv %d %d %d\n", k, t, vp.x, vp.y, vp.z, iv.x, iv.y, iv.z);
            }
            t += step;
            if (t > tend) break;
        }
    }

    struct Cam {
        tgt::vec3 eye, focus, up;
        float fovy;
    };
    const Cam cams[] = {
        // modules/base/test/render/cpuraycaster.vws:96-103 (up not orthogonal to the view)
        {tgt::vec3(1.76227701f, 1.01016998f, 1.77568305f), tgt::vec3(-0.023203f, 0.019314f, 0.012643f),
         tgt::vec3(-0.74486113f, 0.22378904f, 0.62857014f), 45.0f},
        {tgt::vec3(0.9f, 1.4f, -3.1f), tgt::vec3(0.05f, -0.1f, 0.0f), tgt::vec3(0.0f, 1.0f, 0.0f), 30.0f},
        // inside the volume: entry points on the near plane
        {tgt::vec3(0.5f, 0.45f, 0.55f), tgt::vec3(0.5f, 0.5f, 2.0f), tgt::vec3(0.0f, 1.0f, 0.0f), 50.0f},
        // fovy outside [6, 175] degrees: clamped by getProjectionMatrix
        {tgt::vec3(2.0f, 0.3f, 1.5f), tgt::vec3(0.0f), tgt::vec3(0.0f, 0.0f, 1.0f), 3.0f},
        {tgt::vec3(0.5f, 0.45f, 1.15f), tgt::vec3(0.45f, 0.5f, 0.0f), tgt::vec3(0.0f, 1.0f, 0.0f), 179.0f},
    };
    const int sizes[][2] = {{200, 200}, {97, 61}, {61, 97}, {320, 180}};
    int ci = 0;
    for (const Cam& c : cams) {
        for (const auto& wh : sizes) {
            const float pnear = 0.1f, pfar = 50.0f;
            tgt::Camera cam(c.eye, c.focus, c.up, c.fovy, 1.0f, pnear, pfar);
            const tgt::mat4 pv = cam.getProjectionMatrix(tgt::ivec2(wh[0], wh[1])) * cam.getViewMatrix();
            std::printf("cam %d eye %a %a %a focus %a %a %a up %a %a %a fovy %a near %a w %d h %d pv", ci, c.eye.x,
                        c.eye.y, c.eye.z, c.focus.x, c.focus.y, c.focus.z, c.up.x, c.up.y, c.up.z, c.fovy, pnear,
                        wh[0], wh[1]);
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) std::printf(" %a", pv[i][j]);
            std::printf("\n");
        }
        ++ci;
    }
    return 0;
}

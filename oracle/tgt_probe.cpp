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
// [0, dim - 1] and tgt::iround.  Floats are printed as C99 hex literals.
#include <cstdint>
#include <cstdio>

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
    return 0;
}

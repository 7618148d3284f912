// Host-side check of hipdvr::closed_form_steps (voreen_amd/csrc/step_run.h),
// the same source the march kernels compile: for random runs it compares the
// closed form t + k*d with the reference loop's sequential t += step
// (cpuraycaster.cpp:315) for every k <= j (or for sampled k on long runs),
// and checks that every t_k of the jump stays below the bound and that j
// leaves room under the loop cap.  Prints "ok <cases> <jumps>" or the first
// failing case; the exit status is 0 only when every case holds.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>

#include "step_run.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double uniform01() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

static bool check(float t0, float step, float bound, int cap, long& jumps) {
    float d = 0.0f;
    const int j = hipdvr::closed_form_steps(t0, step, bound, cap, d);
    if (j < 0) return false;
    if (j == 0) return true;
    ++jumps;
    if (j > cap - 1) return false;
    const float jumped = t0 + (float)j * d;
    float t = t0;
    const bool every = j <= 4096;
    for (int k = 1; k <= j; ++k) {
        t = t + step;
        if (!(t < bound)) {
            std::printf("t_%d = %a not below bound %a (t0 %a step %a)\n", k, t, bound, t0, step);
            return false;
        }
        if ((every || k == j || (k & 1023) == 0) && t != t0 + (float)k * d) {
            std::printf("t_%d = %a, closed form %a (t0 %a step %a)\n", k, t, t0 + (float)k * d, t0, step);
            return false;
        }
    }
    if (t != jumped) {
        std::printf("jump of %d: %a != %a\n", j, jumped, t);
        return false;
    }
    return true;
}

int main(int argc, char** argv) {
    const long cases = argc > 1 ? std::atol(argv[1]) : 200000;
    long jumps = 0;
    for (long c = 0; c < cases; ++c) {
        // steps log-uniform in [3e-7, 1e-2] (1/(max(dim)*rate) for dims up to
        // 2048 and sampling rates up to 20, and coarse steps); t0 anywhere in
        // (0, 2] with extra weight right below a power of two
        const float step = (float)std::exp(std::log(3e-7) + uniform01() * (std::log(1e-2) - std::log(3e-7)));
        float t0;
        const double u = uniform01();
        if (u < 0.2) {
            const int e = -(int)(uniform01() * 12);
            t0 = (float)(std::ldexp(1.0, e) - uniform01() * 64 * step);
        } else {
            t0 = (float)(uniform01() * 2.0);
        }
        if (!(t0 > 0.0f)) continue;
        const float bound = (float)(t0 + uniform01() * 2.0);
        const int cap = (int)(uniform01() * 65025);
        if (!check(t0, step, bound, cap, jumps)) {
            std::printf("FAIL case %ld\n", c);
            return 1;
        }
    }
    std::printf("ok %ld %ld\n", cases, jumps);
    return 0;
}

// step_run.h -- runs of the ray march's sample recurrence t += step in closed
// form, for the device kernels and for the host-side check of the same code.
#pragma once

#include <hip/hip_runtime.h>

namespace hipdvr {

// The reference loop's float sequence t_{k+1} = t_k + step, k steps at once:
// inside one binade [2^e, 2^(e+1)) every addition of step rounds by the same
// amount (t is on the binade's ulp grid, so (t + step) mod ulp is the same for
// every t), hence t_k = t_0 + k*d with d = (t_0 + step) - t_0, exactly, while
// each sum t_k + step stays below 2^(e+1).  The one exception, a tie resolved
// to even, would alternate the increment; it is excluded by requiring the
// second increment to equal the first.  Returns a number of steps j >= 0 that
// may be taken as t + j*d: every t_k, k <= j, stays below bound and within
// cap_left - 1 steps, with margins of a few steps for the float estimates
// (the two step counts are formed with one reciprocal of d: within 1 ulp of
// the quotients, well inside the margins for counts up to 10^6).
// The product and the sum are exact because t + j*d < 2^(e+1) lies on the
// binade's ulp grid.
__host__ __device__ inline int closed_form_steps(float t, float step, float bound, int cap_left, float& d) {
    if (!(t > 0.0f) || cap_left < 8) return 0;
    const float B = __builtin_bit_cast(float, (__builtin_bit_cast(int, t) & 0x7f800000) + 0x00800000);
    d = (t + step) - t;
    const float t1 = t + d;
    if (((t1 + step) - t1) != d) return 0;
    const float rd = 1.0f / d;
    const float room = (B - step - t) * rd;     // steps k with t_k + step < B: k < room
    const float est = (bound - t) * rd;         // steps k with t_k < bound: k < est
    if (!(room > 6.0f) || !(est > 6.0f)) return 0;
    const int kr = room > 1.0e6f ? 1000000 : (int)room - 3;
    const int ke = est > 1.0e6f ? 1000000 : (int)est - 3;
    int j = kr < ke ? kr : ke;
    j = j < cap_left - 3 ? j : cap_left - 3;
    return j > 0 ? j : 0;
}

}  // namespace hipdvr

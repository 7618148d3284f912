// synth_scratch.h -- set-up of the file-scope scratch of the device
// turbulence generator's no-allocation entry point (internal).
#pragma once

namespace hipdvr {

// Look up the device address of the generator's scratch array on `device`,
// which must be the current device, once per device (hipdvr_create calls it).
void synth_scratch_init(int device);

}  // namespace hipdvr

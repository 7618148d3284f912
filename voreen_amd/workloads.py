"""Seeded benchmark/test workloads of the DVR ray march.

A workload is a volume, a 1D key transfer function, a camera orbit and the
render settings.  Volumes come from libhipdvr's host generators
(include/hipdvr_synth.h); keys and cameras are generated here from SplitMix64
streams so every machine produces the same inputs.

  C0  128^3 uint8 Gaussian blobs, 4-key TF (256 entries), 256^2, step 1/256
  C1  512^3 uint16 nested shells + fBm, 6-key TF (4096 entries, ~70% of the
      range transparent), Phong, 1024^2, step 1/1024, 36 orbit views
"""
import math

import numpy as np

from . import hipdvr as H

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, n):
        return self.next() % n


def random_keys(seed, nkeys):
    """nkeys keys: intensities U[0,1] sorted, RGBA8 U{0..255} (config-0 style)."""
    r = SplitMix64(seed)
    its = sorted(float(np.float32(r.uniform())) for _ in range(nkeys))
    return [(i, tuple(r.randint(256) for _ in range(4))) for i in its]


def shells_keys(seed):
    """6 keys for the shells volume: opaque bumps over [0.42, 0.58] and
    [0.86, 1.0], alpha 0 elsewhere (70% of the range transparent)."""
    r = SplitMix64(seed)

    def col(a):
        return (r.randint(256), r.randint(256), r.randint(256), a)

    a1 = 40 + r.randint(120)
    a2 = 80 + r.randint(150)
    return [(0.0, col(0)), (0.42, col(0)), (0.50, col(a1)), (0.58, col(0)),
            (0.86, col(0)), (1.0, col(a2))]


def turbulence_keys(seed):
    """7 keys for the thresholded turbulence volume: the empty voxels and the
    lowest values transparent, a faint haze over [0.06, 0.22], denser
    structure from 0.3 up."""
    r = SplitMix64(seed)

    def col(a):
        return (r.randint(256), r.randint(256), r.randint(256), a)

    return [(0.0, col(0)), (0.04, col(0)), (0.10, col(12 + r.randint(20))), (0.22, col(4)),
            (0.30, col(60 + r.randint(60))), (0.50, col(140 + r.randint(60))), (1.0, col(255))]


def world_box(dims, spacing=(1.0, 1.0, 1.0)):
    """Proxy box centred at the origin, longest physical extent 1."""
    ext = np.array(dims, np.float64) * np.array(spacing, np.float64)
    ext = ext / ext.max()
    return (-ext / 2).astype(np.float32), (ext / 2).astype(np.float32)


def orbit_camera(view, nviews, radius=3.5, elevation_deg=25.0, fovy=30.0, near=0.1):
    """Camera of view `view` on an orbit around the origin (y up)."""
    th = 2.0 * math.pi * view / nviews
    el = math.radians(elevation_deg)
    eye = (radius * math.sin(th) * math.cos(el), radius * math.sin(el),
           radius * math.cos(th) * math.cos(el))
    cam = H.Camera()
    cam.eye[:] = [float(np.float32(v)) for v in eye]
    cam.focus[:] = [0.0, 0.0, 0.0]
    cam.up[:] = [0.0, 1.0, 0.0]
    cam.fovy_deg = fovy
    cam.near_dist = near
    return cam


def step_for(dims, sampling_rate):
    """cpuraycaster.cpp:104: 1/(max(dim)*samplingRate), in float."""
    return float(np.float32(1.0) / (np.float32(max(dims)) * np.float32(sampling_rate)))


CONFIGS = {
    "c0": dict(dims=(128, 128, 128), dtype=H.U8, image=(256, 256), rate=2.0, views=1,
               shading=H.SHADE_NONE, vol_seed=0xC0, tf_seed=0xC0F, tf_width=256),
    "c1": dict(dims=(512, 512, 512), dtype=H.U16, image=(1024, 1024), rate=2.0, views=36,
               shading=H.SHADE_PHONG, vol_seed=0xC1, tf_seed=0xC1F, tf_width=4096),
    # large-volume stress: 8 GiB of thresholded turbulence made on the device
    "c4": dict(dims=(2048, 2048, 1024), dtype=H.U16, image=(3840, 2160), rate=2.0, views=8,
               shading=H.SHADE_PHONG, vol_seed=0xC4, tf_seed=0xC4F, tf_width=4096, empty=0.4),
}


def make_volume(cfg, dims=None):
    c = CONFIGS[cfg]
    dims = dims or c["dims"]
    if cfg == "c0":
        return H.synth_blobs(c["vol_seed"], dims, H.U8)
    return H.synth_shells(c["vol_seed"], dims)


def make_volume_device(ctx, cfg, dims=None):
    """The turbulence volume of `cfg` made on the device: (device buffer,
    exact empty fraction)."""
    c = CONFIGS[cfg]
    return ctx.synth_turbulence(c["vol_seed"], dims or c["dims"], c["empty"])


def make_tf(cfg):
    c = CONFIGS[cfg]
    keys = (random_keys(c["tf_seed"], 4) if cfg == "c0" else
            turbulence_keys(c["tf_seed"]) if cfg == "c4" else shells_keys(c["tf_seed"]))
    return keys, H.tf_from_keys(keys, c["tf_width"])


def batch_specs(seed=0xC2, n=64, lo=64, hi=384, image=512):
    """Uneven batch: per job dims log-uniform per axis in [lo, hi], uint8 or
    uint16 by coin flip, a blob volume (the single-volume generator), a
    seeded 4-8-key TF and an orbit camera; all seeded from `seed`."""
    r = SplitMix64(seed)
    specs = []
    for j in range(n):
        dims = tuple(int(round(math.exp(math.log(lo) + r.uniform() * (math.log(hi) - math.log(lo)))))
                     for _ in range(3))
        dtype = H.U8 if r.randint(2) == 0 else H.U16
        keys = random_keys(r.next(), 4 + r.randint(5))
        cam = orbit_camera(r.randint(36), 36, elevation_deg=-30.0 + 60.0 * r.uniform())
        specs.append(dict(dims=dims, dtype=dtype, vol_seed=r.next(), keys=keys, camera=cam,
                          image=(image, image)))
    return specs

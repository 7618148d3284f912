"""ctypes wrapper of voreen_amd/hipraycaster/libhipraycaster_shim.so: the
HIPRaycaster processor compiled against the Voreen shim, driven like a network
would drive it (volume inport, entry/exit images, TF keys, isReady/process)."""
import ctypes as C
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "voreen_amd", "hipraycaster", "libhipraycaster_shim.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        from voreen_amd import hipdvr as H
        H.load()
        l = C.CDLL(LIB)
        vp = C.c_void_p
        l.hrc_create.restype = vp
        l.hrc_destroy.argtypes = [vp]
        l.hrc_describe.argtypes = [vp, C.c_char_p, C.c_int]
        l.hrc_set_property.argtypes = [vp, C.c_char_p, C.c_char_p]
        l.hrc_set_volume.argtypes = [vp, vp, vp, C.c_int, C.c_float, C.c_float, vp, vp]
        l.hrc_set_tf.argtypes = [vp, vp, C.c_int, C.c_float, C.c_float]
        l.hrc_set_gradient_volume.argtypes = [vp, vp, vp, C.c_int, C.c_int]
        l.hrc_set_tf2d.argtypes = [vp, vp, C.c_int, C.c_int, vp]
        l.hrc_set_images.argtypes = [vp, vp, vp, C.c_int, C.c_int]
        l.hrc_connect.argtypes = [vp, C.c_int, C.c_int]
        l.hrc_is_ready.argtypes = [vp]
        l.hrc_step.argtypes = [vp]
        l.hrc_step.restype = C.c_float
        l.hrc_render_step.argtypes = [vp]
        l.hrc_render_step.restype = C.c_float
        l.hrc_set_interaction.argtypes = [C.c_int]
        l.hrc_process.argtypes = [vp, vp, C.c_int, C.c_int]
        l.hrc_last_error.argtypes = [vp, C.c_char_p, C.c_int]
        l.hrc_stats.argtypes = [vp, vp]
        l.hrc_learned_stats.argtypes = [vp, vp]
        l.hrc_invalidation.argtypes = [vp]
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Raycaster:
    def __init__(self):
        self.h = C.c_void_p(lib().hrc_create())
        self._keep = []

    def close(self):
        if self.h:
            lib().hrc_destroy(self.h)
            self.h = None

    def describe(self):
        buf = C.create_string_buffer(1 << 16)
        lib().hrc_describe(self.h, buf, len(buf))
        return json.loads(buf.value.decode())

    def set(self, prop, value):
        return lib().hrc_set_property(self.h, prop.encode(), str(value).encode())

    def set_volume(self, vol, rwm=(1.0, 0.0), spacing=(1.0, 1.0, 1.0), offset=(0.0, 0.0, 0.0)):
        vol = np.ascontiguousarray(vol)
        self._keep.append(vol)
        nz, ny, nx = vol.shape
        dims = np.array([nx, ny, nz], np.uint32)
        sp = np.array(spacing, np.float32)
        of = np.array(offset, np.float32)
        lib().hrc_set_volume(self.h, _p(vol), _p(dims), 0 if vol.dtype == np.uint8 else 1,
                             rwm[0], rwm[1], _p(sp), _p(of))

    def set_tf(self, keys, domain=(0.0, 1.0)):
        from voreen_amd import hipdvr as H
        arr, n = H.make_keys(keys)
        self._keep.append(arr)
        lib().hrc_set_tf(self.h, C.cast(arr, C.c_void_p), n, domain[0], domain[1])

    def set_gradient_volume(self, grad):
        """(Z, Y, X, 3 or 4) uint8 / uint16 / float32 on volumehandle.gradientvolumehandle."""
        grad = np.ascontiguousarray(grad)
        self._keep.append(grad)
        nz, ny, nx, ch = grad.shape
        dims = np.array([nx, ny, nz], np.uint32)
        self._keep.append(dims)
        dt = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[grad.dtype]
        return lib().hrc_set_gradient_volume(self.h, _p(grad), _p(dims), dt, ch)

    def set_tf2d(self, rgba, domain=((0.0, 1.0), (0.0, 1.0))):
        rgba = np.ascontiguousarray(rgba, np.float32)
        dom = np.array([domain[0][0], domain[0][1], domain[1][0], domain[1][1]], np.float32)
        return lib().hrc_set_tf2d(self.h, _p(rgba), rgba.shape[1], rgba.shape[0], _p(dom))

    def set_images(self, entry, exit_):
        entry = np.ascontiguousarray(entry, np.float32)
        exit_ = np.ascontiguousarray(exit_, np.float32)
        self.size = (entry.shape[1], entry.shape[0])
        lib().hrc_set_images(self.h, _p(entry), _p(exit_), entry.shape[1], entry.shape[0])

    def connect(self, entry=True, exit_=True):
        lib().hrc_connect(self.h, int(entry), int(exit_))

    def is_ready(self):
        return bool(lib().hrc_is_ready(self.h))

    def step(self):
        return lib().hrc_step(self.h)

    def render_step(self):
        """The step HIPRaycaster marches with."""
        return lib().hrc_render_step(self.h)

    @staticmethod
    def set_interaction(on):
        lib().hrc_set_interaction(int(on))

    def process(self):
        w, h = self.size
        out = np.zeros((h, w, 4), np.float32)
        nerr = lib().hrc_process(self.h, _p(out), w, h)
        return out, nerr

    def last_error(self):
        buf = C.create_string_buffer(4096)
        lib().hrc_last_error(self.h, buf, len(buf))
        return buf.value.decode()

    def stats(self):
        from voreen_amd import hipdvr as H
        st = H.Stats()
        lib().hrc_stats(self.h, C.byref(st))
        return st.as_dict()

    def learned_stats(self):
        """(learned, stats dict) of the last frame."""
        from voreen_amd import hipdvr as H
        st = H.LearnedStats()
        learned = lib().hrc_learned_stats(self.h, C.byref(st))
        return bool(learned), st.as_dict()

    def invalidation(self):
        return lib().hrc_invalidation(self.h)

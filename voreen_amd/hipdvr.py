"""ctypes binding of libhipdvr.so (the C-ABI declared in include/hipdvr.h).

The library is built in-tree by ``python -m voreen_amd.build``.  Loading fails
loudly (``HipDvrError``) when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhipdvr.so")

OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED = 0, 1, 2, 3, 4
U8, U16, F32 = 0, 1, 2
FILTER_NEAREST, FILTER_LINEAR = 0, 1
CLASS_TF, CLASS_PREINT, CLASS_TF2D = 0, 1, 2
GRADIENT_CENTRAL, GRADIENT_FORWARD, GRADIENT_NONE, GRADIENT_FILTERED, GRADIENT_SOBEL = range(5)
# VolumeRaycaster "gradient" option keys (volumeraycaster.cpp:126-131) by GRADIENT_* value
GRADIENT_KEYS = ("central-differences", "forward-differences", "none", "filtered", "sobel")
COMPOSITE_DVR, COMPOSITE_MIP, COMPOSITE_MOP, COMPOSITE_MIDA, COMPOSITE_ISO, COMPOSITE_FHP, COMPOSITE_FHN = range(7)
(SHADE_NONE, SHADE_PHONG, SHADE_PHONG_DIFFUSE, SHADE_PHONG_SPECULAR, SHADE_PHONG_DIFFUSE_AMBIENT,
 SHADE_PHONG_DIFFUSE_SPECULAR, SHADE_TOON, SHADE_COOK_TORRANCE, SHADE_OREN_NAYAR, SHADE_WARD) = range(10)
SHADING_KEYS = ("none", "phong", "phong-diffuse", "phong-specular", "phong-diffuse-ambient",
                "phong-diffuse-specular", "toon", "cook-torrance", "oren-nayar", "ward")
ALPHA_ZERO, ALPHA_USE, ALPHA_ONE = 0, 1, 2
KERNEL_AUTO, KERNEL_REFERENCE, KERNEL_BRICKED, KERNEL_CELL, KERNEL_SLAB = 0, 1, 2, 3, 4

# every symbol include/hipdvr.h and include/hipdvr_synth.h declare
EXPORTS = (
    "hipdvr_create", "hipdvr_destroy", "hipdvr_last_error", "hipdvr_abi_version",
    "hipdvr_set_stream", "hipdvr_synchronize",
    "hipdvr_volume_create", "hipdvr_volume_create_device", "hipdvr_volume_destroy",
    "hipdvr_volume_device_ptr", "hipdvr_volume_set_gradient",
    "hipdvr_tf_from_keys", "hipdvr_tf_create", "hipdvr_tf_destroy", "hipdvr_tf_set_preint",
    "hipdvr_tf2d_create",
    "hipdvr_preint_build",
    "hipdvr_render", "hipdvr_render_device", "hipdvr_entry_exit_device",
    "hipdvr_buffer_alloc", "hipdvr_buffer_free", "hipdvr_memcpy_h2d", "hipdvr_memcpy_d2h",
    "hipdvr_synth_blobs", "hipdvr_synth_shells", "hipdvr_synth_turbulence_device",
    "hipdvr_synth_turbulence_device_noalloc",
    "hipdvr_timing_begin", "hipdvr_timing_read", "hipdvr_copy_device",
    "hipdvr_mlp_create", "hipdvr_mlp_destroy", "hipdvr_mlp_get_params", "hipdvr_mlp_set_params",
    "hipdvr_mlp_forward", "hipdvr_mlp_gradient", "hipdvr_mlp_train_step",
    "hipdvr_learned_create", "hipdvr_learned_destroy", "hipdvr_learned_prepare",
    "hipdvr_learned_trace_reset", "hipdvr_learned_trace", "hipdvr_learned_collect",
    "hipdvr_render_learned_device", "hipdvr_learned_sparse_remarch", "hipdvr_learned_sparse_dims",
    "hipdvr_learned_train",
    "hipdvr_vvd_parse", "hipdvr_vvd_last_error", "hipdvr_volume_load_raw", "hipdvr_volume_load_vvd",
    "hipdvr_volume_minmax", "hipdvr_tf_preint_device",
    "hipdvr_batch_create", "hipdvr_batch_destroy", "hipdvr_batch_set_config", "hipdvr_batch_get_config",
    "hipdvr_batch_autotune", "hipdvr_batch_capture", "hipdvr_batch_launch",
    "hipdvr_batch_set_learned", "hipdvr_batch_launch_learned", "hipdvr_source_digest",
)

MLP_INPUTS, MLP_OUTPUTS, MLP_TARGETS = 16, 2, 3
MLP_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2


class HipDvrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hipdvr error {code}: {msg}")
        self.code = code


class VolumeDesc(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 3), ("dtype", C.c_int32), ("rwm_scale", C.c_float),
                ("rwm_offset", C.c_float), ("spacing", C.c_float * 3), ("offset", C.c_float * 3),
                ("brick_log2", C.c_uint32)]


class Params(C.Structure):
    _fields_ = [("step", C.c_float), ("filter", C.c_int32), ("classification", C.c_int32),
                ("compositing", C.c_int32), ("shading", C.c_int32),
                ("light_pos", C.c_float * 3), ("light_ambient", C.c_float * 3),
                ("light_diffuse", C.c_float * 3), ("light_specular", C.c_float * 3),
                ("shininess", C.c_float), ("camera_pos", C.c_float * 3),
                ("loop_cap", C.c_int32), ("kernel", C.c_int32), ("batch", C.c_int32),
                ("gamma", C.c_float), ("iso_value", C.c_float), ("exact_shading", C.c_int32),
                ("gradient", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("samples", C.c_uint64), ("fetched", C.c_uint64),
                ("shaded", C.c_uint64), ("lane_slots", C.c_uint64), ("shade_slots", C.c_uint64),
                ("kernel_ms", C.c_float)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LearnedOpts(C.Structure):
    _fields_ = [("psnr_floor", C.c_float), ("remarch_stride", C.c_int32), ("skip_logit", C.c_float),
                ("max_log2_step", C.c_int32)]


class TrainOpts(C.Structure):
    _fields_ = [("n_tfs", C.c_uint32), ("n_views", C.c_uint32), ("trace_size", C.c_uint32),
                ("sampling_rate", C.c_float), ("tau_skip", C.c_float), ("tau_step", C.c_float),
                ("min_steps", C.c_uint32), ("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float),
                ("seed", C.c_uint64), ("extra_tfs", C.c_void_p), ("n_extra", C.c_uint32),
                ("extra_width", C.c_uint32), ("max_rows", C.c_uint64)]


class TrainInfo(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("tfs", C.c_uint32), ("views", C.c_uint32), ("trace_size", C.c_uint32),
                ("steps", C.c_uint32), ("loss_first", C.c_float), ("loss_last", C.c_float),
                ("trace_seconds", C.c_float), ("seconds", C.c_float)]


class LearnedStats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("shaded", C.c_uint64), ("sparse_psnr", C.c_float),
                ("fell_back", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchJob(C.Structure):
    _fields_ = [("volume", C.c_void_p), ("tf", C.c_void_p), ("d_entry", C.c_void_p), ("d_exit", C.c_void_p),
                ("d_out", C.c_void_p), ("w", C.c_uint32), ("h", C.c_uint32), ("params", Params)]


class BatchConfig(C.Structure):
    _fields_ = [("tile", C.c_int32), ("brick_log2", C.c_int32), ("batch", C.c_int32), ("resident", C.c_int32)]


TILE_8x8, TILE_16x4, TILE_32x2 = 0, 1, 2


class TFKey(C.Structure):
    _fields_ = [("intensity", C.c_float), ("split", C.c_int32),
                ("color_l", C.c_uint8 * 4), ("color_r", C.c_uint8 * 4)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("focus", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy_deg", C.c_float), ("near_dist", C.c_float)]


_lib = None


def load():
    """Load libhipdvr.so; raises HipDvrError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipDvrError(ERR_HIP, f"{LIB_PATH} is missing: run `python -m voreen_amd.build`")
    lib = C.CDLL(LIB_PATH)
    P, vp, i32, u32 = C.POINTER, C.c_void_p, C.c_int32, C.c_uint32
    sig = {
        "hipdvr_create": ([P(vp), C.c_int], C.c_int),
        "hipdvr_destroy": ([vp], C.c_int),
        "hipdvr_last_error": ([vp], C.c_char_p),
        "hipdvr_abi_version": ([], C.c_int),
        "hipdvr_source_digest": ([], C.c_char_p),
        "hipdvr_set_stream": ([vp, vp], C.c_int),
        "hipdvr_synchronize": ([vp], C.c_int),
        "hipdvr_volume_create": ([vp, P(VolumeDesc), vp, P(vp)], C.c_int),
        "hipdvr_volume_create_device": ([vp, P(VolumeDesc), vp, P(vp)], C.c_int),
        "hipdvr_volume_destroy": ([vp], C.c_int),
        "hipdvr_volume_device_ptr": ([vp, P(vp)], C.c_int),
        "hipdvr_tf_from_keys": ([P(TFKey), i32, i32, C.c_float, C.c_float, C.c_float, i32, vp], C.c_int),
        "hipdvr_tf_create": ([vp, vp, u32, C.c_float, C.c_float, P(vp)], C.c_int),
        "hipdvr_tf_destroy": ([vp], C.c_int),
        "hipdvr_tf2d_create": ([vp, vp, u32, u32, vp, P(vp)], C.c_int),
        "hipdvr_volume_set_gradient": ([vp, vp, i32, u32], C.c_int),
        "hipdvr_tf_set_preint": ([vp, vp, u32], C.c_int),
        "hipdvr_preint_build": ([P(TFKey), i32, C.c_float, C.c_float, C.c_float, i32, u32,
                                 C.c_float, i32, vp], C.c_int),
        "hipdvr_render": ([vp, vp, vp, vp, vp, u32, u32, P(Params), vp, P(Stats)], C.c_int),
        "hipdvr_render_device": ([vp, vp, vp, vp, vp, u32, u32, P(Params), vp, P(Stats)], C.c_int),
        "hipdvr_entry_exit_device": ([vp, P(Camera), vp, vp, u32, u32, i32, vp, vp], C.c_int),
        "hipdvr_buffer_alloc": ([vp, C.c_size_t, P(vp)], C.c_int),
        "hipdvr_buffer_free": ([vp, vp], C.c_int),
        "hipdvr_memcpy_h2d": ([vp, vp, vp, C.c_size_t], C.c_int),
        "hipdvr_memcpy_d2h": ([vp, vp, vp, C.c_size_t], C.c_int),
        "hipdvr_synth_blobs": ([C.c_uint64, u32, u32, u32, i32, vp], C.c_int),
        "hipdvr_synth_shells": ([C.c_uint64, u32, u32, u32, vp], C.c_int),
        "hipdvr_synth_turbulence_device": ([vp, C.c_uint64, u32, u32, u32, C.c_float, vp, P(C.c_double)], C.c_int),
        "hipdvr_synth_turbulence_device_noalloc": ([vp, C.c_uint64, u32, u32, u32, C.c_float, vp, P(C.c_double)],
                                                   C.c_int),
        "hipdvr_timing_begin": ([vp, u32], C.c_int),
        "hipdvr_timing_read": ([vp, C.POINTER(C.c_float), u32, C.POINTER(u32)], C.c_int),
        "hipdvr_copy_device": ([vp, vp, vp, C.c_size_t], C.c_int),
        "hipdvr_mlp_create": ([vp, C.c_uint64, P(vp)], C.c_int),
        "hipdvr_mlp_destroy": ([vp], C.c_int),
        "hipdvr_mlp_get_params": ([vp, vp], C.c_int),
        "hipdvr_mlp_set_params": ([vp, vp], C.c_int),
        "hipdvr_mlp_forward": ([vp, vp, u32, vp], C.c_int),
        "hipdvr_mlp_gradient": ([vp, vp, vp, u32, vp, P(C.c_float)], C.c_int),
        "hipdvr_mlp_train_step": ([vp, vp, vp, u32, C.c_float, P(C.c_float)], C.c_int),
        "hipdvr_learned_create": ([vp, vp, P(vp)], C.c_int),
        "hipdvr_learned_destroy": ([vp], C.c_int),
        "hipdvr_learned_prepare": ([vp, vp, P(Params), vp, P(LearnedOpts)], C.c_int),
        "hipdvr_learned_trace_reset": ([vp], C.c_int),
        "hipdvr_learned_trace": ([vp, vp, P(Params), vp, vp, u32, u32], C.c_int),
        "hipdvr_learned_collect": ([vp, C.c_float, C.c_float, vp, vp, u32, P(u32)], C.c_int),
        "hipdvr_render_learned_device": ([vp, vp, vp, vp, u32, u32, P(Params), P(LearnedOpts), vp,
                                          P(LearnedStats)], C.c_int),
        "hipdvr_learned_sparse_remarch": ([vp, vp, vp, vp, u32, u32, P(Params), P(LearnedOpts), vp,
                                           C.c_uint64, P(u32), P(u32)], C.c_int),
        "hipdvr_learned_sparse_dims": ([u32, u32, P(LearnedOpts), P(u32), P(u32)], C.c_int),
        "hipdvr_learned_train": ([vp, vp, P(TrainOpts), P(TrainInfo)], C.c_int),
        "hipdvr_vvd_parse": ([C.c_char_p, P(VolumeDesc), C.c_char_p, u32], C.c_int),
        "hipdvr_vvd_last_error": ([], C.c_char_p),
        "hipdvr_volume_load_raw": ([vp, C.c_char_p, P(VolumeDesc), C.c_uint64, P(vp)], C.c_int),
        "hipdvr_volume_load_vvd": ([vp, C.c_char_p, u32, P(vp), P(VolumeDesc)], C.c_int),
        "hipdvr_volume_minmax": ([vp, P(C.c_float)], C.c_int),
        "hipdvr_tf_preint_device": ([vp, P(TFKey), i32, C.c_float, C.c_float, C.c_float, i32, u32, C.c_float,
                                     i32], C.c_int),
        "hipdvr_batch_create": ([vp, P(BatchJob), u32, P(vp)], C.c_int),
        "hipdvr_batch_destroy": ([vp], C.c_int),
        "hipdvr_batch_set_config": ([vp, P(BatchConfig)], C.c_int),
        "hipdvr_batch_get_config": ([vp, P(BatchConfig)], C.c_int),
        "hipdvr_batch_autotune": ([vp, u32, P(BatchConfig), P(C.c_float)], C.c_int),
        "hipdvr_batch_capture": ([vp], C.c_int),
        "hipdvr_batch_launch": ([vp, P(Stats)], C.c_int),
        "hipdvr_batch_set_learned": ([vp, P(vp), u32, P(LearnedOpts)], C.c_int),
        "hipdvr_batch_launch_learned": ([vp, P(LearnedStats)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def make_keys(keys):
    """keys: sequence of (intensity, rgba) or (intensity, rgba_left, rgba_right)."""
    arr = (TFKey * max(1, len(keys)))()
    for i, k in enumerate(keys):
        arr[i].intensity = float(k[0])
        cl = k[1]
        cr = k[2] if len(k) > 2 else k[1]
        arr[i].split = int(len(k) > 2)
        for c in range(4):
            arr[i].color_l[c] = int(cl[c])
            arr[i].color_r[c] = int(cr[c])
    return arr, len(keys)


def tf_from_keys(keys, width, gamma=1.0, thr=(0.0, 1.0), alpha_mode=ALPHA_USE):
    """RGBA8 table of a 1D key TF (TransFunc1D::createTextureData semantics)."""
    arr, n = make_keys(keys)
    out = np.zeros((width, 4), np.uint8)
    rc = load().hipdvr_tf_from_keys(arr, n, width, gamma, thr[0], thr[1], alpha_mode, _ptr(out))
    if rc != OK:
        raise HipDvrError(rc, "hipdvr_tf_from_keys: invalid keys or width")
    return out


def preint_build(keys, n, step, use_integral=False, gamma=1.0, thr=(0.0, 1.0),
                 alpha_mode=ALPHA_USE):
    """N x N x RGBA float pre-integration table, element [fe, fs]."""
    arr, nk = make_keys(keys)
    out = np.zeros((n, n, 4), np.float32)
    rc = load().hipdvr_preint_build(arr, nk, gamma, thr[0], thr[1], alpha_mode, n, step,
                                    int(use_integral), _ptr(out))
    if rc != OK:
        raise HipDvrError(rc, "hipdvr_preint_build: invalid arguments")
    return out


def synth_blobs(seed, dims, dtype=U8):
    nx, ny, nz = dims
    out = np.empty((nz, ny, nx), np.uint8 if dtype == U8 else np.uint16)
    rc = load().hipdvr_synth_blobs(seed, nx, ny, nz, dtype, _ptr(out))
    if rc != OK:
        raise HipDvrError(rc, "hipdvr_synth_blobs")
    return out


def synth_shells(seed, dims):
    nx, ny, nz = dims
    out = np.empty((nz, ny, nx), np.uint16)
    rc = load().hipdvr_synth_shells(seed, nx, ny, nz, _ptr(out))
    if rc != OK:
        raise HipDvrError(rc, "hipdvr_synth_shells")
    return out


def default_params(step, filter_=FILTER_LINEAR, classification=CLASS_TF, shading=SHADE_NONE,
                   kernel=KERNEL_AUTO, light=None, camera_pos=(0.0, 0.0, 3.5), batch=0,
                   compositing=COMPOSITE_DVR, gamma=0.0, iso_value=0.5, exact_shading=0,
                   gradient=GRADIENT_CENTRAL):
    """Render parameters with Voreen's VolumeRenderer light defaults
    (src/core/processors/volumerenderer.cpp:41-54)."""
    p = Params()
    p.step = step
    p.filter = filter_
    p.classification = classification
    p.compositing = compositing
    p.gamma = gamma
    p.iso_value = iso_value
    p.exact_shading = exact_shading
    p.gradient = gradient
    p.shading = shading
    lt = dict(pos=(2.3, 1.5, 1.5), ambient=(0.4,) * 3, diffuse=(0.8,) * 3, specular=(0.6,) * 3,
              shininess=60.0)
    if light:
        lt.update(light)
    p.light_pos[:] = list(lt["pos"])
    p.light_ambient[:] = list(lt["ambient"])
    p.light_diffuse[:] = list(lt["diffuse"])
    p.light_specular[:] = list(lt["specular"])
    p.shininess = lt["shininess"]
    p.camera_pos[:] = list(camera_pos)
    p.loop_cap = 0
    p.kernel = kernel
    p.batch = batch
    return p


def vvd_parse(path):
    """Parse a .vvd file (host only): (VolumeDesc, raw file path)."""
    lib = load()
    d = VolumeDesc()
    buf = C.create_string_buffer(4096)
    rc = lib.hipdvr_vvd_parse(os.fsencode(path), C.byref(d), buf, len(buf))
    if rc != OK:
        raise HipDvrError(rc, f"hipdvr_vvd_parse({path}): {lib.hipdvr_vvd_last_error().decode()}")
    return d, buf.value.decode()


class Context:
    """One device context: stream, events, counters (hipdvr_create)."""

    def __init__(self, device=0):
        self.lib = load()
        self.h = C.c_void_p()
        rc = self.lib.hipdvr_create(C.byref(self.h), device)
        if rc != OK:
            raise HipDvrError(rc, f"hipdvr_create(device={device}) failed")

    def check(self, rc, what):
        if rc != OK:
            msg = self.lib.hipdvr_last_error(self.h)
            raise HipDvrError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.hipdvr_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        self.check(self.lib.hipdvr_synchronize(self.h), "hipdvr_synchronize")

    # ---- objects ----
    def volume(self, voxels, rwm=(1.0, 0.0), spacing=(1.0, 1.0, 1.0), offset=(0.0, 0.0, 0.0),
               brick_log2=4):
        return Volume(self, voxels, rwm, spacing, offset, brick_log2)

    def transfunc(self, rgba8, domain=(0.0, 1.0)):
        return TransFunc(self, rgba8, domain)

    def transfunc2d(self, rgba, domain=((0.0, 1.0), (0.0, 1.0))):
        """2D TF (hipdvr_tf2d_create) from a (H, W, 4) float32 texture and the
        (intensity, gradient-magnitude) domains."""
        rgba = np.ascontiguousarray(rgba, np.float32)
        if rgba.ndim != 3 or rgba.shape[2] != 4:
            raise ValueError("2D TF texture must be (H, W, 4)")
        t = TransFunc.__new__(TransFunc)
        t.ctx, t.width, t.h = self, rgba.shape[1], C.c_void_p()
        dom = (C.c_float * 4)(domain[0][0], domain[0][1], domain[1][0], domain[1][1])
        self.check(self.lib.hipdvr_tf2d_create(self.h, _ptr(rgba), rgba.shape[1], rgba.shape[0], dom,
                                               C.byref(t.h)), "hipdvr_tf2d_create")
        return t

    def synth_turbulence(self, seed, dims, empty_fraction=0.4, out=None, noalloc=False):
        """uint16 turbulence volume made on the device (hipdvr_synth_turbulence_device,
        or its no-allocation variant): (device buffer, exact empty fraction);
        dims = (x, y, z); into `out` when given, else a new buffer."""
        nx, ny, nz = dims
        buf = out if out is not None else self.alloc(nx * ny * nz * 2)
        if buf.nbytes < nx * ny * nz * 2:
            raise ValueError("output buffer too small")
        frac = C.c_double()
        fn = "hipdvr_synth_turbulence_device_noalloc" if noalloc else "hipdvr_synth_turbulence_device"
        try:
            self.check(getattr(self.lib, fn)(self.h, seed, nx, ny, nz, empty_fraction, buf.ptr, C.byref(frac)), fn)
        except Exception:
            if out is None:
                buf.free()
            raise
        return buf, frac.value

    def volume_device(self, buf, dims, dtype=U16, rwm=(1.0, 0.0), spacing=(1.0, 1.0, 1.0),
                      offset=(0.0, 0.0, 0.0), brick_log2=4):
        """Volume from device-resident voxels (hipdvr_volume_create_device); dims = (x, y, z)."""
        d = VolumeDesc()
        d.dims[:] = list(dims)
        d.dtype = dtype
        d.rwm_scale, d.rwm_offset = rwm
        d.spacing[:] = list(spacing)
        d.offset[:] = list(offset)
        d.brick_log2 = brick_log2
        h = C.c_void_p()
        self.check(self.lib.hipdvr_volume_create_device(self.h, C.byref(d), buf.ptr, C.byref(h)),
                   "hipdvr_volume_create_device")
        return Volume.from_handle(self, h, d)

    def load_vvd(self, path, brick_log2=4):
        h, d = C.c_void_p(), VolumeDesc()
        self.check(self.lib.hipdvr_volume_load_vvd(self.h, os.fsencode(path), brick_log2, C.byref(h),
                                                   C.byref(d)), "hipdvr_volume_load_vvd")
        d.brick_log2 = brick_log2
        return Volume.from_handle(self, h, d)

    def load_raw(self, path, dims, dtype, rwm=(1.0, 0.0), spacing=(1.0, 1.0, 1.0), offset=(0.0, 0.0, 0.0),
                 brick_log2=4, byte_offset=0):
        d = VolumeDesc()
        d.dims[:] = list(dims)
        d.dtype = dtype
        d.rwm_scale, d.rwm_offset = rwm
        d.spacing[:] = list(spacing)
        d.offset[:] = list(offset)
        d.brick_log2 = brick_log2
        h = C.c_void_p()
        self.check(self.lib.hipdvr_volume_load_raw(self.h, os.fsencode(path), C.byref(d), byte_offset,
                                                   C.byref(h)), "hipdvr_volume_load_raw")
        return Volume.from_handle(self, h, d)

    # ---- buffers ----
    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.lib.hipdvr_buffer_alloc(self.h, nbytes, C.byref(p)), "hipdvr_buffer_alloc")
        return DeviceBuffer(self, p, nbytes)

    def timing_begin(self, capacity):
        self.check(self.lib.hipdvr_timing_begin(self.h, capacity), "hipdvr_timing_begin")

    def timing_read(self, max_n=65536):
        buf = (C.c_float * max_n)()
        n = C.c_uint32()
        self.check(self.lib.hipdvr_timing_read(self.h, buf, max_n, C.byref(n)), "hipdvr_timing_read")
        return list(buf[:min(n.value, max_n)])

    def copy_device(self, dst, src, nbytes):
        self.check(self.lib.hipdvr_copy_device(self.h, dst.ptr, src.ptr, nbytes), "hipdvr_copy_device")

    # ---- rendering ----
    def render(self, vol, tf, entry, exit_, params, stats=True, out=None):
        """hipdvr_render: host entry/exit images in, host image out (into
        `out` when given, a C-contiguous (H, W, 4) float32 array)."""
        entry = np.ascontiguousarray(entry, np.float32)
        exit_ = np.ascontiguousarray(exit_, np.float32)
        if entry.shape != exit_.shape or entry.ndim != 3 or entry.shape[2] != 4:
            raise ValueError("entry/exit must be (H, W, 4) float32 arrays of equal shape")
        h, w = entry.shape[:2]
        if out is None:
            out = np.zeros((h, w, 4), np.float32)
        elif out.shape != entry.shape or out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float32 array of the entry image's shape")
        st = Stats()
        rc = self.lib.hipdvr_render(self.h, vol.h, tf.h, _ptr(entry), _ptr(exit_), w, h,
                                    C.byref(params), _ptr(out), C.byref(st) if stats else None)
        self.check(rc, "hipdvr_render")
        return out, (st.as_dict() if stats else None)

    def render_device(self, vol, tf, d_entry, d_exit, w, h, params, d_out, stats=False):
        st = Stats()
        rc = self.lib.hipdvr_render_device(self.h, vol.h, tf.h, d_entry.ptr, d_exit.ptr, w, h,
                                           C.byref(params), d_out.ptr,
                                           C.byref(st) if stats else None)
        self.check(rc, "hipdvr_render_device")
        return st.as_dict() if stats else None

    def entry_exit_device(self, cam, llf, urb, w, h, d_entry, d_exit, quantize=False):
        llf = np.asarray(llf, np.float32)
        urb = np.asarray(urb, np.float32)
        rc = self.lib.hipdvr_entry_exit_device(self.h, C.byref(cam), _ptr(llf), _ptr(urb), w, h,
                                               int(quantize), d_entry.ptr, d_exit.ptr)
        self.check(rc, "hipdvr_entry_exit_device")


class DeviceBuffer:
    def __init__(self, ctx, ptr, nbytes):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.hipdvr_memcpy_h2d(self.ctx.h, self.ptr, _ptr(arr), arr.nbytes),
                       "hipdvr_memcpy_h2d")

    def download(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.hipdvr_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, out.nbytes),
                       "hipdvr_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            self.ctx.lib.hipdvr_buffer_free(self.ctx.h, self.ptr)
            self.ptr = None


class Volume:
    """Device volume (hipdvr_volume_create); voxels is a (Z, Y, X) array."""

    def __init__(self, ctx, voxels, rwm, spacing, offset, brick_log2):
        voxels = np.ascontiguousarray(voxels)
        if voxels.dtype not in (np.uint8, np.uint16) or voxels.ndim != 3:
            raise ValueError("voxels must be a 3D uint8 or uint16 array")
        self.ctx = ctx
        d = VolumeDesc()
        nz, ny, nx = voxels.shape
        d.dims[:] = [nx, ny, nz]
        d.dtype = U8 if voxels.dtype == np.uint8 else U16
        d.rwm_scale, d.rwm_offset = rwm
        d.spacing[:] = list(spacing)
        d.offset[:] = list(offset)
        d.brick_log2 = brick_log2
        self.desc = d
        self.h = C.c_void_p()
        ctx.check(ctx.lib.hipdvr_volume_create(ctx.h, C.byref(d), _ptr(voxels), C.byref(self.h)),
                  "hipdvr_volume_create")

    @classmethod
    def from_handle(cls, ctx, h, desc):
        v = cls.__new__(cls)
        v.ctx, v.h, v.desc = ctx, h, desc
        return v

    def set_gradient(self, grad):
        """Attach the gradient volume of the 2D TF path: (Z, Y, X, 3 or 4)
        uint8 / uint16 / float32, the volume's dims."""
        grad = np.ascontiguousarray(grad)
        dt = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16, np.dtype(np.float32): F32}.get(grad.dtype)
        if dt is None or grad.ndim != 4:
            raise ValueError("gradient volume must be a (Z, Y, X, C) uint8, uint16 or float32 array")
        if tuple(grad.shape[:3]) != tuple(self.desc.dims[::-1]):
            raise ValueError("gradient volume dimensions differ from intensity volume dimensions")
        self.ctx.check(self.ctx.lib.hipdvr_volume_set_gradient(self.h, _ptr(grad), dt, grad.shape[3]),
                       "hipdvr_volume_set_gradient")

    def minmax(self):
        """(minNorm, maxNorm, min, max) of the whole volume (VolumeMinMax)."""
        out = (C.c_float * 4)()
        self.ctx.check(self.ctx.lib.hipdvr_volume_minmax(self.h, out), "hipdvr_volume_minmax")
        return tuple(out)

    def close(self):
        if self.h:
            self.ctx.lib.hipdvr_volume_destroy(self.h)
            self.h = C.c_void_p()


class TransFunc:
    def __init__(self, ctx, rgba8, domain):
        rgba8 = np.ascontiguousarray(rgba8, np.uint8)
        self.ctx = ctx
        self.width = rgba8.shape[0]
        self.h = C.c_void_p()
        ctx.check(ctx.lib.hipdvr_tf_create(ctx.h, _ptr(rgba8), self.width, domain[0], domain[1],
                                           C.byref(self.h)), "hipdvr_tf_create")

    def preint_device(self, keys, n, step, use_integral=False, gamma=1.0, thr=(0.0, 1.0),
                      alpha_mode=ALPHA_USE):
        """Build the pre-integration table on the device and attach it."""
        arr, nk = make_keys(keys)
        self.ctx.check(self.ctx.lib.hipdvr_tf_preint_device(self.h, arr, nk, gamma, thr[0], thr[1], alpha_mode, n,
                                                            step, int(use_integral)), "hipdvr_tf_preint_device")

    def set_preint(self, table):
        table = np.ascontiguousarray(table, np.float32)
        self.ctx.check(self.ctx.lib.hipdvr_tf_set_preint(self.h, _ptr(table), table.shape[0]),
                       "hipdvr_tf_set_preint")

    def close(self):
        if self.h:
            self.ctx.lib.hipdvr_tf_destroy(self.h)
            self.h = C.c_void_p()


class Mlp:
    """The stepping network (include/hipdvr_learned.h)."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(ctx.lib.hipdvr_mlp_create(ctx.h, seed, C.byref(self.h)), "hipdvr_mlp_create")

    def close(self):
        if self.h:
            self.ctx.lib.hipdvr_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def params(self):
        out = np.zeros(MLP_PARAMS, np.float32)
        self.ctx.check(self.ctx.lib.hipdvr_mlp_get_params(self.h, _ptr(out)), "hipdvr_mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_PARAMS
        self.ctx.check(self.ctx.lib.hipdvr_mlp_set_params(self.h, _ptr(p)), "hipdvr_mlp_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        rows = x.shape[0]
        dx, dy = self.ctx.alloc(max(x.nbytes, 4)), self.ctx.alloc(max(rows * 8, 4))
        try:
            dx.upload(x)
            self.ctx.check(self.ctx.lib.hipdvr_mlp_forward(self.h, dx.ptr, rows, dy.ptr), "hipdvr_mlp_forward")
            return dy.download((rows, MLP_OUTPUTS), np.float32)
        finally:
            dx.free()
            dy.free()

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dx, dt = self.ctx.alloc(x.nbytes), self.ctx.alloc(t.nbytes)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        try:
            dx.upload(x)
            dt.upload(t)
            self.ctx.check(self.ctx.lib.hipdvr_mlp_gradient(self.h, dx.ptr, dt.ptr, x.shape[0], _ptr(g),
                                                            C.byref(loss)), "hipdvr_mlp_gradient")
            return g, loss.value
        finally:
            dx.free()
            dt.free()

    def train_step_device(self, dx, dt, rows, lr, want_loss=False):
        loss = C.c_float()
        self.ctx.check(self.ctx.lib.hipdvr_mlp_train_step(self.h, dx.ptr, dt.ptr, rows, lr,
                                                          C.byref(loss) if want_loss else None),
                       "hipdvr_mlp_train_step")
        return loss.value if want_loss else None


def learned_opts(psnr_floor=45.0, remarch_stride=8, skip_logit=0.0, max_log2_step=3):
    o = LearnedOpts()
    o.psnr_floor, o.remarch_stride, o.skip_logit, o.max_log2_step = (psnr_floor, remarch_stride,
                                                                      skip_logit, max_log2_step)
    return o


class Learned:
    """Learned-stepping state of one volume (include/hipdvr_learned.h)."""

    def __init__(self, ctx, vol):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(ctx.lib.hipdvr_learned_create(ctx.h, vol.h, C.byref(self.h)), "hipdvr_learned_create")

    def close(self):
        if self.h:
            self.ctx.lib.hipdvr_learned_destroy(self.h)
            self.h = C.c_void_p()

    def prepare(self, tf, params, mlp=None, opts=None):
        self.ctx.check(self.ctx.lib.hipdvr_learned_prepare(self.h, tf.h, C.byref(params),
                                                           mlp.h if mlp else None,
                                                           C.byref(opts) if opts else None),
                       "hipdvr_learned_prepare")

    def trace_reset(self):
        self.ctx.check(self.ctx.lib.hipdvr_learned_trace_reset(self.h), "hipdvr_learned_trace_reset")

    def trace(self, tf, params, d_entry, d_exit, w, h):
        self.ctx.check(self.ctx.lib.hipdvr_learned_trace(self.h, tf.h, C.byref(params), d_entry.ptr,
                                                         d_exit.ptr, w, h), "hipdvr_learned_trace")

    def collect(self, tau_skip, tau_step, dx, dt, capacity, count):
        n = C.c_uint32(count)
        self.ctx.check(self.ctx.lib.hipdvr_learned_collect(self.h, tau_skip, tau_step, dx.ptr, dt.ptr,
                                                           capacity, C.byref(n)), "hipdvr_learned_collect")
        return n.value

    def train(self, mlp, extra_tfs=(), **kw):
        """hipdvr_learned_train: trace, collect and Adam steps for this
        volume; keyword arguments are the fields of hipdvr_learned_train_opts
        (n_tfs, n_views, trace_size, sampling_rate, tau_skip, tau_step,
        min_steps, epochs, batch, lr, seed, max_rows).  Returns the info dict."""
        o = TrainOpts()
        for k, v in kw.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown training option {k}")
            setattr(o, k, v)
        keep = None
        if len(extra_tfs):
            keep = np.ascontiguousarray(np.stack([np.asarray(t, np.uint8) for t in extra_tfs]), np.uint8)
            o.extra_tfs = keep.ctypes.data
            o.n_extra = keep.shape[0]
            o.extra_width = keep.shape[1]
        info = TrainInfo()
        self.ctx.check(self.ctx.lib.hipdvr_learned_train(self.h, mlp.h, C.byref(o), C.byref(info)),
                       "hipdvr_learned_train")
        del keep
        return dict(rows=int(info.rows), tfs=int(info.tfs), views=int(info.views), trace_size=int(info.trace_size),
                    steps=int(info.steps), loss_first=float(info.loss_first), loss_last=float(info.loss_last),
                    data_s=round(float(info.trace_seconds), 2), total_s=round(float(info.seconds), 2))

    def render_device(self, tf, d_entry, d_exit, w, h, params, d_out, opts=None, stats=False):
        st = LearnedStats()
        self.ctx.check(self.ctx.lib.hipdvr_render_learned_device(
            self.h, tf.h, d_entry.ptr, d_exit.ptr, w, h, C.byref(params),
            C.byref(opts) if opts else None, d_out.ptr, C.byref(st) if stats else None),
            "hipdvr_render_learned_device")
        return st.as_dict() if stats else None

    def sparse_remarch(self, tf, d_entry, d_exit, w, h, params, d_sparse, opts=None):
        """The judge's plain re-march alone (hipdvr_learned_sparse_remarch) into
        the device buffer d_sparse (its size bounds the grid): (sw, sh)."""
        sw, sh = C.c_uint32(), C.c_uint32()
        self.ctx.check(self.ctx.lib.hipdvr_learned_sparse_remarch(
            self.h, tf.h, d_entry.ptr, d_exit.ptr, w, h, C.byref(params), C.byref(opts) if opts else None,
            d_sparse.ptr, d_sparse.nbytes // 16, C.byref(sw), C.byref(sh)), "hipdvr_learned_sparse_remarch")
        return sw.value, sh.value


def source_digest():
    """Digest of the sources libhipdvr.so was built from
    (hipdvr_source_digest); voreen_amd.build.source_digest() gives the tree's."""
    return load().hipdvr_source_digest().decode()


def learned_sparse_dims(w, h, remarch_stride=0):
    """The judge's sparse grid size for a w x h image (host arithmetic)."""
    o = LearnedOpts()
    o.remarch_stride = remarch_stride
    sw, sh = C.c_uint32(), C.c_uint32()
    rc = load().hipdvr_learned_sparse_dims(w, h, C.byref(o), C.byref(sw), C.byref(sh))
    if rc != OK:
        raise HipDvrError(rc, "hipdvr_learned_sparse_dims")
    return sw.value, sh.value


class Batch:
    """Many volumes in one persistent launch (include/hipdvr_batch.h).
    jobs: sequence of (volume, tf, d_entry, d_exit, d_out, w, h, params)."""

    def __init__(self, ctx, jobs):
        self.ctx = ctx
        arr = (BatchJob * len(jobs))()
        for i, (vol, tf, de, dx, do, w, h, prm) in enumerate(jobs):
            arr[i].volume, arr[i].tf = vol.h, tf.h
            arr[i].d_entry, arr[i].d_exit, arr[i].d_out = de.ptr, dx.ptr, do.ptr
            arr[i].w, arr[i].h, arr[i].params = w, h, prm
        self.n = len(jobs)
        self.h = C.c_void_p()
        ctx.check(ctx.lib.hipdvr_batch_create(ctx.h, arr, len(jobs), C.byref(self.h)), "hipdvr_batch_create")

    def close(self):
        if self.h:
            self.ctx.lib.hipdvr_batch_destroy(self.h)
            self.h = C.c_void_p()

    def set_config(self, tile, brick_log2, batch, resident=0):
        c = BatchConfig(tile, brick_log2, batch, resident)
        self.ctx.check(self.ctx.lib.hipdvr_batch_set_config(self.h, C.byref(c)), "hipdvr_batch_set_config")

    def config(self):
        c = BatchConfig()
        self.ctx.check(self.ctx.lib.hipdvr_batch_get_config(self.h, C.byref(c)), "hipdvr_batch_get_config")
        return (c.tile, c.brick_log2, c.batch, c.resident)

    def autotune(self, reps=3):
        c, ms = BatchConfig(), C.c_float()
        self.ctx.check(self.ctx.lib.hipdvr_batch_autotune(self.h, reps, C.byref(c), C.byref(ms)),
                       "hipdvr_batch_autotune")
        return (c.tile, c.brick_log2, c.batch, c.resident), ms.value

    def capture(self):
        self.ctx.check(self.ctx.lib.hipdvr_batch_capture(self.h), "hipdvr_batch_capture")

    def launch(self, stats=False):
        st = Stats()
        self.ctx.check(self.ctx.lib.hipdvr_batch_launch(self.h, C.byref(st) if stats else None),
                       "hipdvr_batch_launch")
        return st.as_dict() if stats else None

    def set_learned(self, learned, opts=None):
        """learned: one Learned (prepared for that job's TF and step) or None
        per job, in the order the jobs were given."""
        arr = (C.c_void_p * len(learned))(*[l.h.value if l is not None else None for l in learned])
        self.ctx.check(self.ctx.lib.hipdvr_batch_set_learned(self.h, arr, len(learned),
                                                             C.byref(opts) if opts else None),
                       "hipdvr_batch_set_learned")

    def launch_learned(self, stats=False):
        """One batched launch with learned stepping; with stats, one dict per
        job (samples and shaded are the launch's totals)."""
        st = (LearnedStats * self.n)()
        self.ctx.check(self.ctx.lib.hipdvr_batch_launch_learned(self.h, st if stats else None),
                       "hipdvr_batch_launch_learned")
        return [s.as_dict() for s in st] if stats else None

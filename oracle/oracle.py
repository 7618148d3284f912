"""ctypes wrapper of the CPU oracle (oracle/libdvr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker.  The product
package ``voreen_amd`` never imports it.

The oracle restates Voreen's CPURaycaster
(voreen/modules/base/processors/render/cpuraycaster.cpp:90-332) and the
transfer-function, pre-integration and volume-access code it calls; see the
header of dvr_oracle.c for the per-function reference lines and pinning.
"""
import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libdvr_oracle.so")

U8, U16, F32 = 0, 1, 2
FILTER_NEAREST, FILTER_LINEAR = 0, 1
CLASS_TF, CLASS_PREINT, CLASS_TF2D = 0, 1, 2
GRAD_CENTRAL, GRAD_FORWARD, GRAD_NONE, GRAD_FILTERED, GRAD_SOBEL = range(5)
(SHADE_NONE, SHADE_PHONG, SHADE_PHONG_D, SHADE_PHONG_S, SHADE_PHONG_DA, SHADE_PHONG_DS, SHADE_TOON,
 SHADE_COOK_TORRANCE, SHADE_OREN_NAYAR, SHADE_WARD) = range(10)
COMP_DVR, COMP_MIP, COMP_MOP, COMP_MIDA, COMP_ISO, COMP_FHP, COMP_FHN = range(7)
ALPHA_ZERO, ALPHA_USE, ALPHA_ONE = 0, 1, 2


class Volume(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("dtype", C.c_int32), ("data", C.c_void_p),
                ("rwm_scale", C.c_float), ("rwm_offset", C.c_float),
                ("spacing", C.c_float * 3), ("offset", C.c_float * 3),
                ("grad_data", C.c_void_p), ("grad_dtype", C.c_int32), ("grad_channels", C.c_int32)]


class TF(C.Structure):
    _fields_ = [("rgba8", C.c_void_p), ("width", C.c_int32),
                ("domain_lo", C.c_float), ("domain_hi", C.c_float),
                ("rgba2d", C.c_void_p), ("width2d", C.c_int32), ("height2d", C.c_int32),
                ("domain2d", C.c_float * 4)]


class Params(C.Structure):
    _fields_ = [("step", C.c_float), ("filter", C.c_int32), ("classification", C.c_int32),
                ("preint_table", C.c_void_p), ("preint_n", C.c_int32), ("shading", C.c_int32),
                ("light_pos", C.c_float * 3), ("light_ambient", C.c_float * 3),
                ("light_diffuse", C.c_float * 3), ("light_specular", C.c_float * 3),
                ("shininess", C.c_float), ("camera_pos", C.c_float * 3),
                ("loop_cap", C.c_int32), ("compositing", C.c_int32), ("gamma", C.c_float),
                ("iso_value", C.c_float), ("gradient", C.c_int32)]


class TFKey(C.Structure):
    _fields_ = [("intensity", C.c_float), ("split", C.c_int32),
                ("color_l", C.c_uint8 * 4), ("color_r", C.c_uint8 * 4)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("focus", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy_deg", C.c_float), ("near_dist", C.c_float)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(path=None):
    if path is None:
        path = _LIB_PATH
        if not os.path.exists(path):
            build()
    lib = C.CDLL(path)
    P = C.POINTER
    lib.dvr_splitmix64.argtypes = [P(C.c_uint64)]
    lib.dvr_splitmix64.restype = C.c_uint64
    lib.dvr_uniform01.argtypes = [P(C.c_uint64)]
    lib.dvr_uniform01.restype = C.c_double
    lib.dvr_tf_build.argtypes = [P(TFKey), C.c_int32, C.c_int32, C.c_float, C.c_float,
                                 C.c_float, C.c_int32, C.c_void_p]
    lib.dvr_tf_build.restype = None
    lib.dvr_opacity_correct.argtypes = [C.c_float, C.c_float]
    lib.dvr_opacity_correct.restype = C.c_float
    lib.dvr_preint_build.argtypes = [P(TFKey), C.c_int32, C.c_float, C.c_float, C.c_float,
                                     C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_void_p]
    lib.dvr_preint_build.restype = None
    lib.dvr_entry_exit.argtypes = [P(Camera), C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                   C.c_int32, C.c_void_p, C.c_void_p]
    lib.dvr_entry_exit.restype = None
    for fn in (lib.dvr_cpu_render, lib.dvr_fp64_render):
        fn.argtypes = [P(Volume), P(TF), P(Params), C.c_void_p, C.c_void_p, C.c_int32,
                       C.c_int32, C.c_void_p, C.c_int32]
        fn.restype = C.c_int64
    lib.dvr_shade.argtypes = [P(Params), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.dvr_shade.restype = None
    lib.dvr_gradient.argtypes = [P(Volume), P(Params), C.c_void_p, C.c_void_p]
    lib.dvr_gradient.restype = None
    lib.dvr_ray_setup.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p]
    lib.dvr_ray_setup.restype = C.c_float
    lib.dvr_sample_voxel.argtypes = [P(Volume), C.c_void_p, C.c_void_p, C.c_float, C.c_int32, C.c_void_p,
                                     C.c_void_p]
    lib.dvr_sample_voxel.restype = C.c_float
    lib.dvr_sample_linear.argtypes = [P(Volume), C.c_float, C.c_float, C.c_float]
    lib.dvr_sample_linear.restype = C.c_float
    lib.dvr_mlp_forward.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]
    lib.dvr_mlp_forward.restype = None
    lib.dvr_mlp_loss_grad.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                      C.c_void_p]
    lib.dvr_mlp_loss_grad.restype = C.c_double
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def make_keys(keys):
    """keys: list of (intensity, rgba_left[4]) or (intensity, rgba_l, rgba_r)."""
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


def tf_build(keys, width, gamma=1.0, thr=(0.0, 1.0), alpha_mode=ALPHA_USE):
    arr, n = make_keys(keys)
    out = np.zeros((width, 4), np.uint8)
    lib().dvr_tf_build(arr, n, width, gamma, thr[0], thr[1], alpha_mode, _ptr(out))
    return out


def preint_build(keys, n, step, use_integral=False, gamma=1.0, thr=(0.0, 1.0),
                 alpha_mode=ALPHA_USE):
    arr, nk = make_keys(keys)
    out = np.zeros((n, n, 4), np.float32)
    lib().dvr_preint_build(arr, nk, gamma, thr[0], thr[1], alpha_mode, n, step,
                           int(use_integral), _ptr(out))
    return out


def opacity_correct(alpha, step):
    return lib().dvr_opacity_correct(alpha, step)


def entry_exit(eye, focus, up, fovy, near, w, h, llf=(0, 0, 0), urb=(1, 1, 1), quantize=False):
    cam = Camera()
    cam.eye[:] = [float(v) for v in eye]
    cam.focus[:] = [float(v) for v in focus]
    cam.up[:] = [float(v) for v in up]
    cam.fovy_deg = fovy
    cam.near_dist = near
    llf = np.asarray(llf, np.float32)
    urb = np.asarray(urb, np.float32)
    en = np.zeros((h, w, 4), np.float32)
    ex = np.zeros((h, w, 4), np.float32)
    lib().dvr_entry_exit(C.byref(cam), _ptr(llf), _ptr(urb), w, h, int(quantize), _ptr(en), _ptr(ex))
    return en, ex


def _volume(vol, rwm=(1.0, 0.0), spacing=(1.0, 1.0, 1.0), offset=(0.0, 0.0, 0.0)):
    vol = np.ascontiguousarray(vol)
    v = Volume()
    nz, ny, nx = vol.shape
    v.dims[:] = [nx, ny, nz]
    v.dtype = U8 if vol.dtype == np.uint8 else U16
    v.data = vol.ctypes.data
    v.rwm_scale, v.rwm_offset = rwm
    v.spacing[:] = list(spacing)
    v.offset[:] = list(offset)
    return v, vol


def _params(step, filter_, classification, preint, shading, light, compositing=0, gamma=0.0,
            iso_value=0.5):
    p = Params()
    p.compositing = compositing
    p.gamma = gamma
    p.iso_value = iso_value
    p.step = step
    p.filter = filter_
    p.classification = classification
    if preint is not None:
        p.preint_table = preint.ctypes.data
        p.preint_n = preint.shape[0]
    p.shading = shading
    lt = dict(pos=(2.3, 1.5, 1.5), ambient=(0.4,) * 3, diffuse=(0.8,) * 3, specular=(0.6,) * 3,
              shininess=60.0, camera=(0.0, 0.0, 3.5))
    if light:
        lt.update(light)
    p.light_pos[:] = list(lt["pos"])
    p.light_ambient[:] = list(lt["ambient"])
    p.light_diffuse[:] = list(lt["diffuse"])
    p.light_specular[:] = list(lt["specular"])
    p.shininess = lt["shininess"]
    p.camera_pos[:] = list(lt["camera"])
    p.loop_cap = 255 * 255
    return p


def render(vol, tf_rgba8, domain, entry, exit_, step, filter_=FILTER_LINEAR,
           classification=CLASS_TF, preint=None, shading=SHADE_NONE, light=None,
           rwm=(1.0, 0.0), spacing=(1.0, 1.0, 1.0), offset=(0.0, 0.0, 0.0),
           fp64=False, nthreads=0, compositing=COMP_DVR, gamma=0.0, iso_value=0.5,
           gradient=GRAD_CENTRAL, grad_volume=None, tf2d=None, tf2d_domain=((0.0, 1.0), (0.0, 1.0))):
    """Render one frame; vol is a (Z, Y, X) uint8/uint16 array, entry/exit
    (H, W, 4) float32.  The 2D TF path (classification CLASS_TF2D) takes
    grad_volume (Z, Y, X, 3 or 4) uint8/uint16/float32 and tf2d (H2, W2, 4)
    float32 with its (intensity, gradient-magnitude) domains.  Returns
    (rgba (H, W, 4), samples)."""
    v, vol_keep = _volume(vol, rwm, spacing, offset)
    if grad_volume is not None:
        grad_volume = np.ascontiguousarray(grad_volume)
        assert grad_volume.shape[:3] == vol_keep.shape and grad_volume.shape[3] in (3, 4)
        v.grad_data = grad_volume.ctypes.data
        v.grad_dtype = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16,
                        np.dtype(np.float32): F32}[grad_volume.dtype]
        v.grad_channels = grad_volume.shape[3]
    tf_rgba8 = np.ascontiguousarray(np.zeros((2, 4), np.uint8) if tf_rgba8 is None else tf_rgba8, np.uint8)
    t = TF()
    t.rgba8 = tf_rgba8.ctypes.data
    t.width = tf_rgba8.shape[0]
    t.domain_lo, t.domain_hi = domain
    if tf2d is not None:
        tf2d = np.ascontiguousarray(tf2d, np.float32)
        t.rgba2d = tf2d.ctypes.data
        t.height2d, t.width2d = tf2d.shape[:2]
        t.domain2d[:] = [tf2d_domain[0][0], tf2d_domain[0][1], tf2d_domain[1][0], tf2d_domain[1][1]]
    if classification == CLASS_TF2D:
        assert grad_volume is not None and tf2d is not None
    if preint is not None:
        preint = np.ascontiguousarray(preint, np.float32)
    p = _params(step, filter_, classification, preint, shading, light, compositing, gamma, iso_value)
    p.gradient = gradient
    entry = np.ascontiguousarray(entry, np.float32)
    exit_ = np.ascontiguousarray(exit_, np.float32)
    h, w = entry.shape[:2]
    if fp64:
        out = np.zeros((h, w, 4), np.float64)
        n = lib().dvr_fp64_render(C.byref(v), C.byref(t), C.byref(p), _ptr(entry), _ptr(exit_),
                                  w, h, _ptr(out), nthreads)
    else:
        out = np.zeros((h, w, 4), np.float32)
        n = lib().dvr_cpu_render(C.byref(v), C.byref(t), C.byref(p), _ptr(entry), _ptr(exit_),
                                 w, h, _ptr(out), nthreads)
    del vol_keep
    return out, int(n)


def shade(model, g, pos, color, light=None, shininess=60.0):
    """One shading evaluation of the oracle (FP32)."""
    p = _params(1.0, FILTER_LINEAR, CLASS_TF, None, model, light)
    p.shininess = shininess
    g = np.ascontiguousarray(g, np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    color = np.ascontiguousarray(color, np.float32)
    out = np.zeros(3, np.float32)
    lib().dvr_shade(C.byref(p), _ptr(g), _ptr(pos), _ptr(color), _ptr(out))
    return out


def gradient(vol, s, mode, filter_=FILTER_LINEAR, rwm=(1.0, 0.0), spacing=(1.0, 1.0, 1.0)):
    """The oracle's gradient (CALC_GRADIENT of the given mode) at texture-space s."""
    v, keep = _volume(vol, rwm, spacing)
    p = _params(1.0, filter_, CLASS_TF, None, SHADE_PHONG, None)
    p.gradient = mode
    s = np.ascontiguousarray(s, np.float32)
    out = np.zeros(3, np.float32)
    lib().dvr_gradient(C.byref(v), C.byref(p), _ptr(s), _ptr(out))
    del keep
    return out


def ray_setup(first, last, step):
    """(tend, normalized direction) of the oracle's ray setup."""
    f = np.ascontiguousarray(first, np.float32)
    l = np.ascontiguousarray(last, np.float32)
    d = np.zeros(3, np.float32)
    tend = lib().dvr_ray_setup(_ptr(f), _ptr(l), float(step), _ptr(d))
    return np.float32(tend), d


def sample_voxel(dims, first, direction, step, k):
    """(t, clamped voxel position, its rounding) of sample k (dims = x, y, z)."""
    v = Volume()
    v.dims[:] = list(dims)
    f = np.ascontiguousarray(first, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    vp = np.zeros(3, np.float32)
    iv = np.zeros(3, np.int32)
    t = lib().dvr_sample_voxel(C.byref(v), _ptr(f), _ptr(d), float(step), int(k), _ptr(vp), _ptr(iv))
    return np.float32(t), vp, iv


def sample_linear(vol, pos):
    v, keep = _volume(vol)
    return lib().dvr_sample_linear(C.byref(v), float(pos[0]), float(pos[1]), float(pos[2]))


def psnr(a, b):
    """PSNR = 10*log10(1/MSE) over all RGBA channels of images in [0, 1]."""
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0.0 else 10.0 * np.log10(1.0 / mse)


MLP_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2


def mlp_forward(params, x, quantize=True):
    """FP64 forward of the stepping network; params FP32 flat layout."""
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros((x.shape[0], 2), np.float64)
    lib().dvr_mlp_forward(_ptr(params), int(quantize), _ptr(x), x.shape[0], _ptr(y))
    return y


def mlp_loss_grad(params, x, t, quantize=True):
    """FP64 mean loss and its gradient w.r.t. the (FP64) parameters."""
    params = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    g = np.zeros(MLP_PARAMS, np.float64)
    loss = lib().dvr_mlp_loss_grad(_ptr(params), int(quantize), _ptr(x), _ptr(t), x.shape[0], _ptr(g))
    return loss, g


@contextlib.contextmanager
def variant(path):
    """Route every wrapper of this module through another build of the oracle
    (a copy with a planted mistake, built by the mutation tests) for the
    duration of the block."""
    global _lib
    saved = _lib
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = saved

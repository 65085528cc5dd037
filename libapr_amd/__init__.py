"""libapr_amd -- MI355X (gfx950) implementation of LibAPR's APR hot path.

Python binding of the C-ABI in include/aprh.h (libapr_amd/_lib/libaprh.so).
The classes mirror the reference's host interface for this path:

    APRParameters         src/algorithm/APRParameters.hpp:11-70
    APRConverter          src/algorithm/APRConverter.hpp:34-134  (get_apr)
    APR                   src/data_structures/APR/APR.hpp (GenInfo + LinearAccess)
    ParticleData          src/data_structures/APR/particles/ParticleData.hpp (sample_image)

Arrays are numpy (host) or device pointers (``int`` addresses, e.g. from
``torch.Tensor.data_ptr()``); volumes are indexed [z, x, y] with y fastest,
matching PixelData's memory order.

The HIP library is required: importing works without it, but every entry
point raises ``RuntimeError`` when libaprh.so has not been built
(``python -m libapr_amd.build``).  There is no CPU fallback.
"""
import ctypes as C
import weakref
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libaprh.so")

APRH_OK = 0
APRH_ERR_INVALID_ARGUMENT = 1
APRH_ERR_OVERFLOW = 2
APRH_ERR_BAD_DIMS = 3
APRH_ERR_UNSUPPORTED = 4
APRH_ERR_OUT_OF_MEMORY = 5
APRH_ERR_HIP = 6



class AprhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[aprh {code}] {msg}")
        self.code = code


class AprhOverflowError(AprhError, ValueError):
    """u16 IIR under/overflow -- the reference throws std::invalid_argument
    (ComputeGradient.hpp:409-412)."""


class AprhBatchStage(C.Structure):
    _fields_ = [("name", C.c_char_p), ("volumes", C.c_int32), ("tiles", C.c_int64), ("grid", C.c_uint32),
                ("threads", C.c_uint32), ("tuned_ms", C.c_float), ("shape", C.c_char_p), ("n_shapes", C.c_int32)]


class AprhParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("dx", "dy", "dz", "psfx", "psfy", "psfz", "Ip_th", "lambda_", "rel_error",
                                         "sigma_th", "sigma_th_max", "grad_th")] + \
               [(n, C.c_int) for n in ("reflect_bc_lis", "neighborhood_optimization", "constant_intensity_scale")]


class AprhAprInfo(C.Structure):
    _fields_ = [("l_min", C.c_int32), ("l_max", C.c_int32), ("org_dims", C.c_int64 * 3),
                ("total_number_particles", C.c_uint64), ("xz_end_size", C.c_uint64), ("level_xz_size", C.c_uint64),
                ("bspline_offset", C.c_float), ("image_dtype", C.c_int32)]


class AprhTreeInfo(C.Structure):
    _fields_ = [("l_min", C.c_int32), ("l_max", C.c_int32), ("total_number_tree_particles", C.c_uint64),
                ("xz_end_size", C.c_uint64), ("level_xz_size", C.c_uint64)]


class AprhPatch(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("x_begin", "x_end", "y_begin", "y_end", "z_begin", "z_end", "level_delta")]


class AprhRLStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("last_rel_update", C.c_double)]


APRH_DTYPE_U16 = 0
APRH_DTYPE_F32 = 1
CONV_EXACT = 0
CONV_SEPARABLE = 1

_lib = None
vp = C.c_void_p
i64 = C.c_int64


def lib():
    """The loaded libaprh.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libaprh.so not built ({LIB_PATH}); run `python -m libapr_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.aprh_last_error.restype = C.c_char_p
    L.aprh_context_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.aprh_context_destroy.argtypes = [vp]
    L.aprh_context_set_keep_intermediates.argtypes = [vp, C.c_int]
    L.aprh_context_set_timing.argtypes = [vp, C.c_int]
    L.aprh_context_set_conv_mode.argtypes = [vp, C.c_int]
    L.aprh_context_stage_times.argtypes = [vp, C.POINTER(C.c_float), C.c_int]
    L.aprh_stage_name.restype = C.c_char_p
    L.aprh_stage_name.argtypes = [C.c_int]
    L.aprh_context_release_cache.argtypes = [vp]
    L.aprh_get_apr_u16.argtypes = [vp, vp, i64, i64, i64, C.POINTER(AprhParams), C.POINTER(vp)]
    L.aprh_get_apr_sample_u16.argtypes = [vp, vp, i64, i64, i64, C.POINTER(AprhParams), C.POINTER(vp), vp, C.c_uint64,
                                          C.POINTER(C.c_int32)]
    L.aprh_get_apr_f32.argtypes = [vp, vp, i64, i64, i64, C.POINTER(AprhParams), C.POINTER(vp)]
    L.aprh_apr_from_grad_lis.argtypes = [vp, vp, vp, vp, i64, i64, i64, C.POINTER(AprhParams), C.c_float, C.POINTER(vp)]
    L.aprh_apr_info_get.argtypes = [vp, C.POINTER(AprhAprInfo)]
    L.aprh_apr_copy_structure.argtypes = [vp, vp, vp, vp]
    L.aprh_apr_device_structure.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
    L.aprh_apr_intermediate.argtypes = [vp, C.c_int, vp]
    L.aprh_apr_particle_cell_tree.argtypes = [vp, C.c_int, vp, C.POINTER(C.c_int64)]
    L.aprh_apr_destroy.argtypes = [vp]
    L.aprh_sample_image_u16.argtypes = [vp, vp, vp, vp]
    L.aprh_sample_image_u16_noalloc.argtypes = [vp, vp, vp, vp]
    L.aprh_noalloc_scratch_bytes.restype = C.c_size_t
    L.aprh_noalloc_scratch_bytes.argtypes = []
    L.aprh_get_apr_by_block_u16.argtypes = [vp, vp, i64, i64, i64, C.POINTER(AprhParams), i64, i64, C.POINTER(vp)]
    L.aprh_sample_image_blocked_u16.argtypes = [vp, vp, vp, i64, i64, vp]
    L.aprh_bspline_filt_rec_u16.argtypes = [vp, vp, i64, i64, i64, C.c_int, C.c_float, C.c_float]
    L.aprh_bspline_filt_rec_f32.argtypes = [vp, vp, i64, i64, i64, C.c_int, C.c_float, C.c_float]
    L.aprh_inv_bspline_f32.argtypes = [vp, vp, i64, i64, i64, C.c_int]
    L.aprh_gradient_ds_u16.argtypes = [vp, vp, vp, vp, i64, i64, i64, C.c_float, C.c_float, C.c_float]
    L.aprh_sat_mean_f32.argtypes = [vp, vp, i64, i64, i64, C.c_int, C.c_int]
    L.aprh_local_intensity_scale.argtypes = [vp, vp, vp, i64, i64, i64, C.POINTER(AprhParams)]
    L.aprh_synth_blobs_host.argtypes = [vp, i64, i64, i64, C.c_uint64, i64, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_int]
    L.aprh_synth_blobs_device.argtypes = [vp, vp, i64, i64, i64, C.c_uint64, i64, C.c_float, C.c_float, C.c_float,
                                          C.c_float, C.c_float, C.c_int]
    i32p = C.POINTER(C.c_int32)
    L.aprh_apr_from_structure.argtypes = [vp, i64, i64, i64, vp, C.c_uint64, vp, C.c_uint64, C.POINTER(vp)]
    L.aprh_apr_tree_info.argtypes = [vp, vp, C.POINTER(AprhTreeInfo)]
    L.aprh_apr_copy_tree.argtypes = [vp, vp, vp, vp, vp]
    L.aprh_fill_tree_mean.argtypes = [vp, vp, vp, C.c_int, vp]
    L.aprh_downsampled_stencils.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int, vp, i64, vp,
                                            C.POINTER(C.c_int64)]
    L.aprh_gaussian_stencil.argtypes = [vp, vp, C.c_int, vp]
    L.aprh_mean_stencil.argtypes = [vp, vp]
    for f in (L.aprh_convolve_pencil, L.aprh_convolve):
        f.argtypes = [vp, vp, vp, vp, C.c_int32, vp, C.c_int, vp, C.c_int]
    L.aprh_convolve_stencil.argtypes = [vp, vp, vp, C.c_int32, C.c_int32, C.c_int32, vp, C.c_int, vp, C.c_int, C.c_int,
                                        C.c_int]
    L.aprh_richardson_lucy.argtypes = [vp, vp, vp, C.c_int, vp, vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int,
                                       C.c_int, C.c_int, C.c_double, C.POINTER(AprhRLStats)]
    L.aprh_richardson_lucy_tv.argtypes = [vp, vp, vp, C.c_int, vp, vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                          C.c_int, C.c_int, C.c_int]
    L.aprh_reconstruct_constant.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, C.c_int]
    L.aprh_reconstruct_level.argtypes = [vp, vp, vp, C.c_int]
    L.aprh_reconstruct_smooth.argtypes = [vp, vp, vp, C.c_int, C.POINTER(C.c_float), vp, C.c_int]
    L.aprh_patch_check_limits.argtypes = [vp, C.POINTER(AprhPatch)]
    for f in ("aprh_gradient_cfd", "aprh_gradient_sobel"):
        getattr(L, f).argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_float, vp]
    for f in ("aprh_gradient_magnitude_cfd", "aprh_gradient_magnitude_sobel"):
        getattr(L, f).argtypes = [vp, vp, vp, C.c_int, vp, vp]
    L.aprh_local_std.argtypes = [vp, vp, vp, C.c_int, vp, vp]
    L.aprh_face_neighbour_filter.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, vp]
    L.aprh_seperable_face_neighbour_filter.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int32, vp]
    L.aprh_reconstruct_constant_patch.argtypes = [vp, vp, vp, C.c_int, C.POINTER(AprhPatch), vp, C.c_int]
    L.aprh_reconstruct_level_patch.argtypes = [vp, vp, C.POINTER(AprhPatch), vp, C.c_int]
    L.aprh_synth_deconv_device.argtypes = [vp, vp, vp, i64, i64, i64, C.c_uint64, i64, C.c_float, C.c_float, C.c_float,
                                           C.c_float, C.c_float, vp, C.c_int32]
    L.aprh_synth_deconv_host.argtypes = [vp, vp, i64, i64, i64, C.c_uint64, i64, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, vp, C.c_int32]
    L.aprh_sample_image_f32.argtypes = [vp, vp, vp, vp]
    L.aprh_cnn_create.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.aprh_cnn_destroy.argtypes = [vp]
    L.aprh_cnn_get_params.argtypes = [vp, vp, vp]
    L.aprh_cnn_set_params.argtypes = [vp, vp, vp]
    L.aprh_cnn_forward_tiles.argtypes = [vp, vp, vp, C.c_int32, C.c_int32, vp]
    L.aprh_cnn_train_step.argtypes = [vp, vp, vp, vp, C.c_int32, C.c_int32, C.c_float, C.POINTER(C.c_float), vp]
    L.aprh_cnn_extract_tiles.argtypes = [vp, vp, C.c_int, i64, i64, i64, vp, C.c_int32, C.c_int32, C.c_int32, C.c_float, vp]
    L.aprh_cnn_apply.argtypes = [vp, vp, vp, C.c_int, i64, i64, i64, C.c_float, vp]
    L.aprh_cnn_warm_start.argtypes = [vp, vp, vp, vp, C.c_int, vp]
    L.aprh_cnn_set_post_filter.argtypes = [vp, vp, vp]
    L.aprh_mean.argtypes = [vp, vp, C.c_int, C.c_uint64, C.POINTER(C.c_double)]
    L.aprh_batch_create.argtypes = [vp, C.c_int32, vp, C.POINTER(AprhParams), C.POINTER(vp)]
    L.aprh_batch_input.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    L.aprh_batch_run.argtypes = [vp, vp]
    L.aprh_batch_get.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    L.aprh_batch_destroy.argtypes = [vp]
    L.aprh_batch_stage_count.argtypes = [vp]
    L.aprh_batch_stage_info.argtypes = [vp, C.c_int32, C.POINTER(AprhBatchStage)]
    L.aprh_malloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    L.aprh_free.argtypes = [vp, vp]
    L.aprh_memcpy.argtypes = [vp, vp, vp, C.c_size_t]
    _lib = L
    return L


def _check(code):
    if code != APRH_OK:
        msg = lib().aprh_last_error().decode(errors="replace")
        if code == APRH_ERR_OVERFLOW:
            raise AprhOverflowError(code, msg)
        raise AprhError(code, msg)


def _ptr(a):
    """host numpy array -> pointer; int -> device address"""
    if isinstance(a, (int, np.integer)):
        return C.c_void_p(int(a))
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


class APRParameters:
    """APRParameters (APRParameters.hpp:11-70); defaults of the reference."""

    def __init__(self, **kw):
        self.dx = self.dy = self.dz = 1.0
        self.psfx = self.psfy = self.psfz = 2.0
        self.Ip_th = 0.0
        self.lambda_ = 3.0
        self.rel_error = 0.1
        self.sigma_th = 0.0
        self.sigma_th_max = 0.0
        self.grad_th = 0.0
        self.reflect_bc_lis = True
        self.neighborhood_optimization = True
        self.constant_intensity_scale = False
        for k, v in kw.items():
            setattr(self, "lambda_" if k == "lambda" else k, v)

    def c(self):
        return AprhParams(self.dx, self.dy, self.dz, self.psfx, self.psfy, self.psfz, self.Ip_th, self.lambda_,
                          self.rel_error, self.sigma_th, self.sigma_th_max, self.grad_th, int(self.reflect_bc_lis),
                          int(self.neighborhood_optimization), int(self.constant_intensity_scale))


def _register(ctx, obj):
    """obj holds device state of ctx: Context.close() closes it first"""
    deps = getattr(ctx, "_deps", None)
    if deps is not None:
        deps.add(obj)


class Context:
    """One HIP stream on one device (aprh_context)."""

    def __init__(self, device=0, keep_intermediates=False, timing=False):
        h = vp()
        _check(lib().aprh_context_create(device, C.byref(h)))
        self.h = h
        self.device = device
        # objects holding device state of this context: closed before the context is destroyed
        self._deps = weakref.WeakSet()
        if keep_intermediates:
            _check(lib().aprh_context_set_keep_intermediates(h, 1))
        if timing:
            _check(lib().aprh_context_set_timing(h, 1))

    def set_keep_intermediates(self, on):
        _check(lib().aprh_context_set_keep_intermediates(self.h, int(on)))

    def set_timing(self, on):
        _check(lib().aprh_context_set_timing(self.h, int(on)))

    def set_conv_mode(self, mode):
        """CONV_EXACT (reference summation order, bit-equal) or CONV_SEPARABLE (rank-1 stencils
        as three 1-D passes, within float rounding of exact)"""
        _check(lib().aprh_context_set_conv_mode(self.h, int(mode)))

    def stage_times(self):
        n = lib().aprh_stage_count()
        ms = (C.c_float * n)()
        _check(lib().aprh_context_stage_times(self.h, ms, n))
        return {lib().aprh_stage_name(i).decode(): ms[i] for i in range(n)}

    def release_cache(self):
        _check(lib().aprh_context_release_cache(self.h))

    def close(self):
        if getattr(self, "h", None):
            for d in list(self._deps):
                d.close()
            lib().aprh_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class APR:
    """Device-resident APR: GenInfo + LinearAccess (APR.hpp:18-61)."""

    def __init__(self, ctx, handle, owner=None):
        self.ctx, self.h = ctx, handle
        self._owner = owner            # a batch plan that owns the handle (never destroyed here)
        _register(ctx, self)
        info = AprhAprInfo()
        _check(lib().aprh_apr_info_get(handle, C.byref(info)))
        self.l_min, self.l_max = info.l_min, info.l_max
        self.org_dims = tuple(info.org_dims)          # (y, x, z)
        self.total = info.total_number_particles
        self.xz_end_size = info.xz_end_size
        self.level_xz_size = info.level_xz_size
        self.bspline_offset = info.bspline_offset
        self.image_dtype = np.float32 if info.image_dtype == 1 else np.uint16
        self._struct = None

    def total_number_particles(self):
        return self.total

    def level_min(self):
        return self.l_min

    def level_max(self):
        return self.l_max

    def structure(self):
        """(level_xz_vec u64, xz_end_vec u64, y_vec u16) copied to the host"""
        if self._struct is None:
            lxz = np.zeros(self.level_xz_size, np.uint64)
            xze = np.zeros(self.xz_end_size, np.uint64)
            yv = np.zeros(max(self.total, 1), np.uint16)
            _check(lib().aprh_apr_copy_structure(self.h, _ptr(lxz), _ptr(xze), _ptr(yv)))
            self._struct = (lxz, xze, yv[:self.total])
        return self._struct

    def device_structure(self):
        xze, yv = vp(), vp()
        _check(lib().aprh_apr_device_structure(self.h, C.byref(xze), C.byref(yv)))
        return xze.value, yv.value

    def intermediate(self, which):
        y, x, z = self.org_dims
        yd, xd, zd = (y + 1) // 2, (x + 1) // 2, (z + 1) // 2
        shapes = {0: ((z, x, y), np.uint16), 1: ((zd, xd, yd), np.uint16), 2: ((zd, xd, yd), np.float32),
                  3: ((zd, xd, yd), np.uint16), 4: ((zd, xd, yd), np.float32), 5: ((zd, xd, yd), np.float32)}
        shp, dt = shapes[which]
        if which in (0, 1, 3):
            dt = self.image_dtype
        out = np.zeros(shp, dt)
        _check(lib().aprh_apr_intermediate(self.h, which, _ptr(out)))
        return out

    def particle_cell_tree(self, level):
        dims = (C.c_int64 * 3)()
        _check(lib().aprh_apr_particle_cell_tree(self.h, level, None, dims))
        y, x, z = dims
        out = np.zeros((z, x, y), np.uint8)
        _check(lib().aprh_apr_particle_cell_tree(self.h, level, _ptr(out), dims))
        return out

    @classmethod
    def from_structure(cls, dims, xz_end_vec, y_vec, ctx=None):
        """APR from a stored LinearAccess (xz_end_vec u64, y_vec u16) of an image with dims (y, x, z),
        as APRFile::read_apr restores it; host arrays or device addresses (then pass sizes via
        numpy-like objects).  The structure is validated."""
        ctx = ctx or default_context()
        y, x, z = dims
        xze = np.ascontiguousarray(xz_end_vec, np.uint64)
        yv = np.ascontiguousarray(y_vec, np.uint16)
        h = vp()
        _check(lib().aprh_apr_from_structure(ctx.h, y, x, z, _ptr(xze), xze.size, _ptr(yv) if yv.size else None, yv.size,
                                             C.byref(h)))
        return cls(ctx, h)

    def tree_info(self):
        info = AprhTreeInfo()
        _check(lib().aprh_apr_tree_info(self.ctx.h, self.h, C.byref(info)))
        return info

    def total_number_tree_particles(self):
        return self.tree_info().total_number_tree_particles

    def tree(self):
        """tree (interior cell) access: dict with l_min, l_max, total, level_xz_vec, xz_end_vec, y_vec"""
        info = self.tree_info()
        lxz = np.zeros(info.level_xz_size, np.uint64)
        xze = np.zeros(info.xz_end_size, np.uint64)
        yv = np.zeros(max(info.total_number_tree_particles, 1), np.uint16)
        _check(lib().aprh_apr_copy_tree(self.ctx.h, self.h, _ptr(lxz), _ptr(xze), _ptr(yv)))
        return {"l_min": info.l_min, "l_max": info.l_max, "total": info.total_number_tree_particles,
                "level_xz_vec": lxz, "xz_end_vec": xze, "y_vec": yv[:info.total_number_tree_particles]}

    def close(self):
        if getattr(self, "h", None):
            if self._owner is None:
                lib().aprh_apr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class APRConverter:
    """APRConverter<ImageType> (APRConverter.hpp:34-134) for ImageType uint16 (default) or float32."""

    def __init__(self, ctx=None, image_type=np.uint16):
        self.par = APRParameters()
        self.ctx = ctx or default_context()
        self.image_type = np.dtype(image_type)
        if self.image_type not in (np.dtype(np.uint16), np.dtype(np.float32)):
            raise TypeError("APRConverter supports uint16 and float32 images")

    def get_apr(self, image, dims=None):
        """image: numpy [z, x, y] of the converter's image type, or a device address with dims=(y, x, z)"""
        if dims is None:
            z, x, y = image.shape
            if image.dtype != self.image_type:
                raise TypeError(f"APRConverter<{self.image_type}> needs a {self.image_type} image")
        else:
            y, x, z = dims
        h = vp()
        p = self.par.c()
        f = lib().aprh_get_apr_f32 if self.image_type == np.float32 else lib().aprh_get_apr_u16
        _check(f(self.ctx.h, _ptr(image), y, x, z, C.byref(p), C.byref(h)))
        return APR(self.ctx, h)

    def get_apr_and_sample(self, image, parts, capacity, dims=None):
        """get_apr followed by ParticleData.sample_image of the same u16 image in one call.  parts is
        a device address with room for `capacity` u16 particles; returns (apr, sampled), sampled
        False when the APR has more particles (then sample with ParticleData.sample_image)."""
        if self.image_type != np.uint16:
            raise TypeError("get_apr_and_sample converts u16 images")
        if dims is None:
            z, x, y = image.shape
            if image.dtype != np.uint16:
                raise TypeError("get_apr_and_sample needs a uint16 image")
        else:
            y, x, z = dims
        h = vp()
        p = self.par.c()
        done = C.c_int32(0)
        _check(lib().aprh_get_apr_sample_u16(self.ctx.h, _ptr(image), y, x, z, C.byref(p), C.byref(h), parts, capacity,
                                             C.byref(done)))
        return APR(self.ctx, h), bool(done.value)

    def get_apr_custom_grad_scale(self, grad_ds, lis_ds, intensity_ds, dims, bspline_offset=0.0):
        y, x, z = dims
        h = vp()
        p = self.par.c()
        _check(lib().aprh_apr_from_grad_lis(self.ctx.h, _ptr(np.ascontiguousarray(grad_ds, np.uint16)),
                                            _ptr(np.ascontiguousarray(lis_ds, np.float32)),
                                            _ptr(np.ascontiguousarray(intensity_ds, np.float32)), y, x, z, C.byref(p),
                                            bspline_offset, C.byref(h)))
        return APR(self.ctx, h)


class APRBatch:
    """Conversion of a fixed list of u16 volumes of uneven sizes, recorded once into one HIP graph
    (APRConverter<uint16_t>::get_apr per volume, APRConverter.hpp:534-600).

    dims: list of (y, x, z).  set_input(i, image) writes volume i ([z, x, y] numpy u16 or a device
    address); run() replays the graph; get(i) returns volume i's APR (borrowed from the plan: valid
    until the next run() or close())."""

    def __init__(self, dims, par=None, ctx=None):
        self.ctx = ctx or default_context()
        self.dims = [tuple(int(v) for v in d) for d in dims]
        self.par = par or APRParameters()
        d = np.ascontiguousarray(np.array(self.dims, np.int64).reshape(-1))
        p = self.par.c()
        h = vp()
        _check(lib().aprh_batch_create(self.ctx.h, len(self.dims), _ptr(d), C.byref(p), C.byref(h)))
        self.h = h
        _register(self.ctx, self)
        self.voxels = int(sum(y * x * z for y, x, z in self.dims))

    def input_address(self, i):
        out = vp()
        _check(lib().aprh_batch_input(self.h, i, C.byref(out)))
        return out.value

    def set_input(self, i, image):
        y, x, z = self.dims[i]
        if isinstance(image, np.ndarray):
            if image.dtype != np.uint16 or image.shape != (z, x, y):
                raise ValueError(f"volume {i} must be uint16 of shape {(z, x, y)}")
            image = np.ascontiguousarray(image)
        _check(lib().aprh_memcpy(self.ctx.h, self.input_address(i), _ptr(image), 2 * y * x * z))

    def run(self):
        _check(lib().aprh_batch_run(self.ctx.h, self.h))

    def stages(self):
        """the merged stage launches with the autotuned tile shape and grid of each"""
        out = []
        for i in range(lib().aprh_batch_stage_count(self.h)):
            st = AprhBatchStage()
            _check(lib().aprh_batch_stage_info(self.h, i, C.byref(st)))
            out.append({"name": st.name.decode(), "volumes": st.volumes, "tiles": st.tiles, "grid": st.grid,
                        "threads": st.threads, "tuned_ms": round(st.tuned_ms, 4),
                        "shape": st.shape.decode() if st.shape else None, "shapes_tried": st.n_shapes})
        return out

    def get(self, i):
        out = vp()
        _check(lib().aprh_batch_get(self.h, i, C.byref(out)))
        return APR(self.ctx, out, owner=self)

    def close(self):
        if getattr(self, "h", None):
            lib().aprh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class APRConverterBatch:
    """APRConverterBatch<uint16_t> (APRConverterBatch.hpp:32-175): get_apr over z blocks of
    z_block_size planes with ghost_z ghost planes on each side (the reference reads the blocks from
    a TIFF file; here the image is a numpy [z, x, y] uint16 array or a device address with dims)."""

    def __init__(self, ctx=None):
        self.par = APRParameters()
        self.ctx = ctx or default_context()
        self.z_block_size = 128
        self.ghost_z = 32

    def get_apr(self, image, dims=None):
        if dims is None:
            z, x, y = image.shape
            if image.dtype != np.uint16:
                raise TypeError("APRConverterBatch<uint16_t> needs a uint16 image")
        else:
            y, x, z = dims
        h = vp()
        p = self.par.c()
        _check(lib().aprh_get_apr_by_block_u16(self.ctx.h, _ptr(image), y, x, z, C.byref(p), int(self.z_block_size),
                                               int(self.ghost_z), C.byref(h)))
        return APR(self.ctx, h)


class ParticleData:
    """ParticleData<T>::sample_image (ParticleData.hpp:122-127) for uint16 and float images."""

    def __init__(self):
        self.data = None

    def sample_image(self, apr, image, out=None, dtype=None):
        if dtype is None and isinstance(image, (int, np.integer)):
            dtype = "uint16"           # a raw device address: the uint16 image of APRConverter<uint16_t>
        code = _dtype_of(image, dtype)
        if out is None:
            out = np.zeros(max(apr.total, 1), np.uint16 if code == APRH_DTYPE_U16 else np.float32)
        f = lib().aprh_sample_image_u16 if code == APRH_DTYPE_U16 else lib().aprh_sample_image_f32
        _check(f(apr.ctx.h, apr.h, _ptr(image), _ptr(out)))
        self.data = out[:apr.total] if isinstance(out, np.ndarray) else out
        return self.data

    def sample_parts_from_img_blocked(self, apr, image, z_block_size=256, ghost_z=32, out=None):
        """ParticleData::sample_parts_from_img_blocked (ParticleData.hpp:140-141, 267-362) of a uint16
        image (numpy [z, x, y] or a device address)"""
        if out is None:
            out = np.zeros(max(apr.total, 1), np.uint16)
        _check(lib().aprh_sample_image_blocked_u16(apr.ctx.h, apr.h, _ptr(image), int(z_block_size), int(ghost_z), _ptr(out)))
        self.data = out[:apr.total] if isinstance(out, np.ndarray) else out
        return self.data


# ---------------- stage-level entry points ----------------

def bspline_filt_rec(image, axis, lam=3.0, tol=1e-4, ctx=None):
    """in place on a numpy [z, x, y] array (u16 or f32)"""
    ctx = ctx or default_context()
    z, x, y = image.shape
    f = lib().aprh_bspline_filt_rec_u16 if image.dtype == np.uint16 else lib().aprh_bspline_filt_rec_f32
    _check(f(ctx.h, _ptr(image), y, x, z, axis, lam, tol))
    return image


def inv_bspline(image, axis, ctx=None):
    ctx = ctx or default_context()
    z, x, y = image.shape
    _check(lib().aprh_inv_bspline_f32(ctx.h, _ptr(image), y, x, z, axis))
    return image


def gradient_ds(image, h=(1.0, 1.0, 1.0), ctx=None, with_mean=True):
    ctx = ctx or default_context()
    z, x, y = image.shape
    g = np.zeros(((z + 1) // 2, (x + 1) // 2, (y + 1) // 2), np.uint16)
    m = np.zeros(g.shape, np.float32) if with_mean else None
    _check(lib().aprh_gradient_ds_u16(ctx.h, _ptr(image), _ptr(g), _ptr(m) if m is not None else None, y, x, z,
                                      h[0], h[1], h[2]))
    return g, m


def sat_mean(image, axis, offset, ctx=None):
    ctx = ctx or default_context()
    z, x, y = image.shape
    _check(lib().aprh_sat_mean_f32(ctx.h, _ptr(image), y, x, z, axis, offset))
    return image


def local_intensity_scale(intensity, par=None, ctx=None):
    ctx = ctx or default_context()
    par = par or APRParameters()
    z, x, y = intensity.shape
    out = np.zeros_like(intensity, dtype=np.float32)
    p = par.c()
    _check(lib().aprh_local_intensity_scale(ctx.h, _ptr(np.ascontiguousarray(intensity, np.float32)), _ptr(out), y, x,
                                            z, C.byref(p)))
    return out


def synth_blobs(shape_zxy, seed, n_blobs, sigma=(2.0, 6.0), amp=(500.0, 3000.0), background=100.0, noise=True,
                out=None, ctx=None):
    """Synthetic uint16 blob volume; host (numpy) or device (out = device address)."""
    z, x, y = shape_zxy
    if out is None or isinstance(out, np.ndarray):
        if out is None:
            out = np.zeros((z, x, y), np.uint16)
        _check(lib().aprh_synth_blobs_host(_ptr(out), y, x, z, seed, n_blobs, sigma[0], sigma[1], amp[0], amp[1],
                                           background, int(noise)))
        return out
    ctx = ctx or default_context()
    _check(lib().aprh_synth_blobs_device(ctx.h, _ptr(out), y, x, z, seed, n_blobs, sigma[0], sigma[1], amp[0], amp[1],
                                         background, int(noise)))
    return out


# ---------------- particle operations ----------------

def _dtype_of(parts, dtype):
    """aprh dtype code of a particle buffer (numpy array, torch tensor or device address + dtype)"""
    if dtype is None:
        dt = getattr(parts, "dtype", None)
        if dt is None:
            raise TypeError("pass dtype= for a raw device address")
        dtype = str(dt).replace("torch.", "")
    dtype = str(np.dtype(dtype)) if not isinstance(dtype, str) else dtype
    if dtype in ("uint16", "u16"):
        return APRH_DTYPE_U16
    if dtype in ("float32", "f32", "float"):
        return APRH_DTYPE_F32
    raise TypeError(f"particles must be uint16 or float32, not {dtype}")


def _host_parts(parts):
    if isinstance(parts, np.ndarray):
        if parts.dtype not in (np.uint16, np.float32):
            parts = parts.astype(np.float32)
        return np.ascontiguousarray(parts)
    return parts


def _stencil_list(stencils):
    if isinstance(stencils, np.ndarray):
        stencils = [stencils]
    flat, dims = [], []
    for st in stencils:
        st = np.ascontiguousarray(st, np.float32)
        if st.ndim != 3:
            raise ValueError("stencils are 3-D arrays indexed [z, x, y]")
        z, x, y = st.shape
        flat.append(st.reshape(-1))
        dims += [y, x, z]
    return np.concatenate(flat), np.array(dims, np.int32), len(stencils)


class APRTreeNumerics:
    """APRTreeNumerics (APRTreeNumerics.hpp:19, :63-202)."""

    @staticmethod
    def fill_tree_mean(apr, particle_data, tree_data=None, dtype=None):
        parts = _host_parts(particle_data)
        code = _dtype_of(parts, dtype)
        if tree_data is None:
            tree_data = np.zeros(max(apr.total_number_tree_particles(), 1), np.float32)
        _check(lib().aprh_fill_tree_mean(apr.ctx.h, apr.h, _ptr(parts), code, _ptr(tree_data)))
        if isinstance(tree_data, np.ndarray):
            return tree_data[:apr.total_number_tree_particles()]
        return tree_data


class APRStencil:
    """APRStencil (APRStencil.hpp:12-244).  Stencils are float32 arrays indexed [z, x, y]."""

    @staticmethod
    def get_downsampled_stencils(stencil, nlevels, normalize=False):
        st = np.ascontiguousarray(stencil, np.float32)
        z, x, y = st.shape
        need = C.c_int64()
        _check(lib().aprh_downsampled_stencils(_ptr(st), y, x, z, nlevels, int(normalize), None, 0, None, C.byref(need)))
        out = np.zeros(need.value, np.float32)
        dims = np.zeros(3 * max(nlevels, 1), np.int32)
        _check(lib().aprh_downsampled_stencils(_ptr(st), y, x, z, nlevels, int(normalize), _ptr(out), out.size, _ptr(dims),
                                               C.byref(need)))
        res, off = [], 0
        for i in range(max(nlevels, 1)):
            yy, xx, zz = dims[3 * i:3 * i + 3]
            n = int(yy) * int(xx) * int(zz)
            res.append(out[off:off + n].reshape(zz, xx, yy).copy())
            off += n
        return res

    @staticmethod
    def create_gaussian_filter(sigma=(1.0, 1.0, 1.0), size=(5, 5, 5), normalize=True):
        """sigma and size in the reference's (y, x, z) order; result indexed [z, x, y]"""
        sg = np.array(sigma, np.float32)
        sz = np.array(size, np.int32)
        out = np.zeros((int(sz[2]), int(sz[1]), int(sz[0])), np.float32)
        _check(lib().aprh_gaussian_stencil(_ptr(sg), _ptr(sz), int(normalize), _ptr(out)))
        return out

    @staticmethod
    def create_mean_filter(size=(5, 5, 5)):
        sz = np.array(size, np.int32)
        out = np.zeros((int(sz[2]), int(sz[1]), int(sz[0])), np.float32)
        _check(lib().aprh_mean_stencil(_ptr(sz), _ptr(out)))
        return out


class APRFilter:
    """APRFilter::convolve / convolve_pencil (APRFilter.hpp:67-149).

    `stencils` is one [z, x, y] array (then prepared with get_downsampled_stencils unless
    use_stencil_downsample is False) or a list of arrays (stencils[0] at level l_max)."""

    @staticmethod
    def convolve_pencil(apr, stencils, particle_input, particle_output=None, reflect_boundary=True,
                        use_stencil_downsample=True, normalize=False, dtype=None):
        parts = _host_parts(particle_input)
        code = _dtype_of(parts, dtype)
        if particle_output is None:
            particle_output = np.zeros(max(apr.total, 1), np.float32)
        if isinstance(stencils, np.ndarray):
            st = np.ascontiguousarray(stencils, np.float32)
            z, x, y = st.shape
            _check(lib().aprh_convolve_stencil(apr.ctx.h, apr.h, _ptr(st), y, x, z, _ptr(parts), code,
                                               _ptr(particle_output), int(reflect_boundary), int(use_stencil_downsample),
                                               int(normalize)))
        else:
            flat, dims, n = _stencil_list(stencils)
            _check(lib().aprh_convolve_pencil(apr.ctx.h, apr.h, _ptr(flat), _ptr(dims), n, _ptr(parts), code,
                                              _ptr(particle_output), int(reflect_boundary)))
        if isinstance(particle_output, np.ndarray):
            return particle_output[:apr.total]
        return particle_output

    convolve = convolve_pencil


class APRNumerics:
    """APRNumerics::richardson_lucy (APRNumerics.hpp:130-134, :549-593)."""

    @staticmethod
    def _gradient(fn, apr, particle_input, particle_output, dimension, delta, dtype):
        parts = _host_parts(particle_input)
        code = _dtype_of(parts, dtype)
        if particle_output is None:
            particle_output = np.zeros(max(apr.total, 1), np.float32)
        _check(getattr(lib(), fn)(apr.ctx.h, apr.h, _ptr(parts), code, int(dimension), float(delta), _ptr(particle_output)))
        return particle_output[:apr.total] if isinstance(particle_output, np.ndarray) else particle_output

    @staticmethod
    def gradient_cfd(apr, particle_input, particle_output=None, dimension=0, delta=1.0, dtype=None):
        """APRNumerics::gradient_cfd (APRNumerics.hpp:165-184); dimension 0 y, 1 x, 2 z"""
        return APRNumerics._gradient("aprh_gradient_cfd", apr, particle_input, particle_output, dimension, delta, dtype)

    @staticmethod
    def gradient_sobel(apr, particle_input, particle_output=None, dimension=0, delta=1.0, dtype=None):
        """APRNumerics::gradient_sobel (APRNumerics.hpp:187-213)"""
        return APRNumerics._gradient("aprh_gradient_sobel", apr, particle_input, particle_output, dimension, delta, dtype)

    @staticmethod
    def _magnitude(fn, apr, particle_input, particle_output, deltas, dtype):
        parts = _host_parts(particle_input)
        code = _dtype_of(parts, dtype)
        if particle_output is None:
            particle_output = np.zeros(max(apr.total, 1), np.float32)
        d = np.ascontiguousarray(deltas, np.float32)
        if d.size != 3:
            raise ValueError("deltas: three pixel sizes (y, x, z)")
        _check(getattr(lib(), fn)(apr.ctx.h, apr.h, _ptr(parts), code, _ptr(d), _ptr(particle_output)))
        return particle_output[:apr.total] if isinstance(particle_output, np.ndarray) else particle_output

    @staticmethod
    def gradient_magnitude_cfd(apr, particle_input, particle_output=None, deltas=(1.0, 1.0, 1.0), dtype=None):
        """APRNumerics::gradient_magnitude_cfd (APRNumerics.hpp:216-248)"""
        return APRNumerics._magnitude("aprh_gradient_magnitude_cfd", apr, particle_input, particle_output, deltas, dtype)

    @staticmethod
    def gradient_magnitude_sobel(apr, particle_input, particle_output=None, deltas=(1.0, 1.0, 1.0), dtype=None):
        """APRNumerics::gradient_magnitude_sobel (APRNumerics.hpp:251-283)"""
        return APRNumerics._magnitude("aprh_gradient_magnitude_sobel", apr, particle_input, particle_output, deltas, dtype)

    @staticmethod
    def local_std(apr, particle_input, particle_output=None, size=(3, 3, 3), dtype=None):
        """APRNumerics::local_std (APRNumerics.hpp:357-389); size in (y, x, z)"""
        parts = _host_parts(particle_input)
        code = _dtype_of(parts, dtype)
        if particle_output is None:
            particle_output = np.zeros(max(apr.total, 1), np.float32)
        sz = np.ascontiguousarray(size, np.int32)
        if sz.size != 3:
            raise ValueError("size: three window sizes (y, x, z)")
        _check(lib().aprh_local_std(apr.ctx.h, apr.h, _ptr(parts), code, _ptr(sz), _ptr(particle_output)))
        return particle_output[:apr.total] if isinstance(particle_output, np.ndarray) else particle_output

    @staticmethod
    def face_neighbour_filter(apr, particle_input, filter, dimension, particle_output=None, dtype=None):
        """APRNumerics::face_neighbour_filter (APRNumerics.hpp:305-355); filter = (-face, centre, +face)"""
        parts = _host_parts(particle_input)
        code = _dtype_of(parts, dtype)
        if particle_output is None:
            particle_output = np.zeros(max(apr.total, 1), np.float32)
        f = np.ascontiguousarray(filter, np.float32)
        _check(lib().aprh_face_neighbour_filter(apr.ctx.h, apr.h, _ptr(parts), code, _ptr(f), int(dimension),
                                                _ptr(particle_output)))
        return particle_output[:apr.total] if isinstance(particle_output, np.ndarray) else particle_output

    @staticmethod
    def seperable_face_neighbour_filter(apr, particle_input, filter, repeats=1, particle_output=None, dtype=None):
        """APRNumerics::seperable_face_neighbour_filter (APRNumerics.hpp:286-302)"""
        parts = _host_parts(particle_input)
        code = _dtype_of(parts, dtype)
        if particle_output is None:
            particle_output = np.zeros(max(apr.total, 1), np.float32)
        f = np.ascontiguousarray(filter, np.float32)
        _check(lib().aprh_seperable_face_neighbour_filter(apr.ctx.h, apr.h, _ptr(parts), code, _ptr(f), int(repeats),
                                                          _ptr(particle_output)))
        return particle_output[:apr.total] if isinstance(particle_output, np.ndarray) else particle_output

    @staticmethod
    def richardson_lucy(apr, particle_input, particle_output, psf, number_iterations, use_stencil_downsample=True,
                        normalize=False, resume=False, tol=0.0, dtype=None):
        """particle_output: float32[total] (host or device) or None; returns (output, stats).
        tol > 0 stops after the first iteration whose relative update is below tol."""
        parts = _host_parts(particle_input)
        code = _dtype_of(parts, dtype)
        if particle_output is None:
            if resume:
                raise ValueError("resume needs the previous estimate")
            particle_output = np.zeros(max(apr.total, 1), np.float32)
        st = np.ascontiguousarray(psf, np.float32)
        z, x, y = st.shape
        stats = AprhRLStats()
        _check(lib().aprh_richardson_lucy(apr.ctx.h, apr.h, _ptr(parts), code, _ptr(particle_output), _ptr(st), y, x, z,
                                          int(number_iterations), int(use_stencil_downsample), int(normalize),
                                          int(resume), float(tol), C.byref(stats)))
        out = particle_output[:apr.total] if isinstance(particle_output, np.ndarray) else particle_output
        return out, {"iterations": stats.iterations, "last_rel_update": stats.last_rel_update}

    @staticmethod
    def richardson_lucy_tv(apr, particle_input, particle_output, psf, number_iterations, reg_factor, use_stencil_downsample=True,
                           normalize=False, resume=False, dtype=None):
        """APRNumerics::richardson_lucy_tv (APRNumerics.hpp:489-525, :549-572): Richardson-Lucy with
        total-variation regularisation.  particle_output: float32[total] (host or device) or None;
        returns the output."""
        parts = _host_parts(particle_input)
        code = _dtype_of(parts, dtype)
        if particle_output is None:
            if resume:
                raise ValueError("resume needs the previous estimate")
            particle_output = np.zeros(max(apr.total, 1), np.float32)
        st = np.ascontiguousarray(psf, np.float32)
        z, x, y = st.shape
        _check(lib().aprh_richardson_lucy_tv(apr.ctx.h, apr.h, _ptr(parts), code, _ptr(particle_output), _ptr(st), y, x, z,
                                             int(number_iterations), float(reg_factor), int(use_stencil_downsample),
                                             int(normalize), int(resume)))
        return particle_output[:apr.total] if isinstance(particle_output, np.ndarray) else particle_output


class ReconPatch:
    """ReconPatch (APRReconstruction.hpp:20-66): a box in pixels of the image at level
    l_max + level_delta; an end below 0 means the full extent."""

    _names = ("x_begin", "x_end", "y_begin", "y_end", "z_begin", "z_end", "level_delta")

    def __init__(self, x_begin=0, x_end=-1, y_begin=0, y_end=-1, z_begin=0, z_end=-1, level_delta=0):
        self.x_begin, self.x_end, self.y_begin, self.y_end = x_begin, x_end, y_begin, y_end
        self.z_begin, self.z_end, self.level_delta = z_begin, z_end, level_delta

    def size(self):
        return (self.z_end - self.z_begin) * (self.x_end - self.x_begin) * (self.y_end - self.y_begin)

    def _c(self):
        return AprhPatch(*[int(getattr(self, n)) for n in self._names])

    def check_limits(self, apr):
        """clamps the bounds to the APR's image at this level; False when a range is empty
        (the reference prints a message and returns false, :38-64)"""
        c = self._c()
        if lib().aprh_patch_check_limits(apr.h, C.byref(c)) != APRH_OK:
            return False
        for n in self._names:
            setattr(self, n, getattr(c, n))
        return True

    def __repr__(self):
        return "ReconPatch(" + ", ".join(f"{n}={getattr(self, n)}" for n in self._names) + ")"


def _clamped_patch(apr, patch):
    p = ReconPatch(*[getattr(patch, n) for n in ReconPatch._names])
    if not p.check_limits(apr):
        raise ValueError(f"invalid patch {patch!r}: ReconPatch expects begin < end in all dimensions")
    return p


class APRReconstruction:
    """APRReconstruction::reconstruct_constant (APRReconstruction.hpp:80-120, :297-394) and
    reconstruct_level (:131-171), each also over a ReconPatch (:186-196, :296-490), and
    reconstruct_smooth (:270-295)."""

    @staticmethod
    def reconstruct_constant(apr, parts, level_delta=0, out_dtype=None, out=None, dtype=None, patch=None):
        """with `patch` (a ReconPatch; its level_delta replaces the argument) the result is
        the clamped patch, [z, x, y]"""
        parts = _host_parts(parts)
        code = _dtype_of(parts, dtype)
        odt = np.dtype(out_dtype) if out_dtype is not None else (np.dtype(np.uint16) if code == APRH_DTYPE_U16
                                                                  else np.dtype(np.float32))
        ocode = _dtype_of(None, str(odt))
        if patch is not None:
            p = _clamped_patch(apr, patch)
            shape = (p.z_end - p.z_begin, p.x_end - p.x_begin, p.y_end - p.y_begin)
            if out is None:
                out = np.zeros(shape, odt)
            elif out.size != int(np.prod(shape)) or out.dtype != odt or not out.flags.c_contiguous:
                raise ValueError(f"out must be a contiguous {odt} array of {shape}")
            c = p._c()
            _check(lib().aprh_reconstruct_constant_patch(apr.ctx.h, apr.h, _ptr(parts), code, C.byref(c), _ptr(out), ocode))
            return out
        if out is None:
            y, x, z = apr.org_dims
            s = 2 ** (-level_delta)
            out = np.zeros((-(-z // s), -(-x // s), -(-y // s)), odt)
        _check(lib().aprh_reconstruct_constant(apr.ctx.h, apr.h, _ptr(parts), code, int(level_delta), _ptr(out), ocode))
        return out


    @staticmethod
    def reconstruct_level(apr, out_dtype=np.uint16, out=None, patch=None):
        """APRReconstruction::reconstruct_level (APRReconstruction.hpp:131-171): the level of the
        particle covering each pixel, [z, x, y] of the original size; with `patch`, the clamped
        patch (:397-490)"""
        odt = np.dtype(out_dtype)
        ocode = _dtype_of(None, str(odt))
        if patch is not None:
            p = _clamped_patch(apr, patch)
            shape = (p.z_end - p.z_begin, p.x_end - p.x_begin, p.y_end - p.y_begin)
            if out is None:
                out = np.zeros(shape, odt)
            elif out.size != int(np.prod(shape)) or out.dtype != odt or not out.flags.c_contiguous:
                raise ValueError(f"out must be a contiguous {odt} array of {shape}")
            c = p._c()
            _check(lib().aprh_reconstruct_level_patch(apr.ctx.h, apr.h, C.byref(c), _ptr(out), ocode))
            return out
        if out is None:
            y, x, z = apr.org_dims
            out = np.zeros((z, x, y), odt)
        _check(lib().aprh_reconstruct_level(apr.ctx.h, apr.h, _ptr(out), ocode))
        return out

    @staticmethod
    def reconstruct_smooth(apr, parts, scale_d=(2.0, 2.0, 2.0), out_dtype=None, out=None, dtype=None):
        """APRReconstruction::reconstruct_smooth (APRReconstruction.hpp:270-295): the constant
        reconstruction smoothed per axis by running means whose half-width follows the particle
        level (calc_sat_adaptive_y/x/z, :528-876); scale_d in (y, x, z) order; [z, x, y]"""
        parts = _host_parts(parts)
        code = _dtype_of(parts, dtype)
        odt = np.dtype(out_dtype) if out_dtype is not None else (np.dtype(np.uint16) if code == APRH_DTYPE_U16
                                                                  else np.dtype(np.float32))
        ocode = _dtype_of(None, str(odt))
        if out is None:
            y, x, z = apr.org_dims
            out = np.zeros((z, x, y), odt)
        sc = (C.c_float * 3)(*[float(v) for v in scale_d])
        _check(lib().aprh_reconstruct_smooth(apr.ctx.h, apr.h, _ptr(parts), code, sc, _ptr(out), ocode))
        return out


def synth_deconv(shape_zxy, seed, n_blobs, gt_out, noisy_out, sigma=(1.0, 4.0), amp=(200.0, 2000.0), background=100.0,
                 psf_sigma=(1.0, 1.0, 2.0), psf_taps=9, ctx=None):
    """Deconvolution workload on the device: gt_out (f32) = background + blobs, noisy_out (u16) =
    shot noise of gt blurred by a separable Gaussian PSF (psf_sigma in (y, x, z) order).
    gt_out / noisy_out are device addresses (or torch tensors)."""
    ctx = ctx or default_context()
    z, x, y = shape_zxy
    ps = np.array(psf_sigma, np.float32)
    _check(lib().aprh_synth_deconv_device(ctx.h, _ptr(gt_out), _ptr(noisy_out), y, x, z, seed, n_blobs, sigma[0], sigma[1],
                                          amp[0], amp[1], background, _ptr(ps), psf_taps))


def synth_deconv_host(shape_zxy, seed, n_blobs, sigma=(1.0, 4.0), amp=(200.0, 2000.0), background=100.0,
                      psf_sigma=(1.0, 1.0, 2.0), psf_taps=9):
    """the deconvolution workload of synth_deconv computed on the host (same bytes):
    returns (gt f32 (z, x, y), noisy u16 (z, x, y))"""
    z, x, y = shape_zxy
    gt = np.zeros((z, x, y), np.float32)
    noisy = np.zeros((z, x, y), np.uint16)
    ps = np.array(psf_sigma, np.float32)
    _check(lib().aprh_synth_deconv_host(_ptr(gt), _ptr(noisy), y, x, z, seed, n_blobs, sigma[0], sigma[1], amp[0], amp[1],
                                        background, _ptr(ps), psf_taps))
    return gt, noisy


# ---------------- APR files (libaprh_io.so) ----------------

IO_LIB_PATH = os.path.join(_HERE, "_lib", "libaprh_io.so")
_io = None


def io_lib():
    """The loaded libaprh_io.so (raises if it has not been built)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IO_LIB_PATH):
        raise RuntimeError(f"libaprh_io.so not built ({IO_LIB_PATH}); run `python -m libapr_amd.build` "
                           "(needs the system HDF5 and c-blosc)")
    lib()
    L = C.CDLL(IO_LIB_PATH)
    L.aprh_io_last_error.restype = C.c_char_p
    L.aprh_file_open.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp)]
    L.aprh_file_close.argtypes = [vp]
    L.aprh_file_write_apr.argtypes = [vp, vp, vp, C.c_uint64, C.c_char_p, C.c_int, C.POINTER(AprhParams)]
    L.aprh_file_write_structure.argtypes = [vp, C.c_uint64, C.c_char_p, vp, vp, C.c_uint64, vp, C.c_uint64,
                                            C.POINTER(AprhParams)]
    L.aprh_file_read_apr.argtypes = [vp, vp, C.c_uint64, C.c_char_p, C.POINTER(vp)]
    L.aprh_file_read_info.argtypes = [vp, C.c_uint64, C.c_char_p, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.aprh_file_read_structure.argtypes = [vp, C.c_uint64, C.c_char_p, vp, vp]
    L.aprh_file_write_particles.argtypes = [vp, vp, C.c_char_p, vp, C.c_int, C.c_uint64, C.c_int, C.c_uint64, C.c_char_p]
    L.aprh_file_read_particles.argtypes = [vp, vp, C.c_char_p, C.c_int, C.c_uint64, C.c_char_p, vp, C.c_uint64,
                                           C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
    _io = L
    return L


def _io_check(code):
    if code != APRH_OK:
        msg = io_lib().aprh_io_last_error().decode(errors="replace")
        raise AprhOverflowError(code, msg) if code == APRH_ERR_OVERFLOW else AprhError(code, msg)


class APRFile:
    """APRFile (APRFile.hpp:19-141): HDF5 .apr files, interchangeable with the reference's.

        f = APRFile(); f.open("x.apr", "WRITE"); f.write_apr(apr); f.write_particles("particles", parts); f.close()
        f.open("x.apr", "READ"); apr = f.read_apr(); parts = f.read_particles(apr, "particles")
    """

    def __init__(self):
        self.h = None

    def open(self, file_name, read_write="WRITE"):
        h = vp()
        _io_check(io_lib().aprh_file_open(os.fsencode(file_name), read_write.encode(), C.byref(h)))
        self.h = h
        return True

    def close(self):
        if self.h:
            io_lib().aprh_file_close(self.h)
            self.h = None

    def write_apr(self, apr, t=0, channel_name="t", write_tree=True, parameters=None):
        p = C.byref(parameters.c()) if parameters is not None else None
        _io_check(io_lib().aprh_file_write_apr(apr.ctx.h, self.h, apr.h, t, channel_name.encode(), int(write_tree), p))
        return True

    def write_structure(self, dims, xz_end_vec, y_vec, t=0, channel_name="t", parameters=None):
        """the linear access of an image of dims (y, x, z) from host arrays (no device)"""
        d = np.ascontiguousarray(dims, np.int64)
        xze = np.ascontiguousarray(xz_end_vec, np.uint64)
        yv = np.ascontiguousarray(y_vec, np.uint16)
        p = C.byref(parameters.c()) if parameters is not None else None
        _io_check(io_lib().aprh_file_write_structure(self.h, t, channel_name.encode(), _ptr(d), _ptr(xze), xze.size,
                                                      _ptr(yv) if yv.size else None, yv.size, p))

    def read_apr(self, t=0, channel_name="t", ctx=None):
        ctx = ctx or default_context()
        h = vp()
        _io_check(io_lib().aprh_file_read_apr(ctx.h, self.h, t, channel_name.encode(), C.byref(h)))
        return APR(ctx, h)

    def read_info(self, t=0, channel_name="t"):
        """(dims (y, x, z), total particles, xz_end_vec size)"""
        d = np.zeros(3, np.int64)
        tot, nx = C.c_uint64(0), C.c_uint64(0)
        _io_check(io_lib().aprh_file_read_info(self.h, t, channel_name.encode(), _ptr(d), C.byref(tot), C.byref(nx)))
        return tuple(int(v) for v in d), tot.value, nx.value

    def read_structure(self, t=0, channel_name="t"):
        """(dims, xz_end_vec, y_vec) as host arrays (no device)"""
        dims, tot, nx = self.read_info(t, channel_name)
        xze = np.zeros(max(nx, 1), np.uint64)
        yv = np.zeros(max(tot, 1), np.uint16)
        _io_check(io_lib().aprh_file_read_structure(self.h, t, channel_name.encode(), _ptr(xze), _ptr(yv)))
        return dims, xze[:nx], yv[:tot]

    def write_particles(self, particles_name, particles, apr_or_tree=True, t=0, channel_name="t", ctx=None, dtype=None):
        """particles: numpy array (host) or a device address with dtype and ctx"""
        if isinstance(particles, np.ndarray):
            code = _dtype_of(particles, dtype)
            n, cx = particles.size, None
            particles = np.ascontiguousarray(particles)
        else:
            code = _dtype_of(particles, dtype)
            n = getattr(particles, "size", None)
            cx = (ctx or default_context()).h
        if n is None:
            raise ValueError("the number of particles of a device buffer: pass a numpy array or an object with .size")
        _io_check(io_lib().aprh_file_write_particles(cx, self.h, particles_name.encode(), _ptr(particles), code, n,
                                                      int(apr_or_tree), t, channel_name.encode()))
        return True

    def read_particles(self, particles_name, apr_or_tree=True, t=0, channel_name="t"):
        """host numpy array of the dataset (u16 or f32)"""
        dt, n = C.c_int(0), C.c_uint64(0)
        L = io_lib()
        _io_check(L.aprh_file_read_particles(None, self.h, particles_name.encode(), int(apr_or_tree), t, channel_name.encode(),
                                             None, 0, C.byref(dt), C.byref(n)))
        out = np.zeros(max(n.value, 1), np.uint16 if dt.value == APRH_DTYPE_U16 else np.float32)
        _io_check(L.aprh_file_read_particles(None, self.h, particles_name.encode(), int(apr_or_tree), t, channel_name.encode(),
                                             _ptr(out), n.value, None, None))
        return out[:n.value]

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------- learned Richardson-Lucy warm start ----------------

CNN_PARAMS = 14737
CNN_TILE = 64
CNN_HALO = 4


def image_mean(data, n=None, dtype=None, ctx=None):
    """mean of a u16 / f32 array (host or device), accumulated in double on the device"""
    ctx = ctx or default_context()
    code = _dtype_of(data, dtype)
    n = data.size if n is None else n
    m = C.c_double()
    _check(lib().aprh_mean(ctx.h, _ptr(data), code, n, C.byref(m)))
    return m.value


def extract_tiles(image, dims, origins, size, halo, scale, out=None, dtype=None, ctx=None):
    """tiles of side size + 2*halo at origins ((y, x, z) per tile) minus halo, reflect-padded,
    divided by scale; returns float32 [T, s, s, s]"""
    ctx = ctx or default_context()
    code = _dtype_of(image, dtype)
    y, x, z = dims
    org = np.ascontiguousarray(np.asarray(origins, np.int32).reshape(-1, 3))
    T = org.shape[0]
    s = size + 2 * halo
    if out is None:
        out = np.zeros((T, s, s, s), np.float32)
    _check(lib().aprh_cnn_extract_tiles(ctx.h, _ptr(image), code, y, x, z, _ptr(org), T, size, halo, scale, _ptr(out)))
    return out


class WarmStartCNN:
    """The learned warm start for APRNumerics.richardson_lucy: four valid 3x3x3 convolutions
    (1 -> 16 -> 16 -> 16 -> 1, ReLU) with a residual output, bf16 MFMA on the device."""

    def __init__(self, ctx=None, seed=0, params=None):
        self.ctx = ctx or default_context()
        h = vp()
        _check(lib().aprh_cnn_create(self.ctx.h, seed, C.byref(h)))
        self.h = h
        self.post_filter = None
        _register(self.ctx, self)
        if params is not None:
            self.set_params(params)

    def get_params(self):
        p = np.zeros(CNN_PARAMS, np.float32)
        _check(lib().aprh_cnn_get_params(self.ctx.h, self.h, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != CNN_PARAMS:
            raise ValueError(f"expected {CNN_PARAMS} parameters")
        _check(lib().aprh_cnn_set_params(self.ctx.h, self.h, _ptr(p)))

    def set_post_filter(self, sigma=None, size=None):
        """Gaussian (sigma, size in (y, x, z)) applied to the warm-start particles; None: off"""
        if sigma is None:
            _check(lib().aprh_cnn_set_post_filter(self.h, None, None))
            self.post_filter = None
            return
        sg = np.ascontiguousarray(sigma, np.float32)
        sz = np.ascontiguousarray(size, np.int32)
        _check(lib().aprh_cnn_set_post_filter(self.h, _ptr(sg), _ptr(sz)))
        self.post_filter = {"sigma": [float(v) for v in sg], "size": [int(v) for v in sz]}

    def save(self, path):
        np.save(path, self.get_params(), allow_pickle=False)

    @classmethod
    def load(cls, path, ctx=None, post_filter="recorded"):
        """weights from `path`; post_filter "recorded" takes the filter recorded beside the
        weights (the .json of the same name, if it names one), None switches it off, or a dict
        {"sigma": (y, x, z), "size": (y, x, z)}"""
        net = cls(ctx, params=np.load(path, allow_pickle=False))
        if post_filter == "recorded":
            meta = os.path.splitext(path)[0] + ".json"
            post_filter = None
            if os.path.exists(meta):
                import json
                with open(meta) as f:
                    post_filter = json.load(f).get("post_filter")
        if post_filter:
            net.set_post_filter(post_filter["sigma"], post_filter["size"])
        return net

    def forward_tiles(self, x, S):
        x = np.ascontiguousarray(x, np.float32)
        T = x.size // (S + 8) ** 3
        y = np.zeros((T, S, S, S), np.float32)
        _check(lib().aprh_cnn_forward_tiles(self.ctx.h, self.h, _ptr(x), T, S, _ptr(y)))
        return y

    def train_step(self, x, target, S, lr, grads=False, T=None):
        """x: [T, S+8, S+8, S+8], target: [T, S, S, S] (host arrays or device addresses with T)"""
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, np.float32)
            target = np.ascontiguousarray(target, np.float32)
            T = x.size // (S + 8) ** 3
        loss = C.c_float()
        g = np.zeros(CNN_PARAMS, np.float32) if grads else None
        _check(lib().aprh_cnn_train_step(self.ctx.h, self.h, _ptr(x), _ptr(target), T, S, lr, C.byref(loss),
                                         _ptr(g) if grads else None))
        return (loss.value, g) if grads else loss.value

    def apply(self, image, dims=None, scale=0.0, out=None, dtype=None):
        """the network over a whole image [z, x, y] (or a device address with dims=(y, x, z))"""
        code = _dtype_of(image, dtype)
        if dims is None:
            z, x, y = image.shape
        else:
            y, x, z = dims
        if out is None:
            out = np.zeros((z, x, y), np.float32)
        _check(lib().aprh_cnn_apply(self.ctx.h, self.h, _ptr(image), code, y, x, z, scale, _ptr(out)))
        return out

    def warm_start(self, apr, parts, out=None, dtype=None):
        """initial Richardson-Lucy estimate for `apr`'s particles (use with resume=True)"""
        parts = _host_parts(parts)
        code = _dtype_of(parts, dtype)
        if out is None:
            out = np.zeros(max(apr.total, 1), np.float32)
        _check(lib().aprh_cnn_warm_start(self.ctx.h, self.h, apr.h, _ptr(parts), code, _ptr(out)))
        return out[:apr.total] if isinstance(out, np.ndarray) else out

    def close(self):
        if getattr(self, "h", None):
            lib().aprh_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """device memory from the context's allocator (aprh_malloc), freed with the object"""

    def __init__(self, nbytes, ctx=None):
        self.ctx = ctx or default_context()
        p = vp()
        _check(lib().aprh_malloc(self.ctx.h, max(int(nbytes), 1), C.byref(p)))
        self.ptr = p.value
        self.nbytes = int(nbytes)
        _register(self.ctx, self)

    def __int__(self):
        return self.ptr

    def data_ptr(self):
        return self.ptr

    def to_host(self, dtype, count=None):
        dt = np.dtype(dtype)
        count = self.nbytes // dt.itemsize if count is None else count
        h = np.empty(count, dt)
        _check(lib().aprh_memcpy(self.ctx.h, C.c_void_p(h.ctypes.data), C.c_void_p(self.ptr), h.nbytes))
        return h

    def from_host(self, arr):
        arr = np.ascontiguousarray(arr)
        _check(lib().aprh_memcpy(self.ctx.h, C.c_void_p(self.ptr), C.c_void_p(arr.ctypes.data), arr.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            lib().aprh_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

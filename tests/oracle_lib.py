"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference
build (oracle/_ref/libref_ieee.so, libref_rel.so) -- both export the same C
symbols (oracle/oracle_capi.cpp, oracle/ref_driver.cpp).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
i64 = C.c_int64


class AprhParams(C.Structure):
    """mirror of aprh_params (include/aprh_params.h)"""
    _fields_ = [(n, C.c_float) for n in ("dx", "dy", "dz", "psfx", "psfy", "psfz", "Ip_th", "lambda_", "rel_error",
                                         "sigma_th", "sigma_th_max", "grad_th")] + \
               [(n, C.c_int) for n in ("reflect_bc_lis", "neighborhood_optimization", "constant_intensity_scale")]


def params(**kw):
    p = AprhParams(1, 1, 1, 2, 2, 2, 0, 3.0, 0.1, 0, 0, 0, 1, 1, 0)
    for k, v in kw.items():
        setattr(p, "lambda_" if k == "lambda" else k, v)
    return p


def build(ref=True):
    import subprocess
    target = "all" if ref else "liboracle.so"
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, target], check=True)


def _bind(lib):
    lib.orc_bspline_u16.argtypes = [u16p, i64, i64, i64, C.c_int, C.c_float, C.c_float]
    lib.orc_bspline_u16.restype = C.c_uint64
    lib.orc_bspline_f32.argtypes = [f32p, i64, i64, i64, C.c_int, C.c_float, C.c_float]
    lib.orc_bspline_f32.restype = C.c_uint64
    lib.orc_bspline_coeffs.argtypes = [C.c_float, C.c_float, i64, f32p, C.c_int]
    lib.orc_inv_bspline_f32.argtypes = [f32p, i64, i64, i64, C.c_int]
    lib.orc_grad_mag_ds_u16.argtypes = [u16p, u16p, i64, i64, i64, C.c_float, C.c_float, C.c_float]
    lib.orc_grad_mag_ds_f32.argtypes = [f32p, f32p, i64, i64, i64, C.c_float, C.c_float, C.c_float]
    lib.orc_downsample_mean_u16_f32.argtypes = [u16p, f32p, i64, i64, i64]
    lib.orc_downsample_f32.argtypes = [f32p, f32p, i64, i64, i64, C.c_int]
    lib.orc_sat_mean_f32.argtypes = [f32p, i64, i64, i64, C.c_int, i64]
    lib.orc_pad_f32.argtypes = [f32p, f32p, i64, i64, i64, C.c_int, C.c_int, C.c_int]
    lib.orc_local_intensity_scale.argtypes = [f32p, C.c_void_p, i64, i64, i64, C.POINTER(AprhParams)]
    lib.orc_get_apr_u16.argtypes = [u16p, i64, i64, i64, C.POINTER(AprhParams), C.c_int, C.c_char_p, C.c_int]
    lib.orc_get_apr_u16.restype = C.c_void_p
    lib.orc_apr_from_grad_lis.argtypes = [u16p, f32p, f32p, i64, i64, i64, C.POINTER(AprhParams), C.c_float]
    lib.orc_apr_from_grad_lis.restype = C.c_void_p
    lib.orc_apr_info.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]
    lib.orc_apr_get.argtypes = [C.c_void_p, u64p, u64p, u16p]
    lib.orc_apr_intermediate.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    lib.orc_apr_intermediate.restype = C.c_int
    lib.orc_apr_pct.argtypes = [C.c_void_p, C.c_int, u8p]
    lib.orc_apr_pct.restype = C.c_int
    lib.orc_sample_image_u16.argtypes = [C.c_void_p, u16p, u16p]
    if hasattr(lib, "orc_get_apr_by_block_u16"):
        lib.orc_get_apr_by_block_u16.argtypes = [u16p, i64, i64, i64, C.POINTER(AprhParams), i64, i64, C.c_char_p, C.c_int]
        lib.orc_get_apr_by_block_u16.restype = C.c_void_p
        lib.orc_sample_image_blocked_u16.argtypes = [C.c_void_p, u16p, i64, i64, u16p]
    lib.orc_apr_free.argtypes = [C.c_void_p]
    lib.orc_num_threads.restype = C.c_int
    lib.orc_apr_from_structure.argtypes = [i64, i64, i64, u64p, C.c_uint64, u16p, C.c_uint64]
    lib.orc_apr_from_structure.restype = C.c_void_p
    lib.orc_tree_info.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]
    lib.orc_tree_get.argtypes = [C.c_void_p, u64p, u64p, u16p]
    lib.orc_fill_tree_mean_f32.argtypes = [C.c_void_p, f32p, f32p]
    lib.orc_fill_tree_mean_u16.argtypes = [C.c_void_p, u16p, f32p]
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    lib.orc_downsampled_stencils.argtypes = [f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, i64, C.c_void_p]
    lib.orc_downsampled_stencils.restype = i64
    for n in ("orc_convolve_pencil_f32", "orc_convolve_f32"):
        getattr(lib, n).argtypes = [C.c_void_p, f32p, i32p, C.c_int, f32p, f32p, C.c_int]
    lib.orc_convolve_pencil_u16.argtypes = [C.c_void_p, f32p, i32p, C.c_int, u16p, f32p, C.c_int]
    lib.orc_local_std_u16.argtypes = [C.c_void_p, u16p, i32p, f32p]
    lib.orc_local_std_f32.argtypes = [C.c_void_p, f32p, i32p, f32p]
    for n, t in (("orc_richardson_lucy_u16", u16p), ("orc_richardson_lucy_f32", f32p)):
        getattr(lib, n).argtypes = [C.c_void_p, t, f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_double, C.POINTER(C.c_double)]
        getattr(lib, n).restype = C.c_int
    for n, t in (("orc_richardson_lucy_tv_u16", u16p), ("orc_richardson_lucy_tv_f32", f32p)):
        if hasattr(lib, n):
            getattr(lib, n).argtypes = [C.c_void_p, t, f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                        C.c_int]
    lib.orc_gaussian_stencil.argtypes = [f32p, i32p, C.c_int, f32p]
    if hasattr(lib, "orc_sample_image_f32"):
        lib.orc_sample_image_f32.argtypes = [C.c_void_p, f32p, f32p]
    if hasattr(lib, "orc_cnn_forward"):
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        lib.orc_cnn_forward.argtypes = [f32p, f32p, C.c_int, C.c_int, C.c_int, f32p]
        lib.orc_cnn_grad.argtypes = [f32p, f32p, f32p, C.c_int, C.c_int, C.c_int, f64p]
        lib.orc_cnn_grad.restype = C.c_double
    if hasattr(lib, "orc_gradient_u16"):
        lib.orc_gradient_u16.argtypes = [C.c_void_p, u16p, C.c_int, C.c_int, C.c_float, f32p]
        lib.orc_gradient_f32.argtypes = [C.c_void_p, f32p, C.c_int, C.c_int, C.c_float, f32p]
        lib.orc_gradient_magnitude_u16.argtypes = [C.c_void_p, u16p, C.c_int, f32p, f32p]
        lib.orc_gradient_magnitude_f32.argtypes = [C.c_void_p, f32p, C.c_int, f32p, f32p]
    if hasattr(lib, "orc_face_neighbour_filter_f32"):
        lib.orc_face_neighbour_filter_f32.argtypes = [C.c_void_p, f32p, f32p, C.c_int, C.c_int, f32p]
    if hasattr(lib, "orc_sat_adaptive_f32"):
        lib.orc_sat_adaptive_f32.argtypes = [f32p, u16p, C.c_int64, C.c_int64, C.c_int64, C.c_float, C.c_int, C.c_int, C.c_int]
    if hasattr(lib, "orc_reconstruct_smooth"):
        lib.orc_reconstruct_smooth.argtypes = [C.c_void_p, C.c_void_p, C.c_int, f32p, C.c_void_p, C.c_int]
    if hasattr(lib, "orc_reconstruct"):
        lib.orc_reconstruct.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int]
    return lib


_LIBS = {}


def load(which="oracle"):
    """which: 'oracle' (restatement), 'ref' (reference, IEEE build), 'ref_rel' (reference, Release flags)"""
    if which not in _LIBS:
        path = {"oracle": os.path.join(ORACLE_DIR, "liboracle.so"),
                "ref": os.path.join(ORACLE_DIR, "_ref", "libref_ieee.so"),
                "ref_rel": os.path.join(ORACLE_DIR, "_ref", "libref_rel.so")}[which]
        if not os.path.exists(path):
            return None
        _LIBS[which] = _bind(C.CDLL(path))
    return _LIBS[which]


class OracleAPR:
    """Result of get_apr on one of the CPU libraries."""

    def __init__(self, lib, handle, dims, is_float=False):
        self.lib, self.h, self.dims, self.is_float = lib, handle, dims, is_float
        info = np.zeros(6, np.int64)
        lib.orc_apr_info(handle, info)
        self.l_min, self.l_max, self.total = int(info[0]), int(info[1]), int(info[2])
        self.n_xz, self.n_lxz = int(info[3]), int(info[4])
        self.bspline_offset = float(np.frombuffer(info[5:6].tobytes()[:4], np.float32)[0])
        self.level_xz_vec = np.zeros(self.n_lxz, np.uint64)
        self.xz_end_vec = np.zeros(self.n_xz, np.uint64)
        self.y_vec = np.zeros(max(self.total, 1), np.uint16)
        lib.orc_apr_get(handle, self.level_xz_vec, self.xz_end_vec, self.y_vec)
        self.y_vec = self.y_vec[:self.total]

    def intermediate(self, which):
        y, x, z = self.dims
        yd, xd, zd = (y + 1) // 2, (x + 1) // 2, (z + 1) // 2
        shapes = {0: ((z, x, y), np.uint16), 1: ((zd, xd, yd), np.uint16), 2: ((zd, xd, yd), np.float32),
                  3: ((zd, xd, yd), np.uint16), 4: ((zd, xd, yd), np.float32), 5: ((zd, xd, yd), np.float32)}
        shp, dt = shapes[which]
        if self.is_float and which in (0, 1, 3):
            dt = np.float32
        out = np.zeros(shp, dt)
        if self.lib.orc_apr_intermediate(self.h, which, out.ctypes.data) != 0:
            return None
        return out

    def pct(self, level, shape):
        out = np.zeros(shape, np.uint8)
        if self.lib.orc_apr_pct(self.h, level, out.reshape(-1)) != 0:
            return None
        return out

    def sample_blocked(self, img, z_block_size, ghost_z):
        """ParticleData::sample_parts_from_img_blocked (ParticleData.hpp:292-362)"""
        parts = np.zeros(max(self.total, 1), np.uint16)
        self.lib.orc_sample_image_blocked_u16(self.h, np.ascontiguousarray(img).reshape(-1), z_block_size, ghost_z, parts)
        return parts[:self.total]

    def sample(self, img):
        parts = np.zeros(max(self.total, 1), np.uint16)
        self.lib.orc_sample_image_u16(self.h, np.ascontiguousarray(img).reshape(-1), parts)
        return parts[:self.total]

    # ---- particle operations ----
    def tree(self):
        info = np.zeros(5, np.int64)
        self.lib.orc_tree_info(self.h, info)
        tl_min, tl_max, tot, nxz, nlxz = (int(v) for v in info)
        lxz = np.zeros(nlxz, np.uint64)
        xze = np.zeros(nxz, np.uint64)
        yv = np.zeros(max(tot, 1), np.uint16)
        self.lib.orc_tree_get(self.h, lxz, xze, yv)
        return dict(l_min=tl_min, l_max=tl_max, total=tot, level_xz_vec=lxz, xz_end_vec=xze, y_vec=yv[:tot])

    def fill_tree_mean(self, parts):
        t = self.tree()
        out = np.zeros(max(t["total"], 1), np.float32)
        if parts.dtype == np.uint16:
            self.lib.orc_fill_tree_mean_u16(self.h, np.ascontiguousarray(parts), out)
        else:
            self.lib.orc_fill_tree_mean_f32(self.h, np.ascontiguousarray(parts, np.float32), out)
        return out[:t["total"]]

    def convolve_pencil(self, stencils, parts, reflect=True, which="pencil"):
        flat = np.ascontiguousarray(np.concatenate([s.reshape(-1) for s in stencils]).astype(np.float32))
        dims = np.ascontiguousarray(np.array([[s.shape[2], s.shape[1], s.shape[0]] for s in stencils], np.int32).reshape(-1))
        out = np.zeros(max(self.total, 1), np.float32)
        if parts.dtype == np.uint16:
            self.lib.orc_convolve_pencil_u16(self.h, flat, dims, len(stencils), np.ascontiguousarray(parts), out, int(reflect))
        else:
            f = self.lib.orc_convolve_pencil_f32 if which == "pencil" else self.lib.orc_convolve_f32
            f(self.h, flat, dims, len(stencils), np.ascontiguousarray(parts, np.float32), out, int(reflect))
        return out[:self.total]

    def local_std(self, parts, size=(3, 3, 3)):
        """APRNumerics::local_std (APRNumerics.hpp:357-389); size in (y, x, z)"""
        out = np.zeros(max(self.total, 1), np.float32)
        sz = np.ascontiguousarray(size, np.int32)
        if parts.dtype == np.uint16:
            self.lib.orc_local_std_u16(self.h, np.ascontiguousarray(parts), sz, out)
        else:
            self.lib.orc_local_std_f32(self.h, np.ascontiguousarray(parts, np.float32), sz, out)
        return out[:self.total]

    def richardson_lucy(self, parts, psf, niter, use_ds=True, normalize=True, u=None, tol=0.0):
        resume = u is not None
        u = np.ones(max(self.total, 1), np.float32) if u is None else np.ascontiguousarray(u, np.float32).copy()
        rel = C.c_double(0)
        z, x, y = psf.shape
        f = self.lib.orc_richardson_lucy_u16 if parts.dtype == np.uint16 else self.lib.orc_richardson_lucy_f32
        p = np.ascontiguousarray(parts) if parts.dtype == np.uint16 else np.ascontiguousarray(parts, np.float32)
        it = f(self.h, p, u, np.ascontiguousarray(psf, np.float32).reshape(-1), y, x, z, niter, int(use_ds), int(normalize),
               int(resume), tol, C.byref(rel))
        return u[:self.total], it, rel.value

    def richardson_lucy_tv(self, parts, psf, niter, reg, use_ds=True, normalize=True, u=None):
        """APRNumerics::richardson_lucy_tv (APRNumerics.hpp:489-525, :549-572)"""
        resume = u is not None
        u = np.ones(max(self.total, 1), np.float32) if u is None else np.ascontiguousarray(u, np.float32).copy()
        z, x, y = psf.shape
        f = self.lib.orc_richardson_lucy_tv_u16 if parts.dtype == np.uint16 else self.lib.orc_richardson_lucy_tv_f32
        p = np.ascontiguousarray(parts) if parts.dtype == np.uint16 else np.ascontiguousarray(parts, np.float32)
        f(self.h, p, u, np.ascontiguousarray(psf, np.float32).reshape(-1), y, x, z, niter, reg, int(use_ds), int(normalize),
          int(resume))
        return u[:self.total]

    def gradient(self, parts, sobel, dim, delta):
        """APRNumerics::gradient_cfd / gradient_sobel (reference build only)"""
        out = np.zeros(max(self.total, 1), np.float32)
        if parts.dtype == np.uint16:
            self.lib.orc_gradient_u16(self.h, np.ascontiguousarray(parts), int(sobel), dim, delta, out)
        else:
            self.lib.orc_gradient_f32(self.h, np.ascontiguousarray(parts, np.float32), int(sobel), dim, delta, out)
        return out[:self.total]

    def gradient_magnitude(self, parts, sobel, deltas):
        """APRNumerics::gradient_magnitude_cfd / _sobel (reference build only)"""
        out = np.zeros(max(self.total, 1), np.float32)
        d = np.ascontiguousarray(deltas, np.float32)
        if parts.dtype == np.uint16:
            self.lib.orc_gradient_magnitude_u16(self.h, np.ascontiguousarray(parts), int(sobel), d, out)
        else:
            self.lib.orc_gradient_magnitude_f32(self.h, np.ascontiguousarray(parts, np.float32), int(sobel), d, out)
        return out[:self.total]

    def face_neighbour_filter(self, parts, filt, dim=0, repeats=-1):
        """APRNumerics::face_neighbour_filter (repeats < 0, along dim) or
        seperable_face_neighbour_filter (repeats >= 0) of float particles"""
        out = np.zeros(max(self.total, 1), np.float32)
        self.lib.orc_face_neighbour_filter_f32(self.h, np.ascontiguousarray(parts, np.float32),
                                               np.ascontiguousarray(filt, np.float32), dim, repeats, out)
        return out[:self.total]

    def sample_f32(self, img):
        parts = np.zeros(max(self.total, 1), np.float32)
        self.lib.orc_sample_image_f32(self.h, np.ascontiguousarray(img, np.float32).reshape(-1), parts)
        return parts[:self.total]

    def reconstruct_smooth(self, parts, scale=(2.0, 2.0, 2.0), out_dtype=None):
        """APRReconstruction::reconstruct_smooth (APRReconstruction.hpp:270-295): image [z, x, y]"""
        parts = np.ascontiguousarray(parts)
        pdt = 0 if parts.dtype == np.uint16 else 1
        if pdt == 1:
            parts = np.ascontiguousarray(parts, np.float32)
        odt = np.dtype(out_dtype) if out_dtype is not None else parts.dtype
        y, x, z = self.dims
        out = np.zeros((z, x, y), odt)
        self.lib.orc_reconstruct_smooth(self.h, parts.ctypes.data, pdt, np.ascontiguousarray(scale, np.float32), out.ctypes.data,
                                        0 if odt == np.uint16 else 1)
        return out

    def reconstruct(self, parts, level_delta=0, out_dtype=None):
        """piecewise-constant image [z, x, y] at level l_max + level_delta"""
        parts = np.ascontiguousarray(parts)
        pdt = 0 if parts.dtype == np.uint16 else 1
        if pdt == 1:
            parts = np.ascontiguousarray(parts, np.float32)
        odt = np.dtype(out_dtype) if out_dtype is not None else parts.dtype
        y, x, z = self.dims
        s = 2 ** (-level_delta)
        out = np.zeros((-(-z // s), -(-x // s), -(-y // s)), odt)
        self.lib.orc_reconstruct(self.h, parts.ctypes.data, pdt, level_delta, out.ctypes.data, 0 if odt == np.uint16 else 1)
        return out

    def __del__(self):
        try:
            self.lib.orc_apr_free(self.h)
        except Exception:
            pass


def get_apr(img_zxy, par, which="oracle", keep=True):
    """img_zxy: numpy uint16 array of shape (z, x, y) (y fastest)."""
    lib = load(which)
    z, x, y = img_zxy.shape
    err = C.create_string_buffer(512)
    h = lib.orc_get_apr_u16(np.ascontiguousarray(img_zxy).reshape(-1), y, x, z, C.byref(par), int(keep), err, 512)
    if not h:
        raise ValueError(err.value.decode())
    return OracleAPR(lib, h, (y, x, z))


def get_apr_by_block(img_zxy, par, z_block_size, ghost_z, which="oracle"):
    """APRConverterBatch<uint16_t> (APRConverterBatch.hpp:111-175) on a (z, x, y) uint16 image"""
    lib = load(which)
    z, x, y = img_zxy.shape
    err = C.create_string_buffer(512)
    h = lib.orc_get_apr_by_block_u16(np.ascontiguousarray(img_zxy).reshape(-1), y, x, z, C.byref(par), z_block_size, ghost_z,
                                     err, 512)
    if not h:
        raise ValueError(err.value.decode())
    return OracleAPR(lib, h, (y, x, z))


def get_apr_f32(img_zxy, par, which="oracle", keep=True):
    """APRConverter<float>::get_apr; img_zxy float32 (z, x, y)"""
    lib = load(which)
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    lib.orc_get_apr_f32.argtypes = [f32p, C.c_int64, C.c_int64, C.c_int64, C.POINTER(AprhParams), C.c_int, C.c_char_p,
                                    C.c_int]
    lib.orc_get_apr_f32.restype = C.c_void_p
    z, x, y = img_zxy.shape
    err = C.create_string_buffer(512)
    h = lib.orc_get_apr_f32(np.ascontiguousarray(img_zxy, np.float32).reshape(-1), y, x, z, C.byref(par), int(keep), err, 512)
    if not h:
        raise ValueError(err.value.decode())
    return OracleAPR(lib, h, (y, x, z), is_float=True)


def apr_from_grad_lis(grad, lis, intensity, dims, par, offset, which="oracle"):
    lib = load(which)
    y, x, z = dims
    h = lib.orc_apr_from_grad_lis(np.ascontiguousarray(grad, np.uint16).reshape(-1),
                                  np.ascontiguousarray(lis, np.float32).reshape(-1),
                                  np.ascontiguousarray(intensity, np.float32).reshape(-1), y, x, z, C.byref(par), offset)
    return OracleAPR(lib, h, dims)


def apr_from_structure(dims, xz_end_vec, y_vec, which="oracle"):
    """oracle APR handle built from a LinearAccess (e.g. the GPU's)"""
    lib = load(which)
    y, x, z = dims
    xze = np.ascontiguousarray(xz_end_vec, np.uint64)
    yv = np.ascontiguousarray(y_vec, np.uint16)
    h = lib.orc_apr_from_structure(y, x, z, xze, xze.size, yv if yv.size else np.zeros(1, np.uint16), yv.size)
    if not h:
        raise ValueError("structure does not match the dimensions")
    return OracleAPR(lib, h, dims)


def downsampled_stencils(stencil_zxy, nlevels, normalize, which="oracle"):
    lib = load(which)
    z, x, y = stencil_zxy.shape
    st = np.ascontiguousarray(stencil_zxy, np.float32).reshape(-1)
    tot = lib.orc_downsampled_stencils(st, y, x, z, nlevels, int(normalize), None, 0, None)
    out = np.zeros(tot, np.float32)
    dims = np.zeros(3 * max(nlevels, 1), np.int32)
    lib.orc_downsampled_stencils(st, y, x, z, nlevels, int(normalize), out.ctypes.data, tot, dims.ctypes.data)
    res, off = [], 0
    for i in range(max(nlevels, 1)):
        yy, xx, zz = dims[3 * i: 3 * i + 3]
        n = int(yy) * int(xx) * int(zz)
        res.append(out[off:off + n].reshape(int(zz), int(xx), int(yy)))
        off += n
    return res


def gaussian_stencil_lib(sigma, size, normalize=True, which="oracle"):
    """APRStencil::create_gaussian_filter through an oracle library; sigma/size in (y, x, z)"""
    lib = load(which)
    out = np.zeros((size[2], size[1], size[0]), np.float32)
    lib.orc_gaussian_stencil(np.array(sigma, np.float32), np.array(size, np.int32), int(normalize), out.reshape(-1))
    return out


def gaussian_stencil(sigma, size, normalize=True):
    """APRStencil::create_gaussian_filter (APRStencil.hpp:173-228): separable float
    Gaussian, expf on the host, product gauss_y*gauss_x*gauss_z, sum in float."""
    import math
    sy, sx, sz = size
    def g1(s, n):
        if n <= 1:
            return np.ones(1, np.float32)
        v = []
        for i in range(-(n // 2), n // 2 + 1):
            num = np.float32(-i * i) / (np.float32(2) * np.float32(s) * np.float32(s))
            v.append(np.float32(np.exp(np.float32(num), dtype=np.float32)) /
                     (np.float32(s) * np.float32(math.sqrt(2 * math.pi))))
        return np.array(v, np.float32)
    gy, gx, gz = g1(sigma[0], sy), g1(sigma[1], sx), g1(sigma[2], sz)
    st = np.zeros((sz, sx, sy), np.float32)
    tot = np.float32(0)
    for i in range(sy):
        for j in range(sx):
            for k in range(sz):
                v = np.float32(np.float32(gy[i] * gx[j]) * gz[k])
                st[k, j, i] = v
                tot = np.float32(tot + v)
    if normalize:
        st = (st / tot).astype(np.float32)
    return st

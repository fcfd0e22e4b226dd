"""ctypes binding of libvpbs (the C ABI declared in include/vpbs.h).

The library is built in-tree (``__graft_entry__.build()``) as
``vortexpasta.jl_amd/libvpbs.so``. There is no fallback: if the library is
missing or fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes
import math
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libvpbs.so"

ABI_VERSION = 1
F64, F32 = 0, 1
VELOCITY = 1
STREAMFUNCTION = 2
LIA_OFF = 4
LIA_ONLY = 8
NO_SHORTRANGE = 16
NO_LONGRANGE = 32
NEIGHBOUR_STATS = 64
KEEP_UHAT = 128
FFT_FUSED, FFT_ROCFFT, FFT_SEPARATE = 0, 1, 2  # vpbs_params.fft
UHAT_VELOCITY = 1
UHAT_STREAMFUNCTION = 2

EXPORTED = (
    "vpbs_create",
    "vpbs_destroy",
    "vpbs_last_error",
    "vpbs_set_points",
    "vpbs_set_points_device",
    "vpbs_set_filaments",
    "vpbs_get_points",
    "vpbs_advance",
    "vpbs_get_filaments",
    "vpbs_get_uhat",
    "vpbs_energy_spectrum",
    "vpbs_energy_spectrum_noalloc",
    "vpbs_interpolate_band",
    "vpbs_compute",
    "vpbs_get_outputs",
    "vpbs_output_device_ptrs",
    "vpbs_neighbour_stats",
    "vpbs_get_timings",
    "vpbs_get_sizes",
    "vpbs_mlp_create",
    "vpbs_mlp_destroy",
    "vpbs_mlp_last_error",
    "vpbs_mlp_num_params",
    "vpbs_mlp_set_params",
    "vpbs_mlp_get_params",
    "vpbs_mlp_forward",
    "vpbs_mlp_infer",
    "vpbs_mlp_set_output_bound",
    "vpbs_mlp_gradient",
    "vpbs_mlp_train",
    "vpbs_learned_setup",
    "vpbs_compute_learned",
    "vpbs_learned_stats",
    "vpbs_learned_features",
    "vpbs_learned_difference",
)


class VpbsParams(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("Gamma", ctypes.c_double),
        ("a", ctypes.c_double),
        ("Delta", ctypes.c_double),
        ("Ls", ctypes.c_double * 3),
        ("alpha", ctypes.c_double),
        ("rcut", ctypes.c_double),
        ("Ns", ctypes.c_int32 * 3),
        ("nquad", ctypes.c_int32),
        ("nsubdiv", ctypes.c_int32),
        ("nufft_m", ctypes.c_int32),
        ("nufft_sigma", ctypes.c_double),
        ("avoid_explicit_erf", ctypes.c_int32),
        ("truncate_spherical", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("fft", ctypes.c_int32),
    ]


class VpbsTimings(ctypes.Structure):
    _fields_ = [
        ("total", ctypes.c_double),
        ("local_term", ctypes.c_double),
        ("cell_build", ctypes.c_double),
        ("pair", ctypes.c_double),
        ("self", ctypes.c_double),
        ("lr_sort", ctypes.c_double),
        ("spread", ctypes.c_double),
        ("fft_forward", ctypes.c_double),
        ("fourier", ctypes.c_double),
        ("fft_backward", ctypes.c_double),
        ("interp", ctypes.c_double),
        ("merge", ctypes.c_double),
        ("pair_candidates", ctypes.c_int64),
        ("grid", ctypes.c_int64 * 3),
        ("cells", ctypes.c_int64 * 3),
        ("point_data", ctypes.c_double),
        ("node_sort", ctypes.c_double),
        ("features", ctypes.c_double),
        ("mlp", ctypes.c_double),
        ("check", ctypes.c_double),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["grid"] = tuple(self.grid)
        d["cells"] = tuple(self.cells)
        return d


class VpbsLearnedReport(ctypes.Structure):
    _fields_ = [
        ("err_corrected", ctypes.c_double),
        ("err_uncorrected", ctypes.c_double),
        ("nsample", ctypes.c_int64),
        ("fallback", ctypes.c_int32),
        ("total_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class VpbsError(RuntimeError):
    pass


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def load() -> ctypes.CDLL:
    """Load libvpbs.so and declare the signatures. Raises if the library is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise VpbsError(f"libvpbs not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    L = ctypes.CDLL(str(LIB_PATH))
    vp = ctypes.c_void_p
    i64, u32, d = ctypes.c_int64, ctypes.c_uint32, ctypes.c_double
    L.vpbs_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(VpbsParams)]
    L.vpbs_create.restype = ctypes.c_int
    L.vpbs_destroy.argtypes = [vp]
    L.vpbs_destroy.restype = None
    L.vpbs_last_error.argtypes = [vp]
    L.vpbs_last_error.restype = ctypes.c_char_p
    pts_sig = [vp, i64, vp, vp, vp, vp, vp, vp, i64, vp, vp]
    L.vpbs_set_points.argtypes = pts_sig
    L.vpbs_set_points.restype = ctypes.c_int
    L.vpbs_set_points_device.argtypes = pts_sig
    L.vpbs_set_points_device.restype = ctypes.c_int
    L.vpbs_set_filaments.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, d, ctypes.c_int32, vp, vp]
    L.vpbs_set_filaments.restype = ctypes.c_int
    L.vpbs_get_points.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.vpbs_get_points.restype = ctypes.c_int
    L.vpbs_advance.argtypes = [vp, d, ctypes.c_int32]
    L.vpbs_advance.restype = ctypes.c_int
    L.vpbs_get_filaments.argtypes = [vp, vp, vp, vp, vp]
    L.vpbs_get_filaments.restype = ctypes.c_int
    L.vpbs_get_uhat.argtypes = [vp, vp, vp, vp]
    L.vpbs_get_uhat.restype = ctypes.c_int
    L.vpbs_energy_spectrum.argtypes = [vp, i64, d, vp]
    L.vpbs_energy_spectrum.restype = ctypes.c_int
    L.vpbs_energy_spectrum_noalloc.argtypes = [vp, i64, d, vp]
    L.vpbs_energy_spectrum_noalloc.restype = ctypes.c_int
    L.vpbs_interpolate_band.argtypes = [vp, d, d, vp]
    L.vpbs_interpolate_band.restype = ctypes.c_int
    L.vpbs_compute.argtypes = [vp, u32, d, vp, vp]
    L.vpbs_compute.restype = ctypes.c_int
    L.vpbs_get_outputs.argtypes = [vp, vp, vp]
    L.vpbs_get_outputs.restype = ctypes.c_int
    L.vpbs_output_device_ptrs.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.vpbs_output_device_ptrs.restype = ctypes.c_int
    L.vpbs_neighbour_stats.argtypes = [vp, vp, vp]
    L.vpbs_neighbour_stats.restype = ctypes.c_int
    L.vpbs_get_timings.argtypes = [vp, ctypes.POINTER(VpbsTimings)]
    L.vpbs_get_timings.restype = ctypes.c_int
    L.vpbs_get_sizes.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.vpbs_get_sizes.restype = ctypes.c_int
    i32 = ctypes.c_int32
    L.vpbs_mlp_create.argtypes = [ctypes.POINTER(vp), i32, i32, i32, i32, i32]
    L.vpbs_mlp_create.restype = ctypes.c_int
    L.vpbs_mlp_destroy.argtypes = [vp]
    L.vpbs_mlp_destroy.restype = None
    L.vpbs_mlp_last_error.argtypes = [vp]
    L.vpbs_mlp_last_error.restype = ctypes.c_char_p
    L.vpbs_mlp_num_params.argtypes = [vp]
    L.vpbs_mlp_num_params.restype = ctypes.c_int64
    L.vpbs_mlp_set_params.argtypes = [vp, vp, vp, vp, vp]
    L.vpbs_mlp_set_params.restype = ctypes.c_int
    L.vpbs_mlp_get_params.argtypes = [vp, vp, vp, vp, vp]
    L.vpbs_mlp_get_params.restype = ctypes.c_int
    L.vpbs_mlp_forward.argtypes = [vp, i64, vp, vp]
    L.vpbs_mlp_forward.restype = ctypes.c_int
    L.vpbs_mlp_infer.argtypes = [vp, i64, vp, vp]
    L.vpbs_mlp_infer.restype = ctypes.c_int
    L.vpbs_mlp_set_output_bound.argtypes = [vp, vp]
    L.vpbs_mlp_set_output_bound.restype = ctypes.c_int
    L.vpbs_mlp_gradient.argtypes = [vp, i64, vp, vp, ctypes.POINTER(d), vp]
    L.vpbs_mlp_gradient.restype = ctypes.c_int
    L.vpbs_mlp_train.argtypes = [vp, i64, vp, vp, i32, i32, d, d, d, d, ctypes.c_uint64, vp]
    L.vpbs_mlp_train.restype = ctypes.c_int
    L.vpbs_learned_setup.argtypes = [vp, vp, d, d, d, d, d, ctypes.c_uint64]
    L.vpbs_learned_setup.restype = ctypes.c_int
    L.vpbs_compute_learned.argtypes = [vp, vp, ctypes.POINTER(VpbsLearnedReport)]
    L.vpbs_compute_learned.restype = ctypes.c_int
    L.vpbs_learned_stats.argtypes = [vp, vp, vp, vp]
    L.vpbs_learned_stats.restype = ctypes.c_int
    L.vpbs_learned_features.argtypes = [vp, vp, vp]
    L.vpbs_learned_features.restype = ctypes.c_int
    L.vpbs_learned_difference.argtypes = [vp, vp, vp]
    L.vpbs_learned_difference.restype = ctypes.c_int
    _lib = L
    return L


def make_params(params, device: int = 0) -> VpbsParams:
    """Translate a ParamsBiotSavart into the C parameter struct."""
    p = VpbsParams()
    p.abi_version = ABI_VERSION
    p.precision = F32 if params.precision == "f32" else F64
    p.Gamma = params.Γ
    p.a = params.a
    p.Delta = params.Δ
    for d in range(3):
        p.Ls[d] = params.Ls[d]
        p.Ns[d] = params.Ns[d]
    p.alpha = params.α
    p.rcut = params.rcut
    p.nquad = len(params.quadrature)
    p.nsubdiv = params.backend_short.nsubdiv
    p.nufft_m = params.backend_long.m
    p.nufft_sigma = params.backend_long.σ
    p.avoid_explicit_erf = int(params.avoid_explicit_erf)
    p.truncate_spherical = int(params.longrange_truncate_spherical)
    p.device = device
    p.fft = {"fused": FFT_FUSED, "rocfft": FFT_ROCFFT, "separate": FFT_SEPARATE}[params.backend_long.fft]
    return p


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return ctypes.c_void_p(a.ptr)
    return ctypes.c_void_p(a.ctypes.data)


def _expect(a, shape, dtype, name):
    """A host array or DeviceArray of exactly this shape and type (C-contiguous)."""
    ok = tuple(a.shape) == tuple(shape) and np.dtype(a.dtype) == np.dtype(dtype)
    if isinstance(a, np.ndarray):
        ok = ok and a.flags.c_contiguous
    elif not isinstance(a, DeviceArray):
        ok = False
    if not ok:
        raise ValueError(f"{name}: expected a C-contiguous {np.dtype(dtype).name} array of shape {tuple(shape)}")


_hip = None


def hip():
    """The HIP runtime libvpbs is linked against (already loaded with it: found by its soname), for
    device buffers of the caller (training samples, device-side outputs)."""
    global _hip
    if _hip is None:
        load()
        H = ctypes.CDLL("libamdhip64.so.7")
        sz, vp = ctypes.c_size_t, ctypes.c_void_p
        H.hipSetDevice.argtypes = [ctypes.c_int]
        H.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        H.hipMemcpy2D.argtypes = [vp, sz, vp, sz, sz, sz, ctypes.c_int]
        H.hipGetErrorString.argtypes = [ctypes.c_int]
        H.hipGetErrorString.restype = ctypes.c_char_p
        for f in (H.hipSetDevice, H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipMemcpy2D):
            f.restype = ctypes.c_int
        _hip = H
    return _hip


def _hip_check(rc, what):
    if rc != 0:
        raise VpbsError(f"{what} failed ({rc}): {hip().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A C-contiguous array in device memory (hipMalloc), freed with the object.

    Plumbing for the caller's side of the C ABI: training samples and device-resident
    outputs; every computation on it runs in libvpbs.
    """

    def __init__(self, shape, dtype=np.float32, device: int = 0):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.device = device
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        H = hip()
        p = ctypes.c_void_p()
        _hip_check(H.hipSetDevice(device), "hipSetDevice")
        _hip_check(H.hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    def data_ptr(self) -> int:
        return self.ptr

    def __len__(self):
        return self.shape[0]

    @staticmethod
    def from_numpy(a, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.shape, a.dtype, device)
        d.copy_from(a)
        return d

    def copy_from(self, a):
        """Copy a host array (or a DeviceArray) of the same size and type into this one."""
        if isinstance(a, DeviceArray):
            if a.nbytes != self.nbytes:
                raise ValueError("size mismatch")
            src = ctypes.c_void_p(a.ptr)
        else:
            a = np.ascontiguousarray(a, dtype=self.dtype)
            if a.nbytes != self.nbytes:
                raise ValueError("size mismatch")
            src = ctypes.c_void_p(a.ctypes.data)
        _hip_check(hip().hipMemcpy(ctypes.c_void_p(self.ptr), src, self.nbytes, 4), "hipMemcpy")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _hip_check(hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr), self.nbytes, 4),
                   "hipMemcpy")
        return out

    def rows(self, step: int) -> np.ndarray:
        """Every ``step``-th row (first axis) copied to the host (hipMemcpy2D)."""
        row = self.nbytes // self.shape[0]
        n = (self.shape[0] + step - 1) // step
        out = np.empty((n,) + self.shape[1:], self.dtype)
        _hip_check(hip().hipMemcpy2D(ctypes.c_void_p(out.ctypes.data), row, ctypes.c_void_p(self.ptr), row * step,
                                     row, n, 4), "hipMemcpy2D")
        return out

    @staticmethod
    def concatenate(parts, device: int = 0) -> "DeviceArray":
        """Rows of the parts (same trailing shape and type) in one new array."""
        parts = list(parts)
        tail, dt = parts[0].shape[1:], parts[0].dtype
        if any(p.shape[1:] != tail or p.dtype != dt for p in parts):
            raise ValueError("parts differ in trailing shape or type")
        out = DeviceArray((sum(p.shape[0] for p in parts),) + tail, dt, device)
        off = 0
        for p in parts:
            _hip_check(hip().hipMemcpy(ctypes.c_void_p(out.ptr + off), ctypes.c_void_p(p.ptr), p.nbytes, 4),
                       "hipMemcpy")
            off += p.nbytes
        return out

    def free(self):
        if getattr(self, "ptr", None):
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """Owns one vpbs_ctx (one GPU). Not thread-safe: calls must not overlap."""

    def __init__(self, params, device: int = 0):
        self._L = load()
        self._h = ctypes.c_void_p()
        self._p = make_params(params, device)
        rc = self._L.vpbs_create(ctypes.byref(self._h), ctypes.byref(self._p))
        if rc != 0:
            raise VpbsError(f"vpbs_create failed ({rc}): {self._L.vpbs_last_error(None).decode()}")
        self.Np = 0

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.vpbs_last_error(self._h).decode()
            raise VpbsError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if self._h:
            self._L.vpbs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, pd):
        arrs = dict(
            nodes=np.ascontiguousarray(pd.nodes, dtype=np.float64),
            idx=np.ascontiguousarray(pd.node_idx_prev, dtype=np.int32),
            sp=np.ascontiguousarray(pd.sp, dtype=np.float64),
            spp=np.ascontiguousarray(pd.spp, dtype=np.float64),
            dlm=np.ascontiguousarray(pd.dlm, dtype=np.float64),
            dlp=np.ascontiguousarray(pd.dlp, dtype=np.float64),
            pts=np.ascontiguousarray(pd.points, dtype=np.float64),
            chg=np.ascontiguousarray(pd.charges, dtype=np.float64),
        )
        Np = arrs["nodes"].shape[1]
        Nq = arrs["pts"].shape[1]
        rc = self._L.vpbs_set_points(
            self._h, Np, _ptr(arrs["nodes"]), _ptr(arrs["idx"]), _ptr(arrs["sp"]), _ptr(arrs["spp"]),
            _ptr(arrs["dlm"]), _ptr(arrs["dlp"]), Nq, _ptr(arrs["pts"]), _ptr(arrs["chg"]),
        )
        self._check(rc, "vpbs_set_points")
        self.Np = Np

    def set_filaments(self, batch, params):
        """Build the point data on the device from a FilamentBatch (vpbs_set_filaments).

        Takes the batch's nodes, knots, periods, end-to-end offsets and spline
        coefficients; the quadrature rule and ``lia_segment_fraction`` come from
        ``params`` (NoQuadrature → straight segments). With
        ``lia_segment_fraction`` set, the near-singular local integrals are then
        evaluated on the device by ``compute`` (``quadrature_near_singularity``).
        """
        from .params import NoQuadrature

        counts = np.ascontiguousarray(batch.Ns, dtype=np.int64)
        X = np.ascontiguousarray(batch.X.T, dtype=np.float64)
        ts = np.ascontiguousarray(batch.ts, dtype=np.float64)
        T = np.ascontiguousarray(batch.T, dtype=np.float64)
        xoff = np.ascontiguousarray(batch.Xoff.T, dtype=np.float64)
        c = np.ascontiguousarray(batch.c.T, dtype=np.float64)
        quad = params.quadrature
        if isinstance(quad, NoQuadrature):
            zs = ws = None
        else:
            zs, ws = (np.ascontiguousarray(a, dtype=np.float64) for a in quad.rule())
        g = params.lia_segment_fraction
        zn = wn = None
        if g is not None:
            zn, wn = (np.ascontiguousarray(a, dtype=np.float64) for a in params.quadrature_near_singularity.rule())
        rc = self._L.vpbs_set_filaments(
            self._h, len(counts), _ptr(counts), _ptr(X), _ptr(ts), _ptr(T), _ptr(xoff), _ptr(c),
            _ptr(zs), _ptr(ws), math.nan if g is None else float(g), 0 if zn is None else len(zn), _ptr(zn), _ptr(wn),
        )
        self._check(rc, "vpbs_set_filaments")
        self.Np = int(counts.sum())
        self.Nf = len(counts)

    def get_points(self):
        """Device point data as a dict of host arrays (vpbs_get_points)."""
        Np = self.Np
        nq = self._p.nquad
        out = dict(
            nodes=np.empty((3, Np)), node_idx_prev=np.empty(Np, dtype=np.int32), sp=np.empty((3, Np)),
            spp=np.empty((3, Np)), dlm=np.empty(Np), dlp=np.empty(Np), points=np.empty((3, nq * Np)),
            charges=np.empty((3, nq * Np)),
        )
        rc = self._L.vpbs_get_points(self._h, *(_ptr(a) for a in out.values()))
        self._check(rc, "vpbs_get_points")
        return out

    def compute(self, flags: int, integration_cutoff=None, fetch=True):
        """Run vpbs_compute. Returns (v, psi) as (3, Np) arrays (None where not requested / not fetched)."""
        cut = math.nan if integration_cutoff is None else float(integration_cutoff)
        v = np.empty((3, self.Np)) if (fetch and flags & VELOCITY) else None
        psi = np.empty((3, self.Np)) if (fetch and flags & STREAMFUNCTION) else None
        rc = self._L.vpbs_compute(self._h, flags, cut, _ptr(v), _ptr(psi))
        self._check(rc, "vpbs_compute")
        return v, psi

    def advance(self, dt: float, nsteps: int = 1):
        """SSPRK33 steps on the device-resident filaments (vpbs_advance)."""
        self._check(self._L.vpbs_advance(self._h, float(dt), int(nsteps)), "vpbs_advance")

    def get_filaments(self):
        """Device filament state: dict with X (Np, 3), knots (Np,), periods (Nf,), coefs (Np, 3)."""
        X = np.empty((3, self.Np))
        c = np.empty((3, self.Np))
        ts = np.empty(self.Np)
        T = np.empty(self.Nf)
        self._check(self._L.vpbs_get_filaments(self._h, _ptr(X), _ptr(ts), _ptr(T), _ptr(c)), "vpbs_get_filaments")
        return {"X": X.T.copy(), "knots": ts, "periods": T, "coefs": c.T.copy()}

    def uhat_quantity(self):
        """Quantity of the kept Fourier field ("velocity", "streamfunction" or None), without copying it."""
        q = ctypes.c_int32()
        self._check(self._L.vpbs_get_uhat(self._h, ctypes.byref(q), None, None), "vpbs_get_uhat")
        return {UHAT_VELOCITY: "velocity", UHAT_STREAMFUNCTION: "streamfunction"}.get(q.value)

    def get_uhat(self):
        """Kept Fourier field (vpbs_get_uhat): (quantity, array (3, N2, N1, N0//2+1) complex128).

        quantity is "velocity", "streamfunction" or None (nothing kept).
        """
        q = ctypes.c_int32()
        dims = (ctypes.c_int64 * 3)()
        self._check(self._L.vpbs_get_uhat(self._h, ctypes.byref(q), dims, None), "vpbs_get_uhat")
        name = {UHAT_VELOCITY: "velocity", UHAT_STREAMFUNCTION: "streamfunction"}.get(q.value)
        if name is None:
            return None, None
        nkx, n1, n2 = dims
        buf = np.empty((3, n2, n1, nkx, 2))
        self._check(self._L.vpbs_get_uhat(self._h, None, None, _ptr(buf)), "vpbs_get_uhat")
        return name, buf[..., 0] + 1j * buf[..., 1]

    def energy_spectrum(self, nk: int, dk: float) -> np.ndarray:
        Ek = np.empty(int(nk))
        self._check(self._L.vpbs_energy_spectrum(self._h, int(nk), float(dk), _ptr(Ek)), "vpbs_energy_spectrum")
        return Ek

    def energy_spectrum_noalloc(self, nk: int, dk: float, out=None):
        """vpbs_energy_spectrum_noalloc: the same spectrum without device allocation (nk ≤ 1024);
        ``out``: nk float64 values, a numpy array (default) or a DeviceArray."""
        Ek = np.empty(int(nk)) if out is None else out
        _expect(Ek, (int(nk),), np.float64, "out")
        self._check(self._L.vpbs_energy_spectrum_noalloc(self._h, int(nk), float(dk), _ptr(Ek)),
                    "vpbs_energy_spectrum_noalloc")
        return Ek

    def interpolate_band(self, kmin: float, kmax: float) -> np.ndarray:
        out = np.empty((3, self.Np))
        self._check(self._L.vpbs_interpolate_band(self._h, float(kmin), float(kmax), _ptr(out)),
                    "vpbs_interpolate_band")
        return out

    def neighbour_stats(self):
        counts = np.empty(self.Np, dtype=np.int64)
        hashes = np.empty(self.Np, dtype=np.uint64)
        self._check(self._L.vpbs_neighbour_stats(self._h, _ptr(counts), _ptr(hashes)), "vpbs_neighbour_stats")
        return counts, hashes

    def timings(self) -> dict:
        t = VpbsTimings()
        self._check(self._L.vpbs_get_timings(self._h, ctypes.byref(t)), "vpbs_get_timings")
        return t.as_dict()

    # learned correction
    def learned_setup(self, net, alpha_conv: float, rcut_conv: float, feature_radius: float,
                      sample_fraction: float = 0.01, tol: float = 1e-4, seed: int = 0):
        h = net.handle if net is not None else None
        self._check(self._L.vpbs_learned_setup(self._h, h, float(alpha_conv), float(rcut_conv), float(feature_radius),
                                               float(sample_fraction), float(tol), ctypes.c_uint64(seed)),
                    "vpbs_learned_setup")

    def compute_learned(self, fetch=True):
        """Returns (v (3, Np) or None, report dict)."""
        v = np.empty((3, self.Np)) if fetch else None
        rep = VpbsLearnedReport()
        self._check(self._L.vpbs_compute_learned(self._h, _ptr(v), ctypes.byref(rep)), "vpbs_compute_learned")
        return v, rep.as_dict()

    def learned_stats(self):
        """(calls, fallbacks, largest sampled error) of the learned evaluations since learned_setup."""
        calls, fb, err = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        self._check(self._L.vpbs_learned_stats(self._h, ctypes.byref(calls), ctypes.byref(fb), ctypes.byref(err)),
                    "vpbs_learned_stats")
        return {"calls": calls.value, "fallbacks": fb.value, "max_err_corrected": err.value}

    def learned_features(self, X=None, frame=None):
        """Features into X ((Np, 128) float32, numpy or DeviceArray) and frames (Np, 9)."""
        if X is not None:
            _expect(X, (self.Np, 128), np.float32, "X")
        if frame is not None:
            _expect(frame, (self.Np, 9), np.float32, "frame")
        px, pf = _ptr(X), _ptr(frame)
        self._check(self._L.vpbs_learned_features(self._h, px, pf), "vpbs_learned_features")
        return X, frame

    def learned_difference(self, dv=None, T=None):
        """Converged − cheap velocity: dv (3, Np) float64 and/or targets T (Np, 3) float32 (numpy or DeviceArray)."""
        if dv is not None:
            _expect(dv, (3, self.Np), np.float64, "dv")
        if T is not None:
            _expect(T, (self.Np, 3), np.float32, "T")
        pd_, pt = _ptr(dv), _ptr(T)
        self._check(self._L.vpbs_learned_difference(self._h, pd_, pt), "vpbs_learned_difference")
        return dv, T

    def sizes(self):
        g = (ctypes.c_int64 * 3)()
        c = (ctypes.c_int64 * 3)()
        self._check(self._L.vpbs_get_sizes(self._h, g, c), "vpbs_get_sizes")
        return tuple(g), tuple(c)


class MLP:
    """Owns one vpbs_mlp (network of the learned correction) on one GPU.

    Sample arrays are DeviceArrays of float32 rows (n, n_in) / (n, n_out).
    """

    def __init__(self, n_in: int, n_hidden: int, n_layers: int, n_out: int, device: int = 0):
        self._L = load()
        self._h = ctypes.c_void_p()
        rc = self._L.vpbs_mlp_create(ctypes.byref(self._h), n_in, n_hidden, n_layers, n_out, device)
        if rc != 0:
            raise VpbsError(f"vpbs_mlp_create failed ({rc}): invalid sizes or device")
        self.n_in, self.H, self.nl, self.n_out = n_in, n_hidden, n_layers, n_out
        self.nparams = int(self._L.vpbs_mlp_num_params(self._h))

    @property
    def handle(self):
        return self._h

    def _check(self, rc, what):
        if rc != 0:
            raise VpbsError(f"{what} failed ({rc}): {self._L.vpbs_mlp_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            self._L.vpbs_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, in_shift, in_scale, out_scale):
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (params, in_shift, in_scale, out_scale)]
        if a[0].size != self.nparams or a[1].size != self.n_in or a[2].size != self.n_in or a[3].size != self.n_out:
            raise ValueError("parameter or normalisation sizes do not match the network")
        self._check(self._L.vpbs_mlp_set_params(self._h, *(_ptr(x) for x in a)), "vpbs_mlp_set_params")

    def set_output_bound(self, ybound=None):
        """Clamp of the BF16 inference's outputs to ±ybound (n_out values; None: no bound)."""
        b = None if ybound is None else np.ascontiguousarray(ybound, dtype=np.float32)
        if b is not None and b.size != self.n_out:
            raise ValueError("the output bound needs one value per output")
        self._check(self._L.vpbs_mlp_set_output_bound(self._h, _ptr(b)), "vpbs_mlp_set_output_bound")

    def get_params(self):
        out = (np.empty(self.nparams, np.float32), np.empty(self.n_in, np.float32), np.empty(self.n_in, np.float32),
               np.empty(self.n_out, np.float32))
        self._check(self._L.vpbs_mlp_get_params(self._h, *(_ptr(x) for x in out)), "vpbs_mlp_get_params")
        return out

    @staticmethod
    def _dev(t, cols):
        if not (isinstance(t, DeviceArray) and t.dtype == np.float32 and len(t.shape) == 2 and t.shape[1] == cols):
            raise ValueError(f"expected a float32 DeviceArray of shape (n, {cols})")
        return ctypes.c_void_p(t.ptr)

    def _out(self, n):
        return DeviceArray((n, self.n_out), np.float32)

    def forward(self, X):
        Y = self._out(X.shape[0])
        self._check(self._L.vpbs_mlp_forward(self._h, X.shape[0], self._dev(X, self.n_in), self._dev(Y, self.n_out)),
                    "vpbs_mlp_forward")
        return Y

    def infer(self, X):
        Y = self._out(X.shape[0])
        self._check(self._L.vpbs_mlp_infer(self._h, X.shape[0], self._dev(X, self.n_in), self._dev(Y, self.n_out)),
                    "vpbs_mlp_infer")
        return Y

    def gradient(self, X, T):
        g = np.empty(self.nparams, np.float32)
        loss = ctypes.c_double()
        self._check(self._L.vpbs_mlp_gradient(self._h, X.shape[0], self._dev(X, self.n_in), self._dev(T, self.n_out),
                                              ctypes.byref(loss), _ptr(g)), "vpbs_mlp_gradient")
        return loss.value, g

    def train(self, X, T, steps: int, batch: int, lr: float, beta1=0.9, beta2=0.999, eps=1e-8, seed: int = 0,
              record_losses: bool = True):
        losses = np.empty(steps) if record_losses and steps > 0 else None
        self._check(self._L.vpbs_mlp_train(self._h, X.shape[0], self._dev(X, self.n_in), self._dev(T, self.n_out),
                                           int(steps), int(batch), float(lr), float(beta1), float(beta2), float(eps),
                                           ctypes.c_uint64(seed), _ptr(losses)), "vpbs_mlp_train")
        return losses

"""ctypes binding of libcgfdm_amd.so (the C ABI declared in include/cgfdm_amd.h).

The library is the product path: there is no CPU fallback. Importing this
module when the shared library is missing raises immediately.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .model import Model

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libcgfdm_amd.so")

SCHEME_CGFDM = 0
SCHEME_AWENO = 1

ERRORS = {0: "ok", 2: "invalid argument", 3: "HIP error", 4: "switched off", 100: "no params.json", 1001: "parse error",
          110: "CFL violated", 130: "PML too thick", 180: "bad source", 10001: "allocation failure"}


class CgfdmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code} {ERRORS.get(code, '?')}] {msg}")
        self.code = code


class _Desc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("npml", ctypes.c_int), ("free_surface", ctypes.c_int), ("scheme", ctypes.c_int),
                ("precision", ctypes.c_int), ("coef_storage", ctypes.c_int), ("nt", ctypes.c_int),
                ("dh", ctypes.c_double), ("dt", ctypes.c_double)]


_lib = None

EXPORTS = ["cgfdm_create", "cgfdm_destroy", "cgfdm_set_source", "cgfdm_set_stations", "cgfdm_step",
           "cgfdm_set_stream", "cgfdm_synchronize", "cgfdm_step_index", "cgfdm_get_stations",
           "cgfdm_get_pgv", "cgfdm_get_surface_displacement", "cgfdm_get_field", "cgfdm_set_field",
           "cgfdm_get_setup", "cgfdm_get_ranges", "cgfdm_timing", "cgfdm_device_bytes",
           "cgfdm_last_error", "cgfdm_grid_ranges", "cgfdm_forward_record", "cgfdm_adjoint", "cgfdm_get_slice",
           "cgfdm_grid_ranges_noalloc", "cgfdm_kernel_name", "cgfdm_create_from_params", "cgfdm_write_outputs",
           "cgfdm_run_info"]


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not found: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    L.cgfdm_create.argtypes = [ctypes.POINTER(_Desc), vp, vp, ctypes.POINTER(vp)]
    L.cgfdm_destroy.argtypes = [vp]
    L.cgfdm_set_source.argtypes = [vp, i, i, d, vp, vp]
    L.cgfdm_set_stations.argtypes = [vp, i, vp]
    L.cgfdm_step.argtypes = [vp, i]
    L.cgfdm_set_stream.argtypes = [vp, vp]
    L.cgfdm_synchronize.argtypes = [vp]
    L.cgfdm_step_index.argtypes = [vp]
    L.cgfdm_get_stations.argtypes = [vp, vp]
    L.cgfdm_get_pgv.argtypes = [vp, vp]
    L.cgfdm_get_surface_displacement.argtypes = [vp, vp]
    L.cgfdm_get_field.argtypes = [vp, i, vp]
    L.cgfdm_get_slice.argtypes = [vp, i, i, i, vp]
    L.cgfdm_set_field.argtypes = [vp, i, vp]
    L.cgfdm_get_setup.argtypes = [vp, i, vp]
    L.cgfdm_get_ranges.argtypes = [vp, vp]
    L.cgfdm_timing.argtypes = [vp, i, vp]
    L.cgfdm_device_bytes.restype = ll
    L.cgfdm_device_bytes.argtypes = [vp]
    L.cgfdm_last_error.restype = ctypes.c_char_p
    L.cgfdm_grid_ranges.argtypes = [i, i, i, vp, vp, vp]
    L.cgfdm_grid_ranges_noalloc.argtypes = [i, i, i, vp, vp, vp, vp]
    L.cgfdm_kernel_name.restype = ctypes.c_char_p
    L.cgfdm_kernel_name.argtypes = [vp, i]
    L.cgfdm_forward_record.argtypes = [vp, i, i, vp]
    L.cgfdm_adjoint.argtypes = [vp, vp, vp, vp, vp, vp]
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise CgfdmError(rc, lib().cgfdm_last_error().decode())


def grid_ranges(nx, ny, nz, coord, medium):
    """FP64 GPU evaluation of the CFL inputs of a grid (Hmin, Hmax and medium ranges)."""
    L = lib()
    c = np.ascontiguousarray(coord, np.float32)
    m = np.ascontiguousarray(medium, np.float32)
    out = np.empty(8, np.float64)
    _check(L.cgfdm_grid_ranges(int(nx), int(ny), int(nz), _ptr(c), _ptr(m), _ptr(out)))
    return dict(zip(["hmin", "hmax", "vsmin", "vsmax", "vpmin", "vpmax", "rhomin", "rhomax"], out))


class Solver:
    """One single-GPU domain advanced by the HIP kernels.

    scheme: SCHEME_CGFDM (MacCormack + RK4 + CFS-PML, W = velocity/stress) or
    SCHEME_AWENO (WENO5 alternative flux + TVD-RK3 + sponge, W = momentum/strain;
    the model's npml is not used).
    precision: 4 (FP32 wavefield) or 8 (FP64); coef_storage: 0 native, 1 fp16, 2 bf16 (CGFDM).
    """

    def __init__(self, model: Model, precision=4, coef_storage=0, scheme=SCHEME_CGFDM, stream=None):
        self.L = lib()
        self.m = model
        self.scheme = scheme
        desc = _Desc(model.nx, model.ny, model.nz, model.npml, 1 if model.free_surface else 0, scheme,
                     precision, coef_storage, model.nt, float(model.dh), float(model.dt))
        coord = np.ascontiguousarray(model.coord, np.float32)
        med = np.ascontiguousarray(model.medium, np.float32)
        h = ctypes.c_void_p()
        self.h = None
        if stream is not None:
            # create on the caller's stream: set after creation (setup runs on the default stream)
            pass
        _check(self.L.cgfdm_create(ctypes.byref(desc), _ptr(coord), _ptr(med), ctypes.byref(h)))
        self.h = h
        if stream is not None:
            self.set_stream(stream)
        NZ, NY, NX = model.shape_padded
        self.shape = (NZ, NY, NX)
        self.N = NZ * NY * NX
        self.ns = 0
        if model.src_cells is not None and len(model.src_cells):
            cells = np.ascontiguousarray(model.src_cells, np.int64)
            rate = np.ascontiguousarray(model.src_rate, np.float32)
            _check(self.L.cgfdm_set_source(self.h, len(cells), rate.shape[1], float(model.src_dt),
                                           _ptr(cells), _ptr(rate)))
        if model.stations is not None and len(model.stations):
            st = np.ascontiguousarray(model.stations, np.int32)
            _check(self.L.cgfdm_set_stations(self.h, len(st), _ptr(st)))
            self.ns = len(st)

    def close(self):
        if getattr(self, "h", None):
            self.L.cgfdm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream, an int handle or None (default stream)."""
        handle = getattr(stream, "cuda_stream", stream)
        _check(self.L.cgfdm_set_stream(self.h, ctypes.c_void_p(handle or 0)))

    def step(self, n=1):
        _check(self.L.cgfdm_step(self.h, int(n)))

    def synchronize(self):
        _check(self.L.cgfdm_synchronize(self.h))

    @property
    def it(self):
        return self.L.cgfdm_step_index(self.h)

    def stations(self):
        out = np.zeros((self.ns, self.m.nt, 9), np.float32)
        if self.ns:
            _check(self.L.cgfdm_get_stations(self.h, _ptr(out)))
        return out

    def pgv(self):
        out = np.zeros((self.m.ny, self.m.nx, 2), np.float32)
        _check(self.L.cgfdm_get_pgv(self.h, _ptr(out)))
        return out

    def surface_displacement(self):
        out = np.zeros((3, self.m.ny, self.m.nx), np.float32)
        _check(self.L.cgfdm_get_surface_displacement(self.h, _ptr(out)))
        return out

    def field(self, comp):
        out = np.empty(self.N, np.float64)
        _check(self.L.cgfdm_get_field(self.h, int(comp), _ptr(out)))
        return out.reshape(self.shape)

    def wavefield(self):
        return np.stack([self.field(c) for c in range(9)])

    def slice(self, comp, axis, index):
        """One component on an interior plane at padded `index`, float32 in physical
        units (Vx..Tyz for both schemes): axis 0 -> [nz][ny], 1 -> [nz][nx], 2 -> [ny][nx]."""
        NZ, NY, NX = self.shape
        nz, ny, nx = NZ - 6, NY - 6, NX - 6
        shp = {0: (nz, ny), 1: (nz, nx), 2: (ny, nx)}[int(axis)]
        out = np.empty(shp, np.float32)
        _check(self.L.cgfdm_get_slice(self.h, int(comp), int(axis), int(index), _ptr(out)))
        return out

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        _check(self.L.cgfdm_set_field(self.h, int(comp), _ptr(a)))

    def setup(self, what):
        """Setup product `what`. CGFDM: 0..8 metrics, 9 J, 10..12 mu/lambda/(1/rho),
        20..28 / 30..38 surface matrices, 40.. PML profiles. ME-AWENO: the 19
        coefficient planes (0..8 cofactors J*xi_x.., 9 J, 10..12 medium,
        13..18 x_xi, x_eta)."""
        NZ, NY, NX = self.shape
        if what < 13 or (self.scheme == SCHEME_AWENO and what < 19):
            n, shp = self.N, self.shape
        elif what < 40:
            n, shp = NY * NX, (NY, NX)
        else:
            n = (NX, NY, NZ)[(what - 40) // 3]
            shp = (n,)
        out = np.empty(n, np.float64)
        _check(self.L.cgfdm_get_setup(self.h, int(what), _ptr(out)))
        return out.reshape(shp)

    def ranges(self):
        out = np.empty(8, np.float64)
        _check(self.L.cgfdm_get_ranges(self.h, _ptr(out)))
        return dict(zip(["hmin", "hmax", "vsmin", "vsmax", "vpmin", "vpmax", "rhomin", "rhomax"], out))

    def timing(self, enable):
        out = np.zeros(6, np.float64)
        _check(self.L.cgfdm_timing(self.h, 1 if enable else 0, _ptr(out)))
        return dict(interior_ms=out[0], interior_n=int(out[1]), surface_ms=out[2], surface_n=int(out[3]),
                    step_ms=out[4], step_n=int(out[5]))

    def forward_record(self, nsteps, checkpoint_every=100):
        """Run nsteps from the current state keeping checkpoints for `adjoint`;
        returns the station velocity [ns, nsteps, 3] (CGFDM scheme)."""
        out = np.zeros((max(self.ns, 1), nsteps, 3), np.float64)
        _check(self.L.cgfdm_forward_record(self.h, int(nsteps), int(checkpoint_every), _ptr(out)))
        self._fr_n = nsteps
        return out[: self.ns]

    def adjoint(self, src):
        """Adjoint of the last forward_record for the adjoint source
        src[ns, nsteps, 3] = dJ/d(station velocity): gradients of J with respect to
        mu, lambda, 1/rho per padded cell and to the moment-rate table given to
        set_source ([npts, nt, 6])."""
        NZ, NY, NX = self.shape
        src = np.ascontiguousarray(src, np.float64)
        # every element is written by the library
        gmu, glam, gb = (np.empty(self.N) for _ in range(3))
        npts = 0 if self.m.src_cells is None else len(self.m.src_cells)
        nt = 0 if self.m.src_rate is None else self.m.src_rate.shape[1]
        grate = np.zeros((max(npts, 1), max(nt, 1), 6))
        _check(self.L.cgfdm_adjoint(self.h, _ptr(src), _ptr(gmu), _ptr(glam), _ptr(gb), _ptr(grate)))
        return dict(gmu=gmu.reshape(self.shape), glam=glam.reshape(self.shape), gb=gb.reshape(self.shape),
                    grate=grate[:npts])

    def kernel_name(self, which=0):
        """Name of the kernel the solver launches: 0 = the interior (CGFDM) or
        face-flux (ME-AWENO) kernel, 1 = the surface or update kernel."""
        return self.L.cgfdm_kernel_name(self.h, int(which)).decode()

    def device_bytes(self):
        return int(self.L.cgfdm_device_bytes(self.h))

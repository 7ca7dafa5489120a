"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
H = 3  # ghost layers on every side (reference macro.h:39)
_LIB = None


def build(force=False):
    so = os.path.join(_HERE, "liboracle.so")
    srcs = [os.path.join(_HERE, f) for f in ("cgfdm_oracle.cpp", "aweno_oracle.cpp")]
    if force or not os.path.exists(so) or os.path.getmtime(so) < max(os.path.getmtime(f) for f in srcs):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        vp, i, d, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, i, i, i, i, i, d, d, i, vp, vp]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_source.argtypes = [vp, i, i, d, vp, vp]
        L.orc_set_stations.argtypes = [vp, i, vp]
        L.orc_set_threads.argtypes = [i]
        L.orc_step.argtypes = [vp, i]
        L.orc_get_field.argtypes = [vp, i, i, vp]
        L.orc_set_field.argtypes = [vp, i, vp]
        L.orc_get_setup.argtypes = [vp, i, vp]
        L.orc_get_ranges.argtypes = [vp, vp]
        L.orc_get_stations.argtypes = [vp, vp]
        L.orc_get_pgv.argtypes = [vp, vp]
        L.orc_get_surface_dis.argtypes = [vp, vp]
        L.orc_face_count.restype = ll
        L.orc_face_count.argtypes = [vp, i]
        L.orc_get_face.argtypes = [vp, i, i, vp]
        L.orc_get_face_rhs.argtypes = [vp, i, i, vp]
        L.orc_set_face.argtypes = [vp, i, i, vp]
        L.orc_face_info.argtypes = [vp, i, vp]
        L.orc_source_only.argtypes = [vp]
        L.orc_rhs.argtypes = [vp, i, i, i, i]
        L.orc_rk_stage.argtypes = [i, i, ll, vp, vp, vp, vp]
        L.orc_step_index.restype = i
        L.orc_step_index.argtypes = [vp]
        L.orc_adjoint_dot.restype = i
        L.orc_adjoint_dot.argtypes = [vp, i, ctypes.c_ulonglong, vp]
        L.orc_set_medium_params.argtypes = [vp, vp, vp, vp]
        L.orc_set_step_index.argtypes = [vp, i]
        L.orc_adjoint_misfit.restype = d
        L.orc_adjoint_misfit.argtypes = [vp, i, vp, vp, vp, vp, vp]
        L.awo_create.restype = vp
        L.awo_create.argtypes = [i, i, i, i, i, d, d, i, vp, vp]
        L.awo_destroy.argtypes = [vp]
        L.awo_set_source.argtypes = [vp, i, i, d, vp, vp]
        L.awo_set_stations.argtypes = [vp, i, vp]
        L.awo_step.argtypes = [vp, i]
        L.awo_get_field.argtypes = [vp, i, i, vp]
        L.awo_set_field.argtypes = [vp, i, vp]
        L.awo_get_setup.argtypes = [vp, i, vp]
        L.awo_rhs.argtypes = [vp]
        L.awo_char_free_surface.argtypes = [vp]
        L.awo_weno.restype = d
        L.awo_weno.argtypes = [d, d, d, d, d, i]
        L.awo_get_stations.argtypes = [vp, vp]
        L.awo_get_pgv.argtypes = [vp, vp]
        L.awo_get_surface_dis.argtypes = [vp, vp]
        L.awo_step_index.restype = i
        L.awo_step_index.argtypes = [vp]
        L.awo_rhs_dir.argtypes = [vp, i]
        L.awo_source_only.argtypes = [vp]
        L.awo_observe.argtypes = [vp]
        L.awo_sponge.argtypes = [vp]
        L.awo_tvd_stage.argtypes = [i, i, ll, vp, vp, vp]
        L.awo_alpha.restype = d
        L.awo_alpha.argtypes = [vp]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """CPU restatement of the CGFDM step on a `Model` (prec 4 = float, 8 = double)."""

    def __init__(self, model, prec=8, threads=None):
        self.L = lib()
        self.m = model
        self.prec = prec
        if threads:
            self.L.orc_set_threads(int(threads))
        self._coord = np.ascontiguousarray(model.coord, np.float32)
        self._med = np.ascontiguousarray(model.medium, np.float32)
        self.h = self.L.orc_create(prec, model.nx, model.ny, model.nz, model.npml,
                                   1 if model.free_surface else 0, float(model.dh), float(model.dt),
                                   int(model.nt), _ptr(self._coord), _ptr(self._med))
        NZ, NY, NX = model.shape_padded
        self.shape = (NZ, NY, NX)
        self.N = NZ * NY * NX
        if model.src_cells is not None and len(model.src_cells):
            cells = np.ascontiguousarray(model.src_cells, np.int64)
            rate = np.ascontiguousarray(model.src_rate, np.float32)
            self.L.orc_set_source(self.h, len(cells), rate.shape[1], float(model.src_dt), _ptr(cells), _ptr(rate))
        self.ns = 0
        if model.stations is not None and len(model.stations):
            st = np.ascontiguousarray(model.stations, np.int32)
            self.ns = len(st)
            self.L.orc_set_stations(self.h, self.ns, _ptr(st))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def step(self, n=1):
        self.L.orc_step(self.h, int(n))

    @property
    def it(self):
        return self.L.orc_step_index(self.h)

    def field(self, comp, which=0):
        out = np.empty(self.N, np.float64)
        self.L.orc_get_field(self.h, which, comp, _ptr(out))
        return out.reshape(self.shape)

    def wavefield(self, which=0):
        return np.stack([self.field(c, which) for c in range(9)])

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        self.L.orc_set_field(self.h, comp, _ptr(a))

    def setup(self, what):
        NZ, NY, NX = self.shape
        if what < 13:
            n, shp = self.N, self.shape
        elif what < 40:
            n, shp = NY * NX, (NY, NX)
        else:
            ax = (what - 40) // 3
            n = (NX, NY, NZ)[ax]
            shp = (n,)
        out = np.empty(n, np.float64)
        self.L.orc_get_setup(self.h, what, _ptr(out))
        return out.reshape(shp)

    def ranges(self):
        out = np.empty(8, np.float64)
        self.L.orc_get_ranges(self.h, _ptr(out))
        return dict(zip(["hmin", "hmax", "vsmin", "vsmax", "vpmin", "vpmax", "rhomin", "rhomax"], out))

    def stations(self):
        out = np.zeros((self.ns, self.m.nt, 9), np.float32)
        if self.ns:
            self.L.orc_get_stations(self.h, _ptr(out))
        return out

    def pgv(self):
        out = np.zeros((self.m.ny, self.m.nx, 2), np.float32)
        self.L.orc_get_pgv(self.h, _ptr(out))
        return out

    def surface_dis(self):
        NZ, NY, NX = self.shape
        out = np.zeros((3, NY, NX), np.float64)
        self.L.orc_get_surface_dis(self.h, _ptr(out))
        return out

    def rhs(self, fb, with_boundaries=True):
        self.L.orc_rhs(self.h, int(fb[0]), int(fb[1]), int(fb[2]), 1 if with_boundaries else 0)
        return self.wavefield(which=3)

    def face(self, f, comp):
        n = self.L.orc_face_count(self.h, f)
        out = np.empty(n, np.float64)
        self.L.orc_get_face(self.h, f, comp, _ptr(out))
        return out

    def face_rhs(self, f, comp):
        """h_Aux of face f, component comp (after rhs(..., with_boundaries=True))."""
        n = self.L.orc_face_count(self.h, f)
        out = np.empty(n, np.float64)
        self.L.orc_get_face_rhs(self.h, f, comp, _ptr(out))
        return out

    def set_face(self, f, comp, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        assert a.size == self.L.orc_face_count(self.h, f)
        self.L.orc_set_face(self.h, f, comp, _ptr(a))

    def face_info(self, f):
        """(axis, side) of face f in launch order."""
        out = np.zeros(2, np.int32)
        self.L.orc_face_info(self.h, f, _ptr(out))
        return int(out[0]), int(out[1])

    def face_cells(self, f):
        """Padded (i, j, k) of every entry of face f's slab, in slab order."""
        ax, side = self.face_info(f)
        m = self.m
        n0, n1 = {0: (m.npml, m.ny), 1: (m.nx, m.npml), 2: (m.nx, m.ny)}[ax]
        q = np.arange(self.L.orc_face_count(self.h, f))
        i = q % n0 + H
        j = (q // n0) % n1 + H
        k = q // (n0 * n1) + H
        if side:
            if ax == 0:
                i = i + m.nx - m.npml
            elif ax == 1:
                j = j + m.ny - m.npml
            else:
                k = k + m.nz - m.npml
        return i, j, k

    def source_only(self):
        """Inject the source at the current step index without stepping."""
        self.L.orc_source_only(self.h)

    def set_step_index(self, it):
        self.L.orc_set_step_index(self.h, int(it))


def adjoint_dot(o, which, seed=1):
    """(<Op X, Y>, <X, Op^T Y>) for random X, Y; which 0 = stage right-hand side
    with the F/B pattern of the current step index, 1 = the whole RK4 step."""
    out = np.zeros(2)
    rc = lib().orc_adjoint_dot(o.h, int(which), int(seed), _ptr(out))
    assert rc == 0
    return out


def set_step_index(o, it):
    lib().orc_set_step_index(o.h, int(it))


def set_medium_params(o, mu, lam, b):
    """Overwrite (mu, lambda, 1/rho) per padded cell (FP64 arrays)."""
    a = [np.ascontiguousarray(x, np.float64).ravel() for x in (mu, lam, b)]
    lib().orc_set_medium_params(o.h, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]))


def adjoint_misfit(o, nsteps, data, npts=0, nt_src=0):
    """J = 1/2 sum (station velocity - data)^2 over nsteps steps from the current
    state, and its adjoint gradient: dict(J, gmu, glam, gb, grate[npts][nt][6])."""
    data = np.ascontiguousarray(data, np.float64)
    N = o.N
    gmu, glam, gb = np.zeros(N), np.zeros(N), np.zeros(N)
    grate = np.zeros((max(npts, 1), max(nt_src, 1), 6))
    J = lib().orc_adjoint_misfit(o.h, int(nsteps), _ptr(data), _ptr(gmu), _ptr(glam), _ptr(gb),
                                 _ptr(grate) if npts else None)
    return dict(J=J, gmu=gmu.reshape(o.shape), glam=glam.reshape(o.shape), gb=gb.reshape(o.shape),
                grate=grate[:npts])


def rk_stage(prec, irk, h, w, t, m):
    """Apply one low-storage RK4 stage combination in place on flat arrays."""
    L = lib()
    L.orc_rk_stage(prec, irk, h.size, _ptr(h), _ptr(w), _ptr(t), _ptr(m))


class AwenoOracle:
    """CPU restatement of the ME-AWENO (SCFDM) step on a `Model`.

    W = (rho vx, rho vy, rho vz, exx, eyy, ezz, gyz, gxz, gxy); the model's npml
    is ignored (the scheme absorbs with the 20-layer exponential sponge).
    """

    def __init__(self, model, prec=8, threads=None):
        self.L = lib()
        self.m = model
        self.prec = prec
        if threads:
            self.L.orc_set_threads(int(threads))
        self._coord = np.ascontiguousarray(model.coord, np.float32)
        self._med = np.ascontiguousarray(model.medium, np.float32)
        self.h = self.L.awo_create(prec, model.nx, model.ny, model.nz, 1 if model.free_surface else 0,
                                   float(model.dh), float(model.dt), int(model.nt),
                                   _ptr(self._coord), _ptr(self._med))
        NZ, NY, NX = model.shape_padded
        self.shape = (NZ, NY, NX)
        self.N = NZ * NY * NX
        self.ns = 0
        if model.src_cells is not None and len(model.src_cells):
            cells = np.ascontiguousarray(model.src_cells, np.int64)
            rate = np.ascontiguousarray(model.src_rate, np.float32)
            self.L.awo_set_source(self.h, len(cells), rate.shape[1], float(model.src_dt), _ptr(cells), _ptr(rate))
        if model.stations is not None and len(model.stations):
            st = np.ascontiguousarray(model.stations, np.int32)
            self.L.awo_set_stations(self.h, len(st), _ptr(st))
            self.ns = len(st)

    def __del__(self):
        try:
            if self.h:
                self.L.awo_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def step(self, n=1):
        self.L.awo_step(self.h, int(n))

    @property
    def it(self):
        return self.L.awo_step_index(self.h)

    @property
    def alpha(self):
        return self.L.awo_alpha(self.h)

    def field(self, comp, which=0):
        """which: 0 = W, 1 = h_W (last right-hand side), 2 = last face-flux buffer."""
        out = np.empty(self.N, np.float64)
        self.L.awo_get_field(self.h, int(which), int(comp), _ptr(out))
        return out.reshape(self.shape)

    def wavefield(self, which=0):
        return np.stack([self.field(c, which) for c in range(9)])

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        self.L.awo_set_field(self.h, int(comp), _ptr(a))

    def setup(self, what):
        """CJM plane 0..18: cofactors J*xi_x..J*zt_z, J, mu, lambda, 1/rho, x_xi..z_et."""
        out = np.empty(self.N, np.float64)
        self.L.awo_get_setup(self.h, int(what), _ptr(out))
        return out.reshape(self.shape)

    def rhs(self):
        """h_W = DT * (flux-difference right-hand side) of the current W."""
        self.L.awo_rhs(self.h)
        return self.wavefield(1)

    def rhs_dir(self, d):
        """h_W of one direction (0 = x, 1 = y, 2 = z) of the current W."""
        self.L.awo_rhs_dir(self.h, int(d))
        return self.wavefield(1)

    def char_free_surface(self):
        self.L.awo_char_free_surface(self.h)

    def source_only(self):
        self.L.awo_source_only(self.h)

    def observe(self):
        self.L.awo_observe(self.h)

    def sponge(self):
        self.L.awo_sponge(self.h)

    def stations(self):
        out = np.zeros((self.ns, self.m.nt, 9), np.float32)
        if self.ns:
            self.L.awo_get_stations(self.h, _ptr(out))
        return out

    def pgv(self):
        out = np.zeros((self.m.ny, self.m.nx, 2), np.float32)
        self.L.awo_get_pgv(self.h, _ptr(out))
        return out

    def surface_dis(self):
        out = np.zeros((3, self.m.ny, self.m.nx), np.float32)
        self.L.awo_get_surface_dis(self.h, _ptr(out))
        return out


def tvd_stage(prec, irk, h, w, m):
    """One TVD-RK3 stage on float64 arrays (w and m updated in place)."""
    for a in (h, w, m):
        assert a.dtype == np.float64 and a.flags.c_contiguous and a.size % 9 == 0
    lib().awo_tvd_stage(int(prec), int(irk), h.size, _ptr(h), _ptr(w), _ptr(m))


def weno(u, prec=8):
    """WENO5-JS value at the right face of the middle of five samples."""
    return lib().awo_weno(*[float(x) for x in u], int(prec))

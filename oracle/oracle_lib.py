"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the compiled
reference kernels (oracle/_ref/libhbiref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libhbiref.so")
REFERENCE_DIR = "/root/reference"

_FIELDS = ("xcol", "ycol", "zcol", "xel", "xer", "ang", "angd", "dsl", "dsd", "rake",
           "xs1", "xs2", "xs3", "ys1", "ys2", "ys3", "zs1", "zs2", "zs3",
           "ev11", "ev12", "ev13", "ev21", "ev22", "ev23", "ev31", "ev32", "ev33",
           "yel", "yer", "xs4", "zs4", "ds")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OgGeom(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("vcomp", C.c_int), ("md", C.c_int), ("w", C.c_double)] + \
               [(f, _dp) for f in _FIELDS]


def build(force: bool = False) -> None:
    """Compile the oracle (and, where the reference sources exist, the
    reference kernels) with oracle/Makefile."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir(REFERENCE_DIR) and (force or not os.path.exists(REFLIB)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.og_entry.restype = C.c_double
        L.og_entry.argtypes = [C.POINTER(OgGeom), C.c_int, C.c_int]
        L.og_entries.argtypes = [C.POINTER(OgGeom), C.c_int, _ip, _ip, _dp]
        L.og_hmat_build.restype = C.c_void_p
        L.og_hmat_build.argtypes = [C.POINTER(OgGeom), C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]
        L.og_hmat_partition.restype = C.c_void_p
        L.og_hmat_partition.argtypes = [C.POINTER(OgGeom), C.c_int, C.c_double]
        L.og_hmat_nleaves.restype = C.c_int
        L.og_hmat_nleaves.argtypes = [C.c_void_p]
        L.og_hmat_lod.argtypes = [C.c_void_p, _ip]
        L.og_hmat_leaves.argtypes = [C.c_void_p, _ip]
        L.og_hmat_leaf_data.argtypes = [C.c_void_p, C.c_int, _dp]
        L.og_hmat_free.argtypes = [C.c_void_p]
        L.og_hmat_matvec.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.og_dense_cluster.argtypes = [C.c_void_p, _dp]
        L.og_sim_create.restype = C.c_void_p
        L.og_sim_create.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double] + [_dp] * 7 + [C.c_int]
        L.og_sim_free.argtypes = [C.c_void_p]
        L.og_sim_set_opening.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.og_sim_nmatvec.restype = C.c_longlong
        L.og_sim_nmatvec.argtypes = [C.c_void_p]
        L.og_derivs.argtypes = [C.c_void_p, _dp, _dp]
        L.og_rkck.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp]
        L.og_errmax.restype = C.c_double
        L.og_errmax.argtypes = [C.c_int, _dp, _dp, C.c_double]
        L.og_rkqs.restype = C.c_int
        L.og_rkqs.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _ip]
        L.og_run_grid.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_double, _dp, _dp, _dp, _dp, C.c_int,
                                  _ip, _dp, _dp]
        L.og_rkqs_control.restype = C.c_int
        L.og_rkqs_control.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.og_post_step.restype = C.c_double
        L.og_post_step.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.og_run.restype = C.c_int
        L.og_run.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_int, _dp, _dp, _dp, C.c_double,
                             C.c_double, _dp, C.c_int, _ip, _dp, _dp]
        L.og_deriv.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double] + [C.c_double] * 10 + [_dp] * 3
        L.og_tableau_out.argtypes = [_dp]
        L.og_mesh_3ddip.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double] + [_dp] * 5
        L.og_mesh_2dp.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp]
        L.og_dc3d.argtypes = [C.c_double] * 11 + [C.c_double, C.c_int, _dp]
        L.og_tdstressfs.argtypes = [C.c_double] * 3 + [_dp] * 3 + [C.c_double] * 5 + [_dp]
        L.og_tdstresshs.argtypes = [C.c_double] * 3 + [_dp] * 3 + [C.c_double] * 5 + [_dp]
        L.og_td_harfunc.argtypes = [C.c_double] * 3 + [_dp] * 3 + [C.c_double] * 5 + [_dp]
        L.og_angfsc.argtypes = [C.c_double] * 9 + [_dp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def rkqs_control(h, errmax, x, dtmax):
    """Step-size decision of rkqs (main_LH.f90:2335-2357): (status, h_new) with
    status 1 = accepted (h_new is the next trial step), 0 = rejected (h_new is
    the retry), -1 = step-size underflow."""
    hn = np.zeros(1)
    st = lib().og_rkqs_control(float(h), float(errmax), float(x), float(dtmax), _ptr(hn))
    return int(st), float(hn[0])


def post_step(y, a, vref, dtdid, ds, slip, vel):
    """Post-step update (main_LH.f90:1266-1324) in place on slip and vel;
    returns (max |V|, mean slip)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    ds = np.ascontiguousarray(ds, dtype=np.float64)
    ms = np.zeros(1)
    mv = lib().og_post_step(len(slip), _ptr(y), _ptr(a), float(vref), float(dtdid), _ptr(ds), _ptr(slip),
                            _ptr(vel), _ptr(ms))
    return float(mv), float(ms[0])


def _iptr(a):
    return a.ctypes.data_as(_ip)


class Geom:
    """Holds an OgGeom and keeps the numpy arrays it points at alive."""

    def __init__(self, g):
        self.keep = []
        s = OgGeom()
        s.kind, s.n, s.vcomp, s.md, s.w = int(g.kind), int(g.n), int(g.vcomp), int(g.md), float(g.w)
        for f in _FIELDS:
            a = g.arrays.get(f)
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.float64)
                self.keep.append(a)
                setattr(s, f, _ptr(a))
        self.s = s
        self.n = int(g.n)


def entries(g, ii, jj) -> np.ndarray:
    G = Geom(g)
    ii = np.ascontiguousarray(ii, dtype=np.int32)
    jj = np.ascontiguousarray(jj, dtype=np.int32)
    out = np.empty(len(ii))
    lib().og_entries(C.byref(G.s), len(ii), _iptr(ii), _iptr(jj), _ptr(out))
    return out


def dc3d(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, disl1, disl2, fullspace=False):
    u = np.empty(12)
    lib().og_dc3d(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, disl1, disl2, int(fullspace), _ptr(u))
    return u


def tdstressfs(X, Y, Z, P1, P2, P3, ss, ds, ts, mu, lam):
    P1, P2, P3 = (np.ascontiguousarray(p, dtype=np.float64) for p in (P1, P2, P3))
    S = np.empty(6)
    lib().og_tdstressfs(X, Y, Z, _ptr(P1), _ptr(P2), _ptr(P3), ss, ds, ts, mu, lam, _ptr(S))
    return S


def tdstresshs(X, Y, Z, P1, P2, P3, ss, ds, ts, mu, lam, part="total"):
    """Half-space stress of a triangular dislocation (TDstressHS), or only
    its harmonic part (part="harmonic", TDstress_HarFunc)."""
    P1, P2, P3 = (np.ascontiguousarray(p, dtype=np.float64) for p in (P1, P2, P3))
    S = np.empty(6)
    f = lib().og_tdstresshs if part == "total" else lib().og_td_harfunc
    f(X, Y, Z, _ptr(P1), _ptr(P2), _ptr(P3), ss, ds, ts, mu, lam, _ptr(S))
    return S


def angfsc(y1, y2, y3, beta, b1, b2, b3, nu, a) -> np.ndarray:
    """Free-surface correction strains (v11 v22 v33 v12 v13 v23) of an angular dislocation."""
    v = np.empty(6)
    lib().og_angfsc(y1, y2, y3, beta, b1, b2, b3, nu, a, _ptr(v))
    return v


def tableau() -> np.ndarray:
    o = np.empty(29)
    lib().og_tableau_out(_ptr(o))
    return o


class HMatrix:
    """Oracle lattice H-matrix (np=1): partition, ACA+ factors, FP64 matvec."""

    def __init__(self, g, leaf=15, eta=2.0, eps=1e-4, kmax=1000, nthreads=0, partition_only=False):
        self.G = Geom(g)
        L = lib()
        if partition_only:
            self.h = L.og_hmat_partition(C.byref(self.G.s), leaf, eta)
        else:
            self.h = L.og_hmat_build(C.byref(self.G.s), leaf, eta, eps, kmax, nthreads)
        self.n = self.G.n

    def __del__(self):
        if getattr(self, "h", None):
            lib().og_hmat_free(self.h)
            self.h = None

    @property
    def lod(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        lib().og_hmat_lod(self.h, _iptr(out))
        return out

    @property
    def leaves(self) -> np.ndarray:
        nl = lib().og_hmat_nleaves(self.h)
        out = np.empty((nl, 6), dtype=np.int32)
        lib().og_hmat_leaves(self.h, _iptr(out))
        return out

    def leaf_data(self, ip: int) -> np.ndarray:
        lt, _, ndl, _, ndt, kt = self.leaves[ip]
        size = (ndl + ndt) * kt if lt == 1 else ndl * ndt
        out = np.empty(max(size, 1))
        lib().og_hmat_leaf_data(self.h, ip, _ptr(out))
        return out[:size]

    def matvec(self, x, nthreads=1) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        lib().og_hmat_matvec(self.h, _ptr(x), _ptr(y), nthreads)
        return y

    def dense(self) -> np.ndarray:
        A = np.empty((self.n, self.n))
        lib().og_dense_cluster(self.h, _ptr(A))
        return A


class Sim:
    """Oracle ODE system: friction RHS + Cash-Karp RK45 with the reference's controller."""

    def __init__(self, hs: HMatrix, fr, lod, hdiv=0.0, hn: HMatrix = None, backslip=True, limitsigma=True,
                 minsig=1.0, maxsig=300.0, dtmax=1e10, nthreads=1):
        self.hs, self.hn = hs, hn
        self.arr = [np.ascontiguousarray(getattr(fr, k)[lod], dtype=np.float64)
                    for k in ("a", "b", "dc", "f0", "vpl", "taudot", "sigdot")]
        self.s = lib().og_sim_create(hs.h, hn.h if hn is not None else None, int(fr.law), int(backslip),
                                     int(limitsigma), float(fr.vref), minsig, maxsig, dtmax, float(hdiv),
                                     *[_ptr(a) for a in self.arr], int(nthreads))
        self.n = hs.n

    def __del__(self):
        if getattr(self, "s", None):
            lib().og_sim_free(self.s)
            self.s = None

    def set_opening(self, hs2: "HMatrix", hn2: "HMatrix"):
        """Opening (main_LH.f90:705-716, 2096-2143): tensile matrices onto shear and normal traction."""
        self._open = (hs2, hn2)
        if lib().og_sim_set_opening(self.s, hs2.h, hn2.h) != 0:
            raise ValueError("opening needs the normal matrix and matrices of the same size")

    @property
    def nmatvec(self):
        return lib().og_sim_nmatvec(self.s)

    def derivs(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        d = np.empty_like(y)
        lib().og_derivs(self.s, _ptr(y), _ptr(d))
        return d

    def rkck(self, y, h):
        y = np.ascontiguousarray(y, dtype=np.float64)
        yo = np.empty_like(y)
        ye = np.empty_like(y)
        lib().og_rkck(self.s, _ptr(y), h, _ptr(yo), _ptr(ye))
        return yo, ye

    def rkqs(self, y, x, htry, eps):
        """In-place step; returns (status, x, hdid, hnext, errmax, nrjct)."""
        xx = np.array([x], dtype=np.float64)
        hd = np.zeros(1)
        hn = np.zeros(1)
        em = np.zeros(1)
        nr = np.zeros(1, dtype=np.int32)
        st = lib().og_rkqs(self.s, _ptr(y), _ptr(xx), htry, eps, _ptr(hd), _ptr(hn), _ptr(em), _iptr(nr))
        return st, float(xx[0]), float(hd[0]), float(hn[0]), float(em[0]), int(nr[0])

    def run(self, y, x, dt, eps_r, nsteps, slip, vel, ds, velmin=1e-20, velmax=1e7, rec=None, slips=False):
        """Run nsteps accepted steps in place; returns (steps_done, x, dt, hist, vhist),
        and with slips=True also the slip history at rec (steps_done, nrec)."""
        xx = np.array([x], dtype=np.float64)
        dd = np.array([dt], dtype=np.float64)
        hist = np.zeros((nsteps, 6))
        rec = np.ascontiguousarray(rec if rec is not None else [], dtype=np.int32)
        vh = np.zeros((nsteps, max(len(rec), 1)))
        sh = np.zeros((nsteps, max(len(rec), 1))) if slips else None
        done = lib().og_run(self.s, _ptr(y), _ptr(xx), _ptr(dd), eps_r, nsteps, _ptr(slip), _ptr(vel), _ptr(ds),
                            velmin, velmax, _ptr(hist), len(rec), _iptr(rec), _ptr(vh),
                            _ptr(sh) if slips else None)
        out = (done, float(xx[0]), float(dd[0]), hist[:done], vh[:done, :len(rec)])
        return out + (sh[:done, :len(rec)],) if slips else out


    def run_grid(self, y, x, hs, eps_r, slip, vel, ds, rec=None):
        """Replay the step sizes hs in place (every step taken); returns (x,
        hist, vhist, shist) with hist rows as run's (rejections 0)."""
        hs = np.ascontiguousarray(hs, dtype=np.float64)
        n = len(hs)
        xx = np.array([x], dtype=np.float64)
        hist = np.zeros((n, 6))
        rec = np.ascontiguousarray(rec if rec is not None else [], dtype=np.int32)
        vh = np.zeros((n, max(len(rec), 1)))
        sh = np.zeros((n, max(len(rec), 1)))
        lib().og_run_grid(self.s, _ptr(y), _ptr(xx), _ptr(hs), n, eps_r, _ptr(slip), _ptr(vel), _ptr(ds), _ptr(hist),
                          len(rec), _iptr(rec), _ptr(vh), _ptr(sh))
        return float(xx[0]), hist, vh[:, :len(rec)], sh[:, :len(rec)]


# ---------------------------------------------------------------------------
# the compiled reference kernels (only where /root/reference was present at build time)
# ---------------------------------------------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REFLIB)


def ref_entries(g, ii, jj) -> np.ndarray:
    """HACApK_entry_ij of the reference itself, via oracle/_ref/libhbiref.so."""
    global _ref
    if _ref is None:
        _ref = C.CDLL(REFLIB)
        _ref.hbiref_entries.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _ip, _ip, _dp]
    geo = g.geo_matrix()
    ii = np.ascontiguousarray(ii, dtype=np.int32)
    jj = np.ascontiguousarray(jj, dtype=np.int32)
    out = np.empty(len(ii))
    _ref.hbiref_entries(int(g.kind), int(g.n), int(g.vcomp), int(g.md), float(g.w),
                        geo.ctypes.data_as(_dp), len(ii), _iptr(ii), _iptr(jj), _ptr(out))
    return out


def _reflib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REFLIB)
        _ref.hbiref_entries.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _ip, _ip, _dp]
    _ref.hbiref_angfsc.argtypes = [C.c_double] * 9 + [_dp]
    _ref.hbiref_td_hs.argtypes = [C.c_double] * 3 + [_dp] * 3 + [C.c_double] * 5 + [_dp]
    return _ref


def ref_angfsc(y1, y2, y3, beta, b1, b2, b3, nu, a) -> np.ndarray:
    """AngDisStrainFSC of the reference itself."""
    v = np.empty(6)
    _reflib().hbiref_angfsc(y1, y2, y3, beta, b1, b2, b3, nu, a, _ptr(v))
    return v


def ref_td_hs(X, Y, Z, P1, P2, P3, ss, ds, ts, mu, lam) -> np.ndarray:
    """TDstressHS (first 6) and TDstress_HarFunc (last 6) of the reference itself."""
    P1, P2, P3 = (np.ascontiguousarray(p, dtype=np.float64) for p in (P1, P2, P3))
    o = np.empty(12)
    _reflib().hbiref_td_hs(X, Y, Z, _ptr(P1), _ptr(P2), _ptr(P3), ss, ds, ts, mu, lam, _ptr(o))
    return o

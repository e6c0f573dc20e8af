"""ctypes binding of libhbi.so (include/hbi.h).

The HIP library is the product path: there is no CPU fallback.  Loading
raises if hbi_amd/libhbi.so is missing (build it with ``python -m hbi_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libhbi.so")

PREC_F64, PREC_F32, PREC_MIXED = 0, 1, 2
E_ARG, E_STEP = -1, -4

_FIELDS = ("xcol", "ycol", "zcol", "xel", "xer", "ang", "angd", "dsl", "dsd", "rake",
           "xs1", "xs2", "xs3", "ys1", "ys2", "ys3", "zs1", "zs2", "zs3",
           "ev11", "ev12", "ev13", "ev21", "ev22", "ev23", "ev31", "ev32", "ev33",
           "yel", "yer", "xs4", "zs4", "ds")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Geom(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int), ("vcomp", C.c_int), ("md", C.c_int), ("w", C.c_double)] + \
               [(f, _dp) for f in _FIELDS]


class HmatOpts(C.Structure):
    _fields_ = [("leaf", C.c_int), ("eta", C.c_double), ("eps_aca", C.c_double), ("kmax", C.c_int),
                ("bf16_margin", C.c_double), ("tile_rows", C.c_int), ("chunk_kb", C.c_int),
                ("dense_launch", C.c_int)]


class HmatStats(C.Structure):
    _fields_ = [("n", C.c_int), ("nleaf", C.c_int), ("nlr", C.c_int), ("ndense", C.c_int), ("kt_max", C.c_int),
                ("tile_rows", C.c_int), ("kt_mean", C.c_double), ("dense_entries", C.c_longlong),
                ("lr_entries", C.c_longlong), ("lr_dims", C.c_longlong), ("stored_bytes", C.c_longlong * 3),
                ("alg_bytes", C.c_longlong * 3), ("alg_bytes_p1", C.c_longlong * 3),
                ("alg_bytes_p2", C.c_longlong * 3), ("chunk_p1", C.c_int * 3), ("chunk_p2", C.c_int * 3),
                ("dense_launch", C.c_int * 3), ("nbf16_blocks", C.c_int), ("ncapped", C.c_int),
                ("t_partition", C.c_double), ("t_dense", C.c_double), ("t_aca", C.c_double), ("t_pack", C.c_double)]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class RkOpts(C.Structure):
    _fields_ = [("law", C.c_int), ("backslip", C.c_int), ("limitsigma", C.c_int), ("prec", C.c_int),
                ("vref", C.c_double), ("minsig", C.c_double), ("maxsig", C.c_double), ("dtmax", C.c_double),
                ("hdiv", C.c_double), ("eps_r", C.c_double), ("velmin", C.c_double), ("velmax", C.c_double)]


RK_OUTPUT_FULL = 2   # hbi_rk_info.status: output buffers full, drain with get_output


class RkOutOpts(C.Structure):
    _fields_ = [("interval", C.c_int), ("capacity", C.c_int), ("ev_capacity", C.c_int), ("pad", C.c_int),
                ("k0", C.c_longlong), ("velth", C.c_double), ("dtout_inter", C.c_double), ("dtout_co", C.c_double)]


class RkEvent(C.Structure):
    _fields_ = [("count", C.c_longlong), ("k", C.c_longlong), ("onset_time", C.c_double), ("mw", C.c_double),
                ("moment", C.c_double), ("hypoloc", C.c_int), ("pad", C.c_int)]


class RkInfo(C.Structure):
    _fields_ = [("t", C.c_double), ("dt_next", C.c_double), ("dt_last", C.c_double), ("errmax_last", C.c_double),
                ("attempts", C.c_longlong), ("accepted", C.c_longlong), ("rejected", C.c_longlong),
                ("vmax", C.c_double), ("status", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


FV_NEUMANN, FV_DIRICHLET = 0, 1
FV_INJECTIONQ, FV_INJECTIONP, FV_STEADYSTATE, FV_THRUST, FV_OTHER = 0, 1, 2, 3, 4
FV_MAX_CELLS = 8192


class FvOpts(C.Structure):
    _fields_ = [("law", C.c_int), ("backslip", C.c_int), ("prec", C.c_int), ("pressure_dependent", C.c_int),
                ("pfconst", C.c_int), ("bcl", C.c_int), ("bcr", C.c_int), ("setting", C.c_int),
                ("outpertime", C.c_int)] + [(k, C.c_double) for k in (
                    "vref", "eps_r", "hdiv", "ds0", "kL", "kpmin", "s0", "eta", "beta", "Bs", "pbcl", "pbcr", "q0",
                    "pinj", "tinj_s", "dtout_s", "velmin", "velmax", "tmax_s")] + [
                    (k, C.c_int) for k in ("imax", "jmax", "bct", "bcb")] + [
                    (k, C.c_double) for k in ("dtmax", "minsig", "phi0")]


class FvInfo(C.Structure):
    _fields_ = [("t", C.c_double), ("dt_next", C.c_double), ("dt_last", C.c_double), ("errmax_last", C.c_double),
                ("attempts", C.c_longlong), ("accepted", C.c_longlong), ("rejected", C.c_longlong),
                ("vmax", C.c_double), ("niter", C.c_int), ("status", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class HbiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hbi error %d: %s" % (code, msg))
        self.code = code


_lib = None
SYMBOLS = ("hbi_version", "hbi_last_error", "hbi_device_count", "hbi_entries", "hbi_hmat_build", "hbi_hmat_destroy",
           "hbi_hmat_stats_get", "hbi_hmat_lod", "hbi_hmat_leaves", "hbi_hmat_leaf_data", "hbi_hmat_matvec",
           "hbi_hmat_matmul", "hbi_hmat_time_matvec", "hbi_rk_create", "hbi_rk_destroy", "hbi_rk_set_state",
           "hbi_rk_get_state", "hbi_rk_derivs", "hbi_rk_rkck", "hbi_rk_step", "hbi_rk_last_ms",
           "hbi_hmat_time_matmul", "hbi_hmat_matmul_plan", "hbi_mfma_f64_rate", "hbi_mfma_f32_rate", "hbi_hmat_matvec_noalloc", "hbi_ens_create", "hbi_ens_destroy", "hbi_ens_set_state", "hbi_ens_get_state", "hbi_ens_set_clock", "hbi_ens_get_clock",
           "hbi_ens_step", "hbi_ens_last_ms", "hbi_mlp_create", "hbi_mlp_destroy", "hbi_mlp_nparams",
           "hbi_mlp_set_params", "hbi_mlp_get_params", "hbi_mlp_forward", "hbi_mlp_loss_grad", "hbi_mlp_train",
           "hbi_rk_set_controller", "hbi_rk_set_trace", "hbi_rk_get_trace", "hbi_ens_set_controller",
           "hbi_ens_set_trace", "hbi_ens_get_trace", "hbi_ens_set_probes", "hbi_ens_set_step_grid", "hbi_rk_step_many", "hbi_rk_set_output", "hbi_rk_get_output",
           "hbi_rk_set_opening", "hbi_rk_get_opening", "hbi_rk_set_probes", "hbi_rk_get_probes",
           "hbi_fv_create", "hbi_fv_destroy", "hbi_fv_set_state", "hbi_fv_get_state", "hbi_fv_derivs",
           "hbi_fv_diffuse", "hbi_fv_step", "hbi_fv_last_ms", "hbi_fv_set_probes", "hbi_fv_get_probes", "hbi_fv_set_wells",
           "hbi_fv_set_break")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libhbi.so not built: run `python -m hbi_amd.build` (no CPU fallback exists)")
        L = C.CDLL(LIBPATH)
        L.hbi_last_error.restype = C.c_char_p
        L.hbi_entries.argtypes = [C.POINTER(Geom), C.c_int, _ip, _ip, _dp]
        L.hbi_hmat_build.argtypes = [C.POINTER(Geom), C.POINTER(HmatOpts), C.POINTER(C.c_void_p)]
        L.hbi_hmat_destroy.argtypes = [C.c_void_p]
        L.hbi_hmat_stats_get.argtypes = [C.c_void_p, C.POINTER(HmatStats)]
        L.hbi_hmat_lod.argtypes = [C.c_void_p, _ip]
        L.hbi_hmat_leaves.argtypes = [C.c_void_p, _ip]
        L.hbi_hmat_leaf_data.argtypes = [C.c_void_p, C.c_int, _dp]
        L.hbi_hmat_matvec.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int]
        L.hbi_hmat_matmul.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int]
        L.hbi_hmat_time_matvec.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.hbi_rk_create.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(RkOpts)] + [_dp] * 8 + [C.POINTER(C.c_void_p)]
        L.hbi_rk_destroy.argtypes = [C.c_void_p]
        L.hbi_rk_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double]
        L.hbi_rk_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, C.POINTER(RkInfo)]
        L.hbi_rk_derivs.argtypes = [C.c_void_p, _dp, _dp]
        L.hbi_rk_rkck.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp]
        L.hbi_rk_step.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _ip, _dp, C.POINTER(RkInfo)]
        L.hbi_rk_last_ms.argtypes = [C.c_void_p, _dp]
        L.hbi_hmat_time_matmul.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.hbi_hmat_matmul_plan.argtypes = [C.c_void_p, C.c_int, _ip]
        L.hbi_mfma_f64_rate.argtypes = [C.c_int, C.c_int, _dp]
        L.hbi_mfma_f32_rate.argtypes = [C.c_int, C.c_int, _dp]
        _lp = C.POINTER(C.c_longlong)
        L.hbi_hmat_matvec_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.hbi_ens_create.argtypes = [C.c_void_p, C.POINTER(RkOpts), C.c_int] + [_dp] * 7 + [C.POINTER(C.c_void_p)]
        L.hbi_ens_destroy.argtypes = [C.c_void_p]
        L.hbi_ens_set_state.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double]
        L.hbi_ens_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _lp, _lp, _ip]
        L.hbi_ens_set_clock.argtypes = [C.c_void_p, _dp, _dp]
        L.hbi_ens_get_clock.argtypes = [C.c_void_p, _dp, _dp]
        L.hbi_ens_step.argtypes = [C.c_void_p, C.c_int, _dp, _lp]
        L.hbi_ens_last_ms.argtypes = [C.c_void_p, _dp]
        L.hbi_ens_set_probes.argtypes = [C.c_void_p, C.c_int, _ip]
        L.hbi_ens_set_step_grid.argtypes = [C.c_void_p, C.c_longlong, _dp]
        _fp = C.POINTER(C.c_float)
        L.hbi_mlp_create.argtypes = [C.c_ulonglong, C.POINTER(C.c_void_p)]
        L.hbi_mlp_destroy.argtypes = [C.c_void_p]
        L.hbi_mlp_nparams.argtypes = []
        L.hbi_mlp_set_params.argtypes = [C.c_void_p, _fp, _fp, _fp]
        L.hbi_mlp_get_params.argtypes = [C.c_void_p, _fp, _fp, _fp]
        L.hbi_mlp_forward.argtypes = [C.c_void_p, _fp, C.c_int, _fp]
        L.hbi_mlp_loss_grad.argtypes = [C.c_void_p, _fp, _fp, C.c_int, _dp, _fp]
        L.hbi_mlp_train.argtypes = [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _dp]
        L.hbi_rk_set_controller.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.hbi_rk_set_trace.argtypes = [C.c_void_p, C.c_int]
        L.hbi_rk_get_trace.argtypes = [C.c_void_p, _dp, _dp, _ip]
        L.hbi_ens_set_controller.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.hbi_ens_set_trace.argtypes = [C.c_void_p, C.c_int]
        L.hbi_rk_step_many.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(RkInfo)]
        L.hbi_rk_set_output.argtypes = [C.c_void_p, C.POINTER(RkOutOpts)]
        L.hbi_rk_set_opening.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hbi_rk_get_opening.argtypes = [C.c_void_p, _dp, _dp]
        L.hbi_rk_set_probes.argtypes = [C.c_void_p, C.c_int, _ip]
        L.hbi_rk_get_probes.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), _dp]
        L.hbi_fv_create.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(FvOpts)] + [_dp] * 11 + [C.POINTER(C.c_void_p)]
        L.hbi_fv_destroy.argtypes = [C.c_void_p]
        L.hbi_fv_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double]
        L.hbi_fv_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.POINTER(FvInfo)]
        L.hbi_fv_derivs.argtypes = [C.c_void_p, _dp, _dp]
        L.hbi_fv_diffuse.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _ip, _dp]
        L.hbi_fv_step.argtypes = [C.c_void_p, C.c_int, _dp, C.POINTER(FvInfo)]
        L.hbi_fv_last_ms.argtypes = [C.c_void_p, _dp]
        L.hbi_fv_set_probes.argtypes = [C.c_void_p, C.c_int, _ip]
        L.hbi_fv_set_break.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.hbi_fv_get_probes.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), _dp]
        L.hbi_fv_set_wells.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp]
        L.hbi_rk_get_output.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong), _dp, _dp,
                                        C.POINTER(C.c_int), C.POINTER(RkEvent), _dp]
        L.hbi_ens_get_trace.argtypes = [C.c_void_p, _dp, _dp, _ip]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise HbiError(rc, lib().hbi_last_error().decode())


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _ipp(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _GeomHold:
    def __init__(self, g):
        self.keep = []
        s = Geom()
        s.kind, s.n, s.vcomp, s.md, s.w = int(g.kind), int(g.n), int(g.vcomp), int(g.md), float(g.w)
        for f in _FIELDS:
            a = g.arrays.get(f)
            if a is not None:
                a = _f64(a)
                self.keep.append(a)
                setattr(s, f, _p(a))
        self.s = s


def device_count() -> int:
    c = C.c_int(0)
    _check(lib().hbi_device_count(C.byref(c)))
    return c.value


def entries(g, ii, jj) -> np.ndarray:
    """K(i, j) on the GPU (HACApK_entry_ij)."""
    G = _GeomHold(g)
    ii = np.ascontiguousarray(ii, dtype=np.int32)
    jj = np.ascontiguousarray(jj, dtype=np.int32)
    out = np.empty(len(ii))
    _check(lib().hbi_entries(C.byref(G.s), len(ii), _ipp(ii), _ipp(jj), _p(out)))
    return out


class HMatrix:
    """Device lattice H-matrix (HACApK_generate + HACApK_construct_LH)."""

    def __init__(self, g, leaf=15, eta=2.0, eps=1e-4, kmax=1000, bf16_margin=0.0, tile_rows=0, chunk_kb=0,
                 dense_launch=0):
        G = _GeomHold(g)
        o = HmatOpts(leaf=leaf, eta=eta, eps_aca=eps, kmax=kmax, bf16_margin=bf16_margin, tile_rows=tile_rows,
                     chunk_kb=chunk_kb, dense_launch=dense_launch)
        h = C.c_void_p()
        _check(lib().hbi_hmat_build(C.byref(G.s), C.byref(o), C.byref(h)))
        self.h = h
        self.n = int(g.n)

    def close(self):
        if getattr(self, "h", None):
            lib().hbi_hmat_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def stats(self) -> dict:
        s = HmatStats()
        _check(lib().hbi_hmat_stats_get(self.h, C.byref(s)))
        return s.as_dict()

    @property
    def lod(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        _check(lib().hbi_hmat_lod(self.h, _ipp(out)))
        return out

    @property
    def leaves(self) -> np.ndarray:
        nl = self.stats["nleaf"]
        out = np.empty((nl, 6), dtype=np.int32)
        _check(lib().hbi_hmat_leaves(self.h, _ipp(out)))
        return out

    def leaf_data(self, ip: int) -> np.ndarray:
        lt, _, ndl, _, ndt, kt = self.leaves[ip]
        size = (ndl + ndt) * kt if lt == 1 else ndl * ndt
        out = np.empty(max(size, 1))
        _check(lib().hbi_hmat_leaf_data(self.h, int(ip), _p(out)))
        return out[:size]

    def matvec(self, x, prec=PREC_F64) -> np.ndarray:
        x = _f64(x)
        y = np.empty(self.n)
        _check(lib().hbi_hmat_matvec(self.h, _p(x), _p(y), int(prec), 0))
        return y

    def matmul(self, X, prec=PREC_F64) -> np.ndarray:
        X = np.asfortranarray(X, dtype=np.float64)
        nrhs = X.shape[1]
        Y = np.empty((self.n, nrhs), order="F")
        _check(lib().hbi_hmat_matmul(self.h, _p(X), _p(Y), int(nrhs), int(prec)))
        return Y

    def time_matvec(self, prec=PREC_F64, iters=20):
        ms = np.zeros(3)
        _check(lib().hbi_hmat_time_matvec(self.h, int(prec), int(iters), _p(ms)))
        return {"a_ms": ms[0], "b_ms": ms[1], "total_ms": ms[2]}

    def time_matmul(self, prec=PREC_F64, iters=20):
        """Device milliseconds of one 64-column product and of its two phases."""
        ms = np.zeros(3)
        _check(lib().hbi_hmat_time_matmul(self.h, int(prec), int(iters), _p(ms)))
        return {"p1_ms": ms[0], "p2_ms": ms[1], "total_ms": ms[2]}

    def matmul_plan(self, prec=PREC_F64) -> dict:
        """The autotuned 64-column plan: chunk sizes, item counts, matrix cores."""
        info = np.zeros(7, dtype=np.int32)
        _check(lib().hbi_hmat_matmul_plan(self.h, int(prec), info.ctypes.data_as(_ip)))
        return {"kc1": int(info[0]), "cols2": int(info[1]), "items1": int(info[2]), "items2": int(info[3]),
                "zpart_slots": int(info[4]), "cores": ("fp64", "fp32", "bf16")[int(info[5])],
                "xcd_phase1": bool(info[6] & 1), "stages_phase1": int((info[6] >> 1) & 7),
                "stages_phase2": int((info[6] >> 4) & 7)}


def mfma_f64_rate(waves_per_simd=2, iters=4000) -> float:
    """Measured TFLOP/s of the FP64 16x16x4 MFMA with no memory traffic."""
    t = np.zeros(1)
    _check(lib().hbi_mfma_f64_rate(int(waves_per_simd), int(iters), _p(t)))
    return float(t[0])


def mfma_f32_rate(waves_per_simd=2, iters=4000) -> float:
    """Measured TFLOP/s of the FP32 16x16x4 MFMA with no memory traffic."""
    t = np.zeros(1)
    _check(lib().hbi_mfma_f32_rate(int(waves_per_simd), int(iters), _p(t)))
    return float(t[0])


class RK:
    """Device friction + RK45 integrator over one (or two) H-matrices."""

    def __init__(self, hs: HMatrix, fr, lod, area, hdiv=0.0, eps_r=1e-6, prec=PREC_F64, hn: Optional[HMatrix] = None,
                 backslip=True, limitsigma=True, minsig=1.0, maxsig=300.0, dtmax=1e10, velmin=1e-20, velmax=1e7):
        self.hs, self.hn = hs, hn
        lod = np.asarray(lod)
        self.keep = [_f64(getattr(fr, k)[lod]) for k in ("a", "b", "dc", "f0", "vpl", "taudot", "sigdot")]
        self.keep.append(_f64(np.asarray(area)[lod]))
        o = RkOpts(law=int(fr.law), backslip=int(backslip), limitsigma=int(limitsigma), prec=int(prec),
                   vref=float(fr.vref), minsig=minsig, maxsig=maxsig, dtmax=dtmax, hdiv=float(hdiv), eps_r=eps_r,
                   velmin=velmin, velmax=velmax)
        r = C.c_void_p()
        _check(lib().hbi_rk_create(hs.h, hn.h if hn is not None else None, C.byref(o),
                                   *[_p(a) for a in self.keep], C.byref(r)))
        self.r = r
        self.n = hs.n

    def close(self):
        if getattr(self, "r", None):
            lib().hbi_rk_destroy(self.r)
            self.r = None

    def __del__(self):
        self.close()

    def set_state(self, y, t=0.0, dt=1.0, slip=None, vel=None):
        y = _f64(y)
        slip = _f64(slip) if slip is not None else None
        vel = _f64(vel) if vel is not None else None
        _check(lib().hbi_rk_set_state(self.r, _p(y), _p(slip), _p(vel), float(t), float(dt)))

    def get_state(self):
        y = np.empty(3 * self.n)
        slip = np.empty(self.n)
        vel = np.empty(self.n)
        info = RkInfo()
        _check(lib().hbi_rk_get_state(self.r, _p(y), _p(slip), _p(vel), C.byref(info)))
        return y, slip, vel, info.as_dict()

    def derivs(self, y):
        y = _f64(y)
        d = np.empty_like(y)
        _check(lib().hbi_rk_derivs(self.r, _p(y), _p(d)))
        return d

    def rkck(self, y, h):
        y = _f64(y)
        yo = np.empty_like(y)
        ye = np.empty_like(y)
        _check(lib().hbi_rk_rkck(self.r, _p(y), float(h), _p(yo), _p(ye)))
        return yo, ye

    def step(self, nsteps, history=False, rec=None):
        hist = np.zeros((nsteps, 9)) if history else None
        rec = np.ascontiguousarray(rec, dtype=np.int32) if rec is not None and len(rec) else None
        nrec = 0 if rec is None else len(rec)
        vh = np.zeros((nsteps, nrec)) if nrec else None
        info = RkInfo()
        rc = lib().hbi_rk_step(self.r, int(nsteps), _p(hist), nrec, _ipp(rec), _p(vh), C.byref(info))
        _check(rc)
        d = info.as_dict()
        return d, hist, vh

    def last_ms(self) -> float:
        ms = C.c_double(0)
        _check(lib().hbi_rk_last_ms(self.r, C.byref(ms)))
        return ms.value

    def set_output(self, velth=1e-2, dtout_inter=365 * 24 * 3600.0, dtout_co=1000.0, interval=0, capacity=64,
                   ev_capacity=16, k0=0, off=False):
        """Field snapshots and event list decided on the device (main_LH.f90:1339-1407)."""
        if off:
            _check(lib().hbi_rk_set_output(self.r, None))
            self.out_caps = None
            return
        o = RkOutOpts(interval=int(interval), capacity=int(capacity), ev_capacity=int(ev_capacity), k0=int(k0),
                      velth=float(velth), dtout_inter=float(dtout_inter), dtout_co=float(dtout_co))
        _check(lib().hbi_rk_set_output(self.r, C.byref(o)))
        self.out_caps = (int(capacity), int(ev_capacity))

    def set_opening(self, hs2: HMatrix, hn2: HMatrix):
        """Opening (tensile dislocation matrices onto shear and normal traction;
        main_LH.f90:705-716, 2096-2143, 1279-1283).  Before set_output."""
        self._open = (hs2, hn2)
        _check(lib().hbi_rk_set_opening(self.r, hs2.h, hn2.h))

    def get_opening(self):
        """(veln, slipn) after the last accepted step, cluster order."""
        vn, sn = np.empty(self.n), np.empty(self.n)
        _check(lib().hbi_rk_get_opening(self.r, _p(vn), _p(sn)))
        return vn, sn

    def set_probes(self, idx):
        """Per-element local output (main_LH.f90:1099-1122, 1495-1499): after
        every accepted step of the following step calls, record V, slip, tau,
        sigma, mu, psi, veln, slipn at up to 16 cluster-order indices."""
        a = np.ascontiguousarray(idx, dtype=np.int32)
        self._nprobe = int(a.size)
        _check(lib().hbi_rk_set_probes(self.r, int(a.size), a.ctypes.data_as(_ip) if a.size else None))

    def get_probes(self) -> np.ndarray:
        """Probe rows of the last step call: (accepted steps, probes, 8)."""
        nr = C.c_longlong(0)
        _check(lib().hbi_rk_get_probes(self.r, C.byref(nr), None))
        out = np.empty((nr.value, getattr(self, "_nprobe", 0), 8))
        if out.size:
            _check(lib().hbi_rk_get_probes(self.r, C.byref(nr), _p(out)))
        return out

    def get_output(self) -> dict:
        """Drain the held output: snapshot steps, times and fields (vel, slip,
        sigma, tau, and veln, slipn with opening; cluster order), finished
        events and their coseismic slip."""
        cap, ecap = self.out_caps
        n = self.n
        nf = 6 if getattr(self, "_open", None) else 4
        ks = np.zeros(cap, dtype=np.int64)
        ts = np.zeros(cap)
        f = np.zeros((cap, nf, n))
        ev = (RkEvent * ecap)()
        eq = np.zeros((ecap, n))
        ns, ne = C.c_int(0), C.c_int(0)
        _check(lib().hbi_rk_get_output(self.r, C.byref(ns), ks.ctypes.data_as(C.POINTER(C.c_longlong)), _p(ts),
                                       _p(f), C.byref(ne), ev, _p(eq)))
        ns, ne = ns.value, ne.value
        events = [{k: getattr(ev[i], k) for k, _ in RkEvent._fields_ if k != "pad"} for i in range(ne)]
        out = {"k": ks[:ns].copy(), "t": ts[:ns].copy(), "vel": f[:ns, 0].copy(), "slip": f[:ns, 1].copy(),
               "sigma": f[:ns, 2].copy(), "tau": f[:ns, 3].copy(), "events": events, "eqslip": eq[:ne].copy()}
        if nf == 6:
            out["veln"], out["slipn"] = f[:ns, 4].copy(), f[:ns, 5].copy()
        return out

    @staticmethod
    def step_many(runs, nsteps):
        """Advance several runs (own matrices, own step control) nsteps accepted
        steps each, concurrently on the device; returns their info dicts."""
        arr = (C.c_void_p * len(runs))(*[r.r for r in runs])
        infos = (RkInfo * len(runs))()
        _check(lib().hbi_rk_step_many(arr, len(runs), int(nsteps), infos))
        return [i.as_dict() for i in infos]

    def set_controller(self, mode, mlp=None, safety=0.9, ratio_lo=0.125, ratio_hi=8.0):
        """mode 0: rkqs (the reference's controller); 1: learned proposals from mlp."""
        _check(lib().hbi_rk_set_controller(self.r, int(mode), mlp.m if mlp is not None else None, float(safety),
                                           float(ratio_lo), float(ratio_hi)))

    def set_trace(self, cap):
        _check(lib().hbi_rk_set_trace(self.r, int(cap)))
        self._trace_cap = int(cap)

    def get_trace(self):
        """(features (n, 12), raw (n, 2) = (h, errmax)) of the recorded accepted steps."""
        cap = getattr(self, "_trace_cap", 0)
        feat = np.zeros((max(cap, 1), 12))
        raw = np.zeros((max(cap, 1), 2))
        cnt = C.c_int(0)
        _check(lib().hbi_rk_get_trace(self.r, _p(feat), _p(raw), C.byref(cnt)))
        return feat[:cnt.value], raw[:cnt.value]


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


class MLP:
    """Step-size MLP (12 -> 128 -> 128 -> 1, GELU, BF16 weights) on the device."""

    N_IN = 12

    def __init__(self, seed=0):
        m = C.c_void_p()
        _check(lib().hbi_mlp_create(int(seed), C.byref(m)))
        self.m = m
        self.nparams = int(lib().hbi_mlp_nparams())

    def close(self):
        if getattr(self, "m", None):
            lib().hbi_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        self.close()

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        mu = np.empty(12, dtype=np.float32)
        sd = np.empty(12, dtype=np.float32)
        _check(lib().hbi_mlp_get_params(self.m, _fptr(p), _fptr(mu), _fptr(sd)))
        return p, mu, sd

    def set_params(self, params=None, mu=None, sd=None):
        cv = lambda a: np.ascontiguousarray(a, dtype=np.float32) if a is not None else None
        params, mu, sd = cv(params), cv(mu), cv(sd)
        _check(lib().hbi_mlp_set_params(self.m, _fptr(params), _fptr(mu), _fptr(sd)))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty(X.shape[0], dtype=np.float32)
        _check(lib().hbi_mlp_forward(self.m, _fptr(X), X.shape[0], _fptr(Y)))
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        L = C.c_double(0)
        _check(lib().hbi_mlp_loss_grad(self.m, _fptr(X), _fptr(T), X.shape[0], C.byref(L), _fptr(g)))
        return L.value, g

    def train(self, X, T, epochs=20, batch=1024, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(max(epochs, 1))
        _check(lib().hbi_mlp_train(self.m, _fptr(X), _fptr(T), X.shape[0], int(epochs), int(batch), float(lr),
                                   int(seed), _p(hist)))
        return hist[:epochs]


class Ensemble:
    """nreal <= 64 friction realisations on one fault, advanced in lockstep
    (one 64-column H product per RK stage, per-realisation step control)."""

    def __init__(self, hs: HMatrix, frs, lod, hdiv=0.0, eps_r=1e-6, prec=PREC_F64, backslip=True,
                 limitsigma=True, minsig=1.0, maxsig=300.0, dtmax=1e10, velmin=1e-20, velmax=1e7):
        frs = list(frs)
        if not 1 <= len(frs) <= 64:
            raise ValueError("1 <= number of realisations <= 64")
        law, vref = int(frs[0].law), float(frs[0].vref)
        if any(int(f.law) != law or float(f.vref) != vref for f in frs):
            raise ValueError("realisations must share the friction law and vref")
        lod = np.asarray(lod)
        self.hs = hs
        self.nreal = len(frs)
        self.n = hs.n
        self.keep = [np.ascontiguousarray(np.stack([np.asarray(getattr(f, k), dtype=np.float64)[lod] for f in frs]))
                     for k in ("a", "b", "dc", "f0", "vpl", "taudot", "sigdot")]
        o = RkOpts(law=law, backslip=int(backslip), limitsigma=int(limitsigma), prec=int(prec), vref=vref,
                   minsig=minsig, maxsig=maxsig, dtmax=dtmax, hdiv=float(hdiv), eps_r=eps_r, velmin=velmin,
                   velmax=velmax)
        e = C.c_void_p()
        _check(lib().hbi_ens_create(hs.h, C.byref(o), self.nreal, *[_p(a) for a in self.keep], C.byref(e)))
        self.e = e

    def close(self):
        if getattr(self, "e", None):
            lib().hbi_ens_destroy(self.e)
            self.e = None

    def __del__(self):
        self.close()

    def set_state(self, y, t=0.0, dt=1.0):
        """y: (nreal, 3N) interleaved (psi, tau, sigma) per realisation, cluster order."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        if y.shape != (self.nreal, 3 * self.n):
            raise ValueError("y must be (nreal, 3N)")
        _check(lib().hbi_ens_set_state(self.e, _p(y), float(t), float(dt)))

    def get_state(self):
        y = np.empty((self.nreal, 3 * self.n))
        slip = np.empty((self.nreal, self.n))
        vel = np.empty((self.nreal, self.n))
        t = np.empty(self.nreal)
        acc = np.empty(self.nreal, dtype=np.int64)
        rej = np.empty(self.nreal, dtype=np.int64)
        st = np.empty(self.nreal, dtype=np.int32)
        _check(lib().hbi_ens_get_state(self.e, _p(y), _p(slip), _p(vel), _p(t),
                                       acc.ctypes.data_as(C.POINTER(C.c_longlong)),
                                       rej.ctypes.data_as(C.POINTER(C.c_longlong)), _ipp(st)))
        return {"y": y, "slip": slip, "vel": vel, "t": t, "accepted": acc, "rejected": rej, "status": st}

    def set_clock(self, t, dt):
        """Per-realisation t and next step size (after set_state)."""
        t = np.ascontiguousarray(t, dtype=np.float64)
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        if t.shape != (self.nreal,) or dt.shape != (self.nreal,):
            raise ValueError("t and dt must have nreal entries")
        _check(lib().hbi_ens_set_clock(self.e, _p(t), _p(dt)))

    def get_clock(self):
        """(t, next step size) per realisation."""
        t = np.empty(self.nreal)
        dt = np.empty(self.nreal)
        _check(lib().hbi_ens_get_clock(self.e, _p(t), _p(dt)))
        return t, dt

    def set_probes(self, idx):
        """History rows of the following step calls also carry (V, slip) of up
        to 16 elements (cluster-order indices), per realisation."""
        a = np.ascontiguousarray(idx, dtype=np.int32)
        _check(lib().hbi_ens_set_probes(self.e, int(a.size), a.ctypes.data_as(_ip) if a.size else None))
        self._nprobe = int(a.size)

    def set_step_grid(self, h=None):
        """Replay step sizes h (nreal, cap): accepted step k of realisation r
        takes h[r, k], every attempt accepted; None returns to rkqs."""
        if h is None:
            _check(lib().hbi_ens_set_step_grid(self.e, 0, None))
            return
        h = np.ascontiguousarray(h, dtype=np.float64)
        if h.ndim != 2 or h.shape[0] != self.nreal:
            raise ValueError("h must be (nreal, cap)")
        _check(lib().hbi_ens_set_step_grid(self.e, h.shape[1], _p(h)))

    def step(self, nsteps, history=False):
        """Advance every running realisation nsteps accepted steps.  Returns the
        lockstep attempts, and with history=True also (nreal, nsteps, 5 + 2 p)
        rows of (t, dt, errmax, rejections before the step, log10 Vmax) and
        (V, slip) of the p probe elements."""
        att = C.c_longlong(0)
        hist = np.zeros((self.nreal, nsteps, 5 + 2 * getattr(self, "_nprobe", 0))) if history else None
        _check(lib().hbi_ens_step(self.e, int(nsteps), _p(hist), C.byref(att)))
        return (att.value, hist) if history else att.value

    def last_ms(self) -> float:
        ms = C.c_double(0)
        _check(lib().hbi_ens_last_ms(self.e, C.byref(ms)))
        return ms.value

    def set_controller(self, mode, mlp=None, safety=0.9, ratio_lo=0.125, ratio_hi=8.0):
        """mode 0: rkqs for every realisation; 1: learned proposals from mlp."""
        _check(lib().hbi_ens_set_controller(self.e, int(mode), mlp.m if mlp is not None else None, float(safety),
                                            float(ratio_lo), float(ratio_hi)))

    def set_trace(self, cap):
        _check(lib().hbi_ens_set_trace(self.e, int(cap)))
        self._trace_cap = int(cap)

    def get_trace(self):
        """list over realisations of (features (n, 12), raw (n, 2) = (h, errmax))"""
        cap = max(getattr(self, "_trace_cap", 0), 1)
        feat = np.zeros((self.nreal, cap, 12))
        raw = np.zeros((self.nreal, cap, 2))
        cnt = np.zeros(self.nreal, dtype=np.int32)
        _check(lib().hbi_ens_get_trace(self.e, _p(feat), _p(raw), _ipp(cnt)))
        return [(feat[r, :cnt[r]], raw[r, :cnt[r]]) for r in range(self.nreal)]


FV_ARRAYS = ("a", "b", "dc", "f0", "vpl", "kpmaxv", "kTv", "phi", "pfhyd", "qin", "area")


class FV:
    """Device fluid-driven cycle of main_fv.f90 (hbi_fv_*): 4-component
    rate-and-state RK45 on effective normal stress with permeability
    evolution, and the implicit pore-pressure step after every accepted step.
    ``opts``: FvOpts field values; ``arrays``: the FV_ARRAYS in cluster order."""

    def __init__(self, hs: HMatrix, opts: dict, arrays: dict, hn: Optional[HMatrix] = None):
        self.hs, self.hn = hs, hn
        self.n = hs.n
        self.keep = [_f64(arrays[k]) for k in FV_ARRAYS]
        for a in self.keep:
            if a.size != self.n:
                raise ValueError("per-element arrays must have N entries")
        o = FvOpts(**opts)
        f = C.c_void_p()
        _check(lib().hbi_fv_create(hs.h, hn.h if hn is not None else None, C.byref(o),
                                   *[_p(a) for a in self.keep], C.byref(f)))
        self.f = f

    def close(self):
        if getattr(self, "f", None):
            lib().hbi_fv_destroy(self.f)
            self.f = None

    def __del__(self):
        self.close()

    def set_state(self, y, pf, disp=None, t=0.0, dt=1.0):
        y, pf = _f64(y), _f64(pf)
        disp = _f64(disp) if disp is not None else None
        _check(lib().hbi_fv_set_state(self.f, _p(y), _p(pf), _p(disp), float(t), float(dt)))
        self._accepted = 0

    def get_state(self) -> dict:
        n = self.n
        out = {"y": np.empty(4 * n), "pf": np.empty(n), "disp": np.empty(n), "vel": np.empty(n),
               "qflow": np.empty(n)}
        info = FvInfo()
        _check(lib().hbi_fv_get_state(self.f, _p(out["y"]), _p(out["pf"]), _p(out["disp"]), _p(out["vel"]),
                                      _p(out["qflow"]), C.byref(info)))
        out["info"] = info.as_dict()
        return out

    def derivs(self, y):
        y = _f64(y)
        d = np.empty_like(y)
        _check(lib().hbi_fv_derivs(self.f, _p(y), _p(d)))
        return d

    def diffuse(self, pf, y, sigma, h, t, dtnxt):
        """One implicit pore-pressure step: (pfnew, CG iterations, dt bound)."""
        pf, y, sigma = _f64(pf), _f64(y), _f64(sigma)
        out = np.empty(self.n)
        it = C.c_int(0)
        dt = C.c_double(dtnxt)
        _check(lib().hbi_fv_diffuse(self.f, _p(pf), _p(y), _p(sigma), float(h), float(t), _p(out), C.byref(it),
                                    C.byref(dt)))
        return out, it.value, dt.value

    def step(self, nsteps, history=False):
        """nsteps accepted steps; history rows: t, log10 max|V|, mean disp,
        mean mu, max sigma_e, min sigma_e, argmax (1-based), errmax, dt, CG iterations."""
        hist = np.zeros((nsteps, 10)) if history else None
        info = FvInfo()
        _check(lib().hbi_fv_step(self.f, int(nsteps), _p(hist), C.byref(info)))
        d = info.as_dict()
        done = d["accepted"] - getattr(self, "_accepted", 0)
        self._accepted = d["accepted"]
        if history:
            hist = hist[:max(0, min(nsteps, done))]
        return d, hist

    def set_break(self, t=np.inf, vmax_above=np.inf, vmax_below=0.0):
        """Later step calls also end after the first accepted step with time > t,
        max |V| > vmax_above or max |V| < vmax_below."""
        _check(lib().hbi_fv_set_break(self.f, float(t), float(vmax_above), float(vmax_below)))

    def last_ms(self) -> float:
        ms = C.c_double(0)
        _check(lib().hbi_fv_last_ms(self.f, C.byref(ms)))
        return ms.value

    def set_probes(self, idx):
        """Local output (main_fv.f90:1312-1316) at up to 16 cluster-order indices."""
        a = np.ascontiguousarray(idx, dtype=np.int32)
        self._nprobe = int(a.size)
        _check(lib().hbi_fv_set_probes(self.f, int(a.size), a.ctypes.data_as(_ip) if a.size else None))

    def get_probes(self) -> np.ndarray:
        """Rows of the last step call: (steps, probes, 9): V, disp, sigma_e,
        tau, pf, sigma, mu, psi, kp."""
        nr = C.c_longlong(0)
        _check(lib().hbi_fv_get_probes(self.f, C.byref(nr), None))
        out = np.empty((nr.value, getattr(self, "_nprobe", 0), 9))
        if out.size:
            _check(lib().hbi_fv_get_probes(self.f, C.byref(nr), _p(out)))
        return out

    def set_wells(self, wells):
        """Grid mode: wells (i, j, rates, times) with 1-based grid indices and
        schedules of up to 50 entries (times in seconds); [] or None: the
        centre source."""
        wells = list(wells or [])
        nw = len(wells)
        S = 50
        iw = np.array([w[0] for w in wells], dtype=np.int32)
        jw = np.array([w[1] for w in wells], dtype=np.int32)
        kl = np.array([len(w[2]) for w in wells], dtype=np.int32)
        qv = np.zeros((max(nw, 1), S))
        qt = np.zeros((max(nw, 1), S))
        for k, w in enumerate(wells):
            if len(w[2]) != len(w[3]) or not 1 <= len(w[2]) <= S:
                raise ValueError("a well needs 1..50 rates with as many times")
            qv[k, :len(w[2])] = w[2]
            qt[k, :len(w[3])] = w[3]
        ptr = lambda a: a.ctypes.data_as(_ip) if nw else None
        _check(lib().hbi_fv_set_wells(self.f, nw, ptr(iw), ptr(jw), ptr(kl), _p(qv) if nw else None,
                                      _p(qt) if nw else None))

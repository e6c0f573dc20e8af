"""Synthetic fault problems and their geometry, in the reference's conventions.

Host-side logic only (numpy): mesh generators that follow the reference's own
generators, friction-parameter fields for the benchmark configurations, and
writers for the reference's ``.in`` / parameter-file / STL input formats.

Reference conventions followed here:
  * ``coordinate2dp``   main_LH.f90:1701-1718  (2dp / 2dvs)
  * ``coordinate3ddip`` main_LH.f90:1781-1835  (3dph)
  * STL facet reading   main_LH.f90:485-510 and ``evcalc`` main_LH.f90:1940-1990 (3dnt)
  * ``rake = crake/180*pi`` main_LH.f90:518-576
  * ``psi = a ln(2 V0/V sinh(tau/(a sigma)))`` main_LH.f90:1004
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

PI = 4.0 * math.atan(1.0)
RIGID = 32.04      # m_const.f90:6 (GPa)
VS = 3.464         # m_const.f90:7 (km/s)
POIS = 0.25        # m_const.f90:3
# default vref of main_LH.f90:136 is a single-precision literal stored in real(8)
VREF_DEFAULT = float(np.float32(1e-6))

KIND_2DVS = 1
KIND_OKADA_HS = 2     # 3dph / 3dhr
KIND_OKADA_FS = 3     # 3dnr
KIND_TRI_FS = 4       # 3dnt
KIND_TRI_HS = 5       # 3dht
KIND_2DP = 6          # 2dp: in-plane, full space
KIND_3DP = 7          # 3dp: planar quadrilaterals, full space
KIND_2DPH = 8         # 2dph, 2dnh: in-plane dipping segments, half-space
KIND_2DN = 9          # 2dn: in-plane nonplanar segments, full space
KIND_25D = 10         # 25d: 2.5-D rectangles
KIND_2DPFZ = 11       # 2dpfz: antiplane planar segments with ten image pairs at multiples of w
KIND_2DNV = 12        # 2dn_vector: in-plane nonplanar segments (end-point tractions), full space
KIND_2DNA = 13        # 2dna: antiplane nonplanar segments, full space

_GEO_FIELDS = ("xcol", "ycol", "zcol", "xel", "xer", "ang", "angd", "dsl", "dsd", "rake",
               "xs1", "xs2", "xs3", "ys1", "ys2", "ys3", "zs1", "zs2", "zs3",
               "ev11", "ev12", "ev13", "ev21", "ev22", "ev23", "ev31", "ev32", "ev33",
               "yel", "yer", "xs4", "zs4", "ds")


@dataclass
class Geometry:
    """Element geometry in mesh order (the reference's ``st_bemv``)."""
    kind: int
    n: int
    w: float = 0.0            # okada normalisation length (st_bemv%w = ds0)
    vcomp: int = 0            # 0 shear ('s'), 1 normal ('n')
    md: int = 0               # TD slip mode: 0 's', 1 'st', 2 'dp', 3 'open'
    ds: Optional[np.ndarray] = None    # element areas (or lengths in 2-D)
    ds0: float = 0.0
    arrays: Dict[str, np.ndarray] = field(default_factory=dict)

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays", {})
        if name in _GEO_FIELDS:
            return arrays.get(name)
        raise AttributeError(name)

    def field_names(self):
        return _GEO_FIELDS

    def geo_matrix(self) -> np.ndarray:
        """(n, 33) column-major table in the field order of ``_GEO_FIELDS``."""
        g = np.zeros((len(_GEO_FIELDS), self.n), dtype=np.float64)
        for k, f in enumerate(_GEO_FIELDS):
            a = self.arrays.get(f)
            if a is not None:
                g[k] = a
        return np.asfortranarray(g.T)


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def mesh_2dvs(n: int, ds0: float) -> Geometry:
    """2-D antiplane fault in a half-space (coordinate2dp)."""
    i = np.arange(1, n + 1, dtype=np.float64)
    xel = (i - 1) * ds0
    xer = i * ds0
    xcol = 0.5 * (xel + xer)
    z = np.zeros(n)
    g = Geometry(kind=KIND_2DVS, n=n, ds=np.full(n, ds0), ds0=ds0,
                 arrays=dict(xcol=_c(xcol), ycol=_c(z), zcol=_c(z.copy()), xel=_c(xel), xer=_c(xer)))
    return g


def mesh_2dp(n: int, ds0: float) -> Geometry:
    """2-D in-plane planar fault in a full space (coordinate2dp, main_LH.f90:1701-1718)."""
    g = mesh_2dvs(n, ds0)
    g.kind = KIND_2DP
    return g


def mesh_2dpfz(n: int, ds0: float, fwid: float) -> Geometry:
    """2-D antiplane planar fault of n segments (coordinate2dp) with the image
    pairs of matels2dpFZ_ij at multiples of ``fwid`` (st_bemv%w = fwid,
    main_LH.f90:606); the reference's mesh switch has no case for this kind."""
    g = mesh_2dvs(n, ds0)
    g.kind = KIND_2DPFZ
    g.w = float(fwid)
    return g


def mesh_3dp(imax: int, jmax: int, ds0: float) -> Geometry:
    """Planar fault of imax x jmax squares in a full space, in the x-z plane
    (coordinate3dp, main_LH.f90:1752-1779): corners 1..4 at (+,+), (-,+), (-,-), (+,-)."""
    n = imax * jmax
    i = np.repeat(np.arange(1, imax + 1), jmax).astype(np.float64)
    j = np.tile(np.arange(1, jmax + 1), imax).astype(np.float64)
    xcol = (i - imax // 2 - 0.5) * ds0
    zcol = -(j - jmax // 2 - 0.5) * ds0
    h = 0.5 * ds0
    arr = dict(xcol=xcol, ycol=np.zeros(n), zcol=zcol, xs1=xcol + h, xs2=xcol - h, xs3=xcol - h, xs4=xcol + h,
               zs1=zcol + h, zs2=zcol + h, zs3=zcol - h, zs4=zcol - h)
    return Geometry(kind=KIND_3DP, n=n, ds=np.full(n, ds0 * ds0), ds0=ds0, arrays={k: _c(v) for k, v in arr.items()})


def mesh_2dph(n: int, ds0: float, dipangle: float) -> Geometry:
    """2-D in-plane fault dipping from the free surface (y = depth) in a
    half-space (coordinate2dph, main_LH.f90:1720-1737)."""
    i = np.arange(1, n + 1, dtype=np.float64)
    c, s = math.cos(dipangle * PI / 180), math.sin(dipangle * PI / 180)
    return mesh_segments((i - 1) * ds0 * c, i * ds0 * c, (i - 1) * ds0 * s, i * ds0 * s, KIND_2DPH)


def mesh_segments(xel, xer, yel, yer, kind: int = KIND_2DN, w: float = 1e8) -> Geometry:
    """2-D segments from their end points (coordinate2dn, main_LH.f90:1739-1750):
    lengths ds, angles atan2, midpoints; kind 2dn, 2dnh (KIND_2DPH), 2dn_vector
    (KIND_2DNV), 2dna (KIND_2DNA) or 25d (element width ``w``, the reference's fwid)."""
    xel, xer, yel, yer = (_c(a) for a in (xel, xer, yel, yer))
    n = len(xel)
    ds = np.sqrt((xer - xel) ** 2 + (yer - yel) ** 2)
    ang = np.arctan2(yer - yel, xer - xel)
    arr = dict(xcol=0.5 * (xel + xer), ycol=0.5 * (yel + yer), zcol=np.zeros(n), xel=xel, xer=xer, yel=yel, yer=yer,
               ang=ang, ds=ds)
    return Geometry(kind=kind, n=n, w=float(w) if kind == KIND_25D else 0.0, ds=_c(ds), ds0=float(ds.min()),
                    arrays={k: _c(v) for k, v in arr.items()})


def read_segments(path: str) -> np.ndarray:
    """Geometry file of the 2-D problems: one `xel xer yel yer` line per element
    (main_LH.f90:443-452)."""
    return np.loadtxt(path, ndmin=2)[:, :4]


def mesh_rectangles(rows: np.ndarray, halfspace: bool = True, crake: float = 0.0) -> Geometry:
    """Rectangles from the 3dnr / 3dhr geometry file rows `xcol ycol zcol ang
    angd dsl dsd` (main_LH.f90:460-471): ds = dsl dsd, w = ds0 = min dsl."""
    rows = np.asarray(rows, dtype=np.float64)
    n = rows.shape[0]
    arr = dict(xcol=rows[:, 0], ycol=rows[:, 1], zcol=rows[:, 2], ang=rows[:, 3], angd=rows[:, 4], dsl=rows[:, 5],
               dsd=rows[:, 6], rake=np.full(n, crake / 180.0 * PI))
    ds0 = float(rows[:, 5].min())
    return Geometry(kind=KIND_OKADA_HS if halfspace else KIND_OKADA_FS, n=n, w=ds0, ds=_c(rows[:, 5] * rows[:, 6]),
                    ds0=ds0, arrays={k: _c(v) for k, v in arr.items()})


def mesh_3dph(imax: int, jmax: int, ds0: float, dipangle: float = 90.0, crake: float = 0.0,
              fullspace: bool = False) -> Geometry:
    """Planar rectangular fault in a half-space (coordinate3ddip)."""
    cdip = math.cos(dipangle * PI / 180.0)
    sdip = math.sin(dipangle * PI / 180.0)
    yr = [0.0]
    zr = [0.0]
    for j in range(1, jmax + 1):
        yr.append(yr[j - 1] - ds0 * cdip)
        zr.append(zr[j - 1] - ds0 * sdip)
    n = imax * jmax
    xcol = np.empty(n)
    ycol = np.empty(n)
    zcol = np.empty(n)
    angd = np.empty(n)
    ycol_j = [0.5 * (yr[j - 1] + yr[j]) for j in range(1, jmax + 1)]
    zcol_j = [0.5 * (zr[j - 1] + zr[j]) for j in range(1, jmax + 1)]
    angd_j = [math.atan2(zr[j - 1] - zr[j], yr[j - 1] - yr[j]) for j in range(1, jmax + 1)]
    k = 0
    for i in range(1, imax + 1):
        xc = (float(i - imax // 2) - 0.5) * ds0
        xcol[k:k + jmax] = xc
        ycol[k:k + jmax] = ycol_j
        zcol[k:k + jmax] = zcol_j
        angd[k:k + jmax] = angd_j
        k += jmax
    rake = np.full(n, crake / 180.0 * PI)
    g = Geometry(kind=KIND_OKADA_FS if fullspace else KIND_OKADA_HS, n=n, w=ds0, ds=np.full(n, ds0 * ds0),
                 ds0=ds0,
                 arrays=dict(xcol=_c(xcol), ycol=_c(ycol), zcol=_c(zcol), ang=np.zeros(n), angd=_c(angd),
                             dsl=np.full(n, ds0), dsd=np.full(n, ds0), rake=_c(rake)))
    return g


def evcalc(xs1, xs2, xs3, ys1, ys2, ys3, zs1, zs2, zs3):
    """Local frames and areas of triangles (evcalc, main_LH.f90:1940-1990)."""
    vba0, vba1, vba2 = xs2 - xs1, ys2 - ys1, zs2 - zs1
    vca0, vca1, vca2 = xs3 - xs1, ys3 - ys1, zs3 - zs1
    ev31 = vba1 * vca2 - vba2 * vca1
    ev32 = vba2 * vca0 - vba0 * vca2
    ev33 = vba0 * vca1 - vba1 * vca0
    rr = np.sqrt(ev31 * ev31 + ev32 * ev32 + ev33 * ev33)
    ev31, ev32, ev33 = ev31 / rr, ev32 / rr, ev33 / rr
    horiz = ~(np.abs(ev33) < 1.0)
    ev11 = -ev32
    ev12 = ev31.copy()
    ev13 = np.zeros_like(ev31)
    rr2 = np.sqrt(ev11 * ev11 + ev12 * ev12)
    with np.errstate(invalid="ignore", divide="ignore"):
        ev11 = np.where(horiz, 1.0, ev11 / rr2)
        ev12 = np.where(horiz, 0.0, ev12 / rr2)
    ev21 = ev32 * ev13 - ev33 * ev12
    ev22 = ev33 * ev11 - ev31 * ev13
    ev23 = ev31 * ev12 - ev32 * ev11
    t1 = vba0 * vba0 + vba1 * vba1 + vba2 * vba2
    t2 = vca0 * vca0 + vca1 * vca1 + vca2 * vca2
    t3 = vba0 * vca0 + vba1 * vca1 + vba2 * vca2
    ds = 0.5 * np.sqrt(t1 * t2 - t3 * t3)
    return dict(ev11=ev11, ev12=ev12, ev13=ev13, ev21=ev21, ev22=ev22, ev23=ev23,
                ev31=ev31, ev32=ev32, ev33=ev33), ds


def mesh_triangles(tri: np.ndarray, crake: float = 0.0, halfspace: bool = False) -> Geometry:
    """Triangle mesh from an (n, 3, 3) vertex array [element, vertex, xyz]:
    full space (3dnt) or, with ``halfspace``, below a free surface at z = 0 (3dht)."""
    tri = np.asarray(tri, dtype=np.float64)
    n = tri.shape[0]
    xs1, xs2, xs3 = tri[:, 0, 0].copy(), tri[:, 1, 0].copy(), tri[:, 2, 0].copy()
    ys1, ys2, ys3 = tri[:, 0, 1].copy(), tri[:, 1, 1].copy(), tri[:, 2, 1].copy()
    zs1, zs2, zs3 = tri[:, 0, 2].copy(), tri[:, 1, 2].copy(), tri[:, 2, 2].copy()
    xcol = (xs1 + xs2 + xs3) / 3
    ycol = (ys1 + ys2 + ys3) / 3
    zcol = (zs1 + zs2 + zs3) / 3
    ev, ds = evcalc(xs1, xs2, xs3, ys1, ys2, ys3, zs1, zs2, zs3)
    arr = dict(xcol=xcol, ycol=ycol, zcol=zcol, xs1=xs1, xs2=xs2, xs3=xs3, ys1=ys1, ys2=ys2, ys3=ys3,
               zs1=zs1, zs2=zs2, zs3=zs3, rake=np.full(n, crake / 180.0 * PI), **ev)
    arr = {k: _c(v) for k, v in arr.items()}
    return Geometry(kind=KIND_TRI_HS if halfspace else KIND_TRI_FS, n=n, ds=ds, ds0=float(np.sqrt(2 * ds).min()), arrays=arr)


def rough_surface_triangles(length: float, width: float, edge: float, hurst: float = 0.8,
                            alpha: float = 1e-2, seed: int = 0) -> np.ndarray:
    """Self-similar rough fault surface y = h(x, z) from a seeded 2-D FFT,
    meshed as right triangles on a regular (x, depth) grid of spacing ``edge``.
    Amplitude-to-wavelength ratio ``alpha`` is imposed on the rms height over
    the longest dimension.  Returns (n, 3, 3) vertices; the surface spans
    x in [-length/2, length/2], depth z in [-width, 0]."""
    nx = max(2, int(round(length / edge)))
    nz = max(2, int(round(width / edge)))
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx + 1, d=edge)
    kz = np.fft.fftfreq(nz + 1, d=edge)
    KX, KZ = np.meshgrid(kx, kz, indexing="ij")
    k = np.sqrt(KX ** 2 + KZ ** 2)
    k[0, 0] = np.inf
    amp = k ** (-(hurst + 1.0))
    phase = np.exp(2j * PI * rng.random(k.shape))
    noise = rng.standard_normal(k.shape)
    h = np.real(np.fft.ifft2(amp * noise * phase))
    h -= h.mean()
    lmax = max(length, width)
    rms = float(np.sqrt(np.mean(h ** 2)))
    if rms > 0:
        h *= alpha * lmax / rms
    xs = np.linspace(-0.5 * length, 0.5 * length, nx + 1)
    zs = np.linspace(0.0, -width, nz + 1)
    tris = []
    for i in range(nx):
        for j in range(nz):
            p00 = (xs[i], h[i, j], zs[j])
            p10 = (xs[i + 1], h[i + 1, j], zs[j])
            p01 = (xs[i], h[i, j + 1], zs[j + 1])
            p11 = (xs[i + 1], h[i + 1, j + 1], zs[j + 1])
            tris.append((p00, p01, p10))
            tris.append((p10, p01, p11))
    return np.asarray(tris, dtype=np.float64)


def plane_triangles(length: float, width: float, edge: float, dip: float = 90.0) -> np.ndarray:
    """Planar fault from the free surface down dip, meshed as right triangles
    on a regular (along-strike, down-dip) grid of spacing ``edge``: x in
    [-length/2, length/2], down-dip extent ``width`` at ``dip`` degrees
    (y = -s cos(dip), z = -s sin(dip) at down-dip distance s).  Returns
    (n, 3, 3) vertices."""
    nx = max(1, int(round(length / edge)))
    ns = max(1, int(round(width / edge)))
    xs = np.linspace(-0.5 * length, 0.5 * length, nx + 1)
    ss = np.linspace(0.0, width, ns + 1)
    cd, sd = np.cos(dip * PI / 180.0), np.sin(dip * PI / 180.0)
    tris = []
    for i in range(nx):
        for j in range(ns):
            p = {(a, b): (xs[i + a], -ss[j + b] * cd, -ss[j + b] * sd) for a in (0, 1) for b in (0, 1)}
            tris.append((p[0, 0], p[0, 1], p[1, 0]))
            tris.append((p[1, 0], p[0, 1], p[1, 1]))
    return np.asarray(tris, dtype=np.float64)


def write_stl(path: str, tri: np.ndarray) -> None:
    """ASCII STL in the 7-lines-per-facet layout the reference reads
    (main_LH.f90:491-509)."""
    with open(path, "w") as f:
        f.write("solid fault\n")
        for t in tri:
            f.write("facet normal 0 0 0\n outer loop\n")
            for v in t:
                f.write("  vertex %.17e %.17e %.17e\n" % (v[0], v[1], v[2]))
            f.write(" endloop\nendfacet\n")
        f.write("endsolid fault\n")


def read_stl(path: str) -> np.ndarray:
    """Read facets written in the 7-lines-per-facet ASCII layout."""
    verts = []
    with open(path) as f:
        for line in f:
            s = line.split()
            if s and s[0] == "vertex":
                verts.append([float(s[1]), float(s[2]), float(s[3])])
    return np.asarray(verts, dtype=np.float64).reshape(-1, 3, 3)


# ---------------------------------------------------------------------------
# friction fields for the benchmark configurations (mesh order)
# ---------------------------------------------------------------------------
@dataclass
class Friction:
    a: np.ndarray
    b: np.ndarray
    dc: np.ndarray
    f0: np.ndarray
    vpl: np.ndarray
    tau: np.ndarray
    sigma: np.ndarray
    vel: np.ndarray
    taudot: np.ndarray
    sigdot: np.ndarray
    law: int = 0           # 0 aging, 1 slip
    vref: float = 1e-6


def steady_tau(a, b, f0, sigma, v, vref):
    """tau such that psi = f0 - b ln(V/V0) at V: inverts main_LH.f90:1004."""
    psi_ss = f0 - b * np.log(v / vref)
    return a * sigma * np.arcsinh(v / (2 * vref) * np.exp(psi_ss / a))


def friction_c1(g: Geometry) -> Friction:
    """2-D antiplane: a = 0.010 above 15 km, ramp to a-b = +0.004 at 18 km;
    b = 0.015, Dc = 8 mm, sigma 50 MPa, Vpl 1e-9, aging law, V0 = 1e-6."""
    depth = g.xcol
    b = np.full(g.n, 0.015)
    a = np.where(depth < 15.0, 0.010,
                 np.where(depth < 18.0, 0.010 + (depth - 15.0) / 3.0 * 0.009, 0.019))
    return _friction(g, a, b, dc=np.full(g.n, 0.008), sigma=50.0, vpl=1e-9, vref=1e-6, law=0)


def friction_c3(g: Geometry) -> Friction:
    """Rough triangular fault: the a-b depth profile of the 2-D antiplane
    setup mapped onto the surface (a = 0.010 above 15 km, ramp to a - b =
    +0.004 at 18 km and below), b = 0.015, Dc = 8 mm, sigma 50 MPa, Vpl 1e-9,
    slip law, V0 = 1e-6."""
    depth = -g.zcol
    b = np.full(g.n, 0.015)
    a = np.where(depth < 15.0, 0.010,
                 np.where(depth < 18.0, 0.010 + (depth - 15.0) / 3.0 * 0.009, 0.019))
    return _friction(g, a, b, dc=np.full(g.n, 0.008), sigma=50.0, vpl=1e-9, vref=1e-6, law=1)


def friction_c2(g: Geometry, imax: int, jmax: int, ds0: float, seed: int = 1,
                patch=(60.0, 30.0), sigma_log: float = 0.1, dc0: float = 0.02, law: int = 0) -> Friction:
    """3-D planar fault: velocity-weakening patch (a-b = -0.004) centred along
    strike and in the upper part of the fault, a-b = +0.004 elsewhere, a = 0.01,
    Dc = 20 mm, sigma 50 MPa; a and Dc perturbed lognormally from a seed."""
    x = g.xcol
    depth = -g.zcol
    w_total = jmax * ds0
    top = 0.5 * (w_total - patch[1])
    in_patch = (np.abs(x) <= 0.5 * patch[0]) & (depth >= top) & (depth <= top + patch[1])
    a = np.full(g.n, 0.01)
    amb = np.where(in_patch, -0.004, 0.004)
    rng = np.random.default_rng(seed)
    za = rng.standard_normal(g.n)
    zd = rng.standard_normal(g.n)
    a = a * np.exp(sigma_log * za)
    b = a - amb
    dc = dc0 * np.exp(sigma_log * zd)
    return _friction(g, a, b, dc=dc, sigma=50.0, vpl=1e-9, vref=1e-6, law=law)


def friction_ensemble(g: Geometry, imax: int, jmax: int, ds0: float, nreal: int, seed: int = 7,
                      patch=(40.0, 20.0)) -> list:
    """Ensemble realisations for the 128x128 fault: a-b ~ N(-0.003, 0.001) in
    the patch, Dc ~ lognormal(median 15 mm, sigma_log 0.3)."""
    rng = np.random.default_rng(seed)
    x = g.xcol
    depth = -g.zcol
    w_total = jmax * ds0
    top = 0.5 * (w_total - patch[1])
    in_patch = (np.abs(x) <= 0.5 * patch[0]) & (depth >= top) & (depth <= top + patch[1])
    out = []
    for _ in range(nreal):
        amb_p = rng.normal(-0.003, 0.001)
        dc = 0.015 * np.exp(0.3 * rng.standard_normal(g.n))
        a = np.full(g.n, 0.01)
        b = a - np.where(in_patch, amb_p, 0.004)
        out.append(_friction(g, a, b, dc=dc, sigma=50.0, vpl=1e-9, vref=1e-6, law=0))
    return out


def _friction(g, a, b, dc, sigma, vpl, vref, law):
    n = g.n
    f0 = np.full(n, 0.6)
    sig = np.full(n, float(sigma))
    vel = np.full(n, float(vpl))
    tau = steady_tau(a, b, f0, sig, vel, vref)
    return Friction(a=_c(a), b=_c(b), dc=_c(dc), f0=f0, vpl=np.full(n, float(vpl)), tau=_c(tau), sigma=sig,
                    vel=vel, taudot=np.zeros(n), sigdot=np.zeros(n), law=law, vref=vref)


def perturb_state(y: np.ndarray, g: Geometry, lod: np.ndarray, amp: float = 0.02, depth: float = 8.0,
                  width: float = 2.0) -> np.ndarray:
    """Lower psi by a smooth bump (nucleation seed) so the first event does not
    grow from rounding noise; in place, cluster order."""
    d = g.xcol[lod] if g.kind == KIND_2DVS else -g.zcol[lod]
    y[0::3] -= amp * np.exp(-((d - depth) / width) ** 2)
    return y


def init_state(fr: Friction, lod: np.ndarray) -> np.ndarray:
    """Interleaved (psi, tau, sigma) state in cluster order (main_LH.f90:1004, 1208-1215)."""
    a = fr.a[lod]
    tau = fr.tau[lod]
    sig = fr.sigma[lod]
    vel = fr.vel[lod]
    psi = a * np.log(2 * fr.vref / vel * np.sinh(tau / sig / a))
    y = np.empty(3 * len(lod))
    y[0::3] = psi
    y[1::3] = tau
    y[2::3] = sig
    return y


# ---------------------------------------------------------------------------
# the reference's input formats
# ---------------------------------------------------------------------------
def write_param_file(path: str, fr: Friction, columns=("a", "b", "dc", "f0", "tau", "sigma", "vel", "vpl")) -> None:
    """Column-named parameter file read by main_LH.f90:530-581 (header row,
    then one row per element in mesh order)."""
    data = np.stack([getattr(fr, c) for c in columns], axis=1)
    with open(path, "w") as f:
        f.write(" ".join(columns) + "\n")
        for row in data:
            f.write(" ".join("%.17e" % v for v in row) + "\n")


def write_in_file(path: str, keys: Dict[str, object]) -> None:
    """``key value`` input file parsed by main_LH.f90:182-328; a list value
    repeats its key once per item (``outloc``, main_LH.f90:322-324)."""
    with open(path, "w") as f:
        for k, v in keys.items():
            for item in (v if isinstance(v, (list, tuple)) else [v]):
                if isinstance(item, bool):
                    item = "T" if item else "F"
                f.write("%s %s\n" % (k, item))


# ------------------------------------------------------------ fluid drivers --
YEAR_S = 365 * 24 * 3600


def fluid_injection(g: Geometry, lod, setting: str = "injectionq", a=0.007, b=0.005, dc=4e-3, f0=0.6, vref=1e-6,
                    vpl=1e-9, velinit=1e-9, sigmainit=50.0, pfinit=0.0, tauinit=0.0, muinit=0.58, phi0=0.1,
                    beta=1e-8, eta=1e-3, kpmin=1e-13, kpmax=1e-12, kL=1e8, kT=1e12, s0=30.0, q0=1.25e-6,
                    pinj=10.0, tinj=1.0, pbcl=0.0, pbcr=0.0, bcl="Neumann", bcr="Neumann", law=0,
                    pressure_dependent=False, pfconst=False, eps_r=1e-4, backslip=True, prec=0,
                    velmin=1e-16, velmax=1e7, tmax=1e12, outpertime=False, dtout=1e3, Bs=0.0):
    """Options, per-element arrays (cluster order) and the initial state
    (y, pf) of the injection settings of main_fv.f90: initcond_injectionq
    (1596-1643: a source q0 at element N/2, active while t < tinj years) and
    initcond_injectionp (1548-1594: pf = pinj at N/2, held there by the
    penalty of Beuler).  Uniform friction; the defaults of main_fv.f90:130-162
    for kL, kT, Bs, tmax, velmin, velmax, minsig; 2-D grid spacing ds0."""
    n = g.n
    lod = np.asarray(lod)
    one = np.ones(n)
    kind = {"injectionq": 0, "injectionp": 1, "steadystate": 2, "thrust": 3}.get(setting, 4)
    qin = np.zeros(n)
    if setting == "injectionq":
        qin[n // 2 - 1] = q0
    arrays = dict(a=a * one, b=b * one, dc=dc * one, f0=f0 * one, vpl=vpl * one, kpmaxv=kpmax * one, kTv=kT * one,
                  phi=phi0 * one, pfhyd=np.zeros(n), qin=qin, area=np.asarray(g.ds, dtype=np.float64)[lod])
    pf = pfinit * one
    if setting == "injectionp":
        pf[n // 2 - 1] = pinj
    sigma = sigmainit * one
    sige = sigma - pf
    tau = tauinit * one
    if muinit != 0.0 or setting == "injectionp":
        tau = sige * muinit
    vel = velinit * one
    psi = arrays["a"] * np.log(2 * vref / vel * np.sinh(tau / sige / arrays["a"]))
    y = np.empty(4 * n)
    y[0::4], y[1::4], y[2::4], y[3::4] = psi, tau, sige, kpmin * one
    opts = dict(law=law, backslip=int(backslip), prec=prec, pressure_dependent=int(pressure_dependent),
                pfconst=int(pfconst), bcl=0 if bcl == "Neumann" else 1, bcr=0 if bcr == "Neumann" else 1,
                setting=kind, outpertime=int(outpertime), vref=vref, eps_r=eps_r, hdiv=0.0, ds0=float(g.ds0),
                kL=kL, kpmin=kpmin, s0=s0, eta=eta, beta=beta, Bs=Bs, pbcl=pbcl, pbcr=pbcr, q0=q0, pinj=pinj,
                tinj_s=tinj * 365 * 24 * 3600, dtout_s=dtout * 365 * 24 * 3600, velmin=velmin, velmax=velmax,
                tmax_s=tmax * 365 * 24 * 3600)
    return opts, arrays, y, pf


def fluid_thrust(g: Geometry, lod, dipangle: float, model: str = "", a0=0.01, b0=0.015, dc=0.02, f0=0.6,
                 vref=1e-6, vpl=1e-9, velinit=1e-9, phi0=0.1, beta=1e-8, eta=1e-3, kpmin=1e-16, kpmax=1e-13,
                 kL=1e-3, s0=50.0, q0=1e-12, law=0, pressure_dependent=False, eps_r=1e-4, backslip=True, prec=0,
                 velmin=1e-16, velmax=1e7, tmax=1e12, Bs=0.0, ds0=None):
    """Options, per-element arrays (cluster order) and the initial state of
    main_fv's 'thrust' setting on a 2dph fault of at least 2000 elements
    (initcond_thrust, main_fv.f90:1378-1450): a ramps from a0 away from
    element 500 (700 for model B) and is capped at b0 + 0.002; sources q0 at
    element 1950 and 19 q0 at 1600 (and -19 q0 at 1200 for model D), their
    flux q(i) = the sum of the sources past element i; the healing time
    kT = min(exp(1e4 / (300 + 12 depth)), 1e12) and the steady permeability
    ks between kpmin and kpmax; lithostatic sigma and hydrostatic pfhyd
    (g = the reference's default-real 9.80), sigma_e integrated down dip with
    the flux's pressure gradient; Dirichlet pressures pbcl = pfhyd(1), pbcr =
    pf(N).  The element index is the position along the fault (mesh order).
    ds0: the element length of the pressure grid (default: the mesh's)."""
    n = g.n
    ds0 = float(g.ds0 if ds0 is None else ds0)
    assert n >= 2000, "thrust places sources at elements 1200-2000"
    lod = np.asarray(lod)
    grav = float(np.float32(9.80))
    ycol = np.asarray(g.ycol, dtype=np.float64)
    i = np.arange(1, n + 1)
    a = np.maximum(a0, a0 + (np.abs(i - 500) - 300) * 2e-5)
    if model == "B":
        a = np.maximum(a0, a0 + (np.abs(i - 700) - 500) * 2e-5)
    a = np.minimum(a, b0 + float(np.float32(0.002)))
    qin = np.zeros(n)
    qin[1949] = q0
    if model == "D":
        qin[1199] = -19 * q0
    qin[1599] = 19 * q0
    q = np.empty(n)
    qs = float(np.cumsum(qin)[-1])
    for k in range(n):
        q[k] = qs - qin[k]
        qs = q[k]
    kT = np.minimum(np.exp(1e4 / (300.0 + 12.0 * ycol)), 1e12)
    ks = (kpmax * vpl / kL + kpmin / kT) / (vpl / kL + 1 / kT)
    c02 = float(np.float32(0.2))
    sigma = grav * 2.6e3 * (ycol + c02) * 1e-3
    pfhyd = grav * 1.0e3 * (ycol + c02) * 1e-3
    pf = np.empty(n)
    sige = np.empty(n)
    pf[0] = pfhyd[0]
    sige[0] = sigma[0] - pf[0]
    sd = math.sin(dipangle * PI / 180.0) * (2.6e3 - 1.0e3) * grav * 1e-3
    for k in range(1, n):
        sige[k] = sige[k - 1] + ds0 * (sd - q[k] * eta / ks[k] * 1e-3 * math.exp(sige[k - 1] / s0))
        pf[k] = sigma[k] - sige[k]
    tau = sige * (f0 + (a - b0) * np.log(velinit / vref)) * float(np.float32(1.001))
    psi = a * np.log(2 * vref / velinit * np.sinh(tau / sige / a))
    one = np.ones(n)
    arrays = dict(a=a[lod], b=b0 * one, dc=dc * one, f0=f0 * one, vpl=vpl * one, kpmaxv=kpmax * one, kTv=kT[lod],
                  phi=phi0 * one, pfhyd=pfhyd[lod], qin=qin[lod], area=np.asarray(g.ds, dtype=np.float64)[lod])
    y = np.empty(4 * n)
    y[0::4], y[1::4], y[2::4], y[3::4] = psi[lod], tau[lod], sige[lod], ks[lod]
    opts = dict(law=law, backslip=int(backslip), prec=prec, pressure_dependent=int(pressure_dependent), pfconst=0,
                bcl=0, bcr=0, setting=3, outpertime=0, vref=vref, eps_r=eps_r, hdiv=0.0, ds0=ds0, kL=kL,
                kpmin=kpmin, s0=s0, eta=eta, beta=beta, Bs=Bs, pbcl=float(pfhyd[0]), pbcr=float(pf[-1]), q0=q0,
                pinj=0.0, tinj_s=0.0, dtout_s=1e3 * 365 * 24 * 3600, velmin=velmin, velmax=velmax,
                tmax_s=tmax * 365 * 24 * 3600)
    return opts, arrays, y, pf[lod]


BP6A_N, BP6A_DS0 = 4000, 0.01


def fluid_bp6a(g: Geometry, lod):
    """The reference's examples/bp6a.in for main_fv (SEAS BP6, aging law): a
    2dp fault of 4000 elements of 10 m, fluid injected at 1.25e-6 m/s into
    the centre for 0.27397 yr, constant permeability 1e-13, tau 29.2 MPa,
    sigma 50 MPa, V 1e-12 m/s; main_fv's defaults otherwise (injectionq,
    Neumann boundaries, eps_r 1e-4, kL 1e8, kT 1e12).  main_fv leaves vref
    undeclared and backslip unset: vref = 1e-6 (main_LH's) and no back slip;
    its kT default is the single-precision literal 1e12."""
    return fluid_injection(g, lod, setting="injectionq", a=0.007, b=0.005, dc=0.004, f0=0.6, velinit=1e-12,
                           sigmainit=50.0, tauinit=29.2, muinit=0.0, phi0=0.1, beta=1e-8, eta=1e-3, kpmin=1e-13,
                           kpmax=1e-13, q0=1.25e-6, tinj=0.27397, tmax=2.0, backslip=False, vpl=0.0, eps_r=1e-4,
                           kT=float(np.float32(1e12)))


def fluid3dp_injection(imax: int, jmax: int, ds0: float, a=0.013, b=0.0, dc=1e-4, f0=0.6, vref=1e-6, vpl=0.0,
                       velinit=1e-12, sigmainit=30.0, tauinit=0.0, muinit=0.55, phi0=0.2, beta=5e-8, eta=0.89e-3,
                       kpmin=1e-14, kpmax=3e-14, kL=1e-3, kT=1e15, s0=50.0, q0=3e-4, bc="Dirichlet", pbc=0.0,
                       law=0, pressure_dependent=False, pfconst=False, eps_r=1e-4, backslip=True, prec=0,
                       velmin=1e-16, velmax=1e7, tmax=5.0, dtmax=1e6, minsig=0.0, lod=None):
    """A main_fv3dp 'injection' run on a 3dp fault of imax x jmax squares of
    ds0 km (the reference's examples/3dinjection2.in: uniform a 0.013, b 0, Dc
    0.1 mm, mu 0.55 on sigma 30 MPa, V 1e-12 m/s, permeability 1e-14..3e-14,
    Dirichlet boundaries at pf 0, dtmax 1e6 s): geometry, options, per-element
    arrays (cluster order, area 1 so that the monitor means are plain means
    as main_fv3dp.f90:795-799) and the initial state (y, pf) of
    initcond_injection (main_fv3dp.f90:1020-1061).  lod: the matrix's cluster
    permutation (identity if None)."""
    g = mesh_3dp(imax, jmax, ds0)
    n = g.n
    lod = np.arange(n) if lod is None else np.asarray(lod)
    one = np.ones(n)
    arrays = dict(a=a * one, b=b * one, dc=dc * one, f0=f0 * one, vpl=vpl * one, kpmaxv=kpmax * one, kTv=kT * one,
                  phi=phi0 * one, pfhyd=np.zeros(n), qin=np.zeros(n), area=one.copy())
    pf = np.zeros(n)
    sige = sigmainit * one - pf
    tau = tauinit * one
    if muinit != 0.0:
        tau = sige * muinit
    vel = velinit * one
    psi = arrays["a"] * np.log(2 * vref / vel * np.sinh(tau / sige / arrays["a"]))
    y = np.empty(4 * n)
    y[0::4], y[1::4], y[2::4], y[3::4] = psi, tau, sige, kpmin * one
    bcv = 0 if bc == "Neumann" else 1
    opts = dict(law=law, backslip=int(backslip), prec=prec, pressure_dependent=int(pressure_dependent),
                pfconst=int(pfconst), bcl=bcv, bcr=bcv, setting=0, outpertime=0, vref=vref, eps_r=eps_r, hdiv=0.0,
                ds0=float(ds0), kL=kL, kpmin=kpmin, s0=s0, eta=eta, beta=beta, Bs=0.0, pbcl=pbc, pbcr=pbc, q0=q0,
                pinj=0.0, tinj_s=0.0, dtout_s=0.0, velmin=velmin, velmax=velmax, tmax_s=tmax * 365 * 24 * 3600,
                imax=imax, jmax=jmax, bct=bcv, bcb=bcv, dtmax=dtmax, minsig=minsig, phi0=phi0)
    return g, opts, arrays, y, pf


# examples/wells.dat of the reference (input_well, main_fv3dp.f90:1101-1118): three wells on the
# 100 x 100 grid of examples/3dinjection2.in with 8-entry rate schedules (m/s) at times (s); grid
# indices 1-based
_T8 = [0.0, 1e7, 1e7, 2e7, 2e7, 3e7, 3e7, 99e7]
WELLS_3DINJ2 = [
    (25, 50, [3e-4, 3e-4, 6e-4, 6e-4, 3e-4, 3e-4, 0.0, 0.0], _T8),
    (50, 50, [3e-4, 3e-4, 0.0, 0.0, 3e-4, 3e-4, 0.0, 0.0], _T8),
    (75, 50, [3e-4, 3e-4, 0.0, 0.0, 3e-4, 3e-4, 0.0, 0.0], _T8),
]

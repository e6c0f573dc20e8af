"""Write golden kernel vectors from the reference's own entry function.

Runs HACApK_entry_ij of the reference, compiled unmodified from its sources
into oracle/_ref/libhbiref.so by oracle/Makefile (``make -C oracle ref``), on
sampled (receiver, source) pairs of the three benchmark geometries and stores
geometry + pairs + values as compressed .npz fixtures next to this script.

    python tests/golden/make_golden.py                # every fixture
    python tests/golden/make_golden.py --more-kinds   # only 2dpfz, 2dn_vector, 2dna
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hbi_amd import problems as P  # noqa: E402
from oracle import oracle_lib as O  # noqa: E402


def sample_pairs(n, rng, nrand, near=None):
    ii = list(rng.integers(0, n, nrand))
    jj = list(rng.integers(0, n, nrand))
    if near is not None:
        for i, j in near:
            ii.append(i)
            jj.append(j)
    return np.asarray(ii, dtype=np.int32), np.asarray(jj, dtype=np.int32)


def save(name, g, ii, jj, vals, **extra):
    arrays = {("geo_" + k): v for k, v in g.arrays.items()}
    np.savez_compressed(os.path.join(HERE, name), kind=g.kind, n=g.n, w=g.w, vcomp=g.vcomp, md=g.md, ii=ii, jj=jj,
                        vals=vals, **arrays, **extra)
    print(name, g.n, len(ii), "nan:", int(np.isnan(vals).sum()))


def main():
    O.build()
    assert O.ref_available(), "oracle/_ref/libhbiref.so missing (needs /root/reference)"
    rng = np.random.default_rng(20251203)
    # C1: 2dvs, N=2048 over 40 km: all self terms, neighbours and random pairs
    g = P.mesh_2dvs(2048, 40.0 / 2048)
    near = [(i, i) for i in range(0, 2048, 7)] + [(i, i + 1) for i in range(0, 2047, 13)] + \
           [(i + 1, i) for i in range(0, 2047, 17)]
    ii, jj = sample_pairs(g.n, rng, 3000, near)
    save("kernel_entries_2dvs.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    # C2: 3dph 256 x 128, ds 0.390625 km, vertical: self, in-plane neighbours, surface row, random
    imax, jmax, ds0 = 256, 128, 0.390625
    g = P.mesh_3dph(imax, jmax, ds0, 90.0)
    near = []
    for i in rng.integers(0, g.n, 300):
        i = int(i)
        for dj in (0, 1, -1, jmax, -jmax, jmax + 1, 2, 2 * jmax):
            j = i + dj
            if 0 <= j < g.n:
                near.append((i, j))
    near += [(k * jmax, m * jmax) for k in range(0, imax, 17) for m in range(0, imax, 23)]   # surface row
    ii, jj = sample_pairs(g.n, rng, 4000, near)
    save("kernel_entries_3dph.npz", g, ii, jj, O.ref_entries(g, ii, jj), imax=imax, jmax=jmax, ds0=ds0)
    # the same mesh, dip 60 degrees, and normal traction, small
    g = P.mesh_3dph(32, 16, 1.0, 60.0, crake=30.0)
    g.vcomp = 1
    ii, jj = sample_pairs(g.n, rng, 1500, [(i, i) for i in range(0, g.n, 5)])
    save("kernel_entries_3dph_dip60_n.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    # C3: rough triangular surface (Hurst 0.8, 1e-2), small patch
    tri = P.rough_surface_triangles(6.0, 3.0, 0.25, hurst=0.8, alpha=1e-2, seed=11)
    g = P.mesh_triangles(tri, crake=0.0)
    ii, jj = sample_pairs(g.n, rng, 3000, [(i, i) for i in range(0, g.n, 3)] + [(i, i + 1) for i in range(0, g.n - 1, 5)])
    save("kernel_entries_3dnt.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    g.vcomp = 1
    ii, jj = sample_pairs(g.n, rng, 1000)
    save("kernel_entries_3dnt_n.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    # TD slip modes 'st', 'dp', 'open' on the same patch
    for md, name in ((1, "st"), (2, "dp"), (3, "open")):
        g.vcomp = 0
        g.md = md
        ii, jj = sample_pairs(g.n, rng, 500)
        save("kernel_entries_3dnt_md%s.npz" % name, g, ii, jj, O.ref_entries(g, ii, jj))
    # 3dht: the rough surface below a free surface at z = 0 (top edge on it),
    # rake 30 degrees (the source's rake sets the slip there), all modes
    tri = P.rough_surface_triangles(6.0, 3.0, 0.25, hurst=0.8, alpha=1e-2, seed=12)
    g = P.mesh_triangles(tri, crake=30.0, halfspace=True)
    near = [(i, i) for i in range(0, g.n, 3)] + [(i, i + 1) for i in range(0, g.n - 1, 5)]
    ii, jj = sample_pairs(g.n, rng, 2000, near)
    save("kernel_entries_3dht.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    g.vcomp = 1
    ii, jj = sample_pairs(g.n, rng, 800, near[:100])
    save("kernel_entries_3dht_n.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    for md, name in ((1, "st"), (2, "dp"), (3, "open")):
        g.vcomp = 0
        g.md = md
        ii, jj = sample_pairs(g.n, rng, 400)
        save("kernel_entries_3dht_md%s.npz" % name, g, ii, jj, O.ref_entries(g, ii, jj))
    # a planar thrust dipping 30 degrees from the surface: shallow sides in
    # both artefact-free configurations, receivers close to the free surface
    g = P.mesh_triangles(P.plane_triangles(8.0, 4.0, 0.5, dip=30.0), crake=90.0, halfspace=True)
    ii, jj = sample_pairs(g.n, rng, 1500, [(i, i) for i in range(g.n)])
    save("kernel_entries_3dht_dip30.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    g.vcomp = 1
    ii, jj = sample_pairs(g.n, rng, 800)
    save("kernel_entries_3dht_dip30_n.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    # free-surface correction strains of single angular dislocations
    # (AngDisStrainFSC): points below the surface, both signs of beta, each
    # Burgers component, depths of the vertex a
    m = 400
    args = np.empty((m, 9))
    args[:, 0:2] = rng.uniform(-3, 3, (m, 2))
    args[:, 2] = -rng.uniform(0.0, 3.0, m)
    args[:, 3] = rng.uniform(0.05, np.pi - 0.05, m) * rng.choice([-1.0, 1.0], m)
    args[:, 4:7] = np.eye(3)[rng.integers(0, 3, m)]
    args[:, 7] = 0.25
    args[:, 8] = rng.uniform(0.05, 3.0, m)
    vals = np.array([O.ref_angfsc(*a) for a in args])
    np.savez_compressed(os.path.join(HERE, "angfsc_points.npz"), args=args, vals=vals)
    print("angfsc_points.npz", m)
    # the other kernels, on the configurations of the reference's examples
    # (examples/2dp.in: 1000 x 0.1 km; bp3dip30thrust.in: 1600 x 0.025 km at
    # 30 degrees; 3dp squares) and on curved 2-D faults
    def pairs(g, nrand):
        near = [(i, i) for i in range(0, g.n, 4)] + [(i, i + 1) for i in range(0, g.n - 1, 9)]
        return sample_pairs(g.n, rng, nrand, near)

    def curve(n, length, amp, y0):
        x = np.linspace(0.0, length, n + 1)
        y = y0 + amp * np.sin(6 * np.pi * x / length) + 0.3 * amp * np.sin(22 * np.pi * x / length)
        return x[:-1], x[1:], y[:-1], y[1:]

    g = P.mesh_2dp(1000, 0.1)
    ii, jj = pairs(g, 1500)
    save("kernel_entries_2dp.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    g = P.mesh_3dp(32, 16, 0.5)
    ii, jj = pairs(g, 1500)
    save("kernel_entries_3dp.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    g.md, g.vcomp = 3, 1
    ii, jj = pairs(g, 600)
    save("kernel_entries_3dp_mdopen_n.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    g = P.mesh_2dph(1600, 0.025, 30.0)
    ii, jj = pairs(g, 1500)
    save("kernel_entries_2dph.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    g.vcomp = 1
    ii, jj = pairs(g, 600)
    save("kernel_entries_2dph_n.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    for kind, tag, y0, extra in ((P.KIND_2DN, "2dn", 0.0, {}), (P.KIND_2DPH, "2dnh", 2.0, {}),
                                 (P.KIND_25D, "25d", 0.0, {"w": 10.0})):
        g = P.mesh_segments(*curve(600, 30.0, 0.5, y0), kind=kind, **extra)
        ii, jj = pairs(g, 1000)
        save("kernel_entries_%s.npz" % tag, g, ii, jj, O.ref_entries(g, ii, jj))
        g.vcomp = 1
        ii, jj = pairs(g, 500)
        save("kernel_entries_%s_n.npz" % tag, g, ii, jj, O.ref_entries(g, ii, jj))
    g = P.mesh_segments(*curve(600, 30.0, 0.5, 0.0), kind=P.KIND_2DN)
    g.md, g.vcomp = 3, 1
    ii, jj = pairs(g, 500)
    save("kernel_entries_2dn_mdopen_n.npz", g, ii, jj, O.ref_entries(g, ii, jj))
    more_kinds()


def more_kinds():
    """The three kinds of HACApK_entry_ij that main_LH's mesh switch never
    selects: 2dpfz (image pairs at multiples of fwid), 2dn_vector and 2dna,
    on a straight fault and the curved fault above (own seed, so the files
    above do not change when this part is re-run alone)."""
    rng = np.random.default_rng(20261204)

    def pairs(g, nrand):
        near = [(i, i) for i in range(0, g.n, 4)] + [(i, i + 1) for i in range(0, g.n - 1, 9)] + \
               [(i + 1, i) for i in range(0, g.n - 1, 11)]
        return sample_pairs(g.n, rng, nrand, near)

    def curve(n, length, amp, y0):
        x = np.linspace(0.0, length, n + 1)
        y = y0 + amp * np.sin(6 * np.pi * x / length) + 0.3 * amp * np.sin(22 * np.pi * x / length)
        return x[:-1], x[1:], y[:-1], y[1:]

    # 2dpfz: 1000 x 0.1 km, image spacing 0.5 km and 3 km
    for fwid, tag in ((0.5, "2dpfz"), (3.0, "2dpfz_w3")):
        g = P.mesh_2dpfz(1000, 0.1, fwid)
        ii, jj = pairs(g, 1500)
        save("kernel_entries_%s.npz" % tag, g, ii, jj, O.ref_entries(g, ii, jj))
    for kind, tag in ((P.KIND_2DNV, "2dnv"), (P.KIND_2DNA, "2dna")):
        g = P.mesh_segments(*curve(600, 30.0, 0.5, 0.0), kind=kind)
        ii, jj = pairs(g, 1000)
        save("kernel_entries_%s.npz" % tag, g, ii, jj, O.ref_entries(g, ii, jj))
        if kind == P.KIND_2DNV:
            g.vcomp = 1
            ii, jj = pairs(g, 600)
            save("kernel_entries_%s_n.npz" % tag, g, ii, jj, O.ref_entries(g, ii, jj))


if __name__ == "__main__":
    if sys.argv[1:] == ["--more-kinds"]:
        O.build()
        assert O.ref_available(), "oracle/_ref/libhbiref.so missing (needs /root/reference)"
        more_kinds()
    else:
        main()

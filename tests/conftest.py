import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhbi.so on the device)")


def load_golden(name):
    """Geometry + pairs + reference values of a golden kernel fixture."""
    from hbi_amd.problems import Geometry
    d = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    arrays = {k[4:]: np.ascontiguousarray(d[k]) for k in d.files if k.startswith("geo_")}
    g = Geometry(kind=int(d["kind"]), n=int(d["n"]), w=float(d["w"]), vcomp=int(d["vcomp"]), md=int(d["md"]),
                 arrays=arrays)
    return g, d["ii"], d["jj"], d["vals"]


def halfspace_tolerance(g, jj, vals):
    """Per-pair absolute tolerance of a half-space triangle entry (3dht).

    The free-surface correction of a triangle side that dips at beta from the
    vertical grows like cot^2(beta) at each end of the side, and the two ends
    cancel to the O(1) result (TDstressFS.f90:405-558, 629-1011): any FP64
    evaluation, the reference's included, then carries an absolute error near
    2 mu * 1e-12 * cot^2(beta) (checked against 60-digit evaluations of the
    same entries).  Sides steeper than 1e-3 rad contribute nothing."""
    A = g.arrays
    P = np.stack([np.stack([A["xs%d" % k][jj], A["ys%d" % k][jj], A["zs%d" % k][jj]], -1) for k in (1, 2, 3)], 1)
    c2 = np.zeros(len(jj))
    for a, b in ((0, 1), (1, 2), (2, 0)):
        sv = P[:, b] - P[:, a]
        beta = np.arccos(-sv[:, 2] / np.linalg.norm(sv, axis=1))
        beta = np.minimum(beta, np.pi - beta)
        c2 = np.maximum(c2, np.where(beta < 1e-3, 0.0, 1.0 / np.tan(beta) ** 2))
    return 1e-10 * np.nanmax(np.abs(vals)) + 1e-11 * 2 * 32.04 * (1.0 + c2)


GOLDEN_FILES = sorted(f for f in os.listdir(GOLDEN) if f.startswith("kernel_entries_") and f.endswith(".npz"))

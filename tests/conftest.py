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


GOLDEN_FILES = sorted(f for f in os.listdir(GOLDEN) if f.startswith("kernel_entries_") and f.endswith(".npz"))

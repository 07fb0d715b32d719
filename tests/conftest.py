import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "bvh-based-collision-detection-scheme_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    """Import the product package from its (hyphenated) directory."""
    if "bvhcd_amd" in sys.modules:
        return sys.modules["bvhcd_amd"]
    spec = importlib.util.spec_from_file_location("bvhcd_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bvhcd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def bvhcd():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc  # oracle/oracle.py
    orc.lib()
    return orc


def sheet_sphere(pkg, n=24, frames=20, subdiv=2, radius=0.15, seed=7):
    """The batched workload's scene (folding sheet + rigid icosphere on a straight path through
    it) split into its deformable part (the sheet) and its rigid part (the sphere)."""
    spec = pkg.SceneSpec(pkg.SCENE_SHEET_SPHERE, seed=seed, n=n, frames=frames, subdiv=subdiv, radius=radius,
                         amplitude=0.15, compress=0.5, fold=3.7, fold_radius=0.05)
    sc = pkg.Scene(spec)
    f = sc.faces()
    nv, nt = n * n, 2 * (n - 1) ** 2
    return sc, f[:nt].copy(), (f[nt:] - nv).astype(np.int32), nv


def check_cut(fi, fe, bvh, nq):
    """Every query's front covers the deformable leaves [0, n) exactly once."""
    n = len(bvh["leaf_par"])
    lo = np.concatenate([bvh["int_rx"][fi[:, 1]], fe[:, 1]])
    hi = np.concatenate([bvh["int_ry"][fi[:, 1]], fe[:, 1]])
    q = np.concatenate([fi[:, 0], fe[:, 0]])
    order = np.lexsort((lo, q))
    q, lo, hi = q[order], lo[order], hi[order]
    assert np.array_equal(np.unique(q), np.arange(nq))
    starts = np.searchsorted(q, np.arange(nq))
    ends = np.append(starts[1:], q.size)
    assert np.all(lo[starts] == 0) and np.all(hi[ends - 1] == n - 1)
    inner = np.ones(q.size, bool)
    inner[ends - 1] = False
    assert np.all(lo[1:][inner[:-1]] == hi[:-1][inner[:-1]] + 1)

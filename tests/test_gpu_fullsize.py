"""Full-size pair parity on the 1024^2 folding cloth (2,093,058 triangles) and the 4-layer stress
scene (8,372,232 triangles).

The plain and the learned schedule run over a stretch of each sequence; on three folded frames
each, both schedules' narrow-phase pair sets must equal the oracle's BUILD + pure traversal +
narrow phase on the same vertices (the pair set does not depend on the schedule), and the
broad-phase counts must agree.
"""
import os
import time

import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def _weights(pkg):
    z = np.load(os.path.join(os.path.dirname(pkg.__file__), "mlp_scheduler.npz"), allow_pickle=False)
    return z["params"], float(z["threshold"])


@pytest.mark.parametrize("name,first,checks", [
    ("cloth1024", 366, (375, 387, 399)),
    ("layers4", 174, (180, 190, 199)),
])
def test_fullsize_pairs_match_oracle(pkg, oracle, name, first, checks):
    t0 = time.time()
    print(f"[gpu] full size {name} frames {first}..{checks[-1]}", flush=True)
    sc = pkg.Scene.preset(name)
    f = sc.faces()
    params, thr = _weights(pkg)
    plain = pkg.CollisionContext(sc.num_tris, sc.num_verts, scheduler=pkg.SCHED_PLAIN)
    learned = pkg.CollisionContext(sc.num_tris, sc.num_verts, scheduler=pkg.SCHED_LEARNED)
    learned.mlp_set_weights(params)
    learned.mlp_set_threshold(thr)
    for c in (plain, learned):
        c.set_topology(f, sc.num_verts)
    orc = oracle.OracleContext(f, sc.num_verts)
    for fr in range(first, checks[-1] + 1):
        v = sc.frame(fr)
        for c in (plain, learned):
            c.update_vertices(v)
            c.step_scheduled()
        if fr not in checks:
            continue
        orc.update_vertices(v)
        orc.step(0, 0)  # BUILD + pure traversal + narrow phase
        want = oracle.pair_set(orc.pairs(True))
        bp = orc.stats()["bp_pairs"]
        for tag, c in (("plain", plain), ("learned", learned)):
            got = oracle.pair_set(c.pairs(True))
            np.testing.assert_array_equal(got, want, err_msg=f"{name} frame {fr} {tag}")
            assert c.stats()["bp_pairs"] == bp, (name, fr, tag)
        print(f"[gpu]   frame {fr}: {want.size} pairs equal ({time.time() - t0:.0f}s)", flush=True)
    for c in (plain, learned):
        c.close()
    orc.close()

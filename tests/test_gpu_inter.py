"""GPU parity of the inter-object front (a rigid icosphere's triangles against a folding sheet's
tree) through the C ABI.

Every frame, the pair sets (before and after the narrow phase) and the fronts (as multisets of
(query, node) entries) must equal the oracle's restatement under the same operations, while the
sheet's own context runs its plain schedule (so its tree is rebuilt every 12 frames and the
inter front must follow); the rigid object's Morton order must equal the oracle's.
"""
import numpy as np
import pytest

from conftest import check_cut, load_package, sheet_sphere

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def _rows(a):
    a = np.asarray(a).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


@pytest.mark.parametrize("n,subdiv,frames", [(24, 2, 20), (96, 3, 40)])
def test_inter_front_matches_oracle(pkg, oracle, n, subdiv, frames):
    print(f"[gpu] inter front n={n} subdiv={subdiv}", flush=True)
    sc, fa, fb, nv = sheet_sphere(pkg, n=n, frames=frames, subdiv=subdiv)
    nvb = sc.num_verts - nv
    ctx = pkg.CollisionContext(len(fa), nv)
    ctx.set_topology(fa, nv)
    inter = pkg.InterFront(ctx, len(fb), nvb)
    inter.set_topology(fb, nvb)
    dref = oracle.OracleContext(fa, nv)
    iref = oracle.OracleInter(dref, fb, nvb)
    ops = [pkg.FRONT_GENERATE, pkg.FRONT_KEEP, pkg.FRONT_KEEP, pkg.FRONT_UPDATE, pkg.FRONT_KEEP,
           pkg.FRONT_REORDER, pkg.FRONT_PURE_BVH_CD, pkg.FRONT_KEEP]
    hits = 0
    for fr in range(frames):
        v = sc.frame(fr)
        va, vb = v[:nv], v[nv:]
        ctx.update_vertices(va)
        ctx.step_scheduled()
        dref.update_vertices(va)
        dref.step_scheduled()
        rop = pkg.BVH_BUILD if fr == 0 else pkg.BVH_REFIT
        fop = ops[fr % len(ops)]
        inter.update_vertices(vb)
        inter.step(rop, fop)
        iref.update_vertices(vb)
        iref.step(rop, fop)
        assert inter.stats()["last_front_op"] == iref.last_front_op, fr
        for narrow in (False, True):
            np.testing.assert_array_equal(oracle.ordered_pair_set(inter.pairs(narrow)),
                                          oracle.ordered_pair_set(iref.pairs(narrow)),
                                          err_msg=f"frame {fr} narrow {narrow}")
        if iref.last_front_op != pkg.FRONT_PURE_BVH_CD:
            for w in (0, 1):
                np.testing.assert_array_equal(_rows(inter.export_front(w)), _rows(iref.front(w)),
                                              err_msg=f"front {w} frame {fr}")
            check_cut(inter.export_front(0), inter.export_front(1), ctx.export_bvh(), len(fb))
        hits += len(inter.pairs(True))
    codes, idx = inter.export_rigid()
    rc, ri = iref.rigid_order()
    np.testing.assert_array_equal(codes, rc)
    np.testing.assert_array_equal(idx, ri)
    assert hits > 0
    inter.close()
    ctx.close()


def test_inter_front_full_cloth(pkg, oracle):
    """A 20,480-triangle icosphere through the 1024^2 folding cloth (2,093,058 triangles): the
    inter pairs of a few folded frames equal the oracle's."""
    print("[gpu] inter front on the 1024^2 cloth", flush=True)
    cloth = pkg.Scene.preset("cloth1024")
    fa = cloth.faces()
    sph = pkg.Scene.preset("spheres", subdiv=5, radius=0.12, spacing=0.0, speed=0.0, sigma=0.0)
    fb = sph.faces()[: len(sph.faces()) // 2].copy()
    nvb = int(fb.max()) + 1
    ctx = pkg.CollisionContext(len(fa), cloth.num_verts, scheduler=pkg.SCHED_PLAIN)
    ctx.set_topology(fa, cloth.num_verts)
    inter = pkg.InterFront(ctx, len(fb), nvb)
    inter.set_topology(fb, nvb)
    dref = oracle.OracleContext(fa, cloth.num_verts)
    iref = oracle.OracleInter(dref, fb, nvb)
    base = sph.frame(0)[:nvb]
    for k, fr in enumerate((380, 381, 382, 383)):
        va = cloth.frame(fr)
        # the sphere sits on a sheet vertex and sinks through it
        anchor = va[300 * 1024 + 700].astype(np.float64) + np.array([0.0, 0.0, 0.06 - 0.04 * k])
        vb = (base + anchor).astype(np.float32)
        ctx.update_vertices(va)
        ctx.step_scheduled()
        dref.update_vertices(va)
        dref.step_scheduled()
        op = (pkg.FRONT_GENERATE, pkg.FRONT_KEEP, pkg.FRONT_UPDATE, pkg.FRONT_KEEP)[k]
        inter.update_vertices(vb)
        inter.step(pkg.BVH_BUILD if k == 0 else pkg.BVH_REFIT, op)
        iref.update_vertices(vb)
        iref.step(0 if k == 0 else 1, op)
        for narrow in (False, True):
            np.testing.assert_array_equal(oracle.ordered_pair_set(inter.pairs(narrow)),
                                          oracle.ordered_pair_set(iref.pairs(narrow)),
                                          err_msg=f"frame {fr} narrow {narrow}")
    assert len(inter.pairs(False)) > 0
    inter.close()
    ctx.close()

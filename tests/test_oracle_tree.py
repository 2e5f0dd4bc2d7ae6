"""Pins the oracle's octree, neighbour search and Barnes-Hut walk.

Oracle functions pinned here (oracle/sphref.cpp):
  tree_make/tree_assign/tree_create (src/bhtree.cpp:55-107, 136-204)
    - every particle lies in the cube of each node that holds it; children
      partition their parent's particles; leaves hold <= leafParticleNumber
      particles unless at the deepest level; mass and centre of mass are
      the direct sums over the node's particles; the root cube is the
      bounding box (centre, largest extent);
    - a pool of fewer nodes than needed fails with SPHGPU_E_TREE.
  tree_search/neighbor_search (src/bhtree.cpp:114-126, 234-270)
    - the gather and the symmetric sets equal brute force
      (exhaustive_search, src/exhaustive_search.cpp:11-42) exactly.
  tree_gravity (src/bhtree.cpp:301-331) and gravity_force
    - at theta -> 0 the walk equals the direct sum (src/gravity_force.cpp:70-84);
    - at theta = 0.5 it is within 1e-2 RMS of the direct sum;
    - the potential energy of the Evrard sphere (rho ~ 1/r) is -2/3 G M^2/R.
"""
from __future__ import annotations

import numpy as np
import pytest

from _cases import CASES
from _oracle import OracleEngine
from sphcode_amd import _abi, ics
from sphcode_amd.engine import SPHError


def _tree(case):
    p, f = CASES[case]()
    e = OracleEngine(p)
    e.upload(f)
    e.make_tree()
    return p, f, e, e.tree_info()


@pytest.mark.parametrize("case", ["evrard_disph", "khi_disph", "sod_ssph", "rollout_ssph"])
def test_tree_geometry(case):
    p, f, e, t = _tree(case)
    x = np.asarray(f["pos"]).reshape(len(f["mass"]), p.dim)
    m = np.asarray(f["mass"])
    # root: bounding box or periodic box
    if p.periodic_is_valid:
        lo = np.array(p.range_min[:p.dim]); hi = np.array(p.range_max[:p.dim])
    else:
        lo, hi = x.min(0), x.max(0)
    assert np.array_equal(t["center"][0], (hi + lo) * 0.5)
    assert t["edge"][0] == np.max(hi - lo)
    # recover each node's particle set by the reference's slot rule, in
    # depth-first pre-order, and check counts, moments and containment
    n_nodes = t["num_nodes"]
    lev, cnt, edge, cen, mass, com = t["level"], t["count"], t["edge"], t["center"], t["mass"], t["com"]

    def visit(k, members):
        assert cnt[k] == len(members)
        if k > 0:
            assert mass[k] == pytest.approx(m[members].sum(), rel=1e-12)
            c = (x[members] * m[members, None]).sum(0) / m[members].sum()
            assert np.allclose(com[k], c, rtol=1e-10, atol=1e-14)
        assert np.all(np.abs(x[members] - cen[k]) <= edge[k] / 2 * (1 + 1e-12))
        leaf = k + 1 >= n_nodes or lev[k + 1] <= lev[k]
        if leaf:
            assert len(members) <= p.tree_leaf_particle_num or lev[k] == p.tree_max_level + 1
            return k + 1
        assert len(members) > p.tree_leaf_particle_num or k == 0
        slot = np.zeros(len(members), int)
        for d in range(p.dim):
            slot |= (x[members, d] > cen[k][d]).astype(int) << d
        nxt = k + 1
        for s_ in range(1 << p.dim):
            ch = members[slot == s_]
            if ch.size:
                assert lev[nxt] == lev[k] + 1
                assert edge[nxt] == edge[k] * 0.5
                nxt = visit(nxt, ch)
        return nxt

    assert visit(0, np.arange(len(m))) == n_nodes


def test_pool_exhaustion_is_an_error():
    p, f = CASES["evrard_disph"]()
    e = OracleEngine(p)
    e.upload(f)
    e.set_tree_pool(10)
    with pytest.raises(SPHError) as exc:
        e.make_tree()
    assert exc.value.code == _abi.E_TREE


@pytest.mark.parametrize("case", ["evrard_disph", "khi_disph", "khi_ssph", "sod_ssph", "rollout_ssph", "sedov_gsph"])
def test_tree_sets_equal_brute_force(case):
    p, f = CASES[case]()
    tree = OracleEngine(p)
    brute = OracleEngine(p, exhaustive=True)
    for e in (tree, brute):
        e.set_record_sets(True)
        e.upload(f)
        e.make_tree()
        e.pre_interaction()
        e.fluid_force()
    for kind in (0, 1):
        a, b = tree.neighbor_sets(kind), brute.neighbor_sets(kind)
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), (case, kind)
    da, db = tree.download(), brute.download()
    assert np.array_equal(da["neighbor"], db["neighbor"])
    assert np.allclose(da["dens"], db["dens"], rtol=1e-12)
    assert np.allclose(da["acc"], db["acc"], rtol=1e-10, atol=1e-12 * np.abs(db["acc"]).max())


def _gravity(params, fields, exhaustive=False):
    e = OracleEngine(params, exhaustive=exhaustive)
    e.upload(fields)
    e.make_tree()
    e.pre_interaction()
    e.fluid_force()
    hydro = e.download(("acc",))["acc"].copy()
    e.gravity_force()
    d = e.download(("acc", "phi", "mass"))
    return d["acc"] - hydro, d["phi"], d["mass"], e


def test_gravity_theta_to_zero_is_direct_sum():
    p, f = ics.evrard(1500, leaf=8)
    p.gravity_theta = 1e-8
    a_tree, phi_tree, _, _ = _gravity(p, f)
    a_dir, phi_dir, _, _ = _gravity(p, f, exhaustive=True)
    assert np.allclose(phi_tree, phi_dir, rtol=1e-12)
    scale = np.sqrt(np.mean(a_dir ** 2))
    assert np.max(np.abs(a_tree - a_dir)) < 1e-11 * scale


def test_gravity_theta_half_accuracy():
    p, f = ics.evrard(3000, leaf=16)
    a_tree, phi_tree, _, e = _gravity(p, f)
    a_dir, phi_dir, _, _ = _gravity(p, f, exhaustive=True)
    rms = np.sqrt(np.mean(np.sum((a_tree - a_dir) ** 2, 1)) / np.mean(np.sum(a_dir ** 2, 1)))
    assert rms < 1e-2
    st = e.stats()
    assert st["gravity_pn"] > 0 and st["gravity_pp"] > 0


def test_evrard_potential_energy():
    """Sum of m phi / 2 for rho = 1/(2 pi r), M = R = G = 1 is -2/3 (plus the
    small softening correction, positive)."""
    p, f = ics.evrard(20000, leaf=32)
    _, phi, m, e = _gravity(p, f)
    w = 0.5 * np.sum(m * phi)
    assert -0.69 < w < -0.63

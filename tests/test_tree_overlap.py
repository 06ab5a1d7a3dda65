"""The tree's box/simplex overlap test against exact overlap by linear programming.

The reference decides whether a simplex belongs to a child box by
shape_intersection (src/painter/space/shape_intersection.h:326-385): a vertex of
one shape inside the other, else no overlap when one shape's constraints put
all of the other's vertices on one side (:150-189), else the feasibility of
both shapes' constraint systems by its simplex-method LP (:191-278,
src/numerical/simplex.h:299-434), the simplex's hyperplane relaxed by 10
epsilons.  That is exact convex overlap.  This build (GPU and checker alike)
uses a conservative test instead: the simplex's bounding box, its N facet
constraints and its hyperplane, each against the box's centre and half sizes
with the reference's guard distance.  The tree's leaf lists must therefore be
supersets of the reference's: every box/simplex pair that overlaps must pass.

Exact overlap here is decided by an LP solved with scipy's HiGHS (a library
routine): t* = min t subject to lo - t <= V^T l <= hi + t, l >= 0, sum l = 1,
the L-infinity gap between the box [lo, hi] and the simplex with vertices V
(t* <= 0: they overlap).  Pairs with |t*| below 1e-9 are too close to call at
the solver's tolerance and are left out.
"""
import numpy as np
import pytest
from scipy.optimize import linprog

from math_amd import native
from oracle import oracle as O


def box_gap(V, lo, hi):
    """L-infinity gap t* between the box [lo, hi] and conv(V) (negative: overlap depth)."""
    k, n = V.shape
    # variables: l (k), t; minimise t
    c = np.zeros(k + 1)
    c[-1] = 1.0
    A, b = [], []
    for i in range(n):
        A.append(np.concatenate([V[:, i], [-1.0]]))   # x_i - t <= hi_i
        b.append(hi[i])
        A.append(np.concatenate([-V[:, i], [-1.0]]))  # -x_i - t <= -lo_i
        b.append(-lo[i])
    A_eq = [np.concatenate([np.ones(k), [0.0]])]
    r = linprog(c, A_ub=np.array(A), b_ub=np.array(b), A_eq=np.array(A_eq), b_eq=[1.0],
                bounds=[(0, None)] * k + [(None, None)], method="highs")
    assert r.status == 0
    return r.fun


def sample_pairs(m, v, f, rng, count):
    """Boxes of the tree's cells (level 1..depth) around points near simplices:
    the cell containing a point of a random simplex, offset by -1..1 cells."""
    n = v.shape[1]
    lo, size = m.tree_root()
    depth = m.tree_info()["depth"]
    out = []
    for _ in range(count):
        s = int(rng.integers(len(f)))
        lam = rng.dirichlet(np.ones(n))
        p = lam @ v[f[s]].astype(np.float64)
        level = int(rng.integers(1, depth + 1))
        cell = size / 2.0 ** level
        q = np.floor((p - lo) / cell) + rng.integers(-1, 2, size=n)
        q = np.clip(q, 0, 2 ** level - 1)
        # a nearby simplex, or the same one
        s2 = s if rng.random() < 0.5 else int(rng.integers(len(f)))
        out.append((s2 if rng.random() < 0.3 else s, lo + q * cell, cell))
    return out


@pytest.mark.parametrize("n,g", [(4, 3), (5, 2)])
def test_conservative_overlap_is_a_superset_of_exact_overlap(n, g):
    v, f = native.meshgen_cube_sphere(n, g, True, 0.03, 0.0, 7)
    m = O.Mesh(v, f)
    m.build_tree(threads=4)
    rng = np.random.default_rng(n)
    V64 = v.astype(np.float64)
    overlap = separated = admitted = 0
    for s, org, size in sample_pairs(m, v, f, rng, 1500):
        t = box_gap(V64[f[s]], org, org + size)
        if abs(t) < 1e-9:
            continue
        got = m.box_overlap(s, org, size)
        assert got in (0, 1)
        if t < 0:
            overlap += 1
            assert got == 1, (s, org, size, t)
        else:
            separated += 1
            admitted += got
    print(f"n={n}: {overlap} overlapping pairs all admitted; {admitted} of {separated} separated pairs admitted")
    assert overlap > 300 and separated > 100
    # the test is a filter, not the identity: most separated pairs are rejected
    assert admitted < 0.5 * separated


@pytest.mark.parametrize("n,g", [(4, 3), (5, 2)])
def test_leaf_lists_contain_every_overlapping_simplex(n, g):
    # the checker's tree (equal to the GPU tree, test_gpu_parity.py): for sampled
    # leaves and the simplices whose bounding boxes meet them, every simplex that
    # overlaps the leaf box (by the LP) is in the leaf's list
    v, f = native.meshgen_cube_sphere(n, g, True, 0.03, 0.0, 8)
    m = O.Mesh(v, f)
    m.build_tree(threads=4)
    lo, size = m.tree_root()
    dump, leaves, k = m.tree_dump(), [], 0
    while k < len(dump):
        c = int(dump[k + 1 + n])
        leaves.append((int(dump[k]), dump[k + 1:k + 1 + n].astype(np.float64), set(dump[k + 2 + n:k + 2 + n + c].tolist())))
        k += 2 + n + c
    V64 = v.astype(np.float64)
    smin, smax = V64[f].min(axis=1), V64[f].max(axis=1)
    rng = np.random.default_rng(10 + n)
    checked = 0
    for li in rng.choice(len(leaves), size=400, replace=False):
        level, q, members = leaves[li]
        cell = size / 2.0 ** level
        blo, bhi = lo + q * cell, lo + (q + 1) * cell
        near = np.nonzero(np.all(smin <= bhi, axis=1) & np.all(smax >= blo, axis=1))[0]
        for s in near:
            t = box_gap(V64[f[s]], blo, bhi)
            if t < -1e-9:
                assert int(s) in members, (li, level, int(s), t)
                checked += 1
    print(f"n={n}: {checked} overlapping (leaf, simplex) pairs, all listed")
    assert checked > 200

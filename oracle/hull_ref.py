"""CPU restatement of the convex hull and Delaunay construction, for tests only.

This is test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker.  The product path is
math_amd/csrc/nsp_hull.hip.

It follows compute_convex_hull / compute_delaunay (geometry/core/convex_hull.cpp):
- quantise(): shuffle_and_convert_to_unique_integer (convex_hull.cpp:661-711)
  without the shuffle.  The caller passes the insertion order, such as the
  library's `order` output; that order is pinned separately against
  std::shuffle in tests/cpp.
- hull(): the incremental construction (create_convex_hull, convex_hull.cpp:589-645)
  with exact Python integers.  Each step tests the point against every live
  facet instead of keeping conflict lists.  The set of facets a point sees is
  the same either way, so the facets made are the same.
- facet orientation: FacetInteger (geometry/core/facet.h:178-207).
- initial simplex: find_simplex_points (convex_hull.cpp:255-298).
- delaunay(): paraboloid_convex_hull (convex_hull.cpp:723-795).

Pinned by tests/test_hull.py against:
- scipy.spatial.ConvexHull and Delaunay (Qhull) facet sets on points in general position;
- the cross-polytope's 2^n facets and their normals (1, ..., 1)/sqrt(n) up to sign;
- the reference test's own checks (test_convex_hull.cpp:112-141): every ridge
  joins two facets, and no point lies beyond a facet;
- Delaunay's empty-circumsphere property, checked by brute force.
"""

import math
from itertools import combinations

import numpy as np


def quantise(points, order, max_value):
    """Integer points of `points` (float32 [count, n]) in insertion order
    `order` (input indices), scaled as the reference scales them; duplicates
    after rounding are dropped, keeping the first.  Returns (ints, kept_order)."""
    p = np.asarray(points, np.float32)
    mn = p.min(axis=0)
    mx = p.max(axis=0)
    max_d = float(np.max((mx - mn).astype(np.float32)))
    if max_d == 0:
        raise ValueError("all points are equal to each other")
    scale = float(max_value) / max_d
    seen = set()
    ints, kept = [], []
    for r in order:
        diff = (p[r] - mn).astype(np.float32)
        v = []
        for x in diff:
            y = float(x) * scale
            # llround: round half away from zero
            ll = int(math.floor(abs(y) + 0.5)) * (1 if y >= 0 else -1)
            if ll < 0 or ll > max_value:
                raise ValueError("points preprocessing error")
            v.append(ll)
        t = tuple(v)
        if t not in seen:
            seen.add(t)
            ints.append(t)
            kept.append(int(r))
    return ints, kept


def _det(m):
    k = len(m)
    if k == 1:
        return m[0][0]
    if k == 2:
        return m[0][0] * m[1][1] - m[0][1] * m[1][0]
    s = 0
    for j in range(k):
        sub = [row[:j] + row[j + 1:] for row in m[1:]]
        t = m[0][j] * _det(sub)
        s = s - t if j & 1 else s + t
    return s


def _ortho(X, v):
    M = len(X[0])
    d = [[X[v[i]][j] - X[v[0]][j] for j in range(M)] for i in range(1, M)]
    out = []
    for j in range(M):
        m = _det([row[:j] + row[j + 1:] for row in d])
        out.append(-m if j & 1 else m)
    return out


def _visible(X, ortho, v0, q):
    return sum(o * (a - b) for o, a, b in zip(ortho, X[q], X[v0]))


def _independent(d, k, M):
    for cols in combinations(range(M), k):
        if _det([[row[c] for c in cols] for row in d[:k]]) != 0:
            return True
    return False


class _Facet:
    __slots__ = ("v", "ortho", "alive")

    def __init__(self, v, ortho):
        self.v = v
        self.ortho = ortho
        self.alive = True


def _oriented(X, v, inner, horizon):
    o = _ortho(X, v)
    s = _visible(X, o, v[0], inner)
    if s > 0:
        o = [-x for x in o]
    elif s == 0:
        if horizon is None:
            raise ValueError("degenerate initial simplex")
        if any((a > 0 and b < 0) or (a < 0 and b > 0) for a, b in zip(o, horizon.ortho)):
            o = [-x for x in o]
    return o


def hull(X):
    """Facets (vertex tuples, integer normals) of the hull of integer points X
    (list of M-tuples) inserted in list order."""
    P = len(X)
    M = len(X[0])
    if P < M + 1:
        raise ValueError("Error point count")
    simplex = [0]
    d = []
    nxt = 1
    for k in range(1, M + 1):
        p = nxt
        while p < P:
            d_try = d + [[X[p][j] - X[0][j] for j in range(M)]]
            if _independent(d_try, k, M):
                d = d_try
                break
            p += 1
        if p >= P:
            raise ValueError(f"point {k + 1} of {M}-simplex not found")
        simplex.append(p)
        nxt = p + 1
    facets = []
    for f in range(M + 1):
        v = [simplex[i] for i in range(M + 1) if i != f]
        facets.append(_Facet(v, _oriented(X, v, simplex[f], None)))
    # ridge -> facets, kept as the live adjacency
    def ridges(fc):
        for r in range(M):
            yield r, tuple(sorted(fc.v[:r] + fc.v[r + 1:]))

    adj = {}
    for fc in facets:
        for r, key in ridges(fc):
            adj.setdefault(key, []).append(fc)
    in_simplex = set(simplex)
    for p in range(P):
        if p in in_simplex:
            continue
        live = [fc for fc in facets if fc.alive]
        seen = [fc for fc in live if _visible(X, fc.ortho, fc.v[0], p) > 0]
        if not seen:
            continue
        if len(seen) >= len(live):
            raise ValueError("All facets are visible from the point")
        seen_ids = set(id(fc) for fc in seen)
        created = []
        for fc in seen:
            for r, key in ridges(fc):
                other = [g for g in adj[key] if g is not fc]
                assert len(other) == 1
                g = other[0]
                if id(g) in seen_ids:
                    continue
                inner = [x for x in g.v if x not in key][0]
                v = list(fc.v)
                v[r] = p
                created.append(_Facet(v, _oriented(X, v, inner, g)))
        for fc in seen:
            fc.alive = False
            for r, key in ridges(fc):
                adj[key] = [g for g in adj[key] if g is not fc]
        for fc in created:
            for r, key in ridges(fc):
                adj.setdefault(key, []).append(fc)
        facets.extend(created)
    return [(tuple(fc.v), tuple(fc.ortho)) for fc in facets if fc.alive]


def convex_hull(points, order):
    """(vertex index tuples in input numbering, unit normals) of compute_convex_hull
    for points inserted in `order`."""
    ints, kept = quantise(points, order, (1 << 30) - 1)
    out = []
    for v, o in hull(ints):
        ln = math.sqrt(sum(float(x) * float(x) for x in o))
        out.append((tuple(kept[i] for i in v), tuple(float(x) / ln for x in o)))
    return out


def delaunay(points, order):
    """(vertex tuples, per-vertex face normals, integer points) of compute_delaunay."""
    ints, kept = quantise(points, order, (1 << 24) - 1)
    N = len(ints[0])
    lifted = [tuple(list(x) + [sum(c * c for c in x)]) for x in ints]
    out = []
    for v, o in hull(lifted):
        if o[N] >= 0:
            continue
        normals = []
        for r in range(N + 1):
            face = [v[i] for i in range(N + 1) if i != r]
            w = _ortho(ints, face)
            s = _visible(ints, w, face[0], v[r])
            if s == 0:
                raise ValueError("degenerate simplex")
            if s > 0:
                w = [-x for x in w]
            ln = math.sqrt(sum(float(x) * float(x) for x in w))
            normals.append(tuple(float(x) / ln for x in w))
        out.append((tuple(kept[i] for i in v), normals))
    return out, ints, kept

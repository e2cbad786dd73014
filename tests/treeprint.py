"""Index-free tree fingerprint (after tests/test_radix_build.py:250-282 of the reference).

Node indices legitimately differ between builds (the reference's serial split numbers nodes in
stack discovery order, the GPU build in depth-first preorder), so trees are compared by what a
traversal sees: the particle visit order (as input indices) and one record per node in DFS order.
"""

import numpy as np


def fingerprint(t, with_quads=False):
    n = t.NumParticles
    First, Next = np.asarray(t.FirstSubnode), np.asarray(t.NextBranch)
    M, C, D, S, Z = (np.asarray(a) for a in (t.Masses, t.Coordinates, t.Deltas, t.Softenings, t.Sizes))
    Q = np.asarray(t.Quadrupoles).reshape(-1, 9) if with_quads else None
    idx = np.asarray(t.TreewalkIndices)
    order, nodes, st = [], [], [n]
    budget = 8 * t.NumNodes + 64
    while st:
        budget -= 1
        if budget < 0:
            raise AssertionError("link chain does not terminate")
        no = st.pop()
        if no < n:
            order.append(int(idx[no]))
            continue
        kids, c, stop = [], First[no], Next[no]
        while c != -1 and c != stop:
            kids.append(int(c))
            c = Next[c]
            if len(kids) > 8:
                raise AssertionError("more than 8 children")
        rec = [M[no], C[no, 0], C[no, 1], C[no, 2], D[no], S[no], Z[no], float(len(kids)),
               float(sum(1 for k in kids if k < n))]
        if with_quads:
            rec += list(Q[no])
        nodes.append(rec)
        st.extend(reversed(kids))
    return np.array(order, dtype=np.int64), np.array(nodes, dtype=np.float64)


def assert_same_tree(a, b, with_quads=False):
    oa, na = fingerprint(a, with_quads)
    ob, nb = fingerprint(b, with_quads)
    assert len(oa) == len(ob) and np.array_equal(oa, ob), "particle visit order differs"
    assert na.shape == nb.shape, f"node count differs: {na.shape} vs {nb.shape}"
    # bit-exact, NaN-safe
    va, vb = na.view(np.int64), nb.view(np.int64)
    if not np.array_equal(va, vb):
        bad = np.argwhere(va != vb)
        i, j = bad[0]
        raise AssertionError(f"{len(bad)} node fields differ; first at node {i} field {j}: {na[i, j]!r} vs {nb[i, j]!r}")

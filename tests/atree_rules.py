"""Brute-force geometric definitions of DAPIF2's adaptive tree and its lists, independent of
the reference's construction order (dapif1.f:1165-1680, dapif2.f:3-474).

The reference domain is the square [-64, 64]^2; level-1 boxes are its quadrants (width 64),
and a box of level l has width 128 / 2^l.  Two boxes touch when their centres are at most
half the sum of their widths apart in both coordinates.
"""
import numpy as np


def box_levels(t):
    nl = np.asarray(t["nl"])
    lev = np.zeros(len(t["nbat"]), np.int64)
    for l in range(1, len(nl) - 1):
        lev[nl[l - 1] - 1:nl[l] - 1] = l
    return lev


def touch(t, lev, a, b):
    """Boxes a and b (1-based) touch (share an edge or a corner) or overlap."""
    wa = 128.0 / 2 ** lev[a - 1]
    wb = 128.0 / 2 ** lev[b - 1]
    return (abs(t["xb"][a - 1] - t["xb"][b - 1]) <= (wa + wb) / 2
            and abs(t["yb"][a - 1] - t["yb"][b - 1]) <= (wa + wb) / 2)


def children(t, b):
    npar = np.asarray(t["npar"])
    return [int(c) + 1 for c in np.nonzero(npar == b)[0]]


def check_tree(t, napb):
    """Box, point and leaf rules.  Returns nothing; asserts."""
    nbox = len(t["nbat"])
    lev = box_levels(t)
    nlev = int(lev.max())
    xs, ys, ibox = np.asarray(t["xs"]), np.asarray(t["ys"]), np.asarray(t["ibox"])
    # every scaled point lies strictly inside [-64*0.99.., 64*0.99..] and inside its box
    assert np.abs(xs).max() <= 64.0 and np.abs(ys).max() <= 64.0
    w = 128.0 / 2.0 ** lev[ibox - 1]
    cx, cy = np.asarray(t["xb"])[ibox - 1], np.asarray(t["yb"])[ibox - 1]
    assert np.all(xs >= cx - w / 2) and np.all(xs <= cx + w / 2)
    assert np.all(ys >= cy - w / 2) and np.all(ys <= cy + w / 2)
    # counts: leaves hold their points, parents the sum of their children
    nbat = np.asarray(t["nbat"])
    indb = np.asarray(t["indb"])
    assert np.all(nbat > 0)
    cnt = np.bincount(ibox, minlength=nbox + 1)[1:]
    assert np.array_equal(cnt[indb == 1], nbat[indb == 1])
    assert np.all(cnt[indb == 0] == 0)
    for b in range(1, nbox + 1):
        ch = children(t, b)
        if indb[b - 1] == 0:
            assert nbat[b - 1] > napb and ch and sum(nbat[c - 1] for c in ch) == nbat[b - 1]
            assert t["nchi"][b - 1] == ch[0] and ch == list(range(ch[0], ch[0] + len(ch)))
            for c in ch:  # quadrant of the parent, one level down
                assert lev[c - 1] == lev[b - 1] + 1
                d = 128.0 / 2 ** (lev[b - 1] + 2)
                assert abs(abs(t["xb"][c - 1] - t["xb"][b - 1]) - d) == 0
                assert abs(abs(t["yb"][c - 1] - t["yb"][b - 1]) - d) == 0
        else:
            assert not ch and t["nchi"][b - 1] == 0
            assert nbat[b - 1] <= napb or lev[b - 1] == nlev
    # children follow their parents' order, level by level
    npar = np.asarray(t["npar"])
    for l in range(2, nlev + 1):
        p = npar[lev == l]
        assert np.all(np.diff(p) >= 0)
    # particle lists of the leaves: JAT(IAT) = count, then the points in index order
    iat, jat = np.asarray(t["iat"]), np.asarray(t["jat"])
    for b in np.nonzero(indb == 1)[0] + 1:
        a = iat[b - 1]
        assert jat[a - 1] == nbat[b - 1]
        pts = jat[a:a + nbat[b - 1]]
        assert np.array_equal(pts, np.nonzero(ibox == b)[0] + 1)
    for b in np.nonzero(indb == 0)[0] + 1:
        assert iat[b - 1] == -nbat[b - 1]


def lists_bruteforce(t):
    """Sets of lists 1..4 of every box from the geometric definitions."""
    nbox = len(t["nbat"])
    lev = box_levels(t)
    indb = np.asarray(t["indb"])
    npar = np.asarray(t["npar"])
    bylev = {}
    for b in range(1, nbox + 1):
        bylev.setdefault(int(lev[b - 1]), []).append(b)
    list1, list2, list3, list4 = {}, {}, {}, {}
    for b in range(1, nbox + 1):
        same = [c for c in bylev[int(lev[b - 1])] if c != b]
        if lev[b - 1] == 1:
            list1[b] = set()
            list2[b] = set(same)
            continue
        p = int(npar[b - 1])
        col = {c for c in same if touch(t, lev, b, c)}
        l1 = {c for c in same if not touch(t, lev, b, c) and int(npar[c - 1]) != p
              and touch(t, lev, p, int(npar[c - 1]))}
        list1[b], list2[b] = l1, col
    for b in range(1, nbox + 1):
        cols = list2[b]
        l3, l4 = set(), set()
        if indb[b - 1] == 1 and cols and any(indb[c - 1] == 0 for c in cols):
            stack = list(cols)
            while stack:
                d = stack.pop()
                if not touch(t, lev, b, d):
                    l4.add(d)
                elif indb[d - 1] == 1:
                    l3.add(d)
                else:
                    stack.extend(children(t, d))
        list3[b], list4[b] = l3, l4
    return list1, list2, list3, list4


def lists_from_tables(t):
    nbox = len(t["nbat"])
    l1p, l2p, jl = np.asarray(t["l1"]), np.asarray(t["l2"]), np.asarray(t["jl"])
    l3p, l4p, jl3, jl4 = np.asarray(t["l3"]), np.asarray(t["l4"]), np.asarray(t["jl3"]), np.asarray(t["jl4"])
    L1, L2, L3, L4 = {}, {}, {}, {}
    for b in range(1, nbox + 1):
        L1[b] = jl[l1p[b - 1] - 1:l2p[b - 1] - 1].tolist()
        L2[b] = jl[l2p[b - 1] - 1:l1p[b] - 1].tolist()
        L3[b] = jl3[l3p[b - 1] - 1:l3p[b] - 1].tolist()
        L4[b] = jl4[l4p[b - 1] - 1:l4p[b] - 1].tolist()
    return L1, L2, L3, L4

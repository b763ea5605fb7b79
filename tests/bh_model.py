"""Pure-Python models of the Barnes-Hut build, used by the tests as checkers.

* insert_build: an independent restatement (numpy float32 scalars) of the
  reference's sequential insertion build, src/cuda/force_barnes_hut.cu:391-543
  of build-workbench/nbody-gpu-sim, to pin the C oracle's orc_build_tree.
* numbering_build: the rule the GPU's parallel builder uses — per-particle
  octant paths, cells ordered by (k, t1 == k, depth) with
  k(X) = max(t2(parent), t1(X)) — written out in Python, to show that the rule
  reproduces the sequential build's node indices.
Both are slow and meant for small inputs.
"""
import numpy as np

F = np.float32
MAX_DEPTH = 32
NODE_FIELDS = ("center", "half_size", "com", "total_mass", "children", "particle_index",
               "is_leaf", "particle_count")


def new_node():
    return {"center": [F(0), F(0), F(0)], "half_size": F(0), "com": [F(0), F(0), F(0)],
            "total_mass": F(0), "children": [-1] * 8, "particle_index": -1, "is_leaf": True,
            "particle_count": 0}


def octant(p, c):
    return (1 if p[0] >= c[0] else 0) | (2 if p[1] >= c[1] else 0) | (4 if p[2] >= c[2] else 0)


def root_cell(bmin, bmax):
    bmin = [F(v) for v in bmin]
    bmax = [F(v) for v in bmax]
    c = [F((bmin[a] + bmax[a]) * F(0.5)) for a in range(3)]
    ext = [F(bmax[a] - bmin[a]) for a in range(3)]
    big = max(ext)
    return c, F(F(big * F(0.5)) + F(0.001))


def child_center(c, half, o):
    hs = F(half * F(0.5))
    return [F(c[0] + (hs if o & 1 else -hs)), F(c[1] + (hs if o & 2 else -hs)),
            F(c[2] + (hs if o & 4 else -hs))], hs


def insert_build(pos, mass, sorted_idx, bmin, bmax, max_nodes):
    """Reference insertion build. Returns (nodes list, max_depth)."""
    nodes = [new_node()]
    c, h = root_cell(bmin, bmax)
    root = nodes[0]
    root.update(center=c, half_size=h, is_leaf=False, particle_count=len(sorted_idx))

    def leaf(parent, o, pid, p, m):
        pn = nodes[parent]
        nc, hs = child_center(pn["center"], pn["half_size"], o)
        pn["children"][o] = len(nodes)
        pn["is_leaf"] = False
        nd = new_node()
        nd.update(center=nc, half_size=hs, particle_index=pid, particle_count=1,
                  total_mass=F(m), com=[F(v) for v in p])
        nodes.append(nd)

    def merge(nd, pid, p, m):
        tot = F(nd["total_mass"] + F(m))
        nd["com"] = [F(F(F(nd["com"][a] * nd["total_mass"]) + F(F(p[a]) * F(m))) / tot)
                     for a in range(3)]
        nd["total_mass"] = tot
        nd["particle_count"] += 1
        if nd["particle_index"] < 0:
            nd["particle_index"] = pid

    deepest = 0
    for pid in sorted_idx:
        p = [F(v) for v in pos[pid]]
        m = F(mass[pid])
        cur, depth, merged = 0, 0, False
        while depth < MAX_DEPTH:
            nd = nodes[cur]
            if nd["is_leaf"]:
                if len(nodes) + MAX_DEPTH + 1 >= max_nodes:
                    merge(nd, pid, p, m)
                    merged = True
                    break
                old, op, om = nd["particle_index"], list(nd["com"]), nd["total_mass"]
                nd.update(is_leaf=False, particle_index=-1, particle_count=0, total_mass=F(0),
                          com=[F(0), F(0), F(0)])
                oo = octant(op, nd["center"])
                if nd["children"][oo] < 0:
                    leaf(cur, oo, old, op, om)
            o = octant(p, nodes[cur]["center"])
            ch = nodes[cur]["children"][o]
            if ch < 0:
                leaf(cur, o, pid, p, m)
                break
            cur, depth = ch, depth + 1
        if not merged and depth >= MAX_DEPTH:
            lf = cur
            while not nodes[lf]["is_leaf"]:
                lf = next(c for c in nodes[lf]["children"] if c >= 0)
            merge(nodes[lf], pid, p, m)
        deepest = max(deepest, depth)
    for nd in reversed(nodes):
        if nd["is_leaf"]:
            continue
        tm, s, cnt = F(0), [F(0), F(0), F(0)], 0
        for ci in nd["children"]:
            if ci < 0:
                continue
            q = nodes[ci]
            tm = F(tm + q["total_mass"])
            s = [F(s[a] + F(q["com"][a] * q["total_mass"])) for a in range(3)]
            cnt += q["particle_count"]
        nd["total_mass"], nd["particle_count"] = tm, cnt
        if tm > 0:
            nd["com"] = [F(s[a] / tm) for a in range(3)]
    return nodes, deepest


def paths(pos, sorted_idx, bmin, bmax):
    """Octant digits of every particle (by insertion rank) at depths 0..31."""
    c0, h0 = root_cell(bmin, bmax)
    out = []
    for pid in sorted_idx:
        p = [F(v) for v in pos[pid]]
        c, h, digs = list(c0), h0, []
        for _ in range(MAX_DEPTH):
            o = octant(p, c)
            digs.append(o)
            c, h = child_center(c, h, o)
        out.append(tuple(digs))
    return out


def numbering_build(pos, mass, sorted_idx, bmin, bmax):
    """Cells of the PR-octree keyed by path prefix, numbered by (k, t1 == k, depth).

    Returns {prefix tuple: index} for every cell; index 0 is the root."""
    P = paths(pos, sorted_idx, bmin, bmax)
    members = {(): list(range(len(P)))}
    cells = {}
    frontier = [()]
    while frontier:
        nxt = []
        for pre in frontier:
            ranks = members[pre]
            d = len(pre)
            by_oct = {}
            for r in ranks:  # ranks stay ascending inside each cell
                by_oct.setdefault(P[r][d], []).append(r)
            t2 = ranks[1] if len(ranks) > 1 else None
            for o, rs in sorted(by_oct.items()):
                ch = pre + (o,)
                members[ch] = rs
                t1 = rs[0]
                k = t1 if d == 0 else max(t2, t1)
                cells[ch] = (k, 1 if t1 == k else 0, d + 1)
                if len(rs) >= 2 and d + 1 < MAX_DEPTH:
                    nxt.append(ch)
        frontier = nxt
    order = sorted(cells, key=lambda pre: cells[pre])
    index = {(): 0}
    for i, pre in enumerate(order):
        index[pre] = i + 1
    return index, members


def nodes_to_array(nodes, dtype):
    a = np.zeros(len(nodes), dtype)
    for i, nd in enumerate(nodes):
        a[i]["center"] = nd["center"]
        a[i]["half_size"] = nd["half_size"]
        a[i]["com"] = nd["com"]
        a[i]["total_mass"] = nd["total_mass"]
        a[i]["children"] = nd["children"]
        a[i]["particle_index"] = nd["particle_index"]
        a[i]["is_leaf"] = 1 if nd["is_leaf"] else 0
        a[i]["particle_count"] = nd["particle_count"]
    return a


def same_nodes(a, b):
    """Field-wise equality of two node arrays (padding bytes ignored); floats bit-exact."""
    if len(a) != len(b):
        return False, f"node count {len(a)} != {len(b)}"
    for f in NODE_FIELDS:
        x, y = a[f], b[f]
        if x.dtype.kind == "f":
            eq = np.array_equal(x.view(np.uint32), y.view(np.uint32))
        else:
            eq = np.array_equal(x, y)
        if not eq:
            bad = np.nonzero(np.any((x != y).reshape(len(a), -1), axis=1))[0]
            return False, f"field {f} differs at nodes {bad[:8].tolist()}"
    return True, ""

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


def _lcp(a, b):
    k = 0
    while k < MAX_DEPTH and a[k] == b[k]:
        k += 1
    return k


def path_order_build(pos, mass, sorted_idx, bmin, bmax, max_nodes):
    """The GPU's construction from the particles' order by octant path.

    Particles sorted by path (octants compared as 7 - o, ties by rank);
    lcp[j] = digits shared with position j - 1; position j starts the cells at
    depths lcp[j] + 1 .. L[j] = min(32, max(lcp[j], lcp[j + 1]) + 1), numbered
    in preorder as first_pre[j] + d - lcp[j] - 1 with first_pre = 1 + the
    exclusive scan of L - lcp; a cell ending before position e has skip
    first_pre[e]. Children come from skip chains, centres of mass bottom-up in
    octant order, indices from (k, t1 == k, depth). Returns (nodes array in the
    reference numbering, max depth, each node's preorder position), or None when
    the node budget could fire (the sequential builder's case)."""
    n = len(sorted_idx)
    P = paths(pos, sorted_idx, bmin, bmax)
    R = sorted(range(n), key=lambda r: (tuple(7 - d for d in P[r]), r))
    sp = [P[r] for r in R]
    lcp = [0] * (n + 1)
    for j in range(1, n):
        lcp[j] = _lcp(sp[j - 1], sp[j])
    L = [min(MAX_DEPTH, max(lcp[j], lcp[j + 1]) + 1) for j in range(n)]
    first = [1] * (n + 1)
    for j in range(n):
        first[j + 1] = first[j] + L[j] - lcp[j]
    C = first[n]
    if C + MAX_DEPTH + 1 >= max_nodes:
        return None
    c0, h0 = root_cell(bmin, bmax)
    depth, octn, leaf, skip = [0] * C, [0] * C, [False] * C, [0] * C
    box, run = [None] * C, [None] * C
    skip[0], box[0], run[0] = C, (c0, h0), (0, n)
    for j in range(n):
        cc, hh = list(c0), h0
        for d in range(1, L[j] + 1):
            cc, hh = child_center(cc, hh, sp[j][d - 1])
            if d <= lcp[j]:
                continue
            p = first[j] + d - lcp[j] - 1
            e = j + 1
            while e < n and _lcp(sp[j], sp[e]) >= d:
                e += 1
            depth[p], octn[p], leaf[p] = d, sp[j][d - 1], d == L[j]
            skip[p], box[p], run[p] = first[e], (cc, hh), (j, e)
    kids = [[] for _ in range(C)]
    for p in range(C):
        q = p + 1
        while not leaf[p] and q < skip[p]:
            kids[p].append(q)
            q = skip[q]
    com, tm, cnt = [None] * C, [F(0)] * C, [0] * C
    t1, t2, pidx = [0] * C, [0] * C, [-1] * C
    none = 1 << 40
    for p in reversed(range(C)):
        if leaf[p]:
            mem = [R[x] for x in range(*run[p])]  # ranks ascending: insertion order
            q = sorted_idx[mem[0]]
            c, m = [F(v) for v in pos[q]], F(mass[q])
            for r in mem[1:]:
                q2 = sorted_idx[r]
                pm = F(mass[q2])
                tot = F(m + pm)
                c = [F(F(F(c[a] * m) + F(F(pos[q2][a]) * pm)) / tot) for a in range(3)]
                m = tot
            com[p], tm[p], cnt[p], pidx[p] = c, m, len(mem), sorted_idx[mem[0]]
            t1[p], t2[p] = mem[0], (mem[1] if len(mem) > 1 else none)
        else:
            m, s, k, ranks = F(0), [F(0)] * 3, 0, []
            for q in sorted(kids[p], key=lambda q: octn[q]):
                m = F(m + tm[q])
                s = [F(s[a] + F(com[q][a] * tm[q])) for a in range(3)]
                k += cnt[q]
                ranks += [t1[q], t2[q]]
            ranks.sort()
            com[p] = [F(s[a] / m) for a in range(3)] if m > 0 else [F(0)] * 3
            tm[p], cnt[p], t1[p], t2[p] = m, k, ranks[0], ranks[1]
    key = [0] * C
    for p in range(C):
        for q in kids[p]:
            kk = t1[q] if p == 0 else max(t2[p], t1[q])
            key[q] = ((kk + 1) << 7) | ((1 if t1[q] == kk else 0) << 6) | depth[q]
    order = sorted(range(C), key=lambda p: (key[p], p))
    fin = [0] * C
    for f, p in enumerate(order):
        fin[p] = f
    out = []
    for p in order:
        nd = new_node()
        ch = [-1] * 8
        for q in kids[p]:
            ch[octn[q]] = fin[q]
        nd.update(center=box[p][0], half_size=box[p][1], com=com[p], total_mass=tm[p],
                  children=ch, particle_index=pidx[p], is_leaf=leaf[p], particle_count=cnt[p])
        out.append(nd)
    multi = any(leaf[p] and cnt[p] > 1 for p in range(C))
    pre_of = [0] * C  # preorder position of every node, by reference index
    for p in range(C):
        pre_of[fin[p]] = p
    return out, (MAX_DEPTH if multi else max(L) - 1), pre_of


def shared_prefix_runs(pos, sorted_idx, bmin, bmax, digits):
    """Sizes of the groups of particles whose paths share their first `digits` digits."""
    groups = {}
    for p in paths(pos, sorted_idx, bmin, bmax):
        groups[p[:digits]] = groups.get(p[:digits], 0) + 1
    return sorted(groups.values(), reverse=True)


def budget_overrun_input(pairs=5, levels=(4, 5, 6), per_level=6):
    """An input on which the reference's insertion runs out of its 2N nodes.

    Two anchors fix the root cube to [-1, 1]^3. Each pair sits at the low corner
    of its own depth-3 cell, 2^-27 apart, so its insertion splits a chain of
    cells down to depth ~27 (octant 0 at every level, so everything else in the
    cell sorts after it in Morton order). Single bodies at the centres of the
    chain cells' empty sibling octants (depths 4-6) then each create one leaf in
    an empty child of an internal cell, where the insertion makes no budget check
    (force_barnes_hut.cu:368-389 via :478-486). With 5 pairs and 6 siblings per
    level (102 bodies, budget 204) the node count passes the budget.
    """
    pts = [(-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)]
    c0, h0 = root_cell([-1, -1, -1], [1, 1, 1])
    body, singles = [], []
    for k in range(pairs):
        qc, qh = c0, h0
        for o in (7 - k, k, 3):
            qc, qh = child_center(qc, qh, o)
        corner = [float(qc[a] - qh) for a in range(3)]
        c = [v + 2.0 ** -27 for v in corner]
        body += [tuple(c), (c[0] + 1.5 * 2.0 ** -27, c[1] + 2.0 ** -27, c[2] + 0.5 * 2.0 ** -27)]
        cc, hh = qc, qh
        for d in range(4, max(levels) + 1):
            if d in levels:
                for s in range(1, 1 + per_level):
                    singles.append(tuple(child_center(cc, hh, s)[0]))
            cc, hh = child_center(cc, hh, 0)
    return pts + body + singles

"""Independent restatement of arc simplification with explicit cell lists (test infrastructure).

Built from the mathematics of the reference's stages, not from the oracle's code: the only oracle
input is the discrete gradient (``oracle.gradient``, pinned on its own by the fixtures, the
exhaustive lower-star tables and rank invariance). From it this module walks the V-paths, runs the
elder-rule sweep with the thresholded union-find, and merges arcs as plain Python lists:

* raw arcs (reference cpu/arcs_geometry.hxx:56-166): from each valid coface (faces) or vertex of
  a saddle, in the reference's order, follow the V-path, recording faces / vertices, to a critical
  cell (target) or off the grid (target -1, vertex arcs only);
* sweep (reference cpu/persistence.hxx:50-163): saddles by (value, id), ascending on the vertex
  side and descending on the face side in the primal complex (reversed in the dual); the younger
  root dies; a pair with persistence <= threshold is a cancellation whose dead extremum is the
  dying side's root in the thresholded union-find;
* merging (reference cpu/arcs_simplification.hxx:383-449): cancellations by (persistence, saddle
  id, target id); when one current root of the saddle's two arc targets is the dead extremum (and
  the roots differ), it gets the survivor as parent and the path reverse(arc_dead ++
  path(T_dead)) ++ arc_other ++ path(T_other); path(e) concatenates these paths up the parents;
* output (reference cpu/arcs_simplification.hxx:91-141,251-292, cpu/workspace.hxx:260-287): each
  surviving saddle, in vertex-side sweep order, emits arc k = raw arc k ++ path(T_k), or nothing
  when T_k is -1; arc offsets mark the end of arc 0.
"""

import numpy as np

import oracle


class Complex:
    def __init__(self, f, dual):
        self.f = np.ascontiguousarray(f, dtype=np.float32)
        self.H, self.W = self.f.shape
        self.Nx = self.W + 1
        self.dual = dual
        self.pw = oracle.gradient(self.f, dual)
        self.flat = self.f.reshape(-1)

    # total order of (value, id) under std::tie; every comparison reversed in the dual complex
    def key(self, v, i):
        return (float(v), i) if not self.dual else (-float(v), -i)

    def top_pixel(self, c):
        """Pixel carrying the cell's value: its highest vertex by (value, pixel id); the lowest in the dual."""
        H, W, Nx = self.H, self.W, self.Nx
        t, q = c % 4, c // 4
        y, x = q // Nx, q % Nx
        cand = {0: [(y, x)], 1: [(y, x - 1), (y, x)], 2: [(y - 1, x), (y, x)],
                3: [(y - 1, x - 1), (y - 1, x), (y, x - 1), (y, x)]}[t]
        cand = [a * W + b for a, b in cand if 0 <= a < H and 0 <= b < W]
        return max(cand, key=lambda p: self.key(self.flat[p], p))

    def value(self, c):
        return np.float32(self.flat[self.top_pixel(c)])

    def valid(self, c):
        t, q = c % 4, c // 4
        y, x = q // self.Nx, q % self.Nx
        return {0: y < self.H and x < self.W, 1: y < self.H, 2: x < self.W, 3: True}[t]

    def starts(self, s, faces):
        """Valid cofaces (faces) or vertices of saddle edge s, in the reference's order."""
        H, W, Nx = self.H, self.W, self.Nx
        t, q = s % 4, s // 4
        y, x = q // Nx, q % Nx
        if faces:
            return [4 * q + 3, 4 * (q + Nx) + 3 if t == 1 else 4 * (q + 1) + 3]
        if t == 1:
            cand = [(y, x - 1) if x - 1 >= 0 else None, (y, x) if x < W else None]
        else:
            cand = [(y - 1, x) if y - 1 >= 0 else None, (y, x) if y < H else None]
        return [4 * (a * Nx + b) for a, b in (c for c in cand if c is not None)]

    def walk(self, start, faces):
        """V-path from a face (vertex): the cells of its kind in order, and the critical end (-1 off grid)."""
        H, W, Nx = self.H, self.W, self.Nx
        cells, cur = [], start
        while True:
            cells.append(cur)
            e = int(self.pw[cur])
            if e == -1:
                return cells, cur
            t, q = e % 4, e // 4
            ey, ex = q // Nx, q % Nx
            if faces:
                a, b = 4 * q + 3, (4 * (q + Nx) + 3 if t == 1 else 4 * (q + 1) + 3)
            elif t == 1:
                a = 4 * (ey * Nx + ex - 1) if ex - 1 >= 0 else -1
                b = 4 * (ey * Nx + ex) if ex < W else -1
            else:
                a = 4 * ((ey - 1) * Nx + ex) if ey - 1 >= 0 else -1
                b = 4 * (ey * Nx + ex) if ey < H else -1
            cur = b if a == cur else a
            if cur == -1:
                return cells, -1


def simplified_arcs(f, thr, dual):
    """Returns (sad_pts, {"faces": (cells, offsets, arc_offsets), "verts": (...)}, merges) for every
    surviving saddle in output order; merges counts the arc contractions of both sides."""
    cx = Complex(f, dual)
    ids = np.arange(cx.pw.size)
    t = ids % 4
    crit = [int(c) for c in ids[(cx.pw == -1)] if cx.valid(int(c))]
    sads = [c for c in crit if c % 4 in (1, 2)]
    ext = {"verts": sorted(c for c in crit if c % 4 == 0), "faces": sorted(c for c in crit if c % 4 == 3)}
    index = {side: {c: i for i, c in enumerate(ext[side])} for side in ext}

    raw = {}
    for side in ("faces", "verts"):
        faces = side == "faces"
        for s in sads:
            arcs = []
            for st in cx.starts(s, faces):
                cells, end = cx.walk(st, faces)
                arcs.append((cells, index[side][end] if end != -1 else -1))
            while len(arcs) < 2:
                arcs.append(([], -1))
            raw[side, s] = arcs

    thr32 = np.float32(thr)
    cancels = {"faces": [], "verts": []}
    removed = set()
    for side in ("verts", "faces"):
        # extremum age: older = lower on the vertex side (primal), higher on the face side; void oldest
        order = sorted(sads, key=lambda s: cx.key(cx.value(s), s))
        if side == "faces":
            order = order[::-1]
        g_parent, t_parent = {}, {}

        def root(p, a):
            if a == -1:
                return -1
            while p.get(a, a) != a:
                a = p[a]
                if a == -1:
                    return -1
            return a

        def younger(a, b):  # True when extremum index a dies against b
            ka = cx.key(cx.value(ext[side][a]), ext[side][a])
            kb = cx.key(cx.value(ext[side][b]), ext[side][b])
            return ka > kb if side == "verts" else ka < kb

        for s in order:
            # the sweep pairs through each saddle's two V-path ends; a missing vertex is the void
            ends = []
            for st in (cx.starts(s, side == "faces") if side == "faces" else _vertex_slots(cx, s)):
                if st is None:
                    ends.append(-1)
                else:
                    _, e = cx.walk(st, side == "faces")
                    ends.append(index[side][e] if e != -1 else -1)
            r1, r2 = root(g_parent, ends[0]), root(g_parent, ends[1])
            if r1 == r2:
                continue
            if r1 == -1:
                dying, surv = r2, -1
            elif r2 == -1:
                dying, surv = r1, -1
            elif younger(r1, r2):
                dying, surv = r1, r2
            else:
                dying, surv = r2, r1
            sv, dv = cx.value(s), cx.value(ext[side][dying])
            pers = np.float32(sv - dv) if (side == "verts") != dual else np.float32(dv - sv)
            g_parent[dying] = surv
            if pers <= thr32:
                removed.add(s)
                d, v = root(t_parent, ends[0]), root(t_parent, ends[1])
                if dying == r2:
                    d, v = v, d
                if d != -1 and d != v:
                    t_parent[d] = v
                cancels[side].append((pers, s, ext[side][d] if d != -1 else -1, d))

    merges = [0]
    paths = {side: _merge(raw, cancels[side], side, merges) for side in ("faces", "verts")}

    surv_order = sorted((s for s in sads if s not in removed), key=lambda s: cx.key(cx.value(s), s))
    result = {}
    for side in ("faces", "verts"):
        path = paths[side]
        cells, offs, arc_offs = [], [0], []
        for s in surv_order:
            for k in range(2):
                a, T = raw[side, s][k]
                if T != -1:
                    cells += a + path(T)
                if k == 0:
                    arc_offs.append(len(cells))
            offs.append(len(cells))
        result[side] = (np.array(cells, np.int32), np.array(offs, np.int32), np.array(arc_offs, np.int32))
    return np.array(surv_order, np.int32), result, merges[0]


def _vertex_slots(cx, s):
    """Both vertex positions of saddle edge s (None where the vertex is off the grid): the sweep
    keeps the position, a missing vertex being the void."""
    H, W, Nx = cx.H, cx.W, cx.Nx
    t, q = s % 4, s // 4
    y, x = q // Nx, q % Nx
    if t == 1:
        return [4 * (y * Nx + x - 1) if x - 1 >= 0 else None, 4 * (y * Nx + x) if x < W else None]
    return [4 * ((y - 1) * Nx + x) if y - 1 >= 0 else None, 4 * (y * Nx + x) if y < H else None]


def _merge(raw, cancels, side, merges):
    """Arc merging of one side; returns path(e), the cells from extremum e to its final root."""
    parent, weight = {}, {}

    def root(a):
        while a != -1 and parent.get(a, a) != a:
            a = parent[a]
        return a

    def path(e):
        cells = []
        while e != -1 and parent.get(e, e) != e:
            cells += weight[e]
            e = parent[e]
        return cells

    merged = set()
    for pers, s, tcell, d in sorted(cancels, key=lambda c: (c[0], c[1], c[2])):
        if s in merged or d < 0:
            continue
        (a0, T0), (a1, T1) = raw[side, s]
        R0, R1 = root(T0), root(T1)
        full0 = a0 + (path(T0) if T0 != -1 else [])
        full1 = a1 + (path(T1) if T1 != -1 else [])
        if R0 == d and R0 != R1:
            parent[R0], weight[R0] = R1, full0[::-1] + full1
            merged.add(s)
        elif R1 == d and R0 != R1:
            parent[R1], weight[R1] = R0, full1[::-1] + full0
            merged.add(s)
    merges[0] += len(merged)
    return path

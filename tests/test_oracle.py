"""Pins the oracle (oracle/dmsc_oracle.cpp) before it is trusted as the parity target.

* The reference's own fixtures (tests/golden/*.npz, produced by the reference on the 32x32 landscape):
  single-thread outputs must match exactly, field by field; multi-thread and GPU outputs (whose
  extremum lists are in thread order) must match up to the induced relabelling.
* An independent persistence computation (Kruskal sweep with elder rule and a void root, written
  from the mathematics in numpy/Python, walking V-paths itself) on random fields with ties.
* Brute-force V-path labels and the Euler characteristic of the raw complex.
"""

import glob
import os

import numpy as np
import pytest

import oracle
from tests.conftest import GOLDEN

LAND = np.load(os.path.join(GOLDEN, "landscape_32x32.npz"))["field"]


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def run(dual, thr):
    return oracle.compute_dmsc(LAND, thr, return_gradient=True, is_dual=dual)


def strip_dups(a):
    if len(a) == 0:
        return a
    m = np.ones(len(a), bool)
    m[1:] = a[1:] != a[:-1]
    return a[m]


@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("flt", [False, True])
def test_single_thread_fixtures_exact(dual, flt):
    gt = load(f"ms_st_{'dual' if dual else 'primal'}_{'flt' if flt else 'raw'}")
    out = run(dual, 0.15 if flt else -1.0)
    assert set(gt) <= set(out)
    for k, v in gt.items():
        assert out[k].shape == v.shape, k
        assert np.array_equal(out[k], v), k


def relabel_equal(a, b, ext_a, ext_b, name):
    """a, b: output dicts; compare the extremum-indexed fields through cell ids."""
    def edges(d, e, p):
        return sorted((int(d["sad_pts"][s]), int(d[p][m])) for s, m in d[e])
    assert set(a[ext_a].tolist()) == set(b[ext_b].tolist()), name
    return edges


@pytest.mark.parametrize("mode", ["mt", "gpu"])
@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("flt", [False, True])
def test_thread_order_fixtures_up_to_relabelling(mode, dual, flt):
    gt = load(f"ms_{mode}_{'dual' if dual else 'primal'}_{'flt' if flt else 'raw'}")
    out = run(dual, 0.15 if flt else -1.0)
    for k in ("max_pts", "min_pts", "sad_pts"):
        assert sorted(gt[k].tolist()) == sorted(out[k].tolist()), k
    assert np.array_equal(gt["sad_pts"], out["sad_pts"])
    for e, p in (("e_max", "max_pts"), ("e_min", "min_pts")):
        ge = sorted((int(gt["sad_pts"][s]), int(gt[p][m])) for s, m in gt[e])
        oe = sorted((int(out["sad_pts"][s]), int(out[p][m])) for s, m in out[e])
        assert ge == oe, e
    for pf, pp, p in (("p_max", "ppairs_max", "max_pts"), ("p_min", "ppairs_min", "min_pts")):
        assert np.array_equal(gt[pf], out[pf]), pf
        gpp = [int(gt[p][i]) if i >= 0 else -1 for i in gt[pp]]
        opp = [int(out[p][i]) if i >= 0 else -1 for i in out[pp]]
        assert gpp == opp, pp
    for lab, p in (("peaks", "max_pts"), ("basins", "min_pts")):
        g = np.where(gt[lab] >= 0, gt[p][np.maximum(gt[lab], 0)], -1)
        o = np.where(out[lab] >= 0, out[p][np.maximum(out[lab], 0)], -1)
        assert np.array_equal(g, o), lab
    for arc, off in (("ridges", "ridge_offsets"), ("valleys", "valley_offsets")):
        assert len(strip_dups(gt[arc])) == len(strip_dups(out[arc])), arc
        assert np.array_equal(gt[off], out[off]), off


# ------------------------------------------------------------------ independent persistence
def _cells(H, W):
    return H, W, W + 1


def _walk(pw, start, Nx, H, W, faces):
    cur = start
    while pw[cur] != -1:
        e = pw[cur]
        t, q = e % 4, e // 4
        ey, ex = q // Nx, q % Nx
        if faces:
            a = 4 * (ey * Nx + ex) + 3
            b = 4 * ((ey + 1) * Nx + ex) + 3 if t == 1 else 4 * (ey * Nx + ex + 1) + 3
        else:
            if t == 1:
                a, b = 4 * (ey * Nx + ex - 1), 4 * (ey * Nx + ex)
            else:
                a, b = 4 * ((ey - 1) * Nx + ex), 4 * (ey * Nx + ex)
        cur = b if a == cur else a
    return cur


def independent_pairs(f, dual):
    """Elder-rule 0-dim persistence over the saddle / extremum multigraph with a void vertex.

    Extremum age: (value, pixel id) ascending for minima in primal (older = lower), reversed for
    maxima; void is older than everything. Saddles are swept by (value, id) ascending on the min
    side and descending on the max side (dual: reversed). Returns {saddle id: (min dying cell id or
    -1, persistence)} and the same for the max side."""
    H, W = f.shape
    Nx = W + 1
    pw = oracle.gradient(f, dual)
    flat = f.reshape(-1)

    def key(v, i):  # total order of (value, id) under std::tie, reversed in dual
        return (v, i) if not dual else (-v if v == v else v, -i)

    def cell_top(c):  # pixel with the "highest" value among the cell's vertices (reference order)
        t, q = c % 4, c // 4
        y, x = q // Nx, q % Nx
        cand = {0: [(y, x)], 1: [(y, x - 1), (y, x)], 2: [(y - 1, x), (y, x)],
                3: [(y - 1, x - 1), (y - 1, x), (y, x - 1), (y, x)]}[t]
        cand = [(a, b) for a, b in cand if 0 <= a < H and 0 <= b < W]
        best = max(cand, key=lambda p: key(flat[p[0] * W + p[1]], p[0] * W + p[1]))
        return best[0] * W + best[1]

    ids = np.arange(pw.size)
    t, q = ids % 4, ids // 4
    y, x = q // Nx, q % Nx
    valid = ((t == 0) & (y < H) & (x < W)) | ((t == 1) & (y < H)) | ((t == 2) & (x < W)) | (t == 3)
    crit = (pw == -1) & valid
    sads = ids[crit & ((t == 1) | (t == 2))]
    res = {}
    for side in ("min", "max"):
        faces = side == "max"
        parent = {}

        def find(a):
            while parent.get(a, a) != a:
                a = parent[a]
            return a

        # ties in value are broken by the cell id of the saddle / extremum itself
        def age(c):  # larger = younger
            k = key(flat[cell_top(c)], c)
            return k if side == "min" else tuple(-v for v in k)

        order = sorted(sads, key=lambda s: key(flat[cell_top(s)], int(s)))
        if side == "max":
            order = order[::-1]
        out = {}
        for s in order:
            sq = s // 4
            sy, sx = sq // Nx, sq % Nx
            if faces:
                starts = [4 * sq + 3, 4 * (sq + Nx) + 3 if s % 4 == 1 else 4 * (sq + 1) + 3]
            elif s % 4 == 1:
                starts = [4 * (sq - 1) if sx >= 1 else None, 4 * sq if sx < W else None]
            else:
                starts = [4 * (sq - Nx) if sy >= 1 else None, 4 * sq if sy < H else None]
            tg = [_walk(pw, c, Nx, H, W, faces) if c is not None else "void" for c in starts]
            r = [find(c) if c != "void" else "void" for c in tg]
            r = [find("void") if c == "void" else c for c in r]
            if r[0] == r[1]:
                out[int(s)] = (-1, -1.0)
                continue
            if r[0] == "void" or r[1] == "void":
                dying = r[1] if r[0] == "void" else r[0]
                surv = "void"
            else:
                dying, surv = (r[0], r[1]) if age(r[0]) > age(r[1]) else (r[1], r[0])
            sv = np.float32(flat[cell_top(s)])
            dv = np.float32(flat[cell_top(dying)])
            pers = (sv - dv) if (side == "min") != dual else (dv - sv)
            parent[dying] = surv
            out[int(s)] = (int(dying), float(np.float32(pers)))
        res[side] = out
    return res


@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("seed", range(12))
def test_persistence_matches_independent_elder_rule(dual, seed):
    rng = np.random.default_rng(seed)
    H, W = rng.integers(2, 12, size=2)
    f = (rng.integers(0, 5, size=(H, W)) if seed % 2 else rng.random((H, W))).astype(np.float32)
    ref = independent_pairs(f, dual)
    out = oracle.compute_dmsc(f, -1.0, is_dual=dual)
    # raw mode: every saddle survives; sad_pts is in min-side order
    mx, mn = ("min", "max") if dual else ("max", "min")
    for side_out, side_ref, pts in (("max", mx, "max_pts"), ("min", mn, "min_pts")):
        for i, s in enumerate(out["sad_pts"]):
            d, p = ref[side_ref][int(s)]
            pp = int(out["ppairs_" + side_out][i])
            assert (out[pts][pp] if pp >= 0 else -1) == d, (side_out, s)
            assert np.float32(out["p_" + side_out][i]) == np.float32(p), (side_out, s)


def _duality_fields():
    rng = np.random.default_rng(77)
    fields = []
    for k in range(40):
        H, W = (int(v) for v in rng.integers(1, 40, size=2))
        f = rng.random((H, W)).astype(np.float32)
        levels = (0, 2, 3, 16)[k % 4]
        if levels:
            f = (np.floor(f * levels) / levels).astype(np.float32)
        if k % 5 == 1:
            f[rng.random((H, W)) < 0.15] = np.inf
        if k % 5 == 2:
            f[rng.random((H, W)) < 0.15] = -np.inf
        fields.append(f)
    fields += [np.zeros((6, 9), np.float32), np.full((5, 4), np.nan, np.float32), np.full((3, 7), np.inf, np.float32)]
    return fields


@pytest.mark.parametrize("dual", [False, True])
def test_every_saddle_merges_on_exactly_one_side(dual):
    """The GPU persistence rounds drop a saddle from one side's sweep as soon as the other side has
    merged it. That rests on planar duality: with the void vertex the complex is a 2-sphere, the
    minimum and maximum saddle graphs are dual graphs on it, the ascending sweep's merges are a
    minimum spanning forest and the descending sweep's merges of the dual a maximum one, and those
    are complements. Checked on the independent elder-rule sweep and on the oracle, including ties,
    infinities, constant and all-NaN fields and one-pixel-wide shapes."""
    rng = np.random.default_rng(5)
    for seed in range(20):
        H, W = (int(v) for v in rng.integers(1, 12, size=2))
        f = (rng.integers(0, 4, size=(H, W)) if seed % 2 else rng.random((H, W))).astype(np.float32)
        ref = independent_pairs(f, dual)
        assert ref["max"].keys() == ref["min"].keys()
        for s in ref["max"]:
            assert (ref["max"][s][0] != -1) != (ref["min"][s][0] != -1), (seed, s)
    for f in _duality_fields():
        out = oracle.extract(f, -1.0, False, dual, False, False, False, False)
        merged_max, merged_min = out[8] != -1, out[9] != -1  # ppairs by output saddle, raw mode keeps all
        assert len(out[3]) > 0 and np.all(merged_max ^ merged_min), f.shape


@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("seed", range(6))
def test_raw_labels_and_euler_characteristic(dual, seed):
    rng = np.random.default_rng(100 + seed)
    H, W = rng.integers(1, 14, size=2)
    f = rng.random((H, W)).astype(np.float32)
    out = oracle.compute_dmsc(f, -1.0, is_dual=dual)
    assert len(out["min_pts"]) - len(out["sad_pts"]) + len(out["max_pts"]) == 1
    pw = oracle.gradient(f, dual)
    Nx = W + 1
    # primal: peaks on faces (H+1, W+1), basins on pixels; dual swaps
    face_lab, pix_lab = (out["peaks"], out["basins"]) if not dual else (out["basins"], out["peaks"])
    face_pts, pix_pts = (out["max_pts"], out["min_pts"]) if not dual else (out["min_pts"], out["max_pts"])
    for yy in range(H + 1):
        for xx in range(W + 1):
            tcell = _walk(pw, 4 * (yy * Nx + xx) + 3, Nx, H, W, True)
            assert face_pts[face_lab[yy, xx]] == tcell
    for yy in range(H):
        for xx in range(W):
            tcell = _walk(pw, 4 * (yy * Nx + xx), Nx, H, W, False)
            assert pix_pts[pix_lab[yy, xx]] == tcell


def test_fixture_files_present():
    assert len(glob.glob(os.path.join(GOLDEN, "ms_*.npz"))) == 12


def test_thread_count_does_not_change_outputs():
    # the oracle's parallel stages keep the single-thread output order, so full-size parity checks
    # may run it on all cores
    from dmsc_amd import synthetic

    f = synthetic.gaussian_mixture(200, 170, 12, (4.0, 20.0), 0.05, seed=5).numpy()
    for thr, dual in ((0.05, False), (0.2, True), (-1.0, False)):
        a = oracle.extract(f, thr, True, dual, True, True, True, True, nthreads=1)
        b = oracle.extract(f, thr, True, dual, True, True, True, True, nthreads=8)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y), equal_nan=np.asarray(x).dtype.kind == "f")

"""Pins of the oracle where the reference's fixtures say nothing.

* Arc merging with many cancellations (reference cpu/arcs_simplification.hxx:310-504) and the
  thresholded union-find (cpu/persistence.hxx:90-105): ridges, valleys and both offset arrays
  equal an independent list-based restatement (tests/arc_reference.py: explicit concatenation with
  reversal, no DAG, no path compression) on fields with 65-356 contractions, primal and dual,
  thresholds {0, 0.05, 0.2}.
* Tie rules of the gradient (reference cpu/cell_compare.hxx:9-31, gradient.hxx:55-112): the
  lower-star pairing is a function of the (value, pixel id) order only, so replacing values by
  their rank under that order leaves the gradient, critical cells and labels unchanged on heavily
  tied fields.
* Arc layout (reference cpu/arcs_geometry.hxx:56-166, cpu/workspace.hxx:266-285): arc k of a
  saddle starts at its k-th valid coface / vertex, arc 1 is empty for a saddle edge with one
  vertex, consecutive cells are adjacent or equal, the arc offsets split each saddle's cells.
* MSComplex helpers (reference dmsc/dmsc.py:87-156) against closed forms of the cell codec.
"""

import numpy as np
import pytest
import torch

import oracle
from dmsc_amd.dmsc import MSComplex
from dmsc_amd.synthetic import gaussian_mixture
from tests.arc_reference import simplified_arcs


def field(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "gauss":
        return gaussian_mixture(48, 44, 10, (3, 10), 0.08, seed=seed).numpy()
    f = rng.random((40, 36)).astype(np.float32)
    if kind == "quant":  # 8 levels: ties everywhere
        f = (np.floor(f * 8) / 8).astype(np.float32)
    return f


@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("thr", [0.0, 0.05, 0.2])
@pytest.mark.parametrize("kind,seed", [("noise", 7), ("quant", 8), ("gauss", 11)])
def test_arc_merging_matches_independent_lists(kind, seed, thr, dual):
    f = field(kind, seed)
    sad, res, merges = simplified_arcs(f, thr, dual)
    assert merges >= 60  # every case contracts many arcs (up to a few survivors for the smooth field)
    out = oracle.compute_dmsc(f, thr, is_dual=dual)
    assert np.array_equal(sad, out["sad_pts"])
    # primal: ridges are face arcs, valleys vertex arcs; the dual swaps them
    ridge, valley = (res["verts"], res["faces"]) if dual else (res["faces"], res["verts"])
    for got, names in ((ridge, ("ridges", "ridge_offsets", "ridge_arc_offsets")),
                       (valley, ("valleys", "valley_offsets", "valley_arc_offsets"))):
        for a, name in zip(got, names):
            assert np.array_equal(a, out[name]), name


def test_independent_lists_detect_a_missing_reversal():
    """The check is sharp: merging without reversing the dead arc changes the ridges."""
    import tests.arc_reference as ar

    f = field("gauss", 11)
    good = ar.simplified_arcs(f, 0.05, False)[1]["faces"][0]
    orig = ar._merge

    def no_reverse(raw, cancels, side, merges):
        flipped = {k: ([(c[::-1], t) for c, t in v]) for k, v in raw.items()}
        return orig(flipped, cancels, side, merges)

    ar._merge = no_reverse
    try:
        bad = ar.simplified_arcs(f, 0.05, False)[1]["faces"][0]
    finally:
        ar._merge = orig
    assert not np.array_equal(good, bad)


def rank_field(f):
    """Values replaced by their rank under (value, pixel id): a tie-free field with the same order."""
    flat = f.reshape(-1)
    order = np.lexsort((np.arange(flat.size), flat))
    r = np.empty(flat.size, np.float32)
    r[order] = np.arange(flat.size, dtype=np.float32)
    return r.reshape(f.shape)


@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("levels,seed", [(2, 1), (3, 2), (5, 3), (16, 4), (256, 5)])
def test_gradient_depends_only_on_the_value_pixel_order(levels, seed, dual):
    """The lower-star pairing compares vertices by (value, pixel id) only, so the gradient, the
    critical cells and the V-path labels of the raw complex do not change. (The saddle sweep is
    not rank-invariant: it breaks value ties by the saddle's own cell id, which a rank field
    turns into a value order, so sad_pts order and pairs are left to the other pins.)"""
    rng = np.random.default_rng(seed)
    H, W = rng.integers(9, 31, size=2)
    f = (rng.integers(0, levels, size=(H, W)) / levels).astype(np.float32)
    r = rank_field(f)
    assert np.array_equal(oracle.gradient(f, dual), oracle.gradient(r, dual))
    a = oracle.compute_dmsc(f, -1.0, is_dual=dual)
    b = oracle.compute_dmsc(r, -1.0, is_dual=dual)
    for k in ("max_pts", "min_pts", "grad_indices", "peaks", "basins"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(np.sort(a["sad_pts"]), np.sort(b["sad_pts"]))


def test_rank_invariance_is_sharp():
    """Ranking by (value, transposed pixel id) instead changes the gradient of a tied field."""
    rng = np.random.default_rng(9)
    f = (rng.integers(0, 3, size=(12, 14)) / 3).astype(np.float32)
    flat = f.reshape(-1)
    tid = (np.arange(flat.size) % 14) * 12 + np.arange(flat.size) // 14
    order = np.lexsort((tid, flat))
    r = np.empty(flat.size, np.float32)
    r[order] = np.arange(flat.size, dtype=np.float32)
    assert not np.array_equal(oracle.gradient(f, False), oracle.gradient(r.reshape(f.shape), False))


def starts(s, W, H, faces):
    Nx = W + 1
    t, q = s % 4, s // 4
    y, x = q // Nx, q % Nx
    if faces:  # both cofaces of an edge always exist on the (H+1) x (W+1) corner grid
        return [4 * q + 3, 4 * (q + Nx) + 3 if t == 1 else 4 * (q + 1) + 3]
    if t == 1:
        cand = [(y, x - 1) if x >= 1 else None, (y, x) if x < W else None]
    else:
        cand = [(y - 1, x) if y >= 1 else None, (y, x) if y < H else None]
    return [4 * (a * Nx + b) for a, b in (c for c in cand if c)]


def slot_yx(c, Nx):
    return (c // 4) // Nx, (c // 4) % Nx


@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("thr", [-1.0, 0.0, 0.05, 0.2])
@pytest.mark.parametrize("seed", range(3))
def test_arc_layout(seed, thr, dual):
    rng = np.random.default_rng(50 + seed)
    H, W = rng.integers(10, 40, size=2)
    f = (gaussian_mixture(H, W, 6, (2, 8), 0.3, seed=seed).numpy() if seed == 2
         else rng.random((H, W)).astype(np.float32))
    out = oracle.compute_dmsc(f, thr, is_dual=dual)
    Nx = W + 1
    n = len(out["sad_pts"])
    checked = 0
    for arcs, off, aoff in (("ridges", "ridge_offsets", "ridge_arc_offsets"),
                            ("valleys", "valley_offsets", "valley_arc_offsets")):
        cells, o, ao = out[arcs], out[off], out[aoff]
        faces = (arcs == "ridges") != dual
        assert len(o) == n + 1 and len(ao) == n and o[0] == 0 and o[-1] == len(cells)
        assert np.all(o[:-1] <= ao) and np.all(ao <= o[1:])
        assert np.all(cells % 4 == (3 if faces else 0))
        for i, s in enumerate(out["sad_pts"]):
            st = starts(int(s), W, H, faces)
            parts = (cells[o[i]:ao[i]], cells[ao[i]:o[i + 1]])
            for k, arc in enumerate(parts):
                if k >= len(st):
                    assert len(arc) == 0  # a saddle edge with one vertex has one vertex arc
                    continue
                if faces:
                    assert len(arc) > 0  # face paths always end at a critical face
                if len(arc):
                    assert arc[0] == st[k]
                    yx = np.array([slot_yx(int(c), Nx) for c in arc])
                    step = np.abs(np.diff(yx, axis=0)).sum(axis=1)
                    assert np.all(step <= 1)  # adjacent or equal
                    checked += 1
    assert checked > 0


@pytest.mark.parametrize("dual", [False, True])
def test_mscomplex_helpers_against_the_cell_codec(dual):
    f = gaussian_mixture(37, 29, 8, (2, 7), 0.05, seed=3).numpy()
    out = oracle.extract(f, 0.05, True, dual, True, True, True, True)
    ms = MSComplex(*[torch.from_numpy(np.asarray(a)) for a in out])
    H, W = f.shape
    Nx = W + 1
    # every valid cell id of the complex: id = type + 4 (y Nx + x)
    ty, yy, xx = np.meshgrid(np.arange(4), np.arange(H + 1), np.arange(W + 1), indexing="ij")
    ids = ty + 4 * (yy * Nx + xx)
    valid = ((ty == 0) & (yy < H) & (xx < W)) | ((ty == 1) & (yy < H)) | ((ty == 2) & (xx < W)) | (ty == 3)
    ids, ty, yy, xx = ids[valid], ty[valid], yy[valid], xx[valid]
    c = torch.from_numpy(ids.astype(np.int32))
    plain = ms.to_coordinates_yx(c, staggered=False)
    assert plain.dtype == torch.int32
    assert np.array_equal(plain.numpy(), np.stack([yy, xx], -1))
    # staggered = centre of the cell's boundary pixels (reference cell_complex.hxx:13-19 table):
    # vertex (y, x); H-edge (y, x-1..x); V-edge (y-1..y, x); face (y-1..y, x-1..x)
    centre = np.stack([yy - 0.5 * ((ty == 2) | (ty == 3)), xx - 0.5 * ((ty == 1) | (ty == 3))], -1)
    st = ms.to_coordinates_yx(c)
    assert st.dtype == torch.float32 and np.array_equal(st.numpy(), centre.astype(np.float32))
    # every gradient pair is a facet relation: centres 0.5 apart along exactly one axis
    pw = out[10]
    a = np.flatnonzero(pw >= 0)
    d = np.abs(ms.to_coordinates_yx(torch.from_numpy(a.astype(np.int32))).numpy()
               - ms.to_coordinates_yx(torch.from_numpy(pw[a])).numpy())
    assert np.all(np.sort(d, axis=1) == np.array([0.0, 0.5], np.float32))
    for name in ("max_pts", "min_pts", "sad_pts"):
        assert torch.equal(getattr(ms, "offset_" + name + "_yx"), ms.to_coordinates_yx(getattr(ms, name)))
    assert ms.to_coordinates_yx(torch.empty(0, dtype=torch.int32)).shape == (0, 2)
    assert ms.to_coordinates_yx(torch.empty(0, dtype=torch.int32), staggered=False).dtype == torch.int32
    # CSR slices: saddle arcs tile the flat arrays in order, and the split gives arc 0 then arc 1
    for flat, off, aoff, get in ((ms.ridges, ms.ridge_offsets, ms.ridge_arc_offsets, ms.get_ridge),
                                 (ms.valleys, ms.valley_offsets, ms.valley_arc_offsets, ms.get_valley)):
        n = len(ms.sad_pts)
        assert n > 0
        assert torch.equal(torch.cat([get(i) for i in range(n)]), flat)
        for i in range(n):
            a0, a1 = get(i, split_arcs=True)
            assert len(a0) == int(aoff[i] - off[i]) and torch.equal(torch.cat([a0, a1]), get(i))

"""CPU side of the persistence-loss checks: loss and gradient assembled in float64 from the oracle's
pairs (the oracle restates the reference CPU path), and tie-free synthetic fields."""

import numpy as np

import oracle

PAIR_SLOTS = (3, 7, 9, 2, 6, 8, 1)  # sad_pts, p_min, ppairs_min, min_pts, p_max, ppairs_max, max_pts


def tie_free_field(h, w, seed):
    """Gaussian bumps plus noise, every value distinct (checked)."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    f = np.zeros((h, w))
    for _ in range(6):
        cy, cx, s, a = rng.uniform(0, h), rng.uniform(0, w), rng.uniform(2, 6), rng.uniform(-1, 1)
        f += a * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    f += rng.uniform(-0.05, 0.05, size=(h, w))
    f = (f - f.min()) / (f.max() - f.min())
    f = f.astype(np.float32)
    assert np.unique(f).size == f.size, "field has ties"
    return f


def value_pixel(field, cell, is_dual):
    """Pixel carrying the value of a cell: highest vertex (primal) / lowest (dual); tie-free input."""
    h, w = field.shape
    nx = w + 1
    t, q = cell & 3, cell >> 2
    y, x = q // nx, q % nx
    if t == 0:
        cand = [(y, x)]
    elif t == 1:
        cand = [(y, x - 1), (y, x)]
    elif t == 2:
        cand = [(y - 1, x), (y, x)]
    else:
        cand = [(y - 1, x - 1), (y - 1, x), (y, x - 1), (y, x)]
    cand = [(a, b) for a, b in cand if 0 <= a < h and 0 <= b < w]
    vals = [field[a, b] for a, b in cand]
    a, b = cand[int(np.argmin(vals) if is_dual else np.argmax(vals))]
    return a * w + b


def oracle_pairs(field, is_dual):
    out = oracle.extract(np.ascontiguousarray(field), -1.0, False, is_dual, False, False, False, False)
    return [np.asarray(out[k]) for k in PAIR_SLOTS]


def value_pixels(field, cells, is_dual):
    """value_pixel of many cells at once: the highest vertex by (value, pixel id) of each cell, the
    lowest in the dual (the reference's cell value rule, cpu/cell_compare.hxx:33-87)."""
    h, w = field.shape
    ff = field.reshape(-1)
    cells = np.asarray(cells, dtype=np.int64)
    t, q = cells & 3, cells >> 2
    y, x = q // (w + 1), q % (w + 1)
    best = np.full(cells.shape, -1, dtype=np.int64)
    for dy, dx, ok in ((0, 0, t == 0), (0, -1, t == 1), (0, 0, t == 1), (-1, 0, t == 2), (0, 0, t == 2),
                       (-1, -1, t == 3), (-1, 0, t == 3), (0, -1, t == 3), (0, 0, t == 3)):
        yy, xx = y + dy, x + dx
        v = ok & (yy >= 0) & (yy < h) & (xx >= 0) & (xx < w)
        cand = np.where(v, yy * w + xx, 0)
        b = np.maximum(best, 0)
        if is_dual:
            better = (ff[cand] < ff[b]) | ((ff[cand] == ff[b]) & (cand < best))
        else:
            better = (ff[cand] > ff[b]) | ((ff[cand] == ff[b]) & (cand > best))
        best = np.where(v & ((best < 0) | better), cand, best)
    return best


def loss_and_grad(field, max_pers, is_dual):
    """(loss, d loss / d field) in float64 from the oracle's pairs: a min-side pair adds +2 pers at
    the saddle's value pixel and -2 pers at the minimum's, a max-side pair +2 pers at the maximum's
    and -2 pers at the saddle's."""
    sad, pmin, ppmin, mins, pmax, ppmax, maxs = oracle_pairs(field, is_dual)
    g = np.zeros(field.size)
    sp = value_pixels(field, sad, is_dual)
    loss = 0.0
    for p, pp, ext, sgn in ((pmin, ppmin, mins, 1.0), (pmax, ppmax, maxs, -1.0)):
        m = (pp >= 0) & (p >= 0) & (p < max_pers)
        pv = p[m].astype(np.float64)
        loss += float(np.sum(pv * pv))
        np.add.at(g, sp[m], sgn * 2 * pv)
        np.add.at(g, value_pixels(field, ext[pp[m]], is_dual), -sgn * 2 * pv)
    return loss, g.reshape(field.shape)


def pairing_signature(field, max_pers, is_dual):
    """What must not change between f - eps and f + eps for a finite difference to be valid."""
    sad, pmin, ppmin, mins, pmax, ppmax, maxs = oracle_pairs(field, is_dual)
    m1 = (pmin >= 0) & (pmin < max_pers)
    m2 = (pmax >= 0) & (pmax < max_pers)
    return tuple(a.tobytes() for a in (sad, ppmin, mins, ppmax, maxs, m1, m2))

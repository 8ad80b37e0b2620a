"""The active-tile rule of the scan launch before an inner-loop round.

In that launch the waves of the face-list tiles lower cells (the inner
admissibility loop's lowering, src/reconstruct.f90:146-171 with
src/gridBlock.f90:194-235): rc |= cross(inadmissible), rcF |= diamond,
concurrently with every wave reading rc / rcF over its tile's footprint
grown by two cells.  The next iteration's update and post-detect lowering
(src/detect.f90:60-72,219-234) touch cells within two cells of any rcF
cell, so the list must hold every tile whose grown footprint holds a
flag *after* the lowering, whatever order the waves ran in.

This test takes the worst order (every wave reads the flags as they were
before the lowering) and checks, by brute force over random grids, that
the kernel's rule (hmhd_kernels.hip, k_tile_scan: flags read, plus the
per-tile bad_bits words of the 3 x 3 tiles around, within two and, before
an inner-loop round, within four cells of the facing edge) lists a
superset of the exact set.  The rule without the within-four bits misses
a tile (planted check that the test sees the gap).
"""
import numpy as np

T = 16  # kCellT

W2 = {"W": 2, "E": 4, "S": 8, "N": 16}
W4 = {"W": 32, "E": 64, "S": 128, "N": 256}


def bad_bits(rx, ry):
    b = 1
    b |= W2["W"] if rx <= 1 else 0
    b |= W2["E"] if rx >= T - 2 else 0
    b |= W2["S"] if ry <= 1 else 0
    b |= W2["N"] if ry >= T - 2 else 0
    b |= W4["W"] if rx <= 3 else 0
    b |= W4["E"] if rx >= T - 4 else 0
    b |= W4["S"] if ry <= 3 else 0
    b |= W4["N"] if ry >= T - 4 else 0
    return b


def tile_words(bad_sides, mx, my):
    """tbad as the face kernel leaves it: the side of cell (i, j) in
    direction d was written by the visit of the tile owning the face; an
    east / north side is owned by the cell east / north of it, and is
    recorded at offset -1 when that cell starts the next tile."""
    ntx, nty = -(-mx // T), -(-my // T)
    words = np.zeros((nty, ntx), dtype=np.int64)
    for (i, j, d) in bad_sides:
        oi, oj = (i + 1, j) if d == "E" else ((i, j + 1) if d == "N" else (i, j))
        if oi >= mx or oj >= my:
            oi, oj = i, j  # the last column / row: the face is the grid's own edge
        tx, ty = oi // T, oj // T
        words[ty, tx] |= bad_bits(i - tx * T, j - ty * T)
    return words


def dilate(mask, shape):
    """OR over the taxicab diamond of radius 2 (clipped to the grid)."""
    my, mx = mask.shape
    out = np.zeros_like(mask)
    for dj in range(-2, 3):
        for di in range(-2, 3):
            if abs(di) + abs(dj) > 2:
                continue
            src = mask[max(0, -dj):my - max(0, dj), max(0, -di):mx - max(0, di)]
            out[max(0, dj):my - max(0, -dj), max(0, di):mx - max(0, -di)] |= src
    return out


def listed(flags, words, lower, use4=True):
    """Tiles the scan lists from the flags it read and the tbad words."""
    my, mx = flags.shape
    ntx, nty = words.shape[1], words.shape[0]
    out = np.zeros((nty, ntx), dtype=bool)
    for ty in range(nty):
        for tx in range(ntx):
            x0, y0 = tx * T - 2, ty * T - 2
            fp = flags[max(0, y0):y0 + T + 4, max(0, x0):x0 + T + 4]
            hit = bool(fp.any())
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    a, b = tx + dx, ty + dy
                    if not (0 <= a < ntx and 0 <= b < nty):
                        continue
                    w = int(words[b, a])
                    n2 = 1 | (W2["E"] if dx < 0 else 0) | (W2["W"] if dx > 0 else 0) | \
                        (W2["N"] if dy < 0 else 0) | (W2["S"] if dy > 0 else 0)
                    n4 = 1 | (W4["E"] if dx < 0 else 0) | (W4["W"] if dx > 0 else 0) | \
                        (W4["N"] if dy < 0 else 0) | (W4["S"] if dy > 0 else 0)
                    hit |= (w & n2) == n2
                    hit |= use4 and lower and (w & n4) == n4
            out[ty, tx] = hit
    return out


def exact(flags, ntx, nty):
    my, mx = flags.shape
    out = np.zeros((nty, ntx), dtype=bool)
    for ty in range(nty):
        for tx in range(ntx):
            x0, y0 = tx * T - 2, ty * T - 2
            out[ty, tx] = bool(flags[max(0, y0):y0 + T + 4, max(0, x0):x0 + T + 4].any())
    return out


def random_case(rng, mx, my):
    # sparse flags left by an earlier post-detect lowering: diamonds of low cells
    low = rng.random((my, mx)) < 0.002
    rcf = dilate(low, low.shape)
    rc = low.copy()
    # a few inadmissible cells, each with one or more inadmissible sides
    bad = np.zeros((my, mx), dtype=bool)
    sides = []
    for _ in range(rng.integers(1, 6)):
        i, j = int(rng.integers(0, mx)), int(rng.integers(0, my))
        bad[j, i] = True
        for d in rng.choice(list("NSEW"), size=int(rng.integers(1, 5)), replace=False):
            sides.append((i, j, d))
    return rc, rcf, bad, sides


def test_rule_covers_the_lowering_in_any_wave_order():
    rng = np.random.default_rng(12345)
    for trial in range(300):
        mx, my = int(rng.integers(20, 70)), int(rng.integers(20, 70))
        rc, rcf, bad, sides = random_case(rng, mx, my)
        words = tile_words(sides, mx, my)
        before = rc | rcf
        # the lowering: rcF |= diamond(bad), rc |= cross(bad) (within the diamond)
        after = before | dilate(bad, bad.shape)
        need = exact(after, words.shape[1], words.shape[0])
        got = listed(before, words, lower=True)
        missing = need & ~got
        assert not missing.any(), f"trial {trial}: tiles {np.argwhere(missing).tolist()} missed"


def test_rule_without_the_far_bits_misses_a_tile():
    """The advisor's geometry: an inadmissible cell X three cells west of a
    tile edge, nothing else flagged.  The lowering raises rcF on the cell L
    two cells east of X, one cell west of the edge; the tile east of the
    edge then has L within its grown footprint.  Read before the lowering,
    only the within-four bits list it."""
    mx = my = 48
    rc = np.zeros((my, mx), dtype=bool)
    rcf = np.zeros((my, mx), dtype=bool)
    X = (T - 4, 20)  # rx = 12 in tile 0: within four of its east edge, not two
    sides = [(X[0], X[1], "W")]
    bad = np.zeros((my, mx), dtype=bool)
    bad[X[1], X[0]] = True
    words = tile_words(sides, mx, my)
    after = rc | rcf | dilate(bad, bad.shape)
    need = exact(after, words.shape[1], words.shape[0])
    assert need[1, 1], "the east tile needs listing"
    old = listed(rc | rcf, words, lower=True, use4=False)
    new = listed(rc | rcf, words, lower=True)
    assert not old[1, 1], "the rule without the far bits should miss it"
    assert new[1, 1] and not (need & ~new).any()

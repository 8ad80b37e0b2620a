"""The active-tile rules of the launch that ends a MOOD iteration.

Before an inner-loop round:

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
the kernel's rule (hmhd_kernels.hip, k_end_iter: flags read, plus the
per-tile bad_bits words of the 3 x 3 tiles around, within two and, before
an inner-loop round, within four cells of the facing edge) lists a
superset of the exact set.  The rule without the within-four bits misses
a tile (planted check that the test sees the gap).

After a ready pass the same launch lowers after detect (rc = cross(low),
rcF = diamond(low)) and syncs the periodic units while its scan waves run,
so the scan reads `low` instead: a tile is listed when a low cell lies
within four cells of it, or of the images of its band cells.  Brute force:
flags = diamond(low), OR-ed over each periodic unit (a guard cell and its
images N cells away), then the exact set of tiles whose grown footprint
holds a flag must be covered by the rule.
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


def synced_flags(low, s, nx, ny, periodic):
    """diamond(low) after the periodic MOOD sync: every band cell takes the
    OR over its unit (itself and its images N cells away)."""
    f = dilate(low, low.shape)
    if not periodic:
        return f
    my, mx = f.shape
    out = f.copy()
    for j in range(my):
        for i in range(mx):
            bc, br = i < 2 * s or i >= nx, j < 2 * s or j >= ny
            if not (bc or br):
                continue
            pi = (i + nx if i < 2 * s else i - nx) if bc else i
            pj = (j + ny if j < 2 * s else j - ny) if br else j
            out[j, i] = f[j, i] | f[j, pi] | f[pj, i] | f[pj, pi]
    return out


def ready_rule(low, s, nx, ny, periodic):
    """k_end_iter's rule after a ready pass (low_near and its band shifts)."""
    my, mx = low.shape
    ntx, nty = -(-mx // T), -(-my // T)
    out = np.zeros((nty, ntx), dtype=bool)

    def near(X0, Y0):
        xa, ya = max(0, X0 - 4), max(0, Y0 - 4)
        return bool(low[ya:max(0, Y0 + T + 4), xa:max(0, X0 + T + 4)].any())

    for ty in range(nty):
        for tx in range(ntx):
            X0, Y0 = tx * T, ty * T
            hit = near(X0, Y0)
            if periodic and not hit:
                sx = [0] + ([nx] if X0 - 2 < 2 * s else []) + ([-nx] if X0 + T + 1 >= nx else [])
                sy = [0] + ([ny] if Y0 - 2 < 2 * s else []) + ([-ny] if Y0 + T + 1 >= ny else [])
                hit = any(near(X0 + a, Y0 + b) for a in sx for b in sy if (a, b) != (0, 0))
            out[ty, tx] = hit
    return out


def test_ready_rule_covers_the_synced_flags():
    rng = np.random.default_rng(777)
    for trial in range(120):
        R = int(rng.integers(1, 4))
        s = R + 2
        nx, ny = int(rng.integers(2 * s, 60)), int(rng.integers(2 * s, 60))
        mx, my = nx + 2 * s, ny + 2 * s
        periodic = bool(rng.integers(2))
        low = np.zeros((my, mx), dtype=bool)
        inner = rng.random((ny, nx)) < rng.choice([0.002, 0.01, 0.05])
        low[s:s + ny, s:s + nx] = inner
        need = exact(synced_flags(low, s, nx, ny, periodic), -(-mx // T), -(-my // T))
        got = ready_rule(low, s, nx, ny, periodic)
        missing = need & ~got
        assert not missing.any(), f"trial {trial} (R={R}, {nx}x{ny}, periodic={periodic}): {np.argwhere(missing).tolist()}"


def test_ready_rule_needs_the_band_images():
    """A low cell next to the east edge of a periodic grid flags, through
    the sync, guard cells on the west side: the tile there is listed only
    through the band shift."""
    s, nx, ny = 4, 40, 40
    mx, my = nx + 2 * s, ny + 2 * s
    low = np.zeros((my, mx), dtype=bool)
    low[20, nx + s - 1] = True  # last interior column; its image is guard column s - 1
    need = exact(synced_flags(low, s, nx, ny, True), -(-mx // T), -(-my // T))
    got = ready_rule(low, s, nx, ny, True)
    assert need[1, 0] and got[1, 0]
    no_shift = np.zeros_like(got)
    for ty in range(got.shape[0]):
        for tx in range(got.shape[1]):
            X0, Y0 = tx * T, ty * T
            no_shift[ty, tx] = bool(low[max(0, Y0 - 4):Y0 + T + 4, max(0, X0 - 4):X0 + T + 4].any())
    assert not no_shift[1, 0], "without the band shifts the west tile would be missed"

"""Integer identities the sphere-fill kernel relies on (bm_tiles.hip), checked
on the CPU with numpy emulations of the instructions.

The dot-product form of the candidate test: for a voxel v of a fill block
(coordinates 0..15) and a candidate centre c relative to the block origin with
radius at most 64 (c in [-64, 79]),
    |v - c|^2 <= r^2  <=>  dot4(2v, c + 64) - (|v|^2 + 128 (vx + vy + vz)) >= K
with K = |c|^2 - r^2 stored as the low 16 bits of the key r^2 << 16 | K, and
dot4 the unsigned 4 x u8 dot product with a 32-bit wrapping accumulator."""
import numpy as np


def _udot4(a, b, acc):
    s = acc.astype(np.int64)
    for i in range(4):
        s = s + ((a >> (8 * i)) & 255).astype(np.int64) * ((b >> (8 * i)) & 255).astype(np.int64)
    return s % 2**32


def _as_i32(u):
    return np.where(u >= 2**31, u - 2**32, u)


def _as_i16(u):
    return np.where(u >= 2**15, u - 2**16, u)


def _dot_form(v, c, r2):
    v2pk = (2 * v[:, 0]) | ((2 * v[:, 1]) << 8) | ((2 * v[:, 2]) << 16)
    nega = (-(v ** 2).sum(1) - 128 * v.sum(1)) % 2**32
    pk = (c[:, 0] + 64) | ((c[:, 1] + 64) << 8) | ((c[:, 2] + 64) << 16)
    k = (c ** 2).sum(1) - r2
    key = (r2 << 16) | (k & 0xFFFF)
    return _as_i32(_udot4(v2pk, pk, nega)) >= _as_i16(key & 0xFFFF), key


def test_dot_form_equals_distance_test():
    rng = np.random.default_rng(0)
    n = 400_000
    v = rng.integers(0, 16, size=(n, 3))
    c = rng.integers(-64, 80, size=(n, 3))
    d2 = ((v - c) ** 2).sum(1)
    # half random radii, half at the boundary |v - c|^2 - 1 .. + 1
    r2 = np.where(rng.random(n) < 0.5, rng.integers(0, 64 * 64 + 1, size=n),
                  np.clip(d2 + rng.integers(-1, 2, size=n), 0, 64 * 64))
    got, key = _dot_form(v, c, r2)
    np.testing.assert_array_equal(got, d2 <= r2)
    np.testing.assert_array_equal(key >> 16, r2)  # keys order as their r^2


def test_dot_form_extremes():
    # every corner of the voxel and centre ranges at the largest radius
    vs = np.array([[x, y, z] for x in (0, 15) for y in (0, 15) for z in (0, 15)])
    cs = np.array([[x, y, z] for x in (-64, 79) for y in (-64, 79) for z in (-64, 79)])
    v = np.repeat(vs, len(cs), axis=0)
    c = np.tile(cs, (len(vs), 1))
    for r2 in (0, 1, 4095, 4096):
        rr = np.full(len(v), r2)
        got, _ = _dot_form(v, c, rr)
        np.testing.assert_array_equal(got, ((v - c) ** 2).sum(1) <= rr)
    # K = |c|^2 - r^2 spans [-4096, 3 * 79^2]: inside 16 signed bits
    assert -2**15 <= -4096 and (cs ** 2).sum(1).max() < 2**15


def test_never_covered_key():
    """Lanes after the relevant ones carry position 0 and key 0x7FFF: no voxel
    of a block reaches K = 32767 (2 v.c - |v|^2 < 2^15)."""
    v = np.array([[x, y, z] for x in range(16) for y in range(16) for z in range(16)])
    v2pk = (2 * v[:, 0]) | ((2 * v[:, 1]) << 8) | ((2 * v[:, 2]) << 16)
    nega = (-(v ** 2).sum(1) - 128 * v.sum(1)) % 2**32
    res = _as_i32(_udot4(v2pk, np.zeros(len(v), np.int64), nega))
    assert (res < 0x7FFF).all()
    # the largest 2 v.c - |v|^2 over the whole range stays below 2^15 as well
    assert (2 * 15 * 79 * 3) < 2**15


# One-word staged candidates of the fill (radius R <= 45): r^2 << 21 plus the
# block-relative centre c + R in three 7-bit fields, built as
#   word = (r^2 << 21) + local7(entry) + tc,  tc = sum over axes of (corner + R) 2^(7 i)
# where the tile corner (relative to the block origin) + R may be negative; the
# placement decodes the fields back to the packed 10-bit position c + 512.
def _local7(en):
    return (en & 7) | ((en & 0x38) << 4) | ((en & 0x1C0) << 8)


def test_compact_staging_word_round_trip():
    rng = np.random.default_rng(1)
    n = 300_000
    for txn in (1, 2):
        radius = rng.integers(1, 46, size=n)
        k = (radius + 7) // 8
        # tile of the reach box (ix, iy, iz) and the entry's local coordinates
        tile = np.stack([rng.integers(0, 2 * k + txn), rng.integers(0, 2 * k + 1),
                         rng.integers(0, 2 * k + 1)], axis=1)
        loc = rng.integers(0, 8, size=(n, 3))
        corner = (tile - k[:, None]) * 8  # relative to the block origin
        c = corner + loc
        hi = np.array([8 * txn - 1, 7, 7])
        kept = ((c >= -radius[:, None]) & (c <= hi + radius[:, None])).all(1)
        r2 = rng.integers(1, radius ** 2 + 1)
        en = loc[:, 0] | (loc[:, 1] << 3) | (loc[:, 2] << 6) | (r2 << 9)
        tc = ((corner[:, 0] + radius) + (corner[:, 1] + radius) * 128
              + (corner[:, 2] + radius) * 16384) % 2**32
        word = ((r2 << 21) + _local7(en) + tc) % 2**32
        off = 512 - radius
        pos = ((word & 127) + off) | ((((word >> 7) & 127) + off) << 10) | \
              ((((word >> 14) & 127) + off) << 20)
        want = (c[:, 0] + 512) | ((c[:, 1] + 512) << 10) | ((c[:, 2] + 512) << 20)
        assert kept.sum() > n // 4
        np.testing.assert_array_equal(pos[kept], want[kept])
        np.testing.assert_array_equal((word >> 21)[kept], r2[kept])
        assert (c[kept] + radius[kept, None]).max() <= 127 and (c[kept] + radius[kept, None]).min() >= 0

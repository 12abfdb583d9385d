"""Byte-parallel identities the EDT X pass relies on (k_edt_x in
bm_kernels.hip), checked on the CPU with numpy emulations of its 32-bit
operations, and the 16-bit saturation of the Y/Z window search.

zero_bytes(x) = ~(((x & 0x7F7F7F7F) + 0x7F7F7F7F) | x | 0x7F7F7F7F) holds
0x80 in each zero byte of x and 0 elsewhere; byte_msbs(m) gathers bits 7, 15,
23, 31 of such a mask into bits 0..3.  A voxel is a site (not selected) unless
its ROI byte is nonzero and its data byte is nonzero exactly when the phase
wants nonzero (src/volume.rs:84-105)."""
import numpy as np

M32 = 0xFFFFFFFF


def zero_bytes(x):
    x = x.astype(np.uint64)
    return (~(((x & 0x7F7F7F7F) + 0x7F7F7F7F) | x | 0x7F7F7F7F)) & M32


def byte_msbs(m):
    return ((m >> 7) | (m >> 14) | (m >> 21) | (m >> 28)) & 0xF


def _words(rng, n):
    x = rng.integers(0, 2**32, n, dtype=np.uint64)
    x[::3] &= 0xFF00FF00
    x[1::5] &= 0x00FF00FF
    x[2::7] = 0
    x[3::11] |= 0x80808080
    x[4::13] = 0x01010101
    return x


def test_zero_bytes_marks_exactly_the_zero_bytes():
    x = _words(np.random.default_rng(0), 300_000)
    z = zero_bytes(x)
    for k in range(4):
        byte = (x >> (8 * k)) & 0xFF
        np.testing.assert_array_equal((z >> (8 * k + 7)) & 1, (byte == 0).astype(np.uint64))
        assert not ((z >> (8 * k)) & 0x7F).any()


def test_site_bits_equal_the_per_voxel_rule():
    rng = np.random.default_rng(1)
    v, r = _words(rng, 200_000), _words(rng, 200_000)
    for want in (0, 1):
        zv = zero_bytes(v)
        sel = (~zero_bytes(r) & M32) & ((~zv & M32) if want else zv) & 0x80808080
        bits = byte_msbs(sel ^ 0x80808080)
        count = np.array([bin(int(s)).count("1") for s in sel[:2000]])
        ref_bits = np.zeros_like(v)
        ref_count = np.zeros(v.size, np.int64)
        for k in range(4):
            vb, rb = (v >> (8 * k)) & 0xFF, (r >> (8 * k)) & 0xFF
            selected = (rb != 0) & ((vb != 0) == bool(want))
            ref_bits |= (~selected).astype(np.uint64) << k
            ref_count += selected
        np.testing.assert_array_equal(bits, ref_bits)
        np.testing.assert_array_equal(count, ref_count[:2000])


def test_planted_carry_is_caught():
    """Without the 0x7F7F7F7F mask on the addend's input, carries cross bytes:
    the zero-byte test above would fail."""
    x = _words(np.random.default_rng(2), 100_000)
    bad = (~((x + 0x7F7F7F7F) | x | 0x7F7F7F7F)) & M32
    assert (bad != zero_bytes(x)).any()


def test_saturated_pairs_keep_minima_below_2_16():
    """min over candidates of sat16(f + o^2) equals sat16 of the exact minimum
    (saturation is monotone), so rows whose distance is below 0xFFFF are exact
    and the others read 0xFFFF, which the Y/Z passes search again in 32 bits."""
    rng = np.random.default_rng(3)
    f = rng.integers(0, 2**31, (20_000, 16), dtype=np.int64)
    f[:, ::3] = rng.integers(0, 70_000, (20_000, 6))
    o2 = (np.arange(16) ** 2)[None, :]
    exact = (f + o2).min(1)
    sat = np.minimum(np.minimum(f, 0xFFFF) + o2, 0xFFFF).min(1)
    np.testing.assert_array_equal(sat, np.minimum(exact, 0xFFFF))
    assert ((exact >= 0xFFFF) == (sat == 0xFFFF)).all()

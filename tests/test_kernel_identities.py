"""The arithmetic shortcuts the HIP kernels take, checked on the CPU against the
plain definitions they replace (brute force or exhaustive where the domain is
small).  The kernels themselves are checked bit for bit against the oracle by
the GPU parity tests; these pin the reasoning each shortcut rests on, so a
slip there shows up without a GPU.
"""
import numpy as np


def test_meta8_context_class_formula():
    """przm_entropy.hip ctx_meta: max((b + 1) / 2, b - 2) + (p >= 192) with b the bit length
    of p, against the class table of vans_meta.go:45-68 for every byte."""
    def table(p):
        return 0 if p == 0 else 1 if p < 4 else 2 if p < 16 else 3 if p < 32 else \
            4 if p < 64 else 5 if p < 128 else 6 if p < 192 else 7
    for p in range(256):
        b = (2 * p + 1).bit_length() - 1
        assert max((b + 1) >> 1, b - 2) + (1 if p >= 192 else 0) == table(p), p


def _fmaf32(a, b, c):
    """fmaf in float32: the float64 product of two floats is exact (48 bits), so one float64
    add and one rounding to float32 equal the fused operation (the addend 2^-9 and products
    below 2^13 keep the float64 sum exact)."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def test_biased_reciprocal_divmod():
    """przm_entropy.hip divmod: trunc(fma(x, rcp(f), 2^-9)) is floor(x / f) or one more for
    x <= 4096 f - 1, so one downward correction is exact.  Checked for every f with the
    reciprocal rounded correctly and one ulp either way (the device rcp bound), over x at
    the ends of each quotient step and a random sample."""
    rng = np.random.default_rng(1)
    bias = np.float32(2.0 ** -9)
    for f in range(1, 4097):
        r0 = np.float32(1.0) / np.float32(f)
        recips = (np.nextafter(r0, np.float32(0)), r0, np.nextafter(r0, np.float32(1)))
        q = rng.integers(0, 4096, 64)
        x = np.concatenate([q * f, q * f + f - 1, rng.integers(0, 4096 * f, 64), [0, 4096 * f - 1]])
        x = x.astype(np.uint64)
        for r in recips:
            est = (np.float64(np.float32(x.astype(np.float64))) * np.float64(r) + np.float64(bias))
            qq = est.astype(np.float32).astype(np.int64)
            rr = x.astype(np.int64) - qq * f
            low = rr < 0
            qq = np.where(low, qq - 1, qq)
            rr = np.where(low, rr + f, rr)
            assert np.array_equal(qq, (x // f).astype(np.int64)), f
            assert np.array_equal(rr, (x % f).astype(np.int64)), f


def test_identity_entry_step():
    """count_step / rans_step: a step that is not live codes the entry freq 4096, cdf 0.  For
    every state below 2^24 that step renormalises nothing (x <= 4096 * 4096 - 1) and the
    biased-reciprocal divmod gives back x itself, with the device rcp exact or one ulp off."""
    x = np.arange(1 << 24, dtype=np.int64)
    f = 4096
    assert (x <= 4096 * f - 1).all()
    bias = np.float64(2.0 ** -9)
    r0 = np.float32(1.0) / np.float32(f)
    for r in (np.nextafter(r0, np.float32(0)), r0, np.nextafter(r0, np.float32(1))):
        qq = (x.astype(np.float32).astype(np.float64) * np.float64(r) + bias).astype(np.float32).astype(np.int64)
        rr = x - qq * f
        low = rr < 0
        qq = np.where(low, qq - 1, qq)
        rr = np.where(low, rr + f, rr)
        assert np.array_equal(qq * 4096 + 0 + rr, x)


def _sad0(x):
    """v_sad_u8(x, 0, 0): the sum of the four bytes of x"""
    return sum((x >> (8 * i)) & 0xFF for i in range(4))


def test_quad_byte_offsets():
    """rans_step: with each lane's byte count (0..2) as byte j of R (the sum over the quad), lane
    j's offset sad_u8(R & above_j, 0, 0) with above_j = (0xFFFFFF00 << 8j) mod 2^32 is the sum of
    the counts of lanes above j, and sad_u8(R, 0, 0) the step total, for all 81 combinations."""
    for c0 in range(3):
        for c1 in range(3):
            for c2 in range(3):
                for c3 in range(3):
                    ne = [c0, c1, c2, c3]
                    R = sum(n << (8 * j) for j, n in enumerate(ne))
                    assert _sad0(R) == sum(ne)
                    for j in range(4):
                        above = (0xFFFFFF00 << (8 * j)) & 0xFFFFFFFF
                        assert _sad0(R & above) == sum(ne[j + 1:])


def _ffbl(x):
    return (x & -x).bit_length() - 1 if x else 0xFFFFFFFF


def test_first_difference_bit_index():
    """przm_lz.hip diff12 / eq16: min(ffbl(x0), ffbl(x1) | 32, ffbl(x2) | 64 [, ffbl(x3) | 96])
    of the xor words is the bit index of the first differing bit, 0xFFFFFFFF when equal, so
    its eighth is the common prefix in bytes; against a byte-by-byte compare."""
    rng = np.random.default_rng(2)
    for _ in range(3000):
        a = rng.integers(0, 256, 16, dtype=np.uint8)
        b = a.copy()
        cut = int(rng.integers(0, 17))
        if cut < 16:
            b[cut] ^= np.uint8(1 << int(rng.integers(0, 8)))
            if rng.random() < 0.5:
                b[cut + 1:] = rng.integers(0, 256, 15 - cut, dtype=np.uint8)
        wa = a.view('<u4').astype(np.int64)
        wb = b.view('<u4').astype(np.int64)
        x = [int(wa[k] ^ wb[k]) for k in range(4)]
        t12 = min(_ffbl(x[0]), _ffbl(x[1]) | 32, _ffbl(x[2]) | 64)
        t16 = min(t12, _ffbl(x[3]) | 96)
        lcp = next((k for k in range(16) if a[k] != b[k]), 16)
        assert min(t12 >> 3, 12) == min(lcp, 12)
        assert min(t16 >> 3, 16) == lcp
        if lcp < 12:  # the bit index itself: 8 * lcp + the first differing bit of that byte
            diff = int(a[lcp] ^ b[lcp])
            assert t12 == 8 * lcp + ((diff & -diff).bit_length() - 1)


def test_filler_key_leaves_the_bucket():
    """bucket_search_range: a filler word whose hash4 high byte differs from the bucket's
    digit can agree with no bucket element on 4 bytes: hash4(0) = 0 and hash4(1) = 0x1E35,
    so 0 serves every digit but 0 and 1 serves digit 0."""
    h = lambda w: ((w * 0x1E35A7BD) & 0xFFFFFFFF) >> 16
    assert h(0) >> 8 == 0 and h(1) >> 8 == 0x1E

"""The PRZS multi-frame stream (decompress.go:842-866), on the CPU.

  * the header layout against the reference's own readers: ParseStreamHeader's byte offsets
    (decompress.go:849-866), written out by hand here, and the C99 decoder's
    przm_get_uncompressed_size, which reads the u32 at byte 5 of a PRZS stream
    (src/przm_dec.c:1564-1567);
  * the chunk count and total of the reference tests: ceil(len / chunk) chunks and the input
    length (tests/stress/stress_test.go:1986-2010 with a prime size);
  * the library's decoder on streams of the oracle's frames (round trips, every frame also
    through the reference decoder) and on broken streams, which must fail cleanly.
"""
import numpy as np
import pytest

from oracle_lib import oracle, refdec, have_refdec, stream_of

from przm_amd import datagen as D
from przm_amd import decode as X

needs_ref = pytest.mark.skipif(not have_refdec(), reason="reference decoder not built")


def _stream(a, level=6):
    frames, sizes = oracle().encode_buffer(a, level)
    return stream_of(frames, sizes, a.size), frames, sizes


def test_header_bytes_by_hand():
    # ParseStreamHeader: magic u32 at 0, flags at 4, chunk count u32 at 5, total u64 at 9
    a = D.markov(150000, 3)
    s, _, sizes = _stream(a)
    assert len(sizes) == 3
    assert s[:4] == b"PRZS" and s[4] == 0
    assert s[5:9] == bytes([3, 0, 0, 0])
    assert s[9:17] == bytes([0xF0, 0x49, 0x02, 0, 0, 0, 0, 0])  # 150000 = 0x249F0
    assert X.ParseStreamHeader(s) == (3, 150000, 0)
    # the first chunk's length, then the first frame
    assert int.from_bytes(s[17:21], "little") == sizes[0]
    assert s[21:25] == b"PRZM"


@needs_ref
def test_reference_decoder_reads_the_count_at_byte_5():
    a = D.markov(5 * 65536 + 17, 4)
    s, _, sizes = _stream(a)
    assert refdec().uncompressed_size(s) == len(sizes) == 6


def test_chunk_count_and_total_of_a_prime_size():
    # stress_test.go:1986-2010: ceil(payload / chunk) chunks and the payload's length
    rng = np.random.default_rng(999999)
    a = rng.integers(0, 256, 1000007, dtype=np.uint8)
    s, _, sizes = _stream(a)
    cnt, total, flags = X.ParseStreamHeader(s)
    assert cnt == (1000007 + 65535) // 65536 == len(sizes) == 16
    assert total == 1000007 and flags == 0
    assert X.Decompress(s) == a.tobytes()


@pytest.mark.parametrize("n", [0, 1, 4096, 65535, 65536])
def test_one_chunk_stays_a_frame(n):
    a = D.markov(n, 7) if n else np.zeros(0, np.uint8)
    s, frames, _ = _stream(a)
    assert s == frames and s[:4] == b"PRZM"
    assert X.Decompress(s) == a.tobytes()


@pytest.mark.parametrize("name,n", [("markov", 65537), ("markov", 300000), ("mixture", 1 << 20),
                                    ("zeros", 200000)])
def test_round_trip(name, n):
    a = {"markov": lambda: D.markov(n, 11), "mixture": lambda: D.mixture(n, 12),
         "zeros": lambda: np.zeros(n, np.uint8)}[name]()
    s, _, _ = _stream(a)
    assert X.Decompress(s) == a.tobytes()


@needs_ref
def test_every_chunk_decodes_with_the_reference_decoder():
    a = D.mixture(700000, 13)
    s, _, sizes = _stream(a)
    pos, out = 17, []
    for n in sizes:
        assert int.from_bytes(s[pos:pos + 4], "little") == n
        out.append(refdec().decompress(s[pos + 4: pos + 4 + int(n)], 65536))
        pos += 4 + int(n)
    assert pos == len(s) and b"".join(out) == a.tobytes()


def _broken(s, fn):
    b = bytearray(s)
    fn(b)
    with pytest.raises(X.DecodeError) as e:
        X.Decompress(bytes(b))
    return e.value.code


def test_broken_streams_fail_cleanly():
    a = D.markov(200000, 21)
    s, _, sizes = _stream(a)
    # a count one higher: the walk runs out of bytes
    assert _broken(s, lambda b: b.__setitem__(slice(5, 9), (len(sizes) + 1).to_bytes(4, "little"))) == X.ERR_CORRUPT
    # a count one lower: bytes left over
    assert _broken(s, lambda b: b.__setitem__(slice(5, 9), (len(sizes) - 1).to_bytes(4, "little"))) == X.ERR_SIZE
    # a total that the chunks do not produce
    assert _broken(s, lambda b: b.__setitem__(slice(9, 17), (200001).to_bytes(8, "little"))) == X.ERR_SIZE
    # a flag no writer defines
    assert _broken(s, lambda b: b.__setitem__(4, 1)) == X.ERR_UNSUPPORTED
    # a chunk length past the end
    assert _broken(s, lambda b: b.__setitem__(slice(17, 21), (len(s)).to_bytes(4, "little"))) == X.ERR_CORRUPT
    # every truncation fails
    for cut in (1, 4, 16, 17, 20, 21, len(s) // 2, len(s) - 1):
        with pytest.raises(X.DecodeError):
            X.Decompress(s[:cut])

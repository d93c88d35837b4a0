"""GPU parity for result egress: device-formatted text equals C's printf
output byte for byte ("%d\\n" of BFS.cu:41-43 / SSSP.cu:41-43, "%f\\n" of
PageRank.cu:42-44, checked against libc's snprintf), and binary dumps read
back bit-exact."""
import ctypes
import struct
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_libc = ctypes.CDLL(None)
_libc.snprintf.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_double]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def printf_f(values):
    """The bytes C's fprintf(f, "%f\\n", x) writes for each x."""
    buf = ctypes.create_string_buffer(512)
    out = []
    for x in values:
        n = _libc.snprintf(buf, 512, b"%f\n", float(x))
        out.append(buf.raw[:n])
    return b"".join(out)


def doubles(rng, k):
    mags = 10.0 ** rng.uniform(-12, 12, k)
    x = mags * rng.choice([-1.0, 1.0], k)
    ties = np.arange(-2000, 2000) / 128.0            # exact .5 ulp-of-1e-6 ties
    ties2 = (np.arange(1, 4000) * 2.0 + 1) / 2 ** 21
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 5e-324, -5e-324, 2.2250738585072014e-308,
                        0.0000005, 0.0000015, 0.0000025, 0.9999995, 0.99999949999999997, 1e-7, 123456789.1234565,
                        2.0 ** 62, -(2.0 ** 62), 9.2e18, 1.0, -1.0, 0.5, 1e-6, 4.5e-7, 5.5e-7])
    return np.concatenate([x, ties, ties2, special]).astype(np.float64)


def test_text_int32(olv, tmp_path):
    say("egress %d")
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.integers(-2 ** 31, 2 ** 31, 100000), [0, -1, 1, 2 ** 31 - 1, -2 ** 31, 0x7FFFFFFF]])
    a = a.astype(np.int32)
    want = "".join(f"{x}\n" for x in a.tolist()).encode()
    p = tmp_path / "d.txt"
    olv.write_text(p, olv.DeviceBuffer.from_numpy(a), "%d")
    assert p.read_bytes() == want
    olv.write_text(p, a, "%d")  # host input
    assert p.read_bytes() == want


def test_text_f64_matches_printf(olv, tmp_path):
    say("egress %f f64")
    x = doubles(np.random.default_rng(2), 100000)
    p = tmp_path / "f.txt"
    olv.write_text(p, olv.DeviceBuffer.from_numpy(x), "%f")
    got, want = p.read_bytes().split(b"\n"), printf_f(x).split(b"\n")
    bad = [(v, g, w) for v, g, w in zip(x, got, want) if g != w]
    assert not bad, bad[:5]
    assert len(got) == len(want)


def test_text_f32_and_struct_field(olv, tmp_path):
    say("egress %f f32 and strided field")
    rng = np.random.default_rng(3)
    x = (10.0 ** rng.uniform(-9, 6, 50000)).astype(np.float32)
    p = tmp_path / "g.txt"
    olv.write_text(p, olv.DeviceBuffer.from_numpy(x), "%f32")
    assert p.read_bytes() == printf_f(x.astype(np.float64))
    # PR_Vertex {double rank; double delta;} (PageRank.cu:38-44): the rank field
    pr = np.zeros(30000, dtype=[("rank", np.float64), ("delta", np.float64)])
    pr["rank"] = rng.random(30000) / 1000
    pr["delta"] = -7.0
    dev = olv.DeviceBuffer.from_numpy(pr.view(np.uint8))
    olv.write_text(p, dev, "%f", stride=16)
    assert p.read_bytes() == printf_f(pr["rank"])


def test_text_errors(olv, tmp_path):
    with pytest.raises(olv.OliveError) as e:
        olv.write_text(tmp_path / "x.txt", olv.DeviceBuffer.from_numpy(np.array([1.0, 1e19])), "%f")
    assert e.value.code == 5
    with pytest.raises(olv.OliveError) as e:
        olv.write_text(tmp_path / "x.txt", olv.DeviceBuffer.from_numpy(np.zeros(8)), "%f", stride=12)
    assert e.value.code == 1


def test_text_many_chunks(olv, tmp_path):
    """More values than one formatting round holds (2^23)."""
    say("egress chunks")
    a = (np.arange(9_000_000, dtype=np.int64) * 7919 % 1_000_003 - 500_000).astype(np.int32)
    p = tmp_path / "big.txt"
    olv.write_text(p, olv.DeviceBuffer.from_numpy(a), "%d")
    got = np.array(p.read_bytes().split(b"\n")[:-1], dtype=np.int64)
    assert got.size == a.size and np.array_equal(got, a)


def test_binary_round_trip(olv, tmp_path):
    say("egress binary")
    rng = np.random.default_rng(4)
    for a in [rng.integers(0, 2 ** 32, 20_000_000, dtype=np.uint32),  # > one 64 MB chunk
              rng.standard_normal(1001), np.zeros(0, np.int32)]:
        p = tmp_path / "b.bin"
        olv.write_binary(p, olv.DeviceBuffer.from_numpy(a), dtype_tag=7)
        raw = p.read_bytes()
        magic, ver, tag, esz, cnt = struct.unpack("<4sIIIQ", raw[:24])
        assert (magic, ver, tag, esz, cnt) == (b"OLVR", 1, 7, a.dtype.itemsize, a.size)
        assert raw[24:] == a.tobytes()
        back = olv.read_binary(p, a.dtype)
        assert np.array_equal(back.numpy(), a)
        olv.write_binary(p, a)  # host input
        assert p.read_bytes()[24:] == a.tobytes()
    p.write_bytes(b"not a dump at all, definitely")
    with pytest.raises(olv.OliveError):
        olv.read_binary(p, np.int32)

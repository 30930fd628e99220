"""xtc trajectory frames (sre_xtc_*, csrc/xtc.cpp; the reference writes them
through libxdrfile from MDSystem::writeHostDataXtc,
H/src/MDSystem_interface.cu:243-290).

libxdrfile is not part of the reference and no .xtc file ships with it, so
byte identity with libxdrfile's output is unpinned.  What is pinned here:
  * the uncompressed layout of frames of at most 9 atoms, byte for byte
    against struct.pack of the XDR words;
  * the header of compressed frames (bounding box of the integer
    coordinates, first class index from the smallest neighbour distance,
    byte count, 4-byte padding) against the same quantities computed here;
  * the class table: entry i packs three values below it into i bits
    (s^3 <= 2^i) and is within 2 % of 2^(i/3);
  * round trips through every branch of the coder (large-atom mixed radix
    and wide fields, runs of small atoms, the pair swap, class steps up and
    down), exact to the integer lattice of the precision.
"""
import os
import re
import struct

import numpy as np
import pytest

from sre import configs
from sre import io as sio
from sre.engine import InvalidArgument, MDExcptCannotOpenFile

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def lattice(x, prec):
    """The format's integers: FP32 product, +-0.5 in double, back to FP32,
    truncated."""
    p = (np.asarray(x, np.float32) * np.float32(prec)).astype(np.float32)
    lf = np.where(p >= 0, p.astype(np.float64) + 0.5, p.astype(np.float64) - 0.5)
    return np.trunc(lf.astype(np.float32)).astype(np.int64)


def class_table():
    src = open(os.path.join(REPO, "md.error_amd", "csrc", "xtc.cpp")).read()
    body = re.search(r"kClassSize\[\] = \{([^}]*)\}", src).group(1)
    return [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]


def roundtrip(tmp_path, x, prec=1000.0, box=(3.0, 3.0, 12.0), name="t.xtc"):
    path = tmp_path / name
    with sio.XtcFile(path, "w") as f:
        f.write(7, 0.25, box, x, prec)
    frames = sio.read_xtc(path)
    assert len(frames) == 1
    step, t, b, y, p = frames[0]
    assert step == 7 and t == np.float32(0.25) and p == np.float32(prec)
    np.testing.assert_array_equal(b, np.diag(np.asarray(box, np.float32)))
    q = lattice(x, prec)
    # decoded value = integer * (1 / precision) in FP32
    expect = (q.astype(np.float32) * np.float32(1.0 / np.float32(prec))).astype(np.float32)
    np.testing.assert_array_equal(y, expect)
    assert np.all(np.abs(y - x) <= 0.5 / prec + 1e-6 * np.abs(x) + 1e-7)
    return path


def test_class_table_packs_three_values_into_index_bits():
    t = class_table()
    assert len(t) == 73 and t[:9] == [0] * 9
    def icbrt(v):                     # largest m with m^3 <= v
        m = int(round(v ** (1.0 / 3.0)))
        while m ** 3 > v:
            m -= 1
        while (m + 1) ** 3 <= v:
            m += 1
        return m
    # the format keeps three historical entries below the largest cube root
    # (5060 at 37, and 2^19 - 1 and 2^23 - 1 where 2^i is a perfect cube)
    historical = {37: 5060, 57: 524287, 69: 8388607}
    for i in range(9, 73):
        assert t[i] ** 3 <= 2 ** i, i
        assert t[i] == historical.get(i, icbrt(2 ** i)), i
        assert t[i] > t[i - 1] or i == 9


def test_small_frames_are_plain_xdr(tmp_path):
    rng = np.random.default_rng(1)
    box = np.array([[2.5, 0, 0], [0, 3.0, 0], [0, 0, 4.0]], np.float32)
    frames = []
    path = tmp_path / "small.xtc"
    with sio.XtcFile(path, "w") as f:
        for n in (0, 1, 9):
            x = rng.uniform(-1, 5, (n, 3)).astype(np.float32)
            f.write(3 * n, 0.5 * n, box, x, 123.0)
            frames.append(x)
    expect = b""
    for n, x in zip((0, 1, 9), frames):
        expect += struct.pack(">iiif", 1995, n, 3 * n, 0.5 * n)
        expect += struct.pack(">9f", *box.ravel())
        expect += struct.pack(">i", n) + struct.pack(f">{3 * n}f", *x.ravel())
    assert path.read_bytes() == expect
    with sio.XtcFile(path, "r", natoms=0) as f:
        step, t, b, y, p = f.read()
        assert (step, y.shape, p) == (0, (0, 3), 0.0)


def test_compressed_header_fields(tmp_path):
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 3, (50, 3)).astype(np.float32)
    x[10] = x[9] + 0.0015           # closest consecutive pair: |dq|_1 small
    prec = 1000.0
    path = roundtrip(tmp_path, x, prec)
    raw = path.read_bytes()
    q = lattice(x, prec)
    w = struct.unpack(">iiif9fi", raw[:56])
    assert w[:3] == (1995, 50, 7) and w[-1] == 50
    (p,) = struct.unpack(">f", raw[56:60])
    mn = struct.unpack(">3i", raw[60:72])
    mx = struct.unpack(">3i", raw[72:84])
    first, nbytes = struct.unpack(">2i", raw[84:92])
    assert p == prec
    assert list(mn) == list(q.min(0)) and list(mx) == list(q.max(0))
    mindiff = np.abs(np.diff(q, axis=0)).sum(1).min()
    t = class_table()
    ci = next(i for i in range(9, 73) if t[i] >= mindiff)
    assert first == ci
    assert len(raw) == 92 + (nbytes + 3) // 4 * 4
    assert raw[92 + nbytes:] == b"\0" * (len(raw) - 92 - nbytes)


def test_roundtrip_liquid_slab(tmp_path):
    w = configs.slab(n=64000)
    x = (w.xyz[:, :3] * 0.34).astype(np.float32)     # sigma = 0.34 nm
    roundtrip(tmp_path, x, 1000.0, box=tuple(w.boxes[0] * 0.34))


def test_roundtrip_water_like_triples(tmp_path):
    """Molecules of three atoms 0.1 nm apart: the pair swap and short runs."""
    rng = np.random.default_rng(3)
    centres = rng.uniform(0, 3, (400, 3))
    x = np.concatenate([centres[:, None, :] + rng.normal(0, 0.06, (400, 3, 3))],
                       1).reshape(-1, 3).astype(np.float32)
    roundtrip(tmp_path, x, 1000.0)


def test_roundtrip_chains_with_changing_spacing(tmp_path):
    """Random walks whose step size changes along the chain (0.02 to 2 nm):
    long runs of small atoms and class steps in both directions."""
    rng = np.random.default_rng(4)
    steps = []
    for scale in (0.02, 0.3, 0.05, 2.0, 0.01, 0.5, 0.08):
        steps.append(rng.normal(0, scale, (300, 3)))
    x = np.cumsum(np.concatenate(steps), 0).astype(np.float32) + 10.0
    roundtrip(tmp_path, x, 1000.0)
    roundtrip(tmp_path, x, 100.0, name="p100.xtc")
    roundtrip(tmp_path, x, 1e5, name="p1e5.xtc")


def test_roundtrip_wide_range_and_negative(tmp_path):
    """Coordinate ranges above 2^24 at the precision: three fixed-width
    fields per large atom; negative coordinates."""
    rng = np.random.default_rng(5)
    x = rng.uniform(-9000, 12000, (200, 3)).astype(np.float32)
    x[50:60] = x[49] + rng.normal(0, 0.01, (10, 3)).astype(np.float32)
    roundtrip(tmp_path, x, 1000.0)


def test_frames_append_and_end(tmp_path):
    rng = np.random.default_rng(6)
    path = tmp_path / "multi.xtc"
    xs = [rng.uniform(0, 5, (30, 3)).astype(np.float32) for _ in range(4)]
    with sio.XtcFile(path, "w") as f:
        f.write(0, 0.0, (5, 5, 5), xs[0])
        f.write(1, 0.005, (5, 5, 5), xs[1])
    with sio.XtcFile(path, "a") as f:
        for k in (2, 3):
            f.write(k, 0.005 * k, (5, 5, 5), xs[k])
    frames = sio.read_xtc(path)
    assert [fr[0] for fr in frames] == [0, 1, 2, 3]
    for fr, x in zip(frames, xs):
        assert np.all(np.abs(fr[3] - x) <= 0.5e-3 + 1e-6)
    with sio.XtcFile(path, "r") as f:
        assert f.natoms == 30
        assert len(list(f)) == 4 and f.read() is None


def test_errors(tmp_path):
    with pytest.raises(MDExcptCannotOpenFile):
        sio.XtcFile(tmp_path / "missing" / "a.xtc", "w")
    path = tmp_path / "a.xtc"
    x = np.zeros((20, 3), np.float32)
    x[:, 0] = np.arange(20)
    with sio.XtcFile(path, "w") as f:
        f.write(0, 0.0, (30, 1, 1), x)
    with pytest.raises(InvalidArgument):
        with sio.XtcFile(path, "r", natoms=21) as f:
            f.read()
    raw = path.read_bytes()
    (tmp_path / "cut.xtc").write_bytes(raw[:-8])
    with pytest.raises(MDExcptCannotOpenFile):
        sio.read_xtc(tmp_path / "cut.xtc")
    with pytest.raises(InvalidArgument):
        with sio.XtcFile(tmp_path / "big.xtc", "w") as f:
            f.write(0, 0.0, (1, 1, 1), np.full((12, 3), 3e6, np.float32), 1000.0)

"""Film output formats (mcrt_film_write_hdr / _ppm / _tonemap_gamma) against
the compiled reference's FrameBuffer::dumpHDR / dumpPPM / tonemapGamma
(film.cpp:9-108, :195-202): byte-identical files and bit-identical
tone-mapped values, on the committed fixture (tests/golden/film.npz,
written by the reference) and, where oracle/_ref is built, on further random
films.  Host-only functions of the product library: no GPU needed."""
import os

import numpy as np
import pytest

import _libs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bytes(path):
    with open(path, "rb") as fh:
        return np.frombuffer(fh.read(), np.uint8)


def test_film_writers_match_reference_fixture(tmp_path):
    import mcrt_amd
    g = np.load(os.path.join(GOLD, "film.npz"))
    import make_golden
    f = make_golden.film_inputs()
    assert np.array_equal(f.view(np.uint32), g["film"].view(np.uint32)), "film generator drifted"
    mcrt_amd.write_hdr(tmp_path / "a.hdr", f)
    mcrt_amd.write_ppm(tmp_path / "a.ppm", f)
    assert np.array_equal(_bytes(tmp_path / "a.hdr"), g["hdr"])
    assert np.array_equal(_bytes(tmp_path / "a.ppm"), g["ppm"])
    t = mcrt_amd.tonemap_gamma(f, 2.2)
    assert np.array_equal(t.view(np.uint32), g["gamma22"].view(np.uint32))


def test_hdr_layout_decodes():
    """Independent decode of the written RGBE file: header lines, the
    scanline headers, 127-byte literal runs, and the mantissa/exponent
    bytes reconstruct each pixel within the format's 8-bit precision."""
    import tempfile
    import mcrt_amd
    rng = np.random.default_rng(3)
    w, h = 300, 5
    f = rng.uniform(0.01, 100.0, (h, w, 3)).astype(np.float32)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "x.hdr")
        mcrt_amd.write_hdr(p, f)
        b = _bytes(p).tobytes()
    head = b"#?RADIANCE\n# Made with custom writer\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y 5 +X 300\n"
    assert b.startswith(head)
    pos = len(head)
    for j in range(h - 1, -1, -1):
        assert b[pos:pos + 4] == bytes([2, 2, w >> 8, w & 255])
        pos += 4
        planes = []
        for _ in range(4):
            plane = bytearray()
            while len(plane) < w:
                n = b[pos]
                assert 0 < n <= 127
                plane += b[pos + 1:pos + 1 + n]
                pos += 1 + n
            planes.append(np.frombuffer(bytes(plane), np.uint8).astype(np.float64))
        r, g_, bb, e = planes
        scale = np.ldexp(1.0, (e - 136).astype(np.int64))
        dec = np.stack([r, g_, bb], axis=1) * scale[:, None]
        ref = f[j].astype(np.float64)
        assert np.all(np.abs(dec - ref) <= ref.max(axis=1, keepdims=True) / 128.0)
    assert pos == len(b)


def test_film_errors(tmp_path):
    import mcrt_amd
    f = np.zeros((2, 2, 3), np.float32)
    with pytest.raises(mcrt_amd.McrtError) as e:
        mcrt_amd.write_hdr(tmp_path / "no" / "such" / "dir.hdr", f)
    assert e.value.code == 3
    with pytest.raises(mcrt_amd.McrtError) as e:
        mcrt_amd.tonemap_gamma(f, 0.0)
    assert e.value.code == 1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_film_writers_match_compiled_reference(tmp_path, seed):
    R = _libs.reference()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import mcrt_amd
    rng = np.random.default_rng(seed)
    h, w = int(rng.integers(1, 9)), int(rng.integers(1, 400))
    f = (rng.uniform(0.0, 1.0, (h, w, 3)) ** 3 * 10.0 ** rng.uniform(-35, 6, (h, w, 1))).astype(np.float32)
    for ext, ppm in (("hdr", 0), ("ppm", 1)):
        a, r = tmp_path / f"a.{ext}", tmp_path / f"r.{ext}"
        (mcrt_amd.write_ppm if ppm else mcrt_amd.write_hdr)(a, f)
        R.ref_film_dump(str(r).encode(), _libs.ptr(f), w, h, ppm)
        assert np.array_equal(_bytes(a), _bytes(r)), ext
    t = f.copy()
    R.ref_film_tonemap_gamma(_libs.ptr(t), w, h, 1.8)
    assert np.array_equal(mcrt_amd.tonemap_gamma(f, 1.8).view(np.uint32), t.view(np.uint32))

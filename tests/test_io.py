"""Caller-side I/O (no device needed): PGM reading and writing as the
reference's loadPGM/savePGM (src/Image.h:110-169), PFM output, the 8-bit
disparity view of src/Main.cu:14-27, and the command-line tool's behaviour
without a device.

The demo pairs tests/golden/{teddy,cones}_{left,right}.pgm are the
reference's own img/*.pgm, copied as data.
"""
import os
import subprocess

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _parse_pgm(path):
    """Independent reader: header tokens with '#' lines skipped, one
    whitespace byte after maxval, then the raw bytes."""
    data = open(path, "rb").read()
    toks, i = [], 0
    while len(toks) < 4:
        while data[i:i + 1].isspace():
            i += 1
        if data[i:i + 1] == b"#":
            while data[i:i + 1] != b"\n":
                i += 1
            continue
        j = i
        while not data[j:j + 1].isspace():
            j += 1
        toks.append(data[i:j])
        i = j + 1
    assert toks[0] == b"P5"
    w, h = int(toks[1]), int(toks[2])
    return np.frombuffer(data[i:i + w * h], np.uint8).reshape(h, w)


@pytest.mark.parametrize("name,shape", [("teddy_left", (375, 450)), ("cones_right", (750, 900))])
def test_read_reference_demo_images(pkg, name, shape):
    img = pkg.read_pgm(os.path.join(GOLDEN, name + ".pgm"))
    assert img.shape == shape
    assert np.array_equal(img, _parse_pgm(os.path.join(GOLDEN, name + ".pgm")))
    assert img.std() > 10  # a real photograph, not a constant buffer


def test_pgm_round_trip_and_comments(pkg, tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    p = tmp_path / "a.pgm"
    pkg.write_pgm(p, img)
    assert open(p, "rb").read(13) == b"P5\n53 37\n255\n"
    assert np.array_equal(pkg.read_pgm(p), img)
    # comment lines before any header token, data starting with a space byte
    img[0, 0] = 32
    q = tmp_path / "b.pgm"
    q.write_bytes(b"P5\n# a comment\n53\n# another\n 37 255\n" + img.tobytes())
    assert np.array_equal(pkg.read_pgm(q), img)


@pytest.mark.parametrize("blob", [b"P6\n2 2\n255\n" + bytes(12), b"P5\n2 2\n65535\n" + bytes(8),
                                  b"P5\n4 4\n255\n" + bytes(15), b"P5\n-3 2\n255\n" + bytes(6), b""])
def test_malformed_pgm_is_an_error(pkg, tmp_path, blob):
    p = tmp_path / "bad.pgm"
    p.write_bytes(blob)
    with pytest.raises(pkg.ElasError) as e:
        pkg.read_pgm(p)
    assert e.value.code == pkg.ELAS_ERR_IO


def test_missing_file_is_an_error(pkg, tmp_path):
    with pytest.raises(pkg.ElasError) as e:
        pkg.read_pgm(tmp_path / "none.pgm")
    assert e.value.code == pkg.ELAS_ERR_IO


def test_pfm_layout(pkg, tmp_path):
    d = np.arange(12, dtype=np.float32).reshape(3, 4) - 1.5
    p = tmp_path / "d.pfm"
    pkg.write_pfm(p, d)
    raw = open(p, "rb").read()
    head = b"Pf\n4 3\n-1\n"
    assert raw[:len(head)] == head
    got = np.frombuffer(raw[len(head):], "<f4").reshape(3, 4)[::-1]
    assert np.array_equal(got, d)


def test_disparity_view_matches_the_reference_formula(pkg):
    rng = np.random.default_rng(1)
    d = rng.integers(-1, 128, (40, 60)).astype(np.float32)
    d[rng.random(d.shape) < 0.2] = -1.0
    got = pkg.disparity_to_u8(d)
    m = float(d.max())
    exp = np.maximum(255.0 * d.astype(np.float64) / m, 0.0).astype(np.uint8)
    assert np.array_equal(got, exp)
    assert got[d == d.max()].min() == 255 and (got[d < 0] == 0).all()
    assert (pkg.disparity_to_u8(np.full((3, 3), -1, np.float32)) == 0).all()


def test_cli_usage_and_missing_device(pkg, tmp_path):
    """-h prints the usage (exit 0) without touching a device; a pair run on
    a machine without a gfx950 GPU fails with a message (exit 1)."""
    r = subprocess.run([pkg.CLI_PATH, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "left.pgm right.pgm" in r.stdout
    bad = subprocess.run([pkg.CLI_PATH, "a", "b", "c"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    left = os.path.join(GOLDEN, "teddy_left.pgm")
    r = subprocess.run([pkg.CLI_PATH, left, left, "-o", str(tmp_path / "x")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and "no usable gfx950 device" in r.stdout

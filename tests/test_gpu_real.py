"""The reference's own demo pairs through the HIP path.

tests/golden/{teddy,cones}_{left,right}.pgm are the reference's img/*.pgm
(teddy 450x375, cones 900x750), copied as data.  At the reference's
effective parameters (disparities 0..255) every intermediate is bit-exact
with the oracle, and the command-line tool (the counterpart of src/Main.cu)
writes exactly the maps the library computes.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STAGE_KEYS = ["desc0", "desc1", "can_left", "can_lr", "can_incon", "can_final", "sp0", "sp1",
              "grid0", "grid1", "tri0", "tri1", "mean0", "mean1", "d1_raw", "d2_raw"]


def _pair(pkg, name):
    return (pkg.read_pgm(os.path.join(GOLDEN, name + "_left.pgm")),
            pkg.read_pgm(os.path.join(GOLDEN, name + "_right.pgm")))


def _read_pfm(path):
    raw = open(path, "rb").read()
    lines = raw.split(b"\n", 3)
    assert lines[0] == b"Pf" and lines[2] == b"-1"
    w, h = (int(x) for x in lines[1].split())
    return np.frombuffer(lines[3], "<f4").reshape(h, w)[::-1]


@pytest.mark.parametrize("name", ["teddy", "cones"])
def test_demo_pair_bit_exact(pkg, oracle, name):
    left, right = _pair(pkg, name)
    print("demo pair", name, left.shape, flush=True)
    el = pkg.Elas(pkg.default_params())
    el.set_debug(True)
    D1, D2 = el.process(left, right)
    ref = oracle.process(left, right, oracle.params())
    assert el.status == ref["status"] == 0
    for k in STAGE_KEYS:
        got = el.fetch(k)
        if not np.array_equal(got, ref[k]):
            bad = np.argwhere(got != ref[k])
            raise AssertionError("%s: %s differs at %d positions, first %s" % (name, k, len(bad), bad[0].tolist()))
    assert np.array_equal(D1, ref["D1"]) and np.array_equal(D2, ref["D2"])
    assert (D1 >= 0).mean() > 0.5


def test_cli_writes_the_library_maps(pkg, tmp_path):
    left, right = _pair(pkg, "teddy")
    el = pkg.Elas(pkg.default_params())
    D1, D2 = el.process(left, right)
    prefix = str(tmp_path / "teddy")
    print("cli run", prefix, flush=True)
    r = subprocess.run([pkg.CLI_PATH, os.path.join(GOLDEN, "teddy_left.pgm"), os.path.join(GOLDEN, "teddy_right.pgm"),
                        "-o", prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Processing:" in r.stdout and "... done!" in r.stdout
    assert np.array_equal(_read_pfm(prefix + "_D1.pfm"), D1)
    assert np.array_equal(_read_pfm(prefix + "_D2.pfm"), D2)
    assert np.array_equal(pkg.read_pgm(prefix + "_D1.pgm"), pkg.disparity_to_u8(D1))
    assert np.array_equal(pkg.read_pgm(prefix + "_D2.pgm"), pkg.disparity_to_u8(D2))


def test_cli_demo_mode_and_size_mismatch(pkg, tmp_path):
    """demo mode over a directory holding two of the four pairs: those are
    written, the missing ones reported (exit 1); a pair of different sizes is
    refused with the reference's message."""
    d = tmp_path / "img"
    d.mkdir()
    for name in ("teddy", "cones"):
        for side in ("left", "right"):
            (d / ("%s_%s.pgm" % (name, side))).write_bytes(
                open(os.path.join(GOLDEN, "%s_%s.pgm" % (name, side)), "rb").read())
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([pkg.CLI_PATH, "demo", str(d), "-o", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1
    for name in ("teddy", "cones"):
        assert (out / (name + "_D1.pfm")).exists() and (out / (name + "_D2.pgm")).exists()
    assert "Could not read file" in r.stdout
    r = subprocess.run([pkg.CLI_PATH, os.path.join(GOLDEN, "teddy_left.pgm"), os.path.join(GOLDEN, "cones_left.pgm"),
                        "-o", str(tmp_path / "mm")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "Images must be of same size" in r.stdout

"""The reference-compatible C++ layer without a GPU: its own check program
(image file formats, record reading and line fits, JSON keys and
defaults) and the image format shared with the Python reader/writer."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cxx", "dropin_check")


def _need_check():
    if not os.path.exists(CHECK):
        pytest.fail("tests/cxx/dropin_check is not built (run __graft_entry__.build())")


def test_cxx_selftest(tmp_path):
    _need_check()
    r = subprocess.run([CHECK, "selftest", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "selftest ok" in r.stdout


def test_image_file_python_and_header_layout(tmp_path):
    from muontomography_amd._abi import make_grid
    from muontomography_amd.imagefile import read_image, write_image
    g = make_grid((5, 4, 3), (2.5, 3.0, 4.0), (-10.0, -20.0, -30.0))
    img = 1e-7 * (1 + 0.37 * np.arange(60))
    for prec in (4, 8):
        p = tmp_path / f"a{prec}.bin"
        write_image(p, img, g, method="poca", n_ray=77, n_iter=1, precision=prec)
        raw = p.read_bytes()
        assert len(raw) == 64 + 60 * prec
        assert raw[0] == 64 and raw[1] == 1 and raw[2:6] == b"poca" and raw[56:64] == b"\0" * 8
        v, h = read_image(p)
        assert h["method"] == "poca" and h["n_ray"] == 77 and (h["nx"], h["ny"], h["nz"]) == (5, 4, 3)
        ref = img.astype(np.float32).astype(np.float64) if prec == 4 else img
        np.testing.assert_array_equal(v, ref)


def test_records_parsed_by_cxx_reader(tmp_path):
    """Collinear hits on the 6 tracker layers: the C++ reader's fitted rays
    reproduce the lines the hits came from (float32 records)."""
    _need_check()
    from muontomography_amd import RAY_DTYPE
    rng = np.random.default_rng(1)
    zl = [1500.0, 1000.0, 500.0, -500.0, -1000.0, -1500.0]
    n = 50
    recs = np.empty((n, 14), np.float32)
    din = np.c_[rng.normal(0, 0.2, n), rng.normal(0, 0.2, n), np.ones(n)]
    dout = din + np.c_[rng.normal(0, 0.02, n), rng.normal(0, 0.02, n), np.zeros(n)]
    c = np.c_[rng.uniform(-400, 400, n), rng.uniform(-400, 400, n), np.zeros(n)]
    recs[:, 0] = 3.0
    recs[:, 13] = 2.0
    for i, z in enumerate(zl):
        d = din if i < 3 else dout
        t = (z - c[:, 2]) / d[:, 2]
        recs[:, 1 + 2 * i] = c[:, 0] + t * d[:, 0]
        recs[:, 2 + 2 * i] = c[:, 1] + t * d[:, 1]
    f = tmp_path / "m.bin"
    f.write_bytes(recs.tobytes())
    out = tmp_path / "rays.bin"
    r = subprocess.run([CHECK, "rays", str(out), "4", "3"] + [str(z) for z in zl] + ["--", str(f)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    rays = np.frombuffer(out.read_bytes(), RAY_DTYPE)
    assert rays.shape[0] == n
    un = din / np.linalg.norm(din, axis=1)[:, None]
    np.testing.assert_allclose(rays["din"], un, atol=2e-6)
    np.testing.assert_allclose(rays["pin"][:, 2], 500.0)
    np.testing.assert_allclose(rays["pout"][:, 2], -500.0)
    np.testing.assert_allclose(rays["pin"][:, 0], recs[:, 5], atol=2e-3)

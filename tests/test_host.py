"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function include/dachshund_gpu.h declares, and the spatial ordering
and cutoff culling (host code of the library) never drop a nonzero pair."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, fixture_pixels


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(ds_\w+)\s*\(",
                            text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from dachshund_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared surface
    assert sorted(_lib.SIGNATURES) == names


def _nonzero_pairs(ora, pts, c):
    """All (i, j) with S_ij != 0 by the oracle's own evaluation."""
    npx = np.zeros((len(pts), 5))
    npx[:, :3] = pts
    npx[:, 3] = 1.0
    S = ora.A_dense(ora.FORM_SN, npx, c) - np.eye(len(pts))
    return np.argwhere(S != 0.0)


@pytest.mark.parametrize("case", ["fixture", "tiny", "small", "clustered",
                                  "outliers"])
def test_culling_keeps_every_nonzero_pair(ora, case):
    import dachshund_amd as D
    from dachshund_amd import synth
    if case == "fixture":
        pts, (lp, ll) = fixture_pixels()[:, :3], (2.5, 2.1)
    elif case == "clustered":
        # many coincident and near-cutoff points
        rng = np.random.default_rng(8)
        base = rng.uniform(0, 30, (40, 3))
        pts = np.repeat(base, 20, axis=0)
        pts[::3] += rng.normal(0, 3.0, (len(pts[::3]), 3))
        lp, ll = 1.7, 2.9
    elif case == "outliers":
        # far-away points stretch the tile-centre range (the binned
        # candidate search must still find every listed tile)
        w = synth.workload("small")
        pts = synth.make_pixels(w)[:, :3].copy()
        pts[5] = [1e4, -3e3, 5e5]
        pts[700] = [-2e3, 40.0, 7.0]
        pts[1300:1310] = pts[1300] + np.array([0.0, 0.0, 1e3])
        lp, ll = w.l_perp, w.l_para
    else:
        w = synth.workload(case)
        pts, (lp, ll) = synth.make_pixels(w)[:, :3], (w.l_perp, w.l_para)
    sp = D.SignalCovarParams(0.05, lp, ll)
    c = ora.covar(0.05, lp, ll)
    perm, ptr, idx = D.plan_tiles(pts, sp)
    n = len(pts)
    assert sorted(perm[:n]) == list(range(n))
    slot = np.empty(n, dtype=np.int64)
    slot[perm[:n]] = np.arange(n)
    listed = set()
    for I in range(len(ptr) - 1):
        for T in idx[ptr[I]:ptr[I + 1]]:
            listed.add((I, int(T)))
    for i, j in _nonzero_pairs(ora, pts, c):
        assert (slot[i] // 64, slot[j] // 16) in listed, (i, j)
    # and the culling does cull on the spread-out synthetic geometry
    if case == "small":
        assert len(idx) * 64 * 16 < 0.5 * n * n


def test_plan_rejects_bad_input():
    import dachshund_amd as D
    from dachshund_amd import _lib
    sp = D.SignalCovarParams(0.05, 2.5, 2.1)
    pts = np.zeros((10, 3))
    pts[3, 1] = np.nan
    with pytest.raises(_lib.DsError):
        D.plan_tiles(pts, sp)


def test_python_mirror_table_is_the_reference_table(ora):
    import dachshund_amd as D
    for params in [(0.05, 2.5, 2.1), (1.0, 0.8, 1.2)]:
        sp = D.SignalCovarParams(*params)
        c = ora.covar(*params)
        assert np.array_equal(sp.func_table, np.array(c.table[:]))
        assert sp.func_table_dx_inv == c.dx_inv
        assert sp.gaussian_const_perp == c.c_perp
        assert sp.gaussian_const_para == c.c_para


def test_read_pixel_data_roundtrip(tmp_path):
    import dachshund_amd as D
    pix = fixture_pixels()
    p = tmp_path / "pix.bin"
    pix.tofile(p)
    np.testing.assert_array_equal(D.read_pixel_data(str(p), len(pix)), pix)
    with pytest.raises(FileNotFoundError):
        D.read_pixel_data(str(tmp_path / "nope.bin"), 3)
    with pytest.raises(ValueError):
        D.read_pixel_data(str(p), len(pix) + 1)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU's share; asked for more it must say so and
    exit before touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


REFERENCE = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "app")),
                    reason="needs the reference sources (build container)")
def test_reference_callers_compile_against_dropin(tmp_path):
    """The reference's callers -- app/dachshund.cc and its gtest suite
    tests/*.cc -- compile unchanged against the drop-in headers
    (dachshund_amd/cxx/include); the app also links against
    libdachshund.so + libdachshund_gpu.so.  Compile only: nothing built
    here is kept."""
    import subprocess
    inc = os.path.join(ROOT, "dachshund_amd", "cxx", "include")
    libdir = os.path.join(ROOT, "dachshund_amd")
    if not os.path.exists(os.path.join(libdir, "libdachshund.so")):
        pytest.skip("libdachshund.so not built")
    app = subprocess.run(
        ["g++", "-std=c++11", "-w", f"-I{inc}",
         os.path.join(REFERENCE, "app", "dachshund.cc"), "-o",
         str(tmp_path / "app.exe"), f"-L{libdir}", "-ldachshund",
         "-ldachshund_gpu", "-lpthread"], capture_output=True, text=True)
    assert app.returncode == 0, app.stderr[-3000:]
    gt = [f"-I{REFERENCE}/ext", f"-I{REFERENCE}/ext/gtest",
          f"-I{REFERENCE}/ext/gtest/include", f"-I{REFERENCE}/tests"]
    for name in ("wfx_tests.cc", "pixel_tests.cc", "signal_covar_tests.cc",
                 "map_covar_tests.cc", "performance_tests.cc",
                 "test_utils.cc"):
        r = subprocess.run(["g++", "-std=c++11", "-w", "-fsyntax-only",
                            f"-I{inc}", *gt,
                            os.path.join(REFERENCE, "tests", name)],
                           capture_output=True, text=True)
        assert r.returncode == 0, (name, r.stderr[-3000:])

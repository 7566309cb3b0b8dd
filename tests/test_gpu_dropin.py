"""The C++ drop-in (the reference's lib/*.h API over libdachshund_gpu.so) and
its app, run as programs on the GPU and compared with the oracle and the
reference's golden map.  Tolerances as in test_gpu_parity.py: FP64 products
to 1e-12 of sum |A a|; solves judged by accuracy against a tight solution.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden

pytestmark = pytest.mark.gpu

APP = os.path.join(ROOT, "dachshund_amd", "dachshund_gpu.exe")
API = os.path.join(ROOT, "dachshund_amd", "dropin_api.exe")


def _run(cmd, cwd):
    print("running", cmd, flush=True)
    p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True,
                       timeout=240)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return p.stdout


def test_cpp_api_matches_oracle(ora, tmp_path):
    from dachshund_amd import synth
    pix = synth.make_pixels(synth.workload("small"))
    n = len(pix)
    pix.tofile(tmp_path / "pix.bin")
    _run([API, str(tmp_path / "pix.bin"), str(n), str(tmp_path)], tmp_path)
    rd = lambda name: np.fromfile(tmp_path / name)  # noqa: E731
    c = ora.covar(0.05, 2.5, 2.1)
    vec = 0.1 * np.arange(n) - 3.0
    for form, key in ((ora.FORM_W, "w"), (ora.FORM_SN, "sn")):
        fp = ora.to_form(pix, form)
        bound = ora.A_product(form, fp, c, np.abs(vec))
        got = rd(f"{key}_prod.bin")
        assert np.all(np.abs(got - ora.A_product(form, fp, c, vec))
                      <= 1e-12 * bound)
    fw = ora.to_form(pix, ora.FORM_W)
    bound = ora.A_product(ora.FORM_W, fw, c, np.abs(vec)) + np.abs(fw[:, 4])
    assert np.all(np.abs(rd("w_res.bin") - ora.A_residual(ora.FORM_W, fw, c,
                                                         vec)) <= 1e-12 * bound)
    assert np.array_equal(rd("w_pc.bin"), ora.precond(ora.FORM_W, fw, c, vec))
    sc = rd("scalars.bin")
    assert sc[0] == ora.norm_b(ora.FORM_W, fw)
    x_star = ora.solve_cf(ora.FORM_W, fw, c)
    for name in ("w_pcg.bin", "generic_pcg.bin"):
        x = rd(name)
        assert np.linalg.norm(x - x_star) <= 1e-6 * np.linalg.norm(x_star)
    assert sc[2] == 1 and sc[4] == 1
    # solve_cf with the exact covariance (app/dachshund.cc:248)
    xe = ora.solve_cf(ora.FORM_W, fw, c, exact=True)
    np.testing.assert_allclose(rd("cf.bin"), xe, rtol=1e-8, atol=1e-10)
    # map side
    hi = pix[:, :3].max(0)
    mc = ora.map_coords(tuple(hi), (4, 4, 4))
    bound = ora.smp_product(pix[:, :3], mc, c, np.abs(vec))
    assert np.all(np.abs(rd("smp.bin") - ora.smp_product(pix[:, :3], mc, c,
                                                         vec)) <= 1e-12 * bound)
    npx = ora.to_form(pix, ora.FORM_SN)
    so = ora.sigma_m_noise_dom(npx, mc, c)
    assert np.all(np.abs(rd("nd.bin") - so) <= 1e-12 * so)
    s_ref, it_ref, ok_ref = ora.map_covar_diag(npx, mc[:8], c, 1000, 50, 1e-6)
    assert np.all(ok_ref) and np.all(sc[6:14] == 1)
    np.testing.assert_allclose(rd("mcov.bin"), s_ref, rtol=1e-5)


def test_app_end_to_end(ora, tmp_path):
    """The app on the reference's own config format and the 8,192-pixel
    workload: map.bin against the reference app's map.bin, and the
    noise-dominated covariance against the oracle."""
    g = golden("app_c1.npz")
    pix = g["pixels"]
    pix.tofile(tmp_path / "pixel_data.bin")
    cfg = str(g["config"]) + "option_noise_covar = 1\n"
    (tmp_path / "run.cfg").write_text(cfg)
    out = _run([APP, "run.cfg"], tmp_path)
    assert "reached" in out
    m = np.fromfile(tmp_path / "map.bin")
    ref = g["map"]
    assert m.shape == ref.shape
    # both solves stop at |r| <= 1e-5 |b|: the maps agree to that level
    assert np.linalg.norm(m - ref) <= 1e-4 * np.linalg.norm(ref)
    kv = dict(line.split(" = ") for line in str(g["config"]).splitlines())
    c = ora.covar(float(kv["corr_var_s"]), float(kv["corr_l_perp"]),
                  float(kv["corr_l_para"]))
    shape = (int(kv["map_nx"]), int(kv["map_ny"]), int(kv["map_nz"]))
    mc = ora.map_coords((float(kv["lx"]), float(kv["ly"]), float(kv["lz"])),
                        shape)
    nd = np.fromfile(tmp_path / "map_covar_nd.bin")
    so = ora.sigma_m_noise_dom(ora.to_form(pix, ora.FORM_SN), mc, c)
    assert np.all(np.abs(nd - so) <= 1e-12 * so + 1e-300)
    # the mixed-precision solve through the same app
    out = _run([APP, "run.cfg", "--mixed"], tmp_path)
    m2 = np.fromfile(tmp_path / "map.bin")
    assert np.linalg.norm(m2 - ref) <= 1e-4 * np.linalg.norm(ref)


def test_app_rejects_unknown_key(tmp_path):
    (tmp_path / "bad.cfg").write_text("lx = 3\nbogus_key = 1\n")
    p = subprocess.run([APP, "bad.cfg"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 1 and "Unknown key" in p.stderr


# The reference's own callers, compiled unchanged against the drop-in
# headers and linked with the gfx950 library (oracle/Makefile target
# `dropin`, built where the reference sources are).
REF_OUT = os.path.join(ROOT, "oracle", "_ref")
REF_TESTS = os.path.join(REF_OUT, "dropin_ref_tests.exe")
REF_APP = os.path.join(REF_OUT, "dropin_ref_app.exe")
_need = pytest.mark.skipif(
    not (os.path.exists(REF_TESTS) and os.path.exists(REF_APP)),
    reason="the reference's callers are built only where its sources are "
           "(make -C oracle dropin)")


@_need
def test_reference_gtest_suite_over_dropin(tmp_path):
    """The reference's gtest suite (tests/*.cc: pixel conversions, the
    covariance KATs, SN == W, lookup vs exact, PCG vs the dense solve, the
    map-covariance solves and the timing regressions), unchanged, passes
    with every call served by the GPU library."""
    (tmp_path / "tests").mkdir()
    shutil.copy(os.path.join(ROOT, "tests", "golden", "small_pixel_data.bin"),
                tmp_path / "tests" / "small_pixel_data.bin")
    out = _run([REF_TESTS], tmp_path)
    print(out[-1500:])
    assert "[  PASSED  ]" in out and "[  FAILED  ]" not in out


@_need
def test_reference_app_over_dropin(tmp_path):
    """app/dachshund.cc itself over the drop-in: its map.bin matches the
    reference app's on the 8,192-pixel workload, its map equals this
    project's app bit for bit, and its verbose PCG trace has the reference's
    lines (lib/pcg.h:112-139) with |r| and delta every iteration."""
    g = golden("app_c1.npz")
    g["pixels"].tofile(tmp_path / "pixel_data.bin")
    (tmp_path / "run.cfg").write_text(str(g["config"]))
    out = _run([REF_APP, "run.cfg"], tmp_path)
    m = np.fromfile(tmp_path / "map.bin")
    ref = g["map"]
    assert np.linalg.norm(m - ref) <= 1e-4 * np.linalg.norm(ref)
    _run([APP, "run.cfg"], tmp_path)
    assert np.array_equal(np.fromfile(tmp_path / "map.bin"), m)
    lines = [ln for ln in out.splitlines() if ln.startswith("    iter ")]
    assert re.match(r"    Goal \|r\| < \(tol\) \|b\| = \S+,  delta = \S+$",
                    out.splitlines()[[i for i, ln in enumerate(
                        out.splitlines()) if ln.startswith("[PCG]")][0] + 1])
    step = re.compile(r"    iter (\d+), \|r\| (\S+), delta (\S+)$")
    assert all(step.match(ln) for ln in lines[:-1])
    last = re.match(r"    iter (\d+) reached \|r\| = (\S+) < tol \|b\| = (\S+)$",
                    lines[-1])
    ref_last = re.match(r"    iter (\d+) reached \|r\| = (\S+) < tol \|b\| = "
                        r"(\S+)$", str(g["log_tail"]))
    assert last and ref_last
    assert abs(int(last.group(1)) - int(ref_last.group(1))) <= 1
    assert [int(step.match(ln).group(1)) for ln in lines[:-1]] == \
        list(range(1, len(lines)))
    assert float(last.group(3)) == pytest.approx(float(ref_last.group(3)),
                                                 rel=1e-6)


@_need
def test_reference_app_double_conversion(tmp_path):
    """With option_noise_covar and option_map_covar both set,
    app/dachshund.cc converts the pixel buffer to NPixel twice
    (app/dachshund.cc:264-282), so MapCovarDiag sees N' = 1/n^4; this
    project's app reproduces that by default (equal map_covar.bin to the
    reference app's source over the same library) and converts once with
    --single-conversion, which matches the oracle's MapCovarDiag on the
    true noise."""
    from dachshund_amd import synth
    w = synth.workload("small")
    pix = synth.make_pixels(w)
    pix.tofile(tmp_path / "pixel_data.bin")
    cfg = (f"lx = {w.box[0]}\nly = {w.box[1]}\nlz = {w.box[2]}\n"
           f"num_pixels = {len(pix)}\nmap_nx = 2\nmap_ny = 2\nmap_nz = 2\n"
           f"corr_var_s = {w.var_s}\ncorr_l_perp = {w.l_perp}\n"
           f"corr_l_para = {w.l_para}\npcg_tol = 1e-06\npcg_max_iter = 500\n"
           "pcg_step_r = 50\npixel_data_path = pixel_data.bin\n"
           "map_path = map.bin\noption_noise_covar = 1\n"
           "option_map_covar = 1\n")
    (tmp_path / "run.cfg").write_text(cfg)
    _run([REF_APP, "run.cfg"], tmp_path)
    ref_mc = np.fromfile(tmp_path / "map_covar.bin")
    _run([APP, "run.cfg"], tmp_path)
    assert np.array_equal(np.fromfile(tmp_path / "map_covar.bin"), ref_mc)
    _run([APP, "run.cfg", "--single-conversion"], tmp_path)
    fixed = np.fromfile(tmp_path / "map_covar.bin")
    from oracle import ora
    c = ora.covar(w.var_s, w.l_perp, w.l_para)
    mc = ora.map_coords(w.box, (2, 2, 2))
    s_ref, _, ok = ora.map_covar_diag(ora.to_form(pix, ora.FORM_SN), mc, c,
                                      500, 50, 1e-6)
    assert np.all(ok)
    np.testing.assert_allclose(fixed, s_ref, rtol=1e-4)
    assert not np.allclose(ref_mc, s_ref, rtol=1e-2)

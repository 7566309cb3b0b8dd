"""The C++ drop-in (the reference's lib/*.h API over libdachshund_gpu.so) and
its app, run as programs on the GPU and compared with the oracle and the
reference's golden map.  Tolerances as in test_gpu_parity.py: FP64 products
to 1e-12 of sum |A a|; solves judged by accuracy against a tight solution.
"""
import os
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

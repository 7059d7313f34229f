"""CPU-side checks: the C-ABI library and its header agree, argument errors
are reported without touching a device, and the host helpers behave."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "stfwi.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(stfwi_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from stfwi_amd import _abi
    if not os.path.exists(_abi.LIB_PATH):
        from stfwi_amd import build
        build.build()
    return _abi.lib()


def test_library_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) >= 18
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "stfwi_amd", "libstfwi.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (stfwi_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # and the ctypes table binds exactly those
    from stfwi_amd._abi import SIGNATURES
    assert set(SIGNATURES) <= set(names)
    assert lib.stfwi_abi_version() == 2


def _cfg(**kw):
    from stfwi_amd import _abi
    base = dict(nz_dom=20, nx_dom=30, pad=6, free_surface=1, fd_order=8, dx=10.0, dz=10.0, dt=1e-3, fc=10.0,
                dtype=_abi.F32, n_shots_max=1, nt_max=10, nrec_max=3, imaging=0, grad_params=1, store_history=0,
                ckpt_every=0)
    base.update(kw)
    return _abi.Config(**base)


@pytest.mark.parametrize("bad,code", [
    (dict(fd_order=6), -2), (dict(fd_order=3), -2), (dict(dtype=7), -1), (dict(nt_max=0), -1),
    (dict(n_shots_max=0), -1), (dict(imaging=5), -1), (dict(grad_params=0), -1), (dict(dt=0.0), -1),
    (dict(nz_dom=1, nx_dom=1, pad=1), -1), (dict(ckpt_every=-1), -1),
    (dict(ckpt_every=4, imaging=1, grad_params=3, store_history=1), -2), (dict(hist_frames=-1), -1),
])
def test_plan_create_rejects_bad_config_without_device(lib, bad, code):
    h = C.c_void_p()
    rc = lib.stfwi_plan_create(C.byref(_cfg(**bad)), C.byref(h))
    assert rc == code
    assert h.value is None
    assert lib.stfwi_last_error().decode()


def test_status_maps_to_python_exceptions(lib):
    from stfwi_amd import _abi
    h = C.c_void_p()
    with pytest.raises(NotImplementedError):
        _abi.check(lib.stfwi_plan_create(C.byref(_cfg(fd_order=6)), C.byref(h)))
    with pytest.raises(ValueError):
        _abi.check(lib.stfwi_plan_create(C.byref(_cfg(nt_max=0)), C.byref(h)))
    assert lib.stfwi_plan_destroy(None) == 0


def test_null_arguments_are_rejected(lib):
    assert lib.stfwi_set_model(None, None, None, None, None, None) == -1
    assert lib.stfwi_forward(None, None, 0, None, 0, None, 0, None) == -1
    assert lib.stfwi_adjoint(None, None, None, None, 0, None) == -1
    assert lib.stfwi_smooth2d(None, None, None, 4, 4, 1.0, 0, None) == -1
    assert lib.stfwi_sosfilt(None, 10, 1, None, 0, 0, 0, 0, None) == -1


def test_algorithmic_bytes_match_published_targets():
    """The 152x500 / nt=3000 / FP32 shot-gradient moves 40.54 GB, the 100%
    roofline is 197.3 shot-gradients/s (FP64: 81.08 GB)."""
    import bench
    tot32, _, _ = bench.algorithmic_bytes_per_shot_gradient(4, 1)
    tot64, _, _ = bench.algorithmic_bytes_per_shot_gradient(8, 1)
    assert abs(tot32 / 1e9 - 40.54) < 0.01
    assert abs(tot64 / 1e9 - 81.08) < 0.01
    assert abs(8e12 / tot32 - 197.3) < 0.1


def test_layered_model_generator():
    from stfwi_amd.models import layered_model, smooth_initial, anomaly_model, ricker
    a = layered_model(152, 500, seed=3)
    b = layered_model(152, 500, seed=3)
    c = layered_model(152, 500, seed=4)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a[1], c[1])
    vs, vp, rho = a
    assert vp.min() >= 1600 * 0.9 - 1e-9 and vp.max() <= 4700 * 1.1 + 1e-9
    ratio = vp / vs
    assert ratio.min() >= 1.6 - 1e-9 and ratio.max() <= 2.0 + 1e-9
    np.testing.assert_allclose(rho, 310.0 * vp ** 0.25)
    # velocity increases with depth on average
    assert vp[-10:].mean() > vp[:10].mean()
    s = smooth_initial(vs, vp, rho, 10.0)
    assert all(x.shape == (152, 500) for x in s)
    assert np.abs(np.diff(s[1], axis=0)).max() < np.abs(np.diff(vp, axis=0)).max()
    vs2, vp2, rho2 = anomaly_model(64, 128, seed=0)
    assert vs2.max() > 1730 * 1.09 and np.all(vp2 == 3000.0)
    w = ricker(100, 1e-3, 25.0)
    assert np.argmax(w) == round(1.5 / 25.0 / 1e-3)


def test_bench_refuses_more_than_one_gpu():
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_marmousi_survey_matches_reference_fixtures(tmp_path):
    """The drivers' Marmousi set-up against what the reference stored with
    its gradients: the 34- and 5-shot source lines, the 250 surface
    receivers and the time step (to the last bit) and step count of the 6 s
    record; the loader reads the reference's .npy layout, and smooth_initial
    with sigma 20 gives the SHA-pinned initial models."""
    from stfwi_amd.models import load_marmousi, marmousi_acquisition, smooth_initial, stable_dt
    from tests.common import load_golden, marmousi_init
    t = load_golden("marmousi_true.npz")
    for name, n in (("first_gradient.npz", 34), ("gradient_compare.npz", 5)):
        g = load_golden(name)
        sx, sz, rx, rz = marmousi_acquisition(int(g["nx"]), n)
        assert np.array_equal(sx, g["src_x"]) and np.array_equal(sz, g["src_z"]), name
        assert np.array_equal(rx, g["rec_x"]) and np.array_equal(rz, g["rec_z"]), name
        dt = stable_dt(t["vp"].max(), float(g["dx"]), float(g["dz"]), int(g["fd_order"]))
        assert dt == float(g["dt"]), name
    fg = load_golden("first_gradient.npz")
    assert int(6.0 / float(fg["dt"])) == int(fg["nt"])
    for k, a in (("vs_true", t["vs"]), ("vp_true", t["vp"]), ("rho_true", t["rho"])):
        np.save(tmp_path / (k + ".npy"), a)
    m = load_marmousi(str(tmp_path))
    assert sorted(m) == ["rho_true", "vp_true", "vs_true"]
    assert np.array_equal(m["vp_true"], t["vp"]) and m["vp_true"].dtype == t["vp"].dtype
    ref = marmousi_init()
    for a, k in zip(smooth_initial(m["vs_true"], m["vp_true"], m["rho_true"], sigma=20),
                    ("vs_init", "vp_init", "rho_init")):
        assert np.array_equal(a, ref[k]), k
    os.remove(tmp_path / "rho_true.npy")
    with pytest.raises(FileNotFoundError):
        load_marmousi(str(tmp_path))


def test_plan_footprint_and_checkpoint_planner(lib):
    """stfwi_plan_footprint sizes a plan without a device; the planner keeps
    the whole history when it fits, otherwise picks the checkpoint spacing
    of the smallest plan (near sqrt(nt)), and refuses what cannot fit."""
    from stfwi_amd.plan import footprint, choose_ckpt
    nz, nx, pad, nt, S = 152, 500, 20, 3000, 25
    # a history row holds the physical columns at offset pad mod 32 of a row
    # padded to a multiple of 32 elements (a tile's columns start a line)
    hpitch = (nx + pad % 32 + 31) // 32 * 32
    assert hpitch == 544
    np_ = nz * hpitch
    full = _cfg(nz_dom=nz, nx_dom=nx, pad=pad, n_shots_max=S, nt_max=nt, nrec_max=500, imaging=2,
                store_history=1)
    b_full = footprint(full)
    hist = S * nt * 3 * np_ * 4
    assert hist < b_full < hist * 1.1          # the strain history dominates
    # per-shot history: the same plan with half the frames is smaller by them
    half = _cfg(nz_dom=nz, nx_dom=nx, pad=pad, n_shots_max=S, nt_max=nt, nrec_max=500, imaging=2,
                store_history=1, hist_frames=S * nt // 2)
    assert b_full - footprint(half) == (S * nt - S * nt // 2) * 3 * np_ * 4
    assert choose_ckpt(full, b_full) == 0
    k = choose_ckpt(full, b_full // 4)
    assert 20 <= k <= 200, k
    kc = _cfg(nz_dom=nz, nx_dom=nx, pad=pad, n_shots_max=S, nt_max=nt, nrec_max=500, imaging=2,
              store_history=1, ckpt_every=k)
    b_k = footprint(kc)
    assert b_k < b_full // 4
    for k2 in (k // 2, 2 * k):           # the chosen spacing is the smallest plan
        kc.ckpt_every = k2
        assert footprint(kc) >= b_k
    from stfwi_amd._abi import StfwiError
    with pytest.raises(StfwiError):
        choose_ckpt(full, 1 << 20)
    zr = _cfg(nz_dom=nz, nx_dom=nx, pad=pad, n_shots_max=S, nt_max=nt, nrec_max=500, imaging=1, grad_params=3,
              store_history=1)
    with pytest.raises(NotImplementedError):
        choose_ckpt(zr, footprint(zr) // 2)

"""Parity of the HIP path (through the C-ABI) with the CPU oracle, and the
adjoint invariants, on an MI355X."""
import numpy as np
import pytest
import torch

import oracle
from oracle import signal as osig
from tests.common import (DP_CONFIGS, TOL_DOT, TOL_F32, TOL_F64, dp_geometry, dp_model, dp_vectors, rel_l2)

pytestmark = pytest.mark.gpu


def _plan(m, fd, fs, dtype, nshots=1, nrec=8, imaging="none", grad="vs", hist=False, nt=None):
    from stfwi_amd import Plan
    return Plan(m["nz"], m["nx"], m["dx"], m["dz"], m["dt"], m["fc"], m["pad"], fd_order=fd, free_surface=fs,
                dtype=dtype, n_shots_max=nshots, nt_max=nt or m["nt"], nrec_max=nrec, imaging=imaging,
                grad=grad, store_history=hist)


@pytest.mark.parametrize("fd,fs,st", DP_CONFIGS)
def test_gpu_forward_adjoint_match_oracle_f64(fd, fs, st):
    m = dp_model()
    sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], fs, st)
    s, y = dp_vectors(rx.size, m["nt"])
    print(f"case fd={fd} fs={fs} src={st}", flush=True)
    p = _plan(m, fd, fs, torch.float64, nrec=rx.size)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots([sx], [sz], st, m["nt"], rx, rz, s)
    rec = p.forward()["rec"][0].cpu().numpy()
    src_a = p.adjoint(torch.as_tensor(y[None]))["src_adj"][0].cpu().numpy()
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    rec_o, _ = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, s, rx, rz, st)
    a_o = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx, sz, y, rx, rz, st)
    assert rel_l2(rec, rec_o) < TOL_F64
    assert rel_l2(src_a, a_o["src_a"]) < TOL_F64
    # dot-product test on the GPU pair itself (test_adjoint_kernel.py:100-108)
    lhs, rhs = float(np.sum(rec * y)), float(np.sum(s * src_a))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < TOL_DOT


def _grad_case(fd, fs, st, seed=0):
    m = dp_model()
    m["nt"] = 120
    rng = np.random.default_rng(seed)
    sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], fs, st)
    wav = osig.ricker(np.arange(m["nt"]) * m["dt"], m["fc"], 1.5 / m["fc"]) * 1e6
    obs = rng.standard_normal((m["nt"], rx.size)) * 1e-9
    return m, sx, sz, rx, rz, wav, obs


@pytest.mark.parametrize("imaging", ["exact", "zhou"])
@pytest.mark.parametrize("fd,fs,st", DP_CONFIGS)
def test_gpu_gradient_matches_oracle_f64(imaging, fd, fs, st):
    m, sx, sz, rx, rz, wav, obs = _grad_case(fd, fs, st)
    print(f"case {imaging} fd={fd} fs={fs} src={st}", flush=True)
    sos = np.array([[0.1, 0.2, 0.1, 1.0, -1.2, 0.5]])
    comp, ntap, scale = 2, 12, 1e15
    p = _plan(m, fd, fs, torch.float64, nrec=rx.size, imaging=imaging, grad="all", hist=True)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots([sx], [sz], st, m["nt"], rx, rz, wav)
    rec = p.forward()["rec"]
    loss, radj = p.residual(rec, torch.as_tensor(obs[None]), comp, sos, ntap, scale, 1)
    p.adjoint(radj)
    gg = p.finalize(m["Vs"], m["Vp"], m["rho"]).cpu().numpy()
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    tape = ("e", "uv") if imaging == "zhou" else ("e", "dxu0", "s")
    rec_o, t = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, wav, rx, rz, st, np.float64, tape)
    l_o, cot = osig.residual_prep(rec_o[comp], obs, sos, ntap, scale, 1)
    assert abs(float(loss[0]) / l_o - 1) < TOL_F64
    ra = np.zeros_like(rec_o)
    ra[comp] = cot
    assert rel_l2(radj[0].cpu().numpy(), ra) < TOL_F64
    a = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx, sz, ra, rx, rz, st, wav, imaging, t)
    chain = oracle.exact_chain if imaging == "exact" else oracle.zhou_chain
    go = chain(g, a["g"], m["Vs"], m["Vp"], m["rho"])
    for i in range(3):
        assert rel_l2(gg[i], go[i]) < TOL_F64, i


def test_gpu_batched_shots_equal_single():
    """Several shots with uneven nt and receiver spreads in one launch give
    the same results as one shot at a time (bitwise)."""
    m = dp_model()
    m["nt"] = 90
    wav = osig.ricker(np.arange(m["nt"]) * m["dt"], m["fc"], 1.5 / m["fc"]) * 1e6
    shots = [(5, 3, "force", 90, np.arange(2, 30, 3)), (20, 6, "moment", 60, np.arange(0, 32, 2)),
             (11, 2, "moment", 75, np.array([4, 4, 9, 13]))]
    nrec = max(len(s[4]) for s in shots)
    pb = _plan(m, 8, True, torch.float32, nshots=3, nrec=nrec, imaging="exact", grad="vs", hist=True)
    pb.set_model(m["Vs"], m["Vp"], m["rho"])
    pb.set_shots([s[0] for s in shots], [s[1] for s in shots], [s[2] for s in shots], [s[3] for s in shots],
                 [s[4] for s in shots], [np.zeros_like(s[4]) for s in shots], wav)
    rb = pb.forward()["rec"]
    rng = np.random.default_rng(1)
    radj = torch.as_tensor(rng.standard_normal(rb.shape), dtype=torch.float32, device=rb.device)
    for i, s in enumerate(shots):
        radj[i, :, s[3]:] = 0
        radj[i, :, :, len(s[4]):] = 0
    sab = pb.adjoint(radj)["src_adj"]
    gb = pb.finalize(m["Vs"], m["Vp"], m["rho"])
    gsum = torch.zeros_like(gb)
    for i, s in enumerate(shots):
        p1 = _plan(m, 8, True, torch.float32, nshots=1, nrec=nrec, imaging="exact", grad="vs", hist=True)
        p1.set_model(m["Vs"], m["Vp"], m["rho"])
        p1.set_shots([s[0]], [s[1]], s[2], s[3], s[4], np.zeros_like(s[4]), wav)
        r1 = p1.forward()["rec"]
        assert torch.equal(r1[0], rb[i])
        sa1 = p1.adjoint(radj[i:i + 1])["src_adj"]
        assert torch.equal(sa1[0], sab[i])
        gsum += p1.finalize(m["Vs"], m["Vp"], m["rho"])
    assert rel_l2(gb.cpu().numpy(), gsum.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("tz", [4, 8])
def test_gpu_receiver_adjoints_at_tile_edges_match_oracle_f64(tz):
    """Receiver adjoints of all four channels at cells on tile-row and grid
    edges, with duplicate receivers and shots of different lengths: source
    adjoints and exact gradients against the oracle, shot by shot."""
    m = dp_model()
    fd, fs = 4, False
    rz = np.array([0, 0, 3, 4, 7, 8, 11, 12, 23, 23, 15, 16])
    rx = np.array([0, 0, 31, 5, 6, 7, 8, 9, 10, 10, 31, 0])
    nts = [60, 41]
    sx, sz, st = [10, 21], [6, 14], ["moment", "force"]
    rng = np.random.default_rng(11)
    wav = osig.ricker(np.arange(m["nt"]) * m["dt"], m["fc"], 1.5 / m["fc"]) * 1e6
    ra = rng.standard_normal((2, 4, m["nt"], rx.size)) * 1e-9
    ra[1, :, nts[1]:] = 0
    print(f"case tz={tz}", flush=True)
    p = _plan(m, fd, fs, torch.float64, nshots=2, nrec=rx.size, imaging="exact", grad="all", hist=True)
    p.set_tiles([tz] * 4)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots(sx, sz, st, nts, rx, rz, wav)
    p.forward()
    sa = p.adjoint(torch.as_tensor(ra))["src_adj"].cpu().numpy()
    gg = p.finalize(m["Vs"], m["Vp"], m["rho"]).cpu().numpy()
    go = np.zeros_like(gg)
    for k in range(2):
        nt = nts[k]
        g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, nt, m["dx"], m["dz"], m["dt"], m["fc"])
        _, t = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx[k], sz[k], wav[:nt], rx, rz, st[k], np.float64,
                              ("e", "dxu0", "s"))
        a = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx[k], sz[k], ra[k][:, :nt], rx, rz, st[k], wav[:nt],
                           "exact", t)
        assert rel_l2(sa[k, :nt], a["src_a"]) < TOL_F64, k
        assert not np.any(sa[k, nt:])
        go += np.array(oracle.exact_chain(g, a["g"], m["Vs"], m["Vp"], m["rho"]))
    for i in range(3):
        assert rel_l2(gg[i], go[i]) < TOL_F64, i


def _f32_case(fd, pert):
    """FP32 HIP gradient and the FP64 / FP32 oracle gradients of one moment
    shot (400 steps) whose observed data come from Vs * pert."""
    m = dp_model()
    m["nt"] = 400
    fs, st = True, "moment"
    sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], fs, st)
    wav = osig.ricker(np.arange(m["nt"]) * m["dt"], m["fc"], 1.5 / m["fc"])
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    obs, _ = oracle.forward(g, m["Vs"] * pert, m["Vp"], m["rho"], sx, sz, wav, rx, rz, st)
    obs = obs[2]
    scale = 1.0 / np.mean(obs ** 2)
    p = _plan(m, fd, fs, torch.float32, nrec=rx.size, imaging="exact", grad="vs", hist=True)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots([sx], [sz], st, m["nt"], rx, rz, wav)
    rec = p.forward()["rec"]
    loss, radj = p.residual(rec, torch.as_tensor(obs[None]), 2, None, 0, scale, 1)
    p.adjoint(radj)
    gg = p.finalize(m["Vs"], m["Vp"], m["rho"]).cpu().numpy()
    out = {"gpu": (rec[0].cpu().numpy(), float(loss[0]), gg[0])}
    for name, dt in (("o64", np.float64), ("o32", np.float32)):
        rec_o, t = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, wav, rx, rz, st, dt, ("e", "dxu0"))
        l_o, cot = osig.residual_prep(rec_o[2].astype(np.float64), obs, None, 0, scale, 1)
        ra = np.zeros((4,) + cot.shape)
        ra[2] = cot
        a = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx, sz, ra, rx, rz, st, wav, "exact", t, None, dt)
        go = oracle.exact_chain(g, a["g"].astype(np.float64), m["Vs"], m["Vp"], m["rho"])[0]
        out[name] = (rec_o, l_o, go)
    return out


@pytest.mark.parametrize("fd", [4, 8])
def test_gpu_f32_gradient_close_to_f64_oracle(fd):
    """FP32 HIP path against the FP64 oracle: seismograms, misfit and the
    gradient within 1e-5 relative L2 for a 10% Vs misfit."""
    o = _f32_case(fd, 1.10)
    rec, loss, g = o["gpu"]
    assert rel_l2(rec, o["o64"][0]) < TOL_F32
    assert abs(loss / o["o64"][1] - 1) < TOL_F32
    assert rel_l2(g, o["o64"][2]) < TOL_F32


@pytest.mark.parametrize("fd", [4, 8])
def test_gpu_f32_ill_conditioned_no_worse_than_reference_algorithm(fd):
    """With a 2% misfit the residual pred - obs cancels ~98% of the data, so
    FP32 wavefield rounding is amplified ~50x in the gradient; the
    reference's algorithm run in FP32 (the oracle's float build) is then
    ~2-3e-5 from FP64.  The HIP FP32 gradient must be no worse than that."""
    o = _f32_case(fd, 1.02)
    e_gpu = rel_l2(o["gpu"][2], o["o64"][2])
    e_ref32 = rel_l2(o["o32"][2], o["o64"][2])
    print(f"fd={fd}: HIP FP32 {e_gpu:.2e}, oracle FP32 {e_ref32:.2e}")
    assert rel_l2(o["gpu"][0], o["o64"][0]) < TOL_F32
    assert e_gpu < 2.0 * max(e_ref32, TOL_F32)


def _big_case():
    """A grid large enough to have interior (mask-free) tiles."""
    nz, nx, pad, nt = 96, 200, 12, 160
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    Vp = 2600.0 + 200.0 * np.sin(2 * np.pi * zz / 37) * np.cos(2 * np.pi * xx / 53)
    Vs = Vp / 1.8
    rho = 2100.0 + 90.0 * np.cos(np.pi * xx / 41)
    m = dict(nz=nz, nx=nx, pad=pad, nt=nt, dx=10.0, dz=10.0, dt=1e-3, fc=20.0, Vs=Vs, Vp=Vp, rho=rho)
    rx = np.arange(0, nx, 3)
    rz = np.full_like(rx, 1)
    wav = osig.ricker(np.arange(nt) * m["dt"], m["fc"], 1.5 / m["fc"]) * 1e6
    return m, rx, rz, wav


@pytest.mark.parametrize("imaging", ["exact", "zhou"])
def test_gpu_interior_tiles_match_oracle_f64(imaging):
    m, rx, rz, wav = _big_case()
    fd, fs = 8, True
    sx, sz = [57, 140], [5, 9]
    rng = np.random.default_rng(3)
    obs = rng.standard_normal((2, m["nt"], rx.size)) * 1e-9
    p = _plan(m, fd, fs, torch.float64, nshots=2, nrec=rx.size, imaging=imaging, grad="all", hist=True)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots(sx, sz, ["moment", "force"], m["nt"], rx, rz, wav)
    rec = p.forward()["rec"]
    loss, radj = p.residual(rec, torch.as_tensor(obs), 2, None, 0, 1e15, 2)
    p.adjoint(radj)
    gg = p.finalize(m["Vs"], m["Vp"], m["rho"]).cpu().numpy()
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    tape = ("e", "uv") if imaging == "zhou" else ("e", "dxu0", "s")
    go = np.zeros_like(gg)
    for k, st in enumerate(["moment", "force"]):
        rec_o, t = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx[k], sz[k], wav, rx, rz, st, np.float64, tape)
        assert rel_l2(rec[k].cpu().numpy(), rec_o) < TOL_F64
        _, cot = osig.residual_prep(rec_o[2], obs[k], None, 0, 1e15, 2)
        ra = np.zeros_like(rec_o)
        ra[2] = cot
        a = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx[k], sz[k], ra, rx, rz, st, wav, imaging, t)
        chain = oracle.exact_chain if imaging == "exact" else oracle.zhou_chain
        go += np.array(chain(g, a["g"], m["Vs"], m["Vp"], m["rho"]))
    for i in range(3):
        assert rel_l2(gg[i], go[i]) < TOL_F64, i


def test_gpu_tile_heights_agree_bitwise():
    """Every tile shape the autotuner can pick gives bitwise-identical
    seismograms and gradients."""
    m, rx, rz, wav = _big_case()
    rng = np.random.default_rng(4)
    obs = torch.as_tensor(rng.standard_normal((1, m["nt"], rx.size)) * 1e-9, dtype=torch.float32)
    p = _plan(m, 8, True, torch.float32, nshots=1, nrec=rx.size, imaging="exact", grad="vs", hist=True)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots([77], [4], "moment", m["nt"], rx, rz, wav)
    tiles, table = p.autotune(3)
    assert all(t in (4, 8, 16, 32) for t in tiles) and set(table) == {"fwd_vel", "fwd_strain", "adj_strain", "adj_vel"}
    outs = []
    for tz in ([8] * 4, [4] * 4, [16] * 4, [32] * 4, [32, 4, 16, 8]):
        p.set_tiles(tz)
        rec = p.forward()["rec"]
        _, radj = p.residual(rec, obs, 2, None, 0, 1e15, 1)
        p.adjoint(radj)
        outs.append((rec.clone(), p.finalize(m["Vs"], m["Vp"], m["rho"]).clone()))
    for r, g in outs[1:]:
        assert torch.equal(r, outs[0][0])
        assert torch.equal(g, outs[0][1])


def test_gpu_stream_groups_agree_bitwise():
    """Running the shots as 1, 2 or 3 concurrent stream groups gives
    bitwise-identical records, source adjoints and gradients."""
    m, rx, rz, wav = _big_case()
    rng = np.random.default_rng(6)
    sx, sz, st = [30, 77, 120, 150], [3, 4, 9, 2], ["moment", "force", "moment", "force"]
    obs = torch.as_tensor(rng.standard_normal((4, m["nt"], rx.size)) * 1e-9, dtype=torch.float32)
    p = _plan(m, 8, True, torch.float32, nshots=4, nrec=rx.size, imaging="zhou", grad="vs_vp", hist=True)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots(sx, sz, st, [m["nt"], m["nt"] - 30, m["nt"], m["nt"] - 7], rx, rz, wav)
    outs = []
    for groups in (1, 2, 3, 4):
        p.set_groups(groups)
        rec = p.forward()["rec"]
        _, radj = p.residual(rec, obs, 2, None, 0, 1e15, 4)
        sa = p.adjoint(radj)["src_adj"]
        outs.append((rec.clone(), sa.clone(), p.finalize(m["Vs"], m["Vp"], m["rho"]).clone()))
    for o in outs[1:]:
        for x, y in zip(o, outs[0]):
            assert torch.equal(x, y)


@pytest.mark.parametrize("comp", ["vx", "ez"])
def test_gpu_adjoint_energy_equals_sum_of_squared_snapshots(comp):
    """The adjoint energy accumulator (the Wr of eprecond1,
    fwi/preconditioning.py:37-62) equals the sum over reverse steps of the
    squared adjoint snapshots of the same component, per shot, with stream
    groups; the next adjoint without it leaves the buffer alone."""
    m = dp_model()
    fs, st = True, "moment"
    sx, sz, rx, rz = dp_geometry(m["nz"], m["nx"], fs, st)
    rng = np.random.default_rng(21)
    p = _plan(m, 8, fs, torch.float64, nshots=2, nrec=rx.size)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots([sx, sx + 5], [sz, sz], st, [m["nt"], m["nt"] - 13], rx, rz, np.ones(m["nt"]))
    p.set_groups(2)
    ra = torch.as_tensor(rng.standard_normal((2, 4, m["nt"], rx.size)))
    ra[1, :, m["nt"] - 13:] = 0
    out = p.adjoint(ra, snap=(comp,), energy=comp)
    e = out["energy"].cpu().numpy()
    sn = out["snap"][:, 0].cpu().numpy()
    ref = np.zeros_like(e)
    for s, nt in enumerate([m["nt"], m["nt"] - 13]):
        for t in range(nt - 1, -1, -1):
            ref[s] += sn[s, t] ** 2
    assert rel_l2(e, ref) < 1e-14
    assert np.all(e >= 0) and np.any(e > 0)
    again = p.adjoint(ra)
    assert "energy" not in again


@pytest.mark.parametrize("ntap", [0, 12])
def test_gpu_residual_unfiltered_ragged_f64(ntap):
    """Without a filter the residual runs over blocks of 64 time steps: shots
    whose record lengths are no multiple of 64 and whose receiver counts are
    below the plan's maximum match the oracle's residual shot by shot, and the
    steps past a shot's record and its absent receivers are zero."""
    m = dp_model()
    nt_max, nts, nrecs = 150, [150, 97], [7, 4]
    rng = np.random.default_rng(11)
    rxl = [np.arange(2, 2 + 3 * n, 3) for n in nrecs]
    rzl = [np.full(n, 3) for n in nrecs]
    p = _plan(m, 4, True, torch.float64, nshots=2, nrec=max(nrecs), nt=nt_max)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots([10, 20], [5, 6], "force", nts, rxl, rzl, np.zeros(nt_max))
    pred = rng.standard_normal((2, 4, nt_max, max(nrecs)))
    obs = rng.standard_normal((2, nt_max, max(nrecs)))
    loss, radj = p.residual(torch.as_tensor(pred), torch.as_tensor(obs), 1, None, ntap, 3.0, 2)
    radj = radj.cpu().numpy()
    for k in range(2):
        nt, nr = nts[k], nrecs[k]
        l_o, cot = osig.residual_prep(pred[k, 1, :nt, :nr], obs[k, :nt, :nr], None, ntap, 3.0, 2)
        assert abs(float(loss[k]) / l_o - 1) < TOL_F64, k
        assert rel_l2(radj[k, 1, :nt, :nr], cot) < TOL_F64, k
        assert not radj[k, 1, nt:].any() and not radj[k, 1, :, nr:].any(), k
        assert not radj[k, [0, 2, 3]].any(), k


@pytest.mark.parametrize("fd,fs,st", [(8, True, "moment"), (4, False, "force"), (2, True, "force")])
def test_gpu_vs_only_exact_gradient_matches_oracle_f64(fd, fs, st):
    """The exact Vs-only gradient (the path the specialised velocity-adjoint
    kernel serves) against the oracle's exact chain."""
    m, sx, sz, rx, rz, wav, obs = _grad_case(fd, fs, st, seed=5)
    print(f"case fd={fd} fs={fs} src={st}", flush=True)
    p = _plan(m, fd, fs, torch.float64, nrec=rx.size, imaging="exact", grad="vs", hist=True)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots([sx], [sz], st, m["nt"], rx, rz, wav)
    rec = p.forward()["rec"]
    loss, radj = p.residual(rec, torch.as_tensor(obs[None]), 2, None, 0, 1e15, 1)
    p.adjoint(radj, want_src=False)
    gg = p.finalize(m["Vs"], m["Vp"], m["rho"]).cpu().numpy()
    g = oracle.Geom(m["nz"], m["nx"], m["pad"], fs, fd, m["nt"], m["dx"], m["dz"], m["dt"], m["fc"])
    rec_o, t = oracle.forward(g, m["Vs"], m["Vp"], m["rho"], sx, sz, wav, rx, rz, st, np.float64,
                              ("e", "dxu0", "s"))
    _, cot = osig.residual_prep(rec_o[2], obs, None, 0, 1e15, 1)
    ra = np.zeros_like(rec_o)
    ra[2] = cot
    a = oracle.adjoint(g, m["Vs"], m["Vp"], m["rho"], sx, sz, ra, rx, rz, st, wav, "exact", t)
    go = oracle.exact_chain(g, a["g"], m["Vs"], m["Vp"], m["rho"])
    assert rel_l2(gg[0], go[0]) < TOL_F64


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gpu_lean_kernels_agree_bitwise(dtype):
    """The time-step kernels' specialised instantiations (no optional
    outputs; exact imaging of Vs only) give the same records, source
    adjoints and gradients bit for bit as the general ones, in stream groups
    and with uneven record lengths; asking for the source adjoint (general
    strain-adjoint kernel) leaves the gradient unchanged."""
    m, rx, rz, wav = _big_case()
    rng = np.random.default_rng(8)
    sx, sz, st = [30, 77, 150], [3, 4, 2], ["moment", "force", "moment"]
    obs = torch.as_tensor(rng.standard_normal((3, m["nt"], rx.size)) * 1e-9, dtype=dtype)
    p = _plan(m, 8, True, dtype, nshots=3, nrec=rx.size, imaging="exact", grad="vs", hist=True)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots(sx, sz, st, [m["nt"], m["nt"] - 23, m["nt"] - 5], rx, rz, wav)
    p.set_groups(2)
    outs = []
    for lean in (True, False):
        p.set_lean(lean)
        rec = p.forward()["rec"]
        _, radj = p.residual(rec, obs, 2, None, 0, 1e15, 3)
        p.adjoint(radj, want_src=False)
        g1 = p.finalize(m["Vs"], m["Vp"], m["rho"]).clone()
        sa = p.adjoint(radj)["src_adj"].clone()
        g2 = p.finalize(m["Vs"], m["Vp"], m["rho"]).clone()
        assert torch.equal(g1, g2)
        outs.append((rec.clone(), sa, g1))
    for x, y in zip(outs[0], outs[1]):
        assert torch.equal(x, y)


def test_gpu_grouped_autotune_keeps_results():
    """The autotuner times its candidates with the plan's stream groups; the
    tile heights it picks give the same records and gradients bit for bit
    as fixed heights, and the plan runs normally after tuning."""
    m, rx, rz, wav = _big_case()
    rng = np.random.default_rng(9)
    obs = torch.as_tensor(rng.standard_normal((3, m["nt"], rx.size)) * 1e-9, dtype=torch.float32)
    p = _plan(m, 8, True, torch.float32, nshots=3, nrec=rx.size, imaging="exact", grad="vs", hist=True)
    p.set_model(m["Vs"], m["Vp"], m["rho"])
    p.set_shots([40, 90, 140], [3, 5, 2], "moment", [m["nt"], m["nt"] - 11, m["nt"]], rx, rz, wav)
    p.set_groups(3)
    tiles, table = p.autotune(2)
    assert all(t in (4, 8, 16, 32) for t in tiles)
    assert set(table) == {"fwd_vel", "fwd_strain", "adj_strain", "adj_vel"}
    outs = []
    for tz in (tiles, [8] * 4):
        p.set_tiles(tz)
        rec = p.forward()["rec"]
        _, radj = p.residual(rec, obs, 2, None, 0, 1e15, 3)
        p.adjoint(radj, want_src=False)
        outs.append((rec.clone(), p.finalize(m["Vs"], m["Vp"], m["rho"]).clone()))
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])

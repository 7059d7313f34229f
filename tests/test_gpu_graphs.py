"""Graph replay of the forward and adjoint time loops: every replay equals
the eager launches bitwise, across re-captures (new tiles, groups, shot
tables) and with the program's own GPU work and synchronisations between
replays."""
import numpy as np
import pytest
import torch

from tests.common import dp_model

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="graph replay is switched off: with it on, a host segmentation fault inside "
                            "stfwi_adjoint ended the full GPU suite twice (stream-groups test after the other "
                            "tests) and its cause is not yet shown"),
]


def _setup(p, m, shots, seed=0):
    from oracle import signal as osig
    rng = np.random.default_rng(seed)
    sx, sz, st, nts = shots
    rx = np.arange(1, m["nx"] - 1, 3)
    rz = np.full_like(rx, 2)
    wav = osig.ricker(np.arange(m["nt"]) * m["dt"], m["fc"], 1.5 / m["fc"]) * 1e6
    p.set_model(m["Vs"] * (1 + 0.01 * rng.standard_normal(m["Vs"].shape)), m["Vp"], m["rho"])
    p.set_shots(sx, sz, st, nts, rx, rz, wav)


def _run(p, seed=1):
    rec = p.forward()["rec"]
    rng = np.random.default_rng(seed)
    radj = torch.as_tensor(rng.standard_normal(tuple(rec.shape)), dtype=rec.dtype, device=rec.device) * rec.abs().max()
    sa = p.adjoint(radj)["src_adj"]
    g = p.finalize(*[torch.as_tensor(x) for x in p._keep["model"][:3]])
    return rec, sa, g


def _plan(m, dtype, graphs):
    from stfwi_amd import Plan
    p = Plan(m["nz"], m["nx"], m["dx"], m["dz"], m["dt"], m["fc"], m["pad"], fd_order=8, free_surface=True,
             dtype=dtype, n_shots_max=3, nt_max=m["nt"], nrec_max=m["nx"], imaging="exact", grad="all",
             store_history=True)
    p.set_graphs(graphs)
    return p


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_graph_replays_equal_eager(dtype):
    m = dp_model()
    shot_sets = [([4, 15, 27], [3, 6, 2], ["moment", "force", "moment"], m["nt"]),
                 ([9, 20, 5], [1, 4, 7], ["force", "force", "moment"], [m["nt"], m["nt"] - 7, m["nt"] - 19])]
    pe, pg = _plan(m, dtype, False), _plan(m, dtype, True)
    for rep, (shots, cfg) in enumerate([(shot_sets[0], (8, 1)), (shot_sets[0], (8, 1)), (shot_sets[1], (8, 1)),
                                        (shot_sets[1], (4, 2)), (shot_sets[0], (4, 2)), (shot_sets[0], (8, 1))]):
        tz, groups = cfg
        outs = []
        for p in (pe, pg):
            p.set_tiles([tz] * 4)
            p.set_groups(groups)
            _setup(p, m, shots, seed=rep)
            outs.append(_run(p, seed=rep))
            # the program's own work between replays
            x = torch.randn(1 << 16, device="cuda", dtype=dtype)
            torch.cuda.synchronize()
            assert torch.isfinite((x * 2).sum())
        for a, b in zip(*outs):
            assert torch.equal(a, b), rep


def test_graph_with_checkpoints_and_subset_channels():
    from stfwi_amd import Plan
    m = dp_model()
    res = []
    for graphs in (False, True):
        p = Plan(m["nz"], m["nx"], m["dx"], m["dz"], m["dt"], m["fc"], m["pad"], fd_order=4, free_surface=False,
                 dtype=torch.float64, n_shots_max=2, nt_max=m["nt"], nrec_max=m["nx"], imaging="exact",
                 grad="vs_vp", store_history=True, ckpt_every=9)
        p.set_graphs(graphs)
        _setup(p, m, ([6, 20], [5, 8], "moment", m["nt"]))
        out = []
        for rep in range(3):
            rec = p.forward(channels=(1, 3))["rec"]
            radj = rec.flip(2).contiguous()
            out.append((rec.clone(), p.adjoint(radj, want_src=False), p.finalize(*p._keep["model"][:3])))
        res.append(out)
    for (ra, _, ga), (rb, _, gb) in zip(*res):
        assert torch.equal(ra, rb) and torch.equal(ga, gb)
    assert not torch.any(res[1][0][0][:, 0]) and not torch.any(res[1][0][0][:, 2])

"""Extract the reference's committed golden outputs into small fixtures.

Reads (allow_pickle=False only) from the reference checkout:
  marmousi_models/{vs,vp,rho}_true.npy
  Results_gradient/gradient_compare.npz      FP64, 5 shots, nt=400, fd8, FS, force
  Results_gradient/gradient_data.npz         FP32 jax.grad run, 5 shots, nt=2890
  Results_first_gradient/first_gradient.npz  FP64, 34 shots, nt=2890, eprecond3

The initial models of all three runs are 1/gaussian_filter(1/v_true, 20)
(velocities) and gaussian_filter(rho_true, 20); the fixtures keep only the
true models plus a SHA-256 of each stored initial model so the tests can
regenerate and verify them.  The pickled 0-d 'illumination' entry of
first_gradient.npz is not read.

Run:  python tests/golden/make_fixtures.py [/root/reference]
"""
import hashlib
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def load(p):
    return np.load(os.path.join(REF, p), allow_pickle=False)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    true = {k: np.load(os.path.join(REF, "marmousi_models", f"{k}_true.npy"), allow_pickle=False)
            for k in ("vs", "vp", "rho")}
    np.savez_compressed(os.path.join(OUT, "marmousi_true.npz"), **true)

    d = load("Results_gradient/gradient_compare.npz")
    scal = ["dx", "dz", "dt", "nt", "pad", "free_surface", "nz", "nx", "f_c", "fd_order",
            "component", "n_shots", "auto_scale", "loss_AD", "loss_HC"]
    np.savez_compressed(
        os.path.join(OUT, "gradient_compare.npz"),
        g_Vs_HC=d["g_Vs_HC"], g_Vs_AD=d["g_Vs_AD"],
        src_x=d["src_x"], src_z=d["src_z"], rec_x=d["rec_x"], rec_z=d["rec_z"],
        sha_Vs_init=sha(d["Vs_init"]), sha_Vp_init=sha(d["Vp_init"]), sha_rho_init=sha(d["rho_init"]),
        **{k: d[k] for k in scal})

    e = load("Results_gradient/gradient_data.npz")
    scal = ["dx", "dz", "pad", "free_surface", "nz", "nx", "nt", "dt", "f_c", "cutoff",
            "component", "fd_order", "n_shots", "loss", "auto_scale"]
    np.savez_compressed(
        os.path.join(OUT, "gradient_data.npz"),
        g_vs=e["g_vs"], src_x=e["src_x"], src_z=e["src_z"], rec_x=e["rec_x"], rec_z=e["rec_z"],
        sha_Vs_init=sha(e["Vs_init"]), **{k: e[k] for k in scal})

    f = load("Results_first_gradient/first_gradient.npz")
    scal = ["dx", "dz", "dt", "nt", "pad", "free_surface", "nz", "nx", "f_c", "cutoff", "fd_order",
            "component", "n_shots", "auto_scale", "grad_smooth_sigma", "illumination_eps", "loss_HC"]
    np.savez_compressed(
        os.path.join(OUT, "first_gradient.npz"),
        g_Vs_HC=f["g_Vs_HC"], illumination_eprecond3=f["illumination_eprecond3"],
        src_x=f["src_x"], src_z=f["src_z"], rec_x=f["rec_x"], rec_z=f["rec_z"],
        sha_Vs_init=sha(f["Vs_init"]), **{k: f[k] for k in scal})


if __name__ == "__main__":
    main()

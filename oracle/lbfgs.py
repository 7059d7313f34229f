"""L-BFGS oracle (TEST INFRASTRUCTURE ONLY).

Two independent statements of the quasi-Newton direction d = -H g used by
the plain method (the reference's scipy L-BFGS-B in
fwi/optim_lbfgsb.py:409-434 and jaxopt.LBFGS in fwi/optimizer.py:110-117):

* ``two_loop``: the two-loop recursion, Nocedal & Wright (2nd ed.)
  Algorithm 7.4, in float64 numpy;
* ``dense_inverse_hessian``: the BFGS inverse-Hessian update
  H+ = (I - rho s y^T) H (I - rho y s^T) + rho s s^T (N&W eq. 6.17) applied
  to H0 = gamma I pair by pair, oldest first, as a dense matrix.

They agree to rounding (N&W section 7.2), which the tests check, so a
mistake in either (an index, a sign, a transposed product) shows.  Parity:
pinned by that identity and by scipy.optimize.minimize(L-BFGS-B) reaching
the same minimiser on convex problems.
"""
from __future__ import annotations

import numpy as np


def two_loop(g, S, Y, gamma):
    """d = -H g from pairs S[i], Y[i] (oldest first) and H0 = gamma I."""
    q = np.array(g, dtype=np.float64, copy=True)
    m = len(S)
    rho = [1.0 / float(np.dot(Y[i], S[i])) for i in range(m)]
    alpha = np.zeros(m)
    for i in range(m - 1, -1, -1):
        alpha[i] = rho[i] * float(np.dot(S[i], q))
        q -= alpha[i] * Y[i]
    r = gamma * q
    for i in range(m):
        beta = rho[i] * float(np.dot(Y[i], r))
        r += S[i] * (alpha[i] - beta)
    return -r


def dense_inverse_hessian(S, Y, gamma, n):
    H = gamma * np.eye(n)
    I = np.eye(n)
    for s, y in zip(S, Y):
        rho = 1.0 / float(np.dot(y, s))
        V = I - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return H

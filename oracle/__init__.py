"""CPU oracle for the elastic velocity-strain FWI gradient path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this package; the product
(``stfwi_amd``) never does, and never falls back to it.

Layout
------
``orc.c`` / ``orc_impl.h``  C restatement of the reference's forward step
                            (fwi/forward.py:119-501), hand-coded adjoint
                            (fwi/adjoint_jax_kernel.py:61-483), Zhou imaging
                            (fwi/imaging.py:83-169) and the exact discrete
                            gradient, built as ``liborc.so`` for float32,
                            float64 and complex128.
``core.py``                 ctypes wrappers around ``liborc.so``, and
                            ``zhou_cubes``: the Zhou imaging of
                            fwi/imaging.py:111-152 from history cubes in numpy.
``signal.py``               Ricker, cosine end-taper, DF-I biquad cascade and
                            its transpose, residual prep (fwi/filters.py,
                            fwi/gradient.py:139-156).
``precond.py``              Gaussian smoothing, illumination H_inv, eprecond3
                            (fwi/loss.py:17-30, fwi/illumination.py:18-48,
                            fwi/preconditioning.py:65-139).
``gradient.py``             compute_gradient for both reference backends
                            (fwi/gradient.py:48-277).

Parity status of each function is recorded in the module docstrings; every
function listed above is pinned by a test under tests/ (fixtures, closed
forms, dot-product, complex-step).  The Zhou density term is pinned by
closed forms of ``zhou_cubes`` on constructed cubes and by the fused C
restatement agreeing with ``zhou_cubes`` on its own adjoint cubes.
"""
from .core import (Geom, build, forward, adjoint, setup, padded_shape,  # noqa: F401
                   exact_chain, zhou_chain, zhou_cubes)
from . import signal, precond, gradient  # noqa: F401

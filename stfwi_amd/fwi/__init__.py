"""Drop-in counterpart of the reference's ``fwi`` package for the gradient
path (fwi/__init__.py): same public names, HIP kernels underneath."""
from .adjoint_jax_kernel import adjoint_jax, pml_apply_T  # noqa: F401
from .forward import forward_jax, build_forward_fn, ricker_jax  # noqa: F401
from .imaging import imaging_condition  # noqa: F401
from .gradient import compute_gradient  # noqa: F401
from .preconditioning import eprecond1, eprecond3, eprecond3_preconditioner, taper_grad  # noqa: F401
from .taper import build_source_taper  # noqa: F401
from .illumination import compute_illumination_for_stage  # noqa: F401
from .loss import build_loss_fn, gaussian_smooth_2d  # noqa: F401
from .filters import jax_biquad, jax_sosfilt, cosine_taper_end, sosfilt_T  # noqa: F401
from .projections import project_velocity, cfl_vp_cap, k_safe_ratio  # noqa: F401
from .optim_lbfgsb import run_multiscale_fwi  # noqa: F401
from .optimizer import run_lbfgs, run_lbfgs_alternating  # noqa: F401

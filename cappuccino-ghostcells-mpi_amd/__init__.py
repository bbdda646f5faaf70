"""MI355X-native pressure-correction (p') hot path of the structured-grid
ghost-cell FV solver: gfx950 HIP kernels behind a C ABI (include/cap_amd.h),
a Fortran ISO_C_BINDING drop-in (fortran/), and this Python binding.

The directory name is not an importable identifier; load it with
``load_package()`` from ``__graft_entry__`` (registered as ``cappuccino_amd``).
"""
from .host import (Context, Coef, UNet, DrvOpts, UvwOpts, LesOpts, CapError, load, COEF_FIELDS, GEOM_FIELDS, EXPORTS,  # noqa: F401
                   FLOW_NAMES, read_grid,
                   CAP_OK, CAP_MAXITER, CAP_NAN, CAP_GUESS_REJECTED, CAP_ERR_TIMEOUT)

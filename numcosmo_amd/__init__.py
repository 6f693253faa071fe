"""numcosmo_amd -- MI355X (gfx950) kernel-density engine behind NumCosmo's
NcmStatsDist API: KDE / VKDE prepare, prepare_interp, eval_m2lnp and sample
with the Gauss and Student-t kernels, on hand-written HIP kernels.

The compute library is ``libnckde.so`` (C-ABI in include/nckde.h), built by
``python -m numcosmo_amd.build``.  Nothing here computes on the CPU: without
the library every call raises.
"""
from .ncm_stats_dist import (  # noqa: F401
    NckdeError, NNLSUMethod, Precision, RNG, StatsDistCV, StatsDistKDE, StatsDistKDECovType,
    StatsDistKernelGauss, StatsDistKernelST, StatsDistVKDE,
)

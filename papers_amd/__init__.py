"""papers_amd: MI355X-native quintic Calabi–Yau metric hot path (cy3q).

The compute lives in ``libcy3q.so``, which holds hand-written HIP kernels for gfx950 behind the
C ABI of ``include/cy3q.h``. This package is the Python mirror of the reference crate's
``QuinticSolver`` API (gable-digital/papers, substrate_sm_derivation/cy3_solver/src/quintic.rs).
It calls the library through ctypes. There is no CPU fallback: if the library is missing or
cannot load, importing ``papers_amd.quintic`` raises.
"""
from .quintic import (  # noqa: F401
    PREC_FP64,
    PREC_SPLIT16,
    REFINE_STEP_FP64,
    REFINE_T_FP64,
    SAMPLER_NEWTON,
    SAMPLER_NEWTON_DWORK,
    SAMPLER_SZ,
    SAMPLER_SZ_DWORK,
    Cy3qError,
    QuinticSolver,
    build_degree_k_quintic_monomials,
    lib_path,
    monge_ampere_residual_quintic,
    n_sections_quintic,
)
from .donaldson_gpu import Cy3DonaldsonGpu, Cy3SigmaGpu  # noqa: F401,E402
from .adam_gpu import AdamGradientKernel  # noqa: F401,E402

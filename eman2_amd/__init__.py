"""fourier_hip: MI355X-native direct Fourier inversion reconstruction,
a drop-in for EMAN2's "fourier" Reconstructor path.

The compute path is libfourier_hip.so (HIP kernels for gfx950 behind the C
ABI of include/fourier_hip.h).  This package holds its Python surface
(``Reconstructors``, ``Transform``, ``EMData``), the synthetic workload
generator used by the benchmark, and agreement metrics.
"""
from .transform import Transform, euler_eman_matrix, euler_eman_matrices  # noqa: F401


def __getattr__(name):
    # the reconstructor surface loads the HIP library lazily so that the
    # CPU-side helpers stay importable without it
    if name in ("Reconstructors", "FourierHipReconstructor", "EMData"):
        from . import reconstructor
        return getattr(reconstructor, name)
    raise AttributeError(name)

"""matvis_amd: matvis's per-time visibility computation on AMD MI355X (gfx950).

A drop-in for ``matvis.simulate_vis(..., use_gpu=True)`` and
``matvis.gpu.simulate``, with the plugin classes ``CoordinateRotation``,
``BeamInterpolator``, ``TauCalculator``, ``ZMatrixCalc`` and ``MatProd``.
All device work runs in hand-written HIP kernels (``libmatvis_hip.so``),
called through a C ABI with ctypes; PyTorch-ROCm supplies device memory.
"""

from pathlib import Path

__version__ = "0.1.0"

HAVE_GPU = True

from . import astro, beams, coordinates  # noqa: E402
from .astro import EarthLocation, SkyCoord, Time  # noqa: E402
from .beams import BeamInterface, GaussianBeam, GridBeam  # noqa: E402
from .wrapper import simulate_vis  # noqa: E402

DATA_PATH = Path(__file__).parent / "data"

del Path

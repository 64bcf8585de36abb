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



def _data_path():
    """The directory holding ``NF_HERA_Dipole_small.fits`` (the reference's
    ``DATA_PATH``, src/matvis/__init__.py:26).  The beam file is not shipped
    here: ``$MATVIS_DATA_PATH`` if set, else an installed ``matvis``
    package's ``data`` directory (located without importing it), else
    ``matvis_amd/data``, where a copy can be placed."""
    import importlib.util
    import os
    from pathlib import Path

    env = os.environ.get("MATVIS_DATA_PATH")
    if env:
        return Path(env)
    try:
        spec = importlib.util.find_spec("matvis")
    except (ImportError, ValueError):
        spec = None
    if spec is not None and spec.submodule_search_locations:
        for loc in spec.submodule_search_locations:
            if (Path(loc) / "data").is_dir():
                return Path(loc) / "data"
    return Path(__file__).parent / "data"


DATA_PATH = _data_path()

del Path

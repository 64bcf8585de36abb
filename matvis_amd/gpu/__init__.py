"""GPU backend: ``simulate`` and the device implementations of the plugins."""

from .gpu import simulate, simulate_multi  # noqa: F401

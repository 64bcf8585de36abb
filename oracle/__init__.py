"""CPU oracle (test infrastructure only; never imported by ``matvis_amd``)."""

from .matvis_oracle import *  # noqa: F401,F403

"""MI355X (gfx950) lattice Boltzmann solver: HIP kernels behind a C ABI (liblbm_hip.so).

The directory name is not a Python identifier; load it with ``load_package()`` from
tests/conftest.py, bench.py or __graft_entry__.py (importlib, under the name
``cylinder_flow_amd``).
"""
from . import lbm, configs  # noqa: F401
from .lbm import Solver, LbmConfig, LbmError, make_config, default_config, lattice_table  # noqa: F401

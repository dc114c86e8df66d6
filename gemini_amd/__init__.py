"""MI355X-native GEMINI multifluid transport hot path: advection sweeps and TRBDF2 energy
diffusion, the rest of fluid_adv's transport block, the source/loss substep and the plasma
output and input files, behind the reference's C API (libgemhip.so; Python front end in
:mod:`gemini_amd.transport`)."""
from .grid import Grid, cartesian_grid, dipole_grid  # noqa: F401
from .state import State, Atmosphere, synthetic_state, synthetic_atmosphere, empty_state, cfl_dt  # noqa: F401

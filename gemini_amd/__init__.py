"""MI355X-native GEMINI multifluid transport hot path (advection sweeps + TRBDF2)."""
from .grid import Grid, cartesian_grid, dipole_grid  # noqa: F401
from .state import State, Atmosphere, synthetic_state, synthetic_atmosphere, empty_state, cfl_dt  # noqa: F401

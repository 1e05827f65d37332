"""MI355X-native (gfx950 HIP) lid-driven cavity solver: the Eulerian-impulse SSP-RK3 time step of the
reference backend.cu behind the same C-ABI (libcavity_solver.so) that its ctypes binding loads."""
from .python_api import Solver, Config, Report, Diagnostics, DT_REASONS, FIELD_NUMBER, load_library, library_path

__all__ = ['Solver', 'Config', 'Report', 'Diagnostics', 'DT_REASONS', 'FIELD_NUMBER', 'load_library',
           'library_path']

"""MI355X shallow-water dynamical core on the conformal cubed sphere.

Host-side mirror of the reference dynamic core's operator interface
(DwyaneChou/Conformal_Cubed_Sphere, dynamic_core/src): the mesh and metric
tables of ``read_mesh`` (io_mod.F90), the initial states of ``initTestCase``
(test_case_mod.F90), and the time-step procedures ``RK3_TVD`` / ``RK4``
(time_scheme_mod.F90), ``spatial_operator`` (spatial_operators_mod.F90) and
the diagnostics of diag_mod.F90 — all running on the GPU through libccs.so —
and a batched ensemble of independent runs (Ensemble).
"""
from .core import (  # noqa: F401
    CcsError,
    DynamicalCore,
    Ensemble,
    History,
    Mesh,
    State,
    cfl_time_step,
    error_norms,
    init_case,
    perturb_state,
    SPATIAL_OPS,
    read_mesh_file,
    read_mesh_file_ext,
    write_mesh_file,
)

__all__ = ["CcsError", "DynamicalCore", "Ensemble", "History", "Mesh", "SPATIAL_OPS", "State", "cfl_time_step",
           "error_norms", "init_case", "perturb_state", "read_mesh_file", "read_mesh_file_ext", "write_mesh_file"]

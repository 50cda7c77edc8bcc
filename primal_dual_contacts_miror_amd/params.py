"""ctypes mirror of ``pdc_params`` / ``pdc_solve_stats`` / ``pdc_timings`` (include/pdc.h).

The defaults are the reference's code defaults: ``simulation.timestep`` 1e-2 and
gravity (0, -9.81, 0) (src/Dynamics/DynamicSystem.cpp:386, DynamicSystem.h:241),
``solver.tolerance`` 1e-6 and ``solver.iterations`` 100
(src/Solver/PrimalDual/PrimalDualForceSolver.cpp:651-652), ``static_coeff`` 0.5 and
``restitution`` 0.0 (NonSmoothContactForce.cpp:329-333), ``spring_k`` 1e6 and
``spring_damp`` 0.01 (NonSmoothContactForce.cpp:324-327), ``maxCGIt`` 1000
(PrimalDualForceSolver.cpp:381).
"""
import ctypes

SHAPE_BOX = 0
SHAPE_SPHERE = 1
SHAPE_ELLIPSOID = 2
SHAPE_MESH = 3
NON_SMOOTH_CONTACT = 0
ONLY_NORMAL = 1
PCG_FP64 = 0
PCG_FP32 = 1


class PdcParams(ctypes.Structure):
    _fields_ = [
        ("timestep", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("force_model", ctypes.c_int32),
        ("spring_based_force", ctypes.c_int32),
        ("friction_coeff", ctypes.c_double),
        ("restitution", ctypes.c_double),
        ("spring_k", ctypes.c_double),
        ("spring_damp", ctypes.c_double),
        ("tolerance", ctypes.c_double),
        ("iterations", ctypes.c_int32),
        ("pcg_precision", ctypes.c_int32),
        ("pcg_tol", ctypes.c_double),
        ("pcg_max_iter", ctypes.c_int32),
        ("grid_n", ctypes.c_int32),
        ("grid_pad_frac", ctypes.c_double),
        ("sample_spacing", ctypes.c_double),
    ]


class PdcSolveStats(ctypes.Structure):
    _fields_ = [
        ("n_contacts", ctypes.c_int32),
        ("n_local_bodies", ctypes.c_int32),
        ("newton_iterations", ctypes.c_int32),
        ("pcg_iterations", ctypes.c_int32),
        ("pcg_maxiter_hits", ctypes.c_int32),
        ("gd_fallbacks", ctypes.c_int32),
        ("line_search_trials", ctypes.c_int32),
        ("ties", ctypes.c_int32),
        ("n_blocks", ctypes.c_int32),
        ("nonfinite", ctypes.c_int32),
        ("exit_error", ctypes.c_double),
        ("char_mass", ctypes.c_double),
        ("char_speed", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class PdcTimings(ctypes.Structure):
    _fields_ = [
        ("contact_detection", ctypes.c_double),
        ("force_process", ctypes.c_double),
        ("build_system", ctypes.c_double),
        ("solve_system", ctypes.c_double),
        ("line_search", ctypes.c_double),
        ("steps", ctypes.c_int64),
        ("newton_iterations", ctypes.c_int64),
        ("contacts", ctypes.c_int64),
    ]


class PdcDetectionStats(ctypes.Structure):
    """Counts of the last contact detection (pdc_get_detection_stats)."""
    _fields_ = [(name, ctypes.c_int64) for name in
                ("pairs", "raw_hits", "ties", "samples_all", "samples_read", "brick_nodes", "large_pairs",
                 "samples_in_grid")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class PdcShapes(ctypes.Structure):
    """Work shapes per scene size class (bodies < 64, < 256, < 1024, >= 1024)."""
    _fields_ = [
        ("contact_tile", ctypes.c_int32 * 4),
        ("body_waves", ctypes.c_int32 * 4),
        ("pcg_chunk", ctypes.c_int32),
    ]

    def as_dict(self):
        return {"contact_tile": list(self.contact_tile), "body_waves": list(self.body_waves),
                "pcg_chunk": int(self.pcg_chunk)}


def default_params(**overrides):
    """Reference code defaults, with keyword overrides (field names of pdc_params)."""
    p = PdcParams()
    p.timestep = 1e-2
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    p.force_model = NON_SMOOTH_CONTACT
    p.spring_based_force = 0
    p.friction_coeff = 0.5
    p.restitution = 0.0
    p.spring_k = 1e6
    p.spring_damp = 0.01
    p.tolerance = 1e-6
    p.iterations = 100
    p.pcg_precision = PCG_FP64
    p.pcg_tol = 1e-8
    p.pcg_max_iter = 1000
    p.grid_n = 64
    p.grid_pad_frac = 0.1
    p.sample_spacing = 0.1
    for k, v in overrides.items():
        if k == "gravity":
            for i in range(3):
                p.gravity[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def params_to_dict(p):
    d = {}
    for name, _ in PdcParams._fields_:
        v = getattr(p, name)
        d[name] = list(v) if name == "gravity" else v
    return d

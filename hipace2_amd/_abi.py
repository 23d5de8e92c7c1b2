"""ctypes binding of the C ABI in include/hpq.h (libhpq.so, built in-tree by build.py).

The library is the only compute path: if it is missing this module raises, there is
no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhpq.so")

HPQ_OK = 0
ERRORS = {1: "HPQ_E_INVALID", 2: "HPQ_E_HIP", 3: "HPQ_E_MG_NOT_CONVERGED", 4: "HPQ_E_MG_DIVERGED",
          5: "HPQ_E_UNSUPPORTED", 6: "HPQ_E_NOMEM"}

LAYOUT_NAMES = ["ncomp", "next_jx_beam", "next_jy_beam", "chi", "Sy", "Sx", "ExmBy", "EypBx", "Ez", "Bx",
                "By", "Bz", "Psi", "jx_beam", "jy_beam", "jz_beam", "jx", "jy", "rhomjz", "rho",
                "prev_jx_beam", "prev_jy_beam", "ions_rhomjz", "next_jx", "next_jy", "jz", "prev_Bx",
                "prev_By", "prev_jx", "prev_jy", "pciter_Bx", "pciter_By",
                "salame_Ez_target", "salame_Ez_no_salame", "salame_Ez", "salame_jx", "salame_jy",
                "salame_jz_beam", "salame_Bx", "salame_By", "salame_Sy", "salame_Sx", "salame_Sy_back",
                "salame_Sx_back"]

EXPORTS = ["hpq_create", "hpq_destroy", "hpq_last_error", "hpq_set_stream", "hpq_layout", "hpq_store_doubles",
           "hpq_deposit_current", "hpq_explicit_deposition", "hpq_deposit_current_slice",
           "hpq_advance_plasma_particles", "hpq_poisson_dirichlet", "hpq_solve_poisson_psi_exmby_eypbx_ez_bz",
           "hpq_solve_poisson_bxby", "hpq_initialize_sxsy_with_beam", "hpq_mg_solve1", "hpq_plasma_count",
           "hpq_init_plasma", "hpq_begin_step", "hpq_solve_one_slice", "hpq_run_slices", "hpq_set_checksums",
           "hpq_add_capture", "hpq_clear_captures", "hpq_stats", "hpq_trace_enable", "hpq_trace_read",
           "hpq_trace_reset", "hpq_bin_tiles", "hpq_deposit_current_vjp", "hpq_explicit_deposition_vjp",
           "hpq_advance_plasma_particles_vjp", "hpq_adjoint_forward", "hpq_adjoint_backward",
           "hpq_beam_begin_step", "hpq_advance_beam_slice", "hpq_sort_beam_slices", "hpq_run_step_with_beam",
           "hpq_set_insitu", "hpq_set_salame", "hpq_bin_tiles_noalloc"]

TRACE = {"push": 0, "deposit": 1, "explicit_deposit": 2, "beam_deposit": 3, "mg_smooth": 4, "mg_coarse": 5,
         "dst": 6}


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xlo", C.c_double), ("xhi", C.c_double), ("ylo", C.c_double), ("yhi", C.c_double),
                ("zlo", C.c_double), ("zhi", C.c_double),
                ("depos_order", C.c_int), ("derivative_type", C.c_int), ("explicit_solver", C.c_int),
                ("deposit_rho", C.c_int), ("do_beam_jx_jy", C.c_int), ("neutralize", C.c_int),
                ("mg_tol_rel", C.c_double), ("mg_tol_abs", C.c_double), ("mg_maxiter", C.c_int),
                ("particle_bc", C.c_int), ("plo", C.c_double * 2), ("phi", C.c_double * 2),
                ("pc_tol", C.c_double), ("pc_maxiter", C.c_int), ("pc_mix", C.c_double),
                ("plasma_charge", C.c_double), ("plasma_mass", C.c_double), ("max_qsa", C.c_double),
                ("bin_period", C.c_int), ("precision", C.c_int), ("poisson_solver", C.c_int),
                ("salame", C.c_int), ("deposition", C.c_int), ("tile_size", C.c_int),
                ("reorder_period", C.c_int), ("mixed_sums", C.c_int), ("mg_guess", C.c_int),
                ("mg_guess_iters", C.c_int)]

PRECISION = {"fp64": 0, "mixed": 1}
DEPOSITION = {"ordered": 0, "atomic": 1}
MIXED_SUMS = {"fp64": 0, "kahan32": 1}
MG_GUESS = {"previous": 0, "linear": 1, "dst": 2}
POISSON_SOLVER = {"fft": 0, "bf16x3": 1}


class Plasma(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("x", "y", "w", "ux", "uy", "psi", "x_prev", "y_prev", "ux_half_step",
                                           "uy_half_step", "psi_half_step", "valid")] + [("n", C.c_int64)]


ADJ_NAMES = ["x", "y", "ux", "uy", "psi", "ux_half_step", "uy_half_step", "psi_half_step"]


class BeamState(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("x", "y", "z", "ux", "uy", "uz", "w", "valid", "nsubcycles")] + \
               [("off", C.POINTER(C.c_int64)), ("n", C.c_int64), ("charge", C.c_double), ("mass", C.c_double),
                ("external_fields", C.c_double * 24), ("use_external_fields", C.c_int)]


class PlasmaAdj(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ADJ_NAMES]


class Beam(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("x", "y", "z", "ux", "uy", "uz", "w", "valid")] + \
               [("off", C.POINTER(C.c_int64)), ("n", C.c_int64), ("charge", C.c_double), ("mass", C.c_double)]


INSITU_FIELD_NR, INSITU_PLASMA_NR, INSITU_BEAM_NR = 10, 14, 22


class Insitu(C.Structure):
    _fields_ = [("field_rdata", C.c_void_p), ("field_sum", C.c_void_p),
                ("plasma_rdata", C.c_void_p), ("plasma_idata", C.c_void_p), ("plasma_sum", C.c_void_p),
                ("plasma_isum", C.c_void_p), ("plasma_radius", C.c_double),
                ("beam_rdata", C.c_void_p), ("beam_idata", C.c_void_p), ("beam_sum", C.c_void_p),
                ("beam_isum", C.c_void_p), ("beam_radius", C.c_double)]


SALAME_TARGET = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p)


class Salame(C.Structure):
    _fields_ = [("mask", C.c_void_p), ("w", C.c_void_p), ("valid", C.c_void_p), ("n_iter", C.c_int),
                ("do_advance", C.c_int), ("relative_tolerance", C.c_double), ("target", SALAME_TARGET),
                ("user", C.c_void_p), ("weight_factor", C.POINTER(C.c_double))]


class PlasmaProfile(C.Structure):
    _fields_ = [("ppc", C.c_int * 2), ("density", C.c_double), ("ramp_a", C.c_double), ("radius", C.c_double),
                ("u_mean", C.c_double * 3), ("density_table", C.c_void_p)]


class HpqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libhpq.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python hipace2_amd/build.py` "
                               "(or __graft_entry__.build()) before using hipace2_amd")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
        i32p = C.POINTER(C.c_int32)
        L.hpq_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
        L.hpq_destroy.argtypes = [vp]
        L.hpq_last_error.argtypes = [vp]
        L.hpq_last_error.restype = C.c_char_p
        L.hpq_set_stream.argtypes = [vp, vp]
        L.hpq_layout.argtypes = [vp, i32p]
        L.hpq_store_doubles.argtypes = [vp]
        L.hpq_store_doubles.restype = C.c_int64
        L.hpq_deposit_current.argtypes = [vp, C.POINTER(Plasma), vp, i32p, ip, C.POINTER(C.c_int64)]
        L.hpq_explicit_deposition.argtypes = [vp, C.POINTER(Plasma), vp]
        L.hpq_deposit_current_slice.argtypes = [vp, C.POINTER(Beam), ip, vp, i32p]
        L.hpq_advance_plasma_particles.argtypes = [vp, C.POINTER(Plasma), vp, ip]
        L.hpq_poisson_dirichlet.argtypes = [vp, vp, vp]
        L.hpq_solve_poisson_psi_exmby_eypbx_ez_bz.argtypes = [vp, vp]
        L.hpq_solve_poisson_bxby.argtypes = [vp, vp]
        L.hpq_initialize_sxsy_with_beam.argtypes = [vp, vp]
        L.hpq_mg_solve1.argtypes = [vp, vp, vp, vp, C.c_int64, C.c_double, C.c_double, ip, C.POINTER(ip)]
        L.hpq_plasma_count.argtypes = [vp, C.POINTER(PlasmaProfile)]
        L.hpq_plasma_count.restype = C.c_int64
        L.hpq_init_plasma.argtypes = [vp, C.POINTER(PlasmaProfile), C.POINTER(Plasma)]
        L.hpq_begin_step.argtypes = [vp, C.POINTER(Plasma), vp]
        L.hpq_solve_one_slice.argtypes = [vp, C.POINTER(Plasma), C.POINTER(Beam), vp, ip]
        L.hpq_run_slices.argtypes = [vp, C.POINTER(Plasma), C.POINTER(Beam), vp, ip, ip]
        L.hpq_set_checksums.argtypes = [vp, vp]
        L.hpq_add_capture.argtypes = [vp, ip, vp]
        L.hpq_set_insitu.argtypes = [vp, C.POINTER(Insitu)]
        L.hpq_set_salame.argtypes = [vp, C.POINTER(Salame)]
        L.hpq_bin_tiles.argtypes = [vp, C.POINTER(Plasma), ip, vp, vp]
        L.hpq_bin_tiles_noalloc.argtypes = [vp, C.POINTER(Plasma), ip, vp, vp]
        L.hpq_clear_captures.argtypes = [vp]
        L.hpq_stats.argtypes = [vp, dp]
        L.hpq_trace_enable.argtypes = [vp, ip, ip]
        L.hpq_trace_read.argtypes = [vp, ip, dp, C.POINTER(C.c_int64)]
        L.hpq_trace_reset.argtypes = [vp]
        L.hpq_deposit_current_vjp.argtypes = [vp, C.POINTER(Plasma), vp, i32p, C.POINTER(PlasmaAdj)]
        L.hpq_explicit_deposition_vjp.argtypes = [vp, C.POINTER(Plasma), vp, vp, C.POINTER(PlasmaAdj)]
        L.hpq_advance_plasma_particles_vjp.argtypes = [vp, C.POINTER(Plasma), vp, C.POINTER(PlasmaAdj),
                                                       C.POINTER(PlasmaAdj), vp]
        L.hpq_adjoint_forward.argtypes = [vp, C.POINTER(Plasma), C.POINTER(Beam), vp, ip]
        L.hpq_beam_begin_step.argtypes = [vp, C.POINTER(BeamState)]
        L.hpq_advance_beam_slice.argtypes = [vp, C.POINTER(BeamState), ip, vp, C.c_double, ip, ip]
        L.hpq_sort_beam_slices.argtypes = [vp, C.POINTER(BeamState), C.POINTER(C.c_int64)]
        L.hpq_run_step_with_beam.argtypes = [vp, C.POINTER(Plasma), C.POINTER(BeamState), vp, C.c_double, ip, ip]
        L.hpq_adjoint_backward.argtypes = [vp, C.POINTER(Plasma), C.POINTER(Beam), vp, vp, vp, vp, C.c_double]
        _lib = L
    return _lib


def check(ctx, rc):
    if rc != HPQ_OK:
        msg = lib().hpq_last_error(ctx)
        raise HpqError(rc, msg.decode() if msg else "")

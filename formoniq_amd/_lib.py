"""ctypes binding of libfeec.so (the C ABI declared in include/feec.h).

The library is loaded from this package directory only. There is no fallback:
if libfeec.so is missing the import fails, so nothing can silently run on a
CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfeec.so")

i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
i8p = C.POINTER(C.c_int8)
vp = C.c_void_p


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("max_iters", C.c_int64), ("precision", C.c_int32),
                ("max_refinements", C.c_int32), ("inner_rtol", C.c_double),
                ("use_graph", C.c_int32), ("graph_iters", C.c_int32),
                ("cheb_degree", C.c_int32), ("cheb_lower", C.c_double), ("precond_storage", C.c_int32),
                ("preconditioner", C.c_int32), ("mg_sweeps", C.c_int32)]


class Report(C.Structure):
    _fields_ = [("iters", C.c_int64), ("refinements", C.c_int32), ("converged", C.c_int32),
                ("residual", C.c_double), ("precond_residual", C.c_double),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Request(C.Structure):
    _fields_ = [("dim", C.c_int32), ("grade", C.c_int32), ("n", C.c_int32), ("bc", C.c_int32),
                ("jitter", C.c_double), ("mesh_seed", C.c_uint64), ("source_seed", C.c_uint64),
                ("precision", C.c_int32), ("reserved", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("ticket", C.c_int64), ("ndofs", C.c_int64), ("ncells", C.c_int64),
                ("iters", C.c_int64), ("refinements", C.c_int32), ("status", C.c_int32),
                ("residual", C.c_double), ("u_norm", C.c_double), ("u_checksum", C.c_double),
                ("admit_time", C.c_double), ("finish_time", C.c_double), ("latency", C.c_double),
                ("build_seconds", C.c_double), ("solve_seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_inflight", C.c_int32),
                ("reserved", C.c_int32), ("solve", SolveOpts)]


S = C.c_int  # feec_status

# name -> (restype, argtypes); every function of include/feec.h
SIGNATURES = {
    "feec_last_error": (C.c_char_p, []),
    "feec_status_string": (C.c_char_p, [C.c_int]),
    "feec_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "feec_ctx_create": (S, [C.c_int, C.POINTER(vp)]),
    "feec_ctx_destroy": (None, [vp]),
    "feec_ctx_synchronize": (S, [vp]),
    "feec_ctx_memory": (S, [vp, i64p]),
    "feec_ctx_allocations": (S, [vp, i64p]),
    "feec_kuhn_counts": (S, [C.c_int, C.c_int, i64p, i64p]),
    "feec_kuhn_grid": (S, [C.c_int, C.c_int, i64p, f64p]),
    "feec_jitter_interior": (S, [C.c_int, C.c_int, f64p, C.c_double, C.c_uint64, C.c_uint32]),
    "feec_normal_vector": (S, [C.c_int64, C.c_uint64, f64p]),
    "feec_mesh_create": (S, [vp, C.c_int, C.c_int64, C.c_int64, i64p, C.POINTER(vp)]),
    "feec_mesh_destroy": (None, [vp]),
    "feec_mesh_dim": (S, [vp, C.POINTER(C.c_int)]),
    "feec_mesh_counts": (S, [vp, i64p]),
    "feec_mesh_skeleton": (S, [vp, C.c_int, i64p]),
    "feec_mesh_cell_faces": (S, [vp, C.c_int, i64p]),
    "feec_mesh_boundary": (S, [vp, C.c_int, u8p]),
    "feec_mesh_orientation": (S, [vp, i8p]),
    "feec_mesh_coboundary": (S, [vp, C.c_int, C.c_int, C.POINTER(vp)]),
    "feec_geom_from_lengths": (S, [vp, f64p, C.POINTER(vp)]),
    "feec_geom_from_coords": (S, [vp, f64p, C.c_int, C.POINTER(vp)]),
    "feec_geom_destroy": (None, [vp]),
    "feec_geom_lengths": (S, [vp, f64p]),
    "feec_cell_metrics": (S, [vp, f64p, f64p, f64p]),
    "feec_causal_census": (S, [vp, i64p, u8p]),
    "feec_element_shape": (S, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "feec_element_matrices": (S, [vp, C.c_int, C.c_int, f64p]),
    "feec_assemble": (S, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    "feec_csr_shape": (S, [vp, i64p, i64p, i64p]),
    "feec_csr_copy": (S, [vp, i64p, i32p, f64p]),
    "feec_csr_spmv": (S, [vp, f64p, f64p]),
    "feec_csr_cg": (S, [vp, f64p, f64p, C.c_double, C.c_int64, C.POINTER(Report)]),
    "feec_csr_cg_noalloc": (S, [vp, f64p, f64p, C.c_double, C.c_int64, C.POINTER(Report)]),
    "feec_csr_destroy": (None, [vp]),
    "feec_harmonics": (S, [vp, C.c_int, C.c_int, i64p, f64p]),
    "feec_harmonics_integral": (S, [vp, C.c_int, C.c_int, i64p, f64p]),
    "feec_betti_number": (S, [vp, C.c_int, C.c_int, i64p]),
    "feec_cohomology_generators": (S, [vp, C.c_int, C.c_int, i64p, i64p]),
    "feec_homology_generators": (S, [vp, C.c_int, C.c_int, i64p, i64p]),
    "feec_mixed_kkt": (S, [vp, C.c_int, C.c_int, C.POINTER(vp), i64p]),
    "feec_solve_source": (S, [vp, C.c_int, C.c_int, f64p, C.POINTER(SolveOpts), f64p, f64p, f64p,
                              C.POINTER(Report)]),
    "feec_hodge_dirac_assemble": (S, [vp, C.c_int, C.c_double, C.POINTER(vp), C.POINTER(vp), i64p]),
    "feec_solve_dirac_source": (S, [vp, C.c_double, f64p, f64p, C.POINTER(SolveOpts), f64p,
                                    C.POINTER(Report)]),
    "feec_dirac_leapfrog": (S, [vp, C.c_int, C.c_double, C.c_int64, f64p, f64p, f64p,
                                C.POINTER(Report)]),
    "feec_dirac_irk": (S, [vp, C.c_int, C.c_int, C.c_double, C.c_int64, f64p, f64p, f64p,
                           C.POINTER(Report)]),
    "feec_problem_create": (S, [vp, C.c_int, C.c_int, C.POINTER(SolveOpts), C.POINTER(vp)]),
    "feec_problem_destroy": (None, [vp]),
    "feec_problem_assemble": (S, [vp]),
    "feec_problem_solve_device": (S, [vp, vp, C.POINTER(Report)]),
    "feec_problem_solve": (S, [vp, f64p, C.POINTER(Report)]),
    "feec_problem_solve_cochain_device": (S, [vp, vp, C.POINTER(Report)]),
    "feec_problem_solve_cochain": (S, [vp, f64p, C.POINTER(Report)]),
    "feec_csr_view": (S, [vp, i64p, i64p, i64p, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.c_int]),
    "feec_csr_spmv_device": (S, [vp, vp, vp]),
    "feec_vec_dot_device": (S, [vp, C.c_int64, vp, vp, f64p]),
    "feec_vec_axpy_device": (S, [vp, C.c_int64, C.c_double, vp, vp]),
    "feec_load_vector": (S, [vp, C.c_int, f64p, f64p]),
    "feec_quadrature_rule": (S, [C.c_int, C.c_int, C.POINTER(C.c_int), f64p, f64p]),
    "feec_cell_quadrature_points": (S, [vp, C.c_int, vp, C.c_int]),
    "feec_simplex_quadrature_points": (S, [vp, C.c_int, C.c_int, vp, C.c_int]),
    "feec_source_load": (S, [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int]),
    "feec_l2_projection": (S, [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.POINTER(Report), C.c_int]),
    "feec_whitney_at_nodes": (S, [vp, C.c_int, C.c_int, vp, vp, C.c_int]),
    "feec_l2_error": (S, [vp, C.c_int, vp, C.c_int, vp, f64p, C.c_int]),
    "feec_derham_map": (S, [vp, C.c_int, C.c_int, vp, vp, C.c_int]),
    "feec_gmsh_read": (S, [C.c_char_p, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int), i64p, i64p,
                           i64p, f64p]),
    "feec_cbor_read_cells": (S, [C.c_char_p, C.c_int64, C.POINTER(C.c_int), i64p, i64p]),
    "feec_cbor_read_lengths": (S, [C.c_char_p, C.c_int64, i64p, f64p]),
    "feec_cbor_write_cells": (S, [C.c_int, C.c_int64, i64p, C.c_char_p, i64p]),
    "feec_cbor_write_lengths": (S, [C.c_int64, f64p, C.c_char_p, i64p]),
    "feec_tune_mode": (S, [C.c_int]),
    "feec_tune_table": (S, [C.c_char_p, i64p]),
    "feec_ctx_stream_rates": (S, [vp, C.c_int64, f64p]),
    "feec_problem_solution": (S, [vp, f64p, f64p, f64p]),
    "feec_problem_sizes": (S, [vp, i64p, i64p]),
    "feec_problem_source_buffer": (S, [vp, C.POINTER(vp)]),
    "feec_problem_preconditioner": (S, [vp, i64p, C.POINTER(C.c_int32), f64p, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32), i64p, i64p]),
    "feec_problem_kernel_times": (S, [vp, C.c_int, f64p]),
    "feec_problem_mg_kernel_times": (S, [vp, C.c_int, C.c_int, f64p]),
    "feec_problem_apply_preconditioner": (S, [vp, C.c_int32, C.c_int32, f64p, f64p]),
    "feec_problem_upload_source": (S, [vp, f64p]),
    "feec_dirac_element_time": (S, [vp, C.c_int, f64p, i64p]),
    "feec_batch_run": (S, [vp, C.c_int, C.POINTER(Request), C.POINTER(SolveOpts), C.POINTER(Result)]),
    "feec_server_create": (S, [vp, C.POINTER(ServerOpts), C.POINTER(vp)]),
    "feec_server_submit": (S, [vp, C.POINTER(Request), i64p]),
    "feec_server_poll": (S, [vp, C.c_int64, C.POINTER(Result)]),
    "feec_server_solution": (S, [vp, C.c_int64, f64p, i64p]),
    "feec_server_drain": (S, [vp]),
    "feec_server_release": (S, [vp, C.c_int64]),
    "feec_server_held": (S, [vp, i64p, i64p]),
    "feec_server_destroy": (None, [vp]),
}

_LIB = None


def load():
    """Load libfeec.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libfeec.so is not built ({LIB_PATH}); run `python -m formoniq_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


STATUS = {0: "FEEC_OK", 1: "FEEC_E_ARG", 2: "FEEC_E_DEVICE", 3: "FEEC_E_OOM", 4: "FEEC_E_TOPOLOGY",
          5: "FEEC_E_DEGENERATE_CELL", 6: "FEEC_E_NOT_SPD", 7: "FEEC_E_NOT_CONVERGED",
          8: "FEEC_E_UNSUPPORTED", 9: "FEEC_E_ADMISSION", 10: "FEEC_E_AGAIN"}


class FeecError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status
        self.name = STATUS.get(status, str(status))


def check(status):
    if status != 0:
        msg = load().feec_last_error()
        raise FeecError(status, msg.decode() if msg else "")
    return status

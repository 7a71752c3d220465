"""ctypes binding of liblbm_hip.so (declarations: include/lbm_abi.h).

Loading fails loudly: there is no CPU fallback for any kernel.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from ._build import LIB_PATH, STAMP_PATH, source_hash

LBM_BGK, LBM_TRT, LBM_MRT = 0, 1, 2
STORE_F32, STORE_F16_SHIFT, STORE_BF16_SHIFT, STORE_F64 = 0, 1, 2, 3
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1
SC_NONE, SC_C1, SC_C2 = 0, 1, 2
TAU_CONST, TAU_FIELD, TAU_POWER_LAW = 0, 1, 2
NEEM_VELOCITY, NEEM_PRESSURE, RULE_VELOCITY_BB, RULE_OPEN_NEUMANN, RULE_PRESSURE_ABC, RULE_VELOCITY_EQ = \
    1, 2, 3, 4, 5, 6
BC_PERIODIC, BC_INSIDE, BC_BOUNCEBACK, BC_VELOCITY_NEEM, BC_PRESSURE_NEEM, BC_VELOCITY_BB, BC_OPEN_NEUMANN = \
    1, 2, 3, 4, 5, 6, 7
BC_PRESSURE_ABC, BC_VELOCITY_EQ = 8, 9
BC_OPEN_CONVECTIVE1, BC_OPEN_EXTRAPOLATION = 10, 11
RULE_OPEN_CONVECTIVE1 = 7

CODE_BOUNCE_MASK = 0x00FF
CODE_ACTIVE = 0x0100
CODE_KEEP_MACRO = 0x0200
CODE_NEEM_SHIFT = 10
CODE_NEEM_MASK = 0x3C00
CODE_PULLED = 0x4000      # inactive node pulled by an active one (plain pull)
CODE_SC_SOLID = 0x8000    # fluid_strategy false: a Shan-Chen neighbour giving adhesion
PSI_SC, PSI_VDW, PSI_RK, PSI_PR, PSI_CS = 0, 1, 2, 3, 4
PSI_BC_NEUMANN, PSI_BC_CONVECTIVE1, PSI_BC_CONVECTIVE2, PSI_BC_EXTRAPOLATION, PSI_BC_FIXED = 1, 2, 3, 4, 5
MAX_NEEM = 15
MAX_NC = 2
ABI_VERSION = 17


class NeemRule(C.Structure):
    _fields_ = [("kind", C.c_int32), ("direction", C.c_int32), ("ux", C.c_float), ("uy", C.c_float),
                ("rho", C.c_float), ("pad", C.c_int32)]


class Lattice(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nc", C.c_int32),
        ("collision", C.c_int32), ("storage", C.c_int32), ("tau_mode", C.c_int32),
        ("shan_chen", C.c_int32), ("n_neem", C.c_int32),
        ("plane_stride", C.c_int64),
        ("tau", C.c_float * MAX_NC), ("tau_anti", C.c_float * MAX_NC), ("magic", C.c_float * MAX_NC),
        ("s_shear", C.c_float * MAX_NC), ("s_bulk", C.c_float * MAX_NC),
        ("visc_m", C.c_float), ("visc_nm1", C.c_float),
        ("tau_min", C.c_float), ("tau_max", C.c_float),
        ("gx", C.c_float), ("gy", C.c_float),
        ("g_coh", C.c_float * (MAX_NC * MAX_NC)),
        ("psi_rho0", C.c_float * MAX_NC),
        ("psi_kind", C.c_int32 * MAX_NC),
        ("eos", (C.c_float * 6) * MAX_NC),
        ("g_adh", C.c_float * MAX_NC),
        ("sc_adhesion", C.c_int32), ("huang", C.c_int32),
        ("huang_k1", C.c_float * MAX_NC), ("huang_k2", C.c_float * MAX_NC), ("huang_g", C.c_float * MAX_NC),
        ("round_step", C.c_uint32),
        ("rounding", C.c_int32),
        ("pad1", C.c_int32),
    ]


class Fields(C.Structure):
    _fields_ = [
        ("f", C.c_void_p * 2), ("macro", C.c_void_p * 2), ("tau", C.c_void_p),
        ("psi", C.c_void_p * 2), ("body_force", C.c_void_p), ("sc_force", C.c_void_p),
        ("shear_rate", C.c_void_p), ("code", C.c_void_p), ("tau2", C.c_void_p), ("temperature", C.c_void_p),
    ]


class GFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "f", "f2", "rho", "total_rho", "vel", "body_force", "sc_force", "shear_rate", "pressure",
        "total_pressure", "temperature", "psi", "neighbor", "tau", "sc_fluid")]


class IBParticles(C.Structure):
    _fields_ = [("max_number", C.c_int32), ("max_num_ib", C.c_int32), ("num", C.c_int32),
                ("numib_cof", C.c_float)] + [(n, C.c_void_p) for n in (
                    "pos", "vel", "force", "torque", "radius", "angle", "num_ib", "node_pos", "node_vel",
                    "node_vel_desired", "node_force", "node_force_temp", "group", "count", "work")]


class LBMError(RuntimeError):
    """A failure reported by the HIP library (message from lbm_last_error)."""


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("LBM_HIP_LIB", str(LIB_PATH)))
    if not path.exists():
        raise RuntimeError(
            f"HIP library {path} is missing: build it with taichi_lbm._build.build() "
            "(or __graft_entry__.build()); there is no CPU fallback")
    if "LBM_HIP_LIB" not in os.environ:
        want = source_hash()
        have = STAMP_PATH.read_text().strip() if STAMP_PATH.exists() else ""
        if have != want:
            raise RuntimeError(f"HIP library {path} was built from other sources than the ones in this tree: "
                               "rebuild it with taichi_lbm._build.build() (or __graft_entry__.build())")
    L = C.CDLL(str(path))
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    LP, FP, NP, GP = C.POINTER(Lattice), C.POINTER(Fields), C.POINTER(NeemRule), C.POINTER(GFields)
    sig = {
        "lbm_abi_version": ([], C.c_int),
        "lbm_last_error": ([], C.c_char_p),
        "lbm_run": ([LP, FP, NP, I32, I32, I32, P], C.c_int),
        "lbm_run_blocked": ([LP, FP, NP, I32, I32, I32, I32, P, C.POINTER(C.c_int32), P], C.c_int),
        "lbm_run_blocked_noalloc": ([LP, FP, NP, I32, I32, I32, I32, C.POINTER(C.c_int32), P], C.c_int),
        "lbm_tile_rows": ([I32, I32], C.c_int),
        "lbm_finish": ([LP, FP, NP, I32, P], C.c_int),
        "lbm_reduce_macro": ([LP, P, P, D, P], C.c_int),
        "lbm_residual": ([I64, I64, P, P, D, P], C.c_int),
        "lbm_copy_f32": ([P, P, I64, P], C.c_int),
        "lbm_g_boundary": ([LP, GP, I32, NP, P, I32, P], C.c_int),
        "lbm_g_sc_periodic_psi": ([LP, GP, I32, P], C.c_int),
        "lbm_g_sc_psi_boundary": ([LP, GP, I32, I32, C.c_float, P, I32, P], C.c_int),
    }
    for name in ("lbm_g_init_lbm", "lbm_g_density", "lbm_g_pressure", "lbm_g_force_density", "lbm_g_velocity",
                 "lbm_g_shear_rate", "lbm_g_relaxation_time", "lbm_g_collide", "lbm_g_sc_update_psi",
                 "lbm_g_sc_force"):
        sig[name] = ([LP, GP, P, I32, P], C.c_int)
    BP = C.c_void_p
    sig.update({
        "lbm_batch_create": ([LP, FP, NP, I32, C.POINTER(C.c_void_p)], C.c_int),
        "lbm_batch_update": ([BP, I32, LP, FP, NP], C.c_int),
        "lbm_batch_set": ([BP, I32, I32, I32, P], C.c_int),
        "lbm_batch_set_round_step": ([BP, I32, C.c_uint32, P], C.c_int),
        "lbm_batch_commit": ([BP, P], C.c_int),
        "lbm_batch_tiles": ([BP], C.c_int),
        "lbm_batch_capacity": ([BP], C.c_int),
        "lbm_batch_reserve": ([BP, I32], C.c_int),
        "lbm_batch_run": ([BP, I32, I32, I32, P], C.c_int),
        "lbm_batch_finish": ([BP, I32, I32, P], C.c_int),
        "lbm_batch_graph": ([BP, I32, I32, I32, P, C.POINTER(C.c_void_p)], C.c_int),
        "lbm_batch_mode": ([BP], C.c_int),
        "lbm_batch_reserve2": ([BP, I32], C.c_int),
        "lbm_batch_tile_cols": ([BP, I32], C.c_int),
        "lbm_batch_ring": ([BP, I32], C.c_int),
        "lbm_batch_tiles2": ([BP], C.c_int),
        "lbm_batch_release": ([BP], C.c_int),
        "lbm_batch_run_blocked": ([BP, I32, I32, I32, C.POINTER(C.c_int32), P], C.c_int),
        "lbm_batch_graph_blocked": ([BP, I32, I32, I32, P, C.POINTER(C.c_void_p), C.POINTER(C.c_int32)], C.c_int),
        "lbm_graph_launch": ([P, P], C.c_int),
        "lbm_graph_destroy": ([P], C.c_int),
        "lbm_batch_residual": ([BP, I32, D, P], C.c_int),
        "lbm_batch_destroy": ([BP], C.c_int),
    })
    sig.update({
        "lbm_sc_batch_create": ([LP, FP, NP, C.POINTER(C.c_int32), I32, C.POINTER(C.c_void_p)], C.c_int),
        "lbm_sc_batch_run": ([P, I32, I32, I32, C.POINTER(C.c_int32), P], C.c_int),
        "lbm_sc_batch_tiles": ([P], C.c_int),
        "lbm_sc_batch_destroy": ([P], C.c_int),
    })
    IP = C.POINTER(IBParticles)
    sig.update({
        "lbm_ib_work_ints": ([I32, I32], C.c_int64),
        "lbm_ib_init_nodes": ([IP, P], C.c_int),
        "lbm_ib_update": ([LP, IP, P], C.c_int),
        "lbm_ib_reset_force": ([IP, P], C.c_int),
        "lbm_ib_interpolate": ([LP, GP, IP, P], C.c_int),
        "lbm_ib_spread": ([LP, GP, IP, P], C.c_int),
        "lbm_ib_mdf": ([LP, GP, IP, I32, P], C.c_int),
        "lbm_ib_particle_force": ([IP, P], C.c_int),
        "lbm_ib_run": ([LP, FP, IP, GP, I32, I32, I32, I32, C.POINTER(C.c_int32), P], C.c_int),
    })
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if L.lbm_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{path}: ABI version {L.lbm_abi_version()} != {ABI_VERSION}")
    _lib = L
    return L


def exported_symbols():
    """Names of every entry point the header declares (checked by the tests)."""
    return ["lbm_abi_version", "lbm_last_error", "lbm_run", "lbm_run_blocked", "lbm_run_blocked_noalloc",
            "lbm_finish", "lbm_reduce_macro",
            "lbm_residual", "lbm_copy_f32", "lbm_g_init_lbm", "lbm_g_density", "lbm_g_pressure", "lbm_g_force_density",
            "lbm_g_velocity", "lbm_g_shear_rate", "lbm_g_relaxation_time", "lbm_g_collide", "lbm_g_boundary",
            "lbm_g_sc_update_psi", "lbm_g_sc_periodic_psi", "lbm_g_sc_force", "lbm_g_sc_psi_boundary",
            "lbm_batch_create",
            "lbm_batch_update", "lbm_batch_set", "lbm_batch_set_round_step", "lbm_batch_commit", "lbm_batch_tiles", "lbm_batch_capacity",
            "lbm_batch_reserve", "lbm_batch_reserve2", "lbm_batch_run", "lbm_batch_finish", "lbm_batch_graph",
            "lbm_batch_mode", "lbm_batch_tile_cols", "lbm_tile_rows", "lbm_batch_ring", "lbm_batch_tiles2", "lbm_batch_release",
            "lbm_batch_run_blocked", "lbm_batch_graph_blocked", "lbm_graph_launch",
            "lbm_graph_destroy", "lbm_batch_residual", "lbm_batch_destroy",
            "lbm_sc_batch_create", "lbm_sc_batch_run", "lbm_sc_batch_tiles", "lbm_sc_batch_destroy",
            "lbm_ib_work_ints", "lbm_ib_init_nodes", "lbm_ib_update", "lbm_ib_reset_force", "lbm_ib_interpolate",
            "lbm_ib_spread", "lbm_ib_mdf", "lbm_ib_particle_force", "lbm_ib_run"]


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().lbm_last_error().decode(errors="replace")
        raise LBMError(f"{what}: {msg} (status {rc})" if what else f"{msg} (status {rc})")


def ptr(t) -> int:
    """Device address of a torch tensor (None -> NULL)."""
    return 0 if t is None else int(t.data_ptr())

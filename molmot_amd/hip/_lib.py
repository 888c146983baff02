"""ctypes binding of libmolmot_hip.so (the C ABI in include/molmot_hip.h).

torch is imported before the library is loaded: torch ships its own HIP
runtime (libamdhip64.so.7), and loading it first makes the dynamic linker
bind this library to that same runtime, so device pointers and streams
from torch are valid here.  There is no CPU fallback: if the library is
missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libmolmot_hip.so")

_lib = None

c_int, c_long, c_double, c_uint, c_void_p = (ctypes.c_int, ctypes.c_long, ctypes.c_double,
                                             ctypes.c_uint, ctypes.c_void_p)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class MCWFConfig(ctypes.Structure):
    """mm_mcwf_config (include/molmot_hip.h)."""
    _fields_ = [("dt", c_double), ("r_escape", c_double), ("n_steps", c_int), ("save_every", c_int),
                ("n_save", c_int), ("renorm_interval", c_int), ("avg_start", c_int), ("precision", c_int),
                ("jump_cap", c_int), ("jump_rule", c_int), ("seed", ctypes.c_ulonglong)]


class MCWFOutputs(ctypes.Structure):
    """mm_mcwf_outputs (include/molmot_hip.h): device pointers, NULL = not wanted."""
    _fields_ = [(n, c_void_p) for n in ("r", "v", "pop", "force", "force_sum", "psi_final", "r_final",
                                        "v_final", "jump_t", "n_scatters", "escaped", "n_saved", "n_avg",
                                        "jump_ch")]


class MolmotHIPError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status."""


# name -> (restype, argtypes); pointers to device memory are passed as c_void_p
_SIGNATURES = {
    "mm_last_error": (ctypes.c_char_p, []),
    "mm_version": (c_int, []),
    "mm_rate_plan_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int,
                                    _dp, _dp, _dp, _dp, _dp, _ip, _dp, _ip, _dp, c_int,
                                    c_double, c_double, c_double, c_double]),
    "mm_rate_plan_destroy": (c_int, [c_void_p]),
    "mm_rate_plan_n_terms": (c_int, [c_void_p]),
    "mm_rate_eq_map": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_double,
                               c_void_p, c_void_p, c_void_p, c_void_p]),
    "mm_rate_eq_points": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "mm_rate_eq_grad_points": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double,
                                       c_void_p, c_void_p, c_void_p]),
    "mm_traj_euler_1d": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double,
                                 c_double, c_int, c_double, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_double, c_uint, c_uint, c_uint, c_void_p]),
    "mm_traj_euler_workspace_bytes": (c_long, [c_long, c_int]),
    "mm_traj_euler_schedule": (c_int, [c_void_p, c_long, c_int, c_int, c_double]),
    "mm_traj_euler_1d_ws": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double,
                                    c_double, c_int, c_double, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_double, c_uint, c_uint, c_uint, c_int, c_void_p,
                                    c_long, c_void_p]),
    "mm_traj_euler_rf_1d": (c_int, [c_void_p, c_void_p, c_long, c_void_p, c_void_p, c_double, c_double,
                                    c_double, c_int, c_double, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_double, c_uint, c_uint, c_uint, c_void_p]),
    "mm_traj_axis3d": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double,
                               c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mm_traj_axis3d_workspace_bytes": (c_long, [c_long, c_int]),
    "mm_traj_axis3d_ws": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double, c_int,
                                  c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_long,
                                  c_void_p]),
    "mm_traj_vsq_1d": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double,
                               c_int, c_int, c_double, c_void_p, c_void_p, c_void_p]),
    "mm_traj_vsq_workspace_bytes": (c_long, [c_long, c_int]),
    "mm_traj_vsq_1d_ws": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double, c_int,
                                  c_int, c_double, c_void_p, c_void_p, c_int, c_void_p, c_long, c_void_p]),
    "mm_traj_axis3d_rf": (c_int, [c_void_p, c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double,
                                  c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mm_traj_vsq_rf_1d": (c_int, [c_void_p, c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double,
                                  c_int, c_int, c_double, c_void_p, c_void_p, c_void_p]),
    "mm_lindblad_class_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_int, _dp, _dp, _dp, _dp, _dp,
                                         _dp, _dp, c_double, c_double]),
    "mm_lindblad_class_destroy": (c_int, [c_void_p]),
    "mm_lindblad_rk4": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_int, c_int, c_int,
                                c_void_p, c_void_p, c_void_p, c_void_p]),
    "mm_rate_plan_set_block": (c_int, [c_void_p, c_int]),
    "mm_rate_plan_get_block": (c_int, [c_void_p]),
    "mm_lindblad_class_set_lanes": (c_int, [c_void_p, c_int, c_int]),
    "mm_lindblad_class_get_lanes": (c_int, [c_void_p, c_int]),
    "mm_lindblad_rk4_uneven": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_long, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_double, c_int, c_int, c_int, c_void_p]),
    "mm_lindblad_liouvillian": (c_int, [c_void_p, c_long, c_void_p, c_int, c_void_p, c_void_p]),
    "mm_mcwf_problem_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, _dp, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _dp, _dp, c_double, c_double, c_double, c_double, c_double,
                                       c_int, _dp]),
    "mm_mcwf_problem_destroy": (c_int, [c_void_p]),
    "mm_mcwf_field_table": (c_int, [c_void_p, c_long, c_void_p, c_int, c_void_p, c_void_p]),
    "mm_mcwf_run": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_void_p, c_void_p, c_long,
                            ctypes.POINTER(MCWFConfig), ctypes.POINTER(MCWFOutputs), c_void_p]),
    "mm_force3d_plan_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, _dp, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _ip, _dp, _dp, c_double, c_double, c_double, c_double,
                                       c_double]),
    "mm_force3d_plan_destroy": (c_int, [c_void_p]),
    "mm_force3d_plan_set_spin_basis": (c_int, [c_void_p, _dp, _dp, _dp]),
    "mm_force3d_leapfrog_workspace_bytes": (c_long, [c_long]),
    "mm_force3d_leapfrog_ws": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_int, c_double,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_long,
                                       c_void_p]),
    "mm_force3d_points": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p,
                                  c_void_p]),
    "mm_force3d_leapfrog": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_int, c_double,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mm_traj_vjp_workspace_doubles": (c_long, [c_long, c_int, c_int]),
    "mm_traj_vjp_1d": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double, c_int,
                               c_void_p, c_double, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_long, c_void_p]),
    "mm_traj_tangent_1d": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double, c_int,
                                   c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mm_traj_tangent_1d_ws": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double, c_int,
                                      c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                      c_long, c_void_p]),
    "mm_philox4x32_10": (c_int, [c_long, c_void_p, c_uint, c_uint, c_void_p, c_void_p]),
    "mm_traj_queue_set_wait": (c_int, [c_double]),
    "mm_zgesv_batched": (c_int, [c_int, c_int, c_long, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mm_traj_stats_workspace_bytes": (c_long, [c_long, c_int]),
    "mm_traj_stats_1d": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double, c_int,
                                 c_double, c_double, c_void_p, c_int, c_void_p, c_double, c_uint, c_uint, c_uint,
                                 c_void_p, c_long, c_void_p]),
    "mm_rate_plan_specialize": (c_int, [c_void_p, c_int]),
    "mm_rate_plan_is_specialized": (c_int, [c_void_p]),
    "mm_rate_plan_spec_source": (c_long, [c_void_p, c_int, ctypes.c_char_p, c_long]),
    "mm_rate_plan_specialize_rf": (c_int, [c_void_p, c_void_p]),
    "mm_rate_plan_is_specialized_rf": (c_int, [c_void_p, c_void_p]),
    "mm_rate_plan_spec_source_rf": (c_long, [c_void_p, c_void_p, ctypes.c_char_p, c_long]),
    "mm_rate_spec_source_tables": (c_long, [c_int, c_int, c_int, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _ip, _dp,
                                            c_int, c_double, c_double, c_double, c_double, c_int, ctypes.c_char_p,
                                            c_long]),
    "mm_rtc_compile": (c_long, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, c_long]),
    "mm_traj_euler_noalloc_max_bytes": (c_long, []),
    "mm_traj_euler_1d_noalloc": (c_int, [c_void_p, c_long, c_void_p, c_void_p, c_double, c_double, c_double, c_int,
                                         c_double, c_void_p, c_void_p, c_void_p, c_double, c_uint, c_uint, c_uint,
                                         c_int, c_void_p]),
}

# Symbols the header declares; tests check the library exports all of them.
DECLARED_SYMBOLS = tuple(_SIGNATURES)


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    """Load (once) and return the library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MolmotHIPError(
                f"{LIB_PATH} is missing: build the gfx950 kernels first "
                "(python -m molmot_amd.csrc.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().mm_last_error().decode(errors="replace")
        raise MolmotHIPError(f"{what or 'molmot_hip'} failed ({rc}): {msg}")


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise MolmotHIPError("no HIP device visible: the molmot_amd kernels run on MI355X only")


def stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def queue_workspace(nbytes: int, device, stream=None):
    """A work-queue workspace of nbytes (uint8 device tensor, None for 0)
    whose 256-byte control block is zeroed on ``stream`` by a fill kernel:
    the launch clears it again when it runs the queue, and when the auto
    rule picks one thread per particle instead the give-up flag read by
    check_queue stays 0."""
    if nbytes <= 0:
        return None
    ws = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
    s = stream if stream is not None else torch.cuda.current_stream(device)
    with torch.cuda.stream(s):
        ws[:min(256, int(nbytes))].zero_()
    return ws


def queue_timed_out(ws, stream=None) -> bool:
    """True if the work-queue launch that used workspace ``ws`` (a uint8
    device tensor) on ``stream`` (default: the current stream) gave up a
    wait: its 32-bit flag at byte offset 8.  Reads the flag back on that
    stream, so it waits for the launch; False without a workspace or while
    the stream is being captured into a graph (nothing has run yet)."""
    if ws is None or ws.numel() < 12:
        return False
    s = stream if stream is not None else torch.cuda.current_stream(ws.device)
    with torch.cuda.stream(s):
        if torch.cuda.is_current_stream_capturing():
            return False
        return int(ws[8:12].view(torch.int32).item()) != 0


def check_queue(ws, what: str, stream=None) -> None:
    """Raise MolmotHIPError if the queue launch on ``ws`` gave up a wait (its
    counts are then -1 and its state outputs incomplete)."""
    if queue_timed_out(ws, stream):
        raise MolmotHIPError(f"{what}: a work-queue wait timed out; the outputs are incomplete "
                             "(counts set to -1)")


def set_queue_wait(seconds: float | None) -> None:
    """Bound of every later work-queue wait in seconds (None: the default,
    max(10 s, 10 ms x steps per chunk))."""
    check(lib().mm_traj_queue_set_wait(-1.0 if seconds is None else float(seconds)), "mm_traj_queue_set_wait")


def ptr(t) -> int | None:
    return None if t is None else int(t.data_ptr())

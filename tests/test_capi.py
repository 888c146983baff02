"""C ABI surface (CPU only: nothing here launches a kernel)."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT

HEADER_DIR = os.path.join(ROOT, "include")


def declared_in_headers():
    names = set()
    for f in os.listdir(HEADER_DIR):
        if f.endswith(".h"):
            txt = open(os.path.join(HEADER_DIR, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", txt))
    return names


def test_binding_covers_header():
    from molmot_amd.hip import _lib
    assert declared_in_headers() == set(_lib.DECLARED_SYMBOLS)


def _load():
    from molmot_amd.hip import _lib
    if not _lib.available():
        pytest.skip("libmolmot_hip.so not built")
    return _lib.lib()


def test_library_exports_every_declared_symbol():
    L = _load()
    for name in declared_in_headers():
        assert hasattr(L, name), name
    assert L.mm_version() >= 100


def test_invalid_arguments_rejected_before_any_device_call():
    from molmot_amd.hip import _lib
    L = _load()
    h = ctypes.c_void_p()
    z = np.zeros(64)
    dp = z.ctypes.data_as(_lib._dp)
    ip = np.zeros(8, dtype=np.int32).ctypes.data_as(_lib._ip)
    rc = L.mm_rate_plan_create(ctypes.byref(h), 0, 4, 8, dp, dp, dp, dp, dp, ip, dp, None, None, 0,
                               1.0, 1.0, 1.0, 1.0)
    assert rc == -1 and b"n_ground" in L.mm_last_error()
    rc = L.mm_rate_plan_create(ctypes.byref(h), 12, 4, 8, dp, dp, dp, dp, dp, ip, dp, None, None, 0,
                               -1.0, 1.0, 1.0, 1.0)
    assert rc == -1 and b"Gamma" in L.mm_last_error()
    assert L.mm_rate_eq_map(None, None, 1, None, 1, 0.0, None, None, None, None) == -1
    assert L.mm_traj_euler_1d(None, 1, None, None, 0, 1, 1e-6, 10, 0.015, None, None, None, None, None,
                              0.0, 0, 0, 0, None) == -1
    assert L.mm_philox4x32_10(-1, None, 0, 0, None, None) == -1
    assert L.mm_rate_plan_destroy(None) == 0
    # MCWF: size and argument checks precede any allocation
    rc = L.mm_mcwf_problem_create(ctypes.byref(h), 5, 3, 2, dp, dp, dp, dp, dp, dp, dp, dp, dp, 0.1, 1.0, 1.0,
                                  1.0, 1e-3, 1, dp)
    assert rc == -2 and b"5 ground" in L.mm_last_error()
    rc = L.mm_mcwf_problem_create(ctypes.byref(h), 12, 4, 9, dp, dp, dp, dp, dp, dp, dp, dp, dp, 0.1, 1.0, 1.0,
                                  1.0, 1e-3, 1, dp)
    assert rc == -1 and b"frequencies" in L.mm_last_error()
    cfg = _lib.MCWFConfig(dt=0.01, n_steps=10, save_every=0, n_save=1)
    out = _lib.MCWFOutputs()
    assert L.mm_mcwf_run(None, 1, None, None, None, None, 0, ctypes.byref(cfg), ctypes.byref(out), None) == -1
    assert L.mm_mcwf_field_table(None, 4, None, 0, None, None) == -1
    assert L.mm_mcwf_problem_destroy(None) == 0
    # entry points without a reference counterpart: null handles and bad sizes
    assert L.mm_traj_tangent_1d(None, 1, None, None, 0.1, 1.0, 1e-6, 10, None, 0.015, None, None, None, None,
                                None) == -1
    assert L.mm_traj_vjp_workspace_doubles(-1, 10, 5) == -1
    assert L.mm_traj_euler_rf_1d(None, None, 1, None, None, 0.1, 1.0, 1e-6, 10, 0.015, None, None, None, None,
                                 None, 0.0, 0, 0, 0, None) == -1
    assert L.mm_traj_axis3d_rf(None, None, 1, None, None, 0.1, 1.0, 1e-6, 10, 0.015, None, None, None, None,
                               None, None) == -1
    assert L.mm_traj_vsq_rf_1d(None, None, 1, None, None, 0.1, 1.0, 1e-6, 10, 0, 0.015, None, None, None) == -1
    assert L.mm_rate_plan_set_block(None, 64) == -1
    assert L.mm_lindblad_class_set_lanes(None, 0, 64) == -1
    assert L.mm_lindblad_rk4_uneven(None, 1, None, None, None, None, None, None, 1, None, None, None, None, None,
                                    0.05, 10, 5, 0, None) == -1
    assert b"null class" in L.mm_last_error()
    assert L.mm_traj_euler_1d_ws(None, 1, None, None, 0.1, 1.0, 1e-6, 1000, 0.015, None, None, None, None, None,
                                 0.0, 0, 0, 0, 0, None, 0, None) == -1
    assert b"null plan" in L.mm_last_error()
    assert L.mm_traj_euler_schedule(None, 100, 1000, 0, 0.0) == -1


def test_traj_queue_workspace_size():
    """Workspace of the queued trajectory schedule: 256 bytes of counters
    plus 8 per queue item, groups x (chunks - 1), with groups = ceil(n/64),
    chunks of ceil((n_steps-1)/256) steps rounded up to even and at least 128."""
    L = _load()

    def expect(n, n_steps):
        steps = n_steps - 1
        if steps < 256:
            return 0
        chunk = max(-(-steps // 256), 128)
        chunk += chunk & 1
        chunks = -(-steps // chunk)
        return 256 + 8 * (-(-n // 64)) * (chunks - 1)

    # the 1e5 x 30,000 ensemble: 1,563 groups, 235 chunks of 128 steps
    assert L.mm_traj_euler_workspace_bytes(100_000, 30_000) == 256 + 8 * 1563 * 234
    for n, ns in ((1, 257), (64, 257), (65, 1000), (1000, 1500), (100_000, 600), (262_144, 30_000),
                  (10 ** 7, 100_000), (5, 256), (5, 1), (0, 5000)):
        assert L.mm_traj_euler_workspace_bytes(n, ns) == (expect(n, ns) if n > 0 else 0), (n, ns)


def test_philox_restatement_matches_kat(golden):
    """Pure-Python Philox4x32-10 vs rocRAND-generated vectors and the
    published Random123 known answers (kat_vectors: philox4x32 10 rounds)."""
    from oracle.philox import philox4x32_10
    d = golden("philox4x32_10_kat.npz")
    for c, k, o in zip(d["counter"], d["key"], d["out"]):
        assert philox4x32_10(c, k) == tuple(int(x) for x in o)
    assert philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert philox4x32_10((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert philox4x32_10((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == \
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_queue_wait_bound_validation():
    """mm_traj_queue_set_wait: a number of seconds up to one day; < 0 restores
    the default (no device call)."""
    L = _load()
    assert L.mm_traj_queue_set_wait(float("nan")) == -1 and b"seconds" in L.mm_last_error()
    assert L.mm_traj_queue_set_wait(1e6) == -1
    assert L.mm_traj_queue_set_wait(5.0) == 0
    assert L.mm_traj_queue_set_wait(-1.0) == 0


def test_force3d_spin_and_segment_entry_points_validate():
    """mm_force3d_plan_set_spin_basis / mm_force3d_leapfrog_ws reject a null
    plan and bad sizes before any device call; the workspace size is two
    256-byte-padded index arrays after a 256-byte header."""
    L = _load()
    assert L.mm_force3d_plan_set_spin_basis(None, None, None, None) == -1
    assert b"null plan" in L.mm_last_error()
    assert L.mm_force3d_leapfrog_ws(None, 10, None, None, 0.1, 1e-5, 10, 0.015, None, None, None, None, None,
                                    0, None, 0, None) == -1
    assert L.mm_force3d_leapfrog_workspace_bytes(0) == 256
    assert L.mm_force3d_leapfrog_workspace_bytes(64) == 256 + 2 * 256
    assert L.mm_force3d_leapfrog_workspace_bytes(100_000) == 256 + 2 * 400_128
    assert L.mm_force3d_leapfrog_workspace_bytes(-1) == 0

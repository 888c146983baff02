"""FP64 CPU oracle for molmot's rate-equation path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke`` and the CPU
baseline leg of bench.py may import this package; it is the checker, never
the thing measured as the product and never a fallback for it.

``liboracle_rate_eq.so`` (built from rate_eq_oracle.c by the Makefile here)
restates, with hbar as an argument:

* ``rate_eq(..., flavour="jit")``   -- molmot/obe/rate_equations_jit.py:18-145
* ``rate_eq(..., flavour="numpy")`` -- molmot/obe/rate_equations.py:23-191
* ``traj_euler``                    -- molmot/propagation/trajectories_jit.py:18-58

Pinning (tests/test_oracle_golden.py): the NumPy flavour reproduces
validation/three_way_results.npz ``Fz_cpu_s``/``Fv_cpu_s``/``Pz_cpu_s``/
``k_cpu_s``/``beta_cpu_s``; the JIT flavour with the legacy hbar reproduces the
CUDA columns ``Fz_gpu``/``Pz_gpu``/``k_gpu``/``beta_gpu``; the unsaturated
mode reproduces the Julia fixture julia_redmot_results.npz; ``traj_euler``
reproduces the 4 reference trajectories and the 100-particle batch of
validation/trajectory_cpu_vs_gpu_results.npz.  Closed forms (two-level
system), CTMC invariants (M p = 0, sum p = 1) and an mpmath solve pin what the
fixtures leave open.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_rate_eq.so")
_LIB_PATH_V4 = os.path.join(_HERE, "liboracle_rate_eq_v4.so")
_lib = None
_variant = "x86-64-v3"


def host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in line + " "
    except OSError:
        pass
    return False


def use_avx512_build() -> str:
    """Load the x86-64-v4 (AVX-512) build instead of the x86-64-v3 one if
    the host supports it and nothing has been loaded yet; returns the ISA
    level in use.  Same results, used by the CPU baseline."""
    global _LIB_PATH, _variant
    if _lib is None and host_has_avx512() and os.path.exists(_LIB_PATH_V4):
        _LIB_PATH, _variant = _LIB_PATH_V4, "x86-64-v4"
    return _variant


def variant() -> str:
    return _variant

FLAVOURS = {"jit": 0, "numpy": 1}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle library (gcc, see Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        common = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _dp]
        L.oracle_rate_eq.argtypes = common + [ctypes.c_double] * 7 + [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_rate_eq.restype = ctypes.c_int
        L.oracle_rate_eq_batch.argtypes = common + [ctypes.c_long, _dp, _dp] + [ctypes.c_double] * 5 + \
            [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
        L.oracle_rate_eq_batch.restype = ctypes.c_int
        L.oracle_traj_euler.argtypes = common + [ctypes.c_long, _dp, _dp, ctypes.c_double, ctypes.c_int] + \
            [ctypes.c_double] * 7 + [ctypes.c_int, _dp, _dp, _ip, _dp, _dp, ctypes.c_int]
        L.oracle_traj_euler.restype = ctypes.c_int
        L.oracle_traj_axis3d.argtypes = L.oracle_traj_euler.argtypes
        L.oracle_traj_axis3d.restype = ctypes.c_int
        L.oracle_traj_vsq.argtypes = common + [ctypes.c_long, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int] + [ctypes.c_double] * 7 + \
            [ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_long), ctypes.c_int]
        L.oracle_traj_vsq.restype = ctypes.c_int
        L.oracle_div_rn_mismatches.argtypes = [ctypes.c_double, ctypes.c_long, ctypes.c_ulonglong, ctypes.c_int]
        L.oracle_div_rn_mismatches.restype = ctypes.c_long
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


class _Tables:
    """Keeps the contiguous arrays alive while ctypes holds their pointers."""

    def __init__(self, mol, bt):
        self.n_g, self.n_e, self.n_b = int(mol.n_ground), int(mol.n_excited), int(bt.n_beams)
        self.E, self.pE = _d(mol.energies)
        self.D, self.pD = _d(mol.d_squared)
        self.Z, self.pZ = _d(mol.zeeman_z_diag)
        self.dirs, self.pdirs = _d(bt.dirs)
        self.freqs, self.pfreqs = _d(bt.freqs)
        self.pols, self.ppols = _i(bt.pols)
        self.s0s, self.ps0s = _d(bt.s0s)

    def args(self):
        return [self.n_g, self.n_e, self.n_b, self.pE, self.pD, self.pZ,
                self.pdirs, self.pfreqs, self.ppols, self.ps0s]


def rate_eq(mol, bt, v, z, B_gradient, Gamma_eff=1.0, hbar=1.0545718e-34,
            flavour="jit", saturate=True):
    """One evaluation -> (F [N], p [n_g], R [rad/s])."""
    t = _Tables(mol, bt)
    F = ctypes.c_double()
    R = ctypes.c_double()
    p = np.zeros(t.n_g)
    rc = lib().oracle_rate_eq(*t.args(), float(v), float(z), float(B_gradient), float(mol.Gamma),
                              float(mol.k), float(Gamma_eff), float(hbar),
                              FLAVOURS[flavour], int(bool(saturate)),
                              ctypes.byref(F), p.ctypes.data_as(_dp), ctypes.byref(R))
    if rc != 0:
        raise ValueError("oracle_rate_eq: bad sizes")
    return F.value, p, R.value


def rate_eq_points(mol, bt, v, z, B_gradient, Gamma_eff=1.0, hbar=1.0545718e-34,
                   flavour="jit", saturate=True, n_threads=None, want_pop=True):
    """Evaluate at scattered points (v[i], z[i]) -> F (n,), P (n, n_g), R (n,)."""
    t = _Tables(mol, bt)
    v, pv = _d(np.ravel(v))
    z, pz = _d(np.ravel(z))
    if v.shape != z.shape:
        raise ValueError("v and z must have the same length")
    n = v.size
    F = np.zeros(n)
    R = np.zeros(n)
    P = np.zeros((n, t.n_g)) if want_pop else None
    nt = n_threads or os.cpu_count() or 1
    rc = lib().oracle_rate_eq_batch(*t.args(), n, pv, pz, float(B_gradient), float(mol.Gamma),
                                    float(mol.k), float(Gamma_eff), float(hbar),
                                    FLAVOURS[flavour], int(bool(saturate)),
                                    F.ctypes.data_as(_dp),
                                    P.ctypes.data_as(_dp) if want_pop else None,
                                    R.ctypes.data_as(_dp), int(nt))
    if rc != 0:
        raise ValueError("oracle_rate_eq_batch: bad sizes")
    return F, P, R


def rate_eq_map(mol, bt, v_arr, z_arr, B_gradient, **kw):
    """Map over the (v, z) grid, v-major like solve_rate_equations_batch
    (molmot_cuda/rate_equations_cuda.py:448-544) -> F (nv,nz), P (nv,nz,n_g), R (nv,nz)."""
    v_arr = np.asarray(v_arr, dtype=np.float64)
    z_arr = np.asarray(z_arr, dtype=np.float64)
    V, Z = np.meshgrid(v_arr, z_arr, indexing="ij")
    F, P, R = rate_eq_points(mol, bt, V, Z, B_gradient, **kw)
    nv, nz = v_arr.size, z_arr.size
    return F.reshape(nv, nz), (P.reshape(nv, nz, -1) if P is not None else None), R.reshape(nv, nz)


def traj_euler(mol, bt, z0, v0, dt, n_steps, B_gradient, Gamma_eff=1.0,
               hbar=1.0545718e-34, z_escape=0.015, save_full=False, n_threads=None,
               flavour="jit"):
    """Symplectic-Euler 1D trajectories -> (zf, vf, n_actual, z_traj, v_traj)."""
    t = _Tables(mol, bt)
    z0, pz0 = _d(np.ravel(z0))
    v0, pv0 = _d(np.ravel(v0))
    n = z0.size
    zf = np.zeros(n)
    vf = np.zeros(n)
    na = np.zeros(n, dtype=np.int32)
    zt = np.zeros((n, n_steps)) if save_full else None
    vt = np.zeros((n, n_steps)) if save_full else None
    nt = n_threads or os.cpu_count() or 1
    rc = lib().oracle_traj_euler(*t.args(), n, pz0, pv0, float(dt), int(n_steps), float(mol.mass),
                                 float(B_gradient), float(mol.Gamma), float(mol.k), float(Gamma_eff),
                                 float(hbar), float(z_escape), FLAVOURS[flavour],
                                 zf.ctypes.data_as(_dp), vf.ctypes.data_as(_dp), na.ctypes.data_as(_ip),
                                 zt.ctypes.data_as(_dp) if save_full else None,
                                 vt.ctypes.data_as(_dp) if save_full else None, int(nt))
    if rc != 0:
        raise ValueError("oracle_traj_euler: bad sizes")
    return zf, vf, na, zt, vt


def traj_axis3d(mol, bt, r0, v0, dt, n_steps, B_gradient, Gamma_eff=1.0, hbar=1.0545718e-34,
                r_escape=0.015, save_full=False, n_threads=None, flavour="jit"):
    """3D per-axis trajectories (molmot/propagation/trajectories.py:78-156)
    -> (r_final (n,3), v_final (n,3), n_actual, r_traj | None, v_traj | None)."""
    t = _Tables(mol, bt)
    r0, pr0 = _d(np.reshape(r0, (-1, 3)))
    v0, pv0 = _d(np.reshape(v0, (-1, 3)))
    n = r0.shape[0]
    rf = np.zeros((n, 3))
    vf = np.zeros((n, 3))
    na = np.zeros(n, dtype=np.int32)
    rt = np.zeros((n, n_steps, 3)) if save_full else None
    vt = np.zeros((n, n_steps, 3)) if save_full else None
    nt = n_threads or os.cpu_count() or 1
    rc = lib().oracle_traj_axis3d(*t.args(), n, pr0, pv0, float(dt), int(n_steps), float(mol.mass),
                                  float(B_gradient), float(mol.Gamma), float(mol.k), float(Gamma_eff),
                                  float(hbar), float(r_escape), FLAVOURS[flavour],
                                  rf.ctypes.data_as(_dp), vf.ctypes.data_as(_dp), na.ctypes.data_as(_ip),
                                  rt.ctypes.data_as(_dp) if save_full else None,
                                  vt.ctypes.data_as(_dp) if save_full else None, int(nt))
    if rc != 0:
        raise ValueError("oracle_traj_axis3d: bad sizes")
    return rf, vf, na, rt, vt


def traj_vsq(mol, bt, z0, v0, dt, n_steps, i_eq, B_gradient, Gamma_eff=1.0, hbar=1.0545718e-34,
             z_escape=0.015, n_threads=None, flavour="jit"):
    """Per-particle sum of v^2 over steps j >= i_eq and the sample counts
    (molmot/propagation/trajectories.py:159-229)."""
    t = _Tables(mol, bt)
    z0, pz0 = _d(np.ravel(z0))
    v0, pv0 = _d(np.ravel(v0))
    n = z0.size
    vsq = np.zeros(n)
    cnt = np.zeros(n, dtype=np.int64)
    nt = n_threads or os.cpu_count() or 1
    rc = lib().oracle_traj_vsq(*t.args(), n, pz0, pv0, float(dt), int(n_steps), int(i_eq), float(mol.mass),
                               float(B_gradient), float(mol.Gamma), float(mol.k), float(Gamma_eff), float(hbar),
                               float(z_escape), FLAVOURS[flavour], vsq.ctypes.data_as(_dp),
                               cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), int(nt))
    if rc != 0:
        raise ValueError("oracle_traj_vsq: bad sizes")
    return vsq, cnt


def div_rn_mismatches(y, n=2_000_000, seed=1, correct=True):
    """Number of x among n pseudo-random doubles for which the kernels'
    one-correction quotient (correct=False: the uncorrected product) differs
    from the IEEE quotient x / y."""
    return int(lib().oracle_div_rn_mismatches(float(y), int(n), int(seed), int(bool(correct))))

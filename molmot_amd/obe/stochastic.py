"""Monte Carlo wavefunction (quantum jump) solver on MI355X.

Drop-in for molmot/obe/stochastic.py (SSEProblem, SSESolver, run_ensemble)
and molmot_cuda/stochastic_cuda.py (SSESolverCUDA, run_ensemble_cuda).  The
problem set-up (natural units t Gamma, r k, v k / Gamma; rotated six-beam
polarisations; branching ratios; Zeeman blocks) is done here on the host;
every trajectory runs in the gfx950 kernel behind ``mm_mcwf_run``
(molmot_amd/csrc/mcwf.hip), one lane per trajectory, the whole step loop in
one launch.  There is no CPU integrator: without the HIP library every call
raises.

Precision: ``precision="fp32"`` (default) keeps the amplitudes and the
right-hand side in FP32 with r, v, t, the norm, the jump threshold, the
channel search and the accumulated force in FP64; ``"fp64"`` runs everything
in FP64.

Random numbers are Philox4x32-10 streams keyed by (seed, trajectory index),
so a trajectory's jumps do not depend on how the ensemble is batched.  They
are not numpy's PCG64 stream of the CPU reference (stochastic.py:688) nor
cuRAND's (stochastic_cuda.py:455-456): per-trajectory histories differ from
either reference, the ensemble statistics are what agree.

Beyond the reference: SSEResult.forces holds the normalised dipole force
<psi|F|psi>/<psi|psi> (N) at each save point (the reference leaves it None),
and the device API returns the per-trajectory time integral of that force.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import torch

from ..constants import hbar
from ..hip import _lib

_X_HAT = np.array([1.0, 0.0, 0.0])
_Y_HAT = np.array([0.0, 1.0, 0.0])
_Z_HAT = np.array([0.0, 0.0, 1.0])
BEAM_DIRS = (_X_HAT, _Y_HAT, _Z_HAT, -_X_HAT, -_Y_HAT, -_Z_HAT)
PRECISION = {"fp64": 0, "fp32": 1}
# "reference": jump when 1 - |psi|^2 > -ln U, as molmot does
# (stochastic.py:742-746); a threshold above 1 (probability 1/e) then never
# fires, so the ensemble only approximates the master equation.
# "exact": jump when |psi|^2 < U, the waiting-time rule of the unravelling.
JUMP_RULE = {"reference": 0, "exact": 1}


# ---------------------------------------------------------------------------
# polarisation rotation (molmot/obe/stochastic.py:59-156)
# ---------------------------------------------------------------------------

def _wigner_D_1(cos_beta, sin_beta, alpha, gamma):
    """j = 1 rotation matrix D(alpha, beta, gamma), ZYZ convention, rows and
    columns ordered q = -1, 0, +1 (stochastic.py:59-88)."""
    a_m, a_p = np.exp(-1j * alpha), np.exp(1j * alpha)
    g_m, g_p = np.exp(-1j * gamma), np.exp(1j * gamma)
    cb, sb = cos_beta, sin_beta
    s2 = sb / np.sqrt(2)
    return np.array([[0.5 * (1 + cb) * a_m * g_m, -s2 * a_m, 0.5 * (1 - cb) * a_m * g_p],
                     [s2 * g_m, cb, -s2 * g_p],
                     [0.5 * (1 - cb) * a_p * g_m, s2 * a_p, 0.5 * (1 + cb) * a_p * g_p]], dtype=complex)


def _flip_polarization(eps):
    """Helicity flip of a retro-reflected beam: eps[-1] <-> eps[+1],
    eps[0] -> -eps[0] (stochastic.py:91-100)."""
    return np.array([eps[2], -eps[1], eps[0]], dtype=complex)


def rotate_polarization(pol, k_hat):
    """Express a polarisation given about z in the frame whose quantisation
    axis is the beam direction k_hat (stochastic.py:103-156): Euler angles of
    the axis-angle rotation about k_hat x z (y component negated, the sign
    convention of the reference), then D x = pol solved for x."""
    k_hat = np.asarray(k_hat, dtype=float)
    k_hat = k_hat / np.linalg.norm(k_hat)
    theta = np.arccos(np.clip(k_hat[2], -1.0, 1.0))
    if abs(theta) < 1e-12:
        return np.array(pol, dtype=complex)
    if abs(theta - np.pi) < 1e-12:
        alpha, beta, gamma = 0.0, np.pi, 0.0
    else:
        ax = np.cross(k_hat, _Z_HAT)
        x, y, z = ax[0], -ax[1], ax[2]
        ct, st = np.cos(theta), np.sin(theta)
        one = 1 - ct
        a33 = one * z * z + ct
        a31 = one * z * x - y * st
        a32 = one * z * y + x * st
        a13 = one * x * z + y * st
        a23 = one * y * z - x * st
        alpha = np.arctan2(a23, a13)
        beta = np.arctan2(np.sqrt(max(0.0, 1 - a33 * a33)), a33)
        gamma = np.arctan2(a32, -a31)
    D = _wigner_D_1(np.cos(beta), np.sin(beta), alpha, gamma)
    return np.linalg.solve(D, np.asarray(pol, dtype=complex))


def sample_direction_fast(rng):
    """Isotropic unit vector (stochastic.py:159-166)."""
    ct = 2.0 * rng.random() - 1.0
    st = np.sqrt(1.0 - ct * ct)
    ph = 2.0 * np.pi * rng.random()
    return np.array([st * np.cos(ph), st * np.sin(ph), ct])


def compute_fields_6beam(r, t, omega_states, omega_lasers, sats, pols, beam_radius_k, k=1.0,
                         cached_eps=None, _E_kq_buf=None, _a_buf=None):
    """Six-beam field amplitudes (stochastic.py:173-277), host evaluation:
    E_kq (6, 3) complex for beams +x +y +z -x -y -z and a (6, n_f) the
    per-frequency amplitudes sqrt(s)/(2 sqrt 2) gauss exp(i(-+k r + w t))."""
    nf = len(omega_lasers)
    denom = beam_radius_k ** 2 / 2.0
    rp = np.array([r[1] ** 2 + r[2] ** 2, r[0] ** 2 + r[2] ** 2, r[0] ** 2 + r[1] ** 2])
    gauss = np.sqrt(np.exp(-rp / denom))
    eps = cached_eps if cached_eps is not None else polarization_table(pols)
    a = np.zeros((6, nf), dtype=complex) if _a_buf is None else _a_buf
    a[:] = 0.0
    kr = k * np.asarray(r[:3])
    for f in range(nf):
        G = np.sqrt(sats[f]) / (2.0 * np.sqrt(2.0))
        wt = omega_lasers[f] * t
        for ax in range(3):
            a[ax, f] = G * gauss[ax] * np.exp(1j * (-kr[ax] + wt))
            a[ax + 3, f] = G * gauss[ax] * np.exp(1j * (kr[ax] + wt))
    E = np.zeros((6, 3), dtype=complex) if _E_kq_buf is None else _E_kq_buf
    E[:] = np.einsum("bf,bfq->bq", a, np.conj(eps))
    return E, a


def polarization_table(pols):
    """(6, n_f, 3) polarisations in each beam's frame (the rotation already
    reverses the helicity of the backward beams)."""
    eps = np.zeros((6, len(pols), 3), dtype=complex)
    for i, p in enumerate(pols):
        for b, kd in enumerate(BEAM_DIRS):
            eps[b, i] = rotate_polarization(p, kd)
    return eps


# ---------------------------------------------------------------------------
# problem set-up (molmot/obe/stochastic.py:284-555)
# ---------------------------------------------------------------------------

class SSEProblem:
    """One molecule species in the six-beam MOT, natural units.

    Same constructor and attributes as the reference (stochastic.py:284-440):
    mol_data, beam_pairs [{'freq','q_fwd','q_bwd','s0'}], B_gradient [T/m],
    Omega_Gamma (kept, unused: intensities come from 's0'), Gamma_eff_factor
    (kept, unused: the decay rate is Gamma), beam_radius [m],
    add_spontaneous_kick, diffusion_constants [kg^2 m^2 / s^3] (3,)."""

    def __init__(self, mol_data, beam_pairs, B_gradient, Omega_Gamma=1.0, Gamma_eff_factor=1.0,
                 beam_radius=10e-3, add_spontaneous_kick=True, diffusion_constants=None):
        self.mol_data = mol_data
        self.beam_pairs = beam_pairs
        self.B_gradient = B_gradient
        self.Omega_Gamma = Omega_Gamma
        self.Gamma_eff_factor = Gamma_eff_factor
        self.beam_radius = beam_radius
        self.add_spontaneous_kick = add_spontaneous_kick
        self.Gamma = mol_data.Gamma
        self.k = mol_data.k
        self.mass = mol_data.mass
        self.wavelength = mol_data.wavelength
        self.n_ground = mol_data.n_ground
        self.n_excited = mol_data.n_excited
        self.n_states = mol_data.n_states
        self.v_recoil_SI = hbar * self.k / self.mass
        self.t_nat = 1.0 / self.Gamma
        self.r_nat = 1.0 / self.k
        self.v_nat = self.Gamma / self.k
        self.F_nat = hbar * self.k * self.Gamma
        self.v_recoil_nat = self.v_recoil_SI / self.v_nat
        self.mass_nat = self.mass * self.Gamma / (hbar * self.k ** 2)
        if diffusion_constants is not None:
            self.diffusion_nat = np.asarray(diffusion_constants, dtype=float) / ((hbar * self.k) ** 2 * self.Gamma)
        else:
            self.diffusion_nat = np.zeros(3)
        ng, ne = self.n_ground, self.n_excited
        self.d_ge = np.abs(np.asarray(mol_data.tdm)[:ng, ng:ng + ne, :]).astype(float)
        self.d_eg = np.transpose(self.d_ge, (1, 0, 2))
        d2 = self.d_ge ** 2
        self.branching = np.zeros((ne, ng, 3))
        for ie in range(ne):
            tot = 0.0
            for ig in range(ng):
                for q in range(3):
                    tot += d2[ig, ie, q]
            if tot > 1e-30:
                self.branching[ie] = d2[:, ie, :] / tot
        self.omega0s = 2.0 * np.pi * mol_data.energies / self.Gamma
        self.zeeman_z_diag = mol_data.zeeman_z_diag
        self.zeeman_x_ground = np.array(mol_data.zeeman_x[:ng, :ng], dtype=complex)
        self.zeeman_y_ground = np.array(mol_data.zeeman_y[:ng, :ng], dtype=complex)
        self.zeeman_z_ground = np.array(mol_data.zeeman_z[:ng, :ng], dtype=complex)
        self._setup_beam_params()
        self.beam_radius_k = beam_radius * self.k
        self.n_psi = self.n_states
        self.r_idx = 2 * self.n_states
        self.v_idx = self.r_idx + 3
        self.state_size = 2 * self.n_states + 6

    def _setup_beam_params(self):
        n = len(self.beam_pairs)
        self.omega_lasers = np.array([2.0 * np.pi * bp["freq"] / self.Gamma for bp in self.beam_pairs])
        self.sats = np.array([float(bp["s0"]) for bp in self.beam_pairs])
        self.pols = []
        for bp in self.beam_pairs:
            p = np.zeros(3, dtype=complex)
            p[bp["q_fwd"]] = 1.0
            self.pols.append(p)
        assert len(self.pols) == n

    def build_initial_state(self, r0_SI, v0_SI, psi0=None):
        """State vector [Re psi (N), Im psi (N), r (3), v (3)] in natural
        units (stochastic.py:442-474); default psi = first ground state."""
        N = self.n_states
        u0 = np.zeros(self.state_size)
        if psi0 is None:
            u0[0] = 1.0
        else:
            psi0 = np.asarray(psi0, dtype=complex)
            psi0 = psi0 / np.linalg.norm(psi0)
            u0[:N] = psi0.real
            u0[N:2 * N] = psi0.imag
        u0[self.r_idx:self.r_idx + 3] = np.asarray(r0_SI) / self.r_nat
        u0[self.v_idx:self.v_idx + 3] = np.asarray(v0_SI) / self.v_nat
        return u0

    def get_B_field_gauss(self, r_nat):
        """B'(-x/2, -y/2, z) in Gauss at r (natural units) (stochastic.py:526-549)."""
        rs = np.asarray(r_nat) * self.r_nat
        return np.array([-self.B_gradient * rs[0] / 2.0 * 1e4, -self.B_gradient * rs[1] / 2.0 * 1e4,
                         self.B_gradient * rs[2] * 1e4])

    def get_omega0s_zeeman(self, r_nat):
        """Energies with the diagonal Zeeman shift at r (stochastic.py:476-524)."""
        Bx, By, Bz = self.get_B_field_gauss(r_nat)
        ng = self.n_ground
        z = np.zeros(self.n_states)
        z[:ng] = np.real(Bx * np.diag(self.zeeman_x_ground) + By * np.diag(self.zeeman_y_ground)
                         + Bz * np.diag(self.zeeman_z_ground))
        z[ng:] = Bz * self.zeeman_z_diag[ng:]
        return self.omega0s + z


@dataclass
class SSEResult:
    """One trajectory (stochastic.py:556-570)."""
    times: np.ndarray
    positions: np.ndarray
    velocities: np.ndarray
    populations: np.ndarray
    photons_scattered: int = 0
    jump_times: list = field(default_factory=list)
    jump_channels: list = field(default_factory=list)
    forces: Optional[np.ndarray] = None


# ---------------------------------------------------------------------------
# device plan
# ---------------------------------------------------------------------------

class MCWFPlan:
    """Device tables of one SSEProblem (the kernel's McProblem)."""

    def __init__(self, problem: SSEProblem):
        _lib.require_gpu()
        p = problem
        ng, ne, N = p.n_ground, p.n_excited, p.n_states
        for nm in ("zeeman_x_ground", "zeeman_y_ground", "zeeman_z_ground"):
            Z = getattr(p, nm)
            if not np.allclose(Z, Z.conj().T, rtol=0, atol=1e-12 * max(1.0, np.max(np.abs(Z)))):
                raise ValueError(f"{nm} is not Hermitian; the kernel applies H_Z through its upper triangle")
        self.problem = p
        self.shift = (float(np.mean(p.omega0s[:ng])), float(np.mean(p.omega0s[ng:])))
        wbase = p.omega0s.astype(float).copy()
        wbase[:ng] -= self.shift[0]
        wbase[ng:] -= self.shift[1]
        Z = np.stack([p.zeeman_x_ground, p.zeeman_y_ground, p.zeeman_z_ground])
        eps = polarization_table(p.pols)
        G = np.sqrt(p.sats) / (2.0 * np.sqrt(2.0))
        amp = G[None, :, None] * np.conj(eps)                       # (6, nf, 3)
        wl = p.omega_lasers - (self.shift[1] - self.shift[0])
        self.eps = eps
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (wbase, Z.real, Z.imag, np.asarray(p.zeeman_z_diag)[ng:], p.d_ge, p.branching, amp.real,
                 amp.imag, wl, p.diffusion_nat)]
        dp = [a.ctypes.data_as(_lib._dp) for a in arrs]
        self._h = ctypes.c_void_p()
        rc = _lib.lib().mm_mcwf_problem_create(
            ctypes.byref(self._h), ng, ne, len(wl), *dp[:9], float(p.B_gradient), float(p.r_nat),
            float(p.beam_radius_k), float(p.mass_nat), float(p.v_recoil_nat), int(bool(p.add_spontaneous_kick)),
            dp[9])
        _lib.check(rc, "mm_mcwf_problem_create")
        self.N = N

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.check(_lib.lib().mm_mcwf_problem_destroy(self._h), "mm_mcwf_problem_destroy")
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def stage_times(dt_nat: float, n_steps: int) -> np.ndarray:
        """t_k accumulated step by step in FP64 (as the integrator does) and
        t_k + dt/2, interleaved: (2 (n_steps + 1),)."""
        t = np.empty(n_steps + 1)
        t[0] = 0.0
        if n_steps:
            t[1:] = np.cumsum(np.full(n_steps, dt_nat))
        out = np.empty(2 * (n_steps + 1))
        out[0::2] = t
        out[1::2] = t + 0.5 * dt_nat
        return out

    def field_table(self, dt_nat: float, n_steps: int, precision: str, device, stream=None):
        tt = torch.from_numpy(self.stage_times(dt_nat, n_steps)).to(device)
        rdt = torch.float64 if precision == "fp64" else torch.float32
        S = torch.empty((tt.numel(), 18, 2), dtype=rdt, device=device)
        rc = _lib.lib().mm_mcwf_field_table(self._h, tt.numel(), _lib.ptr(tt), PRECISION[precision], _lib.ptr(S),
                                            _lib.stream_handle(stream))
        _lib.check(rc, "mm_mcwf_field_table")
        return S


def mcwf_ensemble_device(plan: MCWFPlan, r0: torch.Tensor, v0: torch.Tensor, dt_nat: float, n_steps: int,
                         save_every: int = 100, seed: int = 0, psi0: Optional[torch.Tensor] = None,
                         precision: str = "fp32", r_escape_nat: float = 0.0, renorm_interval: int = 0,
                         avg_start: int = 0, jump_cap: int = 0, jump_rule: str = "reference",
                         want=("r", "v", "pop", "force", "force_sum", "final"), stream=None):
    """Run len(r0) trajectories on the device (natural units).

    r0, v0: (n, 3) float64 device tensors; psi0: (n, N) complex128 or None
    (first ground state).  Returns a dict of device tensors: n_scatters,
    escaped, n_saved, n_avg (n,) int32 and, as requested by ``want``,
    r, v, force (n, n_save, 3), pop (n, n_save, N), force_sum (n, 3),
    r_final, v_final (n, 3), psi_final (n, N) complex128, jump_t, jump_ch
    (n, jump_cap).  Also 'times' (n_save,) numpy (natural units)."""
    if precision not in PRECISION:
        raise ValueError(f"precision must be one of {tuple(PRECISION)}")
    if jump_rule not in JUMP_RULE:
        raise ValueError(f"jump_rule must be one of {tuple(JUMP_RULE)}")
    dev = r0.device
    n = r0.shape[0]
    N = plan.N
    r0 = r0.contiguous().to(torch.float64)
    v0 = v0.contiguous().to(torch.float64)
    if r0.shape != (n, 3) or v0.shape != (n, 3):
        raise ValueError("r0 and v0 must be (n, 3)")
    n_save = n_steps // save_every + 1
    p0 = None
    if psi0 is not None:
        psi0 = psi0.to(dev)
        if psi0.shape != (n, N):
            raise ValueError(f"psi0 must be (n, {N})")
        p0 = torch.view_as_real(psi0.to(torch.complex128).contiguous()).contiguous()
    S = plan.field_table(dt_nat, n_steps, precision, dev, stream)
    f64 = dict(dtype=torch.float64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    out = {"n_scatters": torch.zeros(n, **i32), "escaped": torch.zeros(n, **i32), "n_saved": torch.zeros(n, **i32),
           "n_avg": torch.zeros(n, **i32)}
    if "r" in want:
        out["r"] = torch.zeros((n, n_save, 3), **f64)
    if "v" in want:
        out["v"] = torch.zeros((n, n_save, 3), **f64)
    if "force" in want:
        out["force"] = torch.zeros((n, n_save, 3), **f64)
    if "pop" in want:
        out["pop"] = torch.zeros((n, n_save, N), **f64)
    if "force_sum" in want:
        out["force_sum"] = torch.zeros((n, 3), **f64)
    if "final" in want:
        out["r_final"] = torch.zeros((n, 3), **f64)
        out["v_final"] = torch.zeros((n, 3), **f64)
        out["psi_final"] = torch.zeros((n, N, 2), **f64)
    if jump_cap > 0:
        out["jump_t"] = torch.zeros((n, jump_cap), **f64)
        out["jump_ch"] = torch.full((n, jump_cap), -1, **i32)
    cfg = _lib.MCWFConfig(dt=float(dt_nat), r_escape=float(r_escape_nat), n_steps=int(n_steps),
                          save_every=int(save_every), n_save=int(n_save), renorm_interval=int(renorm_interval),
                          avg_start=int(avg_start), precision=PRECISION[precision], jump_cap=int(jump_cap),
                          jump_rule=JUMP_RULE[jump_rule],
                          seed=int(seed) & 0xFFFFFFFFFFFFFFFF)
    o = _lib.MCWFOutputs(**{k: (_lib.ptr(out[k]) if k in out else None) for k in
                            ("r", "v", "pop", "force", "force_sum", "psi_final", "r_final", "v_final", "jump_t",
                             "n_scatters", "escaped", "n_saved", "n_avg", "jump_ch")})
    rc = _lib.lib().mm_mcwf_run(plan._h, n, _lib.ptr(r0), _lib.ptr(v0), _lib.ptr(p0), _lib.ptr(S), S.shape[0],
                                ctypes.byref(cfg), ctypes.byref(o), _lib.stream_handle(stream))
    _lib.check(rc, "mm_mcwf_run")
    if "psi_final" in out:
        out["psi_final"] = torch.view_as_complex(out["psi_final"])
    t = MCWFPlan.stage_times(dt_nat, n_steps)[0::2]
    out["times"] = t[::save_every][:n_save]
    return out


# ---------------------------------------------------------------------------
# reference-shaped drivers
# ---------------------------------------------------------------------------

def _dt_steps(p: SSEProblem, t_max, dt):
    dt_SI = 0.01 / p.Gamma if dt is None else dt
    dt_nat = dt_SI / p.t_nat
    return dt_nat, int((t_max / p.t_nat) / dt_nat)


class SSESolverHIP:
    """SSESolverCUDA / SSESolver API (stochastic_cuda.py:861-1116,
    stochastic.py:579-804) on the gfx950 kernel."""

    def __init__(self, problem: SSEProblem, precision: str = "fp32", jump_rule: str = "reference"):
        self.prob = problem
        self.precision = precision
        self.jump_rule = jump_rule
        self.plan = MCWFPlan(problem)

    def run(self, r0, v0, t_max, dt=None, psi0=None, save_every=100, rng_seed=None, r_escape=None,
            renorm_interval=0):
        """One trajectory (SI in and out); jump times and channels are logged."""
        return self.run_ensemble(1, lambda: np.asarray(r0, float), lambda: np.asarray(v0, float), t_max, dt=dt,
                                 psi0=psi0, save_every=save_every, rng_seed=rng_seed, r_escape=r_escape,
                                 renorm_interval=renorm_interval, log_jumps=True)[0]

    def run_ensemble(self, n_particles, r0_sampler, v0_sampler, t_max, dt=None, psi0=None, save_every=100,
                     rng_seed=None, r_escape=None, renorm_interval=0, callback=None, log_jumps=False):
        """All trajectories in one launch; samplers are called on the host
        (r0 then v0 per particle, as stochastic_cuda.py:1006-1020)."""
        p = self.prob
        N = p.n_states
        dt_nat, n_steps = _dt_steps(p, t_max, dt)
        r0 = np.zeros((n_particles, 3))
        v0 = np.zeros((n_particles, 3))
        for i in range(n_particles):
            r0[i] = np.asarray(r0_sampler(), dtype=float) / p.r_nat
            v0[i] = np.asarray(v0_sampler(), dtype=float) / p.v_nat
        dev = torch.device("cuda", torch.cuda.current_device())
        ps = None
        if psi0 is not None:
            pi = np.asarray(psi0, dtype=complex)
            pi = pi / np.linalg.norm(pi)
            ps = torch.from_numpy(np.repeat(pi[None, :], n_particles, 0)).to(dev)
        if rng_seed is None:
            rng_seed = int(np.random.default_rng().integers(0, 2 ** 62))
        out = mcwf_ensemble_device(self.plan, torch.from_numpy(r0).to(dev), torch.from_numpy(v0).to(dev), dt_nat,
                                   n_steps, save_every, rng_seed, ps, self.precision,
                                   0.0 if r_escape is None else r_escape / p.r_nat, renorm_interval, 0,
                                   n_steps if log_jumps else 0, self.jump_rule, want=("r", "v", "pop", "force"))
        h = {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in out.items()}
        res = []
        ng = p.n_ground
        for i in range(n_particles):
            ns = max(int(h["n_saved"][i]), 1)
            jt, jc = [], []
            if log_jumps:
                m = min(int(h["n_scatters"][i]), h["jump_t"].shape[1])
                jt = list(h["jump_t"][i, :m] * p.t_nat)
                jc = [(int(c) // 3 // ng, int(c) // 3 % ng, int(c) % 3) for c in h["jump_ch"][i, :m]]
            r = SSEResult(times=h["times"][:ns] * p.t_nat, positions=h["r"][i, :ns] * p.r_nat,
                          velocities=h["v"][i, :ns] * p.v_nat, populations=h["pop"][i, :ns],
                          photons_scattered=int(h["n_scatters"][i]), jump_times=jt, jump_channels=jc,
                          forces=h["force"][i, :ns] * p.F_nat)
            res.append(r)
            if callback is not None:
                callback(i, r)
        return res


SSESolverCUDA = SSESolverHIP
SSESolver = SSESolverHIP


def run_ensemble_hip(problem, n_particles, r0_sampler, v0_sampler, t_max, dt=None, save_every=100, rng_seed=None,
                     r_escape=None, renorm_interval=0, callback=None, precision="fp32", jump_rule="reference"):
    """run_ensemble_cuda (stochastic_cuda.py:1123-1173) on the gfx950 kernel."""
    return SSESolverHIP(problem, precision, jump_rule).run_ensemble(n_particles, r0_sampler, v0_sampler, t_max, dt=dt,
                                                         save_every=save_every, rng_seed=rng_seed,
                                                         r_escape=r_escape, renorm_interval=renorm_interval,
                                                         callback=callback)


run_ensemble_cuda = run_ensemble_hip


def run_ensemble(problem, n_particles, r0_sampler, v0_sampler, t_max, dt=None, save_every=100, rng_seed=None,
                 callback=None):
    """run_ensemble (stochastic.py:1074-1119) on the gfx950 kernel."""
    return run_ensemble_hip(problem, n_particles, r0_sampler, v0_sampler, t_max, dt=dt, save_every=save_every,
                            rng_seed=rng_seed, callback=callback)


__all__ = ["SSEProblem", "SSEResult", "SSESolver", "SSESolverHIP", "SSESolverCUDA", "MCWFPlan",
           "mcwf_ensemble_device", "run_ensemble", "run_ensemble_hip", "run_ensemble_cuda",
           "rotate_polarization", "compute_fields_6beam", "sample_direction_fast", "polarization_table",
           ]

"""CPU oracle of the molecular-dynamics step -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module.  A plain FP64 restatement (Python floats
and numpy, no contraction into FMA) of RBMD's step either side of the force,
in the reference's order of operations:

* velocity half-kick  v += 0.5 F / m dt ftm2v
  (src/simulate_pipeline/src/controller/velocity_controller/
  velocity_controller_op/rocm/update_velocity_op.hip.cu:5-27);
* drift x += v dt and one periodic image shift with image-flag counting
  (position_controller_op/rocm/update_position_op.hip.cu:17-37,
  src/data_manager/include/model/box.h:133-173);
* temperature T = 0.5 mvv2e sum m v^2 / (tdof kB / 2), tdof = 3N - 3
  (nose_hoover_controller.cpp:256-287, 1013-1025);
* the Nose-Hoover chain (3 links, one loop, no drag; SetUp, ComputeTempTarget,
  NHCTempIntegrate: nose_hoover_controller.cpp:415-497, 732-840) and the NVT
  velocity-Verlet order (nvt_ensemble.cpp:139-167, InitialIntegrate /
  FinalIntegrate: nose_hoover_controller.cpp:633-700);
* the energy-stable rescale of NVE runs with RBL lists
  (energy_stable_scheme_controller.cpp:56-144, nve_ensemble.cpp:58-88) with
  the Hamiltonian of step 1 as reference.

Pins (tests/test_oracle_md.py): with no force the unwrapped drift is exact and
image flags count the crossings; velocity Verlet conserves a harmonic
system's energy to O(dt^2) and is time reversible; the chain's extended
energy (kinetic + potential + the FixNH chain energy) is conserved by
NVT runs of harmonic oscillators, and the chain drives the temperature to its
target; the energy-stable rescale contracts H - H_init by exactly 9/10 per
step when the potential energy is constant (the closed form of
xi^2 = 1 + dt / (10 dt KE) (H_init - H)).
"""
import math

import numpy as np

CHAIN = 3

# kB, ftm2v, mvv2e of RBMD's unit styles (src/common/unit_factor.cpp:23-37)
UNITS = {
    "LJ": (1.0, 1.0, 1.0),
    "REAL": (1.9872067 * 10.0 ** -3.0, 4.186 * 10.0 ** -4.0, 1.0 / (4.186 * 10.0 ** -4.0)),
    "METAL": (8.617343e-5, 1.0 / 1.0364269e-4, 1.0364269e-4),
}


def wrap(x, img, lo, L):
    """One image shift per dimension (ApplyPBC), flags counted; in place."""
    for d in range(3):
        hi = lo[d] + L[d]
        m = x[:, d] >= hi
        x[m, d] -= L[d]
        img[m, d] += 1
        m = x[:, d] < lo[d]
        x[m, d] += L[d]
        img[m, d] -= 1


def half_kick(v, f, m, dt, ftm2v):
    """v + 0.5 F / m dt ftm2v, evaluated left to right as UpdateVelocity."""
    return v + 0.5 * f / m[:, None] * dt * ftm2v


def temp_sum(v, m, mvv2e):
    return float(np.sum(m * np.sum(v * v, axis=1))) * mvv2e


def temperature(ts, tdof, kb):
    return 0.5 * ts / (tdof * kb / 2.0)


def target_temperature(t_start, t_stop, num_steps, step):
    """ComputeTempTarget: linear ramp over num_steps."""
    if t_stop == t_start:
        return t_start
    delta = float(step)
    if delta != 0.0:
        delta = delta / float(num_steps)
    return t_start + delta * (t_stop - t_start)


class Chain:
    """Nose-Hoover chain state (eta, eta_dot, eta_dotdot, eta_mass)."""

    def __init__(self):
        self.eta = [0.0] * CHAIN
        self.eta_dot = [0.0] * (CHAIN + 1)
        self.eta_dotdot = [0.0] * CHAIN
        self.eta_mass = [0.0] * CHAIN
        self.T = 0.0
        self.t_target = 0.0

    def setup(self, T, t_target, tdof, kb, t_freq):
        """SetUp: masses from the target, forces on links 1.. of a chain at rest."""
        self.T = T
        self.t_target = t_target
        ff = t_freq * t_freq
        self.eta_mass[0] = tdof * kb * t_target / ff
        for c in range(1, CHAIN):
            self.eta_mass[c] = kb * t_target / ff
        for c in range(1, CHAIN):
            self.eta_dotdot[c] = (self.eta_mass[c - 1] * self.eta_dot[c - 1] *
                                  self.eta_dot[c - 1] - kb * t_target) / self.eta_mass[c]

    def half(self, dt, tdof, kb, t_freq):
        """NHCTempIntegrate; returns the velocity factor, T rescaled."""
        dthalf, dt4, dt8 = 0.5 * dt, 0.25 * dt, 0.125 * dt
        tt = self.t_target
        tdkb = tdof * kb
        ke_target = tdkb * tt
        ff = t_freq * t_freq
        ke = tdkb * self.T
        self.eta_mass[0] = tdkb * tt / ff
        for c in range(1, CHAIN):
            self.eta_mass[c] = kb * tt / ff
        self.eta_dotdot[0] = (ke - ke_target) / self.eta_mass[0] if self.eta_mass[0] > 0 else 0.0
        ed = self.eta_dot
        for c in range(CHAIN - 1, 0, -1):
            e = math.exp(-dt8 * ed[c + 1])
            ed[c] *= e
            ed[c] += self.eta_dotdot[c] * dt4
            ed[c] *= e
        e0 = math.exp(-dt8 * ed[1])
        ed[0] *= e0
        ed[0] += self.eta_dotdot[0] * dt4
        ed[0] *= e0
        factor = math.exp(-dthalf * ed[0])
        self.T = self.T * factor * factor
        ke = tdkb * self.T
        self.eta_dotdot[0] = (ke - ke_target) / self.eta_mass[0] if self.eta_mass[0] > 0 else 0.0
        for c in range(CHAIN):
            self.eta[c] += dthalf * ed[c]
        ed[0] *= e0
        ed[0] += self.eta_dotdot[0] * dt4
        ed[0] *= e0
        for c in range(1, CHAIN):
            e = math.exp(-dt8 * ed[c + 1])
            ed[c] *= e
            self.eta_dotdot[c] = (self.eta_mass[c - 1] * ed[c - 1] * ed[c - 1] -
                                  kb * tt) / self.eta_mass[c]
            ed[c] += self.eta_dotdot[c] * dt4
            ed[c] *= e
        return factor

    def energy(self, tdof, kb):
        """FixNH chain energy: ke_target eta_0 + kT sum_{i>0} eta_i + sum Q eta_dot^2 / 2."""
        kt = kb * self.t_target
        e = 0.0
        for c in range(CHAIN):
            e += 0.5 * self.eta_mass[c] * self.eta_dot[c] ** 2
            e += (tdof * kt if c == 0 else kt) * self.eta[c]
        return e


def energy_stable_xi(ke, pe, ke_init, pe_init, dt):
    """xi = sqrt(1 + dt / (gamma KE) (H_init - H)), gamma = 10 dt."""
    gamma = 10.0 * dt
    if ke <= 0.0:
        return 1.0
    return math.sqrt(1.0 + dt / (gamma * ke) * ((pe_init + ke_init) - (pe + ke)))


class Md:
    """The reference's step sequence around a force callback.

    force(x, step) -> (F (N,3), pe_per_atom or None).  mode: "NVE",
    "NVE-ES" (energy-stable) or "NVT" (Nose-Hoover).  Arrays are FP64,
    image flags int64; everything is updated in place.
    """

    def __init__(self, force, x, v, img, mass, lo, L, mode, dt, unit="LJ",
                 t_control=(1.0, 1.0, 1.0), num_steps=0):
        self.force, self.x, self.v, self.img, self.m = force, x, v, img, mass
        self.lo, self.L = np.asarray(lo, float), np.asarray(L, float)
        self.mode, self.dt = mode, dt
        self.kb, self.ftm2v, self.mvv2e = UNITS[unit]
        self.n = len(x)
        self.tdof = 3.0 * self.n - 3.0
        self.t_start, self.t_stop, t_damp = t_control
        self.t_freq = 1.0 / t_damp
        self.num_steps = num_steps
        self.chain = Chain()
        self.step = 0
        self.ke_init = self.pe_init = 0.0
        self.xi = 1.0
        # init: wrap, force of step 0, thermostat set-up
        wrap(self.x, self.img, self.lo, self.L)
        self.F, self.pe = force(self.x, 0)
        ts = temp_sum(self.v, self.m, self.mvv2e)
        self.T = temperature(ts, self.tdof, self.kb)
        self.ke = ts / (2.0 * self.n)
        if mode == "NVT":
            self.chain.setup(self.T, target_temperature(self.t_start, self.t_stop,
                                                        num_steps, 0),
                             self.tdof, self.kb, self.t_freq)

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step += 1
            n = self.step
            if self.mode == "NVT":
                self.chain.t_target = target_temperature(self.t_start, self.t_stop,
                                                         self.num_steps, n)
                self.v *= self.chain.half(self.dt, self.tdof, self.kb, self.t_freq)
            self.v[:] = half_kick(self.v, self.F, self.m, self.dt, self.ftm2v)
            self.x += self.v * self.dt
            wrap(self.x, self.img, self.lo, self.L)
            self.F, self.pe = self.force(self.x, n)
            self.v[:] = half_kick(self.v, self.F, self.m, self.dt, self.ftm2v)
            ts = temp_sum(self.v, self.m, self.mvv2e)
            self.T = temperature(ts, self.tdof, self.kb)
            self.ke = ts / (2.0 * self.n)
            if self.mode == "NVT":
                self.chain.T = self.T
                self.v *= self.chain.half(self.dt, self.tdof, self.kb, self.t_freq)
                self.T = self.chain.T
            elif self.mode == "NVE-ES":
                if n == 1:
                    self.ke_init, self.pe_init = self.ke, self.pe
                self.xi = energy_stable_xi(self.ke, self.pe, self.ke_init, self.pe_init,
                                           self.dt)
                self.v *= self.xi

    def chain_energy(self):
        return self.chain.energy(self.tdof, self.kb) if self.mode == "NVT" else 0.0

"""Restatement of the reference's fluid velocity autocorrelation,
Vel_Autocorr_Fluid (star-polymers/lib/Analysis/Vel_Autocorr_Fluid.h:19-116).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
windows and lag sums (k_fluid_vacf); never by the library.  Plain Python
loops over particles, as the reference's functor runs, for small systems.
Pinned by hand-worked sequences in tests/test_observables_cpu.py.
"""
from __future__ import annotations

import numpy as np


class VelAutocorrFluid:
    """Vel_Autocorr_Fluid(N_p, N, dt) (:31-41)."""

    def __init__(self, n_particles: int, entries: int, dt: float):
        self.entries = int(entries)
        self.n = int(n_particles)
        self.dt = float(dt)
        self.particle_counter = 0
        self.entry_counter = 0
        self.counter = 0
        self.velocities = [[np.zeros(3) for _ in range(self.entries)] for _ in range(self.n)]
        self.autocorrelation = np.zeros((self.n, self.entries))
        self.average = np.zeros(self.entries)

    def particle(self, v):
        """operator()(const Particle&) (:44-66): one particle, in order."""
        w = self.velocities[self.particle_counter]
        w.pop(0)
        w.append(np.asarray(v, np.float64).copy())
        self.particle_counter += 1
        if self.particle_counter == self.n and self.entry_counter < self.entries:
            self.entry_counter += 1
        if self.particle_counter == self.n and self.entry_counter == self.entries:
            self.calculate()
        if self.particle_counter == self.n:
            self.particle_counter = 0

    def sample(self, vel):
        """MPC::operator()(functor) over the solvent: vel (N_p, 3), row = particle."""
        for v in vel:
            self.particle(v)

    def calculate(self):
        """calculate() (:68-78)."""
        for i in range(self.n):
            w = self.velocities[i]
            for j in range(self.entries):
                self.autocorrelation[i][j] += float(np.dot(w[0], w[j]))
        self.counter += 1

    def lag_sums(self):
        """sum over particles of the accumulated products, per lag."""
        return self.autocorrelation.sum(axis=0)

    def calculate_average(self):
        """calculate_average() (:80-88)."""
        for i in range(self.entries):
            self.average[i] = 0.0
            for j in range(self.n):
                self.average[i] += self.autocorrelation[j][i] / self.counter
            self.average[i] /= self.n

    def value(self) -> float:
        """value() (:90-97)."""
        d = 0.0
        for el in self.average:
            d += el * self.dt / self.counter
        return d / 3.0

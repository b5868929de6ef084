"""Analysis observers of the reference and its statistics file, on top of
the device samplers of MPCSystem.

The reference applies functors to every solvent particle on the CPU
(`MPC::operator()`, star-polymers/lib/Hydrodynamics/MPC.h:110-116) and keeps
their running results in std::map's.  Here one device pass per sample
returns the per-bin sums (MPCSystem.velocity_profile / fourier_modes /
speed_histogram, kernels in csrc/spmd_kernels.hip), and these classes keep
the running results and print them in the reference's formats.  They also
accept raw samples, so the same accumulation serves any source of samples.

  VelocityProfile   VelocityX         lib/Analysis/Velocity_x.h:22-148
  FourierAutocorr   Fourier_Autocorr  lib/Analysis/Fourier_Autocorr.h:19-127
  SpeedDistribution VelocityDistribution lib/Analysis/Vel_distribution.h:17-57
  FluidVelocityAutocorr Vel_Autocorr_Fluid lib/Analysis/Vel_Autocorr_Fluid.h:19-116
  StatisticsFile    statistics*.dat   src/star-polymers.cpp:307, 466-486
  run_with_output   the driver loop   src/star-polymers.cpp:412-536
  run_mpc_test      the solvent test  src/mpc_test.cpp:42-146
"""
from __future__ import annotations

from collections import deque
from typing import Callable, TextIO

import numpy as np

# the SRD kinematic viscosity at rho = 10, h = 0.1, alpha = 130 deg, kT = 1
# (kinetic + collisional parts), the curve results/plot_MPC_tests.py:107
# draws through the transverse autocorrelation
def srd_viscosity(n: float = 10.0, h: float = 0.1, alpha_deg: float = 130.0, kT: float = 1.0) -> float:
    """Kinematic shear viscosity of SRD with random grid shift, unit cells
    and unit particle mass: nu_kin = kT h / 2 [5 n / ((n - 1 + e^-n)(2 -
    cos a - cos 2a)) - 1], nu_coll = (1 - cos a) / (18 h) (n - 1 + e^-n) / n."""
    a = np.deg2rad(alpha_deg)
    z = n - 1.0 + np.exp(-n)
    nu_kin = kT * h / 2.0 * (5.0 * n / (z * (2.0 - np.cos(a) - np.cos(2.0 * a))) - 1.0)
    nu_coll = (1.0 - np.cos(a)) / (18.0 * h) * z / n
    return float(nu_kin + nu_coll)


class VelocityProfile:
    """VelocityX (Velocity_x.h:22-148): mean solvent velocity relative to
    the star's centre of mass, binned by floor(x/width) and floor(y/width)
    after the Lees-Edwards wrap; map 0 is keyed by x, map 1 by y."""

    def __init__(self, width: float = 0.5):
        self.width = float(width)
        self.maps: list[dict[float, np.ndarray]] = [{}, {}]

    def add(self, sample):
        """sample: [(keys, sums)] per map, sums[:, 0:3] = sum vx, sum vy,
        count.  Bins with no particle do not enter the maps, as in the
        reference's std::map."""
        for a in range(2):
            keys, sums = sample[a]
            m = self.maps[a]
            for k, s in zip(np.asarray(keys, float), np.asarray(sums, float)):
                if s[2] == 0:
                    continue
                k = float(k)
                if k in m:
                    m[k] += s
                else:
                    m[k] = s.copy()

    def sample(self, system, r: int = 0):
        """One application of the functor to the current solvent."""
        self.add(system.velocity_profile(self.width, r))

    def table(self, axis: int):
        """keys, mean vx, mean vy, count of one map in key order."""
        m = self.maps[axis]
        keys = np.array(sorted(m))
        if keys.size == 0:
            return keys, keys.copy(), keys.copy(), keys.copy()
        s = np.array([m[k] for k in keys])
        return keys, s[:, 0] / s[:, 2], s[:, 1] / s[:, 2], s[:, 2]

    def print_result(self, os: TextIO):
        """VelocityX::print_result (Velocity_x.h:98-144): both maps, one bin
        per line (key, mean vx, mean vy, count; %.8e), a blank line between."""
        for a in range(2):
            keys, vx, vy, n = self.table(a)
            for row in zip(keys, vx, vy, n):
                os.write("%.8e %.8e %.8e %.8e\n" % row)
            if a == 0:
                os.write("\n")

    def shear_rate(self, y_max: float | None = None) -> float:
        """Count-weighted least-squares slope of mean vx against the bin
        centre of the y map (the analytic profile is vx = shear * y,
        results/plot_MPC_tests.py:61), over |centre| <= y_max if given."""
        keys, vx, _, n = self.table(1)
        y = keys + 0.5 * self.width
        sel = np.ones_like(y, bool) if y_max is None else np.abs(y) <= y_max
        return weighted_slope(y[sel], vx[sel], n[sel])


def weighted_slope(x, y, w) -> float:
    x, y, w = (np.asarray(v, float) for v in (x, y, w))
    xm = np.sum(w * x) / np.sum(w)
    ym = np.sum(w * y) / np.sum(w)
    return float(np.sum(w * (x - xm) * (y - ym)) / np.sum(w * (x - xm) ** 2))


class FourierAutocorr:
    """Fourier_Autocorr (Fourier_Autocorr.h:19-127): autocorrelation of the
    transverse velocity modes V_j at k = 2 pi / L e_j over a sliding window
    of `entries` samples, `dt` apart.  The window starts filled with zeros;
    from the entries-th sample on, every sample adds Re(V_j(oldest)^H
    V_j(oldest + i)) to lag i (Fourier_Autocorr.h:80-93)."""

    def __init__(self, entries: int = 50, dt: float = 1.0):
        self.entries = int(entries)
        self.dt = float(dt)
        self.window = deque([np.zeros((3, 3), complex)] * self.entries, maxlen=self.entries)
        self.filled = 0
        self.sums = np.zeros((3, self.entries))
        self.count = 0

    def add(self, modes):
        """modes: complex (3, 3), row j = V_j."""
        self.window.append(np.asarray(modes, complex).copy())
        if self.filled < self.entries:
            self.filled += 1
        if self.filled == self.entries:
            w = list(self.window)
            first = w[0]
            for i in range(self.entries):
                self.sums[:, i] += np.real(np.sum(np.conj(first) * w[i], axis=1))
            self.count += 1

    def sample(self, system, r: int = 0):
        self.add(system.fourier_modes(r))

    def correlation(self) -> np.ndarray:
        """(3, entries): the averaged autocorrelation for k along x, y, z."""
        return self.sums / max(self.count, 1)

    def print_result(self, os: TextIO):
        """Fourier_Autocorr::print_result (Fourier_Autocorr.h:113-119)."""
        c = self.correlation()
        for i in range(self.entries):
            os.write("%g %g %g %g\n" % (i * self.dt, c[0, i], c[1, i], c[2, i]))

    def viscosity(self, L: float, t_fit: float = 15.0) -> float:
        """nu from C(k, t) / C(k, 0) = exp(-nu k^2 t) (the curve of
        results/plot_MPC_tests.py:107): least-squares line through
        ln(C/C(0)) of the three directions' mean over 0 <= t <= t_fit."""
        c = self.correlation()
        c = (c / c[:, :1]).mean(axis=0)
        t = np.arange(self.entries) * self.dt
        sel = (t <= t_fit) & (c > 0)
        slope = np.polyfit(t[sel], np.log(c[sel]), 1)[0]
        return float(-slope / (2.0 * np.pi / L) ** 2)


class SpeedDistribution:
    """VelocityDistribution (Vel_distribution.h:17-57): histogram of |v|
    over the solvent with bins floor(|v|/width)."""

    def __init__(self, width: float = 0.5, nbins: int = 1024):
        self.width = float(width)
        self.nbins = int(nbins)
        self.counts = np.zeros(self.nbins + 1, np.uint64)

    def add(self, counts):
        self.counts += np.asarray(counts, np.uint64)

    def sample(self, system, r: int = 0):
        self.add(system.speed_histogram(self.width, self.nbins, r))

    def density(self):
        """bin keys and probability density count / (total * width) of the
        non-empty bins (Vel_distribution.h:42-55)."""
        total = float(self.counts.sum())
        b = np.nonzero(self.counts[: self.nbins])[0]
        return b * self.width, self.counts[b] / (total * self.width)

    def print_result(self, os: TextIO):
        for k, p in zip(*self.density()):
            os.write("%.8e %.8e\n" % (k, p))
        os.write("\n")


class FluidVelocityAutocorr:
    """Vel_Autocorr_Fluid (Vel_Autocorr_Fluid.h:19-116): every solvent
    particle keeps a window of `entries` velocity samples, `dt` apart; once
    the windows are full, every sample adds v(oldest) . v(oldest + i) of
    every particle to lag i (calculate(), :68-78).  The windows (particles
    followed by id, so tracking must be on) and the lag sums stay on the
    device (MPCSystem.fluid_vacf_*); this class reads them back."""

    def __init__(self, entries: int = 50, dt: float = 1.0):
        self.entries = int(entries)
        self.dt = float(dt)
        self.system = None
        self.r = 0

    def sample(self, system, r: int = 0):
        """One application of the functor to the current solvent."""
        system.fluid_vacf_sample(self.entries, r)
        self.system, self.r = system, r

    def average(self) -> np.ndarray:
        """calculate_average (:80-88): the lag sums / Counter / particles."""
        sums, counter = self.system.fluid_vacf_read(self.entries, self.r)
        with np.errstate(divide="ignore", invalid="ignore"):
            return sums / np.float64(counter) / float(self.system.num_fluid(self.r))

    def counter(self) -> int:
        return self.system.fluid_vacf_read(self.entries, self.r)[1]

    def value(self) -> float:
        """value() (:90-97) as the reference computes it: the sum of the
        averaged lags times dt, divided by Counter once more and by 3."""
        avg, n = self.average(), self.counter()
        d = 0.0
        with np.errstate(divide="ignore", invalid="ignore"):
            for a in avg:  # in order, as the reference's loop
                d += float(np.float64(a) * self.dt / np.float64(n))
        return d / 3.0

    def print_result(self, os: TextIO):
        """print_result (:99-106)."""
        avg = self.average()
        os.write("#Diffusion coefficient = %g\n" % self.value())
        for i in range(self.entries):
            os.write("%g %g\n" % (i * self.dt, avg[i]))


STAT_HEADER = ("Step   Epot     Ekin     Temp   n_p s_p  R_gyr   G_xx      G_xy      G_xz       G_yx       "
               "G_yy      G_yz      G_zx     Gzy     Gzz      wx      wy     wz\n")


def statistics_row(step: int, row) -> str:
    """One line of statistics*.dat as star-polymers.cpp:470-482 writes it
    (each value by operator<< with the default 6 significant digits and a
    trailing space)."""
    return f"{int(step)} " + "".join("%g " % float(v) for v in row) + "\n"


def eigen_row(values) -> str:
    """A row vector as Eigen's default IOFormat prints it (Eigen/src/Core/
    IO.h): every coefficient in the stream's default format (6 significant
    digits), right-aligned to the widest of them, separated by spaces."""
    t = ["%g" % float(v) for v in values]
    w = max(len(x) for x in t)
    return " ".join(x.rjust(w) for x in t)


def center_of_mass_text(system, r: int = 0) -> str:
    """Box::print_center_of_mass (Box.cpp:654-660): per molecule the mean
    position and the mean velocity of its monomers (Molecule.cpp:76-92) on
    one line."""
    pos, vel, _ = system.get_monomers(r)
    if len(pos) == 0:
        return ""
    n = len(pos)   # sums in monomer order, as the reference's loop
    return eigen_row(pos.cumsum(axis=0)[-1] / n) + " " + eigen_row(vel.cumsum(axis=0)[-1] / n) + "\n"


def print_fluid_text(system, step: int, z_start: int, z_stop: int, r: int = 0) -> str:
    """MPC::print_fluid (MPC.cpp:533-553): per z plane the cell centres,
    then for x = 0..Lx-1 one row of the x and y components of the mean
    solvent velocity of every cell (x, y, z) of the last collision."""
    cv = system.cell_fluid_velocity(z_start, z_stop, r)   # [z, y, x, 3]
    nz, ly, lx, _ = cv.shape
    out = ["TIME %d \n" % step]
    for k in range(nz):
        out.append("Z %d \n" % (z_start + k))
        out.append("".join("%f " % (x + 0.5) for x in range(lx)) + "\n")
        out.append("".join("%f " % (y + 0.5) for y in range(ly)) + "\n")
        for x in range(lx):
            out.append("".join("%f %f " % (cv[k, y, x, 0], cv[k, y, x, 1]) for y in range(ly)) + "\n")
        out.append("\n")
    out.append("\n")
    return "".join(out)


class StatisticsFile:
    """statistics*.dat (star-polymers.cpp:300-307, 470-482): the header on
    an empty file, then one row per output step."""

    def __init__(self, path: str):
        import os
        self.f = open(path, "a")
        if os.path.getsize(path) == 0:
            self.f.write(STAT_HEADER)

    def write(self, step: int, row):
        self.f.write(statistics_row(step, row))
        self.f.flush()

    def close(self):
        self.f.close()


def run_with_output(system, t0: int, n: int, mpc_dt: float, every: int,
                    on_output: Callable[[int], None], equil: int = -1):
    """The loop of star-polymers.cpp:412-536 over steps t0 .. t0+n-1: at an
    output step t (t > equil and t % every == 0) Thermostat::propagate(true),
    then on_output(t) (the observers), then MPC::step(t, mpc_dt); between
    output steps the graph-replayed MPCSystem.run."""
    t, end = t0, t0 + n
    while t < end:
        nxt = max(t, equil + 1)
        nxt = ((nxt + every - 1) // every) * every
        if nxt >= end:
            system.run(t, end - t, mpc_dt)
            return
        if nxt > t:
            system.run(t, nxt - t, mpc_dt)
        system.propagate(True)
        on_output(nxt)
        system.step(nxt, mpc_dt)
        t = nxt + 1


def mpc_test_tag(n_particles: int, shear: float, h: float, amc: bool) -> str:
    """The file-name suffix of mpc_test.cpp:79-82 (ss_para; operator<< of
    the doubles)."""
    return "_N%d_Shear%g_h%g%s" % (n_particles, shear, h, "_AMC" if amc else "")


def run_mpc_test(system, steps_equil: int, steps_total: int, steps_output: int, step_size: float,
                 outdir: str = ".", r: int = 0):
    """mpc_test.cpp:42-146 on a pure-solvent MPCSystem made with
    step_update = 1 (and the angular-momentum flag if wanted), tracking on:
    MPC::step every step; after the equilibration, at every output step the
    temperature line, VelocityX (bins 0.2) and VelocityDistribution (bins
    0.01), and every 10th step Vel_Autocorr_Fluid and Fourier_Autocorr (50
    entries, dt 1.0); then the five files fluid*, temperature*, velocity*,
    fourier*, maxwell* with the reference's names.  Returns their paths."""
    import os

    p = system.params[r]
    # BoxX*BoxY*BoxZ*N of the requested (not even-rounded) box, as :79
    n_part = int(p.L[0]) * int(p.L[1]) * int(p.L[2]) * int(p.n_c)
    tag = mpc_test_tag(n_part, p.shear, step_size, bool(p.angular_momentum))
    names = {k: os.path.join(outdir, k + tag + ".dat") for k in ("fluid", "temperature", "velocity", "fourier",
                                                                  "maxwell")}
    velocity_average = VelocityProfile(0.2)
    autocorr = FluidVelocityAutocorr(50, 1.0)
    maxwell = SpeedDistribution(0.01, 1024)
    fourier = FourierAutocorr(50, 1.0)
    with open(names["temperature"], "w") as temperature_file:
        for n in range(steps_total):
            system.step(n, step_size)
            if n > steps_equil and not n % steps_output:
                velocity_average.sample(system, r)
                maxwell.sample(system, r)
                temperature_file.write("%d %g\n" % (n, system.temperature(r)))
            if n > steps_equil and not n % 10:
                autocorr.sample(system, r)
                fourier.sample(system, r)
    for key, obs in (("fluid", velocity_average), ("maxwell", maxwell), ("velocity", autocorr),
                     ("fourier", fourier)):
        with open(names[key], "w") as f:
            obs.print_result(f)
    return names

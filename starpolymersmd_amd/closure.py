"""Learned implicit-solvent closure of the MPC solvent.

The star keeps its own velocity-Verlet MD (FENE + LJ forces, exactly as in
explicit runs); what the solvent does to it at each collision -- the
per-monomer velocity change of MPC::collide/thermostat (MPC.cpp:239-299)
with the monomers in the cells -- is replaced by a stochastic kick
v += mu(f) + diag(sigma(f)) xi drawn from a network of the monomer's 24
features f (starpolymersmd_amd/csrc/spmd_kernels.hip, k_closure_features).
The network (MLP 24-128-128-128-6, SiLU) is trained on explicit MPC-MD
runs of the library itself: mean (drift) and log-variance (noise) of the
velocity change, Gaussian negative log-likelihood, Adam, with hand-written
BF16 matrix-core forward/backward kernels (spmd_mlp.hip).

Explicit MPC stays the fallback: a monomer whose regime features (star
extent, bonds, crowding, shear rate; RANGE_CHECKED) leave the trained range
by more than 25% of it sends its replica through explicit intervals.

    from starpolymersmd_amd import closure
    data = closure.generate(shears=[0.01, 0.05], L=32, arms=10, A=5, B=5, intervals=400)
    net = closure.ClosureNet(seed=1)
    net.fit(data, steps=500, batch=65536)
    run = closure.ClosureRunner(net, data.fmin, data.fmax, shear=0.05, arms=10, A=5, B=5)
    run.advance(1000)
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from ._lib import SpmdError, load, ptr
from .system import MPCSystem, make_params

N_FEAT = 24


class DeviceBuffer:
    """A device allocation of the library (spmd_dev_*)."""

    def __init__(self, nbytes: int):
        self._lib = load()
        p = C.c_void_p()
        rc = self._lib.spmd_dev_alloc(int(nbytes), C.byref(p))
        if rc:
            raise SpmdError(rc, f"device allocation of {nbytes} bytes")
        self.p, self.nbytes = p, int(nbytes)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @property
    def ptr(self) -> int:
        return self.p.value

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        rc = self._lib.spmd_dev_upload(C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p), a.nbytes)
        if rc:
            raise SpmdError(rc, "upload")

    def download(self, shape, dtype=np.float32, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert offset + out.nbytes <= self.nbytes
        rc = self._lib.spmd_dev_download(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), out.nbytes)
        if rc:
            raise SpmdError(rc, "download")
        return out

    def close(self):
        if getattr(self, "p", None) is not None and self.p.value:
            self._lib.spmd_dev_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ClosureNet:
    """The network behind spmd_mlp_* (C ABI)."""

    def __init__(self, seed: int = 1):
        self._lib = load()
        h = C.c_void_p()
        rc = self._lib.spmd_mlp_create(seed, C.byref(h))
        if rc:
            raise SpmdError(rc, "spmd_mlp_create")
        self.h = h
        self.seed = seed
        self.n_params = int(self._lib.spmd_mlp_num_params())
        self.norm = None
        self._fit_calls = 0

    def _chk(self, rc):
        if rc:
            raise SpmdError(rc, self._lib.spmd_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self._lib.spmd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters and normalisation
    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        self._chk(self._lib.spmd_mlp_get_params(self.h, ptr(p, C.c_float)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self._chk(self._lib.spmd_mlp_set_params(self.h, ptr(p, C.c_float)))

    def set_norm(self, mu_x, sd_x, sd_y):
        mu_x, sd_x, sd_y = (np.ascontiguousarray(a, np.float32) for a in (mu_x, sd_x, sd_y))
        self._chk(self._lib.spmd_mlp_set_norm(self.h, ptr(mu_x, C.c_float), ptr(sd_x, C.c_float),
                                              ptr(sd_y, C.c_float)))
        self.norm = (mu_x, sd_x, sd_y)

    def forward(self, x: np.ndarray, fused: bool = True) -> np.ndarray:
        """(M, 24) features -> (M, 6): mean and log-variance in units of sd_y
        (fused=False: through the training path's per-layer products)."""
        x = np.ascontiguousarray(x, np.float32)
        xb = DeviceBuffer.from_numpy(x)
        ob = DeviceBuffer(x.shape[0] * 6 * 4)
        fn = self._lib.spmd_mlp_forward if fused else self._lib.spmd_mlp_forward_unfused
        self._chk(fn(self.h, C.c_void_p(xb.ptr), C.c_void_p(ob.ptr), x.shape[0]))
        return ob.download((x.shape[0], 6))

    def gradient(self, x: np.ndarray, y: np.ndarray, fused: bool = True):
        """Mean loss and its gradient (parameter order) over the rows
        (fused=False: through the per-layer products)."""
        xb = DeviceBuffer.from_numpy(np.asarray(x, np.float32))
        yb = DeviceBuffer.from_numpy(np.asarray(y, np.float32))
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double()
        fn = self._lib.spmd_mlp_gradient if fused else self._lib.spmd_mlp_gradient_unfused
        self._chk(fn(self.h, C.c_void_p(xb.ptr), C.c_void_p(yb.ptr), len(x), ptr(g, C.c_float), C.byref(loss)))
        return loss.value, g

    def fit(self, data: "ClosureData", steps: int = 1000, batch: int = 65536, lr: float = 1e-3,
            record_loss: bool = True, features=None) -> np.ndarray:
        """Adam on minibatches drawn with replacement on the device; sets the
        normalisation from the data first.  Returns the per-step loss (NaN
        where not recorded).  features: the indices the network may use;
        the others get a normalisation scale of 1e30, so they enter as zero
        in training and inference alike (FREE_DRAINING restricts it to a
        learned Langevin bath)."""
        sd_x = np.array(data.sd_x, np.float32)
        if features is not None:
            off = np.ones(N_FEAT, bool)
            off[list(features)] = False
            sd_x[off] = 1e30
        self.set_norm(data.mu_x, sd_x, data.sd_y)
        hist = np.full(steps, np.nan)
        self._fit_calls += 1
        self._chk(self._lib.spmd_mlp_fit(self.h, C.c_void_p(data.X.ptr), C.c_void_p(data.Y.ptr), data.n,
                                         min(batch, data.n), steps, lr, self.seed * 1000003 + self._fit_calls,
                                         ptr(hist, C.c_double) if record_loss else None))
        return hist

    def nll(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """Per-sample Gaussian negative log-likelihood (without the constant)
        of velocity changes y (M, 3) under the network's prediction for
        features x: sum over components of ((y/sd_y - mu)^2 exp(-s) + s) / 2,
        the distribution the closure's kick draws from."""
        o = self.forward(x).astype(np.float64)
        yn = np.asarray(y, np.float64) / self.norm[2].astype(np.float64)
        s = np.clip(o[:, 3:], -12.0, 12.0)
        return 0.5 * np.sum((yn - o[:, :3]) ** 2 * np.exp(-s) + s, axis=1)

    def save(self, path):
        mu, sd, sy = self.norm
        np.savez(path, params=self.get_params(), mu_x=mu, sd_x=sd, sd_y=sy)

    def load(self, path):
        z = np.load(path)
        self.set_params(z["params"])
        self.set_norm(z["mu_x"], z["sd_x"], z["sd_y"])


@dataclass
class ClosureData:
    X: DeviceBuffer    # (n, 24) float32 features
    Y: DeviceBuffer    # (n, 3)  float32 velocity change at the collision
    n: int
    mu_x: np.ndarray
    sd_x: np.ndarray
    sd_y: np.ndarray
    fmin: np.ndarray
    fmax: np.ndarray
    intervals: int
    seconds: float

    @classmethod
    def from_numpy(cls, X, Y, intervals=0, seconds=0.0) -> "ClosureData":
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        Xd = X.astype(np.float64)
        mu, sd = Xd.mean(0), Xd.std(0)
        sd = np.where(sd > 1e-3 * np.maximum(1.0, np.abs(mu)), sd, np.maximum(1.0, np.abs(mu)))
        return cls(X=DeviceBuffer.from_numpy(X), Y=DeviceBuffer.from_numpy(Y), n=len(X),
                   mu_x=mu.astype(np.float32), sd_x=sd.astype(np.float32),
                   sd_y=Y.astype(np.float64).std(0).astype(np.float32), fmin=X.min(0), fmax=X.max(0),
                   intervals=intervals, seconds=seconds)


def _star_system(L, n_c, shear, seed, arms, A, B, mass, step_update, md_dt):
    g = MPCSystem(make_params(L=(L, L, L), n_c=n_c, shear=shear, step_update=step_update, md_dt=md_dt, seed=seed))
    g.add_star(A, B, arms, mass, coil=True)
    g.md_init()
    return g


def preshear(system):
    """Add the steady Couette profile v_x += shear y to every replica's
    solvent.  The reference starts the solvent at rest (MPC::initialize);
    the Lees-Edwards boundaries then set up the flow by momentum diffusion,
    which takes about L^2 / (pi^2 nu) to reach the centre of the box (270
    time units, 2,700 collision intervals at L=48): until then a star at the
    centre sits in still solvent.  Starting from the profile skips that
    transient.  The solvent goes back through set_fluid, which numbers the
    particle ids afresh in storage order."""
    for r, p in enumerate(system.params):
        pos, vel = system.get_fluid(r)
        vel[:, 0] += np.float32(p.shear) * pos[:, 1]
        system.set_fluid(pos, vel, r)


def generate(shears, L=32, arms=10, A=5, B=5, mass=10.0, intervals=400, warmup=50, n_c=10, step_update=10,
             md_dt=0.01, seed=11, presheared=True) -> ClosureData:
    """Explicit MPC-MD runs, one per shear rate: after `warmup` intervals,
    each interval records every monomer's features just before the
    collision and its velocity change by the collision (intervals whose
    collision recentred the star are skipped: the recentring subtracts the
    star's COM velocity).  presheared: the solvent starts from the steady
    flow profile (preshear), so the data show the star in developed shear
    flow rather than in the start-up transient."""
    h = md_dt * step_update
    feats, targs = [], []
    t_start = time.perf_counter()
    n_int = 0
    for si, shear in enumerate(shears):
        g = _star_system(L, n_c, shear, seed + 101 * si, arms, A, B, mass, step_update, md_dt)
        g.initialize()
        if presheared:
            preshear(g)
        t = 0
        g.run(t, 1 + warmup * step_update, h)
        t += 1 + warmup * step_update
        nm = g.num_monomers()
        row = nm * N_FEAT * 4
        buf = DeviceBuffer(intervals * row)
        dv = np.empty((intervals, nm, 3))
        keep = np.ones(intervals, bool)
        for k in range(intervals):
            g.run(t, step_update - 1, h)  # the interval's substeps and MPC::step calls without collision
            t += step_update - 1
            g.propagate()
            g.closure_features(buf.ptr + k * row)
            _, v0, _ = g.get_monomers()
            g.step(t, h)                  # the collision
            t += 1
            _, v1, _ = g.get_monomers()
            dv[k] = v1 - v0
            keep[k] = not g.last_collision()[1]
            n_int += 1
        g.close()
        X = buf.download((intervals, nm, N_FEAT))
        buf.close()
        feats.append(X[keep].reshape(-1, N_FEAT))
        targs.append(dv[keep].reshape(-1, 3))
    return ClosureData.from_numpy(np.concatenate(feats), np.concatenate(targs), intervals=n_int,
                                  seconds=time.perf_counter() - t_start)


def generate_batched(shears, stars_per_rate=2, L=96, arms=30, A=20, B=20, mass=10.0, intervals=100, stride=150,
                     warmup=0, n_c=10, step_update=10, md_dt=0.01, seed=11, presheared=True,
                     progress=None) -> ClosureData:
    """As generate, with every (shear rate, star) pair a replica of one
    batched system, and one recorded interval every `stride` intervals
    after `warmup`: the data then span the star's whole history in shear
    flow -- the start-up stretching, the overshoot and the steady state --
    instead of one window of it."""
    h = md_dt * step_update
    t_start = time.perf_counter()
    reps = [(s, k) for s in shears for k in range(stars_per_rate)]
    g = MPCSystem([make_params(L=(L, L, L), n_c=n_c, shear=s, step_update=step_update, md_dt=md_dt,
                               seed=seed + 101 * i) for i, (s, k) in enumerate(reps)])
    for r in range(len(reps)):
        g.add_star(A, B, arms, mass, r=r, coil=True)
    g.md_init()
    g.initialize()
    if presheared:
        preshear(g)
    nm = g.num_monomers(0)
    R = len(reps)
    row = R * nm * N_FEAT * 4
    buf = DeviceBuffer(intervals * row)
    dv = np.empty((intervals, R, nm, 3))
    keep = np.ones((intervals, R), bool)
    t = 0
    n_int = 0
    if warmup:
        g.run(t, 1 + warmup * step_update, h)
        t += 1 + warmup * step_update
    for k in range(intervals):
        n = (stride - 1) * step_update + step_update - 1 + (1 if t == 0 else 0)
        g.run(t, n, h)
        t += n
        g.propagate()
        g.closure_features(buf.ptr + k * row)
        v0 = [g.get_monomers(r)[1] for r in range(R)]
        g.step(t, h)
        t += 1
        for r in range(R):
            dv[k, r] = g.get_monomers(r)[1] - v0[r]
            keep[k, r] = not g.last_collision(r)[1]  # a recentring subtracts the COM velocity
        n_int += stride
        if progress is not None:
            progress(k + 1, intervals)
    g.sync()
    g.close()
    X = buf.download((intervals, R, nm, N_FEAT))
    buf.close()
    return ClosureData.from_numpy(X[keep].reshape(-1, N_FEAT), dv[keep].reshape(-1, 3), intervals=n_int * R,
                                  seconds=time.perf_counter() - t_start)


def shape_stats(G):
    """Rg^2 = tr G, alignment angle chi = 1/2 atan2(2 G_xy, G_xx - G_yy) (rad)
    of a gyration tensor."""
    return float(np.trace(G)), 0.5 * math.atan2(2.0 * G[0, 1], G[0, 0] - G[1, 1])


# features whose trained range defines the closure's regime: the star's
# extent and bond geometry, crowding, shear rate, mass and connectivity.
# Velocity-like features (3-8, 19-21) have Gaussian tails and the type
# flags are categorical; they are not range-checked.
RANGE_CHECKED = (0, 1, 2, 9, 10, 11, 12, 13, 14, 15, 18, 22, 23)

# the velocity relative to the imposed flow and the shear rate: a network
# restricted to these is a learned free-draining Langevin bath (drag toward
# the local flow plus noise, no dependence on the star's configuration)
FREE_DRAINING = (6, 7, 8, 18)
# the free-draining inputs plus the crowding count, the bead type and the
# mean velocity of the monomers within 1.5 relative to the monomer's own:
# drag toward the local flow and a learned coupling to the neighbours
LOCAL_COUPLING = (6, 7, 8, 15, 16, 18, 19, 20, 21)
# the network the library runs by default: local coupling.  It predicts
# held-out collisions better than the free-draining bath, and the
# configuration inputs of the full 24-feature network (position relative to
# the centre, bond vectors) buy no further likelihood and bias the
# alignment angle of sheared stars.
HEADLINE = LOCAL_COUPLING


def regime_range(fmin, fmax, checked=RANGE_CHECKED):
    """fmin/fmax for spmd_closure_run with the unchecked features open."""
    lo = np.full(N_FEAT, -np.inf, np.float32)
    hi = np.full(N_FEAT, np.inf, np.float32)
    for c in checked:
        lo[c], hi[c] = fmin[c], fmax[c]
    return lo, hi


class ClosureRunner:
    """Replicas of one star advanced by MD + the learned solvent kick, with
    explicit MPC fallback for replicas whose features leave the trained
    range, and explicit MPC as the measure of its error: every check_every-th
    interval, check_replicas replicas (in turn) are cross-checked by an
    explicit MPC run at L, rho=n_c_fallback from the same star state beside
    a closure run of the same length (the main run is not touched); the
    Rg^2, alignment angle and rotation frequency of both are recorded
    (cross_check_summary)."""

    def __init__(self, net: ClosureNet, fmin, fmax, shear, arms, A, B, mass=10.0, replicas=1, L=48,
                 step_update=10, md_dt=0.01, seed=5, shears=None, fallback_intervals=20, n_c_fallback=10,
                 box=128, check_every=0, check_replicas=8, check_intervals=20):
        """L: box of the explicit fallback and cross-check runs (larger when
        the star reaches further); box: the closure systems' box, which
        without solvent only frames the Verlet grid (it must hold the
        stretched star: monomers within box/2 of its centre)."""
        self.net = net
        self.fmin, self.fmax = regime_range(np.asarray(fmin, np.float32), np.asarray(fmax, np.float32))
        self.shears = list(shears) if shears is not None else [shear] * replicas
        self.L, self.S, self.dt = L, step_update, md_dt
        self.star = (arms, A, B, mass)
        self.seed = seed
        self.sys = MPCSystem([make_params(L=(box, box, box), n_c=0, shear=s, step_update=step_update, md_dt=md_dt,
                                          seed=seed + r) for r, s in enumerate(self.shears)])
        for r in range(len(self.shears)):
            self.sys.add_star(A, B, arms, mass, r=r, coil=True)
        self.sys.md_init()
        self.fallback_intervals = fallback_intervals
        self.n_c_fallback = n_c_fallback
        self.fallbacks = 0
        # explicit fallback systems kept for reuse, by (box, shear rate):
        # their solvent stays in developed flow between fallbacks
        self._pool = {}
        self.pool_size = 4
        self.t = 0
        self.box = box
        self.check_every, self.check_replicas, self.check_intervals = check_every, check_replicas, check_intervals
        self.checks = []      # per cross-checked replica: (t, replica, shear, closure stats, explicit stats)
        self._check_next = 0  # the replica the next cross-check starts from

    def advance(self, n_intervals: int, chunk: int = 10):
        done = 0
        while done < n_intervals:
            k = min(chunk, n_intervals - done)
            if self.check_every:
                k = min(k, self.check_every - self.t % self.check_every)
            oor = self.sys.closure_run(self.net, k, self.fmin, self.fmax)
            done += k
            self.t += k
            for r in np.nonzero(oor)[0]:
                self._explicit(int(r))
            if self.check_every and self.t % self.check_every == 0:
                self._cross_check()

    def _star_states(self, reps):
        """each replica's star centred, velocities relative to the local flow"""
        out = []
        for r in reps:
            pos, vel, _ = self.sys.get_monomers(r)
            com = pos.mean(axis=0)
            vel = vel.copy()
            vel[:, 0] -= self.shears[r] * com[1]
            out.append((pos - com, vel))
        return out

    def _cross_check(self):
        """Explicit MPC (L, rho) and the closure from the same star states of
        check_replicas replicas over check_intervals intervals: per replica
        the time means of Rg^2, alignment angle and w_z on each side."""
        arms, A, B, mass = self.star
        R = len(self.shears)
        m = min(self.check_replicas, R)
        reps = [(self._check_next + i) % R for i in range(m)]
        self._check_next = (self._check_next + m) % R
        states = self._star_states(reps)
        reach = max(float(np.abs(p).max()) for p, _ in states)
        L = max(self.L, 2 * int(math.ceil(reach + 4.0)))
        h = self.S * self.dt
        seed = self.seed + 7919 * (1 + len(self.checks))
        ex = MPCSystem([make_params(L=(L,) * 3, n_c=self.n_c_fallback, shear=self.shears[r], step_update=self.S,
                                    md_dt=self.dt, seed=seed + i) for i, r in enumerate(reps)])
        cl = MPCSystem([make_params(L=(self.box,) * 3, n_c=0, shear=self.shears[r], step_update=self.S,
                                    md_dt=self.dt, seed=seed + 100 + i) for i, r in enumerate(reps)])
        try:
            for i, (p, v) in enumerate(states):
                for g in (ex, cl):
                    g.add_star(A, B, arms, mass, r=i, coil=True)
                    g.set_monomers(p, v, r=i)
            ex.md_init()
            cl.md_init()
            ex.initialize()
            preshear(ex)  # the flow the closure stands for
            acc_e, acc_c = [], []
            t = 0
            for _ in range(self.check_intervals):
                n = self.S + (1 if t == 0 else 0)
                ex.run(t, n, h)
                t += n
                cl.closure_run(self.net, 1, self.fmin, self.fmax)
                acc_e.append([(*shape_stats(ex.gyration_tensor(i)), ex.rotation_frequency(i)[2]) for i in range(m)])
                acc_c.append([(*shape_stats(cl.gyration_tensor(i)), cl.rotation_frequency(i)[2]) for i in range(m)])
            ex.sync()
            me, mc = np.array(acc_e).mean(axis=0), np.array(acc_c).mean(axis=0)
            for i, r in enumerate(reps):
                self.checks.append((self.t, r, float(self.shears[r]), mc[i], me[i]))
        finally:
            ex.close()
            cl.close()

    def cross_check_summary(self):
        """Over all cross-checks: closure / explicit ratios of the mean Rg^2,
        Rg, alignment angle and w_z with standard errors (from the spread of
        the per-replica differences), and the number of checked replicas."""
        if not self.checks:
            return None
        c = np.array([x[3] for x in self.checks])
        e = np.array([x[4] for x in self.checks])
        n = len(c)
        out = {"checks": n, "L": self.L}
        for k, name in enumerate(("Rg2", "chi", "wz")):
            mc, me = c[:, k].mean(), e[:, k].mean()
            d = c[:, k] - e[:, k]
            se = d.std(ddof=1) / np.sqrt(n) if n > 1 else float("nan")
            out[name + "_closure"], out[name + "_explicit"] = float(mc), float(me)
            out[name + "_ratio"] = float(mc / me) if me != 0 else float("nan")
            out[name + "_ratio_se"] = float(se / abs(me)) if me != 0 else float("nan")
        out["Rg_ratio"] = float(np.sqrt(out["Rg2_ratio"]))
        out["Rg_ratio_se"] = 0.5 * out["Rg2_ratio_se"] / out["Rg_ratio"]
        return out

    def _explicit(self, r):
        """Run the replica's star through explicit MPC intervals, then resume."""
        arms, A, B, mass = self.star
        pos, vel, _ = self.sys.get_monomers(r)
        com = pos.mean(axis=0)
        pos = pos - com
        vel = vel.copy()
        vel[:, 0] -= self.shears[r] * com[1]  # same velocities relative to the local flow
        # a box that holds the star with room to move: monomers must stay
        # inside the Verlet grid (|x| < L/2 - 1/2, MPC::check_bounds)
        reach = float(np.abs(pos).max()) if len(pos) else 0.0
        L = max(self.L, 2 * int(math.ceil(reach + 4.0)))
        h = self.S * self.dt
        key = (L, float(self.shears[r]))
        entry = self._pool.pop(key, None)
        if entry is None:
            g = MPCSystem(make_params(L=(L,) * 3, n_c=self.n_c_fallback, shear=self.shears[r], step_update=self.S,
                                      md_dt=self.dt, seed=self.seed + 1000 + self.fallbacks))
            g.add_star(A, B, arms, mass, coil=True)
            g.set_monomers(pos, vel)
            g.md_init()
            g.initialize()
            preshear(g)  # the flow the closure stands for
            g.run(0, 1 + self.fallback_intervals * self.S, h)
            t = 1 + self.fallback_intervals * self.S
        else:
            # a solvent already in steady flow: put this star in its centre
            g, t = entry
            g.set_monomers(pos, vel)
            g.md_init()
            g.run(t, self.fallback_intervals * self.S, h)
            t += self.fallback_intervals * self.S
        p2, v2, _ = g.get_monomers()
        self._pool[key] = (g, t)  # most recently used last
        while len(self._pool) > self.pool_size:
            old = next(iter(self._pool))
            self._pool.pop(old)[0].close()
        self.sys.set_monomers(p2, v2, r=r)
        self.sys.md_init()
        self.fallbacks += 1

    def gyration(self, r=0):
        return self.sys.gyration_tensor(r)

    def close(self):
        """Release the pooled fallback systems."""
        for g, _ in self._pool.values():
            g.close()
        self._pool.clear()

"""The learned closure on the GPU: the BF16 matrix-core network against
the FP64 oracle (forward within BF16 tolerance, gradient direction and
size), Adam training on a known regression, the feature kernel against the
oracle's definition, the star observables against the oracle, and the
closure end to end against explicit MPC and against the star without it.

Tolerances: BF16 operands carry 8 significant bits (relative 2^-9 per
rounding), accumulations are FP32; outputs of the 4-layer network are
compared at 3e-2 of their scale, gradients by cosine > 0.995 and norm
within 5%."""
import os
import sys

import numpy as np
import pytest

import oracle_binding as ob
from starpolymersmd_amd import MPCSystem, make_params

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp as omlp  # noqa: E402

pytestmark = pytest.mark.gpu


def _net(seed=1):
    from starpolymersmd_amd.closure import ClosureNet

    return ClosureNet(seed)


@pytest.mark.parametrize("fused", [True, False])
def test_forward_matches_oracle(fused):
    net = _net()
    p = (omlp.init_params(7) * 0.7).astype(np.float32)
    net.set_params(p)
    mu = np.linspace(-1, 1, 24).astype(np.float32)
    sd = np.linspace(0.5, 2.0, 24).astype(np.float32)
    net.set_norm(mu, sd, np.ones(3, np.float32))
    rng = np.random.default_rng(0)
    x = (rng.normal(size=(1000, 24)) * sd + mu).astype(np.float32)
    out = net.forward(x, fused=fused)
    ref, _ = omlp.forward(p.astype(np.float64), x.astype(np.float64), mu, sd)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(out - ref) <= 3e-2 * scale + 1e-3), np.abs(out - ref).max(axis=0) / scale


@pytest.mark.parametrize("M", [1, 17, 63, 64, 65, 1000, 307456])
def test_fused_forward_matches_unfused(M):
    """The fused pass reorders each layer's inputs inside the matrix-core
    products; both paths round the same FP32 sums to BF16 between layers,
    so they agree to a few BF16 units on the outputs.  Ragged row counts
    cover the partial last tile; 307,456 rows (256 stars of 1,201 monomers)
    cover the wave-strided loop over the grid."""
    net = _net()
    p = (omlp.init_params(11) * 0.7).astype(np.float32)
    net.set_params(p)
    mu = np.linspace(-0.5, 0.5, 24).astype(np.float32)
    sd = np.linspace(0.5, 1.5, 24).astype(np.float32)
    net.set_norm(mu, sd, np.ones(3, np.float32))
    rng = np.random.default_rng(M)
    x = (rng.normal(size=(M, 24)) * sd + mu).astype(np.float32)
    a = net.forward(x, fused=True)
    b = net.forward(x, fused=False)
    assert np.all(np.isfinite(a))
    scale = np.abs(b).max(axis=0) + 1e-6
    err = np.abs(a - b) / scale
    assert err.max() <= 2e-2, err.max(axis=0)
    assert np.median(err) <= 1e-3


def test_gradient_matches_oracle():
    net = _net()
    p = (omlp.init_params(9) * 0.7).astype(np.float32)
    net.set_params(p)
    net.set_norm(np.zeros(24, np.float32), np.ones(24, np.float32), np.array([1.0, 2.0, 0.5], np.float32))
    rng = np.random.default_rng(4)
    M = 3000
    x = rng.normal(size=(M, 24)).astype(np.float32)
    y = (rng.normal(size=(M, 3)) * [1.0, 2.0, 0.5]).astype(np.float32)
    loss, g = net.gradient(x, y)
    Lr, gr = omlp.loss_and_grad(p.astype(np.float64), x, y, [1.0, 2.0, 0.5])
    assert abs(loss - Lr) <= 2e-2 * abs(Lr)
    for a, b in ((0, 3200), (3200, 19712), (19712, 36224), (36224, 36998)):  # per layer
        ga, gb = g[a:b].astype(np.float64), gr[a:b]
        cos = ga @ gb / (np.linalg.norm(ga) * np.linalg.norm(gb))
        assert cos > 0.995, (a, cos)
        assert abs(np.linalg.norm(ga) / np.linalg.norm(gb) - 1) < 0.05, (a, np.linalg.norm(ga), np.linalg.norm(gb))


@pytest.mark.parametrize("M", [1, 63, 64, 65, 3000, 65536])
def test_fused_gradient_matches_unfused(M):
    """The fused training pass (forward, loss, backward and weight-gradient
    products in one kernel per 64-row chunk) against the per-layer products:
    the same BF16 roundings of activations and gradients up to the order of
    FP32 sums, so per layer the gradients agree to cosine 0.9999 and norm 1%.
    Ragged row counts cover the partial chunk (padding rows contribute
    nothing); 65,536 rows take four chunks per workgroup."""
    net = _net()
    p = (omlp.init_params(13) * 0.7).astype(np.float32)
    net.set_params(p)
    net.set_norm(np.linspace(-0.3, 0.3, 24).astype(np.float32), np.linspace(0.7, 1.3, 24).astype(np.float32),
                 np.array([1.0, 2.0, 0.5], np.float32))
    rng = np.random.default_rng(M)
    x = rng.normal(size=(M, 24)).astype(np.float32)
    y = (rng.normal(size=(M, 3)) * [1.0, 2.0, 0.5]).astype(np.float32)
    lf, gf = net.gradient(x, y)
    lu, gu = net.gradient(x, y, fused=False)
    assert np.all(np.isfinite(gf))
    assert abs(lf - lu) <= 1e-3 * abs(lu) + 1e-6, (lf, lu)
    for a, b in ((0, 3200), (3200, 19712), (19712, 36224), (36224, 36998)):
        ga, gb = gf[a:b].astype(np.float64), gu[a:b].astype(np.float64)
        cos = ga @ gb / (np.linalg.norm(ga) * np.linalg.norm(gb))
        assert cos > 0.9999, (M, a, cos)
        assert abs(np.linalg.norm(ga) / np.linalg.norm(gb) - 1) < 0.01, (M, a)


def test_train_step_keeps_weight_copies_current():
    """The Adam step updates the FP32 parameters and, in place, the BF16
    copies the fused kernels read: the forward pass after training equals,
    bit for bit, the forward pass after the same parameters are set afresh
    (which packs every copy from the FP32 parameters)."""
    from starpolymersmd_amd.closure import ClosureData

    net = _net(5)
    rng = np.random.default_rng(8)
    x = rng.normal(size=(20000, 24)).astype(np.float32)
    y = (x[:, :3] * 0.5 + 0.1 * rng.normal(size=(20000, 3))).astype(np.float32)
    hist = net.fit(ClosureData.from_numpy(x, y), steps=7, batch=4096, lr=1e-3)
    assert np.all(np.isfinite(hist)) and hist[-1] < hist[0]
    out_trained = net.forward(x[:3000])
    net.set_params(net.get_params())
    assert np.array_equal(out_trained, net.forward(x[:3000]))


def test_fit_draws_the_philox_rows():
    """spmd_mlp_fit draws row b of step t's minibatch inside the fused
    kernel as Philox4x32-10(counter (b, t, 7, 0), key = the fit's seed) mod
    N.  The first step's recorded loss comes from the parameters before any
    update, so it equals the FP64 oracle's loss on exactly those rows (BF16
    tolerance); the rows of another step, from these data whose per-row
    scale spreads 30-fold, give a loss farther away than that tolerance."""
    from starpolymersmd_amd.closure import ClosureData, ClosureNet

    rng = np.random.default_rng(21)
    N, B = 5000, 256
    x = rng.normal(size=(N, 24)).astype(np.float32)
    y = (rng.normal(size=(N, 3)) * np.linspace(0.1, 3.0, N)[:, None]).astype(np.float32)
    data = ClosureData.from_numpy(x, y)
    net = ClosureNet(4)
    p = net.get_params().astype(np.float64)
    hist = net.fit(data, steps=1, batch=B, lr=1e-3)
    seed = 4 * 1000003 + 1  # ClosureNet.fit: seed * 1000003 + its call count

    def rows(step):
        out = []
        for b in range(B):
            w = ob.philox([b, step, 7, 0], [seed & 0xFFFFFFFF, seed >> 32])
            out.append(((int(w[0]) << 32) | int(w[1])) % N)
        return np.array(out)

    def loss(r):
        return omlp.loss_and_grad(p, x[r], y[r], data.sd_y, data.mu_x, data.sd_x)[0]

    l0 = loss(rows(0))
    assert abs(hist[0] - l0) <= 2e-2 * abs(l0), (hist[0], l0)
    assert abs(loss(rows(1)) - l0) > 4e-2 * abs(l0)


def test_adam_learns_a_known_map():
    """y = A x (3 of the 24 inputs), noise sd 0.1: the network must reach
    the noise floor's log-likelihood region (mean NLL per component of
    N(0, 0.1^2) in units of sd_y)."""
    from starpolymersmd_amd.closure import ClosureData, ClosureNet

    rng = np.random.default_rng(5)
    M = 65536
    x = rng.normal(size=(M, 24)).astype(np.float32)
    Amat = rng.normal(size=(3, 24)) * 0.3
    y = (x @ Amat.T + 0.1 * rng.normal(size=(M, 3))).astype(np.float32)
    data = ClosureData.from_numpy(x, y)
    sy = data.sd_y
    net = ClosureNet(3)
    hist = net.fit(data, steps=400, batch=8192, lr=3e-3)
    first, last = hist[0], hist[-10:].mean()
    floor = np.sum(0.5 * (1.0 + np.log((0.1 / sy) ** 2)))  # NLL of the exact noise model, summed over components
    assert last < first - 2.0, hist
    assert last < floor + 2.0, (last, floor, hist)


@pytest.mark.parametrize("L,A,B,arms", [(20, 4, 4, 6), (48, 20, 20, 30)])
def test_features_match_oracle(L, A, B, arms):
    """The neighbour features come from a spatial hash on the device and from
    all pairs in the oracle; the second case is the 1,201-monomer star the
    closure is benchmarked on."""
    g = MPCSystem(make_params(L=(L,) * 3, n_c=10, shear=0.03, step_update=10, seed=8, md_dt=0.01))
    g.add_star(A, B, arms, 10.0, coil=True)
    g.md_init()
    g.initialize()
    g.run(0, 31, 0.1)
    g.sync()
    o = ob.Oracle(L=(L,) * 3, n_c=10, shear=0.03)
    o.add_star(A, B, arms, 10.0)
    bonds = [(int(min(a, b)), int(max(a, b))) for a, b in o.bonds()]
    mass, btype, _ = o.monomer_props()
    pos, vel, _ = g.get_monomers()
    nm = g.num_monomers()
    from starpolymersmd_amd.closure import DeviceBuffer

    buf = DeviceBuffer(nm * 24 * 4)
    g.closure_features(buf.ptr)
    f = buf.download((nm, 24))
    ref = omlp.features(pos, vel, mass, btype, bonds, 0.03)
    cols = [c for c in range(24) if c != 15]
    assert np.allclose(f[:, cols], ref[:, cols], rtol=1e-5, atol=1e-5), np.abs(f[:, cols] - ref[:, cols]).max(axis=0)
    assert np.abs(f[:, 15] - ref[:, 15]).max() <= 1  # a pair at exactly r = 1.5 may round either way


def test_star_observables_match_oracle():
    g = MPCSystem(make_params(L=(20, 20, 20), n_c=10, shear=0.02, step_update=10, seed=3))
    g.add_star(5, 5, 10, 10.0, coil=True)
    g.md_init()
    g.initialize()
    g.run(0, 21, 0.1)
    pos, vel, _ = g.get_monomers()
    o = ob.Oracle(L=(20, 20, 20), n_c=10, shear=0.02)
    o.add_star(5, 5, 10, 10.0)
    o.set_monomers(pos, vel)
    _, Go = o.gyration()
    assert np.allclose(g.gyration_tensor(), Go, rtol=1e-12, atol=1e-12)
    assert np.allclose(g.rotation_frequency(), o.rotation_frequency(), rtol=1e-9, atol=1e-12)


def test_closure_end_to_end():
    """Train on explicit MPC-MD of a small star under shear; on collisions
    of an independent run the network (and its free-draining restriction)
    predicts the velocity change far better than the best features-blind
    Gaussian.  Then advance the star with the closure instead of the
    solvent: its kinetic temperature follows explicit MPC and its size
    stays close; without the closure (the star's MD alone) the star heats
    up under its own relaxation and never exchanges momentum with a bath."""
    from starpolymersmd_amd import closure

    A, B, arms, shear = 3, 3, 6, 0.05
    data = closure.generate([0.02, 0.05, 0.1], L=20, arms=arms, A=A, B=B, intervals=1000, warmup=20)
    assert data.n > 100000
    # 300 steps of 16,384 (33 passes over the data): longer training on
    # this little data (3 runs of a 49-monomer star) memorises it and
    # predicts held-out collisions worse than the features-blind Gaussian
    net = closure.ClosureNet(2)
    hist = net.fit(data, steps=300, batch=16384, lr=2e-3)
    assert hist[-10:].mean() < hist[0] - 0.4, hist[::30]
    # held-out collisions (an independent run): the network's likelihood
    # against the best features-blind Gaussian of the training data (its
    # mean and variance per component), both in units of sd_y
    held = closure.generate([0.05], L=20, arms=arms, A=A, B=B, intervals=100, warmup=20, seed=77)
    xh, yh = held.X.download((held.n, 24)), held.Y.download((held.n, 3))
    yt = data.Y.download((data.n, 3)).astype(np.float64) / net.norm[2]
    m0, v0 = yt.mean(axis=0), yt.var(axis=0)
    yn = yh.astype(np.float64) / net.norm[2]
    nll_blind = float(np.mean(0.5 * np.sum((yn - m0) ** 2 / v0 + np.log(v0), axis=1)))
    nll_net = float(net.nll(xh, yh).mean())
    fd = closure.ClosureNet(2)
    fd.fit(data, steps=300, batch=16384, lr=2e-3, features=closure.FREE_DRAINING)
    nll_fd = float(fd.nll(xh, yh).mean())
    print("\nheld-out NLL: network", nll_net, "free-draining network", nll_fd, "features-blind Gaussian", nll_blind)
    assert nll_net < nll_blind - 0.3, (nll_net, nll_blind)
    assert nll_fd < nll_blind - 0.3, (nll_fd, nll_blind)

    def stats(advance, sys_, n=300, every=10, skip=50):
        rg, chi, ek = [], [], []
        for _ in range(n):
            advance(every)
            r2, c = closure.shape_stats(sys_.gyration_tensor())
            rg.append(r2)
            chi.append(c)
            _, v, _ = sys_.get_monomers()
            ek.append(0.5 * 10.0 * np.mean(np.sum((v - v.mean(0)) ** 2, axis=1)) / 1.5)
        return tuple(float(np.mean(a[skip:])) for a in (rg, chi, ek))

    g = MPCSystem(make_params(L=(20, 20, 20), n_c=10, shear=shear, step_update=10, seed=21, md_dt=0.01))
    g.add_star(A, B, arms, 10.0, coil=True)
    g.md_init()
    g.initialize()
    t = [0]

    def adv_explicit(k):
        g.run(t[0], k * 10, 0.1)
        t[0] += k * 10

    ex = stats(adv_explicit, g)
    # without solvent the box only frames the Verlet grid: a larger one keeps
    # a hot, stretched star inside it
    run = closure.ClosureRunner(net, data.fmin, data.fmax, shear=shear, arms=arms, A=A, B=B, L=20, seed=21,
                                box=24)
    cl = stats(lambda k: run.advance(k), run.sys)
    m = MPCSystem(make_params(L=(24, 24, 24), n_c=0, shear=shear, step_update=10, seed=21, md_dt=0.01))
    m.add_star(A, B, arms, 10.0, coil=True)
    m.md_init()

    def adv_md(k):
        for _ in range(k * 10):
            m.propagate()
        p, v, _ = m.get_monomers()  # keep the drifting star inside its Verlet grid
        m.set_monomers(p - p.mean(axis=0), v)
        m.md_init()

    nc = stats(adv_md, m, n=60, skip=10)  # unthermostatted: a short window shows it
    print("\n(Rg^2, chi, T): explicit", ex, "closure", cl, "md only", nc, "fallbacks", run.fallbacks)
    assert run.fallbacks <= 30, run.fallbacks   # explicit fallback only now and then (300 chunks)
    assert abs(cl[2] - ex[2]) < 0.2, (cl, ex)   # kinetic temperature of the bath
    assert abs(nc[2] - ex[2]) > 3 * abs(cl[2] - ex[2]), (nc, cl, ex)
    # size: the per-collision closure (no solvent memory between collisions)
    # keeps the star within 40% of explicit MPC's Rg^2 under shear
    assert abs(cl[0] / ex[0] - 1.0) < 0.4, (cl, ex)


def test_preshear_sets_the_couette_profile():
    """closure.preshear adds v_x = shear y to the solvent: the VelocityX
    profile right after it has the imposed slope."""
    from starpolymersmd_amd import closure

    g = MPCSystem([make_params(L=(16, 16, 16), n_c=10, shear=s, seed=3 + r) for r, s in enumerate((0.01, 0.03))])
    g.initialize()
    closure.preshear(g)
    for r, s in enumerate((0.01, 0.03)):
        p, v = g.get_fluid(r)
        slope = np.polyfit(p[:, 1].astype(np.float64), v[:, 0].astype(np.float64), 1)[0]
        assert abs(slope - s) < 0.003, (r, slope)


def test_closure_tracks_explicit_in_shear_flow():
    """The 1,201-monomer star (f=30, 20A+20B) in shear flow at 0.012: the
    headline network (closure.HEADLINE inputs) trained on explicit runs
    started from the flow profile (8 rates in [4e-3, 1.6e-2], 2 stars each,
    L=80); 64 replicas each way from coils (explicit at L=80, whose Verlet
    grid holds the stretched stars), 400 intervals of warm-up, then Rg^2,
    the alignment angle and the rotation frequency about z averaged over
    800 intervals.  Bar: the closure's Rg and alignment angle within 3% of
    explicit MPC's, each widened by two combined standard errors (replica
    spread), and w_z within 5% plus two standard errors; the explicit
    cross-check of ClosureRunner (every 100th interval, L=48, rho=10)
    agrees on Rg within 1%."""
    from starpolymersmd_amd import closure as cl

    arms, A, B, S, shear, R = 30, 20, 20, 10, 0.012, 64
    parts = [cl.generate(list(np.geomspace(4e-3, 1.6e-2, 8)), L=80, arms=arms, A=A, B=B, intervals=150,
                         warmup=800, seed=11 + 7 * k) for k in range(2)]
    xs = [q.X.download((q.n, cl.N_FEAT)) for q in parts]
    ys = [q.Y.download((q.n, 3)) for q in parts]
    data = cl.ClosureData.from_numpy(np.concatenate(xs), np.concatenate(ys))
    net = cl.ClosureNet(1)
    net.fit(data, steps=2000, batch=65536, lr=2e-3, features=cl.HEADLINE)

    def ensemble(advance, sysm):
        advance(400)
        acc = []
        for _ in range(80):
            advance(10)
            acc.append([(*cl.shape_stats(sysm.gyration_tensor(r)), sysm.rotation_frequency(r)[2]) for r in range(R)])
        pm = np.array(acc).mean(axis=0)
        return pm.mean(axis=0), pm.std(axis=0, ddof=1) / np.sqrt(R)

    e = MPCSystem([make_params(L=(80,) * 3, n_c=10, shear=shear, step_update=S, seed=100 + r) for r in range(R)])
    for r in range(R):
        e.add_star(A, B, arms, 10.0, r=r, coil=True)
    e.md_init()
    e.initialize()
    cl.preshear(e)
    t = [0]

    def adv_e(k):
        n = k * S + (1 if t[0] == 0 else 0)
        e.run(t[0], n, 0.1)
        t[0] += n

    em, es = ensemble(adv_e, e)
    e.sync()  # raises if a monomer left the Verlet grid
    e.close()
    c = cl.ClosureRunner(net, data.fmin, data.fmax, shear=None, arms=arms, A=A, B=B, L=64, shears=[shear] * R,
                         seed=500)
    cm, cs = ensemble(lambda k: c.advance(k, chunk=k), c.sys)
    c.sys.sync()
    c.close()
    ratio = cm / em
    rse = np.abs(ratio) * np.sqrt((cs / cm) ** 2 + (es / em) ** 2)
    rg, rg_se = np.sqrt(ratio[0]), 0.5 * rse[0] / np.sqrt(ratio[0])
    print(f"\n(Rg^2, chi, w_z) explicit {em} +- {es} closure {cm} +- {cs}; Rg ratio {rg:.4f} +- {rg_se:.4f}, "
          f"chi ratio {ratio[1]:.4f} +- {rse[1]:.4f}, w_z ratio {ratio[2]:.4f} +- {rse[2]:.4f}")
    assert abs(rg - 1.0) <= 0.03 + 2 * rg_se, (rg, rg_se)
    assert abs(ratio[1] - 1.0) <= 0.03 + 2 * rse[1], (ratio[1], rse[1])
    assert abs(ratio[2] - 1.0) <= 0.05 + 2 * rse[2], (ratio[2], rse[2])

    # the explicit cross-check inside the run: every 100th interval, 8 replicas
    x = cl.ClosureRunner(net, data.fmin, data.fmax, shear=None, arms=arms, A=A, B=B, L=48, shears=[shear] * 16,
                         seed=700, check_every=100, check_replicas=8, check_intervals=20)
    x.advance(200, chunk=50)
    summ = x.cross_check_summary()
    x.close()
    print("cross-check", summ)
    assert summ["checks"] == 16
    assert abs(summ["Rg_ratio"] - 1.0) <= 0.01 + 2 * summ["Rg_ratio_se"], summ


@pytest.mark.gpu
def test_neighbour_coupling_beats_free_draining_bath_on_heldout_collisions():
    """What the network's neighbour inputs gain over the method without
    them: trained on the same explicit runs of the 1,201-monomer star (8
    rates, 4 independent stars each), the network on the flow-relative
    velocity, shear rate, crowding, type and the neighbours' mean velocity
    (closure.LOCAL_COUPLING) predicts the velocity changes of collisions in
    an independent run better than the same network on the flow-relative
    velocity and shear rate alone (closure.FREE_DRAINING, a learned
    free-draining Langevin bath): lower negative log-likelihood per sample,
    by more than 3 standard errors of interval blocks."""
    from starpolymersmd_amd import closure as cl

    arms, A, B = 30, 20, 20
    shears = list(np.geomspace(4e-3, 1.6e-2, 8))

    def dl(d):
        x, y = d.X.download((d.n, 24)), d.Y.download((d.n, 3))
        d.X.close()
        d.Y.close()
        return x, y

    parts = [dl(cl.generate(shears, L=64, arms=arms, A=A, B=B, intervals=150, warmup=800))]
    parts += [dl(cl.generate(shears, L=80, arms=arms, A=A, B=B, intervals=150, warmup=800, seed=11 + 7 * k))
              for k in (1, 2, 3)]
    data = cl.ClosureData.from_numpy(np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]))
    xh, yh = dl(cl.generate([0.012], L=80, arms=arms, A=A, B=B, intervals=50, warmup=800, seed=99))
    nll = {}
    for key, feats in (("fd", cl.FREE_DRAINING), ("local", cl.LOCAL_COUPLING)):
        net = cl.ClosureNet(1)
        net.fit(data, steps=2000, batch=65536, lr=2e-3, features=feats)
        nll[key] = net.nll(xh, yh)
    g = (nll["fd"] - nll["local"]).reshape(-1, 1 + arms * (A + B)).mean(axis=1)
    gain, se = g.mean(), g.std(ddof=1) / np.sqrt(len(g))
    print(f"\nheld-out NLL free-draining {nll['fd'].mean():.4f} local coupling {nll['local'].mean():.4f} "
          f"gain {gain:.5f} +- {se:.5f}")
    assert gain > 3 * se and gain > 0, (gain, se)

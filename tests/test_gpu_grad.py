"""The device gradient tape against the FP64 oracle adjoint and FP64 finite differences.

Bars:
  * dL/dtheta within 1e-3 of the FP64 oracle adjoint in the vector 2-norm, and per
    component within 1e-3 relative -- or, for the few ill-conditioned components (d/dsigma
    carries 1/sigma^3 ~ 1e6 over sums that cancel), within 3x the deviation of the oracle's
    own f32 adjoint from its FP64 adjoint, i.e. as accurate as f32 arithmetic allows;
  * dL/dtheta within 1e-3 relative of FP64 central differences of the oracle forward, and for
    the L2-to-a-perturbed-target loss, 64 sampled coordinates at eps = 1e-4;
  * dL/dA_0 within 1e-3 of the FP64 oracle's, relative to its largest entry;
  * the checkpoint interval and batching change nothing, bit for bit;
  * the derivative of total mass (conserved by reintegration) is 1 for every cell and ~0
    for every parameter at the full differentiation size (32 members of 512^2 x 2 channels).
"""
import numpy as np
import pytest

import autoverse_amd as av
from autoverse_amd import grad as G
from autoverse_amd import workloads
from autoverse_amd.population import HipPopulation

pytestmark = pytest.mark.gpu

REL = 1e-3


def _say(*a):
    print(*a, flush=True)


def _case(O, size=32, channels=2, seed=3, R=6, warm=3):
    cfg, sd = workloads.multispecies(size=size, channels=channels, seed=seed)
    cfg.kernel_radius = R
    st = O.seed_generate(sd, size, size, 1, channels)
    p = O.Propagator(cfg, np.float32)
    st, _, _ = p.run(st, warm)
    return cfg, [np.asarray(c, np.float32).reshape(-1) for c in st]


def _seed(n, channels, s=0):
    rng = np.random.default_rng(s)
    return [rng.standard_normal(n).astype(np.float32) for _ in range(channels)]


def _gpu_grad(cfg, st, steps, seed, k=2):
    state = av.SimulationState(st, cfg.width, cfg.height)
    return G.grad(cfg, state, steps, seed, checkpoint_every=k)


def _close(got, want, want32, what):
    """got (device f32) against want (FP64 oracle); want32 is the oracle's f32 result."""
    scale = np.abs(want).max()
    err = np.abs(got - want)
    f32_floor = 3.0 * np.abs(want32 - want)
    bad = err > np.maximum(REL * np.abs(want), f32_floor) + 0.1 * REL * scale
    assert not bad.any(), f"{what}: worst {err.max():.3e} vs scale {scale:.3e} at {np.nonzero(bad)[0][:8]}"
    vec = np.linalg.norm(got - want) / np.linalg.norm(want)
    _say(f"  {what}: |g - g64| / |g64| = {vec:.2e} (oracle f32: "
         f"{np.linalg.norm(want32 - want) / np.linalg.norm(want):.2e})")
    assert vec <= REL


@pytest.mark.parametrize("case", [("32x2", 32, 2, 6), ("48x1", 48, 1, 8), ("64x3", 64, 3, 9)])
def test_grad_matches_fp64_oracle_adjoint(oracle, case):
    name, size, ch, R = case
    _say(f"[grad_vs_oracle] {name}")
    cfg, st = _case(oracle, size, ch, seed=size + ch, R=R)
    T = 5
    seed = _seed(size * size, ch, 1)
    final, g0, th = _gpu_grad(cfg, st, T, seed)
    p64 = oracle.Propagator(cfg, np.float64)
    at, o0, oth = p64.grad([a.astype(np.float64) for a in st], T, [a.astype(np.float64) for a in seed])
    _, _, oth32 = oracle.Propagator(cfg, np.float32).grad(st, T, seed)
    for c in range(ch):
        assert np.abs(final.channels[c] - at[c]).max() < 1e-4
        e = np.abs(g0[c] - o0[c]).max() / np.abs(o0[c]).max()
        _say(f"  dL/dA0[{c}] max rel err {e:.2e}")
        assert e <= REL
    _close(th, oth, oth32, "dtheta")


def test_grad_matches_fp64_central_differences(oracle):
    cfg, st = _case(oracle)
    T = 4
    seed = _seed(32 * 32, 2, 2)
    _, _, th = _gpu_grad(cfg, st, T, seed)
    theta = G.theta_of(cfg).astype(np.float32).astype(np.float64)  # the config's own f32 values
    st64 = [a.astype(np.float64) for a in st]
    sd64 = [a.astype(np.float64) for a in seed]
    for j in [0, 1, 2, 4, 5, 12, 13, 14, 20, 22]:
        e = 1e-6 * max(1.0, abs(theta[j]))
        tp, tm = theta.copy(), theta.copy()
        tp[j] += e
        tm[j] -= e
        fd = (oracle.loss(cfg, st64, T, sd64, tp) - oracle.loss(cfg, st64, T, sd64, tm)) / (2 * e)
        _say(f"  theta[{j}]: gpu {th[j]:.6e} fd {fd:.6e}")
        assert abs(th[j] - fd) <= REL * abs(fd) + 1e-6 * np.abs(th).max()


def test_l2_target_loss_matches_fp64_fd_on_64_coordinates(oracle):
    """The differentiation workload's loss at a small size: L = ||A_T - target||^2 with the
    target run under parameters perturbed by N(0, 0.05 |p|); 4 kernels (2 channels), T = 8 with a
    checkpoint every 4.  64 sampled coordinates -- every parameter and the rest initial-state
    cells -- against FP64 central differences with eps = 1e-4.  A coordinate counts when its
    difference quotient is stable (eps and eps/2 agree to 1e-5 relative: no reintegration kink in
    reach); at least 48 of the 64 must be, and each of those matches to 1e-3 relative."""
    cfg, st = _case(oracle, 32, 2, seed=11, R=6)
    T, eps = 8, 1e-4
    theta = G.theta_of(cfg).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(42)
    tgt_theta = theta + rng.normal(size=theta.size) * 0.05 * np.abs(theta)
    st64 = [a.astype(np.float64) for a in st]
    target, _, _ = oracle.Propagator(cfg, np.float64, theta=tgt_theta).run([a.copy() for a in st64], T)

    def L(th, s0):
        out, _, _ = oracle.Propagator(cfg, np.float64, theta=th).run([a.copy() for a in s0], T)
        return sum(float(((o - t) ** 2).sum()) for o, t in zip(out, target))

    a_t, _, _ = oracle.Propagator(cfg, np.float64, theta=theta).run([a.copy() for a in st64], T)
    seed = [(2.0 * (a - t)).astype(np.float32) for a, t in zip(a_t, target)]
    _, g0, th = _gpu_grad(cfg, st, T, seed, k=4)
    n = 32 * 32
    coords = [("theta", j) for j in range(theta.size)]
    cells = rng.choice(2 * n, size=64 - len(coords), replace=False)
    coords += [("state", int(i)) for i in cells]
    assert len(coords) == 64

    def fd(kind, j, e):
        if kind == "theta":
            tp, tm = theta.copy(), theta.copy()
            tp[j] += e
            tm[j] -= e
            return (L(tp, st64) - L(tm, st64)) / (2 * e)
        c, i = divmod(j, n)
        sp = [a.copy() for a in st64]
        sm = [a.copy() for a in st64]
        sp[c][i] += e
        sm[c][i] -= e
        return (L(theta, sp) - L(theta, sm)) / (2 * e)

    gscale = max(np.abs(th).max(), max(np.abs(g).max() for g in g0))
    kept, worst = 0, 0.0
    for kind, j in coords:
        f1, f2 = fd(kind, j, eps), fd(kind, j, eps / 2)
        if abs(f1 - f2) > 1e-5 * max(abs(f1), 1e-6 * gscale):
            continue  # a kink within eps: the difference quotient is not a derivative
        kept += 1
        got = th[j] if kind == "theta" else g0[j // n][j % n]
        err = abs(got - f1) / max(abs(f1), 1e-3 * gscale)
        worst = max(worst, err)
        assert err <= REL, (kind, j, got, f1)
    _say(f"  {kept} of 64 coordinates smooth at eps = 1e-4; worst relative error {worst:.2e}")
    assert kept >= 48


def test_checkpoint_interval_and_batching_are_bitwise_neutral(oracle):
    cfg, st = _case(oracle)
    T = 7
    seed = _seed(32 * 32, 2, 3)
    ref = None
    for k in (1, 3, 7, 20, 0):  # 0: interval chosen from free memory
        _, g0, th = _gpu_grad(cfg, st, T, seed, k=k)
        if ref is None:
            ref = (g0, th)
            continue
        for a, b in zip(g0, ref[0]):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(th, ref[1])
    # the same member inside a population of three
    cfg2, st2 = _case(oracle, 48, 1, seed=9, R=8)
    cfg3, st3 = _case(oracle, 64, 3, seed=4, R=9)
    with HipPopulation([cfg2, cfg, cfg3]) as pop:
        for i, (c, s) in enumerate([(cfg2, st2), (cfg, st), (cfg3, st3)]):
            pop.set_state(i, av.SimulationState(s, c.width, c.height))
        pop.tape_forward(T, 3)
        g0s, ths = pop.tape_backward([_seed(48 * 48, 1, 5), seed, _seed(64 * 64, 3, 6)])
    for a, b in zip(g0s[1], ref[0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ths[1], ref[1])


def test_tape_leaves_final_state_and_repeats(oracle):
    cfg, st = _case(oracle)
    seed = _seed(32 * 32, 2, 4)
    with HipPopulation([cfg]) as pop:
        pop.set_state(0, av.SimulationState(st, 32, 32))
        pop.tape_forward(6, 4)
        s1 = pop.get_state(0)
        g_a, t_a = pop.tape_backward([seed])
        s2 = pop.get_state(0)
        g_b, t_b = pop.tape_backward([seed])
        pop.advance(2)
        s3 = pop.get_state(0)
    for c in range(2):
        np.testing.assert_array_equal(s1.channels[c], s2.channels[c])
        np.testing.assert_array_equal(g_a[0][c], g_b[0][c])
    np.testing.assert_array_equal(t_a[0], t_b[0])
    assert s1.step == 6 and s3.step == 8
    with av.HipPropagator(cfg) as hp:
        ref = av.SimulationState(st, 32, 32)
        hp.run(ref, 8)
    assert max(np.abs(s3.channels[c] - ref.channels[c]).max() for c in range(2)) < 1e-5


def test_zero_steps_returns_the_seed(oracle):
    cfg, st = _case(oracle)
    seed = _seed(32 * 32, 2, 8)
    _, g0, th = _gpu_grad(cfg, st, 0, seed)
    for a, b in zip(g0, seed):
        np.testing.assert_array_equal(a, b)
    assert np.all(th == 0.0)


def test_total_mass_gradient_at_full_differentiation_size():
    """32 members of 512^2 x 2 channels (every cell kept active by a small background),
    checkpoints every 3 (T = 8) and every 8 (T = 64): L = total mass is conserved by
    reintegration, so dL/dA_0 = 1 and dL/dtheta = 0 in exact arithmetic.  In f32 the box
    coordinates near x ~ 512 carry ulp(512) = 6e-5 and the sharp growth functions
    (G' ~ 1000 at sigma ~ 0.01) amplify that noise; the oracle's own f32 adjoint of member 0
    at T = 8 deviates by mean 2.4e-7, 99th percentile 4.4e-6, max 3.1e-3.  Bars: 10x those
    statistics at T = 8; at T = 64 a mean deviation below 1e-4."""
    members = [workloads.differentiation(seed=0x5EED4 + j) for j in range(32)]
    rng = np.random.default_rng(100)
    for T, k in ((8, 3), (64, 8)):
        _say(f"[grad_mass_512] 32 x 512^2 x 2, T={T}, k={k}")
        with HipPopulation([c for c, _ in members]) as pop:
            for i, (c, s) in enumerate(members):
                st = av.SimulationState.from_seed(s, c)
                pop.set_state(i, av.SimulationState([a + np.float32(1e-3) for a in st.channels], c.width, c.height))
            pop.tape_forward(T, k)
            g0, th = pop.tape_backward([[np.ones(512 * 512, np.float32)] * 2 for _ in members])
            _, th_rand = pop.tape_backward([[rng.standard_normal(512 * 512).astype(np.float32) for _ in range(2)]
                                            for _ in members])
        d = np.abs(np.stack([np.stack(g) for g in g0]) - 1.0)
        ratio = max(np.abs(t).max() / np.abs(r).max() for t, r in zip(th, th_rand))
        d0 = d[0]
        _say(f"  |dL/dA0 - 1|: mean {d.mean():.2e} p99 {np.percentile(d, 99):.2e} max {d.max():.2e}; "
             f"member 0: mean {d0.mean():.2e} max {d0.max():.2e}; |dL/dtheta| / random-seed scale {ratio:.2e}")
        assert np.all(np.isfinite(d))
        if T == 8:
            assert d.mean() < 2.4e-6 and np.percentile(d, 99) < 4.4e-5 and d.max() < 3.1e-2
            assert ratio < 1e-2
        else:
            assert d.mean() < 1e-4


@pytest.mark.parametrize("dt", [None, 3.0])
def test_stored_convolution_tape_matches_recompute(dt):
    """The automatic layout keeps every step's state, convolutions, U and source spectra (each
    backward step then reads them straight from their tape slots); a checkpoint every step
    without stored convolutions recomputes the FFT passes in the backward instead.  The two
    forwards run different instantiations of the inverse row pass (with and without the
    convolution store), whose growth epilogues may round differently in the last bit, so the
    bar is closeness: a wrong slot or a stale U would miss it by orders of magnitude.  dt = 3
    pushes flows past the fused stencil's radius, so the taped forward is replayed with
    checked steps and the general path."""
    pop_def = workloads.population(6, seed=0x5EED7)
    cfgs = [c for c, _ in pop_def]
    if dt is not None:
        for c in cfgs:
            c.dt = dt
    rng = np.random.default_rng(5)
    seed = [[rng.standard_normal(c.width * c.height).astype(np.float32) for _ in range(c.channels)] for c in cfgs]
    out = {}
    for k in (0, 1):
        _say(f"[tape_layout] 6 members, T=5, dt={dt}, checkpoint_every={k}")
        with HipPopulation(cfgs) as pop:
            pop.set_seeds([s for _, s in pop_def])
            pop.advance(3)
            used = pop.tape_forward(5, k)
            g0, th = pop.tape_backward(seed)
            after = [pop.get_state(i).channels for i in range(len(cfgs))]
        out[k] = (used, g0, th, after)
    assert out[0][0] == 1 and out[1][0] == 1
    ds = max(float(np.abs(a - b).max()) for m in range(len(cfgs)) for a, b in zip(out[0][3][m], out[1][3][m]))
    dg = max(float(np.abs(a - b).max()) / (1.0 + float(np.abs(b).max()))
             for m in range(len(cfgs)) for a, b in zip(out[0][1][m], out[1][1][m]))
    dt = max(float(np.abs(a - b).max()) / (1e-12 + float(np.abs(b).max())) for a, b in zip(out[0][2], out[1][2]))
    _say(f"  states max|d| {ds:.2e}; dL/dA0 max|d|/(1+max|g|) {dg:.2e}; dL/dtheta max|d|/max|th| {dt:.2e}")
    assert ds <= 1e-5 and dg <= 1e-4 and dt <= 1e-3


def test_flow_lenia_steps_autograd_function_matches_grad(oracle):
    """FlowLeniaSteps (the torch.autograd.Function over the tape) gives grad()'s final state,
    dL/dA_0 and dL/dtheta bit for bit for L = <w, A_T>, with theta at the config's own values."""
    import torch

    cfg, st = _case(oracle, size=48, channels=2, seed=11)
    steps, k = 6, 3
    w = _seed(cfg.width * cfg.height, cfg.channels, s=5)
    final, g0, th = _gpu_grad(cfg, st, steps, w, k=k)
    _say(f"[FlowLeniaSteps] 48^2 x 2, {steps} steps, checkpoint every {k}")
    a0 = torch.from_numpy(np.stack(st).reshape(cfg.channels, cfg.height, cfg.width)).requires_grad_(True)
    theta = torch.from_numpy(G.theta_of(cfg)).requires_grad_(True)
    aT = G.FlowLeniaSteps.apply(a0, theta, cfg, steps, k)
    wt = torch.from_numpy(np.stack(w).reshape(cfg.channels, cfg.height, cfg.width))
    (aT * wt).sum().backward()
    for c in range(cfg.channels):
        np.testing.assert_array_equal(aT[c].detach().numpy().reshape(-1), final.channels[c])
        np.testing.assert_array_equal(a0.grad[c].numpy().reshape(-1), g0[c])
    np.testing.assert_array_equal(theta.grad.numpy(), th)
    assert theta.grad.dtype == torch.float64


def test_grad_at_differentiation_member_size(oracle):
    """One member of the differentiation workload (512^2 x 2 channels, 4 kernels, R = 15, seed as
    in the bench), T = 8 with a checkpoint every 4, against the FP64 oracle adjoint.

    At this size f32 state arithmetic limits every f32 adjoint: the oracle's own f32 adjoint is
    about 3e-3 from its FP64 adjoint in the parameter vector (box coordinates near x ~ 512 carry
    ulp(512) = 6e-5, and G' reaches ~1000).  The bars: dL/dA0 within 1e-3 of the FP64 adjoint's
    largest entry; every dL/dtheta component within 1e-3 relative or 3x the f32 oracle's own
    deviation (as in the smaller cases); the parameter vector's relative error at most twice the
    f32 oracle's; and the FP64 adjoint itself checked by FP64 central differences on four
    parameters (mu, sigma of kernel 0; amplitude and width of its first ring) to 1e-5.  The
    differences use eps = 1e-6 |p|: at this size box edges cross cell edges within any larger
    eps (the quotient for mu_0 reads 28.60, 28.41, 28.48 at eps = 1e-3, 1e-4, 1e-5 |p| against
    28.376 at 1e-6, where it equals the adjoint to 1e-7), so 1e-4 pins nothing here.  dL/dA0 is
    held to the same rule per cell: the f32 oracle's own dL/dA0 is up to 1.7e-2 of the largest
    entry from the FP64 one here (about 0.5% of the cells beyond 1e-3)."""
    cfg, sd = workloads.differentiation(seed=0x5EED4)
    st = [np.asarray(a, np.float32).reshape(-1) for a in oracle.seed_generate(sd, cfg.width, cfg.height, 1,
                                                                               cfg.channels)]
    T = 8
    _say(f"[grad_512] 512^2 x 2, 4 kernels, T={T}, checkpoint every 4")
    seed = _seed(cfg.width * cfg.height, cfg.channels, 9)
    final, g0, th = _gpu_grad(cfg, st, T, seed, k=4)
    p64 = oracle.Propagator(cfg, np.float64)
    st64 = [a.astype(np.float64) for a in st]
    sd64 = [a.astype(np.float64) for a in seed]
    at, o0, oth = p64.grad([a.copy() for a in st64], T, sd64)
    _, o032, oth32 = oracle.Propagator(cfg, np.float32).grad(st, T, seed)
    for c in range(cfg.channels):
        assert np.abs(final.channels[c] - at[c]).max() < 1e-4
        m = np.abs(o0[c]).max()
        d, d32 = np.abs(g0[c] - o0[c]), np.abs(o032[c] - o0[c])
        _say(f"  dL/dA0[{c}] max rel err {d.max() / m:.2e} (oracle f32: {d32.max() / m:.2e}); "
             f"cells beyond 1e-3: {(d > REL * m).sum()} (oracle f32: {(d32 > REL * m).sum()})")
        assert not (d > np.maximum(REL * m, 3.0 * d32) + 0.1 * REL * m).any()
    scale = np.abs(oth).max()
    err = np.abs(th - oth)
    bad = err > np.maximum(REL * np.abs(oth), 3.0 * np.abs(oth32 - oth)) + 0.1 * REL * scale
    assert not bad.any(), f"dtheta components {np.nonzero(bad)[0]}"
    vec = np.linalg.norm(th - oth) / np.linalg.norm(oth)
    vec32 = np.linalg.norm(oth32 - oth) / np.linalg.norm(oth)
    _say(f"  dtheta: |g - g64| / |g64| = {vec:.2e} (oracle f32: {vec32:.2e})")
    assert vec <= max(REL, 2.0 * vec32)
    theta = G.theta_of(cfg).astype(np.float32).astype(np.float64)
    for j in (1, 2, 12, 14):  # kernel 0's mu and sigma; its first ring's amplitude and width
        e = 1e-6 * max(1e-2, abs(theta[j]))
        tp, tm = theta.copy(), theta.copy()
        tp[j] += e
        tm[j] -= e
        fd = (oracle.loss(cfg, st64, T, sd64, tp) - oracle.loss(cfg, st64, T, sd64, tm)) / (2 * e)
        _say(f"  theta[{j}]: FP64 adjoint {oth[j]:.6e}, FP64 central difference {fd:.6e}, gpu {th[j]:.6e} "
             f"(gpu vs FP64 {abs(th[j] - oth[j]) / abs(oth[j]):.1e}, oracle f32 vs FP64 "
             f"{abs(oth32[j] - oth[j]) / abs(oth[j]):.1e})")
        assert abs(oth[j] - fd) <= 1e-5 * abs(fd) + 1e-7 * scale

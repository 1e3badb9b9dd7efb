"""The persistent population step (hp_batch_set_step_mode HP_STEP_PERSISTENT): one launch per
step whose workgroups walk every member's forward row pass, spectral pass, inverse row pass and
stencil tiles with a bounded grid barrier between the phases.

Bars: every member's trajectory equals the multi-launch step's and a lone HipPropagator's with
the same FFT plan (radix order 0) bit for bit, over several launches in a row (the barrier
counters reset between launches) and across both step parities; no barrier times out; the
gradient tape runs from the persistent step's graphs with the same results; sides outside the
compiled lengths are refused."""
import numpy as np
import pytest

from autoverse_amd import _abi
from autoverse_amd.population import HipPopulation
from autoverse_amd.schema import FlowConfig

from test_gpu_population import _lone, _members

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _mixed(seed=31):
    sides = [64, 128, 192, 256, 64, 128, 256, 192]
    chans = [1, 2, 3, 2, 3, 1, 1, 2]
    mem = _members(sides, chans, seed=seed)
    mem[4][0].flow = FlowConfig(1.0, 2.0, 0.75)  # a member whose box area is not a power of two
    mem[5][0].flow = FlowConfig(1.0, 2.0, 0.6)
    return sides, chans, mem


def _run(mem, mode, chunks):
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_plan({"radix_order": 0})
        pop.set_step_mode(mode)
        assert pop.step_mode()[0] == mode
        pop.set_seeds([s for _, s in mem])
        for k in chunks:
            pop.advance(k)
        states = [pop.get_state(i) for i in range(len(mem))]
        mass = pop.mass_f64()
        _, timeouts = pop.step_mode()
        _, g = pop.stats()
        plan = pop.plan()
    return states, mass, timeouts, g, plan


def test_persistent_step_equals_multi_launch_bitwise():
    sides, chans, mem = _mixed()
    _say(f"[persistent step] sides={sides} channels={chans}, 7 + 1 + 12 steps")
    got, gm, timeouts, g, _ = _run(mem, "persistent", [7, 1, 12])
    want, wm, _, wg, _ = _run(mem, "multi_launch", [20])
    _say(f"  barrier timeouts {timeouts}, general-path steps {g}")
    assert timeouts == 0 and g == 0 and wg == 0
    for i, (cfg, _) in enumerate(mem):
        assert got[i].step == 20 and got[i].time == want[i].time
        for c in range(cfg.channels):
            np.testing.assert_array_equal(got[i].channels[c], want[i].channels[c], err_msg=f"member {i} ch {c}")
    np.testing.assert_array_equal(gm, wm)


def test_persistent_step_members_match_lone_propagators(monkeypatch):
    _, _, mem = _mixed(seed=44)
    got, _, timeouts, g, plan = _run(mem, "persistent", [15])
    assert timeouts == 0 and g == 0
    for i in (0, 2, 3, 4):
        cfg, sd = mem[i]
        want, _ = _lone(cfg, sd, 15, monkeypatch, plan)
        for c in range(cfg.channels):
            np.testing.assert_array_equal(got[i].channels[c], want.channels[c], err_msg=f"member {i} ch {c}")


def test_persistent_step_gradient_tape_matches_multi_launch():
    mem = _members([64, 128, 64], [2, 1, 3], seed=8)
    out = []
    for mode in ("multi_launch", "persistent"):
        with HipPopulation([c for c, _ in mem]) as pop:
            pop.set_plan({"radix_order": 0})
            pop.set_step_mode(mode)
            pop.set_seeds([s for _, s in mem])
            pop.tape_forward(6, 3)
            final = [pop.get_state(i) for i in range(len(mem))]
            seed = [[np.ones_like(ch) for ch in st.channels] for st in final]
            res = pop.tape_backward(seed)
            out.append((final, res))
            assert pop.step_mode()[1] == 0
    (fa, ra), (fb, rb) = out
    for i in range(len(mem)):
        for c in range(len(fa[i].channels)):
            np.testing.assert_array_equal(fa[i].channels[c], fb[i].channels[c])
    ga, gb = ra[0], rb[0]
    for i in range(len(mem)):
        for c in range(len(ga[i])):
            np.testing.assert_array_equal(ga[i][c], gb[i][c])
    np.testing.assert_array_equal(np.concatenate([np.ravel(t) for t in ra[1]]),
                                  np.concatenate([np.ravel(t) for t in rb[1]]))


def test_persistent_step_refuses_other_sides():
    mem = _members([64, 96], [1, 1], seed=3)
    with HipPopulation([c for c, _ in mem]) as pop:
        with pytest.raises(RuntimeError, match="64, 128, 192 and 256"):
            pop.set_step_mode("persistent")
        assert pop.step_mode()[0] == "multi_launch"

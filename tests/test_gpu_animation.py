"""FLWA frames recorded straight from device memory (a HipPropagator's or a population member's
resident state, copied asynchronously behind the queued steps) and read back, and the resume
path (a recorded frame loaded into a fresh propagator continues the same trajectory, bit for
bit)."""
import numpy as np
import pytest

import autoverse_amd as av
from autoverse_amd import workloads
from autoverse_amd.animation import AnimationPlayer, AnimationRecorder, CompressionType, RecorderConfig
from autoverse_amd.population import HipPopulation

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("compression", [CompressionType.NONE, CompressionType.LZ4])
def test_record_from_device_and_resume(tmp_path, compression):
    _say(f"[animation device] 128^2 x 3, compression {compression.name}")
    cfg, seed = workloads.multispecies(size=128, channels=3, seed=12)
    st0 = av.SimulationState.from_seed(seed, cfg)
    path = tmp_path / "run.flwa"
    # recorded run: frames queued between advances, no synchronising read in between
    with av.HipPropagator(cfg) as hp:
        hp.set_state(st0)
        rec = AnimationRecorder(path, cfg, RecorderConfig(compression, frame_skip=1))
        rec.record_frame(hp)
        for _ in range(6):
            hp.advance(5)
            rec.record_frame(hp)
        stats = rec.finalize()
    assert stats.frame_count == 7
    # reference run: the same trajectory read back with get_state
    want = []
    with av.HipPropagator(cfg) as hp:
        hp.set_state(st0)
        want.append(hp.get_state())
        for _ in range(6):
            hp.advance(5)
            want.append(hp.get_state())
    with AnimationPlayer.open(path) as pl:
        assert pl.frame_count() == 7 and abs(pl.dt() - cfg.dt) < 1e-7
        for k in range(7):
            f = pl.read_frame(k)
            for c in range(3):
                np.testing.assert_array_equal(f.channels[c], want[k].channels[c], err_msg=f"frame {k} ch {c}")
        # resume from frame 3 (step 15) and run 15 more steps: the state of frame 6, bit for bit
        with av.HipPropagator(cfg) as hp:
            pl.load_into(hp, 3, time=want[3].time, step=want[3].step)
            hp.advance(15)
            got = hp.get_state()
    assert got.step == 30
    for c in range(3):
        np.testing.assert_array_equal(got.channels[c], want[6].channels[c])


def test_record_population_member(tmp_path):
    _say("[animation member]")
    mem = []
    for side, c in ((64, 2), (96, 1)):
        cfg, seed = workloads.multispecies(size=side, channels=c, seed=side)
        mem.append((cfg, seed))
    path = tmp_path / "m.flwa"
    with HipPopulation([c for c, _ in mem]) as pop:
        pop.set_seeds([s for _, s in mem])
        rec = AnimationRecorder(path, mem[1][0], RecorderConfig(frame_skip=2))
        states = []
        for k in range(6):
            pop.advance(1)
            if rec.record_frame(pop, member=1):
                states.append(pop.get_state(1))
        with pytest.raises(Exception, match="shape"):
            rec.record_frame(pop, member=0)  # another member's grid
        assert rec.finalize().frame_count == 3
    with AnimationPlayer.open(path) as pl:
        assert pl.frame_count() == 3
        for k, s in enumerate(states):
            np.testing.assert_array_equal(pl.read_frame(k).channels[0], s.channels[0])

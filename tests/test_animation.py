"""FLWA animation files on the CPU: the native recorder and player (host buffers) against the
reference's own tests (src/animation/{format,recorder,player}.rs) and the format restatement
oracle/animation.py (byte for byte for uncompressed files; LZ4 frames through an independent
decoder and a hand-built block)."""
import struct

import numpy as np
import pytest

from autoverse_amd import _abi
from autoverse_amd.animation import AnimationPlayer, AnimationRecorder, CompressionType, RecorderConfig
from autoverse_amd.propagator import SimulationState
from autoverse_amd.schema import Pattern, Seed, SimulationConfig
from oracle import animation as OA


def _cfg(w=16, h=16, d=1, c=1, dt=0.1):
    """test_config (recorder.rs:265-276, player.rs:248-259)."""
    return SimulationConfig(width=w, height=h, depth=d, channels=c, dt=dt, kernel_radius=4, kernels=[])


def _noise_state(cfg):
    seed = Seed(Pattern("Noise", amplitude=0.5, channel=0, seed=42))
    return SimulationState.from_seed(seed, cfg)


def test_header_layout_and_roundtrip(tmp_path):
    """format.rs:261-292: a 64 x 48 x 32 x 2 header with LZ4 round-trips; its bytes are the
    restatement's (magic, version, flags, sizes, frame count, dt, 16 reserved zeros)."""
    cfg = _cfg(64, 48, 32, 2, 0.1)
    p = tmp_path / "h.flwa"
    rec = AnimationRecorder(p, cfg, RecorderConfig(CompressionType.LZ4))
    st = rec.finalize()
    assert st.frame_count == 0 and st.compression == CompressionType.LZ4
    b = p.read_bytes()
    assert len(b) == OA.HEADER_SIZE
    assert b == OA.encode_header(64, 48, 32, 2, 0, np.float32(0.1), 1)
    with AnimationPlayer.open(p) as pl:
        h = pl.header()
        assert (h.width, h.height, h.depth, h.channels, h.frame_count) == (64, 48, 32, 2, 0)
        assert abs(h.dt - 0.1) < 1e-6 and h.compression == CompressionType.LZ4


def test_recorder_basic_and_file_bytes(tmp_path):
    """recorder.rs:279-306: 10 frames; the file equals the restatement's bytes for them."""
    cfg = _cfg()
    state = _noise_state(cfg)
    p = tmp_path / "t.flwa"
    rec = AnimationRecorder(p, cfg)
    for _ in range(10):
        assert rec.record_frame(state)
    st = rec.finalize()
    assert st.frame_count == 10 and st.total_bytes > 0
    b = p.read_bytes()
    assert len(b) > OA.HEADER_SIZE and st.total_bytes == len(b)
    assert b == OA.write_file(16, 16, 1, 1, np.float32(0.1), [state.channels] * 10)
    assert st.average_frame_size == 16 * 16 * 4


def test_recorder_frame_skip(tmp_path):
    """recorder.rs:309-333: frame_skip 5 over 20 calls records 4 frames."""
    cfg = _cfg()
    state = _noise_state(cfg)
    rec = AnimationRecorder(tmp_path / "s.flwa", cfg, RecorderConfig(frame_skip=5))
    got = [rec.record_frame(state) for _ in range(20)]
    assert rec.finalize().frame_count == 4
    assert [i for i, g in enumerate(got) if g] == [4, 9, 14, 19]


def test_recorder_max_frames(tmp_path):
    """recorder.rs:336-360."""
    cfg = _cfg()
    state = _noise_state(cfg)
    rec = AnimationRecorder(tmp_path / "m.flwa", cfg, RecorderConfig(max_frames=5))
    for _ in range(100):
        rec.record_frame(state)
    assert rec.finalize().frame_count == 5


def test_recorder_and_player_3d(tmp_path):
    """recorder.rs:363-389 and player.rs:359-395."""
    cfg = _cfg(d=8)
    seed = Seed(Pattern("GaussianSphere", center=(0.5, 0.5, 0.5), radius=0.2, amplitude=1.0, channel=0))
    state = SimulationState.from_seed(seed, cfg)
    p = tmp_path / "3d.flwa"
    rec = AnimationRecorder(p, cfg)
    rec.record_frame(state)
    st = rec.finalize()
    assert st.frame_count == 1 and st.average_frame_size >= 16 * 16 * 8 * 4
    with AnimationPlayer.open(p) as pl:
        assert pl.is_3d() and pl.dimensions() == (16, 16, 8)
        got = pl.read_frame(0)
        assert got.channels[0].size == 16 * 16 * 8
        np.testing.assert_array_equal(got.channels[0], state.channels[0])


def test_player_roundtrip_multiple_and_iterator(tmp_path):
    """player.rs:262-356, 398-427: two channels, five distinct frames, the iterator."""
    cfg = _cfg(c=2)
    base = _noise_state(cfg)
    states = []
    for i in range(5):
        s = base.copy()
        s.channels[0] = s.channels[0] * np.float32(i + 1)
        s.channels[1] = np.full(256, np.float32(i) - 2.5, np.float32)
        states.append(s)
    p = tmp_path / "multi.flwa"
    rec = AnimationRecorder(p, cfg)
    for s in states:
        rec.record_frame(s)
    rec.finalize()
    with AnimationPlayer.open(p) as pl:
        assert pl.frame_count() == 5 and pl.channels() == 2 and pl.dimensions() == (16, 16, 1)
        for i, s in enumerate(states):
            f = pl.read_frame(i)
            assert f.time == 0.0 and f.step == 0
            for c in range(2):
                np.testing.assert_array_equal(f.channels[c], s.channels[c])
        frames = list(pl.frames())
        assert len(frames) == 5
        with pytest.raises(_abi.HipError, match="out of range"):
            pl.read_frame(5)


def test_lz4_frames_decode_with_an_independent_decoder(tmp_path):
    """LZ4-compressed frames (repetitive and random data) decode to the recorded values with the
    restatement's decoder, and the native player reads them back; the repetitive frame shrinks."""
    cfg = _cfg(32, 24, 1, 2)
    rng = np.random.default_rng(3)
    frames = [[np.zeros(768, np.float32), np.tile(np.arange(12, dtype=np.float32), 64)],
              [rng.random(768).astype(np.float32), rng.random(768).astype(np.float32)],
              [np.full(768, 0.25, np.float32), (np.arange(768) % 7).astype(np.float32)]]
    p = tmp_path / "z.flwa"
    rec = AnimationRecorder(p, cfg, RecorderConfig(CompressionType.LZ4))
    for fr in frames:
        rec.record_frame(SimulationState(fr, 32, 24))
    st = rec.finalize()
    h, raw = OA.read_file(p.read_bytes())
    assert h["compression"] == 1 and h["frame_count"] == 3
    for fr, r in zip(frames, raw):
        assert r == OA.encode_frame(fr)
    tab = p.read_bytes()[-3 * OA.INDEX_SIZE:]
    sizes = [struct.unpack("<QQ", tab[16 * i:16 * (i + 1)])[1] for i in range(3)]
    assert sizes[0] < 768 * 2 * 4 // 4 and st.frame_count == 3
    with AnimationPlayer.open(p) as pl:
        for i, fr in enumerate(frames):
            got = pl.read_frame(i)
            for c in range(2):
                np.testing.assert_array_equal(got.channels[c], fr[c])


def test_player_reads_hand_built_lz4_blocks(tmp_path):
    """A file whose frames were packed outside the library: one block of literals only, and one
    with an overlapping match (offset 4, copied 40 bytes) followed by the closing literals."""
    a = np.arange(16, dtype=np.float32)  # 64 bytes
    lit = OA.lz4_literal_block(OA.encode_frame([a]))
    pattern = np.tile(np.array([1.5], np.float32), 16)
    raw = OA.encode_frame([pattern])  # 64 bytes: 4-byte pattern repeated
    # token: 4 literals, match length 40 -> 40 - 4 = 36 >= 15: 15 + ext 21; offset 4; then 20 literals
    blk = bytes([(4 << 4) | 15]) + raw[:4] + bytes([4, 0, 21]) + bytes([20 << 4 if 20 < 15 else 15 << 4, 5]) + raw[44:]
    blk = struct.pack("<I", 64) + blk
    assert OA.lz4_decompress(blk) == raw
    p = tmp_path / "hand.flwa"
    p.write_bytes(OA.write_file(4, 4, 1, 1, np.float32(0.2), [[a], [pattern]], packed=[lit, blk]))
    with AnimationPlayer.open(p) as pl:
        np.testing.assert_array_equal(pl.read_frame(0).channels[0], a)
        np.testing.assert_array_equal(pl.read_frame(1).channels[0], pattern)


def test_player_rejects_bad_files(tmp_path):
    cfg = _cfg()
    p = tmp_path / "bad.flwa"
    p.write_bytes(b"XXXX" + bytes(48))
    with pytest.raises(_abi.HipError, match="magic"):
        AnimationPlayer.open(p)
    p.write_bytes(OA.encode_header(16, 16, 1, 1, 0, 0.1)[:4] + struct.pack("<H", 2) + bytes(46))
    with pytest.raises(_abi.HipError, match="version"):
        AnimationPlayer.open(p)
    p.write_bytes(OA.encode_header(16, 16, 1, 1, 3, 0.1))  # three frames claimed, none present
    with pytest.raises(_abi.HipError, match="too short"):
        AnimationPlayer.open(p)
    # a corrupt LZ4 frame: offset pointing before the start of the output
    bad = struct.pack("<I", 64) + bytes([(1 << 4) | 0, 7, 9, 0])
    p.write_bytes(OA.write_file(4, 4, 1, 1, 0.1, [[np.zeros(16, np.float32)]], packed=[bad]))
    with AnimationPlayer.open(p) as pl, pytest.raises(_abi.HipError, match="corrupt LZ4"):
        pl.read_frame(0)
    with pytest.raises(_abi.HipError, match="cannot create"):
        AnimationRecorder(tmp_path / "no" / "such" / "dir.flwa", cfg)

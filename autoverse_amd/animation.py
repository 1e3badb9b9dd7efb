"""FLWA animation files: the reference's AnimationRecorder / AnimationPlayer over the native library.

Mirrors src/animation/ (format.rs CompressionType :12-30 and AnimationHeader :60-176,
recorder.rs RecorderConfig :14-33, AnimationRecorder :47-205, AnimationStats :208-224,
player.rs AnimationPlayer :26-200 and its frame iterator :213-240).  Recording from a
HipPropagator or a population member copies the device-resident state straight into pinned
staging and writes it on a background thread while the device keeps stepping; the player reads
frames back into host states, or loads one into a HipPropagator to resume a run.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from enum import IntEnum
from typing import Iterator, List

import numpy as np

from . import _abi
from .propagator import HipPropagator, SimulationState
from .schema import SimulationConfig


class CompressionType(IntEnum):
    """CompressionType (format.rs:12-30)."""
    NONE = 0
    LZ4 = 1


@dataclass
class RecorderConfig:
    """RecorderConfig (recorder.rs:14-33)."""
    compression: CompressionType = CompressionType.NONE
    frame_skip: int = 1
    max_frames: int = 0


@dataclass
class AnimationStats:
    """AnimationStats (recorder.rs:208-224)."""
    frame_count: int
    total_bytes: int
    average_frame_size: int
    compression: CompressionType


@dataclass
class AnimationHeader:
    """AnimationHeader (format.rs:60-98)."""
    width: int
    height: int
    depth: int
    channels: int
    frame_count: int
    dt: float
    compression: CompressionType
    delta_encoding: bool

    SIZE = 52

    def frame_size(self) -> int:
        return self.width * self.height * self.depth * self.channels * 4

    def is_3d(self) -> bool:
        return self.depth > 1


def _err(lib, st, what):
    raise _abi.HipError(st, f"{what}: {lib.hp_animation_last_error().decode()}")


class AnimationRecorder:
    """AnimationRecorder (recorder.rs:47-205)."""

    def __init__(self, path: str, sim_config: SimulationConfig, config: RecorderConfig | None = None):
        self._lib = _abi.load_library()
        self._cfg = sim_config
        self._c = sim_config.to_c()
        rc = config or RecorderConfig()
        self._rc = _abi.HpRecorderConfig(int(rc.compression), int(rc.frame_skip), int(rc.max_frames))
        p = ctypes.c_void_p()
        st = self._lib.hp_recorder_open(str(path).encode(), ctypes.byref(self._c.cfg), ctypes.byref(self._rc),
                                        ctypes.byref(p))
        if st != _abi.HP_OK:
            _err(self._lib, st, "hp_recorder_open")
        self._ptr = p
        self._written = 0

    def record_frame(self, source, member: int | None = None) -> bool:
        """record_frame (recorder.rs:119-170).  `source` is a SimulationState (host buffers), a
        HipPropagator (its device-resident state), or a HipPopulation with `member`."""
        rec = ctypes.c_int(0)
        if isinstance(source, HipPropagator):
            st = self._lib.hp_recorder_record_ctx(self._ptr, source._ptr, ctypes.byref(rec))
        elif isinstance(source, SimulationState):
            chans = [np.ascontiguousarray(c, dtype=np.float32) for c in source.channels]
            st = self._lib.hp_recorder_record_host(self._ptr, _abi.float_ptr_array(chans), ctypes.byref(rec))
        else:
            if member is None:
                raise ValueError("a population frame needs the member index")
            st = self._lib.hp_recorder_record_member(self._ptr, source._ptr, int(member), ctypes.byref(rec))
        if st != _abi.HP_OK:
            _err(self._lib, st, "record_frame")
        self._written += rec.value
        return bool(rec.value)

    def frames_written(self) -> int:
        return self._written

    def finalize(self) -> AnimationStats:
        """finalize (recorder.rs:175-205): index table, header frame count, close."""
        s = _abi.HpAnimationStats()
        p, self._ptr = self._ptr, None
        st = self._lib.hp_recorder_finalize(p, ctypes.byref(s))
        if st != _abi.HP_OK:
            _err(self._lib, st, "finalize")
        return AnimationStats(int(s.frame_count), int(s.total_bytes), int(s.average_frame_size),
                              CompressionType(s.compression))

    def __del__(self):
        try:
            if getattr(self, "_ptr", None):
                self._lib.hp_recorder_finalize(self._ptr, None)
                self._ptr = None
        except Exception:
            pass


class AnimationPlayer:
    """AnimationPlayer (player.rs:26-240)."""

    def __init__(self, ptr, lib):
        self._ptr, self._lib = ptr, lib
        info = _abi.HpAnimationInfo()
        st = lib.hp_player_info(ptr, ctypes.byref(info))
        if st != _abi.HP_OK:
            _err(lib, st, "hp_player_info")
        self._header = AnimationHeader(info.width, info.height, info.depth, info.channels, int(info.frame_count),
                                       info.dt, CompressionType(info.compression), bool(info.delta_encoding))

    @classmethod
    def open(cls, path: str) -> "AnimationPlayer":
        lib = _abi.load_library()
        p = ctypes.c_void_p()
        st = lib.hp_player_open(str(path).encode(), ctypes.byref(p))
        if st != _abi.HP_OK:
            _err(lib, st, "hp_player_open")
        return cls(p, lib)

    def close(self) -> None:
        if getattr(self, "_ptr", None):
            self._lib.hp_player_close(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def header(self) -> AnimationHeader:
        return self._header

    def frame_count(self) -> int:
        return self._header.frame_count

    def dimensions(self):
        return self._header.width, self._header.height, self._header.depth

    def channels(self) -> int:
        return self._header.channels

    def is_3d(self) -> bool:
        return self._header.is_3d()

    def dt(self) -> float:
        return self._header.dt

    def read_frame_into(self, index: int, channels: List[np.ndarray]) -> None:
        for c in channels:
            if c.dtype != np.float32 or not c.flags["C_CONTIGUOUS"]:
                raise ValueError("channel buffers must be contiguous float32")
        st = self._lib.hp_player_read_frame(self._ptr, int(index), _abi.float_ptr_array(channels))
        if st != _abi.HP_OK:
            _err(self._lib, st, "read_frame")

    def read_frame(self, index: int) -> SimulationState:
        """read_frame (player.rs:101-128): time and step are 0, as the reference returns them."""
        w, h, d = self.dimensions()
        chans = [np.empty(w * h * d, np.float32) for _ in range(self.channels())]
        self.read_frame_into(index, chans)
        return SimulationState(chans, w, h, d)

    def frames(self) -> Iterator[SimulationState]:
        for i in range(self.frame_count()):
            yield self.read_frame(i)

    def load_into(self, prop: HipPropagator, index: int, time: float = 0.0, step: int = 0) -> None:
        """Resume: frame `index` becomes `prop`'s device-resident state with the given time and step."""
        st = self._lib.hp_player_load_ctx(self._ptr, int(index), prop._ptr, float(time), int(step))
        if st != _abi.HP_OK:
            _err(self._lib, st, "load_into")

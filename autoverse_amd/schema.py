"""Configuration and seed types of the Flow Lenia step.

Python mirror of the reference crate's schema (src/schema/config.rs:14-180 and
src/schema/seed.rs:13-152): same field names, same serde JSON layout (``Pattern``
is internally tagged by ``"type"``, seed.rs:36-38), same defaults
(config.rs:37-51, 72-90, 112-120).  ``to_c()`` lowers a value to the POD structs
of include/flowlenia_hip.h; the numeric work (kernel build, seed rasterisation,
the step) happens in the native library.
"""
from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

from . import _abi


class ConfigError(ValueError):
    """SimulationConfig::validate failures (config.rs:169-180)."""

    def __init__(self, status: int, kernel: int | None = None, channel: int | None = None):
        self.status = status
        self.kernel = kernel
        self.channel = channel
        msgs = {
            _abi.HP_ERR_INVALID_DIMENSIONS: "Grid dimensions (width, height, depth) must be non-zero",
            _abi.HP_ERR_INVALID_CHANNELS: "Channel count must be non-zero",
            _abi.HP_ERR_INVALID_TIME_STEP: "Time step must be positive",
            _abi.HP_ERR_INVALID_KERNEL_RADIUS: "Kernel radius must be non-zero",
            _abi.HP_ERR_INVALID_CHANNEL_INDEX: f"Kernel {kernel} references invalid channel {channel}",
        }
        super().__init__(msgs.get(status, _abi.STATUS_NAMES.get(status, str(status))))


@dataclass
class RingConfig:
    """RingConfig (config.rs:92-99)."""
    amplitude: float
    distance: float
    width: float


@dataclass
class KernelConfig:
    """KernelConfig (config.rs:55-90); defaults as KernelConfig::default."""
    radius: float = 1.0
    rings: List[RingConfig] = field(default_factory=lambda: [RingConfig(1.0, 0.5, 0.15)])
    weight: float = 1.0
    mu: float = 0.15
    sigma: float = 0.015
    source_channel: int = 0
    target_channel: int = 0


@dataclass
class FlowConfig:
    """FlowConfig (config.rs:103-120); defaults beta_a=2, n=4, distribution_size=0.5."""
    beta_a: float = 2.0
    n: float = 4.0
    distribution_size: float = 0.5


@dataclass
class EmbeddingConfig:
    """EmbeddingConfig (embedding.rs:326-367): parameter embedding on/off, the softmax
    temperature of the mixing, and linear mass-weighted mixing instead of softmax."""
    enabled: bool = False
    mixing_temperature: float = 1.0
    linear_mixing: bool = False

    @classmethod
    def enabled_softmax(cls) -> "EmbeddingConfig":
        """EmbeddingConfig::enabled (embedding.rs:349-356)."""
        return cls(True, 1.0, False)

    @classmethod
    def enabled_linear(cls) -> "EmbeddingConfig":
        """EmbeddingConfig::enabled_linear (embedding.rs:358-366)."""
        return cls(True, 1.0, True)

    def to_c(self) -> "_abi.HpEmbedding":
        return _abi.HpEmbedding(int(bool(self.enabled)), float(self.mixing_temperature),
                                int(bool(self.linear_mixing)))


@dataclass
class SimulationConfig:
    """SimulationConfig (config.rs:14-51).  `embedding` is read by the embedded-parameter
    propagator only."""
    width: int = 256
    height: int = 256
    depth: int = 1
    channels: int = 1
    dt: float = 0.05
    kernel_radius: int = 13
    kernels: List[KernelConfig] = field(default_factory=lambda: [KernelConfig()])
    flow: FlowConfig = field(default_factory=FlowConfig)
    embedding: EmbeddingConfig = field(default_factory=EmbeddingConfig)

    def is_3d(self) -> bool:
        return self.depth > 1

    def grid_size(self) -> int:
        return self.width * self.height * self.depth

    # ---- JSON (serde field names) -------------------------------------------------
    @classmethod
    def from_dict(cls, d: dict) -> "SimulationConfig":
        kernels = [KernelConfig(radius=k["radius"], rings=[RingConfig(**r) for r in k["rings"]],
                                weight=k["weight"], mu=k["mu"], sigma=k["sigma"],
                                source_channel=k["source_channel"], target_channel=k["target_channel"])
                   for k in d["kernels"]]
        emb = d.get("embedding")
        return cls(width=d["width"], height=d["height"], depth=d.get("depth", 1), channels=d["channels"],
                   dt=d["dt"], kernel_radius=d["kernel_radius"], kernels=kernels,
                   flow=FlowConfig(**d["flow"]),
                   embedding=EmbeddingConfig(**emb) if emb is not None else EmbeddingConfig())

    @classmethod
    def from_json(cls, text: str) -> "SimulationConfig":
        return cls.from_dict(json.loads(text))

    def to_dict(self) -> dict:
        return {
            "width": self.width, "height": self.height, "depth": self.depth, "channels": self.channels,
            "dt": self.dt, "kernel_radius": self.kernel_radius,
            "kernels": [{"radius": k.radius, "rings": [r.__dict__.copy() for r in k.rings],
                         "weight": k.weight, "mu": k.mu, "sigma": k.sigma,
                         "source_channel": k.source_channel, "target_channel": k.target_channel}
                        for k in self.kernels],
            "flow": self.flow.__dict__.copy(),
            "embedding": self.embedding.__dict__.copy(),
        }

    # ---- lowering to the C ABI ------------------------------------------------------
    def to_c(self) -> "CConfig":
        return CConfig(self)

    def validate(self) -> None:
        """SimulationConfig::validate (config.rs:136-164), run by the native library."""
        lib = _abi.load_library()
        c = self.to_c()
        bk, bc = ctypes.c_uint32(0), ctypes.c_uint32(0)
        st = lib.hp_validate_config(ctypes.byref(c.cfg), ctypes.byref(bk), ctypes.byref(bc))
        if st != _abi.HP_OK:
            raise ConfigError(st, bk.value, bc.value)


class CConfig:
    """Owns the ctypes arrays behind one hp_config."""

    def __init__(self, cfg: SimulationConfig):
        rings = []
        ks = []
        for k in cfg.kernels:
            ks.append(_abi.HpKernel(k.radius, k.weight, k.mu, k.sigma, k.source_channel, k.target_channel,
                                    len(rings), len(k.rings)))
            rings.extend(_abi.HpRing(r.amplitude, r.distance, r.width) for r in k.rings)
        self.kernels = (_abi.HpKernel * max(1, len(ks)))(*ks)
        self.rings = (_abi.HpRing * max(1, len(rings)))(*rings)
        self.cfg = _abi.HpConfig(
            cfg.width, cfg.height, cfg.depth, cfg.channels, cfg.dt, cfg.kernel_radius,
            _abi.HpFlow(cfg.flow.beta_a, cfg.flow.n, cfg.flow.distribution_size),
            len(ks), ctypes.cast(self.kernels, ctypes.POINTER(_abi.HpKernel)),
            len(rings), ctypes.cast(self.rings, ctypes.POINTER(_abi.HpRing)))


# ---------------------------------------------------------------------------------------
# Seeds (seed.rs:13-152)
# ---------------------------------------------------------------------------------------
@dataclass
class BlobSpec:
    """BlobSpec (seed.rs:138-143)."""
    center: Tuple[float, float]
    radius: float
    amplitude: float
    channel: int


@dataclass
class SphereSpec:
    """SphereSpec (seed.rs:146-152)."""
    center: Tuple[float, float, float]
    radius: float
    amplitude: float
    channel: int


@dataclass
class Pattern:
    """One variant of the Pattern enum (seed.rs:36-135); ``type`` is the serde tag."""
    type: str
    center: Tuple[float, ...] = (0.5, 0.5)
    radius: float = 0.1
    amplitude: float = 1.0
    channel: Optional[int] = 0
    inner_radius: float = 0.0
    outer_radius: float = 0.0
    major_radius: float = 0.0
    minor_radius: float = 0.0
    seed: int = 0
    blobs: List[BlobSpec] = field(default_factory=list)
    spheres: List[SphereSpec] = field(default_factory=list)
    values: List[tuple] = field(default_factory=list)

    @classmethod
    def from_dict(cls, d: dict) -> "Pattern":
        t = d["type"]
        if t not in _abi.PATTERN_KINDS:
            raise ValueError(f"unknown pattern type {t!r}")
        kw = {k: v for k, v in d.items() if k != "type"}
        if "center" in kw:
            kw["center"] = tuple(kw["center"])
        if "blobs" in kw:
            kw["blobs"] = [BlobSpec(tuple(b["center"]), b["radius"], b["amplitude"], b["channel"])
                           for b in kw["blobs"]]
        if "spheres" in kw:
            kw["spheres"] = [SphereSpec(tuple(s["center"]), s["radius"], s["amplitude"], s["channel"])
                             for s in kw["spheres"]]
        if "values" in kw:
            kw["values"] = [tuple(v) for v in kw["values"]]
        return cls(type=t, **kw)


@dataclass
class Seed:
    """Seed (seed.rs:13-32); default is a GaussianBlob at (0.5, 0.5), radius 0.1."""
    pattern: Pattern = field(default_factory=lambda: Pattern("GaussianBlob"))

    @classmethod
    def from_dict(cls, d: dict) -> "Seed":
        return cls(Pattern.from_dict(d["pattern"]))

    @classmethod
    def from_json(cls, text: str) -> "Seed":
        return cls.from_dict(json.loads(text))

    def to_c(self) -> "CSeed":
        return CSeed(self)

    def generate(self, width: int, height: int, depth: int, channels: int):
        """Seed::generate (seed.rs:160-304) -> list of float32 arrays of W*H*D, native code."""
        import numpy as np
        lib = _abi.load_library()
        out = [np.zeros(width * height * depth, dtype=np.float32) for _ in range(channels)]
        ptrs = _abi.float_ptr_array(out)
        cs = self.to_c()
        st = lib.hp_seed_generate(ctypes.byref(cs.seed), width, height, depth, channels, ptrs)
        if st != _abi.HP_OK:
            raise _abi.HipError(st, "hp_seed_generate")
        return out


class CSeed:
    """Owns the ctypes arrays behind one hp_seed."""

    def __init__(self, seed: Seed):
        p = seed.pattern
        kind = _abi.PATTERN_KINDS[p.type]
        center = list(p.center) + [0.0] * (3 - len(p.center))
        blobs = []
        if p.type == "MultiBlob":
            blobs = [_abi.HpBlob((ctypes.c_float * 3)(b.center[0], b.center[1], 0.0), b.radius, b.amplitude,
                                 b.channel) for b in p.blobs]
        elif p.type == "MultiSphere":
            blobs = [_abi.HpBlob((ctypes.c_float * 3)(*s.center), s.radius, s.amplitude, s.channel)
                     for s in p.spheres]
        values = []
        if p.type == "Custom":
            values = [_abi.HpCustomValue(x, y, 0, c, v) for (x, y, c, v) in p.values]
        elif p.type == "Custom3D":
            values = [_abi.HpCustomValue(x, y, z, c, v) for (x, y, z, c, v) in p.values]
        self.blobs = (_abi.HpBlob * max(1, len(blobs)))(*blobs)
        self.values = (_abi.HpCustomValue * max(1, len(values)))(*values)
        ch = -1 if p.channel is None else int(p.channel)
        self.seed = _abi.HpSeed(kind, (ctypes.c_float * 3)(*center), p.radius, p.inner_radius, p.outer_radius,
                                p.major_radius, p.minor_radius, p.amplitude, ch, p.seed,
                                len(blobs), ctypes.cast(self.blobs, ctypes.POINTER(_abi.HpBlob)),
                                len(values), ctypes.cast(self.values, ctypes.POINTER(_abi.HpCustomValue)))


# ---------------------------------------------------------------------------------------
# The native JSON loader (hp_config_from_json / hp_seed_from_json / hp_load_config_and_seed):
# serde's reading of config.rs / seed.rs restated in C++; these helpers turn its C structs back
# into the Python types above.
# ---------------------------------------------------------------------------------------
_KIND_NAMES = {v: k for k, v in _abi.PATTERN_KINDS.items()}


def config_from_c(c) -> SimulationConfig:
    rings = [c.rings[i] for i in range(c.num_rings)]
    ks = []
    for i in range(c.num_kernels):
        k = c.kernels[i]
        rs = [RingConfig(r.amplitude, r.distance, r.width) for r in rings[k.ring_offset:k.ring_offset + k.ring_count]]
        ks.append(KernelConfig(k.radius, rs, k.weight, k.mu, k.sigma, k.source_channel, k.target_channel))
    return SimulationConfig(width=c.width, height=c.height, depth=c.depth, channels=c.channels, dt=c.dt,
                            kernel_radius=c.kernel_radius, kernels=ks,
                            flow=FlowConfig(c.flow.beta_a, c.flow.n, c.flow.distribution_size))


def seed_from_c(s) -> Seed:
    t = _KIND_NAMES[s.kind]
    blobs = [s.blobs[i] for i in range(s.num_blobs)]
    vals = [s.values[i] for i in range(s.num_values)]
    if t in ("GaussianBlob", "GaussianSphere"):
        n = 2 if t == "GaussianBlob" else 3
        p = Pattern(t, center=tuple(s.center[:n]), radius=s.radius, amplitude=s.amplitude, channel=s.channel)
    elif t == "MultiBlob":
        p = Pattern(t, blobs=[BlobSpec(tuple(b.center[:2]), b.radius, b.amplitude, b.channel) for b in blobs])
    elif t == "MultiSphere":
        p = Pattern(t, spheres=[SphereSpec(tuple(b.center[:3]), b.radius, b.amplitude, b.channel) for b in blobs])
    elif t in ("Ring", "Shell"):
        n = 2 if t == "Ring" else 3
        p = Pattern(t, center=tuple(s.center[:n]), inner_radius=s.inner_radius, outer_radius=s.outer_radius,
                    amplitude=s.amplitude, channel=s.channel)
    elif t == "Torus3D":
        p = Pattern(t, center=tuple(s.center[:3]), major_radius=s.major_radius, minor_radius=s.minor_radius,
                    amplitude=s.amplitude, channel=s.channel)
    elif t == "Noise":
        p = Pattern(t, amplitude=s.amplitude, channel=None if s.channel < 0 else s.channel, seed=s.noise_seed)
    elif t == "Custom":
        p = Pattern(t, values=[(v.x, v.y, v.channel, v.value) for v in vals])
    else:
        p = Pattern(t, values=[(v.x, v.y, v.z, v.channel, v.value) for v in vals])
    return Seed(p)


def _native_error(st: int) -> Exception:
    msg = _abi.load_library().hp_json_last_error().decode()
    return _abi.HipError(st, msg)


def config_from_json_native(text: str) -> SimulationConfig:
    """serde_json::from_str::<SimulationConfig>, by the native loader."""
    lib = _abi.load_library()
    d = ctypes.c_void_p()
    st = lib.hp_config_from_json(text.encode(), ctypes.byref(d))
    if st != _abi.HP_OK:
        raise _native_error(st)
    try:
        cfg = config_from_c(lib.hp_config_doc_config(d).contents)
        e = _abi.HpEmbedding()
        lib.hp_config_doc_embedding(d, ctypes.byref(e))
        cfg.embedding = EmbeddingConfig(bool(e.enabled), float(e.mixing_temperature), bool(e.linear_mixing))
        return cfg
    finally:
        lib.hp_config_doc_free(d)


def embedding_from_json_native(text: str) -> dict:
    """The config's EmbeddingConfig (embedding.rs:326-345) as the native loader reads it."""
    lib = _abi.load_library()
    d = ctypes.c_void_p()
    st = lib.hp_config_from_json(text.encode(), ctypes.byref(d))
    if st != _abi.HP_OK:
        raise _native_error(st)
    try:
        e = _abi.HpEmbedding()
        lib.hp_config_doc_embedding(d, ctypes.byref(e))
        return {"enabled": bool(e.enabled), "mixing_temperature": e.mixing_temperature,
                "linear_mixing": bool(e.linear_mixing)}
    finally:
        lib.hp_config_doc_free(d)


def seed_from_json_native(text: str) -> Seed:
    """serde_json::from_str::<Seed>, by the native loader."""
    lib = _abi.load_library()
    d = ctypes.c_void_p()
    st = lib.hp_seed_from_json(text.encode(), ctypes.byref(d))
    if st != _abi.HP_OK:
        raise _native_error(st)
    try:
        return seed_from_c(lib.hp_seed_doc_seed(d).contents)
    finally:
        lib.hp_seed_doc_free(d)


def load_config_and_seed(config_path: str):
    """load_config_and_seed (main.rs:409-436), by the native loader: (SimulationConfig, Seed)."""
    lib = _abi.load_library()
    c, s = ctypes.c_void_p(), ctypes.c_void_p()
    st = lib.hp_load_config_and_seed(str(config_path).encode(), ctypes.byref(c), ctypes.byref(s))
    if st != _abi.HP_OK:
        raise _native_error(st)
    try:
        return config_from_c(lib.hp_config_doc_config(c).contents), seed_from_c(lib.hp_seed_doc_seed(s).contents)
    finally:
        lib.hp_config_doc_free(c)
        lib.hp_seed_doc_free(s)

"""CT-style volume ingest: RAW slice folders and TIFF stacks into device-resident
occupancy volumes (s2h_volume_probe / s2h_volume_read / s2h_vol_load /
s2h_vol_from_samples).

The input description mirrors the crop pipeline's `input` block
(CropInput / CropRawParams, src/config/crop.rs:4-22; load_input_volume,
src/pipeline/crop.rs:75-103): `type` raw|tiff, `path`, inclusive
`slice_start`/`slice_end` (-1 = first/last), and for RAW `width`, `height`,
`bits`, `signed`, `byte_order` (little|le|big|be).  Samples are z-major as
Volume3D (src/io/volume.rs:45-53); the occupancy grid keeps the axes:
S2 (x, y, z) = file (x, y, slice), dims = (width, height, depth).

Which samples are occupied is a Phase: a threshold window lo <= v <= hi, or
foreground v != background with the background given or detected as the modal
boundary value (detect_background_mode, src/pipeline/crop.rs:338-373).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from .s2 import Context, Volume

_KINDS = {"raw": 0, "tiff": 1, "tif": 1}
_DTYPES = {(8, False): "u1", (8, True): "i1", (16, False): "u2", (16, True): "i2", (32, False): "u4",
           (32, True): "i4"}


def parse_byte_order(value: Optional[str]) -> bool:
    """parse_byte_order (src/pipeline/crop.rs:34-47); returns True for big endian."""
    v = (value if value is not None else "little").strip().lower()
    if v in ("little", "le"):
        return False
    if v in ("big", "be"):
        return True
    raise ValueError(f"Invalid byte_order: {v}. Expected little or big")


@dataclass
class VolumeSource:
    path: str
    kind: str = "raw"
    slice_start: int = -1
    slice_end: int = -1
    width: int = 0
    height: int = 0
    bits: int = 8
    signed: bool = False
    byte_order: str = "little"

    @staticmethod
    def from_config(cfg: dict) -> "VolumeSource":
        """The crop config's `input` mapping (src/config/crop.rs:4-22)."""
        kind = str(cfg["type"]).strip().lower()
        if kind not in _KINDS:
            raise ValueError(f"Unsupported input.type: {kind}. Expected raw or tiff")
        src = VolumeSource(path=str(cfg["path"]), kind=kind,
                           slice_start=int(cfg.get("slice_start", -1) if cfg.get("slice_start") is not None else -1),
                           slice_end=int(cfg.get("slice_end", -1) if cfg.get("slice_end") is not None else -1))
        if _KINDS[kind] == 0:
            raw = cfg.get("raw")
            if raw is None:
                raise ValueError("input.raw is required when input.type=raw (width/height/bits/signed/byte_order)")
            src.width, src.height = int(raw["width"]), int(raw["height"])
            src.bits, src.signed = int(raw["bits"]), bool(raw["signed"])
            src.byte_order = raw.get("byte_order") or "little"
        return src

    def _c(self):
        s = L.VolumeSource()
        self._path_bytes = str(self.path).encode()
        s.path = self._path_bytes
        s.kind = _KINDS[self.kind.strip().lower()]
        s.slice_start, s.slice_end = int(self.slice_start), int(self.slice_end)
        s.width, s.height, s.bits = int(self.width), int(self.height), int(self.bits)
        s.is_signed = int(bool(self.signed))
        s.big_endian = int(parse_byte_order(self.byte_order))
        return s


@dataclass
class Phase:
    mode: str = "foreground_auto"   # "range" | "foreground" | "foreground_auto"
    lo: int = 0
    hi: int = 0
    background: int = 0

    @staticmethod
    def threshold(lo: int, hi: Optional[int] = None) -> "Phase":
        return Phase("range", int(lo), int(hi) if hi is not None else (1 << 63) - 1)

    @staticmethod
    def foreground(background: Optional[int] = None) -> "Phase":
        return Phase("foreground_auto") if background is None else Phase("foreground", background=int(background))

    def _c(self):
        p = L.Phase()
        p.mode = {"range": 0, "foreground": 1, "foreground_auto": 2}[self.mode]
        p.lo, p.hi, p.background = int(self.lo), int(self.hi), int(self.background)
        return p


@dataclass
class VolumeInfo:
    width: int
    height: int
    depth: int
    bits: int
    signed: bool
    big_endian: bool

    @property
    def dtype(self) -> np.dtype:
        return np.dtype(_DTYPES[(self.bits, self.signed)]).newbyteorder(">" if self.big_endian else "<")

    @staticmethod
    def _from(i) -> "VolumeInfo":
        return VolumeInfo(int(i.width), int(i.height), int(i.depth), int(i.bits), bool(i.is_signed),
                          bool(i.big_endian))


@dataclass
class LoadedVolume:
    volume: Volume
    info: VolumeInfo
    background: Optional[int]
    occupied: int


def probe(src: VolumeSource) -> VolumeInfo:
    info = L.VolumeInfo()
    cs = src._c()
    L.check(L.lib().s2h_volume_probe(C.byref(cs), C.byref(info)))
    return VolumeInfo._from(info)


def read(src: VolumeSource) -> np.ndarray:
    """The samples on the host, shape (depth, height, width), in the staged dtype."""
    info = probe(src)
    out = np.empty((info.depth, info.height, info.width), dtype=info.dtype)
    got = L.VolumeInfo()
    cs = src._c()
    L.check(L.lib().s2h_volume_read(C.byref(cs), out.ctypes.data_as(C.c_void_p), out.nbytes, C.byref(got)))
    return out


def load(ctx: Context, src: VolumeSource, phase: Phase) -> LoadedVolume:
    """Read, upload and bit-pack a RAW folder or TIFF stack on the device."""
    h = C.c_void_p()
    info = L.VolumeInfo()
    bg = C.c_int64(0)
    occ = C.c_uint64(0)
    cs, cp = src._c(), phase._c()
    L.check(L.lib().s2h_vol_load(ctx.handle, C.byref(cs), C.byref(cp), C.byref(h), C.byref(info), C.byref(bg),
                                 C.byref(occ)))
    return LoadedVolume(Volume(ctx, h), VolumeInfo._from(info), bg.value if phase.mode != "range" else None,
                        occ.value)


def from_samples(ctx: Context, samples: np.ndarray, phase: Phase) -> LoadedVolume:
    """samples: z-major array (depth, height, width) of int8/uint8/int16/uint16/int32/uint32."""
    a = np.asarray(samples)
    if a.ndim != 3:
        raise ValueError("samples must be (depth, height, width)")
    key = (a.dtype.itemsize * 8, a.dtype.kind == "i")
    if a.dtype.kind not in "iu" or key not in _DTYPES:
        raise ValueError(f"unsupported sample type {a.dtype}")
    big = a.dtype.byteorder == ">" or (a.dtype.byteorder == "=" and not np.little_endian)
    a = np.ascontiguousarray(a)
    info = L.VolumeInfo()
    info.depth, info.height, info.width = a.shape
    info.bits, info.is_signed, info.big_endian = key[0], int(key[1]), int(big and a.dtype.itemsize > 1)
    h = C.c_void_p()
    bg = C.c_int64(0)
    occ = C.c_uint64(0)
    cp = phase._c()
    L.check(L.lib().s2h_vol_from_samples(ctx.handle, a.ctypes.data_as(C.c_void_p), a.nbytes, C.byref(info),
                                         C.byref(cp), C.byref(h), C.byref(bg), C.byref(occ)))
    return LoadedVolume(Volume(ctx, h), VolumeInfo._from(info), bg.value if phase.mode != "range" else None,
                        occ.value)

"""Training checkpoints as a version-1 PLY manifest (photara/src/splat/checkpoint.cpp).

A checkpoint is two files in one directory:
  * the model, a 3DGS PLY named ``<stem>.model.<clock>.ply`` (save_gaussians_ply);
  * the manifest at the checkpoint path, a short text file (checkpoint.cpp:266-287):
        PHOTARA_PLY_CHECKPOINT
        1 <iteration> <total_iterations> <gaussian_count> <fingerprint>
        "<model file name>"                      (C++ std::quoted: \\ and " escaped)
        <valid> <center x3> <half_extent x3> <axes x9>   (floats as %.9g)

Writing follows the reference's order: the model first, then the manifest into ``<path>.tmp``,
then an fsync of the model and its directory, then an atomic rename over the manifest; on any
failure the partial files are removed and the previous generation stays valid. After a successful
publication the previous generation's model is deleted when this checkpoint manages it (same
directory, named ``<stem>.model.*.ply``). A lease (``<path>.lock`` under an exclusive
non-blocking flock) keeps two writers apart.

``training_checkpoint_info_bytes`` / ``decode_training_checkpoint_info`` are the binary metadata
record the reference stores in project archives (checkpoint.cpp:179-190, 97-124). The project
archive container itself (project/archive.hpp) is not part of this library.
"""
from __future__ import annotations

import errno
import fcntl
import math
import os
import re
import struct
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Optional, Union

from .ply import load_gaussians_ply, save_gaussians_ply
from .rasterizer import GaussianModel

VERSION = 1
MODEL_MAGIC = b"PHOTARA_PLY_CHECKPOINT\n"
MAX_GAUSSIANS = 50_000_000
_INFO = struct.Struct("<IIIQQ?3f3f9f")  # version, iteration, total, count, fingerprint, valid, ...

PathLike = Union[str, os.PathLike]


@dataclass
class FocusBounds:
    """mvs::OrientedBoundingBox: a centre, positive half extents and a 3x3 axis matrix (row-major
    as stored)."""
    valid: bool = False
    center: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    half_extent: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    axes: List[float] = field(default_factory=lambda: [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0])


@dataclass
class TrainingCheckpointInfo:
    """TrainingCheckpointInfo (photara/include/splat/checkpoint.hpp:28-35)."""
    iteration: int = 0
    total_iterations: int = 0
    gaussian_count: int = 0
    fingerprint: int = 0
    focus_bounds: FocusBounds = field(default_factory=FocusBounds)
    model_path: Optional[Path] = None


def _check_bounds(b: FocusBounds) -> None:
    if b.valid and (not all(math.isfinite(v) for v in list(b.center) + list(b.half_extent) + list(b.axes))
                    or any(v <= 0.0 for v in b.half_extent)):
        raise RuntimeError("Invalid checkpoint focus bounds")


def _check_counts(info: TrainingCheckpointInfo) -> bool:
    return bool(info.total_iterations) and info.iteration <= info.total_iterations and \
        bool(info.gaussian_count) and info.gaussian_count <= MAX_GAUSSIANS


def training_checkpoint_info_bytes(info: TrainingCheckpointInfo) -> bytes:
    """The packed metadata record (checkpoint.cpp:179-190): u32 version, u32 iteration, u32 total,
    u64 count, u64 fingerprint, bool valid, float centre[3], half extent[3], axes[9]."""
    b = info.focus_bounds
    return _INFO.pack(VERSION, int(info.iteration), int(info.total_iterations), int(info.gaussian_count),
                      int(info.fingerprint), bool(b.valid), *map(float, b.center), *map(float, b.half_extent),
                      *map(float, b.axes))


def decode_training_checkpoint_info(data: bytes) -> TrainingCheckpointInfo:
    """Inverse of training_checkpoint_info_bytes, with decode_info's checks (checkpoint.cpp:97-124):
    version 1, exact length, finite bounds with positive extents when valid, 0 < total,
    iteration <= total, 0 < count <= 50,000,000."""
    if len(data) < 4:
        raise RuntimeError("Truncated training checkpoint")
    if struct.unpack_from("<I", data)[0] != VERSION:
        raise RuntimeError("Unsupported training checkpoint version")
    if len(data) < _INFO.size:
        raise RuntimeError("Truncated training checkpoint")
    if len(data) > _INFO.size:
        raise RuntimeError("Unexpected checkpoint data")
    v = _INFO.unpack(data)
    info = TrainingCheckpointInfo(v[1], v[2], v[3], v[4],
                                  FocusBounds(bool(v[5]), list(v[6:9]), list(v[9:12]), list(v[12:21])))
    _check_bounds(info.focus_bounds)
    if not _check_counts(info):
        raise RuntimeError("Invalid training checkpoint metadata")
    return info


def _f32(v) -> float:
    return struct.unpack("<f", struct.pack("<f", float(v)))[0]


def _quote(s: str) -> str:
    return '"' + s.replace("\\", "\\\\").replace('"', '\\"') + '"'


_NUM = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?\Z")


class _Tokens:
    """Whitespace-separated reads in the manner of std::istream >> (and >> std::quoted)."""

    def __init__(self, text: str):
        self.s, self.i = text, 0

    def _skip(self):
        while self.i < len(self.s) and self.s[self.i].isspace():
            self.i += 1

    def word(self) -> str:
        self._skip()
        j = self.i
        while j < len(self.s) and not self.s[j].isspace():
            j += 1
        if j == self.i:
            raise ValueError("missing field")
        w, self.i = self.s[self.i:j], j
        return w

    def unsigned(self, bits: int) -> int:
        w = self.word()
        if not w.isdigit() or int(w) >= 1 << bits:
            raise ValueError(w)
        return int(w)

    def number(self) -> float:
        w = self.word()
        if not _NUM.match(w):
            raise ValueError(w)
        return float(w)

    def quoted(self) -> str:
        self._skip()
        if self.i >= len(self.s) or self.s[self.i] != '"':
            return self.word()
        out, j = [], self.i + 1
        while j < len(self.s):
            c = self.s[j]
            if c == "\\" and j + 1 < len(self.s):
                out.append(self.s[j + 1])
                j += 2
                continue
            if c == '"':
                self.i = j + 1
                return "".join(out)
            out.append(c)
            j += 1
        raise ValueError("unterminated quoted string")


def read_model_checkpoint(path: PathLike) -> Optional[TrainingCheckpointInfo]:
    """The manifest at `path` (checkpoint.cpp:127-155): None when the file does not start with the
    magic line; RuntimeError when it does but a field is missing or invalid (schema 1, 0 < total,
    iteration <= total, 0 < count <= 50,000,000, a model file name, valid focus bounds). A relative
    model path is taken relative to the manifest's directory."""
    path = Path(path)
    with open(path, "rb") as f:
        head = f.read(len(MODEL_MAGIC))
        if head != MODEL_MAGIC:
            return None
        text = f.read(1 << 16).decode("utf-8", errors="strict")
    t = _Tokens(text)
    try:
        schema = t.unsigned(32)
        info = TrainingCheckpointInfo(t.unsigned(32), t.unsigned(32), t.unsigned(64), t.unsigned(64))
        model = t.quoted()
        valid = t.word()
        if valid not in ("0", "1"):
            raise ValueError(valid)
        nums = [t.number() for _ in range(15)]
    except ValueError:
        raise RuntimeError("Invalid PLY checkpoint manifest") from None
    if schema != 1 or not _check_counts(info) or not model:
        raise RuntimeError("Invalid PLY checkpoint manifest")
    info.focus_bounds = FocusBounds(valid == "1", nums[0:3], nums[3:6], nums[6:15])
    _check_bounds(info.focus_bounds)
    mp = Path(model)
    info.model_path = mp if mp.is_absolute() else path.parent / mp
    return info


def _managed(checkpoint: Path, model: Path) -> bool:
    """checkpoint.cpp:157-163: the model sits beside the manifest as <stem>.model.*.ply."""
    return (os.path.normpath(model.parent) == os.path.normpath(checkpoint.parent)
            and model.name.startswith(checkpoint.stem + ".model.") and model.suffix == ".ply")


def _fsync_file_and_dir(path: Path) -> None:
    fd = os.open(path, os.O_RDONLY)
    try:
        os.fsync(fd)
    finally:
        os.close(fd)
    dfd = os.open(path.parent if str(path.parent) else ".", os.O_RDONLY | os.O_DIRECTORY)
    try:
        os.fsync(dfd)
    finally:
        os.close(dfd)


def inspect_training_checkpoint(path: PathLike) -> Optional[TrainingCheckpointInfo]:
    """The manifest's info when it is valid and its model file exists, else None (never raises)."""
    path = Path(path)
    if not str(path) or not path.exists():
        return None
    try:
        info = read_model_checkpoint(path)
    except (OSError, RuntimeError, UnicodeDecodeError):
        return None
    if info is None or not info.model_path.is_file():
        return None
    return info


def checkpoint_model(path: PathLike, device="cpu") -> GaussianModel:
    """The model a manifest references (checkpoint.cpp:218-224)."""
    info = read_model_checkpoint(path)
    if info is None:
        raise RuntimeError("Unsupported training checkpoint format; start a new training run.")
    return load_gaussians_ply(info.model_path, device)


def save_model_checkpoint(path: PathLike, info: TrainingCheckpointInfo, model: GaussianModel,
                          render_method: str = "3dgs", filter_mode: str = "classic") -> Path:
    """Publish one checkpoint generation (detail::save_model_checkpoint, checkpoint.cpp:256-307).
    The caller holds the lease. Returns the new model file's path."""
    path = Path(path)
    if str(path.parent):
        path.parent.mkdir(parents=True, exist_ok=True)
    previous = inspect_training_checkpoint(path)
    model_path = path.parent / f"{path.stem}.model.{time.monotonic_ns()}.ply"
    temporary = path.with_name(path.name + ".tmp")
    created = False
    try:
        save_gaussians_ply(model, model_path, render_method, filter_mode)
        b = info.focus_bounds
        # the bounds are float32 in the reference, printed at precision 9
        floats = " ".join("%.9g" % _f32(v) for v in list(b.center) + list(b.half_extent) + list(b.axes))
        body = (f"1 {int(info.iteration)} {int(info.total_iterations)} {int(info.gaussian_count)} "
                f"{int(info.fingerprint)}\n{_quote(model_path.name)}\n{1 if b.valid else 0} {floats}\n")
        try:
            out = open(temporary, "wb")
        except OSError as e:
            raise RuntimeError("Cannot write PLY checkpoint manifest") from e
        created = True
        with out:
            out.write(MODEL_MAGIC + body.encode("utf-8"))
        _fsync_file_and_dir(model_path)
        os.replace(temporary, path)
    except BaseException:
        if created:
            try:
                os.remove(temporary)
            except OSError:
                pass
        try:
            os.remove(model_path)
        except OSError:
            pass
        raise
    if previous is not None and _managed(path, previous.model_path):
        try:
            os.remove(previous.model_path)
        except OSError:
            pass
    return model_path


def remove_training_checkpoint(path: PathLike) -> None:
    """Delete a manifest and the model it manages (checkpoint.cpp:238-253), under the lease."""
    path = Path(path)
    if not str(path) or not path.exists():
        return
    with TrainingCheckpointLease(path):
        info = read_model_checkpoint(path)
        if info is None:
            raise RuntimeError("Unsupported training checkpoint format")
        os.remove(path)
        if _managed(path, info.model_path):
            try:
                os.remove(info.model_path)
            except FileNotFoundError:
                pass


class TrainingCheckpointLease:
    """Exclusive lease on a checkpoint path: ``<path>.lock`` held under a non-blocking flock
    (checkpoint.cpp:310-331). A second lease on the same path fails while one is held."""

    def __init__(self, path: PathLike):
        self.path = Path(path)
        self.fd = -1
        if not str(self.path):
            return
        if str(self.path.parent):
            self.path.parent.mkdir(parents=True, exist_ok=True)
        lock = self.path.with_name(self.path.name + ".lock")
        fd = os.open(lock, os.O_CREAT | os.O_RDWR, 0o600)
        try:
            fcntl.flock(fd, fcntl.LOCK_EX | fcntl.LOCK_NB)
        except OSError as e:
            os.close(fd)
            if e.errno in (errno.EWOULDBLOCK, errno.EAGAIN, errno.EACCES):
                raise RuntimeError("Training checkpoint is in use by another process or cannot be opened") from None
            raise
        self.fd = fd

    def release(self) -> None:
        if self.fd >= 0:
            os.close(self.fd)
            self.fd = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()

    def __del__(self):
        self.release()


def training_checkpoint_in_use(path: PathLike) -> bool:
    """True while another lease on `path` is held (checkpoint.cpp:192-201)."""
    if not str(path):
        return False
    try:
        TrainingCheckpointLease(path).release()
        return False
    except RuntimeError:
        return True

"""3DGS PLY model files: save_gaussians_ply / load_gaussians_ply
(photara/src/splat/trainer.cpp:2977-3350).

Format: binary little-endian float vertices. The properties are, in order: x y z nx ny nz
f_dc_0..2 f_rest_0..3(K-1)-1 opacity scale_0..2 rot_0..3 [filter_3D]. f_rest holds all remaining
SH bases for R, then G, then B; opacity is the logit, scale the log-scale and rot the scalar-first
quaternion.

The writer emits the reference's header lines, including its render-method and filter-mode
comments, and refuses non-finite parameters. The reader locates properties by name, so any order
and extra float properties are accepted. It infers the SH degree (0..3) from the f_rest count,
maps an infinite opacity logit to +/-104, and rejects non-finite parameters, truncated payloads
and non-float or non-binary files. All errors raise RuntimeError, like the reference's
std::runtime_error.

Optional learned normal features (gaussian_features_0..3) follow filter_3D. This library does not
render the reference's normal channel, so they travel beside the model: ``save_gaussians_ply``
takes them as an argument, and ``load_gaussians_ply(..., return_extras=True)`` returns them (for
a file without them, the reference's smallest-scale-axis seed) with the render method and filter
mode of the header comments.
"""
from __future__ import annotations

import math
import os
from typing import Optional, Union

import numpy as np
import torch

from .rasterizer import GaussianModel

_RENDER_METHODS = ("3dgs", "3dgut")
_FILTER_MODES = ("classic", "mip2d", "mip3d", "mip")


def _host(t: Optional[torch.Tensor]) -> Optional[np.ndarray]:
    return None if t is None else t.detach().to("cpu", torch.float32).contiguous().numpy()


def save_gaussians_ply(model: GaussianModel, path: Union[str, os.PathLike], render_method: str = "3dgs",
                       filter_mode: str = "classic", normal_features: Optional[torch.Tensor] = None) -> None:
    """Write ``model`` as a 3DGS PLY; ``normal_features`` ([N,4]) adds gaussian_features_0..3."""
    if render_method not in _RENDER_METHODS:
        raise ValueError(f"unknown render method {render_method!r}")
    if filter_mode not in _FILTER_MODES:
        raise ValueError(f"unknown filter mode {filter_mode!r}")
    n = model.size()
    means, log_scales, quats = _host(model.means), _host(model.log_scales), _host(model.quaternions)
    logits, sh = _host(model.opacity_logits).reshape(n), _host(model.sh)
    filt = _host(model.filter_3d)
    has_filter = filt is not None and filt.size == n
    normals = _host(normal_features)
    has_normals = normals is not None and normals.shape == (n, 4)
    for name, arr in (("means", means), ("log_scales", log_scales), ("quaternions", quats),
                      ("opacity_logits", logits), ("SH", sh)) + ((("normal_features", normals),) if has_normals else ()) \
            + ((("filter_3D", filt),) if has_filter else ()):
        bad = np.flatnonzero(~np.isfinite(arr.reshape(-1)))
        if bad.size:
            raise RuntimeError(f"Refusing to write non-finite splat parameter {name} at scalar index {int(bad[0])}")
    bases = sh.shape[1]
    header = ["ply", "format binary_little_endian 1.0",
              "comment Photara splat (GGGS-derived geometry; 3DGS-compatible SH layout)",
              f"comment photara_render_method {render_method}", f"comment photara_filter_mode {filter_mode}",
              f"element vertex {n}"]
    props = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    props += [f"f_rest_{i}" for i in range((bases - 1) * 3)]
    props += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    if has_filter:
        props.append("filter_3D")
    if has_normals:
        props += [f"gaussian_features_{i}" for i in range(4)]
    header += [f"property float {p}" for p in props] + ["end_header"]
    rows = np.empty((n, len(props)), np.float32)
    rows[:, 0:3] = means
    rows[:, 3:6] = 0.0
    rows[:, 6:9] = sh[:, 0, :]
    rest = sh[:, 1:, :].transpose(0, 2, 1).reshape(n, -1)  # R bases, then G, then B
    c = 9 + rest.shape[1]
    rows[:, 9:c] = rest
    rows[:, c] = logits
    rows[:, c + 1:c + 4] = log_scales
    rows[:, c + 4:c + 8] = quats
    if has_filter:
        rows[:, c + 8] = filt.reshape(n)
    if has_normals:
        rows[:, len(props) - 4:] = normals
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        f.write(rows.astype("<f4", copy=False).tobytes())


def _header(data: bytes, path) -> tuple:
    """Header lines up to ``end_header`` (a trailing CR stripped) and the payload offset."""
    lines, pos = [], 0
    while True:
        nl = data.find(b"\n", pos)
        if nl < 0:
            raise RuntimeError(f"Invalid Gaussian PLY header: {path}")
        line = data[pos:nl].decode("ascii", errors="replace")
        pos = nl + 1
        if line.endswith("\r"):
            line = line[:-1]
        if not lines and line != "ply":
            raise RuntimeError(f"Invalid Gaussian PLY header: {path}")
        if line == "end_header":
            return lines, pos
        lines.append(line)


def load_gaussians_ply(path: Union[str, os.PathLike], device: Union[str, torch.device] = "cpu",
                       return_extras: bool = False):
    """Read a 3DGS PLY. With ``return_extras`` also returns a dict with ``render_method``,
    ``filter_mode`` and ``normal_features`` ([N,4] float32: the file's gaussian_features_0..3, or
    for a file without them the unit axis of the smallest scale with w = 1)."""
    with open(path, "rb") as f:
        data = f.read()
    lines, start = _header(data, path)
    payload = memoryview(data)[start:]
    count, binary_le, in_vertex = 0, False, False
    render_method, filter_mode = "3dgs", "classic"
    props = []
    for line in lines[1:]:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "comment" and len(tok) >= 2:
            value = tok[2] if len(tok) > 2 else ""
            if tok[1] == "photara_render_method":
                if value not in _RENDER_METHODS:
                    raise RuntimeError(f"Unknown Gaussian PLY rendering method: {value}")
                render_method = value
            elif tok[1] == "photara_filter_mode":
                if value not in _FILTER_MODES:
                    raise RuntimeError(f"Unknown Gaussian PLY filter mode: {value}")
                filter_mode = value
        elif tok[0] == "format":
            binary_le = len(tok) > 1 and tok[1] == "binary_little_endian"
        elif tok[0] == "element":
            in_vertex = len(tok) > 1 and tok[1] == "vertex"
            if in_vertex:
                try:
                    count = int(tok[2])
                except (IndexError, ValueError):
                    count = 0
        elif tok[0] == "property" and in_vertex:
            if len(tok) < 3 or tok[1] not in ("float", "float32"):
                raise RuntimeError(f"Gaussian PLY requires float vertex properties: {path}")
            props.append(tok[2])
    if not binary_le or count == 0 or not props:
        raise RuntimeError(f"Gaussian PLY must contain binary little-endian vertices: {path}")
    index = {p: i for i, p in enumerate(props)}

    def need(name):
        if name not in index:
            raise RuntimeError(f"Gaussian PLY is missing property {name}: {path}")
        return index[name]

    rest = 0
    while f"f_rest_{rest}" in index:
        rest += 1
    if rest % 3:
        raise RuntimeError(f"Gaussian PLY has an invalid SH property count: {path}")
    bases = 1 + rest // 3
    degree = int(round(math.sqrt(bases))) - 1
    if (degree + 1) ** 2 != bases or degree > 3:
        raise RuntimeError(f"Gaussian PLY has unsupported SH degree: {path}")
    row_bytes = 4 * len(props)
    if count > len(payload) // row_bytes:
        raise RuntimeError(f"Gaussian PLY ended inside vertex data: {path}")
    rows = np.frombuffer(payload, dtype="<f4", count=count * len(props)).reshape(count, len(props))
    col = lambda names: rows[:, [need(x) for x in names]].astype(np.float32)
    means = col(["x", "y", "z"])
    log_scales = col(["scale_0", "scale_1", "scale_2"])
    quats = col(["rot_0", "rot_1", "rot_2", "rot_3"])
    logits = col(["opacity"])
    dc = col(["f_dc_0", "f_dc_1", "f_dc_2"])
    sh = np.empty((count, bases, 3), np.float32)
    sh[:, 0, :] = dc
    if bases > 1:
        r = col([f"f_rest_{i}" for i in range(rest)]).reshape(count, 3, bases - 1)
        sh[:, 1:, :] = r.transpose(0, 2, 1)
    filt = col(["filter_3D"]) if "filter_3D" in index else None
    feature_names = [f"gaussian_features_{i}" for i in range(4)]
    if all(x in index for x in feature_names):
        normals = col(feature_names)
    else:
        normals = _smallest_axis_normals(log_scales, quats)
    # a saturated opacity may be stored as an infinite logit: the float32 sigmoid endpoints
    logits = np.where(np.isinf(logits), np.copysign(np.float32(104.0), logits), logits).astype(np.float32)
    for arr in (means, log_scales, quats, logits, sh, normals) + ((filt,) if filt is not None else ()):
        if not np.isfinite(arr).all():
            raise RuntimeError(f"Gaussian PLY contains non-finite parameters: {path}")
    dev = torch.device(device)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    model = GaussianModel(means=t(means), log_scales=t(log_scales), quaternions=t(quats), opacity_logits=t(logits),
                          sh=t(sh), filter_3d=None if filt is None else t(filt), sh_degree=degree)
    if not return_extras:
        return model
    return model, {"render_method": render_method, "filter_mode": filter_mode, "normal_features": t(normals)}


def _smallest_axis_normals(log_scales: np.ndarray, quats: np.ndarray) -> np.ndarray:
    """The reference's legacy-file normal field (trainer.cpp:3284-3309): the rotation's column
    for the smallest scale axis of the normalised quaternion, with w = 1."""
    q = quats.astype(np.float32)
    norm = np.sqrt((q * q).sum(axis=1, dtype=np.float32)).astype(np.float32)
    q = q * (np.float32(1.0) / np.maximum(norm, np.float32(1e-12)))[:, None]
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    one, two = np.float32(1.0), np.float32(2.0)
    cols = np.stack([
        np.stack([one - two * (y * y + z * z), two * (x * y + w * z), two * (x * z - w * y)], 1),
        np.stack([two * (x * y - w * z), one - two * (x * x + z * z), two * (y * z + w * x)], 1),
        np.stack([two * (x * z + w * y), two * (y * z - w * x), one - two * (x * x + y * y)], 1)], 1)
    axis = np.argmin(log_scales, axis=1)  # first minimum on ties
    out = np.ones((len(q), 4), np.float32)
    out[:, :3] = cols[np.arange(len(q)), axis]
    return out

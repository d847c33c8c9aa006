"""Scene I/O: Stanford PLY files to and from GaussianScene.

write_ply_ascii / read_ply_ascii follow the Mojo core
(src/scene/ply.mojo:33-136): the vertex columns are x y z f_dc_0 f_dc_1
f_dc_2 opacity scale_0 scale_1 scale_2 rot_0 rot_1 rot_2 rot_3 (rot_0 = w),
values are stored as the scene holds them (activated opacity, linear scale),
and the reader takes the first 14 tokens of every non-empty line after
end_header that has at least 14 of them, by position, with the higher SH
coefficients zero.  Numbers are parsed the way parse_float
(src/core/types.mojo:66-117) does it: digits accumulated in float32
(val * 10 + d before the point, val + d / div after it, div *= 10 per
digit), the sign applied last, and an exponent applied as
val * 10^(sign * e) with float32 pow.  The parse runs column-wise over all
tokens at once.

read_ply reads any PLY whose vertex element names its properties (ascii or
binary_little_endian; float, double and integer types), by property name:
x y z, f_dc_0..2, f_rest_0..44 (channel-major, as 3DGS training writes them:
f_rest[c*15 + k-1] is SH coefficient k of channel c), opacity, scale_0..2,
rot_0..3.  Binary files are memory-mapped.  activation="3dgs" applies the
sigmoid to opacity and exp to the scales (3DGS stores logits and log-scales);
the default takes the values as stored, as the Mojo reader does.
write_ply_binary writes that layout (binary_little_endian, all 48 SH
coefficients, values as stored), which read_ply reads back bit for bit: the
loader for resident serving scenes.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .scene import GaussianScene, SH_COEFFS

MOJO_COLUMNS = ("x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
                "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3")

_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2",
              "int16": "i2", "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4",
              "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4", "double": "f8",
              "float64": "f8"}


def _columns(scene: GaussianScene) -> dict:
    p = scene.planes
    cols = {"x": p[0], "y": p[1], "z": p[2], "scale_0": p[3], "scale_1": p[4], "scale_2": p[5],
            "rot_0": p[6], "rot_1": p[7], "rot_2": p[8], "rot_3": p[9], "opacity": p[10]}
    for c in range(3):
        cols[f"f_dc_{c}"] = scene.sh[:, c]
    for c in range(3):
        for k in range(1, 16):
            cols[f"f_rest_{c * 15 + k - 1}"] = scene.sh[:, 3 * k + c]
    return cols


def write_ply_ascii(scene: GaussianScene, path: str) -> None:
    """write_ply_ascii (src/scene/ply.mojo:33-72): header, then one line of
    the 14 Mojo columns per Gaussian (%.9g, which a correctly rounded reader
    maps back to the same float32)."""
    cols = _columns(scene)
    data = np.stack([cols[k] for k in MOJO_COLUMNS], axis=1).astype(np.float32)
    with open(path, "w", encoding="ascii") as f:
        f.write("ply\nformat ascii 1.0\n")
        f.write("comment Created by lumen_amd\n")
        f.write(f"element vertex {scene.count}\n")
        for k in MOJO_COLUMNS:
            f.write(f"property float {k}\n")
        f.write("end_header\n")
        if scene.count:
            np.savetxt(f, data.astype(np.float64), fmt="%.9g", delimiter=" ")


def parse_float_tokens(tokens: np.ndarray) -> np.ndarray:
    """parse_float (src/core/types.mojo:66-117) over an array of byte strings."""
    tokens = np.asarray(tokens, dtype=np.bytes_)
    n = tokens.size
    width = tokens.dtype.itemsize
    out = np.zeros(n, np.float32)
    if n == 0 or width == 0:
        return out
    chars = np.frombuffer(tokens.tobytes(), np.uint8).reshape(n, width)
    f32 = np.float32
    sign = np.ones(n, f32)
    val = np.zeros(n, f32)
    div = np.ones(n, f32)
    in_dec = np.zeros(n, bool)
    in_exp = np.zeros(n, bool)
    exp_sign = np.ones(n, f32)
    exp_val = np.zeros(n, f32)
    exp_first = np.zeros(n, bool)
    for j in range(width):
        c = chars[:, j]
        live = c != 0  # numpy pads short strings with NUL bytes
        skip = np.zeros(n, bool)
        if j == 0:
            minus = live & (c == 45)
            sign[minus] = f32(-1.0)
            skip = minus | (live & (c == 43))
        is_e = live & ~skip & ((c == 101) | (c == 69))
        in_exp |= is_e
        exp_first = np.where(is_e, True, exp_first)
        rest = live & ~skip & ~is_e
        # exponent digits (after an optional sign right after the 'e')
        ex = rest & in_exp
        first = ex & exp_first
        exp_sign[first & (c == 45)] = f32(-1.0)
        ex_skip = first & ((c == 45) | (c == 43))
        exp_first = np.where(ex, False, exp_first)
        d = ex & ~ex_skip & (c >= 48) & (c <= 57)
        exp_val = np.where(d, exp_val * f32(10.0) + (c.astype(f32) - f32(48.0)), exp_val)
        # mantissa
        mt = rest & ~in_exp
        in_dec |= mt & (c == 46)
        dig = mt & (c >= 48) & (c <= 57)
        digit = c.astype(f32) - f32(48.0)
        frac = dig & in_dec
        div = np.where(frac, div * f32(10.0), div)
        val = np.where(frac, val + digit / div, val)
        whole = dig & ~in_dec
        val = np.where(whole, val * f32(10.0) + digit, val)
    total = sign * val
    if in_exp.any():
        with np.errstate(over="ignore", under="ignore"):
            scale = np.power(f32(10.0), exp_sign * exp_val, dtype=f32)
        total = np.where(in_exp, total * scale, total).astype(f32)
    return total


def read_ply_ascii(path: str) -> GaussianScene:
    """read_ply_ascii (src/scene/ply.mojo:75-136)."""
    with open(path, "rb") as f:
        content = f.read()
    rows: List[List[bytes]] = []
    in_header = True
    for raw in content.split(b"\n"):
        line = raw.strip()
        if not line:
            continue
        if in_header:
            if line == b"end_header":
                in_header = False
            continue
        tok = line.split(b" ")
        if len(tok) >= 14:
            rows.append(tok[:14])
    n = len(rows)
    if n == 0:
        return GaussianScene(np.zeros((11, 0), np.float32), np.zeros((0, SH_COEFFS), np.float32))
    vals = parse_float_tokens(np.array(rows, dtype=np.bytes_).reshape(-1)).reshape(n, 14)
    x, y, z, d0, d1, d2, op, s0, s1, s2, rw, rx, ry, rz = vals.T
    planes = np.stack([x, y, z, s0, s1, s2, rw, rx, ry, rz, op])
    sh = np.zeros((n, SH_COEFFS), np.float32)
    sh[:, 0], sh[:, 1], sh[:, 2] = d0, d1, d2
    return GaussianScene(planes, sh)


def _read_header(f) -> Tuple[str, List[Tuple[str, int, List[Tuple[str, str]]]], int]:
    magic = f.readline()
    if magic.strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements: List[Tuple[str, int, List[Tuple[str, str]]]] = []
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        parts = line.decode("ascii", "replace").split()
        if not parts:
            continue
        if parts[0] == "format":
            fmt = parts[1]
        elif parts[0] == "element":
            elements.append((parts[1], int(parts[2]), []))
        elif parts[0] == "property":
            if not elements:
                raise ValueError("PLY property before any element")
            if parts[1] == "list":
                raise ValueError("PLY list properties are not supported in the vertex element")
            if parts[1] not in _PLY_TYPES:
                raise ValueError(f"unknown PLY property type {parts[1]}")
            elements[-1][2].append((parts[2], _PLY_TYPES[parts[1]]))
        elif parts[0] == "end_header":
            break
    if fmt is None:
        raise ValueError("PLY header has no format line")
    return fmt, elements, f.tell()


def read_ply(path: str, activation: str = "stored") -> GaussianScene:
    """A Gaussian scene from an ascii or binary_little_endian PLY (by property
    name; see the module docstring)."""
    with open(path, "rb") as f:
        fmt, elements, offset = _read_header(f)
    if not elements or elements[0][0] != "vertex":
        raise ValueError("the first PLY element must be 'vertex'")
    _, n, props = elements[0]
    names = [p[0] for p in props]
    for need in ("x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2",
                 "rot_0", "rot_1", "rot_2", "rot_3"):
        if need not in names:
            raise ValueError(f"PLY vertex element lacks '{need}'")
    if fmt == "binary_little_endian":
        dt = np.dtype([(name, "<" + t) for name, t in props])
        data = np.memmap(path, dtype=dt, mode="r", offset=offset, shape=(n,))
        col = lambda k: np.asarray(data[k], np.float32)  # noqa: E731
    elif fmt == "ascii":
        with open(path, "rb") as f:
            f.seek(offset)
            table = np.loadtxt(f, dtype=np.float64, max_rows=n, ndmin=2) if n else \
                np.zeros((0, len(props)))
        if table.shape != (n, len(props)):
            raise ValueError("PLY vertex rows do not match the header")
        col = lambda k: table[:, names.index(k)].astype(np.float32)  # noqa: E731
    else:
        raise ValueError(f"PLY format {fmt} is not supported")
    planes = np.stack([col("x"), col("y"), col("z"), col("scale_0"), col("scale_1"),
                       col("scale_2"), col("rot_0"), col("rot_1"), col("rot_2"), col("rot_3"),
                       col("opacity")])
    sh = np.zeros((n, SH_COEFFS), np.float32)
    for c in range(3):
        if f"f_dc_{c}" in names:
            sh[:, c] = col(f"f_dc_{c}")
        for k in range(1, 16):
            key = f"f_rest_{c * 15 + k - 1}"
            if key in names:
                sh[:, 3 * k + c] = col(key)
    if activation == "3dgs":
        planes[10] = 1.0 / (1.0 + np.exp(-planes[10].astype(np.float64)))
        planes[3:6] = np.exp(planes[3:6].astype(np.float64))
    elif activation != "stored":
        raise ValueError("activation must be 'stored' or '3dgs'")
    return GaussianScene(planes, sh)


def write_ply_binary(scene: GaussianScene, path: str) -> None:
    """binary_little_endian PLY with x y z, f_dc_0..2, f_rest_0..44, opacity,
    scale_0..2, rot_0..3 as float32 (values as stored)."""
    cols = _columns(scene)
    order = (["x", "y", "z"] + [f"f_dc_{c}" for c in range(3)] +
             [f"f_rest_{i}" for i in range(45)] + ["opacity", "scale_0", "scale_1", "scale_2",
                                                   "rot_0", "rot_1", "rot_2", "rot_3"])
    dt = np.dtype([(k, "<f4") for k in order])
    rec = np.empty(scene.count, dt)
    for k in order:
        rec[k] = cols[k]
    with open(path, "wb") as f:
        head = ["ply", "format binary_little_endian 1.0", f"element vertex {scene.count}"]
        head += [f"property float {k}" for k in order] + ["end_header"]
        f.write(("\n".join(head) + "\n").encode("ascii"))
        rec.tofile(f)

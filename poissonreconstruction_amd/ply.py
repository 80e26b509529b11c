"""PLY I/O of the reference (src/ply.rs), byte for byte.

read_oriented_points (src/ply.rs:13-71): ASCII only; header lines up to
`end_header`; `element vertex N` parsed strictly; the cloud counts as having
normals when any header line contains "nx" or "normal_x"; then every body
line with at least 6 whitespace-separated fields gives one point from its
first 6 fields (a field that is not a number reads as 0.0), whatever N says.

write_mesh (src/ply.rs:73-107): ASCII header with float x/y/z and
`property list uchar int vertex_indices`, vertices printed as Rust prints an
f64 (shortest round-trip digits, plain decimal, no exponent), faces as
"3 a b c".
"""
from __future__ import annotations

import re
from decimal import Decimal

import numpy as np

_RUST_FLOAT = re.compile(r"^[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf|infinity|nan)$", re.IGNORECASE)


def _rust_f64(tok: str) -> float:
    """str::parse::<f64>() with unwrap_or(0.0)."""
    if _RUST_FLOAT.match(tok) is None:
        return 0.0
    return float(tok)


def _rust_usize(tok: str):
    if re.fullmatch(r"\+?\d+", tok) is None:
        return None
    return int(tok)


def read_oriented_points(path: str) -> np.ndarray:
    """(n, 6) float64 [px, py, pz, nx, ny, nz]; raises ValueError with the reference's messages."""
    try:
        fh = open(path, "r", encoding="utf-8", newline=None)
    except OSError as e:
        raise ValueError(f"Cannot open {path}: {e}") from None
    with fh:
        has_normals = False
        while True:
            line = fh.readline()
            if line == "":
                raise ValueError("Unexpected end of file")
            line = line.rstrip("\r\n") if line.endswith("\n") else line
            if line.startswith("end_header"):
                break
            if line.startswith("element vertex "):
                if _rust_usize(line[len("element vertex "):]) is None:
                    raise ValueError("Invalid vertex count")
            elif "nx" in line or "normal_x" in line:
                has_normals = True
        if not has_normals:
            raise ValueError("PLY file does not contain normal data (nx, ny, nz)")
        rows = []
        try:
            for line in fh:
                parts = line.split()
                if len(parts) < 6:
                    continue
                rows.append([_rust_f64(p) for p in parts[:6]])
        except UnicodeDecodeError as e:
            raise ValueError(f"Read error: {e}") from None
    return np.array(rows, dtype=np.float64).reshape(-1, 6)


def rust_f64_display(x: float) -> str:
    """format!("{}", x) for an f64."""
    x = float(x)
    if x != x:
        return "NaN"
    if x in (float("inf"), float("-inf")):
        return "inf" if x > 0 else "-inf"
    t = format(Decimal(repr(x)), "f")
    if "." in t:
        t = t.rstrip("0").rstrip(".")
    return t


def write_mesh(path: str, vertices, triangles) -> None:
    V = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    F = np.asarray(triangles).reshape(-1, 3)
    try:
        fh = open(path, "w", encoding="utf-8", newline="\n")
    except OSError as e:
        raise ValueError(f"Cannot create {path}: {e}") from None
    with fh:
        fh.write(f"ply\nformat ascii 1.0\nelement vertex {len(V)}\nproperty float x\nproperty float y\nproperty float z\n"
                 f"element face {len(F)}\nproperty list uchar int vertex_indices\nend_header\n")
        fh.writelines(f"{rust_f64_display(a)} {rust_f64_display(b)} {rust_f64_display(c)}\n" for a, b, c in V.tolist())
        fh.writelines(f"3 {int(a)} {int(b)} {int(c)}\n" for a, b, c in F.tolist())

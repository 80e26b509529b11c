"""PLY I/O and the command line (src/ply.rs, src/main.rs) on the host."""
import os
import subprocess
import sys

import numpy as np
import pytest

from poissonreconstruction_amd import ply

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("x,s", [(1.0, "1"), (0.1, "0.1"), (1e-7, "0.0000001"), (1.5e21, "1500000000000000000000"),
                                 (-0.0, "-0"), (123.456, "123.456"), (-2.5, "-2.5"), (float("nan"), "NaN"),
                                 (float("inf"), "inf"), (float("-inf"), "-inf"), (1 / 3, "0.3333333333333333"),
                                 (5e-324, "0." + "0" * 323 + "5")])
def test_rust_f64_display(x, s):
    # Rust's Display for f64: shortest round-trip digits, never an exponent
    assert ply.rust_f64_display(x) == s


def write(tmp_path, text, name="c.ply"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


HDR = "ply\nformat ascii 1.0\nelement vertex {n}\nproperty float x\nproperty float y\nproperty float z\n" \
      "property float nx\nproperty float ny\nproperty float nz\nend_header\n"


def test_read_points_and_quirks(tmp_path):
    body = "0 0 0 0 0 1\n1 2 3 0.5 0.5 0.70710678\n7 8 9\nabc 1 2 3 4 5 extra\n3 4 5 6 7 8\n"
    pts = ply.read_oriented_points(write(tmp_path, HDR.format(n=2) + body))
    # every line with >= 6 fields, regardless of the declared count; a bad field reads 0.0
    assert pts.shape == (4, 6)
    assert np.array_equal(pts[2], [0, 1, 2, 3, 4, 5])
    assert np.array_equal(pts[3], [3, 4, 5, 6, 7, 8])


def test_read_errors(tmp_path):
    with pytest.raises(ValueError, match="does not contain normal data"):
        ply.read_oriented_points(write(tmp_path, "ply\nelement vertex 1\nproperty float x\nend_header\n1 2 3 4 5 6\n"))
    with pytest.raises(ValueError, match="Invalid vertex count"):
        ply.read_oriented_points(write(tmp_path, "ply\nelement vertex 3 \nproperty float nx\nend_header\n"))
    with pytest.raises(ValueError, match="Unexpected end of file"):
        ply.read_oriented_points(write(tmp_path, "ply\nproperty float nx\n"))
    with pytest.raises(ValueError, match="Cannot open"):
        ply.read_oriented_points(str(tmp_path / "missing.ply"))
    # "normal_x" (and any header line containing "nx") declares normals
    pts = ply.read_oriented_points(write(tmp_path, "ply\nproperty float normal_x\nend_header\n1 2 3 4 5 6\n"))
    assert pts.shape == (1, 6)


def test_write_mesh_format_and_round_trip(tmp_path):
    V = np.array([[0.1, 1.0, -2.5e-7], [1 / 3, 2.0, 3.0], [0.0, 0.5, 1e21]])
    F = np.array([[0, 1, 2], [2, 1, 0]], dtype=np.uint64)
    p = str(tmp_path / "m.ply")
    ply.write_mesh(p, V, F)
    lines = open(p).read().split("\n")
    assert lines[:10] == ["ply", "format ascii 1.0", "element vertex 3", "property float x", "property float y",
                          "property float z", "element face 2", "property list uchar int vertex_indices", "end_header",
                          "0.1 1 -0.00000025"]
    assert lines[10] == "0.3333333333333333 2 3"
    assert lines[12:14] == ["3 0 1 2", "3 2 1 0"]
    back = np.array([[float(t) for t in l.split()] for l in lines[9:12]])
    assert np.array_equal(back, V)


def cli(*args):
    return subprocess.run([sys.executable, "-m", "poissonreconstruction_amd", *args], capture_output=True, text=True, cwd=ROOT)


def test_cli_error_paths(tmp_path):
    r = cli("-i", str(tmp_path / "nope.ply"))
    assert r.returncode == 1 and r.stderr.startswith("Error: Cannot open")
    r = cli("-i", write(tmp_path, "ply\nproperty float x\nend_header\n"))
    assert r.returncode == 1 and "does not contain normal data" in r.stderr
    r = cli("-i", write(tmp_path, HDR.format(n=0)))
    assert r.returncode == 1 and r.stderr.strip() == "No points found."

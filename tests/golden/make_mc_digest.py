"""Write mc_tables_digest.json: SHA-256 digests of the reference's marching-cubes
tables, read as text from the reference crate.

Run in a container that holds the reference sources:
    python tests/golden/make_mc_digest.py [/path/to/reference]

It reads
  src/tri_table.inl      TRI_TABLE, 256 cases of up to 5 triangles (edge ids, -1 ends)
  src/fem_tree.rs:13-30  EDGE_TABLE, the cut-edge mask of every case
  src/fem_tree.rs:587-588  the cube corner offsets `co` and the edge corner pairs `ev`
and stores only digests of their canonical forms (see canonical_* in
tests/test_mc_oracle.py, which recomputes them from csrc/mc_cases.h), plus the
triangle count, so the fixture holds no table text.
"""
import hashlib
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def canonical_tri(cases):
    return ";".join(",".join(str(e) for e in c) for c in cases)


def canonical_edge_masks(masks):
    return ",".join(str(m) for m in masks)


def canonical_topology(corners, edges):
    return "|".join(",".join(str(v) for v in c) for c in corners) + "#" + "|".join(f"{a},{b}" for a, b in edges)


def sha(s):
    return hashlib.sha256(s.encode()).hexdigest()


def main(ref):
    tri_src = open(os.path.join(ref, "src", "tri_table.inl")).read()
    rows = re.findall(r"\[([-0-9,\s]+)\]", tri_src[tri_src.index("=") + 1:])
    cases = []
    for r in rows:
        vals = [int(v) for v in r.replace(" ", "").split(",") if v]
        assert len(vals) == 16
        cut = vals.index(-1) if -1 in vals else 16
        cases.append(vals[:cut])
    assert len(cases) == 256

    fem = open(os.path.join(ref, "src", "fem_tree.rs")).read()
    et = fem[fem.index("EDGE_TABLE"):]
    et = et[et.index("[", et.index("=")) + 1:et.index("];")]
    masks = [int(v, 16) for v in re.findall(r"0x[0-9a-fA-F]+", et)]
    assert len(masks) == 256

    co = re.search(r"let co: \[\[u32; 3\]; 8\] = \[(.*?)\];", fem).group(1)
    corners = [tuple(int(v) for v in g.split(",")) for g in re.findall(r"\[([0-9,]+)\]", co)]
    ev = re.search(r"let ev: \[\[usize; 2\]; 12\] = \[(.*?)\];", fem).group(1)
    edges = [tuple(int(v) for v in g.split(",")) for g in re.findall(r"\[([0-9,]+)\]", ev)]
    assert len(corners) == 8 and len(edges) == 12

    out = {
        "source": "src/tri_table.inl:1-258, src/fem_tree.rs:13-30, src/fem_tree.rs:587-588",
        "tri_table_sha256": sha(canonical_tri(cases)),
        "edge_table_sha256": sha(canonical_edge_masks(masks)),
        "topology_sha256": sha(canonical_topology(corners, edges)),
        "n_triangles": sum(len(c) // 3 for c in cases),
    }
    with open(os.path.join(HERE, "mc_tables_digest.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

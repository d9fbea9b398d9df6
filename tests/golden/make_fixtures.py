"""Extract the reference's unit-test data for the explicit path into JSON fixtures.

Reads (text only, nothing is executed) from the reference tree:
  input/testing/unit/nodes/nodes_tp.data          -> nodes_tp.json      (test_unit_nodes, tp)
  input/testing/unit/nodes/nodes_si.data          -> nodes_si.json      (test_unit_nodes, simplex: EQ/AO/WSH/WV)
  input/testing/unit/nodes/nodes_pyr.data         -> nodes_pyr.json     (test_unit_nodes, pyramid: GLe/GLoL/WV)
  input/testing/unit/operators/operators_tp.data  -> operators_tp.json  (test_unit_operators_tp.c:75-176)
  input/testing/unit/containers/matrix.data       -> matrix.json        (test_unit_containers, mm/mv/mm_diag)
  input/testing/unit/nodes/nodes_correspondence.data -> nodes_correspondence.json (test_unit_nodes,
                                                     face_correspondence: line_3/4, quad_9/16, tri_6/10)
The JSON files hold the numbers only (inputs and expected outputs); they are committed next to this
script.  Run: python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _tokens(path):
    with open(path) as f:
        return [ln.rstrip("\n") for ln in f]


def _read_matrix(lines, i):
    """At lines[i] == 'container Matrix_d' (or Multiarray_d); returns (layout, shape, values, next index)."""
    hdr = lines[i + 1].split()
    layout = hdr[0]
    if len(hdr) == 3:
        shape = [int(hdr[1]), int(hdr[2])]
    else:  # Multiarray: layout order ext...
        shape = [int(v) for v in hdr[2:]]
    n = 1
    for s in shape:
        n *= s
    vals = []
    j = i + 2
    while len(vals) < n:
        vals.extend(float(t) for t in lines[j].split())
        j += 1
    return layout, shape, vals, j


def _to_rowmajor(layout, shape, vals):
    """Matrix text rows always list the matrix row by row (both 'R' and 'C' files print rows)."""
    r, c = shape[0], shape[1] if len(shape) > 1 else 1
    return [vals[k * c:(k + 1) * c] for k in range(r)]


def parse_nodes(path):
    lines = _tokens(path)
    out = {}
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("Nodes_d"):
            name = ln
            has_w = int(lines[i + 2].split(":")[1])
            i += 3
            assert lines[i].strip() == "rst:"
            lay, shape, vals, i = _read_matrix(lines, i + 1)
            rst = _to_rowmajor(lay, shape, vals)
            w = None
            if has_w:
                assert lines[i].strip() == "w:", lines[i]
                lay, shape, vals, i = _read_matrix(lines, i + 1)
                w = [row[0] for row in _to_rowmajor(lay, shape, vals)]
            out[name] = {"rst": rst, "w": w}
            continue
        i += 1
    return out


def parse_operators(path):
    lines = _tokens(path)
    out = {}
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.endswith("_ops_tp:"):
            key = ln[:-1]
            ext = int(lines[i + 3].split()[1])
            i += 4
            mats = []
            for _ in range(ext):
                while lines[i].strip() != "container Matrix_d":
                    i += 1
                hdr = lines[i + 1].split()
                if len(hdr) < 3:          # an empty (NULL) sub-operator
                    mats.append(None)
                    i += 2
                    continue
                lay, shape, vals, i = _read_matrix(lines, i)
                mats.append(_to_rowmajor(lay, shape, vals))
            out[key] = mats
            continue
        if ln.endswith("_op_std:"):
            key = ln[:-1]
            lay, shape, vals, i = _read_matrix(lines, i + 1)
            out[key] = _to_rowmajor(lay, shape, vals)
            continue
        if ln.endswith("_b:") or ln.endswith("_c:"):
            key = ln[:-1]
            hdr = lines[i + 2].split()     # e.g. "C 2 15 2": column-major (rows, cols)
            rows, cols = int(hdr[2]), int(hdr[3])
            vals = []
            j = i + 3
            while len(vals) < rows * cols:
                vals.extend(float(t) for t in lines[j].split())
                j += 1
            # text lists one matrix row per line
            out[key] = [vals[k * cols:(k + 1) * cols] for k in range(rows)]
            i = j
            continue
        i += 1
    return out


def parse_matrix(path):
    lines = _tokens(path)
    out = {}
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.endswith("_Matrix:"):
            key = ln[:-1]
            lay, shape, vals, i = _read_matrix(lines, i + 1)
            out[key] = _to_rowmajor(lay, shape, vals)
            continue
        if ln.endswith("_Vector:"):
            key = ln[:-1]
            toks = lines[i + 2].split()
            n = int(toks[0])
            out[key] = [float(t) for t in toks[1:1 + n]]
            i += 3
            continue
        i += 1
    return out


def parse_correspondence(path):
    """The Multiarray_Vector_i blocks: name -> list of index vectors (one per permutation case)."""
    lines = _tokens(path)
    out = {}
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.endswith(":") and i + 1 < len(lines) and lines[i + 1].strip() == "container Multiarray_Vector_i":
            name = ln[:-1]
            n = int(lines[i + 3].split()[1])
            assert lines[i + 4].strip() == "ext_0/data"
            vecs = []
            for k in range(n):
                t = lines[i + 5 + k].split()
                vecs.append([int(v) for v in t[1:1 + int(t[0])]])
            out[name] = vecs
            i += 5 + n
        else:
            i += 1
    return out


def main(ref="/root/reference"):
    unit = os.path.join(ref, "input", "testing", "unit")
    fixtures = {
        "nodes_tp.json": parse_nodes(os.path.join(unit, "nodes", "nodes_tp.data")),
        "nodes_si.json": parse_nodes(os.path.join(unit, "nodes", "nodes_si.data")),
        "nodes_pyr.json": parse_nodes(os.path.join(unit, "nodes", "nodes_pyr.data")),
        "operators_tp.json": parse_operators(os.path.join(unit, "operators", "operators_tp.data")),
        "matrix.json": parse_matrix(os.path.join(unit, "containers", "matrix.data")),
        "nodes_correspondence.json": parse_correspondence(os.path.join(unit, "nodes", "nodes_correspondence.data")),
    }
    for name, data in fixtures.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, indent=0)
        print(name, sorted(data.keys()))


if __name__ == "__main__":
    main(*sys.argv[1:])

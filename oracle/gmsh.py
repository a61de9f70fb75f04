"""Restatement of the reference gmsh 2.2 ASCII reader (source/io.d:689-990).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Returns raw mesh arrays:
nodes (n,2) f64, elem_offsets u32, elem_nodes u32 (1-based), bnodes (nb,2) u32
(0-based), btags list[str], bgroup_start u64 -- the inputs of buildMesh.
"""
from types import SimpleNamespace

import numpy as np


def _clean(line):
    # io.d:662-666 readCleanLine: strip, chomp, detab, split on ' ', drop empties
    return [t for t in line.strip().expandtabs().split(" ") if t != ""]


def parse_gmsh(path):
    with open(path, "r") as f:
        lines = f.readlines()
    pos = 0

    def readline():
        nonlocal pos
        if pos >= len(lines):
            return None
        s = lines[pos]
        pos += 1
        return _clean(s)

    btags, nodes, elements = [], [], []
    bnodes_l, bgroups = [], []
    while True:
        sec = readline()
        while sec is not None and len(sec) == 0:
            sec = readline()
        if sec is None:
            break
        assert len(sec) == 1 and sec[0][0] == "$", "bad section header %r" % sec
        name = sec[0][1:]
        if name == "MeshFormat":
            fmt = readline()
            assert int(fmt[1]) == 0, "only ASCII gmsh is supported"
            assert "$End" in readline()[0]
        elif name == "PhysicalNames":
            n = int(readline()[0])
            for _ in range(n):
                ln = readline()
                assert len(ln) == 3
                if int(ln[0]) == 1:
                    btags.append(ln[2].strip('"'))
            assert "$End" in readline()[0]
        elif name == "Nodes":
            n = int(readline()[0])
            for i in range(n):
                ln = readline()
                assert int(ln[0]) == i + 1
                nodes.append([float(v) for v in ln[1:-1]])
            assert "$End" in readline()[0]
        elif name == "Elements":
            n = int(readline()[0])
            current = -1
            for i in range(n):
                ln = [int(v) for v in readline()]
                assert len(ln) > 4 and ln[0] == i + 1
                etype, ntags = ln[1], ln[2]
                assert ntags >= 1
                if etype == 1:
                    g = ln[3]
                    bn = [v - 1 for v in ln[3 + ntags:]]
                    if g != current:
                        current = g
                        bgroups.append([])
                    bgroups[-1].append(g)
                    bnodes_l.append(bn)
                elif etype in (2, 3):
                    elements.append(ln[3 + ntags:])
                else:
                    raise ValueError("Unsuported element type: %d" % etype)
            assert "$End" in readline()[0]
        else:
            while True:
                ln = readline()
                assert ln is not None, "Unexpected end of file"
                if ln and "$End" in ln[0]:
                    break

    # io.d:908-960 regroup boundary groups that gmsh split apart
    if len(bgroups) != len(btags):
        newbn, newbg = [], []
        for i, grp in enumerate(bgroups):
            cur = grp[0]
            old = next((k for k, g in enumerate(newbg) if g[0] == cur), -1)
            start = sum(len(g) for g in bgroups[:i])
            end = sum(len(g) for g in bgroups[: i + 1])
            if old >= 0:
                newend = sum(len(g) for g in newbg[: old + 1])
                newbg[old] = newbg[old] + grp
                newbn = newbn[:newend] + bnodes_l[start:end] + newbn[newend:]
            else:
                newbg.append(list(grp))
                newbn = newbn + bnodes_l[start:end]
        assert len(newbn) == len(bnodes_l)
        bnodes_l, bgroups = newbn, newbg
    assert len(bgroups) == len(btags), "Number of boundary groups does not match the number of boundary tags"

    # io.d:939-960 order boundary nodes by physical tag number i+1
    out_bn, starts = [], []
    for i in range(len(bgroups)):
        gi = next((k for k, g in enumerate(bgroups) if g[0] == i + 1), -1)
        assert gi >= 0, "Could not determine boundary group"
        end = sum(len(g) for g in bgroups[: gi + 1])
        start = sum(len(g) for g in bgroups[:gi])
        starts.append(len(out_bn))
        out_bn.extend(bnodes_l[start:end])

    # io.d:969-987 make every element counter-clockwise
    for i, el in enumerate(elements):
        area = 0.0
        m = len(el)
        for j in range(m):
            n1, n2 = el[j] - 1, el[(j + 1) % m] - 1
            area += nodes[n1][0] * nodes[n2][1] - nodes[n2][0] * nodes[n1][1]
        if area < 0:
            elements[i] = el[::-1]

    offs = np.zeros(len(elements) + 1, dtype=np.uint32)
    offs[1:] = np.cumsum([len(e) for e in elements])
    return SimpleNamespace(
        nodes=np.array(nodes, dtype=np.float64).reshape(-1, 2),
        elem_offsets=offs,
        elem_nodes=np.array([v for e in elements for v in e], dtype=np.uint32),
        bnodes=np.array(out_bn, dtype=np.uint32).reshape(-1, 2),
        btags=btags,
        bgroup_start=np.array(starts, dtype=np.uint64),
    )


def parse_gri(path):
    """Restatement of the XFlow .gri reader (source/io.d:992-1121)."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    pos = 0
    nn, ne, nd = (int(v) for v in lines[pos][:3]); pos += 1
    assert nd == 2
    nodes = [[float(v) for v in lines[pos + i][:2]] for i in range(nn)]
    pos += nn
    ng = int(lines[pos][0]); pos += 1
    bnodes, starts, tags = [], [], []
    for _ in range(ng):
        hdr = lines[pos]; pos += 1
        nf = int(hdr[0])
        if len(hdr) == 3:
            tags.append(hdr[2])
        starts.append(len(bnodes))
        for k in range(nf):
            bn = [int(v) for v in lines[pos + k]]
            bnodes.append([bn[0] - 1, bn[-1] - 1])
        pos += nf
    elements, found = [], 0
    while found < ne:
        hdr = lines[pos]; pos += 1
        sub, q = int(hdr[0]), int(hdr[1])
        faces = 3 if "Tri" in hdr[2] else 4
        for k in range(sub):
            e = [int(v) for v in lines[pos + k]]
            if q == 1:
                elements.append(e)
            elif q == 2:
                elements.append([e[0], e[2], e[5]] if faces == 3 else [e[0], e[2], e[8], e[6]])
            else:
                elements.append([e[0], e[3], e[9]] if faces == 3 else [e[0], e[3], e[15], e[12]])
        pos += sub
        found += sub
    offs = np.zeros(len(elements) + 1, dtype=np.uint32)
    offs[1:] = np.cumsum([len(e) for e in elements])
    return SimpleNamespace(nodes=np.array(nodes, np.float64).reshape(-1, 2), elem_offsets=offs,
                           elem_nodes=np.array([v for e in elements for v in e], np.uint32),
                           bnodes=np.array(bnodes, np.uint32).reshape(-1, 2), btags=tags,
                           bgroup_start=np.array(starts, np.uint64))

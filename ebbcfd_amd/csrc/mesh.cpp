// Mesh ingest (gmsh 2.2 ASCII) and geometry precompute.
//
// parseGmsh follows source/io.d:689-990 (section handling, boundary-group
// regrouping, ordering by physical tag number, clockwise->counter-clockwise fix).
// buildMesh follows source/mesh.d:217-643: edges in first-visit order, owner =
// first visitor (flux multiplier +1), interior normal = normalised (y1-y0, x0-x1),
// boundary normal = (1/len)*(y1-y0, x0-x1), shoelace area/centroid, d = 2A/perim,
// one ghost cell per boundary edge with the owner centroid reflected across the
// edge line, neighbour list in edge order and gradMat = (A^T A)^-1 A^T.
// Every floating-point expression keeps the reference's operation order so the
// geometry is bit-identical to the CPU restatement.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <algorithm>

#include "host.hpp"

namespace ebb {

namespace {

std::vector<std::string> cleanLine(const std::string& line)
{
    std::vector<std::string> out;
    std::string cur;
    for (char ch : line) {
        if (ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n') {
            if (!cur.empty()) { out.push_back(cur); cur.clear(); }
        } else {
            cur.push_back(ch);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

struct EdgeMap {
    std::vector<uint64_t> keys;
    std::vector<uint32_t> vals;
    uint64_t mask = 0;
    explicit EdgeMap(uint64_t n)
    {
        uint64_t cap = 64;
        while (cap < 2 * n + 64) cap <<= 1;
        keys.assign(cap, ~0ull);
        vals.assign(cap, 0);
        mask = cap - 1;
    }
    static uint64_t key(uint32_t a, uint32_t b)
    {
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        return (uint64_t(lo) << 32) | hi;
    }
    static uint64_t mix(uint64_t k)
    {
        k ^= k >> 31; k *= 0x9e3779b97f4a7c15ull; k ^= k >> 29;
        return k;
    }
    uint32_t get(uint32_t a, uint32_t b) const
    {
        uint64_t k = key(a, b), i = mix(k) & mask;
        while (keys[i] != ~0ull) {
            if (keys[i] == k) return vals[i];
            i = (i + 1) & mask;
        }
        return 0;
    }
    void put(uint32_t a, uint32_t b, uint32_t v)
    {
        uint64_t k = key(a, b), i = mix(k) & mask;
        while (keys[i] != ~0ull && keys[i] != k) i = (i + 1) & mask;
        keys[i] = k;
        vals[i] = v;
    }
};

}  // namespace

bool parseGmsh(const std::string& path, RawMesh& out, std::string& err)
{
    std::ifstream f(path);
    if (!f) { err = "cannot open mesh file " + path; return false; }
    std::string line;
    auto next = [&](std::vector<std::string>& tok) -> bool {
        if (!std::getline(f, line)) return false;
        tok = cleanLine(line);
        return true;
    };
    std::vector<std::string> tok;
    std::vector<std::vector<uint32_t>> bnodes;         // per line element
    std::vector<std::vector<int>> bgroups;             // physical tag per line element, grouped by runs
    std::vector<std::vector<uint32_t>> elements;
    out = RawMesh();
    while (true) {
        bool ok = next(tok);
        while (ok && tok.empty()) ok = next(tok);
        if (!ok) break;
        if (tok.size() != 1 || tok[0][0] != '$') { err = "Incorrectly formatted section header: " + line; return false; }
        std::string sec = tok[0].substr(1);
        if (sec == "MeshFormat") {
            if (!next(tok) || tok.size() < 2) { err = "bad $MeshFormat"; return false; }
            if (std::stoi(tok[1]) != 0) { err = "Unsuported mesh format. Expected 0, got " + tok[1]; return false; }
            if (!next(tok) || tok.empty() || tok[0].find("$End") == std::string::npos) { err = "Expected section end"; return false; }
        } else if (sec == "PhysicalNames") {
            if (!next(tok) || tok.size() != 1) { err = "bad $PhysicalNames count"; return false; }
            int n = std::stoi(tok[0]);
            for (int i = 0; i < n; i++) {
                if (!next(tok) || tok.size() != 3) { err = "Expected 3 entries in $PhysicalNames"; return false; }
                if (std::stoi(tok[0]) == 1) {
                    std::string t = tok[2];
                    t.erase(std::remove(t.begin(), t.end(), '"'), t.end());
                    out.btags.push_back(t);
                }
            }
            if (!next(tok) || tok.empty() || tok[0].find("$End") == std::string::npos) { err = "Expected section end"; return false; }
        } else if (sec == "Nodes") {
            if (!next(tok) || tok.size() != 1) { err = "bad $Nodes count"; return false; }
            long n = std::stol(tok[0]);
            out.nodes.reserve(2 * n);
            for (long i = 0; i < n; i++) {
                if (!next(tok) || tok.size() < 3) { err = "Unexpected end of file"; return false; }
                if (std::stol(tok[0]) != i + 1) { err = "Unexpected node number"; return false; }
                out.nodes.push_back(std::strtod(tok[1].c_str(), nullptr));
                out.nodes.push_back(std::strtod(tok[2].c_str(), nullptr));
            }
            if (!next(tok) || tok.empty() || tok[0].find("$End") == std::string::npos) { err = "Expected section end"; return false; }
        } else if (sec == "Elements") {
            if (!next(tok) || tok.size() != 1) { err = "bad $Elements count"; return false; }
            long n = std::stol(tok[0]);
            int current = -1;
            for (long i = 0; i < n; i++) {
                if (!next(tok) || tok.size() <= 4) { err = "Element line must have at least 4 entries"; return false; }
                std::vector<long> v(tok.size());
                for (size_t k = 0; k < tok.size(); k++) v[k] = std::stol(tok[k]);
                if (v[0] != i + 1) { err = "Unexpected element number"; return false; }
                long type = v[1], ntags = v[2];
                if (ntags < 1) { err = "Expected at least one element tag"; return false; }
                if (type == 1) {
                    int g = int(v[3]);
                    std::vector<uint32_t> bn;
                    for (size_t k = 3 + ntags; k < v.size(); k++) bn.push_back(uint32_t(v[k] - 1));
                    if (g != current) { current = g; bgroups.emplace_back(); }
                    bgroups.back().push_back(g);
                    bnodes.push_back(bn);
                } else if (type == 2 || type == 3) {
                    std::vector<uint32_t> el;
                    for (size_t k = 3 + ntags; k < v.size(); k++) el.push_back(uint32_t(v[k]));
                    elements.push_back(el);
                } else {
                    err = "Unsuported element type: " + std::to_string(type);
                    return false;
                }
            }
            if (!next(tok) || tok.empty() || tok[0].find("$End") == std::string::npos) { err = "Expected section end"; return false; }
        } else {
            while (true) {
                if (!next(tok)) { err = "Unexpected end of file"; return false; }
                if (!tok.empty() && tok[0].find("$End") != std::string::npos) break;
            }
        }
    }
    auto groupLen = [](const std::vector<std::vector<int>>& g, size_t a, size_t b) {
        size_t s = 0;
        for (size_t i = a; i < b; i++) s += g[i].size();
        return s;
    };
    // io.d:908-936 regroup split groups
    if (bgroups.size() != out.btags.size()) {
        std::vector<std::vector<uint32_t>> nbn;
        std::vector<std::vector<int>> nbg;
        for (size_t i = 0; i < bgroups.size(); i++) {
            int cur = bgroups[i][0];
            long old = -1;
            for (size_t k = 0; k < nbg.size(); k++) if (nbg[k][0] == cur) { old = long(k); break; }
            size_t start = groupLen(bgroups, 0, i), end = groupLen(bgroups, 0, i + 1);
            if (old >= 0) {
                size_t nend = groupLen(nbg, 0, size_t(old) + 1);
                nbg[old].insert(nbg[old].end(), bgroups[i].begin(), bgroups[i].end());
                nbn.insert(nbn.begin() + nend, bnodes.begin() + start, bnodes.begin() + end);
            } else {
                nbg.push_back(bgroups[i]);
                nbn.insert(nbn.end(), bnodes.begin() + start, bnodes.begin() + end);
            }
        }
        bnodes.swap(nbn);
        bgroups.swap(nbg);
    }
    if (bgroups.size() != out.btags.size()) { err = "Number of boundary groups does not match the number of boundary tags"; return false; }
    // io.d:939-960 order by physical tag number i+1
    for (size_t i = 0; i < bgroups.size(); i++) {
        long gi = -1;
        for (size_t k = 0; k < bgroups.size(); k++) if (bgroups[k][0] == int(i + 1)) { gi = long(k); break; }
        if (gi < 0) { err = "Could not determine boundary group"; return false; }
        size_t end = groupLen(bgroups, 0, size_t(gi) + 1), start = groupLen(bgroups, 0, size_t(gi));
        out.bgroupStart.push_back(out.bnodes.size() / 2);
        for (size_t b = start; b < end; b++) {
            if (bnodes[b].size() != 2) { err = "boundary line element must have 2 nodes"; return false; }
            out.bnodes.push_back(bnodes[b][0]);
            out.bnodes.push_back(bnodes[b][1]);
        }
    }
    // io.d:969-987 clockwise -> counter-clockwise
    out.eoff.assign(1, 0);
    for (auto& el : elements) {
        double area = 0;
        size_t m = el.size();
        for (size_t j = 0; j < m; j++) {
            uint32_t n1 = el[j] - 1, n2 = el[(j + 1) % m] - 1;
            area += out.nodes[2 * n1] * out.nodes[2 * n2 + 1] - out.nodes[2 * n2] * out.nodes[2 * n1 + 1];
        }
        if (area < 0) std::reverse(el.begin(), el.end());
        out.enodes.insert(out.enodes.end(), el.begin(), el.end());
        out.eoff.push_back(uint32_t(out.enodes.size()));
    }
    return true;
}

// XFlow .gri reader (source/io.d:992-1121): "nNodes nElems 2", node lines,
// boundary groups "nFaces nodesPerFace [tag]" whose faces keep their first and
// last node, element groups "count q TriLagrange|QuadLagrange" with q = 1..3
// (corner nodes picked from the higher-order node lists).  No orientation fix is
// applied (the reference applies it to gmsh meshes only).  The reference gives
// every cell the edge count of the last element group, so it cannot read a mesh
// that mixes triangles and quads; such a file is rejected here.
bool parseGri(const std::string& path, RawMesh& out, std::string& err)
{
    std::ifstream f(path);
    if (!f) { err = "cannot open mesh file " + path; return false; }
    std::string line;
    std::vector<std::string> tok;
    auto next = [&]() -> bool {
        while (std::getline(f, line)) {
            tok = cleanLine(line);
            if (!tok.empty()) return true;
        }
        return false;
    };
    out = RawMesh();
    if (!next() || tok.size() < 3) { err = "bad .gri header"; return false; }
    const long nNodes = std::stol(tok[0]), nElems = std::stol(tok[1]), nDims = std::stol(tok[2]);
    if (nDims != 2) { err = std::to_string(nDims) + " dimensional meshes not supported"; return false; }
    for (long i = 0; i < nNodes; i++) {
        if (!next() || tok.size() < 2) { err = "Unexpected end of file in nodes"; return false; }
        out.nodes.push_back(std::strtod(tok[0].c_str(), nullptr));
        out.nodes.push_back(std::strtod(tok[1].c_str(), nullptr));
    }
    if (!next()) { err = "missing boundary group count"; return false; }
    const long nGroups = std::stol(tok[0]);
    for (long g = 0; g < nGroups; g++) {
        if (!next() || tok.size() < 2) { err = "bad boundary group header"; return false; }
        const long nf = std::stol(tok[0]);
        if (tok.size() == 3) out.btags.push_back(tok[2]);
        out.bgroupStart.push_back(out.bnodes.size() / 2);
        for (long k = 0; k < nf; k++) {
            if (!next() || tok.size() < 2) { err = "Unexpected end of file in boundary faces"; return false; }
            out.bnodes.push_back(uint32_t(std::stol(tok.front()) - 1));
            out.bnodes.push_back(uint32_t(std::stol(tok.back()) - 1));
        }
    }
    out.eoff.assign(1, 0);
    long found = 0;
    int faces = 0, firstFaces = 0;
    while (found < nElems) {
        if (!next() || tok.size() < 3) { err = "bad element group header"; return false; }
        const long sub = std::stol(tok[0]);
        const long q = std::stol(tok[1]);
        if (q >= 4 || q == 0) { err = "Unsuported q"; return false; }
        if (tok[2].find("Tri") != std::string::npos) faces = 3;
        else if (tok[2].find("Quad") != std::string::npos) faces = 4;
        else { err = "Unsuported cell type"; return false; }
        if (firstFaces == 0) firstFaces = faces;
        if (faces != firstFaces) { err = "mixed triangle/quad .gri meshes are not supported"; return false; }
        for (long i = 0; i < sub; i++) {
            if (!next()) { err = "Unexpected end of file in elements"; return false; }
            std::vector<uint32_t> e;
            for (auto& t : tok) e.push_back(uint32_t(std::stol(t)));
            std::vector<uint32_t> c;
            if (q == 1) c = e;
            else if (q == 2) c = faces == 3 ? std::vector<uint32_t>{e[0], e[2], e[5]} : std::vector<uint32_t>{e[0], e[2], e[8], e[6]};
            else c = faces == 3 ? std::vector<uint32_t>{e[0], e[3], e[9]} : std::vector<uint32_t>{e[0], e[3], e[15], e[12]};
            if (int(c.size()) != faces) { err = "element has the wrong number of nodes"; return false; }
            out.enodes.insert(out.enodes.end(), c.begin(), c.end());
            out.eoff.push_back(uint32_t(out.enodes.size()));
        }
        found += sub;
    }
    return true;
}

bool buildMesh(const RawMesh& raw, BuiltMesh& m, std::string& err)
{
    m = BuiltMesh();
    const uint32_t nElem = uint32_t(raw.eoff.size() - 1);
    const uint32_t nB = uint32_t(raw.bnodes.size() / 2);
    const uint32_t nTags = uint32_t(raw.btags.size());
    const uint32_t nNodes = uint32_t(raw.nodes.size() / 2);
    const double* nd = raw.nodes.data();
    m.nElem = nElem;
    const uint64_t maxEdges = raw.eoff[nElem];
    for (uint32_t v : raw.enodes)
        if (v < 1 || v > nNodes) { err = "element references a node out of range"; return false; }
    for (uint32_t v : raw.bnodes)
        if (v >= nNodes) { err = "boundary line references a node out of range"; return false; }

    EdgeMap edgeHash(maxEdges), boundaryHash(nB);
    for (uint32_t i = 0; i < nB; i++) boundaryHash.put(raw.bnodes[2 * i], raw.bnodes[2 * i + 1], i + 1);

    m.nEdgesCell.assign(nElem, 0);
    m.cellEdges.assign(size_t(nElem) * MAXE, 0);
    m.fluxMult.assign(size_t(nElem) * MAXE, 1);
    m.nbr.assign(size_t(nElem) * MAXE, 0);
    m.gradMat.assign(size_t(nElem) * 2 * MAXE, 0.0);
    m.area.assign(nElem, 0.0); m.d.assign(nElem, 0.0); m.perim.assign(nElem, 0.0);
    m.cx.assign(nElem, 0.0); m.cy.assign(nElem, 0.0);
    m.bGroups.assign(nTags, {});
    auto reserveE = [&](auto& v) { v.reserve(maxEdges / 2 + nB + 16); };
    reserveE(m.eNode0); reserveE(m.eNode1); reserveE(m.eCell0); reserveE(m.eCell1);
    reserveE(m.len); reserveE(m.nx); reserveE(m.ny); reserveE(m.mx); reserveE(m.my); reserveE(m.eBGroup);

    for (uint32_t i = 0; i < nElem; i++) {
        const uint32_t ne = raw.eoff[i + 1] - raw.eoff[i];
        if (ne < 3 || ne > uint32_t(MAXE)) { err = "element " + std::to_string(i) + " has an unsupported number of edges"; return false; }
        m.nEdgesCell[i] = uint8_t(ne);
        const uint32_t* el = raw.enodes.data() + raw.eoff[i];
        double area = 0.0, perim = 0.0, c0 = 0.0, c1 = 0.0;
        for (uint32_t j = 0; j < ne; j++) {
            const uint32_t n0 = el[j] - 1, n1 = el[(j + 1) % ne] - 1;
            uint32_t edgeidx;
            const uint32_t bval = boundaryHash.get(n0, n1);
            if (bval == 0) {
                const uint32_t found = edgeHash.get(n0, n1);
                if (found != 0) {
                    edgeidx = found - 1;
                    m.eCell1[edgeidx] = i;
                    m.fluxMult[size_t(i) * MAXE + j] = -1;
                } else {
                    const double ddx = nd[2 * n0] - nd[2 * n1], ddy = nd[2 * n0 + 1] - nd[2 * n1 + 1];
                    const double len = std::sqrt(ddx * ddx + ddy * ddy);
                    const double vx = nd[2 * n1 + 1] - nd[2 * n0 + 1], vy = nd[2 * n0] - nd[2 * n1];
                    const double mag = std::sqrt(vx * vx + vy * vy);
                    edgeidx = uint32_t(m.len.size());
                    m.eNode0.push_back(n0); m.eNode1.push_back(n1);
                    m.eCell0.push_back(i); m.eCell1.push_back(0);
                    m.len.push_back(len);
                    m.nx.push_back(vx / mag); m.ny.push_back(vy / mag);
                    m.mx.push_back(0.5 * (nd[2 * n0] + nd[2 * n1]));
                    m.my.push_back(0.5 * (nd[2 * n0 + 1] + nd[2 * n1 + 1]));
                    m.eBGroup.push_back(-1);
                    edgeHash.put(n0, n1, edgeidx + 1);
                    m.interiorEdges.push_back(edgeidx);
                }
            } else {
                const double ddx = nd[2 * n0] - nd[2 * n1], ddy = nd[2 * n0 + 1] - nd[2 * n1 + 1];
                const double len = std::sqrt(ddx * ddx + ddy * ddy);
                const double inv = 1.0 / len;
                const uint32_t bNodeIdx = bval - 1;
                int32_t g = 0;
                for (uint32_t t = 0; t + 1 < nTags; t++)
                    if (bNodeIdx >= raw.bgroupStart[t] && bNodeIdx < raw.bgroupStart[t + 1]) g = int32_t(t);
                if (nTags > 0 && bNodeIdx >= raw.bgroupStart[nTags - 1]) g = int32_t(nTags) - 1;
                edgeidx = uint32_t(m.len.size());
                m.eNode0.push_back(n0); m.eNode1.push_back(n1);
                m.eCell0.push_back(i); m.eCell1.push_back(0);
                m.len.push_back(len);
                m.nx.push_back(inv * (nd[2 * n1 + 1] - nd[2 * n0 + 1]));
                m.ny.push_back(inv * (nd[2 * n0] - nd[2 * n1]));
                m.mx.push_back(0.5 * (nd[2 * n0] + nd[2 * n1]));
                m.my.push_back(0.5 * (nd[2 * n0 + 1] + nd[2 * n1 + 1]));
                m.eBGroup.push_back(g);
                if (nTags > 0) m.bGroups[g].push_back(edgeidx);
                m.boundaryEdges.push_back(edgeidx);
            }
            m.cellEdges[size_t(i) * MAXE + j] = edgeidx;
            perim += m.len[edgeidx];
            const double x0 = nd[2 * n0], x1 = nd[2 * n1], y0 = nd[2 * n0 + 1], y1 = nd[2 * n1 + 1];
            area += x0 * y1 - x1 * y0;
            c0 += (x0 + x1) * (x0 * y1 - x1 * y0);
            c1 += (y0 + y1) * (x0 * y1 - x1 * y0);
        }
        area *= 0.5;
        if (!(area > 0)) { err = "Computed cell " + std::to_string(i) + " area negative"; return false; }
        m.area[i] = area;
        m.perim[i] = perim;
        m.d[i] = (2 * area) / perim;
        const double f = 1 / (6 * area);
        m.cx[i] = c0 * f;
        m.cy[i] = c1 * f;
    }
    m.nEdges = uint32_t(m.len.size());

    // ghost cells (mesh.d:329-381)
    const uint32_t nBE = uint32_t(m.boundaryEdges.size());
    m.nCells = nElem + nBE;
    m.area.resize(m.nCells, 0.0); m.d.resize(m.nCells, 0.0); m.perim.resize(m.nCells, 0.0);
    m.cx.resize(m.nCells, 0.0); m.cy.resize(m.nCells, 0.0);
    for (uint32_t b = 0; b < nBE; b++) {
        const uint32_t e = m.boundaryEdges[b];
        const uint32_t o = m.eCell0[e];
        const double xa = nd[2 * m.eNode0[e]], ya = nd[2 * m.eNode0[e] + 1];
        const double xb = nd[2 * m.eNode1[e]], yb = nd[2 * m.eNode1[e] + 1];
        double xr, yr;
        if (xa != xb) {
            const double x1 = m.cx[o], y1 = m.cy[o];
            const double mm = (ya - yb) / (xa - xb);
            const double c = ya - mm * xa;
            const double dd = (x1 + (y1 - c) * mm) / (1.0 + mm * mm);
            xr = 2.0 * dd - x1;
            yr = 2.0 * dd * mm - y1 + 2.0 * c;
        } else {
            const double dd = m.cx[o] - xa;
            const double x1 = m.cx[o];
            xr = -(2.0 * dd - x1);
            yr = m.cy[o];
        }
        m.cx[nElem + b] = xr;
        m.cy[nElem + b] = yr;
        m.eCell1[e] = nElem + b;
    }

    // neighbours + LSQ matrix (mesh.d:481-642)
    for (uint32_t i = 0; i < nElem; i++) {
        const uint32_t ne = m.nEdgesCell[i];
        double A[MAXE][2] = {};
        for (uint32_t j = 0; j < ne; j++) {
            const uint32_t e = m.cellEdges[size_t(i) * MAXE + j];
            const uint32_t nb = (m.eCell0[e] == i) ? m.eCell1[e] : m.eCell0[e];
            m.nbr[size_t(i) * MAXE + j] = nb;
            A[j][0] = m.cx[nb] - m.cx[i];
            A[j][1] = m.cy[nb] - m.cy[i];
        }
        double b[2][2];
        for (int r = 0; r < 2; r++)
            for (int c = 0; c < 2; c++) {
                double s = 0.0;
                for (int k = 0; k < MAXE; k++) s += A[k][r] * A[k][c];
                b[r][c] = s;
            }
        const double det = b[0][0] * b[1][1] - b[0][1] * b[1][0];
        const double id = 1.0 / det;
        const double inv[2][2] = {{b[1][1] * id, -b[0][1] * id}, {-b[1][0] * id, b[0][0] * id}};
        for (int r = 0; r < 2; r++)
            for (int k = 0; k < MAXE; k++)
                m.gradMat[size_t(i) * 2 * MAXE + r * MAXE + k] = inv[r][0] * A[k][0] + inv[r][1] * A[k][1];
    }
    return true;
}

}  // namespace ebb

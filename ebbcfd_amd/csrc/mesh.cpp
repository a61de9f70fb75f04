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
#include <sched.h>

#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <algorithm>
#include <cerrno>
#include <atomic>
#include <thread>

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
    // full-avalanche 64-bit mixer (splitmix64 finaliser): the table index uses the low
    // bits, which after a single multiply would depend on the low key bits only and
    // cluster structured node numberings into long probe runs
    static uint64_t mix(uint64_t k)
    {
        k ^= k >> 30; k *= 0xbf58476d1ce4e5b9ull;
        k ^= k >> 27; k *= 0x94d049bb133111ebull;
        k ^= k >> 31;
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

// Line reader over a file held in memory: tokens split on ' ', '\t', '\r' (as
// cleanLine does) without allocating per line.
struct LineTokens {
    const char* p = nullptr;
    const char* end = nullptr;
    const char* lineBeg = nullptr;
    const char* lineEnd = nullptr;
    static constexpr int MAXT = 64;
    const char* tb[MAXT];
    const char* te[MAXT];
    int n = 0;
    bool next()
    {
        if (p >= end) return false;
        lineBeg = p;
        const char* e = static_cast<const char*>(std::memchr(p, '\n', size_t(end - p)));
        if (!e) e = end;
        lineEnd = e;
        p = e < end ? e + 1 : end;
        n = 0;
        const char* q = lineBeg;
        while (q < lineEnd) {
            while (q < lineEnd && (*q == ' ' || *q == '\t' || *q == '\r')) q++;
            if (q >= lineEnd) break;
            const char* t0 = q;
            while (q < lineEnd && !(*q == ' ' || *q == '\t' || *q == '\r')) q++;
            if (n < MAXT) { tb[n] = t0; te[n] = q; }
            n++;
        }
        return true;
    }
    std::string str(int k) const { return std::string(tb[k], te[k]); }
    std::string line() const { return std::string(lineBeg, lineEnd); }
    bool startsEnd() const { return n > 0 && str(0).find("$End") != std::string::npos; }
    // integer prefix of token k, like std::stol (optional sign, digits)
    bool num(int k, long& v) const
    {
        const char* a = tb[k];
        if (a < te[k] && *a == '+') a++;
        const auto r = std::from_chars(a, te[k], v);
        return r.ec == std::errc();
    }
};

bool readFile(const std::string& path, std::string& buf)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    if (n < 0) return false;
    f.seekg(0, std::ios::beg);
    buf.resize(size_t(n));
    f.read(&buf[0], n);
    return bool(f) || f.eof();
}

bool parseGmsh(const std::string& path, RawMesh& out, std::string& err)
{
    std::string buf;
    if (!readFile(path, buf)) { err = "cannot open mesh file " + path; return false; }
    LineTokens L;
    L.p = buf.data();
    L.end = buf.data() + buf.size();
    std::vector<std::vector<uint32_t>> bnodes;         // per line element
    std::vector<std::vector<int>> bgroups;             // physical tag per line element, grouped by runs
    std::vector<uint32_t> elemNodes, elemOff(1, 0);    // area elements, 1-based nodes
    out = RawMesh();
    long v[LineTokens::MAXT];
    while (true) {
        bool ok = L.next();
        while (ok && L.n == 0) ok = L.next();
        if (!ok) break;
        if (L.n != 1 || *L.tb[0] != '$') { err = "Incorrectly formatted section header: " + L.line(); return false; }
        const std::string sec = L.str(0).substr(1);
        if (sec == "MeshFormat") {
            long fmt;
            if (!L.next() || L.n < 2 || !L.num(1, fmt)) { err = "bad $MeshFormat"; return false; }
            if (fmt != 0) { err = "Unsuported mesh format. Expected 0, got " + L.str(1); return false; }
            if (!L.next() || !L.startsEnd()) { err = "Expected section end"; return false; }
        } else if (sec == "PhysicalNames") {
            long n;
            if (!L.next() || L.n != 1 || !L.num(0, n)) { err = "bad $PhysicalNames count"; return false; }
            for (long i = 0; i < n; i++) {
                long dim;
                if (!L.next() || L.n != 3 || !L.num(0, dim)) { err = "Expected 3 entries in $PhysicalNames"; return false; }
                if (dim == 1) {
                    std::string t = L.str(2);
                    t.erase(std::remove(t.begin(), t.end(), '"'), t.end());
                    out.btags.push_back(t);
                }
            }
            if (!L.next() || !L.startsEnd()) { err = "Expected section end"; return false; }
        } else if (sec == "Nodes") {
            long n;
            if (!L.next() || L.n != 1 || !L.num(0, n) || n < 0) { err = "bad $Nodes count"; return false; }
            out.nodes.reserve(2 * size_t(n));
            for (long i = 0; i < n; i++) {
                long id;
                if (!L.next() || L.n < 3) { err = "Unexpected end of file"; return false; }
                if (!L.num(0, id) || id != i + 1) { err = "Unexpected node number"; return false; }
                // strtod on the buffer: it stops at the whitespace after the number
                out.nodes.push_back(std::strtod(L.tb[1], nullptr));
                out.nodes.push_back(std::strtod(L.tb[2], nullptr));
            }
            if (!L.next() || !L.startsEnd()) { err = "Expected section end"; return false; }
        } else if (sec == "Elements") {
            long n;
            if (!L.next() || L.n != 1 || !L.num(0, n) || n < 0) { err = "bad $Elements count"; return false; }
            int current = -1;
            for (long i = 0; i < n; i++) {
                if (!L.next() || L.n <= 4) { err = "Element line must have at least 4 entries"; return false; }
                if (L.n > LineTokens::MAXT) { err = "element line has too many entries"; return false; }
                for (int k = 0; k < L.n; k++)
                    if (!L.num(k, v[k])) { err = "bad number in element line"; return false; }
                if (v[0] != i + 1) { err = "Unexpected element number"; return false; }
                const long type = v[1], ntags = v[2];
                if (ntags < 1) { err = "Expected at least one element tag"; return false; }
                if (3 + ntags > L.n) { err = "Element line must have at least 4 entries"; return false; }
                if (type == 1) {
                    const int g = int(v[3]);
                    std::vector<uint32_t> bn;
                    for (int k = int(3 + ntags); k < L.n; k++) bn.push_back(uint32_t(v[k] - 1));
                    if (g != current) { current = g; bgroups.emplace_back(); }
                    bgroups.back().push_back(g);
                    bnodes.push_back(bn);
                } else if (type == 2 || type == 3) {
                    for (int k = int(3 + ntags); k < L.n; k++) elemNodes.push_back(uint32_t(v[k]));
                    elemOff.push_back(uint32_t(elemNodes.size()));
                } else {
                    err = "Unsuported element type: " + std::to_string(type);
                    return false;
                }
            }
            if (!L.next() || !L.startsEnd()) { err = "Expected section end"; return false; }
        } else {
            while (true) {
                if (!L.next()) { err = "Unexpected end of file"; return false; }
                if (L.startsEnd()) break;
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
    const size_t nNodes = out.nodes.size() / 2;
    for (uint32_t x : elemNodes)
        if (x < 1 || x > nNodes) { err = "element references a node out of range"; return false; }
    out.eoff.assign(1, 0);
    out.enodes.reserve(elemNodes.size());
    out.eoff.reserve(elemOff.size());
    for (size_t e = 0; e + 1 < elemOff.size(); e++) {
        uint32_t* el = elemNodes.data() + elemOff[e];
        const size_t m = elemOff[e + 1] - elemOff[e];
        double area = 0;
        for (size_t j = 0; j < m; j++) {
            uint32_t n1 = el[j] - 1, n2 = el[(j + 1) % m] - 1;
            area += out.nodes[2 * n1] * out.nodes[2 * n2 + 1] - out.nodes[2 * n2] * out.nodes[2 * n1 + 1];
        }
        if (area < 0) std::reverse(el, el + m);
        out.enodes.insert(out.enodes.end(), el, el + m);
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
    // whole-token integer in [lo, hi], without exceptions
    auto num = [&](size_t k, long lo, long hi, long& v) -> bool {
        if (k >= tok.size()) return false;
        const char* b = tok[k].c_str();
        char* e = nullptr;
        errno = 0;
        const long x = std::strtol(b, &e, 10);
        if (e == b || *e != '\0' || errno == ERANGE || x < lo || x > hi) return false;
        v = x;
        return true;
    };
    constexpr long MAXN = 0x7fffffffL;
    out = RawMesh();
    long nNodes = 0, nElems = 0, nDims = 0;
    if (!next() || !num(0, 0, MAXN, nNodes) || !num(1, 0, MAXN, nElems) || !num(2, 0, MAXN, nDims)) {
        err = "bad .gri header";
        return false;
    }
    if (nDims != 2) { err = std::to_string(nDims) + " dimensional meshes not supported"; return false; }
    for (long i = 0; i < nNodes; i++) {
        if (!next() || tok.size() < 2) { err = "Unexpected end of file in nodes"; return false; }
        out.nodes.push_back(std::strtod(tok[0].c_str(), nullptr));
        out.nodes.push_back(std::strtod(tok[1].c_str(), nullptr));
    }
    long nGroups = 0;
    if (!next() || !num(0, 0, MAXN, nGroups)) { err = "missing boundary group count"; return false; }
    for (long g = 0; g < nGroups; g++) {
        long nf = 0, npf = 0;
        if (!next() || !num(0, 0, MAXN, nf) || !num(1, 2, 16, npf)) { err = "bad boundary group header"; return false; }
        if (tok.size() == 3) out.btags.push_back(tok[2]);
        out.bgroupStart.push_back(out.bnodes.size() / 2);
        for (long k = 0; k < nf; k++) {
            long a = 0, b = 0;
            if (!next() || tok.size() < 2 || !num(0, 1, nNodes, a) || !num(tok.size() - 1, 1, nNodes, b)) {
                err = "bad boundary face";
                return false;
            }
            out.bnodes.push_back(uint32_t(a - 1));
            out.bnodes.push_back(uint32_t(b - 1));
        }
    }
    out.eoff.assign(1, 0);
    long found = 0;
    int faces = 0, firstFaces = 0;
    while (found < nElems) {
        long sub = 0, q = 0;
        if (!next() || tok.size() < 3 || !num(0, 0, MAXN, sub) || !num(1, 0, MAXN, q)) {
            err = "bad element group header";
            return false;
        }
        if (q >= 4 || q == 0) { err = "Unsuported q"; return false; }
        if (tok[2].find("Tri") != std::string::npos) faces = 3;
        else if (tok[2].find("Quad") != std::string::npos) faces = 4;
        else { err = "Unsuported cell type"; return false; }
        if (firstFaces == 0) firstFaces = faces;
        if (faces != firstFaces) { err = "mixed triangle/quad .gri meshes are not supported"; return false; }
        // nodes per element of a Lagrange triangle / quadrilateral of order q
        const size_t need = faces == 3 ? size_t((q + 1) * (q + 2) / 2) : size_t((q + 1) * (q + 1));
        for (long i = 0; i < sub; i++) {
            if (!next()) { err = "Unexpected end of file in elements"; return false; }
            if (tok.size() != need) { err = "element has the wrong number of nodes"; return false; }
            std::vector<uint32_t> e(need);
            for (size_t k = 0; k < need; k++) {
                long v = 0;
                if (!num(k, 1, nNodes, v)) { err = "bad element node"; return false; }
                e[k] = uint32_t(v);
            }
            std::vector<uint32_t> c;
            if (q == 1) c = e;
            else if (q == 2) c = faces == 3 ? std::vector<uint32_t>{e[0], e[2], e[5]} : std::vector<uint32_t>{e[0], e[2], e[8], e[6]};
            else c = faces == 3 ? std::vector<uint32_t>{e[0], e[3], e[9]} : std::vector<uint32_t>{e[0], e[3], e[15], e[12]};
            out.enodes.insert(out.enodes.end(), c.begin(), c.end());
            out.eoff.push_back(uint32_t(out.enodes.size()));
        }
        found += sub;
    }
    return true;
}

// Threads for the geometry precompute: the CPUs this process may run on, at most
// OMP_NUM_THREADS when that is set, at most 32.
unsigned hostThreads()
{
    unsigned T = std::thread::hardware_concurrency();
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) T = unsigned(CPU_COUNT(&set));
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v > 0 && unsigned(v) < T) T = unsigned(v);
    }
    return T == 0 ? 1 : (T > 32 ? 32 : T);
}

// Run f(begin, end) over [0, n) on the host's cores (serially for small n).
template <typename F>
void parallelFor(uint32_t n, F&& f)
{
    static const unsigned hw = hostThreads();
    unsigned T = n < 65536 ? 1 : hw;
    if (T == 1) { f(0u, n); return; }
    std::vector<std::thread> th;
    const uint32_t chunk = (n + T - 1) / T;
    for (unsigned t = 0; t < T; t++) {
        const uint32_t b = t * chunk, e = std::min<uint32_t>(n, b + chunk);
        if (b < e) th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& x : th) x.join();
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
    for (uint32_t i = 0; i < nElem; i++) {
        const uint32_t ne = raw.eoff[i + 1] - raw.eoff[i];
        if (ne < 3 || ne > uint32_t(MAXE)) { err = "element " + std::to_string(i) + " has an unsupported number of edges"; return false; }
    }

    EdgeMap boundaryHash(nB);
    for (uint32_t i = 0; i < nB; i++) boundaryHash.put(raw.bnodes[2 * i], raw.bnodes[2 * i + 1], i + 1);
    // Edge discovery without a global hash: every cell-edge visit v (cell i, local
    // edge j, v = eoff[i] + j) is bucketed under its lower node, in visit order.  The
    // first visit of a node pair is the first bucket entry with the same upper node,
    // which is exactly the first-visit order the reference's map produces.
    std::vector<uint32_t> bstart(size_t(nNodes) + 1, 0);
    FillVec<uint32_t> bhi(maxEdges), bvis(maxEdges);
    for (uint32_t i = 0; i < nElem; i++) {
        const uint32_t ne = raw.eoff[i + 1] - raw.eoff[i];
        const uint32_t* el = raw.enodes.data() + raw.eoff[i];
        for (uint32_t j = 0; j < ne; j++) {
            const uint32_t n0 = el[j] - 1, n1 = el[(j + 1) % ne] - 1;
            bstart[size_t(n0 < n1 ? n0 : n1) + 1]++;
        }
    }
    for (uint32_t k = 0; k < nNodes; k++) bstart[k + 1] += bstart[k];
    {
        std::vector<uint32_t> cur(bstart.begin(), bstart.end() - 1);
        for (uint32_t i = 0; i < nElem; i++) {
            const uint32_t ne = raw.eoff[i + 1] - raw.eoff[i];
            const uint32_t* el = raw.enodes.data() + raw.eoff[i];
            for (uint32_t j = 0; j < ne; j++) {
                const uint32_t n0 = el[j] - 1, n1 = el[(j + 1) % ne] - 1;
                const uint32_t lo = n0 < n1 ? n0 : n1, hi = n0 < n1 ? n1 : n0;
                const uint32_t q = cur[lo]++;
                bhi[q] = hi;
                bvis[q] = raw.eoff[i] + j;
            }
        }
    }
    auto firstVisit = [&](uint32_t n0, uint32_t n1) {
        const uint32_t lo = n0 < n1 ? n0 : n1, hi = n0 < n1 ? n1 : n0;
        for (uint32_t q = bstart[lo];; q++)
            if (bhi[q] == hi) return bvis[q];
    };

    // A: classify every visit (boundary line / first visit of an interior edge / repeat)
    enum : uint8_t { REPEAT = 0, FIRST = 1, BOUNDARY = 2 };
    // Written where read: kind for every visit, firstOf for non-boundary visits,
    // bvalOf for boundary visits, edgeOfVisit for non-repeat visits.
    FillVec<uint8_t> kind(maxEdges);
    FillVec<uint32_t> firstOf(maxEdges), edgeOfVisit(maxEdges), bvalOf(maxEdges);
    parallelFor(nElem, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; i++) {
            const uint32_t ne = raw.eoff[i + 1] - raw.eoff[i];
            const uint32_t* el = raw.enodes.data() + raw.eoff[i];
            for (uint32_t j = 0; j < ne; j++) {
                const uint32_t n0 = el[j] - 1, n1 = el[(j + 1) % ne] - 1, v = raw.eoff[i] + j;
                const uint32_t bval = boundaryHash.get(n0, n1);
                if (bval) { kind[v] = BOUNDARY; bvalOf[v] = bval; continue; }
                const uint32_t f = firstVisit(n0, n1);
                kind[v] = f == v ? FIRST : REPEAT;
                firstOf[v] = f;
            }
        }
    });
    // B: edge numbers in visit order, boundary groups
    uint32_t nEdges = 0;
    for (uint64_t v = 0; v < maxEdges; v++)
        if (kind[v] != REPEAT) edgeOfVisit[v] = nEdges++;
    m.nEdges = nEdges;
    // Every edge has exactly one non-repeat visit, which writes all its fields below.
    m.eNode0.resize(nEdges); m.eNode1.resize(nEdges); m.eCell0.resize(nEdges); m.eCell1.resize(nEdges);
    m.len.resize(nEdges); m.nx.resize(nEdges); m.ny.resize(nEdges);
    m.mx.resize(nEdges); m.my.resize(nEdges); m.eBGroup.resize(nEdges);
    m.bGroups.assign(nTags, {});
    for (uint64_t v = 0; v < maxEdges; v++) {
        if (kind[v] == FIRST) {
            m.eBGroup[edgeOfVisit[v]] = -1;
            m.interiorEdges.push_back(edgeOfVisit[v]);
        } else if (kind[v] == BOUNDARY) {
            const uint32_t edgeidx = edgeOfVisit[v], bNodeIdx = bvalOf[v] - 1;
            int32_t g = 0;
            for (uint32_t t = 0; t + 1 < nTags; t++)
                if (bNodeIdx >= raw.bgroupStart[t] && bNodeIdx < raw.bgroupStart[t + 1]) g = int32_t(t);
            if (nTags > 0 && bNodeIdx >= raw.bgroupStart[nTags - 1]) g = int32_t(nTags) - 1;
            m.eBGroup[edgeidx] = g;
            if (nTags > 0) m.bGroups[g].push_back(edgeidx);
            m.boundaryEdges.push_back(edgeidx);
        }
    }

    // C: per cell, its edges (geometry of the edges it creates) and its own geometry.
    // The per-cell arrays are sized without a fill: this loop and the ghost and LSQ
    // loops below write every slot, so each page is first touched by its thread.
    const uint32_t nBE = uint32_t(m.boundaryEdges.size());
    m.nCells = nElem + nBE;
    m.nEdgesCell.resize(nElem);
    m.cellEdges.resize(size_t(nElem) * MAXE);
    m.fluxMult.resize(size_t(nElem) * MAXE);
    m.nbr.resize(size_t(nElem) * MAXE);
    m.gradMat.resize(size_t(nElem) * 2 * MAXE);
    m.area.resize(m.nCells); m.d.resize(m.nCells); m.perim.resize(m.nCells);
    m.cx.resize(m.nCells); m.cy.resize(m.nCells);
    std::atomic<uint32_t> badArea(~0u);
    parallelFor(nElem, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; i++) {
            const uint32_t ne = raw.eoff[i + 1] - raw.eoff[i];
            m.nEdgesCell[i] = uint8_t(ne);
            const uint32_t* el = raw.enodes.data() + raw.eoff[i];
            for (uint32_t j = ne; j < uint32_t(MAXE); j++) {
                m.cellEdges[size_t(i) * MAXE + j] = 0;
                m.fluxMult[size_t(i) * MAXE + j] = 1;
            }
            double area = 0.0, perim = 0.0, c0 = 0.0, c1 = 0.0;
            for (uint32_t j = 0; j < ne; j++) {
                const uint32_t n0 = el[j] - 1, n1 = el[(j + 1) % ne] - 1, v = raw.eoff[i] + j;
                const double ddx = nd[2 * n0] - nd[2 * n1], ddy = nd[2 * n0 + 1] - nd[2 * n1 + 1];
                const double len = std::sqrt(ddx * ddx + ddy * ddy);   // the same for either direction
                uint32_t edgeidx;
                if (kind[v] == REPEAT) {
                    edgeidx = edgeOfVisit[firstOf[v]];
                    m.fluxMult[size_t(i) * MAXE + j] = -1;
                } else {
                    edgeidx = edgeOfVisit[v];
                    m.fluxMult[size_t(i) * MAXE + j] = 1;
                    m.eNode0[edgeidx] = n0; m.eNode1[edgeidx] = n1;
                    m.eCell0[edgeidx] = i;
                    m.eCell1[edgeidx] = 0;   // set below: the last other visitor, or the ghost
                    m.len[edgeidx] = len;
                    if (kind[v] == FIRST) {
                        const double vx = nd[2 * n1 + 1] - nd[2 * n0 + 1], vy = nd[2 * n0] - nd[2 * n1];
                        const double mag = std::sqrt(vx * vx + vy * vy);
                        m.nx[edgeidx] = vx / mag; m.ny[edgeidx] = vy / mag;
                    } else {
                        const double inv = 1.0 / len;
                        m.nx[edgeidx] = inv * (nd[2 * n1 + 1] - nd[2 * n0 + 1]);
                        m.ny[edgeidx] = inv * (nd[2 * n0] - nd[2 * n1]);
                    }
                    m.mx[edgeidx] = 0.5 * (nd[2 * n0] + nd[2 * n1]);
                    m.my[edgeidx] = 0.5 * (nd[2 * n0 + 1] + nd[2 * n1 + 1]);
                }
                m.cellEdges[size_t(i) * MAXE + j] = edgeidx;
                perim += len;
                const double x0 = nd[2 * n0], x1 = nd[2 * n1], y0 = nd[2 * n0 + 1], y1 = nd[2 * n1 + 1];
                area += x0 * y1 - x1 * y0;
                c0 += (x0 + x1) * (x0 * y1 - x1 * y0);
                c1 += (y0 + y1) * (x0 * y1 - x1 * y0);
            }
            area *= 0.5;
            if (!(area > 0)) {
                uint32_t cur = badArea.load();
                while (i < cur && !badArea.compare_exchange_weak(cur, i)) {}
                continue;
            }
            m.area[i] = area;
            m.perim[i] = perim;
            m.d[i] = (2 * area) / perim;
            const double f = 1 / (6 * area);
            m.cx[i] = c0 * f;
            m.cy[i] = c1 * f;
        }
    });
    if (badArea.load() != ~0u) { err = "Computed cell " + std::to_string(badArea.load()) + " area negative"; return false; }
    // D: the other cell of each interior edge; the last visitor wins, as in the reference's loop
    for (uint32_t i = 0; i < nElem; i++)
        for (uint32_t v = raw.eoff[i]; v < raw.eoff[i + 1]; v++)
            if (kind[v] == REPEAT) m.eCell1[edgeOfVisit[firstOf[v]]] = i;

    // ghost cells (mesh.d:329-381)
    parallelFor(nBE, [&](uint32_t b0, uint32_t b1) {
    for (uint32_t b = b0; b < b1; b++) {
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
        m.area[nElem + b] = 0.0; m.d[nElem + b] = 0.0; m.perim[nElem + b] = 0.0;
        m.cx[nElem + b] = xr;
        m.cy[nElem + b] = yr;
        m.eCell1[e] = nElem + b;
    }
    });

    // neighbours + LSQ matrix (mesh.d:481-642)
    parallelFor(nElem, [&](uint32_t i0, uint32_t i1) {
    for (uint32_t i = i0; i < i1; i++) {
        const uint32_t ne = m.nEdgesCell[i];
        double A[MAXE][2] = {};
        for (uint32_t j = 0; j < ne; j++) {
            const uint32_t e = m.cellEdges[size_t(i) * MAXE + j];
            const uint32_t nb = (m.eCell0[e] == i) ? m.eCell1[e] : m.eCell0[e];
            m.nbr[size_t(i) * MAXE + j] = nb;
            A[j][0] = m.cx[nb] - m.cx[i];
            A[j][1] = m.cy[nb] - m.cy[i];
        }
        for (uint32_t j = ne; j < uint32_t(MAXE); j++) m.nbr[size_t(i) * MAXE + j] = 0;
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
    });
    return true;
}

// The display mesh ebb-solve writes at start-up (app/solver.d:364-372): every cell split
// into a fan of triangles (element[0], element[j+1], element[j+2]) as source/mesh.d:678-715
// triangulate does, that mesh built again (buildMesh, its own edge numbering), and its
// nodes, triangles, interior edges (nodes, both cells) and boundary edges (nodes, owner,
// boundary tag) written as source/io.d:534-643 saveMatlabMesh does: big-endian u64 counts,
// every index 1-based and stored as a double, the tag names as u32 length + bytes.
bool saveMatlabMesh(const RawMesh& raw, const std::string& path, std::string& err)
{
    RawMesh t;
    t.nodes = raw.nodes;
    t.bnodes = raw.bnodes;
    t.btags = raw.btags;
    t.bgroupStart = raw.bgroupStart;
    t.eoff.assign(1, 0);
    const size_t ne = raw.eoff.empty() ? 0 : raw.eoff.size() - 1;
    for (size_t e = 0; e < ne; e++) {
        const uint32_t* n = raw.enodes.data() + raw.eoff[e];
        const uint32_t m = raw.eoff[e + 1] - raw.eoff[e];
        if (m == 3) {
            t.enodes.insert(t.enodes.end(), n, n + 3);
            t.eoff.push_back(uint32_t(t.enodes.size()));
        } else {
            for (uint32_t j = 0; j + 2 < m; j++) {
                t.enodes.push_back(n[0]);
                t.enodes.push_back(n[j + 1]);
                t.enodes.push_back(n[j + 2]);
                t.eoff.push_back(uint32_t(t.enodes.size()));
            }
        }
    }
    BuiltMesh b;
    if (!buildMesh(t, b, err)) return false;
    std::vector<uint8_t> buf;
    auto u64 = [&](uint64_t v) { for (int i = 7; i >= 0; i--) buf.push_back(uint8_t(v >> (8 * i))); };
    auto f64 = [&](double d) { uint64_t v; std::memcpy(&v, &d, 8); u64(v); };
    const size_t nn = t.nodes.size() / 2, nt = t.eoff.size() - 1;
    buf.reserve(40 + nn * 16 + nt * 24 + size_t(b.nEdges) * 32);
    u64(nn);
    for (size_t i = 0; i < nn; i++) { f64(t.nodes[2 * i]); f64(t.nodes[2 * i + 1]); }
    u64(nt);
    for (size_t e = 0; e < nt; e++)
        for (int k = 0; k < 3; k++) f64(double(t.enodes[3 * e + size_t(k)]));   // elements are 1-based already
    uint64_t nie = 0;
    for (uint32_t e = 0; e < b.nEdges; e++) nie += b.eBGroup[e] < 0 ? 1 : 0;
    u64(nie);
    for (uint32_t e = 0; e < b.nEdges; e++)
        if (b.eBGroup[e] < 0) {
            f64(double(b.eNode0[e]) + 1); f64(double(b.eNode1[e]) + 1);
            f64(double(b.eCell0[e]) + 1); f64(double(b.eCell1[e]) + 1);
        }
    u64(uint64_t(b.nEdges) - nie);
    for (uint32_t e = 0; e < b.nEdges; e++)
        if (b.eBGroup[e] >= 0) {
            f64(double(b.eNode0[e]) + 1); f64(double(b.eNode1[e]) + 1);
            f64(double(b.eCell0[e]) + 1);
            // the index of the edge's tag among the tags (countUntil) + 1
            const std::string& tag = t.btags[size_t(b.eBGroup[e])];
            size_t ti = 0;
            while (ti < t.btags.size() && t.btags[ti] != tag) ti++;
            f64(double(ti) + 1);
        }
    u64(t.btags.size());
    for (const std::string& tag : t.btags) {
        const uint32_t L = uint32_t(tag.size());
        for (int i = 3; i >= 0; i--) buf.push_back(uint8_t(L >> (8 * i)));
        buf.insert(buf.end(), tag.begin(), tag.end());
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = "cannot write " + path; return false; }
    const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    if (w != buf.size()) { err = "short write to " + path; return false; }
    return true;
}

}  // namespace ebb

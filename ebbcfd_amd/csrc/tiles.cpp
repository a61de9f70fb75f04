// Host-side device layout and tile builder (pure host code).
#include "tiles.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>

namespace ebb {

void makeLayout(const BuiltMesh& b, Layout& L)
{
    L = Layout{};
    L.N = b.nElem;
    L.G = b.nCells - b.nElem;
    L.Fi = uint32_t(b.interiorEdges.size());
    L.Fb = uint32_t(b.boundaryEdges.size());
    uint32_t maxE = 3;
    for (uint32_t i = 0; i < L.N; i++) maxE = std::max<uint32_t>(maxE, b.nEdgesCell[i]);
    L.S = maxE <= 3 ? 3 : (maxE == 4 ? 4 : 6);
    const uint32_t S = L.S, N = L.N, F = L.Fi + L.Fb;
    L.faceOfEdge.assign(b.nEdges, 0);
    L.edgeOfFace.assign(F, 0);
    for (uint32_t k = 0; k < L.Fi; k++) { L.faceOfEdge[b.interiorEdges[k]] = k; L.edgeOfFace[k] = b.interiorEdges[k]; }
    for (uint32_t k = 0; k < L.Fb; k++) {
        L.faceOfEdge[b.boundaryEdges[k]] = L.Fi + k;
        L.edgeOfFace[L.Fi + k] = b.boundaryEdges[k];
    }
    L.nbr.resize(size_t(N) * S);
    L.cface.resize(size_t(N) * S);
    L.refId.resize(N);
    L.gm.resize(size_t(N) * 2 * S);
    L.nia.resize(N); L.cd.resize(N); L.cper.resize(N);
    for (uint32_t i = 0; i < N; i++) {
        const uint32_t ne = b.nEdgesCell[i];
        for (uint32_t j = 0; j < S; j++) {
            if (j < ne) {
                L.nbr[size_t(i) * S + j] = b.nbr[size_t(i) * MAXE + j];
                const uint32_t f = L.faceOfEdge[b.cellEdges[size_t(i) * MAXE + j]];
                L.cface[size_t(i) * S + j] = f | (b.fluxMult[size_t(i) * MAXE + j] < 0 ? 0x80000000u : 0u);
            } else {
                L.nbr[size_t(i) * S + j] = i;   // self: du = 0, dx = dy = 0
                L.cface[size_t(i) * S + j] = F; // all-zero padding face
            }
            for (int r = 0; r < 2; r++)
                L.gm[size_t(i) * 2 * S + r * S + j] =
                    j < uint32_t(MAXE) ? b.gradMat[size_t(i) * 2 * MAXE + r * MAXE + j] : 0.0;
        }
        L.nia[i] = -(1.0 / b.area[i]);
        L.cd[i] = b.d[i];
        L.cper[i] = b.perim[i];
        L.refId[i] = i;
    }
    L.cen.resize(b.nCells);
    L.fn.resize(F); L.fmid.resize(F); L.fcell.resize(F); L.flen.resize(F); L.fnode.resize(F);
    for (uint32_t k = 0; k < b.nCells; k++) L.cen[k] = make_double2(b.cx[k], b.cy[k]);
    for (uint32_t f = 0; f < F; f++) {
        const uint32_t e = L.edgeOfFace[f];
        L.fcell[f] = make_uint2(b.eCell0[e], b.eCell1[e]);
        L.fn[f] = make_double2(b.nx[e], b.ny[e]);
        L.fmid[f] = make_double2(b.mx[e], b.my[e]);
        L.flen[f] = b.len[e];
        L.fnode[f] = make_uint2(b.eNode0[e], b.eNode1[e]);
    }
}

void permuteLayout(const Layout& ref, const std::vector<uint32_t>& order, Layout& L)
{
    const uint32_t N = ref.N, S = ref.S, F = ref.Fi + ref.Fb;
    std::vector<uint32_t> dev(N);
    for (uint32_t d = 0; d < N; d++) dev[order[d]] = d;
    auto map = [&](uint32_t c) { return c < N ? dev[c] : c; };
    L = Layout{};
    L.N = ref.N; L.G = ref.G; L.Fi = ref.Fi; L.Fb = ref.Fb; L.S = ref.S;
    L.faceOfEdge = ref.faceOfEdge;
    L.edgeOfFace = ref.edgeOfFace;
    L.nbr.resize(size_t(N) * S); L.cface.resize(size_t(N) * S); L.refId.resize(N);
    L.gm.resize(size_t(N) * 2 * S); L.nia.resize(N); L.cd.resize(N); L.cper.resize(N);
    L.cen = ref.cen;
    for (uint32_t d = 0; d < N; d++) {
        const uint32_t r = order[d];
        for (uint32_t j = 0; j < S; j++) {
            L.nbr[size_t(d) * S + j] = map(ref.nbr[size_t(r) * S + j]);
            L.cface[size_t(d) * S + j] = ref.cface[size_t(r) * S + j];
        }
        for (uint32_t k = 0; k < 2 * S; k++) L.gm[size_t(d) * 2 * S + k] = ref.gm[size_t(r) * 2 * S + k];
        L.nia[d] = ref.nia[r]; L.cd[d] = ref.cd[r]; L.cper[d] = ref.cper[r];
        L.refId[d] = ref.refId[r];
        L.cen[d] = ref.cen[r];
    }
    L.fn = ref.fn; L.fmid = ref.fmid; L.flen = ref.flen; L.fnode = ref.fnode;
    L.fcell.resize(F);
    for (uint32_t f = 0; f < F; f++) L.fcell[f] = make_uint2(map(ref.fcell[f].x), map(ref.fcell[f].y));
}

namespace {

uint64_t spread21(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// interior cells in Morton order of their centroids, cut every T cells
void mortonTiles(const Layout& L, uint32_t T, std::vector<uint32_t>& order, std::vector<uint32_t>& start)
{
    const uint32_t N = L.N;
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (uint32_t i = 0; i < N; i++) {
        x0 = std::min(x0, L.cen[i].x); x1 = std::max(x1, L.cen[i].x);
        y0 = std::min(y0, L.cen[i].y); y1 = std::max(y1, L.cen[i].y);
    }
    const double sx = x1 > x0 ? 2097151.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 2097151.0 / (y1 - y0) : 0.0;
    std::vector<uint64_t> key(N);
    for (uint32_t i = 0; i < N; i++) {
        const uint64_t ux = uint64_t(std::min(2097151.0, std::max(0.0, (L.cen[i].x - x0) * sx)));
        const uint64_t uy = uint64_t(std::min(2097151.0, std::max(0.0, (L.cen[i].y - y0) * sy)));
        key[i] = spread21(ux) | (spread21(uy) << 1);
    }
    order.resize(N);
    for (uint32_t i = 0; i < N; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] != key[b] ? key[a] < key[b] : a < b; });
    start.clear();
    for (uint32_t k = 0; k < N; k += T) start.push_back(k);
    start.push_back(N);
}

// aligned blocks of 16 consecutive cells in Morton order of their centroids, T/16 per tile
void blockTiles(const Layout& L, uint32_t T, std::vector<uint32_t>& order, std::vector<uint32_t>& start)
{
    const uint32_t N = L.N, nb = (N + 15) / 16, per = std::max<uint32_t>(1, T / 16);
    std::vector<double2> bc(nb);
    for (uint32_t b = 0; b < nb; b++) {
        double x = 0, y = 0;
        const uint32_t e = std::min(N, 16 * b + 16);
        for (uint32_t i = 16 * b; i < e; i++) { x += L.cen[i].x; y += L.cen[i].y; }
        bc[b] = make_double2(x / (e - 16 * b), y / (e - 16 * b));
    }
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (uint32_t b = 0; b < nb; b++) {
        x0 = std::min(x0, bc[b].x); x1 = std::max(x1, bc[b].x);
        y0 = std::min(y0, bc[b].y); y1 = std::max(y1, bc[b].y);
    }
    const double sx = x1 > x0 ? 2097151.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 2097151.0 / (y1 - y0) : 0.0;
    std::vector<uint64_t> key(nb);
    for (uint32_t b = 0; b < nb; b++) {
        const uint64_t ux = uint64_t(std::min(2097151.0, std::max(0.0, (bc[b].x - x0) * sx)));
        const uint64_t uy = uint64_t(std::min(2097151.0, std::max(0.0, (bc[b].y - y0) * sy)));
        key[b] = spread21(ux) | (spread21(uy) << 1);
    }
    const bool shortLast = (N % 16) != 0;
    std::vector<uint32_t> blocks;
    for (uint32_t b = 0; b < nb; b++)
        if (!(shortLast && b == nb - 1)) blocks.push_back(b);
    std::sort(blocks.begin(), blocks.end(), [&](uint32_t a, uint32_t b) { return key[a] != key[b] ? key[a] < key[b] : a < b; });
    if (shortLast) blocks.push_back(nb - 1);
    order.clear();
    order.reserve(N);
    start.clear();
    for (size_t k = 0; k < blocks.size(); k++) {
        if (k % per == 0) start.push_back(uint32_t(order.size()));
        const uint32_t b = blocks[k], e = std::min(N, 16 * b + 16);
        for (uint32_t i = 16 * b; i < e; i++) order.push_back(i);
    }
    start.push_back(N);
}

// Regions of up to T cells grown greedily over interior neighbours: the next cell
// is the frontier cell with the most neighbours already in the region (oldest
// first among equals), which keeps regions compact and fills concave gaps.  The
// next seed is the unassigned neighbour of the previous region with the most
// assigned neighbours, so consecutive tiles are adjacent.
void growTiles(const Layout& L, uint32_t T, std::vector<uint32_t>& order, std::vector<uint32_t>& start)
{
    const uint32_t N = L.N, S = L.S;
    std::vector<uint8_t> taken(N, 0);
    std::vector<uint8_t> cnt(N, 0);          // neighbours inside the region being grown
    std::vector<uint32_t> cntStamp(N, ~0u);
    std::vector<std::vector<uint32_t>> bucket(S + 1);
    std::vector<size_t> head(S + 1, 0);
    order.clear();
    order.reserve(N);
    start.clear();
    uint32_t scan = 0, t = 0;
    auto takenNbrs = [&](uint32_t c) {
        uint32_t k = 0;
        for (uint32_t j = 0; j < S; j++) {
            const uint32_t nb = L.nbr[size_t(c) * S + j];
            if (nb < N && nb != c && taken[nb]) k++;
        }
        return k;
    };
    while (order.size() < N) {
        // seed: best unassigned neighbour of the previous region
        uint32_t seed = N, best = 0;
        if (!start.empty()) {
            for (size_t q = start.back(); q < order.size(); q++)
                for (uint32_t j = 0; j < S; j++) {
                    const uint32_t nb = L.nbr[size_t(order[q]) * S + j];
                    if (nb < N && !taken[nb]) {
                        const uint32_t k = takenNbrs(nb);
                        if (seed == N || k > best || (k == best && nb < seed)) { seed = nb; best = k; }
                    }
                }
        }
        if (seed == N) {
            while (taken[scan]) scan++;
            seed = scan;
        }
        const size_t b0 = order.size();
        start.push_back(uint32_t(b0));
        for (uint32_t k = 0; k <= S; k++) { bucket[k].clear(); head[k] = 0; }
        uint32_t c = seed;
        for (;;) {
            taken[c] = 1;
            order.push_back(c);
            if (order.size() - b0 >= T) break;
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t nb = L.nbr[size_t(c) * S + j];
                if (nb >= N || nb == c || taken[nb]) continue;
                if (cntStamp[nb] != t) { cntStamp[nb] = t; cnt[nb] = 0; }
                cnt[nb]++;
                bucket[cnt[nb]].push_back(nb);
            }
            c = N;
            for (uint32_t k = S; k >= 1 && c == N; k--)
                while (head[k] < bucket[k].size()) {
                    const uint32_t x = bucket[k][head[k]++];
                    if (!taken[x] && cnt[x] == k) { c = x; break; }
                }
            if (c == N) break;   // region enclosed
        }
        t++;
    }
    start.push_back(N);
}

}  // namespace

void tileOrder(const Layout& L, uint32_t T, int kind, std::vector<uint32_t>& order, std::vector<uint32_t>& start)
{
    if (L.N == 0 || T == 0) { order.clear(); start.assign(1, 0); return; }
    if (kind == TILE_MORTON) mortonTiles(L, T, order, start);
    else if (kind == TILE_BLOCKS) blockTiles(L, T, order, start);
    else growTiles(L, T, order, start);
}

uint32_t maxTileFaces(const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start)
{
    const uint32_t S = L.S, F = L.Fi + L.Fb;
    if (start.size() < 2) return 0;
    std::vector<uint32_t> fstamp(F, ~0u);
    uint32_t best = 0;
    for (uint32_t t = 0; t + 1 < start.size(); t++) {
        uint32_t nF = 0;
        for (uint32_t k = start[t]; k < start[t + 1]; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t f = L.cface[size_t(order[k]) * S + j] & 0x7fffffffu;
                if (f == F || fstamp[f] == t) continue;
                fstamp[f] = t;
                nF++;
            }
        best = std::max(best, nF);
    }
    return best;
}

bool makeFaceTiles(const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start,
                   FaceTileHost& h, std::string& err)
{
    const uint32_t N = L.N, S = L.S, F = L.Fi + L.Fb;
    h = FaceTileHost{};
    if (N == 0 || start.size() < 2) return true;
    const uint32_t nT = uint32_t(start.size() - 1);
    h.nTiles = nT;
    h.desc.resize(nT);
    std::vector<uint32_t> fstamp(F, ~0u), flidx(F, 0), fl;
    for (uint32_t t = 0; t < nT; t++) {
        const uint32_t c0 = start[t], nC = start[t + 1] - c0;
        fl.clear();
        for (uint32_t k = 0; k < nC; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t f = L.cface[size_t(order[c0 + k]) * S + j] & 0x7fffffffu;
                if (f == F || fstamp[f] == t) continue;
                fstamp[f] = t;
                flidx[f] = uint32_t(fl.size());
                fl.push_back(f);
            }
        const uint32_t nF = uint32_t(fl.size());
        if (nF >= 32767) { err = "tile has too many faces; lower tileCells"; return false; }
        h.desc[t] = make_uint4(nC, nF, uint32_t(h.ids.size()), uint32_t(h.fref.size()));
        for (uint32_t k = 0; k < nC; k++) {
            const uint32_t c = order[c0 + k];
            h.ids.push_back(c);
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t fe = L.cface[size_t(c) * S + j], f = fe & 0x7fffffffu;
                h.cf.push_back(f == F ? uint16_t(nF) : uint16_t(flidx[f] | ((fe >> 31) ? 0x8000u : 0u)));
            }
        }
        for (uint32_t k = 0; k < nF; k++) {
            const uint32_t f = fl[k];
            const uint32_t ca = L.fcell[f].x, cb = f < L.Fi ? L.fcell[f].y : N + (f - L.Fi);
            h.fref.push_back(make_uint4(ca, cb, L.fnode[f].x, L.fnode[f].y));
            h.fid.push_back(f);
        }
        h.maxF = std::max(h.maxF, nF);
        h.sumF += nF;
    }
    // fixed strides (see FaceTileArgs): tile t moves to t * cellStride and t * faceStride,
    // padding slots hold cell 0 and an all-zero face record (loaded, never used)
    uint32_t CS = 0;
    for (uint32_t t = 0; t < nT; t++) CS = std::max(CS, h.desc[t].x);
    const uint32_t FS = std::max(h.maxF, 1u), pad = 2 * FACE_TILE_LANES;
    std::vector<uint32_t> ids(size_t(nT) * CS + pad, 0u), fid(size_t(nT) * FS + pad, 0u);
    std::vector<uint16_t> cf(size_t(nT) * CS * S + size_t(pad) * S, uint16_t(0));
    std::vector<uint4> fref(size_t(nT) * FS + pad, make_uint4(0, 0, 0, 0));
    for (uint32_t t = 0; t < nT; t++) {
        const uint4 d = h.desc[t];
        std::copy(h.ids.begin() + d.z, h.ids.begin() + d.z + d.x, ids.begin() + size_t(t) * CS);
        std::copy(h.cf.begin() + size_t(d.z) * S, h.cf.begin() + size_t(d.z + d.x) * S, cf.begin() + size_t(t) * CS * S);
        std::copy(h.fref.begin() + d.w, h.fref.begin() + d.w + d.y, fref.begin() + size_t(t) * FS);
        std::copy(h.fid.begin() + d.w, h.fid.begin() + d.w + d.y, fid.begin() + size_t(t) * FS);
        h.desc[t] = make_uint4(d.x, d.y, t * CS, t * FS);
    }
    h.ids.swap(ids); h.cf.swap(cf); h.fref.swap(fref); h.fid.swap(fid);
    h.cellStride = CS;
    h.faceStride = FS;
    return true;
}

bool addGradTiles(const Layout& L, FaceTileHost& h, std::string& err)
{
    const uint32_t N = L.N, nT = h.nTiles;
    h.fslot.assign(h.fref.size(), 0u);
    h.nHalo.assign(nT, 0u);
    h.halo.clear();
    h.border.clear();
    h.maxG = 0;
    h.sumH = 0;
    if (nT == 0) { h.haloStride = 1; h.halo.assign(1, 0u); return true; }
    std::vector<uint32_t> stamp(N, ~0u), slot(N, 0u), ring, roff(nT + 1, 0u);
    std::vector<uint8_t> inRing(N, 0);
    uint32_t HS = 1;
    for (uint32_t t = 0; t < nT; t++) {
        const uint4 d = h.desc[t];
        const uint32_t nC = d.x, nF = d.y;
        for (uint32_t k = 0; k < nC; k++) {
            const uint32_t c = h.ids[d.z + k];
            stamp[c] = t;
            slot[c] = k;
        }
        uint32_t nH = 0;
        auto rowOf = [&](uint32_t c) {
            if (stamp[c] != t) {   // a ring cell, first seen
                stamp[c] = t;
                slot[c] = nC + nH++;
                ring.push_back(c);
                inRing[c] = 1;
            }
            return slot[c];
        };
        for (uint32_t k = 0; k < nF; k++) {
            const uint4 r = h.fref[d.w + k];
            const uint32_t la = rowOf(r.x);
            const uint32_t lb = r.y < N ? rowOf(r.y) : la;   // a boundary face: the ghost has no row
            h.fslot[d.w + k] = la | (lb << 16);
        }
        if (nC + nH > 65535) { err = "gradient tile has too many cells; lower tileCells"; return false; }
        h.nHalo[t] = nH;
        roff[t + 1] = roff[t] + nH;
        HS = std::max(HS, nH);
        h.maxG = std::max(h.maxG, nC + nH);
        h.sumH += nH;
    }
    // each own cell's neighbour row, at its position in ids (padding slots: cell 0's)
    h.nbrT.assign(h.ids.size() * L.S, 0u);
    for (size_t x = 0; x < h.ids.size(); x++)
        for (uint32_t j = 0; j < L.S; j++) h.nbrT[x * L.S + j] = L.nbr[size_t(h.ids[x]) * L.S + j];
    h.haloStride = HS;
    h.halo.assign(size_t(nT) * HS, 0u);
    for (uint32_t t = 0; t < nT; t++)
        std::copy(ring.begin() + roff[t], ring.begin() + roff[t + 1], h.halo.begin() + size_t(t) * HS);
    for (uint32_t c = 0; c < N; c++)
        if (inRing[c]) h.border.push_back(c);
    return true;
}

uint32_t gradTileLds(const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start,
                     uint32_t rowBytes, uint32_t faceBytes)
{
    const uint32_t N = L.N, S = L.S, F = L.Fi + L.Fb;
    if (N == 0 || start.size() < 2) return 0;
    std::vector<uint32_t> stamp(N, ~0u), fstamp(F, ~0u);
    uint32_t maxRows = 0, maxF = 0;
    for (uint32_t t = 0; t + 1 < start.size(); t++) {
        const uint32_t c0 = start[t], nC = start[t + 1] - c0;
        for (uint32_t k = 0; k < nC; k++) stamp[order[c0 + k]] = t;
        uint32_t nF = 0, nH = 0;
        for (uint32_t k = 0; k < nC; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t f = L.cface[size_t(order[c0 + k]) * S + j] & 0x7fffffffu;
                if (f == F || fstamp[f] == t) continue;
                fstamp[f] = t;
                nF++;
                if (f < L.Fi) {
                    for (uint32_t c : {L.fcell[f].x, L.fcell[f].y})
                        if (stamp[c] != t) { stamp[c] = t; nH++; }
                }
            }
        maxRows = std::max(maxRows, nC + nH);
        maxF = std::max(maxF, nF);
    }
    // the kernel's layout: the largest tile's rows, then the most faces of any tile
    const uint64_t lds = (uint64_t(maxRows) * rowBytes + 15) / 16 * 16 + (uint64_t(faceBytes) * (maxF + 1) + 15) / 16 * 16;
    return uint32_t(std::min<uint64_t>(lds, 0xFFFFFFFFull));
}

bool makeTiles(const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start, TileHost& h,
               std::string& err)
{
    const uint32_t N = L.N, S = L.S, F = L.Fi + L.Fb, Ctot = L.N + L.G;
    h = TileHost{};
    if (N == 0 || start.size() < 2) return true;
    const uint32_t nT = uint32_t(start.size() - 1);
    h.nTiles = nT;
    h.desc.resize(2 * size_t(nT));
    std::vector<uint32_t> stamp(Ctot, ~0u), lidx(Ctot, 0), fstamp(F, ~0u), flidx(F, 0), C, fl;
    std::vector<uint16_t> gown;
    std::vector<uint8_t> img;
    for (uint32_t t = 0; t < nT; t++) {
        C.clear();
        gown.clear();
        auto add = [&](uint32_t c) { stamp[c] = t; lidx[c] = uint32_t(C.size()); C.push_back(c); };
        for (uint32_t k = start[t]; k < start[t + 1]; k++) add(order[k]);
        const uint32_t nC = uint32_t(C.size());
        for (uint32_t k = 0; k < nC; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t nb = L.nbr[size_t(C[k]) * S + j];
                if (nb < N && stamp[nb] != t) add(nb);
            }
        const uint32_t nG = uint32_t(C.size());
        for (uint32_t k = nC; k < nG; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t nb = L.nbr[size_t(C[k]) * S + j];
                if (nb < N && stamp[nb] != t) add(nb);
            }
        const uint32_t nQ = uint32_t(C.size());
        for (uint32_t k = 0; k < nG; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t nb = L.nbr[size_t(C[k]) * S + j];
                if (nb >= N && stamp[nb] != t) { add(nb); gown.push_back(uint16_t(k)); }
            }
        const uint32_t nL = uint32_t(C.size());
        if (nL >= 65536) { err = "tile halo too large; lower tileCells"; return false; }
        // faces of the own cells, in first-seen order
        fl.clear();
        for (uint32_t k = 0; k < nC; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t f = L.cface[size_t(C[k]) * S + j] & 0x7fffffffu;
                if (f == F || fstamp[f] == t) continue;
                fstamp[f] = t;
                flidx[f] = uint32_t(fl.size());
                fl.push_back(f);
            }
        const uint32_t nF = uint32_t(fl.size());
        if (nF >= 32767) { err = "tile has too many faces; lower tileCells"; return false; }
        const TileImg I = tileImage(nC, nG, nQ, nL, nF, int(S));
        img.assign(I.end, 0);
        auto put = [&](uint32_t off, const void* p, size_t n) { std::memcpy(img.data() + off, p, n); };
        for (uint32_t k = 0; k < nL; k++) put(I.cs + 16 * k, &L.cen[C[k]], 16);
        for (uint32_t r = 0; r < 2; r++)
            for (uint32_t j = 0; j < S; j++)
                for (uint32_t k = 0; k < nG; k++)
                    put(I.gm + 8 * ((r * S + j) * nG + k), &L.gm[size_t(C[k]) * 2 * S + r * S + j], 8);
        for (uint32_t k = 0; k < nF; k++) {
            const uint32_t f = fl[k];
            const double g5[5] = {L.fn[f].x, L.fn[f].y, L.fmid[f].x, L.fmid[f].y, L.flen[f]};
            put(I.fs + 40 * k, g5, 40);
            const uint32_t ca = L.fcell[f].x, cb = f < L.Fi ? L.fcell[f].y : N + (f - L.Fi);
            if (stamp[ca] != t || stamp[cb] != t || lidx[ca] >= nG || (f < L.Fi && lidx[cb] >= nG)) {
                err = "tile builder: face cell outside the tile";
                return false;
            }
            const uint32_t ab[2] = {f, lidx[ca] | (lidx[cb] << 16)};
            put(I.fc + 8 * k, ab, 8);
        }
        for (uint32_t k = 0; k < nC; k++) {
            const double m3[3] = {L.cper[C[k]], L.cd[C[k]], L.nia[C[k]]};
            put(I.met + 24 * k, m3, 24);
        }
        put(I.ids, C.data(), 4 * size_t(nL));
        for (uint32_t k = 0; k < nG; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint16_t v = uint16_t(lidx[L.nbr[size_t(C[k]) * S + j]]);
                put(I.nb + 2 * (k * S + j), &v, 2);
            }
        for (uint32_t k = 0; k < nC; k++)
            for (uint32_t j = 0; j < S; j++) {
                const uint32_t fe = L.cface[size_t(C[k]) * S + j], f = fe & 0x7fffffffu;
                const uint16_t v = f == F ? uint16_t(nF) : uint16_t(flidx[f] | ((fe >> 31) ? 0x8000u : 0u));
                put(I.cf + 2 * (k * S + j), &v, 2);
            }
        if (!gown.empty()) put(I.gown, gown.data(), 2 * gown.size());
        const uint32_t boff = uint32_t(h.blob.size()), blen = I.end / 16;
        h.blob.resize(h.blob.size() + blen);
        std::memcpy(h.blob.data() + boff, img.data(), I.end);
        h.desc[2 * size_t(t)] = make_uint4(nC, nG, nQ, nL);
        h.desc[2 * size_t(t) + 1] = make_uint4(nF, boff, blen, 0);
        h.maxDyn[0] = std::max(h.maxDyn[0], tileDyn(nG, nL, 4).end);
        h.maxDyn[1] = std::max(h.maxDyn[1], tileDyn(nG, nL, 8).end);
        h.sumC += nC; h.sumG += nG; h.sumQ += nQ; h.sumL += nL; h.sumF += nF;
        h.maxL = std::max(h.maxL, nL);
        h.maxBlob = std::max(h.maxBlob, I.end);
    }
    if (h.blob.size() >= (uint64_t(1) << 32)) { err = "tile data too large"; return false; }
    for (uint32_t t = 0; t + 1 < nT; t++) {   // blob header: the next tile's descriptor
        const uint32_t boff = h.desc[2 * size_t(t) + 1].y;
        h.blob[boff] = h.desc[2 * size_t(t + 1)];
        h.blob[boff + 1] = h.desc[2 * size_t(t + 1) + 1];
    }
    return true;
}

}  // namespace ebb

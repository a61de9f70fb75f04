// Host-only driver for tests/test_grad_tiles_tables.py: builds the gradient-tile tables
// (tiles.cpp: makeFaceTiles + addGradTiles) of each mesh for several tile sizes and tile
// orders and checks what the tile kernel assumes of them:
//   * each face's two LDS rows name its two cells (own rows 0..nC-1 in tile order, ring rows
//     nC.. in the tile's ring list); a boundary face repeats cell a's row;
//   * every row index is below the tile's rows (nC + nH) and below the rows LDS is sized for;
//   * ring cells are not own cells of the tile; the border is exactly the union of the rings,
//     ascending;
//   * each own cell's copied neighbour row equals the mesh's;
//   * the LDS estimate the tile-size search uses equals the size of the built tables.
// Exit status: the number of violations.
#include <algorithm>
#include <cstdio>
#include <set>
#include <string>

#include "../ebbcfd_amd/csrc/host.hpp"
#include "../ebbcfd_amd/csrc/tiles.hpp"

using namespace ebb;

int main(int argc, char** argv)
{
    int bad = 0;
    for (int a = 1; a < argc; a++) {
        RawMesh raw;
        BuiltMesh b;
        std::string err;
        if (!parseGmsh(argv[a], raw, err) || !buildMesh(raw, b, err)) {
            std::printf("load failed %s: %s\n", argv[a], err.c_str());
            bad++;
            continue;
        }
        Layout L;
        makeLayout(b, L);
        for (int T : {1, 7, 32, 256, 768}) {
            for (int kind : {TILE_GROW, TILE_MORTON, TILE_BLOCKS}) {
                std::vector<uint32_t> order, start;
                tileOrder(L, uint32_t(T), kind, order, start);
                FaceTileHost h;
                if (!makeFaceTiles(L, order, start, h, err) || !addGradTiles(L, h, err)) {
                    std::printf("build failed %s T %d: %s\n", argv[a], T, err.c_str());
                    bad++;
                    continue;
                }
                int before = bad;
                std::set<uint32_t> ring;
                uint32_t maxF = 0;
                for (uint32_t t = 0; t < h.nTiles; t++) {
                    const uint4 d = h.desc[t];
                    const uint32_t nC = d.x, nF = d.y, nH = h.nHalo[t];
                    maxF = std::max(maxF, nF);
                    if (nC + nH > h.maxG || nH > h.haloStride) bad++;
                    auto cellOf = [&](uint32_t r) {
                        return r < nC ? h.ids[d.z + r] : h.halo[size_t(t) * h.haloStride + r - nC];
                    };
                    for (uint32_t k = 0; k < nF; k++) {
                        const uint4 f = h.fref[d.w + k];
                        const uint32_t sl = h.fslot[d.w + k], la = sl & 0xffffu, lb = sl >> 16;
                        if (la >= nC + nH || lb >= nC + nH) { bad++; continue; }
                        if (cellOf(la) != f.x) bad++;
                        if (f.y < L.N ? cellOf(lb) != f.y : lb != la) bad++;
                    }
                    std::set<uint32_t> own;
                    for (uint32_t k = 0; k < nC; k++) {
                        const uint32_t c = h.ids[d.z + k];
                        own.insert(c);
                        for (uint32_t j = 0; j < L.S; j++)
                            if (h.nbrT[size_t(d.z + k) * L.S + j] != L.nbr[size_t(c) * L.S + j]) bad++;
                    }
                    for (uint32_t x = 0; x < nH; x++) {
                        const uint32_t c = h.halo[size_t(t) * h.haloStride + x];
                        if (own.count(c)) bad++;
                        ring.insert(c);
                    }
                }
                if (std::vector<uint32_t>(ring.begin(), ring.end()) != h.border) bad++;
                if (h.fslot.size() != h.fref.size() || h.nbrT.size() != h.ids.size() * L.S) bad++;
                for (const uint32_t rowBytes : {48u, 96u})
                    for (const uint32_t faceBytes : {28u, 40u}) {
                        const uint32_t built = (h.maxG * rowBytes + 15) / 16 * 16 + (faceBytes * (maxF + 1) + 15) / 16 * 16;
                        if (gradTileLds(L, order, start, rowBytes, faceBytes) != built) bad++;
                    }
                std::printf("%s T %d order %d: tiles %u, border %zu of %u cells, %s\n", argv[a], T, kind, h.nTiles,
                            h.border.size(), L.N, bad == before ? "ok" : "VIOLATIONS");
            }
        }
    }
    return bad;
}

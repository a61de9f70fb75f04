// Host-side device layout of a built mesh and the tiles of the fused stage kernel.
// Pure host code (no HIP calls), so it also runs where there is no GPU.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "dev.hpp"
#include "host.hpp"

namespace ebb {

// Arrays in device order (see dev.hpp): cells 0..N-1 interior, N.. ghosts;
// faces 0..Fi-1 interior, Fi..Fi+Fb-1 boundary, Fi+Fb the zero padding face.
struct Layout {
    uint32_t N = 0, G = 0, Fi = 0, Fb = 0, S = 3;
    std::vector<uint32_t> faceOfEdge, edgeOfFace;
    std::vector<uint32_t> nbr, cface, refId;        // N*S, N*S, N
    std::vector<double> gm, nia, cd, cper;           // N*2S, N, N, N
    std::vector<double2> cen, fn, fmid;              // N+G, F, F
    std::vector<uint2> fcell;                        // F
    std::vector<double> flen;                        // F
    std::vector<uint2> fnode;                        // F: the edge's two mesh nodes, in mesh order
};
void makeLayout(const BuiltMesh& b, Layout& L);
// The same layout with interior cells renumbered: device cell d is cell order[d]
// of `ref` (ghosts and faces keep their numbers); refId holds the reference ids.
void permuteLayout(const Layout& ref, const std::vector<uint32_t>& order, Layout& dev);

enum TileOrder { TILE_GROW = 0, TILE_MORTON = 1, TILE_BLOCKS = 2 };

struct TileHost {
    std::vector<uint4> desc, blob;
    uint32_t nTiles = 0;
    uint32_t maxDyn[2] = {0, 0};   // largest working set: scalar limiter layout, LP layout
    // totals for statistics
    uint64_t sumC = 0, sumG = 0, sumQ = 0, sumL = 0, sumF = 0;
    uint32_t maxL = 0;
    uint32_t maxBlob = 0;          // bytes
};

struct FaceTileHost {
    std::vector<uint4> desc, fref;
    std::vector<uint32_t> ids, fid;
    std::vector<uint16_t> cf;
    uint32_t nTiles = 0, maxF = 0;
    uint32_t cellStride = 0, faceStride = 0;   // tile t: cells from t * cellStride, faces from t * faceStride
    uint64_t sumF = 0;
    // gradient tiles (addGradTiles): LDS rows of each face's cells, each tile's ring, the border
    std::vector<uint32_t> fslot, halo, nHalo, border, nbrT;
    uint32_t haloStride = 0, maxG = 0;
    uint64_t sumH = 0;
};

// Cut the interior cells into tiles of at most T cells (TILE_GROW: compact regions
// grown greedily over the cell graph; TILE_MORTON: runs of the Morton order of the
// centroids; TILE_BLOCKS: runs of T/16 aligned blocks of 16 consecutive cells in
// Morton order of the block centroids, each block in its own order, the short last
// block last): tile t is cells order[start[t] .. start[t+1]).
void tileOrder(const Layout& L, uint32_t T, int kind, std::vector<uint32_t>& order, std::vector<uint32_t>& start);

// Largest number of distinct faces of one tile (the face tile kernel's LDS holds 40 bytes
// per face).
uint32_t maxTileFaces(const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start);

// Face tiles: each cell group with the faces of its cells.
bool makeFaceTiles(const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start,
                   FaceTileHost& h, std::string& err);

// Gradient tiles on face tiles built by makeFaceTiles: each tile's own cells take LDS rows
// 0..nC-1 in tile order, the cells across its faces that other tiles own (its ring) rows
// nC.. in order of first appearance among its faces; the border is every ring cell.
bool addGradTiles(const Layout& L, FaceTileHost& h, std::string& err);
// Per-workgroup LDS of gradient tiles with `rowBytes` per gradient row: the rows of the tile
// with the most rows plus the face results (`faceBytes` per face and a padding face) of the
// tile with the most faces, without building the tiles.
uint32_t gradTileLds(const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start,
                     uint32_t rowBytes, uint32_t faceBytes);

// Fused-kernel tiles: each cell group with its halo, tables and static data packed.
bool makeTiles(const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start, TileHost& h,
               std::string& err);

}  // namespace ebb

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
};
void makeLayout(const BuiltMesh& b, Layout& L);

enum TileOrder { TILE_GROW = 0, TILE_MORTON = 1 };

struct TileHost {
    std::vector<uint4> desc, blob;
    uint32_t nTiles = 0;
    uint32_t maxEnd[2] = {0, 0};   // largest LDS image: scalar limiter layout, LP layout
    // totals for statistics
    uint64_t sumC = 0, sumG = 0, sumQ = 0, sumL = 0, sumF = 0;
    uint32_t maxL = 0;
    uint32_t maxBlob = 0;          // bytes
};

// Cut the interior cells into tiles of at most T cells (TILE_GROW: compact regions
// grown greedily over the cell graph; TILE_MORTON: runs of the Morton order of
// the centroids) and pack each tile's halo, tables and static data.
bool makeTiles(const Layout& L, uint32_t T, int order, TileHost& h, std::string& err);

}  // namespace ebb

// Host-side data model: raw mesh, built geometry (reference numbering), config.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace ebb {

constexpr int MAXE = 6;            // source/mesh.d:26 MAX_EDGES
constexpr double GAMMA = 1.4;      // source/euler.d:89

enum BoundaryType { FullState = 0, ConstPressure, InviscidWall, ViscousWall, Symmetry, TempPresInflow, Dirichlet };
enum FluxKind { RusanovFlux = 0, RoeFlux = 1 };
enum IntegratorKind { IntEuler = 0, IntRK2 = 1, IntRK2TVD = 2, IntRK4 = 3 };

// Mesh as read from a file (source/io.d:689-990) -- buildMesh inputs.
struct RawMesh {
    std::vector<double> nodes;          // 2 per node
    std::vector<uint32_t> eoff;         // element offsets, size nElem+1
    std::vector<uint32_t> enodes;       // 1-based node ids, counter-clockwise
    std::vector<uint32_t> bnodes;       // 2 per boundary line, 0-based
    std::vector<std::string> btags;     // boundary tag names
    std::vector<uint64_t> bgroupStart;  // first bnode of each tag
};

// Allocator whose resize() leaves new elements uninitialised, so that large arrays
// every element of which is written anyway are first touched by the threads that
// fill them instead of being zero-filled by one.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    template <typename U> struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <typename U> NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    template <typename U> void construct(U* p) noexcept { ::new (static_cast<void*>(p)) U; }
    template <typename U, typename... A> void construct(U* p, A&&... a) { ::new (static_cast<void*>(p)) U(std::forward<A>(a)...); }
};
template <typename T> using FillVec = std::vector<T, NoInitAlloc<T>>;

// buildMesh output (source/mesh.d:217-643), reference cell/edge numbering.
struct BuiltMesh {
    uint32_t nElem = 0, nCells = 0, nEdges = 0;
    // cells (ghosts have index >= nElem, one per boundary edge in boundaryEdges order)
    FillVec<uint8_t> nEdgesCell;         // nElem
    FillVec<uint32_t> cellEdges;         // MAXE per interior cell (0 past the cell's edges)
    FillVec<int8_t> fluxMult;            // MAXE per interior cell (+1 owner, -1 otherwise)
    FillVec<uint32_t> nbr;               // MAXE per interior cell (0 past the cell's edges)
    FillVec<double> gradMat;             // 2*MAXE per interior cell, row-major (dir, nbr)
    FillVec<double> area, d, perim;      // nCells (ghosts zero)
    FillVec<double> cx, cy;              // nCells (ghost centroids reflected)
    // edges
    FillVec<uint32_t> eNode0, eNode1, eCell0, eCell1;
    FillVec<double> len, nx, ny, mx, my;
    FillVec<int32_t> eBGroup;            // -1 interior
    std::vector<uint32_t> interiorEdges, boundaryEdges;
    std::vector<std::vector<uint32_t>> bGroups;
};

struct BCConfig {
    std::string tag;
    int type = FullState;
    std::vector<double> data;
};

// source/config.d:72-94 + loadConfig defaults (config.d:112-313)
struct Config {
    std::string meshFile, flux, integrator = "Euler";
    double dt = 0, tEnd = 0, CFL = 0.5, lpThresh = -1;
    long long saveIter = -1, plotIter = -1, order = 2;
    std::vector<std::string> forceBoundary;
    bool dynamicDt = true, limited = true, localTimestep = false, multistageLimiting = true;
    bool cflAdjust = true, viscosity = false;
    double ic[4] = {0, 0, 0, 0};
    double Pr = 0.7, R = 8.3145, gamma = 1.4, mu = 0, Re = 1000, L = 1;
    std::vector<BCConfig> bcs;
    // extensions (namespaced key "extensions"; absent from reference configs)
    bool mms = false;                 // manufactured solution: dirichlet = exact, source per stage
    long long maxSteps = -1;          // stop after this many steps (-1: only tEnd)
    double rmsDrop = 0;               // stop once the rho RMS residual drops by this factor below its
                                      // first-step value (0: never)
    int precision = 0;                // 0 = fp64, 1 = mixed fp32 flux/reconstruction/limiter
    int limiter = 0;                  // 0: the reference's scalar limiter; 1: Venkatakrishnan
    double venkatK = 1.0;             // Venkatakrishnan constant K (eps^2 = (K d)^3)
    int stagePath = 1;                // 0: gradient, face, update kernels; 1: gradient + face/update tiles;
                                      // 2: one fused kernel per stage; 3: border gradients + gradient tiles
    long long tileCells = -1;         // cells per tile of the tiled paths (-1: automatic)
    bool spreadTiles = true;          // automatic tiles: smaller tiles on small meshes, so they reach
                                      // every CU
    int64_t spreadOver = 0;           // cells of other solves whose tiles share this one's launches
                                      // (batched members): counted with the mesh's own when spreading
    int kernelOrder = 0;              // StageArgs::order: which kernels walk their work backwards
    int tileOrder = 2;                // 0: greedily grown regions, 1: Morton order of centroids,
                                      // 2: Morton order of aligned 16-cell blocks
    bool renumber = false;            // tiled paths: number device cells tile by tile
    std::vector<std::string> warnings;
    int fluxKind() const { return flux == "roeFlux" ? RoeFlux : RusanovFlux; }
    int integratorKind() const {
        if (integrator == "RK2") return IntRK2;
        if (integrator == "RK2_TVD") return IntRK2TVD;
        if (integrator == "RK4") return IntRK4;
        return IntEuler;
    }
};

// mesh.cpp
bool parseGmsh(const std::string& path, RawMesh& out, std::string& err);
bool parseGri(const std::string& path, RawMesh& out, std::string& err);
bool buildMesh(const RawMesh& raw, BuiltMesh& out, std::string& err);
// triangulate + buildMesh + saveMatlabMesh (source/mesh.d:678-715, source/io.d:534-643)
bool saveMatlabMesh(const RawMesh& raw, const std::string& path, std::string& err);

// config.cpp
bool parseConfig(const std::string& json, Config& out, std::string& err);

// mms.cpp (source/manufacturedsolution.d)
void mmsSolution(double x, double y, double q[4]);
void mmsSolutionGradient(double x, double y, double g[8]);
void mmsSourceTerm(double x, double y, bool viscosity, double R, double mu, double Pr, double S[4]);
void freestreamState(double M, double aoaDeg, double rho, double q[4]);
double aoaRadians(double aoaDeg);

// io.cpp (.esln v2, .frc)
uint32_t crc32(const uint8_t* p, size_t n);

}  // namespace ebb

// Device data layout shared by the kernels (kernels.hip) and the host driver (solver.cpp).
//
// Cells: interior cells 0..N-1 in reference order, ghosts N..N+G-1 in reference
// boundary-edge order (source/mesh.d:329-381).  Faces: interior faces 0..Fi-1 in
// reference interior-edge order, boundary faces Fi..Fi+Fb-1 in boundary-edge
// order (boundary face b owns ghost N+b), then one all-zero padding face used by
// cells with fewer than S edges.  Every per-cell array is fixed stride S (3 for
// all-triangle meshes, 4 with quads, 6 otherwise); a short cell repeats itself as
// neighbour and points at the padding face, which leaves every sum unchanged.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace ebb {

struct BFace {
    int32_t type;          // BoundaryType
    int32_t vertical;      // InviscidWall/Symmetry: edge is vertical (reflection = x-flip)
    double R[4];           // reflection matrix about the edge line (row-major)
    double far[4];         // FullState far state / Dirichlet edge state (edge.q[1])
    double gq[4];          // Dirichlet ghost state at the ghost centroid
    double gg[8];          // Dirichlet ghost gradient (var, dir)
    double nrm[2];         // edge unit normal (owner outward)
    double data[4];        // ConstP: pb; TPInflow: Tt, Pt, cos(aoa), sin(aoa)
    double Rgas;           // gas constant (TPInflow)
    double pad;
};

// Device-resident scalars of the time loop (one per solver).
struct DevScalars {
    unsigned long long newDtBits;   // min over cells of CFL*d/sAve, as ordered bits
    unsigned long long rmaxBits;    // max |R| before scaling, as ordered bits
    unsigned long long errEdge;     // (stage << 32) | edge key, min wins; ~0 = none
    unsigned long long errCell;     // (step << 32) | reference cell, min wins; ~0 = none
    double dt;                      // dt used by the stages of the current step
    double t;                       // simulation time
    double CFL;                     // current CFL (cflAdjust lowers it)
    double residLast;               // residMax of the previous step (cflAdjust)
    int64_t incIters;               // cflAdjust counter
    int64_t steps;                  // completed steps
    int32_t active;                 // 0 once t >= tEnd, maxSteps, convergence or an error was recorded
    int32_t converged;              // 1 once the rho RMS dropped by rmsDrop below its first-step value
    double rms[4];
    double force[2];
    double fsum[2];                 // boundary force sums of the current step, before the AoA rotation
    int32_t forcePending;           // 1: the last step's boundary force is still to be summed
    int32_t pad2;
    int64_t forceSlot;              // its history row (index into the history buffer), -1: none
    unsigned long long errBC;       // (stage << 32) | boundary face whose total-conditions inflow has no
                                    // non-negative Mach root (solve.d:648-671), min wins; ~0 = none
    double rms0;                    // rho RMS residual of the first step (convergence stop)
};

// The ordered boundary-force sum of a time-loop step (mesh.d:645-671), when it is
// left to the next gradient kernel: one extra workgroup of that kernel sums it while
// the others compute gradients, so its long dependent chain of additions is off the
// step's critical path.  The face fluxes it reads are only rewritten by the tile
// kernel that follows.
struct ForceArgs {
    const uint32_t* faces;   // force-boundary faces in bGroups order
    uint32_t n;              // 0: no force boundary (no extra workgroup)
    const double* lenF;      // len*flux per face
    double rot[4];           // freestream AoA rotation (row-major)
    double* hist;            // history rows (HIST_FIELDS doubles each)
};

struct DevMesh {
    uint32_t N, G, Fi, Fb;
    const uint32_t* __restrict__ nbr;    // N*S neighbour cells (ghosts >= N)
    const uint32_t* __restrict__ cface;  // N*S face ids, bit 31 set where the cell is not the owner
    const double2* __restrict__ cen;     // N+G centroids
    const double* __restrict__ nia;      // N: -(1.0/area)
    const double* __restrict__ cd;       // N: d = 2A/perim
    const double* __restrict__ cper;     // N: perim
    const uint2* __restrict__ fcell;     // Fi+Fb: (owner, other)
    const double2* __restrict__ fn;      // Fi+Fb: unit normal (owner outward)
    const double2* __restrict__ fmid;    // Fi+Fb: midpoint
    const double* __restrict__ flen;     // Fi+Fb: length
    const BFace* __restrict__ bface;     // Fb
    const double4* __restrict__ src;     // N MMS source (null if none)
    const uint32_t* __restrict__ refId;  // N: reference index of each device cell
    const double2* __restrict__ node;    // mesh node coordinates
};

struct PhysParams {
    int flux;            // FluxKind
    int viscous;
    int order;
    int limited;
    double Pr, mu, Rgas, lpThresh;
    int localTimestep;
    int mixed;           // 1: fp32 flux/reconstruction/limiter
    int venkat;          // 1: Venkatakrishnan limiter in place of the reference's scalar limiter
    double venkatK;      // its constant K: eps^2 = (K d)^3 with d = 2 area / perimeter of the cell
};

// One stage of an integrator.  `kind` selects the update applied to the residual:
enum StageKind {
    SK_RESIDUAL = 0,   // write R only (ufvmSolver)
    SK_EULER,          // qout = qn + h*R
    SK_RK2_1,          // k1 = R; T = qn + (dt*R)
    SK_RK2_2,          // qout = qn + (h/2)*(k1 + R)
    SK_TVD_1,          // T = qn + h*R
    SK_TVD_2,          // qout = 0.5*((qn + T) + h*R)
    SK_RK4_1,          // acc = R; T = qn + ((dt/2)*R)
    SK_RK4_2,          // acc = acc + 2R; T = qn + ((dt/2)*R)
    SK_RK4_3,          // acc = acc + 2R; T = qn + (dt*R)
    SK_RK4_4,          // qout = qn + (h/6)*(acc + R)
};

struct StageArgs {
    double4* qin;          // stage input (ghost slots written by the gradient kernel)
    const double4* qn;     // state at the start of the step
    double4* qout;         // stage output (T or q^{n+1})
    double4* acc;          // k1 (RK2) or the RK4 accumulator
    double4* R;            // residual output (SK_RESIDUAL) or null
    double* grad;          // N*8 (var, dir)
    double* lim;           // N*4 (scalar) or N*8 (LP)
    double* lenF;          // (F+1)*4: len*flux per face
    double* lenS;          // (F+1): len*sMax per face
    double4* estate;       // optional (F)*2 edge states for .esln, null otherwise
    // mixed precision (PhysParams::mixed): FP32 shadow of the stage input (ghost slots written
    // by the gradient kernel) and of the stage output, FP32 gradient (N*8) and limiter (N*4)
    // rows; null on the FP64 path
    float4* qin32;
    float4* qout32;
    float* grad32;
    float* lim32;
    double* celldt;        // N local time step (localTimestep)
    double* partial;       // 2 per update workgroup: min dt candidate, max |R|
    double* rmsBlocks;     // last stage of a time-loop step: per 16-cell reference block, the 8 leaf
                           // sums of the RMS pairwise sum (null: not computed here)
    DevScalars* sc;
    int kind;              // StageKind
    int limit;             // limiter on for this stage
    int dtUpdate;          // local dt update on this stage
    int first;             // first stage of a step (resets Rmax)
    int last;              // last stage of a step (NaN scan of the new state)
    int checkActive;       // honour sc->active (time loop) or always run
    int deferReduce;       // leave the partials for the step's RMS kernel instead of reducing them now
    int recordErrors;
    uint32_t stageId;      // global stage counter for error ordering
    uint32_t stepId;
    int order;             // bit 0: the face/update tile kernel walks each XCD's slice of tiles
                           // last to first; bit 1: the gradient kernel its slice of cell blocks
    ForceArgs force;       // gradient kernels: the pending boundary force, if any
};

// Tiles for the fused stage kernel: each tile is a compact group of interior
// cells plus the halo it needs: local cells [own | ring 1 | ring 2 | ghosts of
// own + ring 1].  Gradients are computed for own + ring 1, fluxes for every face
// of an own cell, updates for own cells; faces shared by two tiles are computed
// by both, identically.  Everything a tile reads besides the states is packed
// into one contiguous blob that is copied into LDS as is (layout: tileImage).
struct TileArgs {
    uint32_t nTiles;
    uint32_t nBlocks;              // persistent workgroups, each walking a contiguous run of tiles
    uint32_t slotBytes;            // LDS per blob buffer (largest blob); two buffers alternate
    uint32_t ldsBytes;             // LDS per workgroup: two blob buffers + the largest working set
    const uint4* desc;             // 2 per tile: (nC, nG, nQ, nL), (nF, blob offset, blob length, 0) in 16 B;
                                   // each blob starts with the descriptor of the tile after it
    const uint4* blob;
};

// Byte offsets inside a tile's blob, which is copied into LDS as is.
struct TileImg {
    uint32_t hdr, cs, gm, fs, met, fc, ids, nb, cf, gown, end;
};
__host__ __device__ inline TileImg tileImage(uint32_t nC, uint32_t nG, uint32_t nQ, uint32_t nL, uint32_t nF, int S)
{
    auto al = [](uint32_t x) { return (x + 15u) & ~15u; };
    TileImg I;
    uint32_t o = 0;
    I.hdr = o; o += 32;                             // the next tile's descriptor (2 x uint4)
    I.cs = o; o = al(o + 16 * nL);                  // centroids (double2)
    I.gm = o; o = al(o + 8 * 2 * S * nG);           // LSQ weights, (dir, nbr)-major over own + ring 1
    I.fs = o; o = al(o + 40 * (nF + 1));            // faces: normal, midpoint, length; then len*flux, len*sMax
    I.met = o; o = al(o + 24 * nC);                 // own cells: perimeter, d, -1/area
    I.fc = o; o = al(o + 8 * nF);                   // faces: (global face id, local a | local b << 16)
    I.ids = o; o = al(o + 4 * nL);                  // global cell ids
    I.nb = o; o = al(o + 2 * S * nG);               // local neighbours of own + ring 1
    I.cf = o; o = al(o + 2 * S * nC);               // own cells' local faces, bit 15: not the owner
    I.gown = o; o = al(o + 2 * (nL - nQ));          // ghosts: local index of the owner cell
    I.end = o;
    return I;
}

// Byte offsets of a tile's working set (computed in the kernel, not stored).
struct TileDyn {
    uint32_t qs, gs, ls, end;
};
__host__ __device__ inline TileDyn tileDyn(uint32_t nG, uint32_t nL, int LW)
{
    TileDyn D;
    D.qs = 0;                                       // states
    D.gs = D.qs + 32 * nL;                          // gradients (var, dir)
    D.ls = D.gs + 64 * nG;                          // limiters
    D.end = D.ls + 8 * LW * nG;
    return D;
}

// Face tiles for the face+update kernel: groups of interior cells; a workgroup
// computes every face of its cells into LDS (faces between two tiles are
// computed by both, identically) and then updates its cells.  Gradients,
// limiters and ghost states come from the gradient kernel.  Per tile, the face
// references are stored contiguously.
// Lanes of a face tile workgroup.  Tile t's cells start at t * cellStride and its faces at
// t * faceStride, and the arrays carry 2 * FACE_TILE_LANES padding entries after the last
// tile, so a workgroup's first loads (two rounds of cell ids, one of face records) need not
// wait for the descriptor and stay in bounds.
constexpr uint32_t FACE_TILE_LANES = 256;
// LDS per face slot of a face tile's results: len*flux and len*sMax in FP64 (40 B), or on the
// mixed path the FP32 flux and sMax with the FP64 length (28 B); a tile's LDS holds
// faceStride + 1 slots (the padding face last)
__host__ __device__ constexpr uint32_t faceLdsBytes(int mixed) { return mixed ? 28u : 40u; }

struct FaceTileArgs {
    uint32_t nTiles;
    uint32_t ldsBytes;
    uint32_t cellStride, faceStride;
    const uint4* desc;             // per tile: nC, nF, first cell (t * cellStride), first face (t * faceStride)
    const uint32_t* ids;           // own cells' device ids
    const uint16_t* cf;            // S per own cell: local face, bit 15 set where the cell is not the owner;
                                   // nF (an all-zero slot) for padding
    const uint4* fref;             // per face: cell a, cell b, node 0, node 1 (normal, midpoint and length are
                                   // recomputed from the nodes); boundary face b has the ghost N + b as cell b
    const uint32_t* fid;           // per face: global face id, read only to report a NaN flux
    // Gradient tiles (the same tiles, gradients computed inside the tile kernel): each tile's
    // own cells' gradient and limiter rows are computed into LDS slots 0..nC-1; the rows of
    // its ring (the cells across its faces that other tiles own) are copied into slots
    // nC..nC+nH-1 from what the border kernel computed.  Null on the face-tile path.
    const uint32_t* fslot;         // per face: LDS row of cell a (low 16 bits) and of cell b (high 16 bits;
                                   // a boundary face repeats cell a's)
    const uint32_t* halo;          // per tile, from t * haloStride: the ring cells, in slot order
    const uint32_t* nHalo;         // per tile: ring cells
    const uint32_t* border;        // every cell in some tile's ring, ascending
    const uint32_t* nbrT;          // S per own cell, laid out like ids: the cell's neighbour ids
    uint32_t haloStride, nBorder;
    uint32_t glDoubles;            // LDS offset, in doubles, of the face results after the gradient rows
    uint32_t pad;
};

// Launchers (kernels.hip).  All run on `stream`.
void launchFaceTiles(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, int S,
                     hipStream_t stream);
void launchStageTiles(const DevMesh& m, const PhysParams& p, const StageArgs& a, const TileArgs& t, int S,
                      hipStream_t stream);
// gradient tiles: the border kernel (ring cells' gradients; the force and Rmax duties of the
// gradient kernel) then the tile kernel (own gradients, faces and update); parts 1: the border
// kernel only, 2: the tile kernel (and the stage reduction) only
void launchGradTiles(const DevMesh& m, const PhysParams& p, const StageArgs& a, const FaceTileArgs& t, int S,
                     hipStream_t stream, int parts = 3);
void launchGradient(const DevMesh& m, const PhysParams& p, const StageArgs& a, int S, hipStream_t stream);
void launchFaces(const DevMesh& m, const PhysParams& p, const StageArgs& a, hipStream_t stream);
void launchUpdate(const DevMesh& m, const PhysParams& p, const StageArgs& a, int S, hipStream_t stream);

struct DiagArgs {
    const double4* qnew;     // state after the step (device order)
    const double4* qold;     // state before the step
    uint32_t N, L;           // interior cells, cells incl. ghosts (reference array length)
    const uint32_t* perm;    // reference index -> device index (null = identity)
    const double* blockSums; // per 16-element block, its 8 leaf sums from the last stage (null: read the states)
    double* chunkTrees;      // nChunks * 8 * 16 partial pairwise trees
    uint32_t* chunkMask;     // nChunks * 8 bit masks of the levels present
    const uint32_t* forceFaces; uint32_t nForce;
    const double* lenF;
    double rot[4];           // freestream AoA rotation (row-major)
    double tEnd;
    int dynamicDt, cflAdjust;
    double* hist;            // per step: t, dt, rms[4], Rmax, drag, lift, CFL
    int64_t histBase;        // step index of hist[0]
    int64_t histCap;
    int64_t maxSteps;        // stop after this many completed steps (<0: none)
    // stage partials left by the step's stages (deferReduce): stage s holds nPart
    // (min dt candidate, max |R|) pairs at partial + s * pstride; nStages 0: none
    const double* partial;
    uint32_t nPart, pstride, nStages;
    int deferForce;          // 1: the force is left pending for the next gradient kernel
    int histRing;            // 1: history row = steps % histCap (batched runs) instead of steps - histBase
    double rmsDrop;          // > 0: stop once rho RMS <= rmsDrop * its first-step value
    DevScalars* sc;
};
constexpr int HIST_FIELDS = 10;
constexpr int HIST_RING_ROWS = 256;   // rows of a solver's device history ring

// Batched runs (many meshes per launch): the per-mesh arguments of every kernel of a
// time-loop step, kept in device memory, one record per mesh.  A work item (mesh, local
// index) names a cell block of the gradient kernel (FORCE_ITEM: that mesh's pending
// boundary force), a face tile, or a workgroup of the RMS kernel; each workgroup runs the
// solo kernels' code on its item, so every mesh's results equal its solo run's.
struct MemberArgs {
    DevMesh m;
    PhysParams p;
    FaceTileArgs t;
    StageArgs st[2][4];      // [parity][stage in step]: parity 0 reads the state buffer qA
    DiagArgs d[2];           // [parity] the step's RMS kernel and epilogue
    ForceArgs f;             // the pending boundary force at the end of a run
    uint32_t parityFlip;     // 1: the mesh's current state is in qB (its parity is the batch's ^ 1)
    uint32_t nRmsFull;       // full RMS chunks (k_rms_blocks' nfull)
    uint32_t pad[2];
};
constexpr uint32_t FORCE_ITEM = 0xFFFFFFFFu;
constexpr uint32_t RMS_CHUNK_BLOCKS = 1024;     // 16-element blocks per workgroup
void launchRmsChunks(const DiagArgs& d, hipStream_t stream);
// sums a pending boundary force (end of a time-loop batch); no-op when none is pending
void launchPendingForce(const ForceArgs& f, DevScalars* sc, hipStream_t stream);
void launchFinalize(const DiagArgs& d, hipStream_t stream);
void launchInitScalars(DevScalars* sc, double dt, double t, double CFL, hipStream_t stream);

// Batched launches over many meshes (one launch per kernel for all of them).  The kernel
// shape (S, flux, viscous, order, limiter mode, precision) is the group's; items are
// (mesh record, local index).  parity selects each mesh's state buffers (see MemberArgs).
struct GroupShape {
    int S, flux, viscous, order, limited, venkat, mixed;
    int gradTiles;   // 1: gradient items are border blocks and tiles compute their own gradients
};
void launchGradientBatch(const GroupShape& g, int limitStage, const MemberArgs* mem, const uint2* items,
                         uint32_t nItems, uint32_t parity, uint32_t stage, hipStream_t stream);
void launchFaceTilesBatch(const GroupShape& g, int limitStage, const MemberArgs* mem, const uint2* items,
                          uint32_t nItems, uint32_t ldsBytes, uint32_t parity, uint32_t stage, hipStream_t stream);
void launchRmsBatch(const MemberArgs* mem, const uint2* items, uint32_t nItems, uint32_t parity, hipStream_t stream);
void launchFinalizeBatch(const MemberArgs* mem, const uint32_t* members, uint32_t n, uint32_t parity,
                         hipStream_t stream);
void launchPendingForceBatch(const MemberArgs* mem, const uint32_t* members, uint32_t n, hipStream_t stream);
// copies each listed mesh's DevScalars into out[k] (one host copy for all)
void launchGatherScalars(const MemberArgs* mem, const uint32_t* members, uint32_t n, DevScalars* out,
                         hipStream_t stream);
// FP32 shadow q32[0..n) of q[0..n) (mixed precision)
void launchShadow(const double4* q, float4* q32, uint32_t n, hipStream_t stream);

}  // namespace ebb

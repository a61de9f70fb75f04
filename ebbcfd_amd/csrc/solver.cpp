// C ABI: mesh, config and the device-resident solver.
//
// Set-up follows source/solve.d:175-279 (ufvmSetup); the per-stage launch
// sequence follows the integrators (source/integrators.d:22-270); ebb_run is
// the driver loop of app/solver.d:119-250 with every per-step quantity (dt, t,
// RMS residuals, forces, cflAdjust) kept on the device, so the host only syncs
// once per batch of steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "solver_impl.hpp"

namespace ebb {
thread_local std::string lastDeviceError;
}  // namespace ebb

namespace {

void putBE32(std::vector<uint8_t>& b, uint32_t v)
{
    for (int i = 3; i >= 0; i--) b.push_back(uint8_t(v >> (8 * i)));
}
void putBE64(std::vector<uint8_t>& b, uint64_t v)
{
    for (int i = 7; i >= 0; i--) b.push_back(uint8_t(v >> (8 * i)));
}
void putBEd(std::vector<uint8_t>& b, double d)
{
    uint64_t v;
    std::memcpy(&v, &d, 8);
    putBE64(b, v);
}
uint32_t getBE32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }
double getBEd(const uint8_t* p)
{
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v = (v << 8) | p[i];
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}

}  // namespace

namespace ebb {
uint32_t crc32(const uint8_t* p, size_t n)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}
}  // namespace ebb

namespace ebb {

int fail(ebb_solver* s, int code, const std::string& msg)
{
    if (s) { s->errKind = code; s->errMsg = msg; }
    return code;
}

template <typename T>
hipError_t trackAlloc(ebb_solver* s, DBuf<T>& b, size_t n)
{
    hipError_t e = b.alloc(n);
    s->deviceBytes += int64_t((n ? n : 1) * sizeof(T));
    return e;
}
template <typename T>
hipError_t trackUpload(ebb_solver* s, DBuf<T>& b, const std::vector<T>& v)
{
    hipError_t e = b.upload(v);
    s->deviceBytes += int64_t((v.size() ? v.size() : 1) * sizeof(T));
    return e;
}

// Rotation of the boundary force into drag and lift (app/solver.d:178-193).
void aoaRotation(const ebb_solver* s, double rot[4])
{
    const long double a = (long double)s->cfg.ic[1] * (3.14159265358979323846264338327950288L / 180.0L);
    rot[0] = double(cosl(a)); rot[1] = -double(sinl(a)); rot[2] = double(sinl(a)); rot[3] = double(cosl(a));
}

// Paths whose stages start with the gradient kernel leave each step's force sum to the
// next gradient kernel (the fused path sums it in the RMS kernel).
bool deferForce(const ebb_solver* s) { return !s->useTiles && s->d_forceFaces.n > 0; }

ForceArgs forceArgs(const ebb_solver* s)
{
    ForceArgs f{};
    f.faces = s->d_forceFaces.p; f.n = uint32_t(s->d_forceFaces.n);
    f.lenF = s->d_lenF.p;
    aoaRotation(s, f.rot);
    f.hist = s->d_hist.p;
    return f;
}

// The FP32 shadow paired with a state buffer (mixed precision), null otherwise.
float4* shadowOf(ebb_solver* s, const double4* q)
{
    if (!s->phys.mixed || !q) return nullptr;
    if (q == s->d_qA.p) return s->d_qA32.p;
    if (q == s->d_qB.p) return s->d_qB32.p;
    if (q == s->d_T.p) return s->d_T32.p;
    if (q == s->d_T2.p) return s->d_T232.p;
    return nullptr;
}

// the mixed-precision pointers of a stage reading qin and writing qout
void mixArgs(ebb_solver* s, StageArgs& a, const double4* qin, const double4* qout)
{
    a.qin32 = shadowOf(s, qin);
    a.qout32 = shadowOf(s, qout);
    a.grad32 = s->phys.mixed ? s->d_grad32.p : nullptr;
    a.lim32 = s->phys.mixed ? s->d_lim32.p : nullptr;
}

// refresh the FP32 shadow of the current state (after an upload, a load or set-up)
void refreshShadow(ebb_solver* s)
{
    if (s->phys.mixed) launchShadow(s->qcur(), shadowOf(s, s->qcur()), uint32_t(s->b.nCells), s->stream);
}

// the arguments every stage shares
static StageArgs baseStageArgs(ebb_solver* s, int kind, double4* qin, const double4* qn, double4* qout, int limit,
                               int dtUpdate, int first, int last, uint32_t stageId, uint32_t stepId)
{
    StageArgs a{};
    a.qin = qin; a.qn = qn; a.qout = qout; a.acc = s->d_acc.p;
    a.grad = s->d_grad.p; a.lim = s->d_lim.p; a.lenF = s->d_lenF.p; a.lenS = s->d_lenS.p;
    a.celldt = s->d_celldt.p; a.partial = s->d_partial.p; a.sc = s->d_sc;
    a.kind = kind; a.limit = limit; a.dtUpdate = dtUpdate; a.first = first; a.last = last;
    a.recordErrors = 1; a.stageId = stageId; a.stepId = stepId;
    a.order = s->cfg.kernelOrder;
    mixArgs(s, a, qin, qout);
    return a;
}

StageArgs loopStageArgs(ebb_solver* s, int kind, double4* qin, const double4* qn, double4* qout, int limit,
                        int dtUpdate, int first, int last, uint32_t stageInStep, uint32_t stageId, uint32_t stepId)
{
    StageArgs a = baseStageArgs(s, kind, qin, qn, qout, limit, dtUpdate, first, last, stageId, stepId);
    a.checkActive = 1;
    a.rmsBlocks = (last && s->rmsFused) ? s->d_rmsBlocks.p : nullptr;
    // time loop: each stage leaves its partials in its own region; the step's RMS
    // kernel reduces them (nothing in the step reads newDt or Rmax before that)
    a.deferReduce = 1;
    a.partial = s->d_partial.p + size_t(stageInStep) * s->pstride;
    // the gradient kernels of the time loop also sum the previous step's force
    if (deferForce(s)) a.force = forceArgs(s);
    return a;
}

// one residual evaluation + stage update
void stage(ebb_solver* s, int kind, double4* qin, const double4* qn, double4* qout, int limit, int dtUpdate,
           int first, int last, int checkActive, double4* R, double4* estate, uint32_t stepId)
{
    const uint32_t stageId = s->stageCounter++;
    StageArgs a;
    if (checkActive) {
        a = loopStageArgs(s, kind, qin, qn, qout, limit, dtUpdate, first, last, s->stageInStep, stageId, stepId);
        s->stageInStep++;
    } else {
        a = baseStageArgs(s, kind, qin, qn, qout, limit, dtUpdate, first, last, stageId, stepId);
    }
    a.R = R;
    a.estate = estate;
    s->stageStep.push_back(stepId);
    if (s->useTiles && !estate && kind != SK_RESIDUAL) {
        launchStageTiles(s->dm, s->phys, a, s->tiles, int(s->S), s->stream);
    } else if (s->useGradTiles && !estate && kind != SK_RESIDUAL) {
        launchGradTiles(s->dm, s->phys, a, s->ftiles, int(s->S), s->stream);
    } else if (s->useFaceTiles && !estate && kind != SK_RESIDUAL) {
        launchGradient(s->dm, s->phys, a, int(s->S), s->stream);
        launchFaceTiles(s->dm, s->phys, a, s->ftiles, int(s->S), s->stream);
    } else {
        launchGradient(s->dm, s->phys, a, int(s->S), s->stream);
        launchFaces(s->dm, s->phys, a, s->stream);
        launchUpdate(s->dm, s->phys, a, int(s->S), s->stream);
    }
    s->lastInput = qin;
    s->lastLimit = limit;
    s->haveStage = true;
}

// Integrator.step (source/integrators.d:22-270) from qcur() into qnext(): each stage's
// kind, input and output buffers, limiter and local-dt flags
std::vector<StageDesc> stepSequence(ebb_solver* s)
{
    const Config& c = s->cfg;
    const int msl = c.multistageLimiting ? 1 : 0, lts = c.localTimestep ? 1 : 0;
    double4 *q = s->qcur(), *nx = s->qnext(), *T = s->d_T.p, *T2 = s->d_T2.p;
    switch (c.integratorKind()) {
    case IntEuler: return {{SK_EULER, q, nx, 1, 1, 1, 1}};
    case IntRK2: return {{SK_RK2_1, q, T, 1, 1, 1, 0}, {SK_RK2_2, T, nx, msl, 0, 0, 1}};
    case IntRK2TVD:   // stage 2 unlimited (integrators.d:190)
        return {{SK_TVD_1, q, T, 1, lts, 1, 0}, {SK_TVD_2, T, nx, 0, lts ? 0 : 1, 0, 1}};
    default:
        return {{SK_RK4_1, q, T, 1, 1, 1, 0}, {SK_RK4_2, T, T2, msl, 0, 0, 0}, {SK_RK4_3, T2, T, msl, 0, 0, 0},
                {SK_RK4_4, T, nx, msl, 0, 0, 1}};
    }
}

// Integrator.step: all stages of one step from qcur() into qnext()
void stepStages(ebb_solver* s, int checkActive, uint32_t stepId)
{
    s->stageInStep = 0;
    const double4* q = s->qcur();
    for (const StageDesc& x : stepSequence(s))
        stage(s, x.kind, x.in, q, x.out, x.limit, x.dtUpdate, x.first, x.last, checkActive, nullptr, nullptr, stepId);
}

int stagesPerStep(const Config& c)
{
    switch (c.integratorKind()) {
    case IntEuler: return 1;
    case IntRK4: return 4;
    default: return 2;
    }
}

DiagArgs diagArgs(ebb_solver* s, const double4* qnew, const double4* qold, int64_t histBase, int64_t maxSteps)
{
    DiagArgs d{};
    d.qnew = qnew; d.qold = qold; d.N = s->N; d.L = s->N + s->G; d.perm = s->refOf.empty() ? nullptr : s->d_devOf.p;
    d.blockSums = s->rmsFused ? s->d_rmsBlocks.p : nullptr;
    d.chunkTrees = s->d_chunkTrees.p; d.chunkMask = s->d_chunkMask.p;
    d.forceFaces = s->d_forceFaces.p; d.nForce = uint32_t(s->d_forceFaces.n);
    d.lenF = s->d_lenF.p;
    aoaRotation(s, d.rot);
    d.tEnd = s->cfg.tEnd;
    d.dynamicDt = s->cfg.dynamicDt; d.cflAdjust = s->cfg.cflAdjust;
    d.hist = s->d_hist.p; d.histBase = histBase; d.histCap = s->histCap;
    d.maxSteps = maxSteps;
    d.partial = s->d_partial.p; d.nPart = s->nPart; d.pstride = s->pstride; d.nStages = s->stageInStep;
    d.deferForce = deferForce(s) ? 1 : 0;
    d.histRing = 0;
    d.rmsDrop = s->cfg.rmsDrop;
    d.sc = s->d_sc;
    return d;
}

int readScalars(ebb_solver* s)
{
    HIPCHK(hipMemcpyAsync(s->h_sc, s->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EBB_OK;
}

// a new window of stages whose errors the next checkErrors reports
static void beginErrorWindow(ebb_solver* s)
{
    s->stageBase = s->stageCounter;
    s->stageStep.clear();
}

static int64_t stepOfStage(const ebb_solver* s, uint32_t stage)
{
    const uint32_t k = stage - s->stageBase;
    return k < s->stageStep.size() ? s->stageStep[k] : -1;
}

int checkErrors(ebb_solver* s)
{
    const DevScalars& h = *s->h_sc;
    // the reference updates the ghosts before its edge loop and checks the cells after the
    // step: a boundary failure comes first unless a NaN edge came in an earlier stage
    if (h.errBC != ~0ull && (h.errEdge == ~0ull || (h.errBC >> 32) <= (h.errEdge >> 32))) {
        const uint64_t key = h.errBC & 0xFFFFFFFFull;
        s->errKind = EBB_ERR_BC;
        s->errIndex = int64_t(s->b.boundaryEdges[key]);
        s->errStep = stepOfStage(s, uint32_t(h.errBC >> 32));
        s->errMsg = "Both computed mach numbers are negative";
        return EBB_ERR_BC;
    }
    if (h.errEdge != ~0ull) {
        const uint64_t key = h.errEdge & 0xFFFFFFFFull;
        int64_t edge = key < s->Fb ? int64_t(s->b.boundaryEdges[key]) : int64_t(s->b.interiorEdges[key - s->Fb]);
        s->errKind = EBB_ERR_EDGE_NAN;
        s->errIndex = edge;
        s->errStep = stepOfStage(s, uint32_t(h.errEdge >> 32));
        s->errMsg = key < s->Fb ? "Got NaN on boundary edge" : "Got NaN on interior edge";
        return EBB_ERR_EDGE_NAN;
    }
    if (h.errCell != ~0ull) {
        s->errKind = EBB_ERR_CELL_NAN;
        s->errIndex = int64_t(h.errCell & 0xFFFFFFFFull);
        s->errStep = int64_t(h.errCell >> 32);
        s->errMsg = "Got nan on cell average value";
        return EBB_ERR_CELL_NAN;
    }
    return EBB_OK;
}

int clearErrors(ebb_solver* s)
{
    // keep t/dt/CFL/counters, clear the error words and reactivate
    HIPCHK(hipMemcpyAsync(s->h_sc, s->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    beginErrorWindow(s);
    s->h_sc->errEdge = ~0ull;
    s->h_sc->errCell = ~0ull;
    s->h_sc->errBC = ~0ull;
    HIPCHK(hipMemcpyAsync(s->d_sc, s->h_sc, sizeof(DevScalars), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EBB_OK;
}

// Down to 32 cells, in whole 16-cell blocks: a tile's workgroup takes about the same time
// for any tile size up to one round of lanes, so smaller tiles only buy more workgroups.
uint32_t spreadTileCells(uint64_t cells, int cus)
{
    const uint64_t share = cells / (2 * uint64_t(std::max(cus, 1))) / 16 * 16;
    return uint32_t(std::max<uint64_t>(32, std::min<uint64_t>(768, share)));
}

// Tiled paths: choose the tile size (tileCells, or automatically: 768 cells per
// face tile; for the fused kernel the largest multiple of 16 up to 128 whose two
// blob buffers and working set fit 64 KB of LDS) and cut the tiles on the
// reference numbering; with renumbering the interior cells are then numbered tile
// by tile.  `lay` receives the device-order layout, tile t is the device cells
// tileCellsOrder[start[t] .. start[t+1]).
int planTiles(ebb_solver* s, Layout& ref, Layout& lay, std::vector<uint32_t>& start,
              std::vector<uint32_t>& tileCellsOrder)
{
    const Config& c = s->cfg;
    start.clear();
    s->refOf.clear();
    s->devOf.clear();
    if (c.stagePath == 0 || ref.N == 0) { lay = std::move(ref); return EBB_OK; }
    const long long want = c.tileCells;
    if (want != -1 && (want < 1 || want > 1024)) return fail(s, EBB_ERR_ARG, "tileCells must be -1 or in 1..1024");
    std::vector<uint32_t> order;
    uint32_t T = want > 0 ? uint32_t(want) : (c.stagePath == 1 || c.stagePath == 3 ? 768 : 128);
    if (want <= 0 && c.spreadTiles && (c.stagePath == 1 || c.stagePath == 3)) {
        // a mesh too small to give every CU two tiles of 768 cells gets smaller tiles, so it
        // still spreads over the GPU.  A batched member counts the cells running beside it:
        // the launches are shared
        T = std::min(T, spreadTileCells(uint64_t(ref.N) + uint64_t(std::max<int64_t>(c.spreadOver, 0)), s->cus));
    }
    if (c.stagePath == 3 && want <= 0) {
        // automatic gradient tiles: the largest multiple of 16 up to 768 cells whose busiest
        // tile keeps its gradient rows (12 values per own or ring cell) and face results
        // within a third of the CU's 160 KB of LDS, so three workgroups stay resident
        constexpr uint32_t LDS_TARGET = 53248;
        const uint32_t rowBytes = 12 * (c.precision ? 4 : 8);
        for (;;) {
            tileOrder(ref, T, c.tileOrder, order, start);
            const uint32_t need = gradTileLds(ref, order, start, rowBytes, faceLdsBytes(c.precision));
            if (need <= LDS_TARGET || T <= 16) break;
            const uint32_t next = uint32_t(uint64_t(T) * LDS_TARGET / need) / 16 * 16;
            T = std::max<uint32_t>(16, std::min(next, T - 16));
        }
    } else if (c.stagePath == 1 && want <= 0) {
        // automatic face tiles: the largest multiple of 16 up to 768 cells whose busiest
        // tile keeps its face results (40 B per face) within a third of the CU's 160 KB
        // of LDS, so three workgroups stay resident (768 fits triangles; quads and larger
        // polygons have more faces per cell); mixed precision keeps 28 B per face and fits
        // four workgroups, its registers' limit
        const uint32_t LDS_TARGET = c.precision ? 40960 : 53248;
        for (;;) {
            tileOrder(ref, T, c.tileOrder, order, start);
            const uint32_t need = faceLdsBytes(c.precision) * (maxTileFaces(ref, order, start) + 1);
            if (need <= LDS_TARGET || T <= 16) break;
            const uint32_t next = uint32_t(uint64_t(T) * LDS_TARGET / need) / 16 * 16;
            T = std::max<uint32_t>(16, std::min(next, T - 16));
        }
    } else if (c.stagePath == 2 && want <= 0) {
        const int lp = (c.order > 1 && c.limited && c.lpThresh > 0) ? 1 : 0;
        for (;; T -= 16) {
            tileOrder(ref, T, c.tileOrder, order, start);
            TileHost h;
            std::string err;
            if (!makeTiles(ref, order, start, h, err)) return fail(s, EBB_ERR_MESH, err);
            if (2 * h.maxBlob + h.maxDyn[lp] <= 65536 || T <= 32) break;
        }
    } else {
        tileOrder(ref, T, c.tileOrder, order, start);
    }
    s->tileCells = T;
    if (!c.renumber) {   // tiles list their cells by reference number
        lay = std::move(ref);
        tileCellsOrder = std::move(order);
        return EBB_OK;
    }
    permuteLayout(ref, order, lay);
    s->refOf = order;
    s->devOf.assign(ref.N, 0);
    for (uint32_t d = 0; d < ref.N; d++) s->devOf[order[d]] = d;
    tileCellsOrder.resize(ref.N);
    for (uint32_t d = 0; d < ref.N; d++) tileCellsOrder[d] = d;
    return EBB_OK;
}

int buildFaceTiles(ebb_solver* s, const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start)
{
    FaceTileHost h;
    std::string err;
    if (!makeFaceTiles(L, order, start, h, err)) return fail(s, EBB_ERR_MESH, err);
    const bool grad = s->cfg.stagePath == 3;
    // face results of the largest tile plus the padding face, 16-byte aligned
    uint32_t lds = (faceLdsBytes(s->cfg.precision) * (h.faceStride + 1) + 15) / 16 * 16, glDoubles = 0;
    if (grad) {
        if (!addGradTiles(L, h, err)) return fail(s, EBB_ERR_MESH, err);
        const uint32_t rowBytes = 12 * (s->cfg.precision ? 4 : 8);
        glDoubles = (h.maxG * rowBytes + 15) / 16 * 2;
        lds += 8 * glDoubles;
    }
    if (lds > uint32_t(s->ldsPerBlock)) return EBB_OK;   // does not fit: split kernels
    HIPCHK(trackUpload(s, s->d_ftDesc, h.desc));
    HIPCHK(trackUpload(s, s->d_ftRef, h.fref));
    HIPCHK(trackUpload(s, s->d_ftIds, h.ids));
    HIPCHK(trackUpload(s, s->d_ftCf, h.cf));
    HIPCHK(trackUpload(s, s->d_ftFid, h.fid));
    FaceTileArgs ft{};
    ft.nTiles = h.nTiles;
    ft.ldsBytes = lds;
    ft.cellStride = h.cellStride;
    ft.faceStride = h.faceStride;
    ft.desc = s->d_ftDesc.p; ft.ids = s->d_ftIds.p; ft.cf = s->d_ftCf.p; ft.fref = s->d_ftRef.p; ft.fid = s->d_ftFid.p;
    if (grad) {
        HIPCHK(trackUpload(s, s->d_ftSlot, h.fslot));
        HIPCHK(trackUpload(s, s->d_ftHalo, h.halo));
        HIPCHK(trackUpload(s, s->d_ftNbr, h.nbrT));
        HIPCHK(trackUpload(s, s->d_ftNHalo, h.nHalo));
        if (h.border.empty()) h.border.push_back(0);   // a valid pointer; nBorder stays 0
        const uint32_t nB = h.sumH ? uint32_t(h.border.size()) : 0u;
        HIPCHK(trackUpload(s, s->d_border, h.border));
        ft.fslot = s->d_ftSlot.p; ft.halo = s->d_ftHalo.p; ft.nHalo = s->d_ftNHalo.p; ft.border = s->d_border.p;
        ft.nbrT = s->d_ftNbr.p;
        ft.haloStride = h.haloStride;
        ft.nBorder = nB;
        ft.glDoubles = glDoubles;
        s->gradTileStats[0] = h.maxG;
        s->gradTileStats[1] = double(h.sumH);
        s->gradTileStats[2] = nB;
    }
    s->ftiles = ft;
    s->useFaceTiles = true;
    s->useGradTiles = grad;
    return EBB_OK;
}

// Build and upload the tiles of the chosen path on the device-order layout.
int buildTiles(ebb_solver* s, const Layout& L, const std::vector<uint32_t>& order, const std::vector<uint32_t>& start)
{
    s->useTiles = false;
    s->tiles = TileArgs{};
    s->useFaceTiles = false;
    s->useGradTiles = false;
    s->ftiles = FaceTileArgs{};
    if (L.N == 0 || s->cfg.stagePath == 0) return EBB_OK;
    if (s->cfg.stagePath == 1 || s->cfg.stagePath == 3) return buildFaceTiles(s, L, order, start);
    const int lp = (s->phys.order > 1 && s->phys.limited && s->phys.lpThresh > 0) ? 1 : 0;
    TileHost h;
    std::string err;
    if (!makeTiles(L, order, start, h, err)) return fail(s, EBB_ERR_MESH, err);
    if (2 * h.maxBlob + h.maxDyn[lp] > 65536) return EBB_OK;   // does not fit: split kernels
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    HIPCHK(trackUpload(s, s->d_tDesc, h.desc));
    HIPCHK(trackUpload(s, s->d_tBlob, h.blob));
    TileArgs ta{};
    ta.nTiles = h.nTiles;
    ta.nBlocks = std::min<uint32_t>(h.nTiles, 2 * uint32_t(std::max(cus, 1)));   // two resident per CU
    ta.slotBytes = h.maxBlob;
    ta.ldsBytes = 2 * h.maxBlob + h.maxDyn[lp];
    ta.desc = s->d_tDesc.p;
    ta.blob = s->d_tBlob.p;
    s->tiles = ta;
    s->useTiles = true;
    return EBB_OK;
}

// Cut the face tiles again with `tileCells` cells per tile (-1: automatic) and upload them
// in place of the current ones; the stage partials grow with the tile count.  Only the
// face-tile path in the reference numbering (the batched runs' layout) is re-tiled.
int retileFaceTiles(ebb_solver* s, long long tileCells)
{
    if (!s->useFaceTiles || !s->refOf.empty()) return fail(s, EBB_ERR_STATE, "only face tiles in the reference numbering can be re-tiled");
    auto drop = [&](auto& b) { s->deviceBytes -= int64_t((b.n ? b.n : 1) * sizeof(*b.p)); b.release(); };
    drop(s->d_ftDesc); drop(s->d_ftRef); drop(s->d_ftIds); drop(s->d_ftCf); drop(s->d_ftFid);
    drop(s->d_ftSlot); drop(s->d_ftHalo); drop(s->d_ftNHalo); drop(s->d_border); drop(s->d_ftNbr);
    s->cfg.tileCells = tileCells;
    Layout ref, lay;
    makeLayout(s->b, ref);
    std::vector<uint32_t> start, ord;
    int rc = planTiles(s, ref, lay, start, ord);
    if (rc) return rc;
    rc = buildTiles(s, lay, ord, start);
    if (rc) return rc;
    if (!s->useFaceTiles) return fail(s, EBB_ERR_MESH, "face tiles do not fit the workgroup's LDS");
    s->nPart = s->ftiles.nTiles;
    const uint32_t need = 2 * uint32_t(std::max<size_t>({(s->N + 255) / 256, s->ftiles.nTiles})) + 2;
    if (need > s->pstride) {
        s->deviceBytes -= int64_t(s->d_partial.n * sizeof(double));
        s->d_partial.release();
        s->pstride = need;
        HIPCHK((hipError_t)trackAlloc(s, s->d_partial, size_t(s->pstride) * 4));
    }
    return EBB_OK;
}

// Device bytes buildDevice allocates for this mesh and config (the tile arrays estimated
// at their usual fill), without allocating: the admission test of batched runs.
int64_t estimateDeviceBytes(const BuiltMesh& b, const Config& c)
{
    const int64_t N = b.nElem, C = b.nCells, G = C - N, E = b.nEdges, F = E;
    int maxE = 3;
    for (uint32_t i = 0; i < b.nElem; i++) maxE = std::max<int>(maxE, b.nEdgesCell[i]);
    const int64_t S = maxE <= 3 ? 3 : (maxE == 4 ? 4 : 6);
    const int nst = stagesPerStep(c);
    const int k = c.integratorKind();
    const bool lp = c.order > 1 && c.limited && c.lpThresh > 0;
    int64_t bytes = 0;
    bytes += N * S * 4 * 2 + N * 4;                        // nbr, cface, refId
    bytes += int64_t(b.nCells) * 0 + int64_t(b.nEdges) * 0;
    bytes += 3 * N * 8 + C * 16;                           // nia, cd, cper; centroids
    bytes += F * (8 + 16 + 16 + 8);                        // fcell, fn, fmid, flen
    bytes += G * int64_t(sizeof(BFace));
    if (c.mms) bytes += N * 32;
    int64_t nodes = 0;
    for (uint32_t e = 0; e < b.nEdges; e++) nodes = std::max<int64_t>(nodes, std::max(b.eNode0[e], b.eNode1[e]) + 1);
    bytes += nodes * 16;
    bytes += C * 32 * (2 + (nst > 1 ? 1 : 0) + (k == IntRK4 ? 1 : 0));   // qA, qB, T, T2
    if (k == IntRK2 || k == IntRK4) bytes += N * 32;                       // acc
    if (c.precision) bytes += C * 16 * (2 + (nst > 1 ? 1 : 0) + (k == IntRK4 ? 1 : 0)) + N * 48;
    else bytes += N * 64 + N * (lp ? 64 : 32);
    bytes += (F + 1) * 40;                                 // lenF, lenS
    if (c.localTimestep) bytes += N * 8;
    // face tiles: ~1.1 cells' and ~1.7 faces' worth of records per cell, plus the padding
    bytes += int64_t(1.1 * double(N)) * (4 + 2 * S) + int64_t(1.7 * double(N)) * (16 + 4) + 2 * 256 * 20;
    // gradient tiles: an LDS-row pair per face record, each tile's ring and the border (~0.3 N each)
    if (c.stagePath == 3) bytes += int64_t(1.7 * double(N)) * 4 + int64_t(0.6 * double(N)) * 4 + int64_t(1.1 * double(N)) * 4 * S + 4096;
    bytes += ((N + 15) / 16) * 64;                         // RMS leaf sums
    // tiles: ~700 cells each, or at least two per CU on small meshes (down to 32 cells)
    const int64_t tiles = std::max<int64_t>(N / (c.stagePath == 3 ? 250 : 700) + 1,
                                            c.spreadTiles ? std::min<int64_t>(N / 32 + 1, 512) : 0);
    bytes += (2 * std::max<int64_t>((N + 255) / 256, tiles) + 2) * 4 * 8;   // stage partials
    bytes += 256 * HIST_FIELDS * 8;
    const int64_t nchunks = (C / 16 + RMS_CHUNK_BLOCKS - 1) / RMS_CHUNK_BLOCKS + 1;
    bytes += nchunks * (128 * 8 + 32);
    return bytes;
}

// Interior-cell arrays between the reference numbering and the device numbering.
template <typename T>
void toDevice(const ebb_solver* s, const T* ref, T* dev, size_t per = 1)
{
    if (s->refOf.empty()) { std::copy(ref, ref + size_t(s->N) * per, dev); return; }
    for (uint32_t d = 0; d < s->N; d++) std::copy(ref + size_t(s->refOf[d]) * per, ref + size_t(s->refOf[d] + 1) * per, dev + size_t(d) * per);
}
template <typename T>
void toReference(const ebb_solver* s, const T* dev, T* ref, size_t per = 1)
{
    if (s->refOf.empty()) { std::copy(dev, dev + size_t(s->N) * per, ref); return; }
    for (uint32_t d = 0; d < s->N; d++) std::copy(dev + size_t(d) * per, dev + size_t(d + 1) * per, ref + size_t(s->refOf[d]) * per);
}

int buildDevice(ebb_solver* s)
{
    const BuiltMesh& b = s->b;
    const Config& c = s->cfg;
    Layout lay;
    std::vector<uint32_t> tileStart, tileOrd;
    {
        Layout ref;
        makeLayout(b, ref);
        int rc_ = planTiles(s, ref, lay, tileStart, tileOrd);
        if (rc_) return rc_;
    }
    s->N = lay.N; s->G = lay.G; s->Fi = lay.Fi; s->Fb = lay.Fb; s->S = lay.S;
    const uint32_t N = s->N, F = s->Fi + s->Fb;
    s->faceOfEdge = lay.faceOfEdge;
    s->edgeOfFace = lay.edgeOfFace;
    const std::vector<uint32_t>&nbr = lay.nbr, &cface = lay.cface, &refId = lay.refId;
    const std::vector<double>&nia = lay.nia, &cd = lay.cd, &cper = lay.cper, &flen = lay.flen;
    const std::vector<double2>&cen = lay.cen, &fn = lay.fn, &fmid = lay.fmid;
    const std::vector<uint2>& fcell = lay.fcell;
    // boundary binding (solve.d:222-278)
    s->bfaces.assign(s->Fb, BFace{});
    std::vector<int> bcOfTag(s->raw.btags.size(), -1);
    for (size_t t = 0; t < s->raw.btags.size(); t++) {
        for (size_t k = 0; k < c.bcs.size(); k++)
            if (c.bcs[k].tag == s->raw.btags[t]) { bcOfTag[t] = int(k); break; }
        if (bcOfTag[t] < 0 && !b.bGroups[t].empty()) {
            return fail(s, EBB_ERR_BC, "Could not find matching boundary condition tag " + s->raw.btags[t]);
        }
    }
    std::vector<uint32_t> posOfBoundary(b.nEdges, 0);
    for (uint32_t k = 0; k < s->Fb; k++) posOfBoundary[b.boundaryEdges[k]] = k;
    const double* nd = s->raw.nodes.data();
    for (uint32_t k = 0; k < s->Fb; k++) {
        const uint32_t e = b.boundaryEdges[k];
        BFace& bf = s->bfaces[k];
        const int32_t g = b.eBGroup[e];
        const BCConfig& bc = c.bcs[bcOfTag[g]];
        bf.type = bc.type;
        bf.nrm[0] = b.nx[e];
        bf.nrm[1] = b.ny[e];
        bf.Rgas = c.R;
        const double x1 = nd[2 * b.eNode0[e]], y1 = nd[2 * b.eNode0[e] + 1];
        const double x2 = nd[2 * b.eNode1[e]], y2 = nd[2 * b.eNode1[e] + 1];
        bf.vertical = (x1 != x2) ? 0 : 1;
        if (x1 != x2) {
            const double m = (y2 - y1) / (x2 - x1);
            const double den = 1 + m * m;
            bf.R[0] = (1 - m * m) * 1.0 / den;
            bf.R[1] = (2.0 * m) * 1.0 / den;
            bf.R[2] = (2.0 * m) * 1.0 / den;
            bf.R[3] = (m * m - 1) * 1.0 / den;
        }
        if (bc.type == FullState) {
            if (bc.data.size() < 4) return fail(s, EBB_ERR_BC, "fullState boundary data needs 4 values");
            freestreamState(bc.data[0], bc.data[1], bc.data[3], bf.far);
        } else if (bc.type == ConstPressure) {
            if (bc.data.size() != 1) return fail(s, EBB_ERR_BC, "Constant pressure boundary data should only have one element; the constant pressure");
            bf.data[0] = bc.data[0];
        } else if (bc.type == TempPresInflow) {
            if (bc.data.size() != 3) return fail(s, EBB_ERR_BC, "Constant temperature/pressure/aoa boundary data should have three elements");
            const double aoa = aoaRadians(bc.data[2]);
            bf.data[0] = bc.data[0];
            bf.data[1] = bc.data[1];
            bf.data[2] = double(cosl((long double)aoa));
            bf.data[3] = double(sinl((long double)aoa));
        } else if (bc.type == Dirichlet) {
            if (!c.mms) return fail(s, EBB_ERR_BC, "dirichlet boundary needs the manufactured solution (extensions.mms)");
            mmsSolution(b.mx[e], b.my[e], bf.far);
            const uint32_t gc = b.eCell1[e];
            mmsSolution(b.cx[gc], b.cy[gc], bf.gq);
            mmsSolutionGradient(b.cx[gc], b.cy[gc], bf.gg);
        }
    }
    // physics parameters
    PhysParams& p = s->phys;
    p.flux = c.fluxKind();
    p.viscous = c.viscosity ? 1 : 0;
    p.order = int(c.order);
    p.limited = c.limited ? 1 : 0;
    p.Pr = c.Pr;
    // the reference sets mu only for viscous runs (solve.d:186-190); otherwise it keeps D's NaN default
    p.mu = c.viscosity ? (c.ic[3] * 1.0 * c.L) / c.Re : std::nan("");
    s->cfg.mu = p.mu;
    p.Rgas = c.R;
    p.lpThresh = c.lpThresh;
    p.localTimestep = c.localTimestep ? 1 : 0;
    p.mixed = c.precision;
    p.venkat = (c.limiter == 1) ? 1 : 0;
    p.venkatK = c.venkatK;
    // MMS source at centroids
    std::vector<double4> src;
    if (c.mms) {
        src.resize(N);
        for (uint32_t i = 0; i < N; i++) {
            double S4[4];
            const uint32_t r = refId[i];
            mmsSourceTerm(b.cx[r], b.cy[r], c.viscosity, c.R, p.mu, c.Pr, S4);
            src[i] = make_double4(S4[0], S4[1], S4[2], S4[3]);
        }
    }
    // force faces in bGroups order of the forceBoundary tags (mesh.d:645-671)
    std::vector<uint32_t> ff;
    for (auto& tag : c.forceBoundary) {
        auto it = std::find(s->raw.btags.begin(), s->raw.btags.end(), tag);
        if (it == s->raw.btags.end()) continue;
        for (uint32_t e : b.bGroups[size_t(it - s->raw.btags.begin())]) ff.push_back(s->faceOfEdge[e]);
    }
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIPCHK(trackUpload(s, s->d_nbr, nbr));
    HIPCHK(trackUpload(s, s->d_cface, cface));
    HIPCHK(trackUpload(s, s->d_refId, refId));
    {
        std::vector<double2> nodes(s->raw.nodes.size() / 2);
        for (size_t k = 0; k < nodes.size(); k++) nodes[k] = make_double2(s->raw.nodes[2 * k], s->raw.nodes[2 * k + 1]);
        HIPCHK(trackUpload(s, s->d_node, nodes));
    }
    HIPCHK(trackUpload(s, s->d_nia, nia));
    HIPCHK(trackUpload(s, s->d_cd, cd));
    HIPCHK(trackUpload(s, s->d_cper, cper));
    HIPCHK(trackUpload(s, s->d_cen, cen));
    HIPCHK(trackUpload(s, s->d_fcell, fcell));
    HIPCHK(trackUpload(s, s->d_fn, fn));
    HIPCHK(trackUpload(s, s->d_fmid, fmid));
    HIPCHK(trackUpload(s, s->d_flen, flen));
    HIPCHK(trackUpload(s, s->d_bface, s->bfaces));
    HIPCHK(trackUpload(s, s->d_forceFaces, ff));
    s->d_forceFaces.n = ff.size();
    if (c.mms) HIPCHK(trackUpload(s, s->d_src, src));
    const size_t C = b.nCells;
    HIPCHK((hipError_t)trackAlloc(s, s->d_qA, C));
    HIPCHK((hipError_t)trackAlloc(s, s->d_qB, C));
    const int nst = stagesPerStep(c);
    if (nst > 1) HIPCHK((hipError_t)trackAlloc(s, s->d_T, C));
    if (c.integratorKind() == IntRK4) HIPCHK((hipError_t)trackAlloc(s, s->d_T2, C));
    if (c.integratorKind() == IntRK2 || c.integratorKind() == IntRK4) HIPCHK((hipError_t)trackAlloc(s, s->d_acc, N));
    const bool lp = c.order > 1 && c.limited && c.lpThresh > 0;
    if (p.mixed) {
        HIPCHK((hipError_t)trackAlloc(s, s->d_qA32, C));
        HIPCHK((hipError_t)trackAlloc(s, s->d_qB32, C));
        if (nst > 1) HIPCHK((hipError_t)trackAlloc(s, s->d_T32, C));
        if (c.integratorKind() == IntRK4) HIPCHK((hipError_t)trackAlloc(s, s->d_T232, C));
        HIPCHK((hipError_t)trackAlloc(s, s->d_grad32, size_t(N) * 8));
        HIPCHK((hipError_t)trackAlloc(s, s->d_lim32, size_t(N) * 4));
    } else {
        HIPCHK((hipError_t)trackAlloc(s, s->d_grad, size_t(N) * 8));
        HIPCHK((hipError_t)trackAlloc(s, s->d_lim, size_t(N) * (lp ? 8 : 4)));
    }
    HIPCHK((hipError_t)trackAlloc(s, s->d_lenF, size_t(F + 1) * 4));
    HIPCHK((hipError_t)trackAlloc(s, s->d_lenS, size_t(F + 1)));
    if (c.localTimestep) HIPCHK((hipError_t)trackAlloc(s, s->d_celldt, N));
    // physics parameters are set above; the tile builder needs them for the LDS size
    { int rc_ = buildTiles(s, lay, tileOrd, tileStart); if (rc_) return rc_; }
    if (!s->refOf.empty()) HIPCHK(trackUpload(s, s->d_devOf, s->devOf));
    // leaf blocks need 16 aligned consecutive reference cells per 16 lanes of the last stage
    s->rmsFused = !s->useTiles && ((s->useFaceTiles || !s->refOf.empty()) ? c.tileOrder == TILE_BLOCKS : true);
    if (s->rmsFused) HIPCHK((hipError_t)trackAlloc(s, s->d_rmsBlocks, size_t((N + 15) / 16) * 8));
    // the stage kernel's workgroup count: one (min dt, max |R|) pair each
    s->nPart = s->useTiles ? s->tiles.nBlocks : (s->useFaceTiles ? s->ftiles.nTiles : uint32_t((N + 255) / 256));
    s->pstride = 2 * uint32_t(std::max<size_t>({(N + 255) / 256, s->tiles.nBlocks, s->ftiles.nTiles})) + 2;
    HIPCHK((hipError_t)trackAlloc(s, s->d_partial, size_t(s->pstride) * 4));   // up to 4 stages per step
    HIPCHK((hipError_t)trackAlloc(s, s->d_hist, size_t(s->histCap) * HIST_FIELDS));
    const uint64_t nb = uint64_t(C) / 16;
    const uint64_t nchunks = (nb + RMS_CHUNK_BLOCKS - 1) / RMS_CHUNK_BLOCKS + 1;
    HIPCHK((hipError_t)trackAlloc(s, s->d_chunkTrees, size_t(nchunks) * 128));
    HIPCHK((hipError_t)trackAlloc(s, s->d_chunkMask, size_t(nchunks) * 8));
    HIPCHK(hipMalloc((void**)&s->d_sc, sizeof(DevScalars)));
    HIPCHK(hipHostMalloc((void**)&s->h_sc, sizeof(DevScalars), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&s->h_hist, sizeof(double) * size_t(s->histCap) * HIST_FIELDS, hipHostMallocDefault));

    DevMesh& dm = s->dm;
    dm.N = N; dm.G = s->G; dm.Fi = s->Fi; dm.Fb = s->Fb;
    dm.nbr = s->d_nbr.p; dm.cface = s->d_cface.p; dm.cen = s->d_cen.p;
    dm.nia = s->d_nia.p; dm.cd = s->d_cd.p; dm.cper = s->d_cper.p;
    dm.fcell = s->d_fcell.p; dm.fn = s->d_fn.p; dm.fmid = s->d_fmid.p; dm.flen = s->d_flen.p;
    dm.bface = s->d_bface.p; dm.src = c.mms ? s->d_src.p : nullptr; dm.refId = s->d_refId.p;
    dm.node = s->d_node.p;

    // initial state (solve.d:194-207; app/ms.d:295-298)
    std::vector<double4> q0(C, make_double4(0, 0, 0, 0));
    double fs[4];
    freestreamState(c.ic[0], c.ic[1], c.ic[3], fs);
    for (uint32_t i = 0; i < N; i++) {
        if (c.mms) {
            double qq[4];
            mmsSolution(b.cx[refId[i]], b.cy[refId[i]], qq);
            q0[i] = make_double4(qq[0], qq[1], qq[2], qq[3]);
        } else {
            q0[i] = make_double4(fs[0], fs[1], fs[2], fs[3]);
        }
    }
    HIPCHK(hipMemcpy(s->d_qA.p, q0.data(), C * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(nullptr));   // landed before the solver's stream reads it
    s->cur = 0;
    refreshShadow(s);
    s->t0 = 0;
    s->dt0 = c.dt;
    launchInitScalars(s->d_sc, c.dt, 0.0, c.CFL, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return EBB_OK;
}

int createSolver(const ebb_mesh* m, const Config& cfg, int device, ebb_solver** out)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
        std::fprintf(stderr, "ebb_solver_create: no HIP device %d\n", device);
        return EBB_ERR_DEVICE;
    }
    // released to the caller on success; destroyed (device buffers included) on any failure
    std::unique_ptr<ebb_solver, void (*)(ebb_solver*)> s(new ebb_solver(), ebb_solver_destroy);
    s->raw = m->raw;
    s->b = m->b;
    s->cfg = cfg;
    s->device = device;
    {
        int lds = 0, cus = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
            s->ldsPerBlock = lds;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
            s->cus = cus;
    }
    // gradient tiles need gradients (order 2) in rows of 4 scalar limiters: other
    // configurations take the face-tile path
    if (s->cfg.stagePath == 3 && (s->cfg.order < 2 || (s->cfg.limited && s->cfg.lpThresh > 0))) s->cfg.stagePath = 1;
    if (s->cfg.limiter == 1 && (s->cfg.stagePath == 2 || (s->cfg.order > 1 && s->cfg.limited && s->cfg.lpThresh > 0))) {
        std::fprintf(stderr, "ebb_solver_create: the Venkatakrishnan limiter runs on the face-tile and split stage "
                             "paths, without the LP vector limiter\n");
        return EBB_ERR_ARG;
    }
    if (s->cfg.precision != 0) {
        // mixed precision runs the gradient and face kernels in FP32: not the LP vector
        // limiter (an FP64 simplex) and not the single fused stage kernel
        const Config& c = s->cfg;
        const char* why = nullptr;
        if (c.order > 1 && c.limited && c.lpThresh > 0) why = "the LP vector limiter (lpThresh > 0)";
        else if (c.stagePath == 2) why = "stagePath \"fused\"";
        if (why) {
            std::fprintf(stderr, "ebb_solver_create: precision \"mixed\" does not support %s\n", why);
            return EBB_ERR_ARG;
        }
    }
    int rc = buildDevice(s.get());
    if (rc != EBB_OK) {
        std::fprintf(stderr, "ebb_solver_create: %s %s\n", s->errMsg.c_str(), lastDeviceError.c_str());
        return rc;
    }
    *out = s.release();
    return EBB_OK;
}

}  // namespace ebb

// ==================================================================== C ABI
extern "C" {

const char* ebb_error_string(int code)
{
    switch (code) {
    case EBB_OK: return "ok";
    case EBB_ERR_ARG: return "bad argument";
    case EBB_ERR_IO: return "i/o error";
    case EBB_ERR_PARSE: return "parse error";
    case EBB_ERR_MESH: return "mesh error";
    case EBB_ERR_BC: return "boundary condition error";
    case EBB_ERR_DEVICE: return "device error";
    case EBB_ERR_EDGE_NAN: return "NaN flux on edge";
    case EBB_ERR_CELL_NAN: return "NaN state in cell";
    case EBB_ERR_STATE: return "invalid solver state";
    case EBB_ERR_ALLOC: return "out of memory";
    }
    return "unknown";
}

int ebb_mesh_load(const char* path, ebb_mesh** out)
{
    return guarded([&]() -> int {
        if (!path || !out) return EBB_ERR_ARG;
        std::string p(path);
        const bool gri = p.find(".gri") != std::string::npos;   // io.d:673-684 dispatch order
        if (!gri && p.find(".msh") == std::string::npos && p.find(".gmsh") == std::string::npos) return EBB_ERR_PARSE;
        auto m = std::make_unique<ebb_mesh>();
        std::string err;
        {
            std::ifstream f(p);
            if (!f) return EBB_ERR_IO;
        }
        const bool ok = gri ? parseGri(p, m->raw, err) : parseGmsh(p, m->raw, err);
        if (!ok) { std::fprintf(stderr, "ebb_mesh_load: %s\n", err.c_str()); return EBB_ERR_PARSE; }
        if (!buildMesh(m->raw, m->b, err)) { std::fprintf(stderr, "ebb_mesh_load: %s\n", err.c_str()); return EBB_ERR_MESH; }
        *out = m.release();
        return EBB_OK;
    });
}

int ebb_mesh_create(uint32_t n_nodes, const double* xy, uint32_t n_elem, const uint32_t* eoff, const uint32_t* enodes,
                    uint32_t n_bnodes, const uint32_t* bnodes, uint32_t n_tags, const char* const* tags,
                    const uint64_t* bgroup_start, ebb_mesh** out)
{
    return guarded([&]() -> int {
        if (!out || !xy || !eoff || !enodes || (n_bnodes && !bnodes) || (n_tags && (!tags || !bgroup_start))) return EBB_ERR_ARG;
        auto m = std::make_unique<ebb_mesh>();
        m->raw.nodes.assign(xy, xy + 2 * size_t(n_nodes));
        m->raw.eoff.assign(eoff, eoff + n_elem + 1);
        m->raw.enodes.assign(enodes, enodes + eoff[n_elem]);
        m->raw.bnodes.assign(bnodes, bnodes + 2 * size_t(n_bnodes));
        for (uint32_t t = 0; t < n_tags; t++) m->raw.btags.emplace_back(tags[t]);
        m->raw.bgroupStart.assign(bgroup_start, bgroup_start + n_tags);
        std::string err;
        if (!buildMesh(m->raw, m->b, err)) { std::fprintf(stderr, "ebb_mesh_create: %s\n", err.c_str()); return EBB_ERR_MESH; }
        *out = m.release();
        return EBB_OK;
    });
}

int ebb_mesh_info(const ebb_mesh* m, int64_t info[8])
{
    if (!m || !info) return EBB_ERR_ARG;
    const BuiltMesh& b = m->b;
    uint32_t maxE = 0;
    for (auto v : b.nEdgesCell) maxE = std::max<uint32_t>(maxE, v);
    info[0] = b.nCells; info[1] = b.nElem; info[2] = b.nCells - b.nElem; info[3] = b.nEdges;
    info[4] = int64_t(b.interiorEdges.size()); info[5] = int64_t(b.boundaryEdges.size());
    info[6] = maxE; info[7] = int64_t(m->raw.btags.size());
    return EBB_OK;
}

int ebb_mesh_cell_geometry(const ebb_mesh* m, double* area, double* d, double* perim, double* centroid,
                           double* grad_mat, uint32_t* edges, int8_t* flux_mult, uint32_t* neighbors, uint8_t* n_edges)
{
    return guarded([&]() -> int {
        if (!m) return EBB_ERR_ARG;
        const BuiltMesh& b = m->b;
        for (uint32_t i = 0; i < b.nCells; i++) {
            if (area) area[i] = b.area[i];
            if (d) d[i] = b.d[i];
            if (perim) perim[i] = b.perim[i];
            if (centroid) { centroid[2 * i] = b.cx[i]; centroid[2 * i + 1] = b.cy[i]; }
        }
        if (grad_mat) std::memcpy(grad_mat, b.gradMat.data(), b.gradMat.size() * sizeof(double));
        if (edges) std::memcpy(edges, b.cellEdges.data(), b.cellEdges.size() * sizeof(uint32_t));
        if (flux_mult) std::memcpy(flux_mult, b.fluxMult.data(), b.fluxMult.size());
        if (neighbors) std::memcpy(neighbors, b.nbr.data(), b.nbr.size() * sizeof(uint32_t));
        if (n_edges) std::memcpy(n_edges, b.nEdgesCell.data(), b.nEdgesCell.size());
        return EBB_OK;
    });
}

int ebb_mesh_edge_geometry(const ebb_mesh* m, uint32_t* nodes, uint32_t* cells, double* len, double* normal,
                           double* mid, int32_t* bgroup)
{
    return guarded([&]() -> int {
        if (!m) return EBB_ERR_ARG;
        const BuiltMesh& b = m->b;
        for (uint32_t e = 0; e < b.nEdges; e++) {
            if (nodes) { nodes[2 * e] = b.eNode0[e]; nodes[2 * e + 1] = b.eNode1[e]; }
            if (cells) { cells[2 * e] = b.eCell0[e]; cells[2 * e + 1] = b.eCell1[e]; }
            if (len) len[e] = b.len[e];
            if (normal) { normal[2 * e] = b.nx[e]; normal[2 * e + 1] = b.ny[e]; }
            if (mid) { mid[2 * e] = b.mx[e]; mid[2 * e + 1] = b.my[e]; }
            if (bgroup) bgroup[e] = b.eBGroup[e];
        }
        return EBB_OK;
    });
}

int ebb_mesh_tile_stats(const ebb_mesh* m, int tile_cells, int order, double out[8])
{
    return guarded([&]() -> int {
        if (!m || !out || tile_cells <= 0 || tile_cells > 1024 || order < 0 || order > 2) return EBB_ERR_ARG;
        Layout L;
        makeLayout(m->b, L);
        TileHost h;
        std::string err;
        std::vector<uint32_t> ord, start;
        tileOrder(L, uint32_t(tile_cells), order, ord, start);
        if (!makeTiles(L, ord, start, h, err)) {
            std::fprintf(stderr, "ebb_mesh_tile_stats: %s\n", err.c_str());
            return EBB_ERR_MESH;
        }
        const double n = L.N ? double(L.N) : 1.0, nt = h.nTiles ? double(h.nTiles) : 1.0;
        out[0] = double(h.nTiles);
        out[1] = double(h.sumC) / nt;
        out[2] = double(h.sumG) / n; out[3] = double(h.sumQ) / n; out[4] = double(h.sumL) / n;
        out[5] = double(h.sumF) / n;
        out[6] = h.maxL;
        out[7] = 2.0 * h.maxBlob + h.maxDyn[0];
        return EBB_OK;
    });
}

int ebb_mesh_save_mmsh(const ebb_mesh* m, const char* path)
{
    return guarded([&]() -> int {
        if (!m || !path) return EBB_ERR_ARG;
        std::string err;
        if (!saveMatlabMesh(m->raw, path, err)) {
            std::fprintf(stderr, "ebb_mesh_save_mmsh: %s\n", err.c_str());
            return err.rfind("cannot write", 0) == 0 || err.rfind("short write", 0) == 0 ? EBB_ERR_IO : EBB_ERR_MESH;
        }
        return EBB_OK;
    });
}

void ebb_mesh_destroy(ebb_mesh* m) { delete m; }

int ebb_config_parse(const char* json, ebb_config** out)
{
    return guarded([&]() -> int {
        if (!json || !out) return EBB_ERR_ARG;
        auto c = std::make_unique<ebb_config>();
        std::string err;
        if (!parseConfig(json, c->c, err)) { std::fprintf(stderr, "ebb_config_parse: %s\n", err.c_str()); return EBB_ERR_PARSE; }
        *out = c.release();
        return EBB_OK;
    });
}

static double* cfgField(Config& c, const std::string& k, bool& isInt, long long*& ip, bool*& bp)
{
    isInt = false; ip = nullptr; bp = nullptr;
    if (k == "dt") return &c.dt;
    if (k == "tEnd") return &c.tEnd;
    if (k == "CFL") return &c.CFL;
    if (k == "lpThresh") return &c.lpThresh;
    if (k == "Pr") return &c.Pr;
    if (k == "R") return &c.R;
    if (k == "Re") return &c.Re;
    if (k == "L") return &c.L;
    if (k == "rmsDrop") return &c.rmsDrop;
    if (k == "venkatK") return &c.venkatK;
    if (k.size() == 3 && k[0] == 'i' && k[1] == 'c' && k[2] >= '0' && k[2] <= '3') return &c.ic[k[2] - '0'];
    isInt = true;
    if (k == "saveIter") ip = &c.saveIter;
    else if (k == "plotIter") ip = &c.plotIter;
    else if (k == "order") ip = &c.order;
    else if (k == "maxSteps") ip = &c.maxSteps;
    else if (k == "tileCells") ip = &c.tileCells;
    else if (k == "limited") bp = &c.limited;
    else if (k == "dynamicDt") bp = &c.dynamicDt;
    else if (k == "localTimestep") bp = &c.localTimestep;
    else if (k == "multistageLimiting") bp = &c.multistageLimiting;
    else if (k == "cflAdjust") bp = &c.cflAdjust;
    else if (k == "viscosity") bp = &c.viscosity;
    else if (k == "mms") bp = &c.mms;
    return nullptr;
}

int ebb_config_get(const ebb_config* c, const char* key, double* value)
{
    return guarded([&]() -> int {
        if (!c || !key || !value) return EBB_ERR_ARG;
        bool isInt; long long* ip; bool* bp;
        double* dp = cfgField(const_cast<Config&>(c->c), key, isInt, ip, bp);
        if (dp) *value = *dp;
        else if (ip) *value = double(*ip);
        else if (bp) *value = *bp ? 1.0 : 0.0;
        else if (std::string(key) == "precision") *value = c->c.precision;
        else if (std::string(key) == "limiter") *value = c->c.limiter;
        else if (std::string(key) == "stagePath") *value = c->c.stagePath;
        else if (std::string(key) == "kernelOrder") *value = c->c.kernelOrder;
        else if (std::string(key) == "integrator") *value = c->c.integratorKind();
        else if (std::string(key) == "flux") *value = c->c.fluxKind();
        else return EBB_ERR_ARG;
        return EBB_OK;
    });
}

int ebb_config_set(ebb_config* c, const char* key, double value)
{
    return guarded([&]() -> int {
        if (!c || !key) return EBB_ERR_ARG;
        bool isInt; long long* ip; bool* bp;
        double* dp = cfgField(c->c, key, isInt, ip, bp);
        if (dp) *dp = value;
        else if (ip) *ip = (long long)value;
        else if (bp) *bp = value != 0.0;
        else {
            // enumerated keys take one of their values' numbers (checked before converting)
            const std::string k(key);
            auto take = [&](int hi, int& dst) {
                if (!(value >= 0 && value <= hi && value == std::floor(value))) return false;
                dst = int(value);
                return true;
            };
            const bool ok = (k == "precision" && take(1, c->c.precision)) || (k == "limiter" && take(1, c->c.limiter)) ||
                            (k == "stagePath" && take(3, c->c.stagePath)) ||
                            (k == "kernelOrder" && take(3, c->c.kernelOrder));
            if (!ok) return EBB_ERR_ARG;
        }
        return EBB_OK;
    });
}

int ebb_config_mesh_file(const ebb_config* c, char* buf, size_t len)
{
    if (!c || !buf || len == 0) return EBB_ERR_ARG;
    std::snprintf(buf, len, "%s", c->c.meshFile.c_str());
    return c->c.meshFile.size() < len ? EBB_OK : EBB_ERR_ARG;
}

void ebb_config_destroy(ebb_config* c) { delete c; }

void ebb_solver_destroy(ebb_solver* s);

int ebb_solver_create(const ebb_mesh* m, const ebb_config* c, int device, ebb_solver** out)
{
    return guarded([&]() -> int {
        if (!m || !c || !out) return EBB_ERR_ARG;
        return createSolver(m, c->c, device, out);
    });
}

// Copy an N-row array (rows of `per` doubles) between host (reference numbering)
// and device (device numbering).
static int uploadCells(ebb_solver* s, const double* host, void* devPtr, size_t per)
{
    const size_t n = size_t(s->N) * per;
    std::vector<double> tmp;
    const double* src = host;
    if (!s->refOf.empty()) {
        tmp.resize(n);
        toDevice(s, host, tmp.data(), per);
        src = tmp.data();
    }
    HIPCHK(hipMemcpyAsync(devPtr, src, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EBB_OK;
}

static int downloadCells(ebb_solver* s, const void* devPtr, double* host, size_t per)
{
    const size_t n = size_t(s->N) * per;
    if (s->refOf.empty()) {
        HIPCHK(hipMemcpyAsync(host, devPtr, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return EBB_OK;
    }
    std::vector<double> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), devPtr, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    toReference(s, tmp.data(), host, per);
    return EBB_OK;
}

// The same for FP32 rows (mixed precision), widened to double.
static int downloadCellsF(ebb_solver* s, const float* devPtr, double* host, size_t per)
{
    const size_t n = size_t(s->N) * per;
    std::vector<float> f(n);
    HIPCHK(hipMemcpyAsync(f.data(), devPtr, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    std::vector<double> w(f.begin(), f.end());
    toReference(s, w.data(), host, per);
    return EBB_OK;
}

int ebb_host_alloc(size_t bytes, void** out)
{
    return guarded([&]() -> int {
        if (!out) return EBB_ERR_ARG;
        *out = nullptr;
        if (bytes == 0) return EBB_ERR_ARG;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
            (void)hipGetLastError();
            return EBB_ERR_DEVICE;   // page-locked memory is the HIP runtime's
        }
        void* p = nullptr;
        const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();   // the failure is reported here, not by the next call
            return EBB_ERR_ALLOC;
        }
        *out = p;
        return EBB_OK;
    });
}

void ebb_host_free(void* p)
{
    if (p) (void)hipHostFree(p);
}

int ebb_state_upload(ebb_solver* s, const double* q)
{
    return guarded([&]() -> int {
        if (!s || !q) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        int rc = uploadCells(s, q, s->qcur(), 4);
        if (rc) return rc;
        refreshShadow(s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->stream));
        return EBB_OK;
    });
}

int ebb_state_download(ebb_solver* s, double* q)
{
    return guarded([&]() -> int {
        if (!s || !q) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        return downloadCells(s, s->qcur(), q, 4);
    });
}

int ebb_state_download_all(ebb_solver* s, double* q)
{
    return guarded([&]() -> int {
        if (!s || !q) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        int rc = downloadCells(s, s->qcur(), q, 4);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(q + size_t(s->N) * 4, s->qcur() + s->N, size_t(s->G) * sizeof(double4), hipMemcpyDeviceToHost,
                              s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return EBB_OK;
    });
}

int ebb_residual(ebb_solver* s, int limit, int dt_update, double* R, double* new_dt, double* rmax)
{
    return guarded([&]() -> int {
        if (!s) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        if (!s->d_R.p) HIPCHK((hipError_t)trackAlloc(s, s->d_R, s->N));
        if (!s->d_estate.p) HIPCHK((hipError_t)trackAlloc(s, s->d_estate, size_t(s->Fi + s->Fb) * 2));
        { int rc_ = readScalars(s); if (rc_) return rc_; }
        s->h_sc->rmaxBits = 0;
        if (rmax && *rmax > 0) std::memcpy(&s->h_sc->rmaxBits, rmax, 8);
        beginErrorWindow(s);
        s->h_sc->errEdge = ~0ull;
        s->h_sc->errCell = ~0ull;
        s->h_sc->errBC = ~0ull;
        HIPCHK(hipMemcpyAsync(s->d_sc, s->h_sc, sizeof(DevScalars), hipMemcpyHostToDevice, s->stream));
        stage(s, SK_RESIDUAL, s->qcur(), s->qcur(), nullptr, limit, dt_update, 0, 0, 0, s->d_R.p, s->d_estate.p, 0);
        HIPCHK(hipGetLastError());
        { int rc_ = readScalars(s); if (rc_) return rc_; }
        if (R) { int rc_ = downloadCells(s, s->d_R.p, R, 4); if (rc_) return rc_; }
        double nd, rm;
        std::memcpy(&nd, &s->h_sc->newDtBits, 8);
        std::memcpy(&rm, &s->h_sc->rmaxBits, 8);
        if (new_dt) *new_dt = nd;
        if (rmax) *rmax = rm;
        return checkErrors(s);
    });
}

int ebb_step(ebb_solver* s, double* dt, double* rmax)
{
    return guarded([&]() -> int {
        if (!s || !dt) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        { int rc_ = readScalars(s); if (rc_) return rc_; }
        DevScalars saved = *s->h_sc;
        s->h_sc->dt = *dt;
        beginErrorWindow(s);
        s->h_sc->errEdge = ~0ull;
        s->h_sc->errCell = ~0ull;
        s->h_sc->errBC = ~0ull;
        HIPCHK(hipMemcpyAsync(s->d_sc, s->h_sc, sizeof(DevScalars), hipMemcpyHostToDevice, s->stream));
        stepStages(s, 0, uint32_t(s->stepsTotal));
        HIPCHK(hipGetLastError());
        { int rc_ = readScalars(s); if (rc_) return rc_; }
        s->cur ^= 1;
        double nd, rm;
        std::memcpy(&nd, &s->h_sc->newDtBits, 8);
        std::memcpy(&rm, &s->h_sc->rmaxBits, 8);
        *dt = nd;
        if (rmax) *rmax = rm;
        int rc = checkErrors(s);
        // the run loop's time/dt are not advanced by a bare integrator step
        s->h_sc->dt = saved.dt;
        HIPCHK(hipMemcpyAsync(s->d_sc, s->h_sc, sizeof(DevScalars), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return rc;
    });
}

int ebb_run(ebb_solver* s, int64_t max_steps, double* t, double* dt, int64_t* steps_done)
{
    return guarded([&]() -> int {
        if (!s) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        { int rc_ = readScalars(s); if (rc_) return rc_; }
        int64_t done = 0;
        int rc = EBB_OK;
        if (s->h_sc->errEdge != ~0ull || s->h_sc->errCell != ~0ull || s->h_sc->errBC != ~0ull) return fail(s, EBB_ERR_STATE, "solver stopped on an error");
        if (!(s->h_sc->t < s->cfg.tEnd)) goto out;
        {
            // (re)activate: the device stops itself at tEnd, after maxSteps or on an error
            const int64_t cfgMax = s->cfg.maxSteps;
            int64_t limit = max_steps;
            if (cfgMax >= 0) {
                const int64_t left = std::max<int64_t>(0, cfgMax - s->stepsTotal);
                limit = (limit < 0) ? left : std::min(limit, left);
            }
            if (limit == 0) goto out;
            s->h_sc->active = 1;
            HIPCHK(hipMemcpyAsync(&s->d_sc->active, &s->h_sc->active, sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
            while (limit < 0 || done < limit) {
                const int64_t batch = limit < 0 ? s->histCap : std::min<int64_t>(s->histCap, limit - done);
                const int64_t base = s->h_sc->steps;
                beginErrorWindow(s);   // no error is pending: the previous batch would have stopped
                const int64_t stopAt = base + batch;
                int c0 = s->cur;
                for (int64_t k = 0; k < batch; k++) {
                    const double4* qold = s->qcur();
                    stepStages(s, 1, uint32_t(base + k));
                    const double4* qnew = s->qnext();
                    DiagArgs d = diagArgs(s, qnew, qold, base, stopAt);
                    launchRmsChunks(d, s->stream);
                    launchFinalize(d, s->stream);
                    s->cur ^= 1;
                }
                // the batch's last step (or the one that ended the loop) still has its force pending
                if (deferForce(s)) launchPendingForce(forceArgs(s), s->d_sc, s->stream);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(s->h_hist, s->d_hist.p, sizeof(double) * size_t(batch) * HIST_FIELDS,
                                      hipMemcpyDeviceToHost, s->stream));
                { int rc_ = readScalars(s); if (rc_) return rc_; }
                const int64_t ran = s->h_sc->steps - base;
                s->cur = (c0 ^ int(ran & 1));
                s->history.insert(s->history.end(), s->h_hist, s->h_hist + ran * HIST_FIELDS);
                s->stepsTotal += ran;
                done += ran;
                rc = checkErrors(s);
                if (rc != EBB_OK) break;
                if (ran < batch) break;              // t >= tEnd
                if (!(s->h_sc->t < s->cfg.tEnd)) break;
                s->h_sc->active = 1;
                HIPCHK(hipMemcpyAsync(&s->d_sc->active, &s->h_sc->active, sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
            }
        }
    out:
        if (t) *t = s->h_sc->t;
        if (dt) *dt = s->h_sc->dt;
        if (steps_done) *steps_done = done;
        return rc;
    });
}

int ebb_run_state(ebb_solver* s, double* t, double* dt, double* cfl, int64_t* steps)
{
    return guarded([&]() -> int {
        if (!s) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        { int rc_ = readScalars(s); if (rc_) return rc_; }
        if (t) *t = s->h_sc->t;
        if (dt) *dt = s->h_sc->dt;
        if (cfl) *cfl = s->h_sc->CFL;
        if (steps) *steps = s->stepsTotal;
        return EBB_OK;
    });
}

int ebb_run_reset(ebb_solver* s, double t, double dt)
{
    return guarded([&]() -> int {
        if (!s) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        launchInitScalars(s->d_sc, dt, t, s->cfg.CFL, s->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->stream));
        s->history.clear();
        s->stepsTotal = 0;
        s->errKind = 0;
        return EBB_OK;
    });
}

int64_t ebb_history_length(ebb_solver* s) { return s ? int64_t(s->history.size() / HIST_FIELDS) : 0; }

int ebb_history(ebb_solver* s, int64_t first, int64_t count, double* out)
{
    return guarded([&]() -> int {
        if (!s || !out || first < 0 || count < 0) return EBB_ERR_ARG;
        if ((first + count) * HIST_FIELDS > int64_t(s->history.size())) return EBB_ERR_ARG;
        std::memcpy(out, s->history.data() + first * HIST_FIELDS, sizeof(double) * size_t(count) * HIST_FIELDS);
        return EBB_OK;
    });
}

// Recompute the edge states of the last residual evaluation into d_estate.
static int refreshEdgeStates(ebb_solver* s)
{
    if (!s->d_estate.p) HIPCHK((hipError_t)trackAlloc(s, s->d_estate, size_t(s->Fi + s->Fb) * 2));
    if (!s->haveStage) return EBB_OK;
    // re-run the last stage's gradient + face kernels on its (unchanged) input;
    // the update kernel is not re-run, so no state or scalar changes
    StageArgs a{};
    a.qin = s->lastInput; a.qn = s->lastInput; a.grad = s->d_grad.p; a.lim = s->d_lim.p;
    a.lenF = s->d_lenF.p; a.lenS = s->d_lenS.p; a.estate = s->d_estate.p; a.celldt = s->d_celldt.p;
        a.sc = s->d_sc; a.kind = SK_RESIDUAL; a.limit = s->lastLimit; a.first = 0; a.checkActive = 0; a.recordErrors = 0;
        mixArgs(s, a, s->lastInput, nullptr);
    // keep the device accumulators: the gradient kernel resets newDt, restore it afterwards
    { int rc_ = readScalars(s); if (rc_) return rc_; }
    DevScalars saved = *s->h_sc;
    launchGradient(s->dm, s->phys, a, int(s->S), s->stream);
    launchFaces(s->dm, s->phys, a, s->stream);
    HIPCHK(hipGetLastError());
    *s->h_sc = saved;
    HIPCHK(hipMemcpyAsync(s->d_sc, s->h_sc, sizeof(DevScalars), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return EBB_OK;
}

int ebb_edge_states(ebb_solver* s, double* eq, double* flux, double* smax)
{
    return guarded([&]() -> int {
        if (!s) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        int rc = refreshEdgeStates(s);
        if (rc) return rc;
        const uint32_t F = s->Fi + s->Fb;
        std::vector<double4> es(size_t(F) * 2);
        std::vector<double> lf(size_t(F) * 4), ls(F);
        HIPCHK(hipMemcpy(es.data(), s->d_estate.p, es.size() * sizeof(double4), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(lf.data(), s->d_lenF.p, lf.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ls.data(), s->d_lenS.p, ls.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (uint32_t f = 0; f < F; f++) {
            const uint32_t e = s->edgeOfFace[f];
            if (eq) {
                const double4 a = es[2 * f], b = es[2 * f + 1];
                double* o = eq + size_t(e) * 8;
                o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
            }
            const double len = s->b.len[e];
            // flux = (len*flux)/len is exact only up to rounding; report len*flux/len
            if (flux) for (int k = 0; k < 4; k++) flux[size_t(e) * 4 + k] = lf[size_t(f) * 4 + k] / len;
            if (smax) smax[e] = ls[f] / len;
        }
        return EBB_OK;
    });
}

int ebb_cell_gradients(ebb_solver* s, double* grad, double* lim)
{
    return guarded([&]() -> int {
        if (!s) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        // the fused and gradient-tile kernels keep gradients in LDS: recompute them for the last
        // stage input
        if (s->useTiles || s->useGradTiles) { int rc_ = refreshEdgeStates(s); if (rc_) return rc_; }
        if (grad) {
            int rc_ = s->phys.mixed ? downloadCellsF(s, s->d_grad32.p, grad, 8) : downloadCells(s, s->d_grad.p, grad, 8);
            if (rc_) return rc_;
        }
        if (lim && !(s->phys.limited && s->lastLimit)) {
            // an unlimited evaluation resets every limiter to 1 (solve.d:324-340) and stores none
            std::fill(lim, lim + size_t(s->N) * 8, 1.0);
        } else if (lim) {
            const bool lp = s->phys.order > 1 && s->phys.limited && s->phys.lpThresh > 0;
            std::vector<double> l(size_t(s->N) * (lp ? 8 : 4));
            {
                int rc_ = s->phys.mixed ? downloadCellsF(s, s->d_lim32.p, l.data(), 4)
                                        : downloadCells(s, s->d_lim.p, l.data(), lp ? 8 : 4);
                if (rc_) return rc_;
            }
            for (uint32_t i = 0; i < s->N; i++)
                for (int k = 0; k < 4; k++) {
                    lim[size_t(i) * 8 + 2 * k] = lp ? l[size_t(i) * 8 + 2 * k] : l[size_t(i) * 4 + k];
                    lim[size_t(i) * 8 + 2 * k + 1] = lp ? l[size_t(i) * 8 + 2 * k + 1] : l[size_t(i) * 4 + k];
                }
        }
        return EBB_OK;
    });
}

// source/io.d:68-161
static int saveSolutionImpl(ebb_solver* s, const char* path, const double* tdt);

int ebb_save_solution(ebb_solver* s, const char* path)
{
    return guarded([&]() -> int { return saveSolutionImpl(s, path, nullptr); });
}

int ebb_save_solution_as(ebb_solver* s, const char* path, double t, double dt)
{
    return guarded([&]() -> int {
        const double tdt[2] = {t, dt};
        return saveSolutionImpl(s, path, tdt);
    });
}

static int saveSolutionImpl(ebb_solver* s, const char* path, const double* tdt)
{
    if (!s || !path) return EBB_ERR_ARG;
    HIPCHK(hipSetDevice(s->device));
    int rc = refreshEdgeStates(s);
    if (rc) return rc;
    const uint32_t F = s->Fi + s->Fb, N = s->N;
    std::vector<double4> es(size_t(F) * 2), q(N);
    HIPCHK(hipMemcpy(es.data(), s->d_estate.p, es.size() * sizeof(double4), hipMemcpyDeviceToHost));
    { int rc_ = downloadCells(s, s->qcur(), reinterpret_cast<double*>(q.data()), 4); if (rc_) return rc_; }
    { int rc_ = readScalars(s); if (rc_) return rc_; }
    if (!s->haveStage) {
        // no residual evaluated yet: edge states are the set-up values (far states)
        for (uint32_t k = 0; k < s->Fb; k++) {
            const BFace& bf = s->bfaces[k];
            es[2 * (s->Fi + k) + 1] = make_double4(bf.far[0], bf.far[1], bf.far[2], bf.far[3]);
        }
    }
    std::vector<uint8_t> buf;
    buf.reserve(64 + size_t(N) * 150);
    putBE32(buf, 0xEA98E1F5u);
    putBE32(buf, 2);
    putBE32(buf, N);
    putBEd(buf, tdt ? tdt[0] : s->h_sc->t);
    putBEd(buf, tdt ? tdt[1] : s->h_sc->dt);
    putBE32(buf, 4);
    putBE32(buf, uint32_t(s->cfg.order));
    for (uint32_t i = 0; i < N; i++) {
        const uint32_t ne = s->b.nEdgesCell[i];
        putBE32(buf, i);
        buf.push_back(uint8_t(ne));
        putBEd(buf, q[i].x); putBEd(buf, q[i].y); putBEd(buf, q[i].z); putBEd(buf, q[i].w);
        for (uint32_t j = 0; j < ne; j++) {
            const uint32_t e = s->b.cellEdges[size_t(i) * MAXE + j];
            const uint32_t f = s->faceOfEdge[e];
            const double4 v = (s->b.eCell0[e] == i) ? es[2 * f] : es[2 * f + 1];
            putBEd(buf, v.x); putBEd(buf, v.y); putBEd(buf, v.z); putBEd(buf, v.w);
        }
    }
    // the reference computes the CRC over its buffer including the 4 not yet written CRC bytes (zero here)
    const size_t body = buf.size();
    buf.resize(body + 4, 0);
    const uint32_t crc = crc32(buf.data(), buf.size());
    buf[body + 0] = uint8_t(crc);
    buf[body + 1] = uint8_t(crc >> 8);
    buf[body + 2] = uint8_t(crc >> 16);
    buf[body + 3] = uint8_t(crc >> 24);
    FILE* f = std::fopen(path, "wb");
    if (!f) return EBB_ERR_IO;
    const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return w == buf.size() ? EBB_OK : EBB_ERR_IO;
}

// source/io.d:218-366 (single rank: data points must equal the interior cell count)
int ebb_load_solution(ebb_solver* s, const char* path)
{
    return guarded([&]() -> int {
        if (!s || !path) return EBB_ERR_ARG;
        std::ifstream f(path, std::ios::binary);
        if (!f) return EBB_ERR_IO;
        std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        if (buf.size() < 28 || getBE32(buf.data()) != 0xEA98E1F5u) return EBB_ERR_PARSE;
        const uint32_t ver = getBE32(buf.data() + 4), np = getBE32(buf.data() + 8);
        if (np != s->N) return EBB_ERR_PARSE;
        const double t = getBEd(buf.data() + 12), dt = getBEd(buf.data() + 20);
        size_t off = 28;
        std::vector<double4> q(s->N);
        if (ver == 1) {
            if (buf.size() < off + size_t(np) * 32) return EBB_ERR_PARSE;
            for (uint32_t i = 0; i < np; i++, off += 32)
                q[i] = make_double4(getBEd(&buf[off]), getBEd(&buf[off + 8]), getBEd(&buf[off + 16]), getBEd(&buf[off + 24]));
        } else if (ver == 2) {
            off += 8;  // M, order
            for (uint32_t i = 0; i < np; i++) {
                if (buf.size() < off + 5 + 32) return EBB_ERR_PARSE;
                const uint8_t ne = buf[off + 4];
                off += 5;
                q[i] = make_double4(getBEd(&buf[off]), getBEd(&buf[off + 8]), getBEd(&buf[off + 16]), getBEd(&buf[off + 24]));
                off += 32 + size_t(ne) * 32;
                if (off > buf.size()) return EBB_ERR_PARSE;
            }
        } else {
            return EBB_ERR_PARSE;
        }
        HIPCHK(hipSetDevice(s->device));
        { int rc_ = uploadCells(s, reinterpret_cast<const double*>(q.data()), s->qcur(), 4); if (rc_) return rc_; }
        refreshShadow(s);
        return ebb_run_reset(s, t, dt);
    });
}

int ebb_write_forces(ebb_solver* s, const char* path, int64_t first, int64_t count, int append)
{
    return guarded([&]() -> int {
        if (!s || !path || first < 0 || count < 0) return EBB_ERR_ARG;
        if ((first + count) * HIST_FIELDS > int64_t(s->history.size())) return EBB_ERR_ARG;
        std::vector<uint8_t> buf;
        buf.reserve(size_t(count) * 24);
        for (int64_t k = first; k < first + count; k++) {
            const double* h = s->history.data() + k * HIST_FIELDS;
            putBEd(buf, h[0]);
            putBEd(buf, h[7]);
            putBEd(buf, h[8]);
        }
        FILE* f = std::fopen(path, append ? "ab" : "wb");
        if (!f) return EBB_ERR_IO;
        const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
        std::fclose(f);
        return w == buf.size() ? EBB_OK : EBB_ERR_IO;
    });
}

int ebb_last_error(ebb_solver* s, int* kind, int64_t* index, int64_t* step, char* msg, size_t len)
{
    if (!s) return EBB_ERR_ARG;
    if (kind) *kind = s->errKind;
    if (index) *index = s->errIndex;
    if (step) *step = s->errStep;
    if (msg && len) std::snprintf(msg, len, "%s%s%s", s->errMsg.c_str(), lastDeviceError.empty() ? "" : " ", lastDeviceError.c_str());
    return EBB_OK;
}

int ebb_time_steps(ebb_solver* s, int steps, double out[6])
{
    return guarded([&]() -> int {
        if (!s || !out || steps <= 0) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(s->device));
        { int rc_ = readScalars(s); if (rc_) return rc_; }
        DevScalars saved = *s->h_sc;
        const int nst = stagesPerStep(s->cfg);
        std::vector<hipEvent_t> ev(size_t(steps) * nst * 3 + 1);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        std::vector<int> part;   // which figure each event interval adds to
        // same launch sequence as stepStages, with an event after each kernel
        HIPCHK(hipEventRecord(ev[0], s->stream));
        size_t k = 1;
        for (int st = 0; st < steps; st++) {
            double4* q = s->qcur();
            const std::vector<StageDesc> seq = stepSequence(s);
            for (auto& x : seq) {
                StageArgs a{};
                a.qin = x.in; a.qn = q; a.qout = x.out; a.acc = s->d_acc.p; a.grad = s->d_grad.p; a.lim = s->d_lim.p;
                a.lenF = s->d_lenF.p; a.lenS = s->d_lenS.p; a.celldt = s->d_celldt.p; a.partial = s->d_partial.p; a.sc = s->d_sc;
                a.kind = x.kind; a.limit = x.limit; a.dtUpdate = x.dtUpdate; a.first = x.first; a.last = x.last;
                a.rmsBlocks = (x.last && s->rmsFused) ? s->d_rmsBlocks.p : nullptr;
                a.checkActive = 0; a.recordErrors = 1; a.stageId = s->stageCounter++; a.stepId = uint32_t(s->stepsTotal);
                a.order = s->cfg.kernelOrder;
                mixArgs(s, a, x.in, x.out);
                s->stageStep.push_back(s->stepsTotal);
                if (s->useTiles) {
                    launchStageTiles(s->dm, s->phys, a, s->tiles, int(s->S), s->stream);
                    HIPCHK(hipEventRecord(ev[k++], s->stream));
                    part.push_back(3);
                } else if (s->useGradTiles) {
                    launchGradTiles(s->dm, s->phys, a, s->ftiles, int(s->S), s->stream, 1);
                    HIPCHK(hipEventRecord(ev[k++], s->stream));
                    launchGradTiles(s->dm, s->phys, a, s->ftiles, int(s->S), s->stream, 2);
                    HIPCHK(hipEventRecord(ev[k++], s->stream));
                    part.push_back(0); part.push_back(3);
                } else if (s->useFaceTiles) {
                    launchGradient(s->dm, s->phys, a, int(s->S), s->stream);
                    HIPCHK(hipEventRecord(ev[k++], s->stream));
                    launchFaceTiles(s->dm, s->phys, a, s->ftiles, int(s->S), s->stream);
                    HIPCHK(hipEventRecord(ev[k++], s->stream));
                    part.push_back(0); part.push_back(3);
                } else {
                    launchGradient(s->dm, s->phys, a, int(s->S), s->stream);
                    HIPCHK(hipEventRecord(ev[k++], s->stream));
                    launchFaces(s->dm, s->phys, a, s->stream);
                    HIPCHK(hipEventRecord(ev[k++], s->stream));
                    launchUpdate(s->dm, s->phys, a, int(s->S), s->stream);
                    HIPCHK(hipEventRecord(ev[k++], s->stream));
                    part.push_back(0); part.push_back(1); part.push_back(2);
                }
                s->lastInput = x.in;
                s->lastLimit = x.limit;
                s->haveStage = true;
            }
            s->cur ^= 1;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventSynchronize(ev[k - 1]));
        double tp[4] = {0, 0, 0, 0};
        for (size_t j = 1; j < k; j++) {
            float a;
            HIPCHK(hipEventElapsedTime(&a, ev[j - 1], ev[j]));
            tp[part[j - 1]] += a;
        }
        float total;
        HIPCHK(hipEventElapsedTime(&total, ev[0], ev[k - 1]));
        const double n = double(steps) * nst;
        out[0] = tp[0] / n; out[1] = tp[1] / n; out[2] = tp[2] / n; out[3] = tp[3] / n;
        out[4] = double(total) / steps;
        out[5] = nst;
        for (auto& e : ev) (void)hipEventDestroy(e);
        // timing steps advance the state but not the run loop's clock
        *s->h_sc = saved;
        HIPCHK(hipMemcpyAsync(s->d_sc, s->h_sc, sizeof(DevScalars), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return EBB_OK;
    });
}

int64_t ebb_solver_device_bytes(const ebb_solver* s) { return s ? s->deviceBytes : 0; }

void ebb_solver_destroy(ebb_solver* s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->d_nbr.release(); s->d_cface.release(); s->d_refId.release(); s->d_forceFaces.release();
    s->d_nia.release(); s->d_cd.release(); s->d_cper.release(); s->d_flen.release();
    s->d_grad.release(); s->d_lim.release(); s->d_lenF.release(); s->d_lenS.release(); s->d_celldt.release();
    s->d_hist.release(); s->d_chunkTrees.release(); s->d_chunkMask.release(); s->d_partial.release();
    s->d_cen.release(); s->d_fn.release(); s->d_fmid.release(); s->d_fcell.release(); s->d_bface.release();
    s->d_src.release(); s->d_qA.release(); s->d_qB.release(); s->d_T.release(); s->d_T2.release();
    s->d_acc.release(); s->d_R.release(); s->d_estate.release();
    s->d_tDesc.release(); s->d_tBlob.release();
    s->d_ftDesc.release(); s->d_ftRef.release(); s->d_ftIds.release(); s->d_ftCf.release(); s->d_ftFid.release(); s->d_node.release();
    s->d_ftSlot.release(); s->d_ftHalo.release(); s->d_ftNHalo.release(); s->d_border.release(); s->d_ftNbr.release();
    s->d_devOf.release(); s->d_rmsBlocks.release();
    s->d_qA32.release(); s->d_qB32.release(); s->d_T32.release(); s->d_T232.release();
    s->d_grad32.release(); s->d_lim32.release();
    if (s->d_sc) (void)hipFree(s->d_sc);
    if (s->h_sc) (void)hipHostFree(s->h_sc);
    if (s->h_hist) (void)hipHostFree(s->h_hist);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"

// Internal definitions shared by the C ABI's translation units (solver.cpp: one mesh per
// solver; batch.cpp: many solvers advanced by the same launches).  Not part of the ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/ebbcfd.h"
#include "dev.hpp"
#include "host.hpp"
#include "tiles.hpp"

using namespace ebb;

struct ebb_mesh {
    RawMesh raw;
    BuiltMesh b;
};

struct ebb_config {
    Config c;
};

namespace ebb {

extern thread_local std::string lastDeviceError;

#define HIPCHK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            lastDeviceError = std::string(#x) + ": " + hipGetErrorString(e_); \
            return EBB_ERR_DEVICE;                                    \
        }                                                             \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count)
    {
        n = count;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e == hipSuccess) e = hipMemset(p, 0, count * sizeof(T));
        // the fill runs on the null stream, which the library's non-blocking streams do not
        // wait for: it must finish before any of them writes the buffer
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        return e;
    }
    hipError_t upload(const std::vector<T>& v)
    {
        hipError_t e = alloc(v.size());
        if (e == hipSuccess && !v.empty()) e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        // a copy from pageable memory may return before its transfer lands; the kernels that
        // read the buffer run on non-blocking streams
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace ebb

struct ebb_solver {
    // host copies
    RawMesh raw;
    BuiltMesh b;
    Config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t N = 0, G = 0, Fi = 0, Fb = 0, S = 3;
    std::vector<uint32_t> faceOfEdge;   // reference edge -> device face
    std::vector<uint32_t> edgeOfFace;   // device face -> reference edge
    std::vector<BFace> bfaces;
    PhysParams phys{};
    DevMesh dm{};
    // device buffers
    DBuf<uint32_t> d_nbr, d_cface, d_refId, d_forceFaces;
    DBuf<double> d_nia, d_cd, d_cper, d_flen, d_grad, d_lim, d_lenF, d_lenS, d_celldt, d_hist, d_chunkTrees;
    DBuf<double> d_partial;
    DBuf<double2> d_cen, d_fn, d_fmid;
    DBuf<uint2> d_fcell;
    DBuf<BFace> d_bface;
    DBuf<double4> d_src, d_qA, d_qB, d_T, d_T2, d_acc, d_R, d_estate;
    DBuf<uint32_t> d_chunkMask;
    // mixed precision: FP32 shadows of the state buffers (paired with qA, qB, T, T2) and FP32
    // gradient and limiter rows
    DBuf<float4> d_qA32, d_qB32, d_T32, d_T232;
    DBuf<float> d_grad32, d_lim32;
    // fused-stage tiles
    DBuf<uint4> d_tDesc, d_tBlob;
    TileArgs tiles{};
    bool useTiles = false;
    uint32_t tileCells = 0;
    DBuf<uint4> d_ftDesc, d_ftRef;
    DBuf<uint32_t> d_ftIds, d_ftFid;
    DBuf<uint16_t> d_ftCf;
    DBuf<double2> d_node;
    FaceTileArgs ftiles{};
    bool useFaceTiles = false;
    // gradient tiles (face tiles whose kernel computes its own cells' gradients)
    DBuf<uint32_t> d_ftSlot, d_ftHalo, d_ftNHalo, d_border, d_ftNbr;
    bool useGradTiles = false;
    double gradTileStats[3] = {0, 0, 0};   // largest rows per tile, ring cells summed over tiles, border cells
    int ldsPerBlock = 65536;               // hipDeviceProp_t::sharedMemPerBlock of the solver's device
    int cus = 256;                         // its compute units
    // device cell numbering of the tiled paths: device cell d is reference cell
    // refOf[d] (empty: identical numbering); d_devOf maps back for the RMS sums
    std::vector<uint32_t> refOf, devOf;
    DBuf<uint32_t> d_devOf;
    // RMS leaf sums computed by the stage that completes a step (split path, and
    // face tiles cut along aligned 16-cell blocks)
    bool rmsFused = false;
    DBuf<double> d_rmsBlocks;
    DevScalars* d_sc = nullptr;
    DevScalars* h_sc = nullptr;  // pinned
    double* h_hist = nullptr;    // pinned
    int64_t histCap = HIST_RING_ROWS;
    // time-loop bookkeeping
    int cur = 0;                 // 0: qA holds the current state, 1: qB
    double4* lastInput = nullptr;
    int lastLimit = 1;
    bool haveStage = false;
    uint32_t stageCounter = 0;
    // the step of each stage launched since the error words were last cleared (stage
    // stageBase + k ran in step stageStep[k]): edge and boundary errors carry their stage
    uint32_t stageBase = 0;
    std::vector<int64_t> stageStep;
    // stage partials: pairs per stage (the stage kernel's workgroups), region stride,
    // and the stage of the current time-loop step
    uint32_t nPart = 0, pstride = 0, stageInStep = 0;
    int64_t stepsTotal = 0;
    double t0 = 0, dt0 = 0;
    std::vector<double> history;     // HIST_FIELDS per completed step of ebb_run
    // errors
    int errKind = 0;
    int64_t errIndex = -1, errStep = -1;
    std::string errMsg;

    double4* qcur() { return cur == 0 ? d_qA.p : d_qB.p; }
    double4* qnext() { return cur == 0 ? d_qB.p : d_qA.p; }
    int64_t deviceBytes = 0;
};

namespace ebb {

// set-up (ufvmSetup + device upload) of a solver whose raw mesh, built mesh, config and
// device are filled in; EBB_OK or an error code (s->errMsg says why)
int buildDevice(ebb_solver* s);
// the arguments of one stage of a time-loop step (checkActive, deferred reductions)
StageArgs loopStageArgs(ebb_solver* s, int kind, double4* qin, const double4* qn, double4* qout, int limit,
                        int dtUpdate, int first, int last, uint32_t stageInStep, uint32_t stageId, uint32_t stepId);
// the stage sequence of one integrator step from qcur() to qnext() (source/integrators.d:22-270)
struct StageDesc {
    int kind;
    double4 *in, *out;
    int limit, dtUpdate, first, last;
};
std::vector<StageDesc> stepSequence(ebb_solver* s);
DiagArgs diagArgs(ebb_solver* s, const double4* qnew, const double4* qold, int64_t histBase, int64_t maxSteps);
ForceArgs forceArgs(const ebb_solver* s);
bool deferForce(const ebb_solver* s);
int stagesPerStep(const Config& c);
void refreshShadow(ebb_solver* s);
// the mixed-precision pointers of a stage reading qin and writing qout (null on the FP64 path)
void mixArgs(ebb_solver* s, StageArgs& a, const double4* qin, const double4* qout);
int readScalars(ebb_solver* s);
int checkErrors(ebb_solver* s);
// device bytes a solver for this mesh and config would allocate (admission by footprint)
int64_t estimateDeviceBytes(const BuiltMesh& b, const Config& c);
// face tiles cut again with tileCells per tile (batched runs' tuning)
int retileFaceTiles(ebb_solver* s, long long tileCells);
// automatic face-tile size that spreads `cells` over two tiles per CU (16-cell blocks, 32-768)
uint32_t spreadTileCells(uint64_t cells, int cus);
int fail(ebb_solver* s, int code, const std::string& msg);
// creates a solver (ufvmSetup + upload) as ebb_solver_create does, with config overrides applied
int createSolver(const ebb_mesh* m, const Config& c, int device, ebb_solver** out);

}  // namespace ebb

// Every C entry point that parses, allocates or touches the device runs its body
// through guarded(): an exception (std::bad_alloc, a length_error from a size read
// out of a file, ...) becomes a status code instead of unwinding into the caller.
template <typename F>
auto guarded(F&& f) -> decltype(f())
{
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return EBB_ERR_ALLOC;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ebbcfd: %s\n", e.what());
        return EBB_ERR_PARSE;
    } catch (...) {
        return EBB_ERR_PARSE;
    }
}


// ebb_residual_noalloc: the residual evaluation of ebb_residual (source/solve.d:322-1122)
// for callers that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid it).  It allocates nothing: R goes to a device buffer the
// caller passes, no edge states are stored, and the per-workgroup (min dt, max |R|)
// partials of the update kernel live in `scratch`, a file-scope device array of fixed size
// (each GPU the library runs on holds one until the process ends).  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race.  Each workgroup writes its
// own partial before the reduction reads it, so nothing has to start at zero.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "solver_impl.hpp"

namespace {

// one (min dt candidate, max |R|) pair per update workgroup of 256 cells: 65,536 pairs
// (1 MiB) cover 16,777,216 interior cells
constexpr uint32_t SCRATCH_PAIRS = 65536;
__device__ __align__(256) double scratch[2 * SCRATCH_PAIRS];

constexpr int MAX_DEVICES = 64;
double* scratchPtr[MAX_DEVICES] = {};
std::mutex scratchMu;

// The device address of `scratch` on `device` (which must be current), looked up once per
// device: the host-side symbol is only a stand-in and must never reach a kernel.
int scratchFor(int device, double** out)
{
    if (device < 0 || device >= MAX_DEVICES) return EBB_ERR_DEVICE;
    std::lock_guard<std::mutex> g(scratchMu);
    if (!scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            lastDeviceError = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return EBB_ERR_DEVICE;
        }
        scratchPtr[device] = static_cast<double*>(p);
    }
    *out = scratchPtr[device];
    return EBB_OK;
}

// Rmax starts from the caller's value (the reference max-accumulates it, solve.d:1117);
// the error words are cleared for this evaluation.  A kernel, not a host copy, so the call
// needs no synchronisation before its launches.
__global__ void k_prep_residual(DevScalars* sc, unsigned long long rmaxBits)
{
    sc->rmaxBits = rmaxBits;
    sc->errEdge = ~0ull;
    sc->errCell = ~0ull;
    sc->errBC = ~0ull;
}

}  // namespace

extern "C" int ebb_residual_noalloc(ebb_solver* s, int limit, int dt_update, double* R_dev, double* new_dt,
                                    double* rmax)
{
    return guarded([&]() -> int {
        if (!s || !R_dev) return EBB_ERR_ARG;
        // R is written in device cell order, which is the reference order unless the solver
        // renumbers its cells (extensions.cellOrder = "tiles")
        if (!s->refOf.empty()) return EBB_ERR_ARG;
        const uint32_t pairs = (s->N + 255) / 256;
        if (pairs > SCRATCH_PAIRS) {
            std::fprintf(stderr, "ebb_residual_noalloc: %u interior cells exceed the %u this entry point's fixed "
                                 "scratch holds\n", s->N, SCRATCH_PAIRS * 256u);
            return EBB_ERR_ALLOC;
        }
        HIPCHK(hipSetDevice(s->device));
        double* ptr = nullptr;
        { int rc = scratchFor(s->device, &ptr); if (rc) return rc; }
        unsigned long long bits = 0;
        if (rmax && *rmax > 0) std::memcpy(&bits, rmax, 8);
        hipLaunchKernelGGL(k_prep_residual, dim3(1), dim3(1), 0, s->stream, s->d_sc, bits);
        // the split kernels of ebb_residual, with R to the caller's buffer and the partials in scratch
        StageArgs a{};
        a.qin = s->qcur(); a.qn = s->qcur(); a.qout = nullptr; a.acc = s->d_acc.p;
        a.R = reinterpret_cast<double4*>(R_dev);
        a.grad = s->d_grad.p; a.lim = s->d_lim.p; a.lenF = s->d_lenF.p; a.lenS = s->d_lenS.p;
        a.celldt = s->d_celldt.p; a.partial = ptr; a.sc = s->d_sc;
        a.kind = SK_RESIDUAL; a.limit = limit; a.dtUpdate = dt_update; a.first = 0; a.last = 0;
        a.checkActive = 0; a.recordErrors = 1; a.stageId = 0; a.stepId = 0;
        mixArgs(s, a, a.qin, nullptr);
        launchGradient(s->dm, s->phys, a, int(s->S), s->stream);
        launchFaces(s->dm, s->phys, a, s->stream);
        launchUpdate(s->dm, s->phys, a, int(s->S), s->stream);   // + the stage reduction of the partials
        HIPCHK(hipGetLastError());
        s->lastInput = a.qin;
        s->lastLimit = limit;
        s->haveStage = true;
        if (!new_dt && !rmax) return EBB_OK;   // results stay in the solver's device scalars
        HIPCHK(hipMemcpyAsync(s->h_sc, s->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        double nd, rm;
        std::memcpy(&nd, &s->h_sc->newDtBits, 8);
        std::memcpy(&rm, &s->h_sc->rmaxBits, 8);
        if (new_dt) *new_dt = nd;
        if (rmax) *rmax = rm;
        return checkErrors(s);   // the NaN edge (or boundary failure) of this evaluation, as ebb_residual reports it
    });
}

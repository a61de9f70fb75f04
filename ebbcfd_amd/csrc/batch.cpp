// Batched solves: many meshes advanced by the same launches.
//
// The reference runs one mesh per process: its integrator scratch is static and sized by
// the first mesh (source/integrators.d:56-75).  Here each admitted solve is an ordinary
// solver (its state, geometry and scratch resident in HBM), and a time-loop step for all of
// them is one launch per kernel: per RK stage the gradient kernel over every mesh's cell
// blocks and the face/update kernel over every mesh's face tiles, then one RMS kernel and
// one epilogue for all meshes.  Every workgroup runs the solo kernels' code on one
// (mesh, block/tile) item, so each mesh's states and history equal its solo run's bit
// for bit.  The step sequence is captured once as a HIP graph of `graphSteps` steps and
// replayed; meshes that stop (tEnd, maxSteps, the rho-RMS drop of extensions.rmsDrop, an
// error) drop out of the work lists at the next run call.  Admission is by device
// footprint against the batch's byte limit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "solver_impl.hpp"

using namespace ebb;

namespace {

struct BatchMember {
    int64_t id = 0;
    ebb_solver* s = nullptr;
    int64_t bytes = 0;         // device bytes held
    bool running = true;       // still stepping (no stop condition met)
    int errKind = 0;
    int64_t errIndex = -1;
};

// Members that share one kernel shape and stage sequence run in the same launches.
struct BatchGroup {
    GroupShape shape{};
    int nst = 0;
    int limit[4] = {0, 0, 0, 0};
    uint32_t lds = 0;
    std::vector<uint2> grad, tile;
    DBuf<uint2> dGrad, dTile;
};

using GroupKey = std::tuple<int, int, int, int, int, int, int, int, int, int, int, int, int>;

GroupKey groupKey(const ebb_solver* s, const std::vector<StageDesc>& seq)
{
    int lim[4] = {0, 0, 0, 0};
    for (size_t k = 0; k < seq.size() && k < 4; k++) lim[k] = seq[k].limit;
    return GroupKey(int(s->S), s->phys.flux, s->phys.viscous, s->phys.order, s->phys.limited, s->phys.mixed,
                    int(seq.size()), lim[0], lim[1], lim[2], lim[3], s->phys.venkat, s->useGradTiles ? 1 : 0);
}

template <typename T>
hipError_t uploadVec(DBuf<T>& b, const std::vector<T>& v)
{
    b.release();
    return b.upload(v);
}

}  // namespace

struct ebb_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t limit = 0, used = 0;
    int graphSteps = 8;
    bool keepHistory = false;
    long long tileCells = -1;
    int64_t nextId = 0;
    std::vector<BatchMember> members;
    std::vector<BatchGroup> groups;
    std::vector<uint32_t> slotOf;                 // running member index of each slot
    std::vector<MemberArgs> hmem;
    DBuf<MemberArgs> dmem;
    std::vector<uint2> rms;
    DBuf<uint2> dRms;
    std::vector<uint32_t> fin, force;
    DBuf<uint32_t> dFin, dForce;
    DBuf<DevScalars> dScal;
    DevScalars* hScal = nullptr;
    size_t hScalCap = 0;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int maxStages = 0;
    std::string err;

    BatchMember* find(int64_t id)
    {
        for (auto& m : members) if (m.id == id) return &m;
        return nullptr;
    }
};

namespace {

void dropGraph(ebb_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Per-mesh launch records of a running member, for both parities of the batch step.
MemberArgs memberArgs(ebb_solver* s)
{
    MemberArgs M{};
    M.m = s->dm;
    M.p = s->phys;
    M.t = s->ftiles;
    const int saved = s->cur;
    for (int P = 0; P < 2; P++) {
        s->cur = P;
        const std::vector<StageDesc> seq = stepSequence(s);
        const double4* q = s->qcur();
        for (size_t k = 0; k < seq.size(); k++) {
            const StageDesc& x = seq[k];
            M.st[P][k] = loopStageArgs(s, x.kind, x.in, q, x.out, x.limit, x.dtUpdate, x.first, x.last, uint32_t(k),
                                       uint32_t(k), 0);
        }
        s->stageInStep = uint32_t(seq.size());
        M.d[P] = diagArgs(s, s->qnext(), s->qcur(), 0, s->cfg.maxSteps);
        M.d[P].histRing = 1;
    }
    s->cur = saved;
    M.parityFlip = uint32_t(saved);
    M.f = forceArgs(s);
    const uint64_t nb = uint64_t(s->b.nCells) / 16;
    M.nRmsFull = uint32_t(nb / RMS_CHUNK_BLOCKS);
    return M;
}

// Automatic tiles follow the cells running together: a member whose tiles were sized for
// less than half or more than twice the cells now sharing the launches is cut again (the
// results do not depend on the tile size).  Only members below ~400k cells ever change.
int followSpread(ebb_batch* b)
{
    if (b->tileCells != -1) return EBB_OK;
    int64_t total = 0;
    for (auto& mb : b->members) if (mb.running) total += int64_t(mb.s->N);
    for (auto& mb : b->members) {
        ebb_solver* s = mb.s;
        if (!mb.running || s->cfg.tileCells != -1 || !s->cfg.spreadTiles) continue;
        const uint32_t was = spreadTileCells(uint64_t(s->N) + uint64_t(std::max<int64_t>(s->cfg.spreadOver, 0)), s->cus);
        const uint32_t now = spreadTileCells(uint64_t(total), s->cus);
        if (now <= 2 * was && was <= 2 * now) continue;
        s->cfg.spreadOver = total - int64_t(s->N);
        int rc = retileFaceTiles(s, -1);
        if (rc) return rc;
        const int64_t nb = std::max(mb.bytes, s->deviceBytes);
        b->used += nb - mb.bytes;
        mb.bytes = nb;
    }
    return EBB_OK;
}

// Work lists and launch records of the running members.
int rebuild(ebb_batch* b)
{
    dropGraph(b);
    { int rc = followSpread(b); if (rc) return rc; }
    b->groups.clear();
    b->hmem.clear();
    b->slotOf.clear();
    b->rms.clear();
    b->fin.clear();
    b->force.clear();
    b->maxStages = 0;
    std::map<GroupKey, size_t> gidx;
    for (size_t k = 0; k < b->members.size(); k++) {
        BatchMember& mb = b->members[k];
        if (!mb.running) continue;
        ebb_solver* s = mb.s;
        const uint32_t slot = uint32_t(b->hmem.size());
        b->hmem.push_back(memberArgs(s));
        b->slotOf.push_back(uint32_t(k));
        const std::vector<StageDesc> seq = stepSequence(s);
        const GroupKey key = groupKey(s, seq);
        auto it = gidx.find(key);
        if (it == gidx.end()) {
            BatchGroup g;
            g.shape = GroupShape{int(s->S), s->phys.flux, s->phys.viscous, s->phys.order, s->phys.limited,
                                 s->phys.venkat, s->phys.mixed, s->useGradTiles ? 1 : 0};
            g.nst = int(seq.size());
            for (size_t q = 0; q < seq.size() && q < 4; q++) g.limit[q] = seq[q].limit;
            it = gidx.emplace(key, b->groups.size()).first;
            b->groups.push_back(std::move(g));
        }
        BatchGroup& g = b->groups[it->second];
        b->maxStages = std::max(b->maxStages, g.nst);
        g.lds = std::max(g.lds, s->ftiles.ldsBytes);
        if (deferForce(s)) {
            g.grad.push_back(make_uint2(slot, FORCE_ITEM));
            b->force.push_back(slot);
        }
        // gradient items: cell blocks, or on gradient tiles the border blocks (at least one,
        // whose block 0 resets the step's Rmax)
        const uint32_t nblk = s->useGradTiles ? std::max<uint32_t>(1u, (s->ftiles.nBorder + 255) / 256) : (s->N + 255) / 256;
        for (uint32_t x = 0; x < nblk; x++) g.grad.push_back(make_uint2(slot, x));
        for (uint32_t x = 0; x < s->ftiles.nTiles; x++) g.tile.push_back(make_uint2(slot, x));
        for (uint32_t x = 0; x < b->hmem.back().nRmsFull + 3; x++) b->rms.push_back(make_uint2(slot, x));
        b->fin.push_back(slot);
    }
    HIPCHK(hipSetDevice(b->device));
    for (auto& g : b->groups) {
        HIPCHK(uploadVec(g.dGrad, g.grad));
        HIPCHK(uploadVec(g.dTile, g.tile));
    }
    HIPCHK(uploadVec(b->dmem, b->hmem));
    HIPCHK(uploadVec(b->dRms, b->rms));
    HIPCHK(uploadVec(b->dFin, b->fin));
    HIPCHK(uploadVec(b->dForce, b->force));
    if (b->hScalCap < b->fin.size()) {
        if (b->hScal) (void)hipHostFree(b->hScal);
        b->hScal = nullptr;
        b->hScalCap = std::max<size_t>(b->fin.size(), 16);
        HIPCHK(hipHostMalloc((void**)&b->hScal, b->hScalCap * sizeof(DevScalars), hipHostMallocDefault));
        b->dScal.release();
        HIPCHK(b->dScal.alloc(b->hScalCap));
    }
    b->dirty = false;
    return EBB_OK;
}

// one time-loop step of every running member: parity 0 reads each member's current state
void launchStep(ebb_batch* b, uint32_t parity)
{
    for (int st = 0; st < b->maxStages; st++) {
        for (auto& g : b->groups)
            if (st < g.nst)
                launchGradientBatch(g.shape, g.limit[st], b->dmem.p, g.dGrad.p, uint32_t(g.grad.size()), parity,
                                    uint32_t(st), b->stream);
        for (auto& g : b->groups)
            if (st < g.nst)
                launchFaceTilesBatch(g.shape, g.limit[st], b->dmem.p, g.dTile.p, uint32_t(g.tile.size()), g.lds,
                                     parity, uint32_t(st), b->stream);
    }
    launchRmsBatch(b->dmem.p, b->dRms.p, uint32_t(b->rms.size()), parity, b->stream);
    launchFinalizeBatch(b->dmem.p, b->dFin.p, uint32_t(b->fin.size()), parity, b->stream);
}

int captureGraph(ebb_batch* b)
{
    dropGraph(b);
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < b->graphSteps; k++) launchStep(b, uint32_t(k & 1));
    HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return EBB_OK;
}

// The error a stopped member recorded, as the solo run reports it (solver.cpp checkErrors).
void decodeError(BatchMember& mb, const DevScalars& h)
{
    ebb_solver* s = mb.s;
    if (h.errBC != ~0ull && (h.errEdge == ~0ull || (h.errBC >> 32) <= (h.errEdge >> 32))) {
        mb.errKind = EBB_ERR_BC;
        mb.errIndex = int64_t(s->b.boundaryEdges[h.errBC & 0xFFFFFFFFull]);
    } else if (h.errEdge != ~0ull) {
        const uint64_t key = h.errEdge & 0xFFFFFFFFull;
        mb.errKind = EBB_ERR_EDGE_NAN;
        mb.errIndex = key < s->Fb ? int64_t(s->b.boundaryEdges[key]) : int64_t(s->b.interiorEdges[key - s->Fb]);
    } else if (h.errCell != ~0ull) {
        mb.errKind = EBB_ERR_CELL_NAN;
        mb.errIndex = int64_t(h.errCell & 0xFFFFFFFFull);
    }
}

// Advance every running member by `steps` steps (fewer where a member stops); then read
// back each member's scalars (and history rows).
int runSteps(ebb_batch* b, int64_t steps)
{
    if (b->dirty) { int rc = rebuild(b); if (rc) return rc; }
    const size_t n = b->fin.size();
    if (n == 0 || steps <= 0) return EBB_OK;
    std::vector<int64_t> before(n);
    for (size_t k = 0; k < n; k++) before[k] = b->members[b->slotOf[k]].s->stepsTotal;
    int64_t done = 0;
    const int G = b->graphSteps;
    if (G >= 2) {
        if (!b->exec) { int rc = captureGraph(b); if (rc) return rc; }
        for (; done + G <= steps; done += G) HIPCHK(hipGraphLaunch(b->exec, b->stream));
    }
    for (; done < steps; done++) launchStep(b, uint32_t(done & 1));
    launchPendingForceBatch(b->dmem.p, b->dForce.p, uint32_t(b->force.size()), b->stream);
    launchGatherScalars(b->dmem.p, b->dFin.p, uint32_t(n), b->dScal.p, b->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->hScal, b->dScal.p, n * sizeof(DevScalars), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    bool flips = false;
    for (size_t k = 0; k < n; k++) {
        BatchMember& mb = b->members[b->slotOf[k]];
        ebb_solver* s = mb.s;
        const DevScalars& h = b->hScal[k];
        *s->h_sc = h;
        const int64_t ran = h.steps - before[k];
        if (b->keepHistory && ran > 0) {
            // the ring holds the last histCap rows; runs are cut to histCap steps when history is kept
            const size_t old = s->history.size();
            s->history.resize(old + size_t(ran) * HIST_FIELDS);
            for (int64_t st = before[k]; st < h.steps; st++) {
                const int64_t r = st % s->histCap;
                HIPCHK(hipMemcpyAsync(s->history.data() + old + size_t(st - before[k]) * HIST_FIELDS,
                                      s->d_hist.p + r * HIST_FIELDS, HIST_FIELDS * sizeof(double), hipMemcpyDeviceToHost,
                                      b->stream));
            }
        }
        s->stepsTotal = h.steps;
        if (ran & 1) { s->cur ^= 1; flips = true; }
        b->hmem[k].parityFlip = uint32_t(s->cur);
        if (!h.active) {
            mb.running = false;
            decodeError(mb, h);
            b->dirty = true;
        }
    }
    if (b->keepHistory) HIPCHK(hipStreamSynchronize(b->stream));
    if (flips && !b->dirty) {
        HIPCHK(hipMemcpyAsync(b->dmem.p, b->hmem.data(), b->hmem.size() * sizeof(MemberArgs), hipMemcpyHostToDevice,
                              b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return EBB_OK;
}

// Snapshot of the running members (current state and scalars) for the tuner, which runs
// steps it must not keep.
struct Snapshot {
    std::vector<DBuf<double4>> q;
    std::vector<DevScalars> sc;
    std::vector<int> cur;
    std::vector<int64_t> steps;
    std::vector<size_t> hist;
};

int takeSnapshot(ebb_batch* b, Snapshot& sn)
{
    for (auto& mb : b->members) {
        ebb_solver* s = mb.s;
        sn.q.emplace_back();
        if (mb.running) {
            HIPCHK(sn.q.back().alloc(s->b.nCells));
            HIPCHK(hipMemcpyAsync(sn.q.back().p, s->qcur(), size_t(s->b.nCells) * sizeof(double4),
                                  hipMemcpyDeviceToDevice, b->stream));
        }
        HIPCHK(hipMemcpyAsync(s->h_sc, s->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        sn.sc.push_back(*s->h_sc);
        sn.cur.push_back(s->cur);
        sn.steps.push_back(s->stepsTotal);
        sn.hist.push_back(s->history.size());
    }
    return EBB_OK;
}

// puts every member back as the snapshot found it (the snapshot stays for another restore)
int restoreSnapshot(ebb_batch* b, const Snapshot& sn)
{
    for (size_t k = 0; k < b->members.size(); k++) {
        BatchMember& mb = b->members[k];
        ebb_solver* s = mb.s;
        s->cur = sn.cur[k];
        s->stepsTotal = sn.steps[k];
        s->history.resize(sn.hist[k]);
        *s->h_sc = sn.sc[k];
        HIPCHK(hipMemcpyAsync(s->d_sc, s->h_sc, sizeof(DevScalars), hipMemcpyHostToDevice, b->stream));
        if (sn.q[k].p) {
            HIPCHK(hipMemcpyAsync(s->qcur(), sn.q[k].p, size_t(s->b.nCells) * sizeof(double4), hipMemcpyDeviceToDevice,
                                  b->stream));
            mb.running = true;
            mb.errKind = 0;
            mb.errIndex = -1;
        }
        HIPCHK(hipStreamSynchronize(b->stream));
        refreshShadow(s);
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    b->dirty = true;
    return EBB_OK;
}

void releaseSnapshot(Snapshot& sn)
{
    for (auto& q : sn.q) q.release();
    sn = Snapshot{};
}

}  // namespace

extern "C" {

int ebb_estimate_device_bytes(const ebb_mesh* m, const ebb_config* c, int64_t* bytes)
{
    return guarded([&]() -> int {
        if (!m || !c || !bytes) return EBB_ERR_ARG;
        *bytes = estimateDeviceBytes(m->b, c->c);
        return EBB_OK;
    });
}

int ebb_batch_create(int device, int64_t device_bytes_limit, int graph_steps, int keep_history, ebb_batch** out)
{
    return guarded([&]() -> int {
        // a graph may not outrun a member's history ring (runs are cut to whole graphs of it)
        if (!out || device_bytes_limit <= 0 || graph_steps < 0 || (graph_steps & 1) || graph_steps > HIST_RING_ROWS)
            return EBB_ERR_ARG;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return EBB_ERR_DEVICE;
        auto b = std::make_unique<ebb_batch>();
        b->device = device;
        b->limit = device_bytes_limit;
        b->graphSteps = graph_steps;
        b->keepHistory = keep_history != 0;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        *out = b.release();
        return EBB_OK;
    });
}

int ebb_batch_add(ebb_batch* b, const ebb_mesh* m, const ebb_config* c, int64_t* id)
{
    return guarded([&]() -> int {
        if (!b || !m || !c || !id) return EBB_ERR_ARG;
        Config cfg = c->c;
        // batched members run a face-tile stage path (gradient tiles where asked for, else
        // the gradient kernel and face tiles) in the reference numbering with 16-cell-block
        // tiles (the RMS leaf sums come from the last stage), no LP limiter
        if (cfg.order > 1 && cfg.limited && cfg.lpThresh > 0) return EBB_ERR_ARG;
        if (cfg.stagePath == 2) return EBB_ERR_ARG;
        cfg.stagePath = cfg.stagePath == 3 ? 3 : 1;
        cfg.tileOrder = 2;
        cfg.renumber = false;
        // automatic tiles are sized for the cells of the running members and this one
        // together: a member admitted alone or beside small ones spreads over the GPU as a
        // solo run does; one joining members that already fill it takes full-size tiles
        cfg.spreadTiles = true;
        cfg.spreadOver = 0;
        for (auto& mb : b->members) if (mb.running) cfg.spreadOver += int64_t(mb.s->N);
        if (b->tileCells != -1) cfg.tileCells = b->tileCells;
        const int64_t need = estimateDeviceBytes(m->b, cfg);
        if (b->used + need > b->limit) return EBB_ERR_ALLOC;
        ebb_solver* s = nullptr;
        int rc = createSolver(m, cfg, b->device, &s);
        if (rc) return rc;
        if (!s->useFaceTiles || !s->rmsFused) {
            ebb_solver_destroy(s);
            return EBB_ERR_MESH;   // tiles do not fit a workgroup's LDS
        }
        BatchMember mb;
        mb.id = b->nextId++;
        mb.s = s;
        mb.bytes = std::max(need, s->deviceBytes);
        b->used += mb.bytes;
        b->members.push_back(mb);
        b->dirty = true;
        *id = mb.id;
        return EBB_OK;
    });
}

int ebb_batch_remove(ebb_batch* b, int64_t id)
{
    return guarded([&]() -> int {
        if (!b) return EBB_ERR_ARG;
        for (size_t k = 0; k < b->members.size(); k++) {
            if (b->members[k].id != id) continue;
            HIPCHK(hipStreamSynchronize(b->stream));
            b->used -= b->members[k].bytes;
            ebb_solver_destroy(b->members[k].s);
            b->members.erase(b->members.begin() + long(k));
            b->dirty = true;
            return EBB_OK;
        }
        return EBB_ERR_ARG;
    });
}

int ebb_batch_run(ebb_batch* b, int64_t steps, int64_t* running)
{
    return guarded([&]() -> int {
        if (!b || steps < 0) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(b->device));
        int64_t left = steps;
        while (left > 0) {
            // with history kept, a run call is cut so no member overruns its history ring
            int64_t chunk = left;
            if (b->keepHistory) {
                int64_t cap = 1 << 30;
                for (auto& mb : b->members) if (mb.running) cap = std::min<int64_t>(cap, mb.s->histCap);
                chunk = std::min(chunk, cap - (cap % std::max(2, b->graphSteps)));
            }
            int rc = runSteps(b, chunk);
            if (rc) return rc;
            left -= chunk;
            bool any = false;
            for (auto& mb : b->members) any = any || mb.running;
            if (!any) break;
        }
        if (running) {
            int64_t r = 0;
            for (auto& mb : b->members) r += mb.running ? 1 : 0;
            *running = r;
        }
        return EBB_OK;
    });
}

int ebb_batch_status(ebb_batch* b, int64_t id, double out[10])
{
    return guarded([&]() -> int {
        if (!b || !out) return EBB_ERR_ARG;
        BatchMember* mb = b->find(id);
        if (!mb) return EBB_ERR_ARG;
        const DevScalars& h = *mb->s->h_sc;
        out[0] = h.t; out[1] = h.dt; out[2] = double(mb->s->stepsTotal); out[3] = mb->running ? 1.0 : 0.0;
        out[4] = h.converged ? 1.0 : 0.0; out[5] = h.rms0; out[6] = h.rms[0];
        out[7] = mb->errKind; out[8] = double(mb->errIndex); out[9] = double(mb->bytes);
        return EBB_OK;
    });
}

int ebb_batch_state_download(ebb_batch* b, int64_t id, double* q)
{
    return guarded([&]() -> int {
        if (!b || !q) return EBB_ERR_ARG;
        BatchMember* mb = b->find(id);
        if (!mb) return EBB_ERR_ARG;
        HIPCHK(hipStreamSynchronize(b->stream));
        return ebb_state_download(mb->s, q);
    });
}

int64_t ebb_batch_history_length(ebb_batch* b, int64_t id)
{
    if (!b) return 0;
    BatchMember* mb = b->find(id);
    return mb ? int64_t(mb->s->history.size() / HIST_FIELDS) : 0;
}

int ebb_batch_history(ebb_batch* b, int64_t id, int64_t first, int64_t count, double* out)
{
    return guarded([&]() -> int {
        if (!b) return EBB_ERR_ARG;
        BatchMember* mb = b->find(id);
        if (!mb) return EBB_ERR_ARG;
        return ebb_history(mb->s, first, count, out);
    });
}

int ebb_batch_info(ebb_batch* b, int64_t info[6])
{
    if (!b || !info) return EBB_ERR_ARG;
    int64_t r = 0;
    for (auto& mb : b->members) r += mb.running ? 1 : 0;
    info[0] = int64_t(b->members.size());
    info[1] = r;
    info[2] = b->used;
    info[3] = b->limit;
    info[4] = b->tileCells;
    info[5] = b->graphSteps;
    return EBB_OK;
}

int ebb_batch_tune(ebb_batch* b, int steps, double out[4])
{
    return guarded([&]() -> int {
        if (!b || !out || steps <= 0) return EBB_ERR_ARG;
        HIPCHK(hipSetDevice(b->device));
        bool any = false;
        for (auto& mb : b->members) any = any || mb.running;
        if (!any) return EBB_ERR_STATE;
        Snapshot sn;
        { int rc = takeSnapshot(b, sn); if (rc) { releaseSnapshot(sn); return rc; } }
        const bool keep = b->keepHistory;
        b->keepHistory = false;
        const long long tiles[] = {-1, 512, 384, 256};
        const int graphs[] = {0, 2, 8, 32};
        double best = 1e300, base = 0;
        long long bestT = b->tileCells;
        int bestG = b->graphSteps;
        int rc = EBB_OK;
        for (long long T : tiles) {
            for (auto& mb : b->members)
                if (!rc) rc = retileFaceTiles(mb.s, T);
            b->tileCells = T;
            for (int G : graphs) {
                if (rc) break;
                b->graphSteps = G;
                if ((rc = restoreSnapshot(b, sn))) break;
                if ((rc = runSteps(b, std::max(2, G)))) break;   // warm-up, graph capture
                const int64_t n = std::max<int64_t>(steps, 2 * int64_t(std::max(2, G)));
                const auto t0 = std::chrono::steady_clock::now();
                if ((rc = runSteps(b, n))) break;
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / double(n);
                if (T == -1 && G == 8) base = ms;
                if (ms < best) { best = ms; bestT = T; bestG = G; }
            }
        }
        for (auto& mb : b->members)
            if (!rc) rc = retileFaceTiles(mb.s, bestT);
        b->tileCells = bestT;
        b->graphSteps = bestG;
        b->keepHistory = keep;
        if (!rc) rc = restoreSnapshot(b, sn);
        releaseSnapshot(sn);
        if (rc) return rc;
        out[0] = double(bestT); out[1] = bestG; out[2] = best; out[3] = base;
        return EBB_OK;
    });
}

int64_t ebb_batch_device_bytes(const ebb_batch* b) { return b ? b->used : 0; }

void ebb_batch_destroy(ebb_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    dropGraph(b);
    for (auto& mb : b->members) ebb_solver_destroy(mb.s);
    for (auto& g : b->groups) { g.dGrad.release(); g.dTile.release(); }
    b->dmem.release(); b->dRms.release(); b->dFin.release(); b->dForce.release(); b->dScal.release();
    if (b->hScal) (void)hipHostFree(b->hScal);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"

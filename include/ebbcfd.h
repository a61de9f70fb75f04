/*
 * ebbcfd.h -- C ABI of the MI355X explicit residual-and-update solver.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (EBB_OK = 0) instead of the reference's exceptions.  Every call on one
 * solver handle is serialised by the caller; several handles may coexist.
 * Cell arrays use the reference's numbering (elements in file order, ghosts
 * appended in boundary-edge order); states are 4 doubles per cell
 * (rho, rho u, rho v, rho E).  Each entry point names the reference
 * interface it replaces.
 */
#ifndef EBBCFD_H
#define EBBCFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ebb_mesh ebb_mesh;
typedef struct ebb_config ebb_config;
typedef struct ebb_solver ebb_solver;

enum {
    EBB_OK = 0,
    EBB_ERR_ARG = 1,        /* bad argument / handle */
    EBB_ERR_IO = 2,         /* file could not be read or written */
    EBB_ERR_PARSE = 3,      /* malformed mesh or config */
    EBB_ERR_MESH = 4,       /* mesh topology/geometry rejected (e.g. negative area) */
    EBB_ERR_BC = 5,         /* missing or inconsistent boundary condition, or a tpInflow boundary
                               with no non-negative Mach root (source/solve.d:661-670) */
    EBB_ERR_DEVICE = 6,     /* HIP runtime error or no device */
    EBB_ERR_EDGE_NAN = 7,   /* NaN flux on an edge (EdgeException, source/solve.d:806) */
    EBB_ERR_CELL_NAN = 8,   /* NaN state in a cell (CellException, app/solver.d:148) */
    EBB_ERR_STATE = 9,      /* call not valid in the solver's current state */
    EBB_ERR_ALLOC = 10      /* host or device memory exhausted, or a request larger than the memory allowed it */
};

/* Text for a status code. */
const char* ebb_error_string(int code);

/* ---------------------------------------------------------------- mesh */

/* importMesh + buildMesh: gmsh 2.2 ASCII (source/io.d:689-990) or XFlow .gri
 * (source/io.d:992-1121) reader chosen by file name (io.d:669-687), followed
 * by the geometry precompute (source/mesh.d:217-643). */
int ebb_mesh_load(const char* path, ebb_mesh** out);

/* buildMesh from arrays (source/mesh.d:217): xy[2*n_nodes]; elements as CSR
 * (elem_offsets[n_elem+1], elem_nodes 1-based, counter-clockwise);
 * bnodes[2*n_bnodes] 0-based boundary node pairs grouped by tag;
 * tags[n_tags] names; bgroup_start[n_tags] first bnode of each tag. */
int ebb_mesh_create(uint32_t n_nodes, const double* xy, uint32_t n_elem, const uint32_t* elem_offsets,
                    const uint32_t* elem_nodes, uint32_t n_bnodes, const uint32_t* bnodes, uint32_t n_tags,
                    const char* const* tags, const uint64_t* bgroup_start, ebb_mesh** out);

/* info[0..7]: cells (incl. ghosts), interior cells, ghost cells, edges,
 * interior edges, boundary edges, max edges per cell, boundary tags. */
int ebb_mesh_info(const ebb_mesh* m, int64_t info[8]);

/* UCell2 geometry (source/mesh.d:73-94); every output may be NULL.
 * area/d/perim/centroid over all cells (centroid 2 per cell); per interior
 * cell 6 slots each of edges, flux multipliers, neighbours and 12 gradMat
 * entries (2 x 6, row-major); n_edges per interior cell. */
int ebb_mesh_cell_geometry(const ebb_mesh* m, double* area, double* d, double* perim, double* centroid,
                           double* grad_mat, uint32_t* edges, int8_t* flux_mult, uint32_t* neighbors,
                           uint8_t* n_edges);

/* Edge geometry (source/mesh.d:38-65): node pair, cell pair (owner first),
 * length, unit normal, midpoint, boundary group (-1 for interior edges). */
int ebb_mesh_edge_geometry(const ebb_mesh* m, uint32_t* nodes, uint32_t* cells, double* len, double* normal,
                           double* mid, int32_t* bgroup);

/* Tiles the fused stage kernel would use for this mesh (no equivalent in the
 * reference; host only, needs no GPU).  tile_cells > 0; order 0: greedily
 * grown regions, 1: Morton order of cells, 2: Morton order of 16-cell blocks.  out[0] tiles, out[1] mean own cells per tile,
 * out[2..5] cells with gradients (own + ring 1), cells loaded (own + 2 rings),
 * local cells incl. ghosts and faces, each summed over tiles and divided by the
 * interior cell count; out[6] largest local cell count; out[7] LDS bytes per
 * workgroup (scalar limiter layout). */
int ebb_mesh_tile_stats(const ebb_mesh* m, int tile_cells, int order, double out[8]);

/* The display mesh ebb-solve writes at start-up as <meshbase>_<rank>.mmsh
 * (app/solver.d:364-372): triangulate (source/mesh.d:678-715), buildMesh on the triangles,
 * saveMatlabMesh (source/io.d:534-643; big-endian, indices 1-based as doubles). */
int ebb_mesh_save_mmsh(const ebb_mesh* m, const char* path);

void ebb_mesh_destroy(ebb_mesh* m);

/* ---------------------------------------------------------------- config */

/* loadConfig (source/config.d:112-313): the reference JSON keys and defaults
 * ("boudaryConditions", "dt" as a JSON float, ...).  Extra options live in an
 * optional "extensions" object: {"mms": bool, "maxSteps": int,
 * "precision": "fp64"|"mixed", "stagePath": "faceTiles" (default: gradient
 * kernel + one face/update kernel over tiles of cells) | "split" (gradient,
 * face and update kernels) | "fused" (one kernel per stage over tiles with
 * halos), "tileCells": int (cells per tile, -1 automatic), "tileOrder":
 * "blocks" (default: Morton order of aligned 16-cell blocks) | "morton" |
 * "grow", "cellOrder": "reference" (default) | "tiles" (renumber the device
 * cells tile by tile; every interface still uses the reference numbering)}. */
int ebb_config_parse(const char* json, ebb_config** out);

/* Read / override a numeric field by its JSON key: "dt", "tEnd", "CFL",
 * "lpThresh", "saveIter", "plotIter", "order", "maxSteps", "limited",
 * "dynamicDt", "localTimestep", "multistageLimiting", "cflAdjust",
 * "viscosity", "mms", "tileCells", "Pr", "R", "Re", "L", "ic0".."ic3". */
int ebb_config_get(const ebb_config* c, const char* key, double* value);
int ebb_config_set(ebb_config* c, const char* key, double value);
/* Mesh file named by the config ("mesh"), copied into buf. */
int ebb_config_mesh_file(const ebb_config* c, char* buf, size_t len);
void ebb_config_destroy(ebb_config* c);

/* ---------------------------------------------------------------- solver */

/* ufvmSetup (source/solve.d:175-279): binds boundary conditions, sets the
 * freestream (or manufactured-solution) initial state, and uploads the mesh
 * to HIP device `device`.  The solver keeps its own copy of mesh and config. */
int ebb_solver_create(const ebb_mesh* m, const ebb_config* c, int device, ebb_solver** out);

/* Page-locked host memory for the arrays a caller hands to ebb_state_upload,
 * ebb_state_download and ebb_residual: copies from and to it run at the PCIe DMA rate
 * instead of through the runtime's staging of pageable memory.  No counterpart in the
 * reference (a D caller that keeps its integrator allocates its q and R arrays with
 * these instead of the GC).  *out is NULL on failure: EBB_ERR_DEVICE without a HIP
 * device, EBB_ERR_ALLOC when the memory cannot be had. */
int ebb_host_alloc(size_t bytes, void** out);
void ebb_host_free(void* p);

/* Interior states in reference cell order, n_interior x 4. */
int ebb_state_upload(ebb_solver* s, const double* q);
int ebb_state_download(ebb_solver* s, double* q);
/* All cells (interior + ghost slots of the current state buffer), n_cells x 4. */
int ebb_state_download_all(ebb_solver* s, double* q);

/* ufvmSolver (source/solve.d:322-1122) on the current state: writes the ghost
 * states, R (n_interior x 4, already scaled by -1/area, MMS source added when
 * configured), new_dt (min CFL*d/sAve) and max-accumulates *rmax. */
int ebb_residual(ebb_solver* s, int limit, int dt_update, double* R, double* new_dt, double* rmax);

/* ufvmSolver as ebb_residual computes it, for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid it); it has no
 * counterpart in the reference.  It allocates nothing: R_dev is a device buffer of
 * n_interior x 4 doubles the caller owns (reference cell order; solvers that renumber their
 * cells are rejected), no edge states are stored, and the update kernel's per-workgroup
 * partials live in a fixed device array of 65,536 pairs, so meshes above 16,777,216
 * interior cells are rejected (EBB_ERR_ALLOC).  That array is shared by every call on the
 * GPU: calls that overlap on one GPU race.  With new_dt and rmax both NULL the call only
 * queues work on the solver's stream (no synchronisation; results stay in the solver's
 * device scalars); otherwise it waits and returns them as ebb_residual does.  The first
 * call on a device looks the array's address up: make it before any capture. */
int ebb_residual_noalloc(ebb_solver* s, int limit, int dt_update, double* R_dev, double* new_dt, double* rmax);

/* Integrator.step (source/integrators.d:22-270): one step with *dt; on return
 * *dt is the last stage's new time step and *rmax the step's max |R|. */
int ebb_step(ebb_solver* s, double* dt, double* rmax);

/* runIntegrator (app/solver.d:38-277) on the device: steps while t < tEnd,
 * at most max_steps (< 0: no limit) more steps, keeping t, dt, CFL and the
 * RMS/cflAdjust state between calls.  Per-step diagnostics are appended to the
 * history (ebb_history).  Returns EBB_ERR_EDGE_NAN/EBB_ERR_CELL_NAN when the
 * reference would have thrown (details in ebb_last_error). */
int ebb_run(ebb_solver* s, int64_t max_steps, double* t, double* dt, int64_t* steps_done);

/* Time-loop state: t, dt, CFL, steps done (total over all ebb_run calls). */
int ebb_run_state(ebb_solver* s, double* t, double* dt, double* cfl, int64_t* steps);
/* Reset the time loop to (t, dt) with the configured CFL (as after ufvmSetup). */
int ebb_run_reset(ebb_solver* s, double t, double dt);

/* History rows of steps [first, first+count): t, dt, rho/u/v/E RMS, Flux_R
 * (Rmax), drag, lift, CFL -- the printed line of app/solver.d:220-230 and the
 * .frc record of app/solver.d:178-193. */
int ebb_history(ebb_solver* s, int64_t first, int64_t count, double* out);
int64_t ebb_history_length(ebb_solver* s);

/* Edge data of the last residual evaluation (Edge.q[2], flux, sMax): eq is
 * n_edges x 2 x 4, flux n_edges x 4, smax n_edges; any output may be NULL. */
int ebb_edge_states(ebb_solver* s, double* eq, double* flux, double* smax);

/* Per interior cell gradient (4 x 2) and limiter (4 x 2) of the last residual evaluation. */
int ebb_cell_gradients(ebb_solver* s, double* grad, double* lim);

/* saveSolution / loadSolution (source/io.d:68-161, 218-366): .esln v2,
 * big-endian, edge states of the last residual evaluation.  Loading restores
 * the interior state and the time loop's t and dt. */
int ebb_save_solution(ebb_solver* s, const char* path);
/* The same with an explicit header time and time step: the reference's periodic
 * saves run before t += dt (app/solver.d:232-247), its final save after. */
int ebb_save_solution_as(ebb_solver* s, const char* path, double t, double dt);
int ebb_load_solution(ebb_solver* s, const char* path);

/* Append the boundary-force records (t, drag, lift; big-endian f64) of history
 * rows [first, first+count) to a .frc file (app/solver.d:178-193, 252-260). */
int ebb_write_forces(ebb_solver* s, const char* path, int64_t first, int64_t count, int append);

/* Error details: kind (EBB_ERR_*), reference edge or cell index, step. */
int ebb_last_error(ebb_solver* s, int* kind, int64_t* index, int64_t* step, char* msg, size_t len);

/* Device timing of `steps` integrator steps with HIP events on the solver's
 * stream.  out[0..2]: average ms per stage of the gradient, face and update
 * kernels; out[3]: average ms per stage of the tile kernel (face/update or
 * fused, with its stage reduction); kernels the path does not use read 0.
 * out[4]: ms per step of the stage kernels; out[5]: stages per step. */
int ebb_time_steps(ebb_solver* s, int steps, double out[6]);

/* Device bytes held by the solver. */
int64_t ebb_solver_device_bytes(const ebb_solver* s);

void ebb_solver_destroy(ebb_solver* s);

/* The 1-D slope limiters of source/limiters.d:7-37 (kind 0 firstOrderUpwindS,
 * 1 laxWendroffS, 2 minmodS, 3 harmonicS, 4 geometricS, 5 superbeeS) applied on HIP device
 * `device` to n pairs: out[i] = limiter(sp[i], sm[i]).  Host arrays; ebb-solve never calls
 * these limiters (the solver uses the scalar and LP limiters of source/solve.d). */
int ebb_limiter1d(int device, int kind, int64_t n, const double* sp, const double* sm, double* out);

/* ---------------------------------------------------------------- batched solves
 * No counterpart in the reference, whose integrator scratch is process-static and sized by
 * the first mesh (source/integrators.d:56-75), so one process holds one mesh.  A batch
 * holds many solvers (each the result of ufvmSetup on its own mesh and config, state
 * resident on the device) and advances all of them with one launch per kernel: per RK
 * stage one gradient launch and one face/update launch over every mesh's blocks and tiles,
 * then one RMS launch and one epilogue launch, the whole step sequence captured as a HIP
 * graph.  Each mesh's states and history equal its solo ebb_run's bit for bit.  Members
 * use the face-tile stage path (or the gradient tiles, extensions.stagePath = "gradTiles")
 * in the reference cell numbering; the LP vector limiter and the fused stage path are not
 * batched. */
typedef struct ebb_batch ebb_batch;

/* Device bytes ebb_solver_create would allocate for this mesh and config (an estimate
 * within a few percent; admission by footprint uses it). */
int ebb_estimate_device_bytes(const ebb_mesh* m, const ebb_config* c, int64_t* bytes);

/* A batch on HIP device `device` admitting solvers while their device bytes stay within
 * device_bytes_limit.  graph_steps: steps per captured HIP graph (even, at most 256; 0:
 * plain launches).  keep_history: also copy each step's history row to the host. */
int ebb_batch_create(int device, int64_t device_bytes_limit, int graph_steps, int keep_history, ebb_batch** out);

/* Admit a solve (ufvmSetup on its own device state).  EBB_ERR_ALLOC when its footprint
 * does not fit what the limit leaves; *id names it in the calls below. */
int ebb_batch_add(ebb_batch* b, const ebb_mesh* m, const ebb_config* c, int64_t* id);

/* Release a solve's device state (finished or not). */
int ebb_batch_remove(ebb_batch* b, int64_t id);

/* Advance every running solve by `steps` time-loop steps (app/solver.d:119-250); a solve
 * stops by itself at tEnd, maxSteps, the rho-RMS drop of extensions.rmsDrop or an error.
 * *running: solves still stepping afterwards. */
int ebb_batch_run(ebb_batch* b, int64_t steps, int64_t* running);

/* out[0..9]: t, dt, steps, running (1/0), converged (1/0), first-step rho RMS, last rho
 * RMS, error kind (EBB_ERR_*, 0: none), reference edge/cell of the error, device bytes. */
int ebb_batch_status(ebb_batch* b, int64_t id, double out[10]);
/* Interior states of a solve, reference cell order, n_interior x 4. */
int ebb_batch_state_download(ebb_batch* b, int64_t id, double* q);
/* History rows of a solve (as ebb_history), kept when the batch was created with keep_history. */
int64_t ebb_batch_history_length(ebb_batch* b, int64_t id);
int ebb_batch_history(ebb_batch* b, int64_t id, int64_t first, int64_t count, double* out);
/* info[0..5]: solves held, solves running, device bytes admitted, limit, cells per tile
 * (-1: automatic), graph steps. */
int ebb_batch_info(ebb_batch* b, int64_t info[6]);

/* Autotuner: times `steps` steps of the running solves for each tile size (automatic, 512,
 * 384, 256 cells) and graph length (0, 2, 8, 32 steps), keeps the fastest, and restores
 * every solve's state as it was.  out: best tile cells, best graph steps, its ms per step,
 * ms per step of the default (automatic tiles, 8-step graphs). */
int ebb_batch_tune(ebb_batch* b, int steps, double out[4]);

int64_t ebb_batch_device_bytes(const ebb_batch* b);
void ebb_batch_destroy(ebb_batch* b);

/* ---------------------------------------------------------------- serving loop
 * No counterpart in the reference (one mesh per process).  A server owns a batch and a
 * loop thread: requests (mesh + config) are queued as they arrive; each round the loop
 * admits queued requests in arrival order while their device footprint fits
 * device_bytes_limit and fewer than max_solves run, advances every running solve by
 * steps_per_round steps with the batched launches, and retires the solves that stopped
 * (extensions.rmsDrop, maxSteps, tEnd, error), freeing their memory.  Requests arriving
 * during a round join at the next round boundary.  The mesh and config of a request must
 * stay valid until the request is done.  Submission, polling and statistics may be called
 * from any thread. */
typedef struct ebb_server ebb_server;
int ebb_server_create(int device, int64_t device_bytes_limit, int max_solves, int steps_per_round, int graph_steps,
                      int keep_results, ebb_server** out);
/* Queue a request; *ticket names it.  Its arrival time is now.  EBB_ERR_STATE once the
 * loop has ended on a host error (every request not yet done then failed with
 * EBB_ERR_ALLOC). */
int ebb_server_submit(ebb_server* sv, const ebb_mesh* m, const ebb_config* c, int64_t* ticket);
/* out[0..15]: state (0 queued, 1 running, 2 done, 3 failed), error code, arrival, admission
 * and completion times (s since the server started; -1 before), interior cells, then the
 * ten ebb_batch_status fields at completion. */
int ebb_server_poll(ebb_server* sv, int64_t ticket, double out[16]);
/* Final interior states of a done request (server created with keep_results), n_interior x 4. */
int ebb_server_result(ebb_server* sv, int64_t ticket, double* q);
/* Block until at least n_done requests are done or failed (EBB_ERR_STATE on timeout). */
int ebb_server_wait(ebb_server* sv, int64_t n_done, double timeout_s);
/* out[0..11]: submitted, done, failed, running, queued, p50 and p99 time to converged (s,
 * arrival to completion, done requests), solves/s (done / span from the first arrival to
 * the last completion), device bytes in use, seconds in batched runs, rounds, server age (s). */
int ebb_server_stats(ebb_server* sv, double out[12]);
/* Stops the loop after its current round and releases every solve. */
void ebb_server_destroy(ebb_server* sv);

#ifdef __cplusplus
}
#endif

#endif /* EBBCFD_H */

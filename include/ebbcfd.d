// D binding of the C ABI in ebbcfd.h (libebbcfd.so), for D host code that keeps the
// reference's mesh.d / solve.d / integrators.d and calls the MI355X path instead of
// ufvmSolver.  Link with "libs": ["ebbcfd"] and "lflags": ["-L<dir of libebbcfd.so>"].
// Types map as uint32_t -> uint, int64_t -> long, uint64_t -> ulong, int8_t -> byte,
// uint8_t -> ubyte; handles are opaque structs.
module ebb.gpu;
extern(C) @nogc nothrow:

struct ebb_mesh; struct ebb_config; struct ebb_solver;

int ebb_mesh_create(uint nNodes, const(double)* xy, uint nElem, const(uint)* elemOffsets,
                    const(uint)* elemNodes, uint nBnodes, const(uint)* bnodes, uint nTags,
                    const(char*)* tags, const(ulong)* bgroupStart, ebb_mesh** mesh);
int ebb_mesh_load(const(char)* path, ebb_mesh** mesh);
int ebb_config_parse(const(char)* json, ebb_config** cfg);
int ebb_solver_create(const(ebb_mesh)* mesh, const(ebb_config)* cfg, int device, ebb_solver** s);
int ebb_state_upload(ebb_solver* s, const(double)* q);        // q: nInterior x 4, cell order of mesh.d
int ebb_state_download(ebb_solver* s, double* q);
int ebb_residual(ebb_solver* s, int limit, int dtUpdate, double* R, double* newDt, double* Rmax);
int ebb_step(ebb_solver* s, double* dt, double* Rmax);
int ebb_run(ebb_solver* s, long maxSteps, double* t, double* dt, long* stepsDone);
int ebb_history(ebb_solver* s, long first, long count, double* rows);   // 10 doubles per step
int ebb_save_solution(ebb_solver* s, const(char)* path);
int ebb_load_solution(ebb_solver* s, const(char)* path);
int ebb_last_error(ebb_solver* s, int* kind, long* index, long* step, char* msg, size_t len);
void ebb_solver_destroy(ebb_solver* s);
void ebb_mesh_destroy(ebb_mesh* m);
void ebb_config_destroy(ebb_config* c);

// the rest of include/ebbcfd.h: queries, diagnostics and file output
const(char)* ebb_error_string(int code);
int ebb_mesh_info(const(ebb_mesh)* m, long* info);                          // 8 entries
int ebb_mesh_cell_geometry(const(ebb_mesh)* m, double* area, double* d, double* perim, double* centroid,
                           double* gradMat, uint* edges, byte* fluxMult, uint* neighbors, ubyte* nEdges);
int ebb_mesh_edge_geometry(const(ebb_mesh)* m, uint* nodes, uint* cells, double* len, double* normal,
                           double* mid, int* bgroup);
int ebb_mesh_tile_stats(const(ebb_mesh)* m, int tileCells, int order, double* out8);
int ebb_mesh_save_mmsh(const(ebb_mesh)* m, const(char)* path);          // <meshbase>_<rank>.mmsh
int ebb_config_get(const(ebb_config)* c, const(char)* key, double* value);
int ebb_config_set(ebb_config* c, const(char)* key, double value);
int ebb_config_mesh_file(const(ebb_config)* c, char* buf, size_t len);
int ebb_state_download_all(ebb_solver* s, double* q);                      // interior + ghosts
int ebb_run_state(ebb_solver* s, double* t, double* dt, double* cfl, long* steps);
int ebb_run_reset(ebb_solver* s, double t, double dt);
long ebb_history_length(ebb_solver* s);
int ebb_edge_states(ebb_solver* s, double* eq, double* flux, double* smax); // Edge.q[2], flux, sMax
int ebb_cell_gradients(ebb_solver* s, double* grad, double* lim);
int ebb_residual_noalloc(ebb_solver* s, int limit, int dtUpdate, double* RDev, double* newDt, double* Rmax);
int ebb_save_solution_as(ebb_solver* s, const(char)* path, double t, double dt);
int ebb_write_forces(ebb_solver* s, const(char)* path, long first, long count, int append);
int ebb_time_steps(ebb_solver* s, int steps, double* out6);
long ebb_solver_device_bytes(const(ebb_solver)* s);
int ebb_host_alloc(size_t bytes, void** p);                                // page-locked q / R arrays
void ebb_host_free(void* p);
int ebb_limiter1d(int device, int kind, long n, const(double)* sp, const(double)* sm, double* result);

// batched solves (many meshes per launch) and the serving loop
struct ebb_batch; struct ebb_server;
int ebb_estimate_device_bytes(const(ebb_mesh)* m, const(ebb_config)* c, long* bytes);
int ebb_batch_create(int device, long deviceBytesLimit, int graphSteps, int keepHistory, ebb_batch** b);
int ebb_batch_add(ebb_batch* b, const(ebb_mesh)* m, const(ebb_config)* c, long* id);
int ebb_batch_remove(ebb_batch* b, long id);
int ebb_batch_run(ebb_batch* b, long steps, long* running);
int ebb_batch_status(ebb_batch* b, long id, double* out10);
int ebb_batch_state_download(ebb_batch* b, long id, double* q);
long ebb_batch_history_length(ebb_batch* b, long id);
int ebb_batch_history(ebb_batch* b, long id, long first, long count, double* rows);
int ebb_batch_info(ebb_batch* b, long* info6);
int ebb_batch_tune(ebb_batch* b, int steps, double* out4);
long ebb_batch_device_bytes(const(ebb_batch)* b);
void ebb_batch_destroy(ebb_batch* b);
int ebb_server_create(int device, long deviceBytesLimit, int maxSolves, int stepsPerRound, int graphSteps,
                      int keepResults, ebb_server** sv);
int ebb_server_submit(ebb_server* sv, const(ebb_mesh)* m, const(ebb_config)* c, long* ticket);
int ebb_server_poll(ebb_server* sv, long ticket, double* out16);
int ebb_server_result(ebb_server* sv, long ticket, double* q);
int ebb_server_wait(ebb_server* sv, long nDone, double timeoutS);
int ebb_server_stats(ebb_server* sv, double* out12);
void ebb_server_destroy(ebb_server* sv);

// ebb-solve: command-line driver with the reference's interface (app/solver.d:430-470):
//   ebb-solve -c <config.json> [-s <save prefix>] [--device N]
// Outputs as the reference: <meshbase>_0.mmsh at start-up (app/solver.d:364-372), stdout
// lines every plotIter steps (app/solver.d:220-230),
// save_<n>_0.esln every saveIter steps, final_0.esln, boundaryForces.frc, and the
// timing summary of app/solver.d:380-390.  SIGINT/SIGUSR1 stop the loop cleanly
// and still write final_0.esln (app/solver.d:30-36).
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ebbcfd.h"

static volatile std::sig_atomic_t interrupted = 0;
static void handle(int) { std::printf("Signal received\n"); interrupted = 1; }

static int die(const char* what, int rc)
{
    std::fprintf(stderr, "%s: %s\n", what, ebb_error_string(rc));
    return 1;
}

int main(int argc, char** argv)
{
    std::string cfgFile, saveFile;
    int device = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if ((a == "-c" || a == "--config") && i + 1 < argc) cfgFile = argv[++i];
        else if ((a == "-s" || a == "--save") && i + 1 < argc) saveFile = argv[++i];
        else if (a == "--device" && i + 1 < argc) device = std::atoi(argv[++i]);
    }
    if (cfgFile.empty()) {
        std::fprintf(stderr, "usage: ebb-solve -c config.json [-s save_prefix]\n");
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::signal(SIGINT, handle);
    std::signal(SIGUSR1, handle);
    std::ifstream f(cfgFile);
    if (!f) { std::fprintf(stderr, "cannot read %s\n", cfgFile.c_str()); return 1; }
    std::stringstream ss;
    ss << f.rdbuf();
    ebb_config* cfg = nullptr;
    int rc = ebb_config_parse(ss.str().c_str(), &cfg);
    if (rc) return die("config", rc);
    char meshFile[4096];
    ebb_config_mesh_file(cfg, meshFile, sizeof(meshFile));
    if (!saveFile.empty()) saveFile += "0.esln";
    ebb_mesh* mesh = nullptr;
    if ((rc = ebb_mesh_load(meshFile, &mesh))) return die("mesh", rc);
    {
        // the display mesh, named as config.meshFile.split('.')[0] ~ "_<rank>.mmsh" (app/solver.d:372)
        const std::string mf(meshFile);
        const std::string mmsh = mf.substr(0, mf.find('.')) + "_0.mmsh";
        if ((rc = ebb_mesh_save_mmsh(mesh, mmsh.c_str()))) return die("mmsh", rc);
    }
    ebb_solver* s = nullptr;
    if ((rc = ebb_solver_create(mesh, cfg, device, &s))) return die("solver", rc);
    if (!saveFile.empty() && (rc = ebb_load_solution(s, saveFile.c_str()))) return die("Failed to load solution", rc);
    double limited, order, plotIter, saveIter, integ, flux;
    ebb_config_get(cfg, "limited", &limited);
    ebb_config_get(cfg, "order", &order);
    ebb_config_get(cfg, "plotIter", &plotIter);
    ebb_config_get(cfg, "saveIter", &saveIter);
    ebb_config_get(cfg, "integrator", &integ);
    ebb_config_get(cfg, "flux", &flux);
    static const char* integNames[] = {"Euler", "RK2", "RK2_TVD", "RK4"};
    std::printf("Running 2D finite volume solver\n-limited: %s\n-order: %d\n-flux: %s\n-integrator: %s\n",
                limited != 0 ? "true" : "false", int(order), flux == 1 ? "roeFlux" : "rusanovFlux", integNames[int(integ)]);
    std::remove("boundaryForces.frc");
    FILE* touch = std::fopen("boundaryForces.frc", "wb");
    if (touch) std::fclose(touch);
    const auto ts = std::chrono::steady_clock::now();
    int64_t it = 0, saveItr = 0;
    const int64_t plot = int64_t(plotIter), save = int64_t(saveIter);
    double h[10] = {0};
    int status = 0;
    while (!interrupted) {
        // run up to the next step that needs a save, or a bounded batch
        int64_t want = 256;
        if (save > 0) {
            const int64_t next = (it % save == 0) ? it : (it / save + 1) * save;
            want = next - it + 1;
        }
        double t, dt;
        int64_t ran = 0;
        rc = ebb_run(s, want, &t, &dt, &ran);
        for (int64_t k = 0; k < ran; k++) {
            ebb_history(s, it + k, 1, h);
            if (plot != -1 && plot > 0 && (it + k) % plot == 0) {
                std::printf("lift force = %f\t drag force = %f\t t = %f\n", h[8], h[7], h[0]);
                std::printf("rho_RMS = %.10e\tu_RMS = %.10e\tv_RMS = %.10e\tE_RMS = %.10e\tFlux_R = %.10e\t dt = %10.10f\n",
                            h[2], h[3], h[4], h[5], h[6], h[1]);
            }
        }
        if (ran) ebb_write_forces(s, "boundaryForces.frc", it, ran, 1);
        it += ran;
        if (rc) {
            int kind; int64_t idx, step;
            char msg[512];
            ebb_last_error(s, &kind, &idx, &step, msg, sizeof(msg));
            std::printf("Solver encountered an error: %s (index %lld)\n", msg, (long long)idx);
            status = 1;
            break;
        }
        if (save > 0 && ran > 0 && (it - 1) % save == 0) {
            char name[512];
            std::snprintf(name, sizeof(name), "save_%lld_0.esln", (long long)saveItr++);
            // saved before t += dt in the reference loop (app/solver.d:232-247)
            ebb_save_solution_as(s, name, h[0], h[1]);
        }
        if (ran < want) break;   // t >= tEnd or maxSteps reached
    }
    ebb_save_solution(s, "final_0.esln");
    std::printf("lift force = %f\t drag force = %f\t t = %f\n", h[8], h[7], h[0]);
    std::printf("rho_RMS = %.10e\tu_RMS = %.10e\tv_RMS = %.10e\tE_RMS = %.10e\tFlux_R = %.10e\t dt = %10.10f\n",
                h[2], h[3], h[4], h[5], h[6], h[1]);
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
    int64_t info[8];
    ebb_mesh_info(mesh, info);
    std::printf("total solver time: %g s\n", elapsed);
    if (it > 0) {
        const double tpi = elapsed / double(it);
        std::printf("average time per iteration: %g s\n", tpi);
        std::printf("average time per cell: %g us\n", tpi / double(info[0]) * 1e6);
    }
    std::printf("exiting\n");
    ebb_solver_destroy(s);
    ebb_mesh_destroy(mesh);
    ebb_config_destroy(cfg);
    std::printf("total run time: %g\n", std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    return status;
}

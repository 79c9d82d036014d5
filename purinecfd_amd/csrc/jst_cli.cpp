// jst-gpu — command-line driver with the flags and outputs of jst-rs's CLI
// (jst-rs/src/bin/jst.rs:10-123): --mesh --config --steps --out --reslog --quiet --threads
// --no-output, plus --precision fp64|mixed, --wall PATH (wall Cp/Cf table), --forces
// (CL and CD of the final state) and --profile N (N more steps after the outputs are
// written, timed with HIP events: ms per step and the mean stage-kernel launch).  stdout lines, res.log and result.csv use the
// same text formats (Rust `{:6}`, `{:.6e}`, `{:.8e}`), so scripts that read the CPU
// solver's output read this one's.  Built on the public C ABI only (include/jst_gpu.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>

#include "../../include/jst_gpu.h"

namespace {

// Rust's LowerExp with a fixed precision: mantissa as printf, exponent without padding or
// '+' ("1.234560e-5", "2.000000e0").
std::string rust_exp(double v, int prec)
{
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    snprintf(buf, sizeof buf, "%.*e", prec, v);
    std::string s(buf);
    const size_t e = s.find('e');
    return s.substr(0, e) + "e" + std::to_string(atoi(s.c_str() + e + 1));
}

struct Cli {
    std::string mesh = "fangdata.txt", config = "config.json", out = "result.csv", reslog = "res.log";
    std::string wall; // empty: no wall table
    long steps = 0;
    bool quiet = false, no_output = false, forces = false;
    long threads = 0;
    long profile = 0; // steps timed after the run (0: none)
    jst_precision prec = JST_FP64;
};

int usage(const char *msg)
{
    if (msg) fprintf(stderr, "error: %s\n", msg);
    fprintf(stderr,
            "Usage: jst-gpu [--mesh PATH] [--config PATH] [--steps N] [--out PATH] [--reslog PATH]\n"
            "               [--quiet] [--threads N] [--no-output] [--precision fp64|mixed]\n"
            "               [--wall PATH] [--forces] [--profile N]\n");
    return 2;
}

struct LogCtx {
    FILE *log;
    bool quiet;
};

void on_step(long step, double res, void *user)
{
    LogCtx *c = (LogCtx *)user;
    if (!c->quiet) printf("step:%6ld | residual:%s\n", step, rust_exp(res, 6).c_str());
    if (c->log) fprintf(c->log, "%ld,%s\n", step, rust_exp(res, 6).c_str());
}

} // namespace

int main(int argc, char **argv)
{
    Cli cli;
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        auto val = [&](const char *name) -> const char * {
            if (a + 1 >= argc) {
                usage((std::string("missing value for ") + name).c_str());
                exit(2);
            }
            return argv[++a];
        };
        if (s == "--mesh") cli.mesh = val("--mesh");
        else if (s == "--config") cli.config = val("--config");
        else if (s == "--steps") cli.steps = atol(val("--steps"));
        else if (s == "--out") cli.out = val("--out");
        else if (s == "--reslog") cli.reslog = val("--reslog");
        else if (s == "--quiet") cli.quiet = true;
        else if (s == "--threads") cli.threads = atol(val("--threads"));
        else if (s == "--no-output") cli.no_output = true;
        else if (s == "--wall") cli.wall = val("--wall");
        else if (s == "--forces") cli.forces = true;
        else if (s == "--profile") {
            cli.profile = atol(val("--profile"));
            if (cli.profile <= 0 || cli.profile > 1000000) return usage("--profile takes a step count (1..1000000)");
        }
        else if (s == "--precision") {
            std::string p = val("--precision");
            if (p == "fp64") cli.prec = JST_FP64;
            else if (p == "mixed") cli.prec = JST_MIXED;
            else return usage("--precision is fp64 or mixed");
        } else if (s == "-h" || s == "--help") {
            usage(nullptr);
            return 0;
        } else {
            return usage(("unexpected argument " + s).c_str());
        }
    }
    jst_ctx *ctx = nullptr;
    if (jst_create_from_paths(cli.config.c_str(), cli.mesh.c_str(), cli.prec, &ctx) != JST_OK) {
        fprintf(stderr, "Error: %s\n", jst_last_error());
        return 1;
    }
    int ni = 0, nj = 0;
    jst_dims(ctx, &ni, &nj);
    const long cells = (long)ni * nj;
    printf("mesh %s: %d rings x %d points -> %ld cells\n", cli.mesh.c_str(), ni + 1, nj, cells);
    // the thread count of the CPU solver (bin/jst.rs:52-57) is reported for output parity;
    // the GPU run does not use it
    long threads = cli.threads;
    if (threads <= 0) {
        const long avail = std::max(1u, std::thread::hardware_concurrency());
        threads = std::min(std::max((cells + 24575) / 24576, 1L), avail);
    }
    printf("threads: %ld\n", threads);
    LogCtx lc{nullptr, cli.quiet};
    if (!cli.no_output) {
        lc.log = fopen(cli.reslog.c_str(), "w");
        if (!lc.log) {
            fprintf(stderr, "Error: cannot create %s\n", cli.reslog.c_str());
            jst_destroy(ctx);
            return 1;
        }
        fprintf(lc.log, "step,residual\n");
    }
    const auto t0 = std::chrono::steady_clock::now();
    jst_report rep{};
    jst_nonphysical np{};
    const jst_status st = jst_run(ctx, cli.steps, on_step, &lc, &rep, &np);
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (lc.log) fclose(lc.log);
    if (st != JST_OK) {
        if (st == JST_ERR_NONPHYSICAL)
            fprintf(stderr, "Error: NonPhysical { i: %ld, j: %ld, rho: %s, p: %s }\n", np.i, np.j,
                    rust_exp(np.rho, 6).c_str(), rust_exp(np.p, 6).c_str());
        else
            fprintf(stderr, "Error: %s\n", jst_last_error());
        jst_destroy(ctx);
        return 1;
    }
    printf("steps: %ld  final residual: %s%s\n", rep.steps, rust_exp(rep.residual, 6).c_str(),
           rep.converged ? "  (converged)" : "");
    printf("totaltime (physical): %s s\n", rust_exp(rep.totaltime, 6).c_str());
    const double per = rep.steps > 0 ? wall / (double)rep.steps : 0.0;
    printf("wall clock: %.6f s  (%.4f ms/step, %.1f ns/cell/step)\n", wall, per * 1e3,
           per / (double)cells * 1e9);
    if (!cli.no_output) {
        if (jst_write_csv(ctx, cli.out.c_str()) != JST_OK) {
            fprintf(stderr, "Error: %s\n", jst_last_error());
            jst_destroy(ctx);
            return 1;
        }
        printf("all data is written in %s\n", cli.out.c_str());
    }
    if (cli.forces) {
        double cl = 0.0, cd = 0.0;
        if (jst_forces(ctx, nullptr, nullptr, &cl, &cd) != JST_OK) {
            fprintf(stderr, "Error: %s\n", jst_last_error());
            jst_destroy(ctx);
            return 1;
        }
        printf("CL: %s  CD: %s\n", rust_exp(cl, 8).c_str(), rust_exp(cd, 8).c_str());
    }
    if (!cli.wall.empty()) {
        if (jst_write_wall_csv(ctx, cli.wall.c_str()) != JST_OK) {
            fprintf(stderr, "Error: %s\n", jst_last_error());
            jst_destroy(ctx);
            return 1;
        }
        printf("wall data is written in %s\n", cli.wall.c_str());
    }
    if (cli.profile > 0) {
        // after every output is written: the timed steps advance the state further
        double total_ms = 0.0, stage_ms = 0.0;
        if (jst_time_steps(ctx, (int)cli.profile, &total_ms, &stage_ms) != JST_OK) {
            fprintf(stderr, "Error: %s\n", jst_last_error());
            jst_destroy(ctx);
            return 1;
        }
        const double ms = total_ms / (double)cli.profile;
        printf("profile: %ld steps (HIP events): %.4f ms/step, stage kernel %.2f us, %.3e cell-stage updates/s\n",
               cli.profile, ms, stage_ms * 1e3, 5.0 * (double)cells / (ms * 1e-3));
    }
    jst_destroy(ctx);
    return 0;
}

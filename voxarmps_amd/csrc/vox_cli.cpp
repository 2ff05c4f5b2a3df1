// vox_cli.cpp — voxmps, the reference's command line (main.cu:60-209) over
// libvoxarmps: the same 16 positional arguments,
//
//   voxmps [test] [viscoplastic] [turbulence] [compressibility] [matrix opt]
//          [source opt] [code opt] [dimensions] [print all] [print fluid]
//          [rho1] [nu1] [rho2] [nu2] [total time] [time step]
//
// e.g. voxmps dam novisc noturb nocomp nomatopt nosrcopt sequential 2d printall no
//             1000.0 0.000001 500.0 0.001 1 0.005
//
// then the time loop of main.cu:1027-1363 on the GPU, with the reference's
// VTU output per step (output/outN.vtu for every particle, output_fluid/outN.vtu
// for the fluid; inOut.cpp:104-336). Options after the 16 arguments (none in
// the reference, whose input paths are fixed Windows paths):
//   --inputs DIR     directory of the .grd inputs (default "inputs")
//   --outdir DIR     directory that receives output/ and output_fluid/ (".")
//   --max-steps N    stop after N steps even if the total time is not reached
//   --every N        write the VTU files every N steps (default 1)
// Host code only: it links libvoxarmps and includes nothing but its C header.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/voxarmps.h"

namespace {

bool is_one_of(const std::string& s, std::initializer_list<const char*> words)
{
    for (const char* w : words)
        if (s == w) return true;
    return false;
}

int die(const std::string& what)
{
    std::fprintf(stderr, "voxmps: %s\n", what.c_str());
    return 1;
}

int check(int rc, const char* what)
{
    if (rc != VOX_OK) {
        std::fprintf(stderr, "voxmps: %s failed: %s\n", what, vox_last_error());
        return 1;
    }
    return 0;
}

void usage()
{
    std::fprintf(stderr,
                 "Usage: voxmps [test] [viscoplastic (in second fluid)] [turbulence] "
                 "[compressibility] [matrix optimization] [source optimization] [performance "
                 "optimization] [dimensions] [print all particles] [print only fluid] "
                 "[densityfluid1] [viscosityfluid1] [densityfluid2] [viscosityfluid2] "
                 "[totalsimulationtime] [timestepduration] [--inputs DIR] [--outdir DIR] "
                 "[--max-steps N] [--every N]\n"
                 "Example: voxmps dam novisc noturb nocomp nomatopt nosrcopt sequential 2d no "
                 "printfluid 1000.0 0.000001 500.0 0.001 1 0.0005\n");
}

}  // namespace

int main(int argc, char** argv)
{
    const int nargs = 16;
    if (argc - 1 < nargs) {
        std::fprintf(stderr, "Unexpected number of arguments: %d (expected %d)\n", argc - 1, nargs);
        usage();
        return 1;
    }
    std::vector<std::string> a(argv + 1, argv + 1 + nargs);
    std::string inputs = "inputs", outdir = ".";
    long max_steps = -1, every = 1;
    for (int k = 1 + nargs; k < argc; ++k) {
        const std::string o = argv[k];
        if (k + 1 >= argc) return die("option " + o + " needs a value");
        const std::string v = argv[++k];
        if (o == "--inputs") inputs = v;
        else if (o == "--outdir") outdir = v;
        else if (o == "--max-steps") max_steps = std::atol(v.c_str());
        else if (o == "--every") every = std::max(1L, std::atol(v.c_str()));
        else return die("unknown option " + o);
    }
    // the words of main.cu:88-207
    const std::string test = a[0];
    const bool visc = is_one_of(a[1], {"visc", "viscosity", "viscoplasticity", "true", "yes"});
    const bool turb = is_one_of(a[2], {"turb", "turbulence", "true", "yes"});
    const bool comp = is_one_of(a[3], {"comp", "compressible", "true", "yes"});
    const bool matopt = is_one_of(a[4], {"matopt", "matrixoptimization", "true", "yes"});
    const bool srcopt = is_one_of(a[5], {"srcopt", "sourceoptimization", "true", "yes"});
    const std::string code = a[6];
    bool three = is_one_of(a[7], {"3d", "3D", "threedim", "true", "yes"});
    const bool out_all = is_one_of(a[8], {"outAll", "printall", "PRINTALL", "printAll", "true", "yes"});
    const bool out_fluid =
        is_one_of(a[9], {"outFluid", "printfluid", "PRINTFLUID", "printFluid", "true", "yes"});
    const double rho1 = std::atof(a[10].c_str()), nu1 = std::atof(a[11].c_str());
    const double rho2 = std::atof(a[12].c_str()), nu2 = std::atof(a[13].c_str());
    const double T = std::atof(a[14].c_str()), dt = std::atof(a[15].c_str());
    const bool known = is_one_of(test, {"oil", "drop", "shear", "turb"});
    if (three && known) {
        std::printf("The %s test does not support 3D, a 2D simulation will be generated.\n",
                    test.c_str());
        three = false;
    }
    const int dim = three ? 3 : 2;

    // input (main.cu:496-682): count, then rows x y [z] PTYPE u v [w] p
    const std::string path = inputs + "/" + vox_main_input(test.c_str(), three ? 1 : 0);
    int tp = 0;
    if (vox_grd_count(path.c_str(), &tp) != VOX_OK)
        return die("unable to read " + path + (test == "oil" ? " (write it with: python -m "
                                                                "voxarmps_amd.generators oil DIR)"
                                                              : ""));
    std::vector<double> x(tp), y(tp), z(tp, 0.0), u(tp), v(tp), w(tp, 0.0), p(tp);
    std::vector<int32_t> ptype(tp);
    int gp = 0, wp = 0, fp = 0;
    if (check(vox_grd_read(path.c_str(), dim, tp, x.data(), y.data(), z.data(), u.data(), v.data(),
                           w.data(), p.data(), ptype.data(), &gp, &wp, &fp), "reading the input"))
        return 1;
    if (test == "drop")                                             // main.cu:645-651
        for (int i = 0; i < tp; ++i) {
            u[i] = -1 * x[i];
            v[i] = 1 * y[i];
            w[i] = 0.0;
        }

    vox_params P;
    if (check(vox_params_main(&P, test.c_str(), visc, turb, comp, matopt, srcopt, code.c_str(),
                              three ? 1 : 0, rho1, nu1, rho2, nu2, dt, tp, x.data(), y.data(),
                              z.data()), "setting up the run"))
        return 1;
    std::printf("    %s, %s, %s, %dD, test %s\n", P.method == 3 ? "WEAKLY COMPRESSIBLE" : "FULLY INCOMPRESSIBLE",
                P.ref_gpu ? "the reference's CUDA path" : "the reference's CPU path",
                P.turb ? "SPS turbulence" : "no turbulence model", dim,
                known ? test.c_str() : "dam");
    std::printf("    Total Particle Number: %d (ghost %d, wall %d, fluid %d)\n", tp, gp, wp, fp);
    std::printf("    TOTAL SIMULATION TIME: %g seconds, TIME-STEP DURATION: %g seconds\n", T, dt);

    vox_ctx* ctx = nullptr;
    if (check(vox_create(&ctx, &P, tp, gp, wp, fp), "vox_create")) return 1;
    int rc = 0;
    auto write = [&](long step) {
        const std::string n = std::to_string(step);
        if (out_all) {
            mkdir((outdir + "/output").c_str(), 0755);
            rc |= check(vox_write_vtu(ctx, (outdir + "/output/out" + n + ".vtu").c_str(), 0), "VTU");
        }
        if (out_fluid) {
            mkdir((outdir + "/output_fluid").c_str(), 0755);
            rc |= check(vox_write_vtu(ctx, (outdir + "/output_fluid/out" + n + ".vtu").c_str(), 1),
                        "VTU");
        }
    };
    rc |= check(vox_upload_state(ctx, x.data(), y.data(), dim == 3 ? z.data() : nullptr, u.data(),
                                 v.data(), dim == 3 ? w.data() : nullptr, p.data(), ptype.data()),
                "upload");
    if (!rc) write(0);
    if (!rc) rc |= check(vox_init(ctx), "vox_init");
    const auto start = std::chrono::steady_clock::now();
    double t = 0.0;
    long step = 1;
    for (; !rc && t <= T && (max_steps < 0 || step <= max_steps); ++step) {   // main.cu:1027
        rc |= check(vox_step(ctx, 1), "time step");
        if (rc) break;
        if (step % every == 0) write(step);
        vox_params cur;
        vox_get_params(ctx, &cur);
        t += cur.dt;
        if (step == 1) {
            vox_sync(ctx);
            const double ms = std::chrono::duration<double, std::milli>(
                                  std::chrono::steady_clock::now() - start).count();
            std::printf("     Estimated FPS: %g (first step %.3f ms)\n", 1000.0 / ms, ms);
        }
    }
    vox_sync(ctx);
    const double total = std::chrono::duration<double, std::milli>(
                             std::chrono::steady_clock::now() - start).count();
    vox_status st;
    vox_get_status(ctx, &st);
    std::printf("     Total running time: %.1f milliseconds\n", total);
    std::printf("     Total time steps: %ld (simulated time %.6g s)\n", step - 1, st.time);
    std::printf("     Average FPS: %g\n", (step - 1) / (total / 1000.0));
    std::printf("     Last PPE solve: %d iterations, relative residual %.3g; errors: %d NaN, %d "
                "outside the box\n", st.last_iters, st.last_relres, st.nan_count, st.out_of_box);
    vox_destroy(ctx);
    return rc;
}

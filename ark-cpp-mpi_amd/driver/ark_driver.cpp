// One rank's box through the C++ host interface: the reference's time loop
// (source/main.cpp:80-104) with its phase functions running on the GPU.
//
//   ark_driver [--phases] [--mixed] [--corrected] N1 N2 N3 STEPS [OUTDIR]
//   ark_driver --channel [--phases] STEPS OUTDIR [NPRINT [PRINT_FROM [AVG_FROM]]]
//
// --channel runs the reference's own program (source/main.cpp:53-116): its
// Input() and InitializeData(), the first WriteDataParaView, then STEPS steps
// with WriteDataParaView and the progress line "step: .. dt:.. time:.." after
// every step n >= PRINT_FROM with n % NPRINT == 0 (defaults 1000 and 70000,
// the reference's); with AVG_FROM, AverageValues after every step from
// AVG_FROM on and WriteAverageValues at the end.  Files go to OUTDIR, named
// as the reference names them.
//
// Taylor-Green vortex on a box of side 2*pi: c = 10, nu = 1/1600, CFL 0.3.
// --phases  call TimeStepSize, Phase1, StressTensor, UseForces, Phase2,
//           Phase1, UseForces once per step, as the reference's loop does
//           (otherwise Solver::Step runs the fused kernel);
// --mixed   mixed precision (fused steps only);
// --corrected  the corrected CABARET variant (gq from Q on every axis).
// Prints one JSON line: final dt, TIME, nStep, the kinetic energy as
// WriteEnergy defines it (source/main.cpp:2851-2876), mass, and the
// Mcell-steps/s of the step loop.  With OUTDIR every array is written as
// OUTDIR/<name>.f64 (Arr3d order, little endian) with OUTDIR/meta.txt.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "ark/ark_host.hpp"

namespace {

const char* kNames[ARK_NUM_FIELDS] = {
    "roCon", "u1Con", "u2Con", "u3Con", "tCon", "ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon",
    "f1", "f2", "f3", "condition",
    "ro1", "t1", "u11", "u21", "u31", "p1", "ro2", "t2", "u12", "u22", "u32", "p2",
    "ro3", "t3", "u13", "u23", "u33", "p3",
    "sigm11", "sigm21", "sigm31", "sigm12", "sigm22", "sigm32", "sigm13", "sigm23", "sigm33"};

int usage() {
    fprintf(stderr,
            "usage: ark_driver [--phases] [--mixed] [--corrected] N1 N2 N3 STEPS [OUTDIR]\n"
            "       ark_driver --channel [--phases] STEPS OUTDIR [NPRINT [PRINT_FROM [AVG_FROM]]]\n");
    return 2;
}

void dump(ark::Solver& s, const std::string& dir) {
    mkdir(dir.c_str(), 0755);
    for (int f = 0; f < ARK_NUM_FIELDS; ++f) {
        ark::Arr3d a(s.extent(f, 0), s.extent(f, 1), s.extent(f, 2));
        a.bind(&s, f);
        a.download();
        std::string path = dir + "/" + kNames[f] + ".f64";
        FILE* fh = fopen(path.c_str(), "wb");
        if (!fh) throw ark::Error(ARK_ERR_INVALID, "cannot write " + path);
        fwrite(a.ref(), sizeof(double), (size_t)a.n1() * a.n2() * a.n3(), fh);
        fclose(fh);
    }
    const ark_params& p = s.params();
    FILE* fh = fopen((dir + "/meta.txt").c_str(), "w");
    if (!fh) throw ark::Error(ARK_ERR_INVALID, "cannot write meta.txt");
    ark_diag d = s.Diagnostics();
    fprintf(fh, "N1 %d\nN2 %d\nN3 %d\n", p.N1, p.N2, p.N3);
    fprintf(fh, "dx1 %a\ndx2 %a\ndx3 %a\ndt %a\nCFL %a\nVIS %a\nsound %a\nro0 %a\nt0 %a\nTIME %a\nnStep %lld\n",
            p.dx1, p.dx2, p.dx3, s.dt(), p.CFL, p.VIS, p.sound, p.ro0, p.t0, d.time, (long long)d.nstep);
    fclose(fh);
}

// The reference's program on its own problem: Input/InitializeData, the
// first output, the loop of source/main.cpp:80-110 with its outputs.
int run_channel(bool phases, int steps, const std::string& dir, int nprint, int print_from, int avg_from) {
    const std::string out = dir.empty() || dir.back() == '/' ? dir : dir + "/";
    mkdir(out.c_str(), 0755);
    ark::Solver s(ark::Solver::channel());
    s.InitChannel(1.0);
    s.WriteDataParaView(out);
    s.Synchronize();
    const auto t0 = std::chrono::steady_clock::now();
    double start_time = 0.0;
    int n = 0;
    while (n < steps) {
        // the next step with an output or an average; the steps before it
        // run in one call (the library replays them from a graph)
        int next = steps;
        if (avg_from >= 0) next = std::min(next, std::max(n + 1, avg_from));
        if (nprint > 0) {
            const int from = std::max(n + 1, print_from);
            next = std::min(next, (from + nprint - 1) / nprint * nprint);
        }
        if (phases) {
            for (; n < next; ++n) {
                s.TimeStepSize();
                s.NextStep();
                s.HalfTime();
                s.Phase1();
                s.StressTensor();
                s.UseForces();
                s.Phase2();
                s.Phase1();
                s.UseForces();
                s.HalfTime();
            }
        } else {
            s.Step(next - n);
            n = next;
        }
        if (avg_from >= 0 && n >= avg_from) {
            if (n == avg_from) start_time = s.Diagnostics().time;
            s.AverageValues();
        }
        if (nprint > 0 && n >= print_from && n % nprint == 0) {
            s.WriteDataParaView(out);
            printf("step: %d dt:%E time:%8.4f\n", n, s.dt(), s.Diagnostics().time);
        }
    }
    if (avg_from >= 0) s.WriteAverageValues(out, start_time);
    s.Synchronize();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const ark_diag d = s.Diagnostics();
    const ark_params& p = s.params();
    const double cells = (double)p.N1 * p.N2 * p.N3;
    printf("{\"problem\": \"channel\", \"grid\": [%d, %d, %d], \"steps\": %d, \"mode\": \"%s\", \"dt\": %.17g, "
           "\"time\": %.17g, \"nstep\": %lld, \"mass\": %.17g, \"nonfinite\": %lld, \"loop_seconds\": %.6f, "
           "\"mcell_steps_per_s\": %.3f}\n",
           p.N1, p.N2, p.N3, steps, phases ? "phases" : "step", s.dt(), d.time, (long long)d.nstep, d.mass,
           (long long)d.nonfinite, secs, secs > 0 ? cells * steps / secs / 1e6 : 0.0);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    bool phases = false, mixed = false, corrected = false, channel = false;
    int a = 1;
    for (; a < argc && argv[a][0] == '-' && argv[a][1] == '-'; ++a) {
        if (!strcmp(argv[a], "--phases")) phases = true;
        else if (!strcmp(argv[a], "--mixed")) mixed = true;
        else if (!strcmp(argv[a], "--corrected")) corrected = true;
        else if (!strcmp(argv[a], "--channel")) channel = true;
        else return usage();
    }
    if (channel) {
        if (mixed || corrected || argc - a < 2) return usage();
        const int steps = atoi(argv[a]);
        const int nprint = argc - a > 2 ? atoi(argv[a + 2]) : 1000;
        const int print_from = argc - a > 3 ? atoi(argv[a + 3]) : 70000;
        const int avg_from = argc - a > 4 ? atoi(argv[a + 4]) : -1;
        if (steps < 0 || nprint < 0) return usage();
        try {
            return run_channel(phases, steps, argv[a + 1], nprint, print_from, avg_from);
        } catch (const ark::Error& e) {
            fprintf(stderr, "ark_driver: %s (status %d)\n", e.what(), (int)e.status);
            return 1;
        }
    }
    if (argc - a < 4) return usage();
    const int N1 = atoi(argv[a]), N2 = atoi(argv[a + 1]), N3 = atoi(argv[a + 2]), steps = atoi(argv[a + 3]);
    const char* outdir = argc - a > 4 ? argv[a + 4] : nullptr;
    if (N1 < 2 || N2 < 2 || N3 < 2 || steps < 0) return usage();
    if (phases && mixed) {
        fprintf(stderr, "ark_driver: the phase functions run in FP64 only\n");
        return 2;
    }
    try {
        const double two_pi = 2.0 * M_PI;
        ark_params p = ark::Solver::periodic_box(N1, N2, N3, two_pi - 0.0, two_pi - 0.0, two_pi - 0.0, 0.3,
                                                 1.0 / 1600.0, 10.0);
        if (mixed) p.precision = ARK_PREC_MIXED;
        if (corrected) p.scheme = ARK_SCHEME_CORRECTED;
        ark::Solver s(p);
        s.InitTGV(1.0);
        s.Synchronize();
        const auto t0 = std::chrono::steady_clock::now();
        if (phases) {
            for (int n = 0; n < steps; ++n) {
                s.TimeStepSize();
                s.NextStep();
                s.HalfTime();
                s.Phase1();
                s.StressTensor();
                s.UseForces();
                s.Phase2();
                s.Phase1();
                s.UseForces();
                s.HalfTime();
            }
        } else {
            s.Step(steps);
        }
        s.Synchronize();
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        ark_diag d = s.Diagnostics();
        const double cells = (double)N1 * N2 * N3;
        printf("{\"grid\": [%d, %d, %d], \"steps\": %d, \"mode\": \"%s\", \"precision\": \"%s\", "
               "\"dt\": %.17g, \"time\": %.17g, \"nstep\": %lld, \"ke_reference\": %.17g, \"mass\": %.17g, "
               "\"nonfinite\": %lld, \"loop_seconds\": %.6f, \"mcell_steps_per_s\": %.3f}\n",
               N1, N2, N3, steps, phases ? "phases" : "fused", mixed ? "mixed" : "fp64", s.dt(), d.time,
               (long long)d.nstep, d.ke_reference, d.mass, (long long)d.nonfinite, secs,
               secs > 0 ? cells * steps / secs / 1e6 : 0.0);
        if (outdir) dump(s, outdir);
    } catch (const ark::Error& e) {
        fprintf(stderr, "ark_driver: %s (status %d)\n", e.what(), (int)e.status);
        return 1;
    }
    return 0;
}

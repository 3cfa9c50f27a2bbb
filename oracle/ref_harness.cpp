// Test infrastructure only: a driver around the UNMODIFIED reference solver.
//
// This translation unit compiles the reference's own source/main.cpp (read in
// place from the reference checkout by oracle/Makefile) with its entry point and
// its hard-coded channel set-up renamed, and drives the reference's own phase
// functions (TimeStepSize, Phase1, StressTensor, UseForces, Phase2;
// source/main.cpp:374-2495) on a triply periodic box.  Its outputs are the
// golden vectors that pin the project's CPU restatement (oracle/ark_oracle.c)
// and, through it, the HIP kernels.  Nothing in the product links or runs it.
//
// What the driver adds, and why:
//   * a Cartesian communicator with periods={1,1,1} (the reference uses
//     periods={0,0,0}, source/main.cpp:55-61) so the reference's own MPI
//     halo exchange delivers periodic ghosts;
//   * coordX=coordY=coordZ=-1, which disables every physical-boundary branch
//     (source/main.cpp:803-845, 1633-1698, 1731-1745);
//   * allocation with the reference's extents (source/main.cpp:198-277) and a
//     deterministic initial state (Taylor-Green vortex, or a state read from
//     files), including every ghost and the arrays the reference leaves
//     uninitialised;
//   * Con := nCon at t=0 (the first TimeStepSize reads Con, source/main.cpp:382).
//
// Modes:
//   tgv     N1 N2 N3 STEPS OUTDIR [DUMPSTEP...]  Taylor-Green run, any rank count
//   phase   NAME INDIR OUTDIR                     one reference phase on a saved state (1 rank)
//   run     INDIR STEPS OUTDIR                    STEPS full steps from a saved state (1 rank)
//   channel STEPS OUTDIR [DUMPSTEP...]            the reference's own problem (1 rank): its
//           Input() and InitializeData() (source/main.cpp:118-372), periods={0,0,0}
//           as in its main (source/main.cpp:55-61), so every physical-boundary
//           branch is live; arrays Input() allocates and InitializeData() leaves
//           unset are zero-filled first, and Con := nCon as above.
// Options (before the mode): -nu V  -cfl V  -c V  -quiet
//   channel only:  -vtk      the reference's WriteDataParaView() at each dump step
//                            (OUTDIR/out_<step>.vtk, source/main.cpp:2577-2745)
//                  -avg S    AverageValues() after every step from step S on, the
//                            start time taken at step S as the reference's main
//                            does (source/main.cpp:93-96), and WriteAverageValues()
//                            at the end (OUTDIR/average.vtk, 2747-2849)
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <cmath>

#define main ark_ref_main
#define Input ark_ref_Input
#define InitializeData ark_ref_InitializeData
#include REF_MAIN_CPP
#undef main
#undef Input
#undef InitializeData

namespace {

struct Named { const char* name; Arr3d** arr; bool face; };

Named g_arrays[] = {
    {"roCon", &roCon, false}, {"u1Con", &u1Con, false}, {"u2Con", &u2Con, false},
    {"u3Con", &u3Con, false}, {"tCon", &tCon, false},
    {"ronCon", &ronCon, false}, {"u1nCon", &u1nCon, false}, {"u2nCon", &u2nCon, false},
    {"u3nCon", &u3nCon, false}, {"tnCon", &tnCon, false},
    {"f1", &f1, false}, {"f2", &f2, false}, {"f3", &f3, false},
    {"condition", &condition, false},
    {"ro1", &ro1, true}, {"t1", &t1, true}, {"u11", &u11, true}, {"u21", &u21, true},
    {"u31", &u31, true}, {"p1", &p1, true},
    {"ro2", &ro2, true}, {"t2", &t2, true}, {"u12", &u12, true}, {"u22", &u22, true},
    {"u32", &u32, true}, {"p2", &p2, true},
    {"ro3", &ro3, true}, {"t3", &t3, true}, {"u13", &u13, true}, {"u23", &u23, true},
    {"u33", &u33, true}, {"p3", &p3, true},
    {"sigm11", &sigm11, true}, {"sigm21", &sigm21, true}, {"sigm31", &sigm31, true},
    {"sigm12", &sigm12, true}, {"sigm22", &sigm22, true}, {"sigm32", &sigm32, true},
    {"sigm13", &sigm13, true}, {"sigm23", &sigm23, true}, {"sigm33", &sigm33, true},
};
const int kNumArrays = sizeof(g_arrays) / sizeof(g_arrays[0]);

bool g_quiet = false;

long cell_count() { return (long)(n1 + 1) * (n2 + 1) * (n3 + 1); }
long face_count() { return (long)(n1 + 2) * (n2 + 2) * (n3 + 2); }
long count_of(const Named& a) { return a.face ? face_count() : cell_count(); }

// Allocation with the reference's extents (source/main.cpp:198-277).
void allocate_state() {
    x1 = new double[n1 + 2];
    x2 = new double[n2 + 2];
    x3 = new double[n3 + 2];
    for (int a = 0; a < kNumArrays; ++a) {
        if (g_arrays[a].face) *g_arrays[a].arr = new Arr3d(n1 + 2, n2 + 2, n3 + 2);
        else *g_arrays[a].arr = new Arr3d(n1 + 1, n2 + 1, n3 + 1);
        double* p = (*g_arrays[a].arr)->ref();
        for (long q = 0; q < count_of(g_arrays[a]); ++q) p[q] = 0.0;
    }
    averageU1 = new Arr3d(n1 + 1, n2 + 2, n3 + 1);
    averageU2 = new Arr3d(n1 + 1, n2 + 2, n3 + 1);
    averageU3 = new Arr3d(n1 + 1, n2 + 2, n3 + 1);
    int nmax = (int)max3d(n1, n2, n3);
    rBuf = new double[nmax + 2];
    qBuf = new double[nmax + 1];
    tfBuf = new double[nmax + 2];
    tbBuf = new double[nmax + 2];
    u2fBuf = new double[nmax + 2];
    u2bBuf = new double[nmax + 2];
    u3fBuf = new double[nmax + 2];
    u3bBuf = new double[nmax + 2];
}

void setup_comm(int periodic) {
    dims[0] = 0; dims[1] = 0; dims[2] = 1;   // the reference never splits z (main.cpp:54)
    int periods[3] = {periodic, periodic, periodic};
    MPI_Comm_size(MPI_COMM_WORLD, &nproc);
    MPI_Dims_create(nproc, 3, dims);
    MPI_Cart_create(MPI_COMM_WORLD, 3, dims, periods, 0, &newComm);
    MPI_Comm_rank(newComm, &rank);
    MPI_Cart_coords(newComm, rank, 3, coords);
    MPI_Cart_shift(newComm, 0, 1, &left, &right);
    MPI_Cart_shift(newComm, 1, 1, &down, &up);
    MPI_Cart_shift(newComm, 2, 1, &bottom, &top);
    px = dims[0]; py = dims[1]; pz = dims[2];
}

// Periodic wrap of a global cell index into 1..N (ghost 0 -> N, N+1 -> 1) and
// of a global face index into 1..N (face N+1 -> 1, ghost 0 -> N, N+2 -> 2).
int wrap(int g, int N) { return ((g - 1) % N + N) % N + 1; }

struct Tgv { double p0_amp, U0, rho0, c; };

void tgv_point(const Tgv& t, double x, double y, double z, double* u, double* v, double* w, double* p) {
    *u = t.U0 * sin(x) * cos(y) * cos(z);
    *v = -t.U0 * cos(x) * sin(y) * cos(z);
    *w = 0.0;
    *p = t.p0_amp * (cos(2.0 * x) + cos(2.0 * y)) * (cos(2.0 * z) + 2.0);
}

void init_tgv(int N1, int N2, int N3) {
    Tgv t;
    t.U0 = 1.0; t.rho0 = ro0_g; t.c = sound;
    t.p0_amp = t.rho0 * t.U0 * t.U0 / 16.0;
    int ox = coords[0] * (n1 - 1), oy = coords[1] * (n2 - 1), oz = coords[2] * (n3 - 1);
    // cells (including ghosts, edges and corners): value at the wrapped global cell
    for (int k = 0; k <= n3; ++k)
        for (int j = 0; j <= n2; ++j)
            for (int i = 0; i <= n1; ++i) {
                double x = (wrap(ox + i, N1) - 0.5) * dx1;
                double y = (wrap(oy + j, N2) - 0.5) * dx2;
                double z = (wrap(oz + k, N3) - 0.5) * dx3;
                double u, v, w, p;
                tgv_point(t, x, y, z, &u, &v, &w, &p);
                u1nCon->elem(i, j, k) = u; u2nCon->elem(i, j, k) = v; u3nCon->elem(i, j, k) = w;
                ronCon->elem(i, j, k) = ro0_g + p / (sound * sound);
                tnCon->elem(i, j, k) = t0;
                u1Con->elem(i, j, k) = u; u2Con->elem(i, j, k) = v; u3Con->elem(i, j, k) = w;
                roCon->elem(i, j, k) = ronCon->elem(i, j, k);
                tCon->elem(i, j, k) = t0;
                condition->elem(i, j, k) = 0.0;
                f1->elem(i, j, k) = 0.0; f2->elem(i, j, k) = 0.0; f3->elem(i, j, k) = 0.0;
            }
    // faces: axis d face index uses the face wrap, the other two use the cell wrap
    Arr3d* U[3][3] = {{u11, u21, u31}, {u12, u22, u32}, {u13, u23, u33}};
    Arr3d* P[3] = {p1, p2, p3};
    Arr3d* R[3] = {ro1, ro2, ro3};
    Arr3d* T[3] = {t1, t2, t3};
    for (int d = 0; d < 3; ++d)
        for (int k = 0; k <= n3 + 1; ++k)
            for (int j = 0; j <= n2 + 1; ++j)
                for (int i = 0; i <= n1 + 1; ++i) {
                    double x = d == 0 ? (wrap(ox + i, N1) - 1.0) * dx1 : (wrap(ox + i, N1) - 0.5) * dx1;
                    double y = d == 1 ? (wrap(oy + j, N2) - 1.0) * dx2 : (wrap(oy + j, N2) - 0.5) * dx2;
                    double z = d == 2 ? (wrap(oz + k, N3) - 1.0) * dx3 : (wrap(oz + k, N3) - 0.5) * dx3;
                    double u, v, w, p;
                    tgv_point(t, x, y, z, &u, &v, &w, &p);
                    U[d][0]->elem(i, j, k) = u;
                    U[d][1]->elem(i, j, k) = v;
                    U[d][2]->elem(i, j, k) = w;
                    P[d]->elem(i, j, k) = p;
                    R[d]->elem(i, j, k) = ro0_g + p / (sound * sound);
                    T[d]->elem(i, j, k) = t0;
                }
}

void write_meta(const std::string& dir, int N1, int N2, int N3) {
    std::string fn = dir + "/meta.txt";
    FILE* f = fopen(fn.c_str(), "w");
    if (!f) { fprintf(stderr, "cannot write %s\n", fn.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
    fprintf(f, "N1 %d\nN2 %d\nN3 %d\n", N1, N2, N3);
    fprintf(f, "dx1 %a\ndx2 %a\ndx3 %a\n", dx1, dx2, dx3);
    fprintf(f, "dt %a\nCFL %a\nVIS %a\nsound %a\nro0 %a\nt0 %a\nTIME %a\nnStep %d\n",
            dt, CFL, VIS, sound, ro0_g, t0, TIME, nStep);
    fclose(f);
}

bool read_meta(const std::string& dir, int* N1, int* N2, int* N3) {
    std::string fn = dir + "/meta.txt";
    FILE* f = fopen(fn.c_str(), "r");
    if (!f) return false;
    char key[64], val[128];
    while (fscanf(f, "%63s %127s", key, val) == 2) {
        std::string k(key);
        if (k == "N1") *N1 = atoi(val);
        else if (k == "N2") *N2 = atoi(val);
        else if (k == "N3") *N3 = atoi(val);
        else if (k == "nStep") nStep = atoi(val);
        else {
            double v = strtod(val, nullptr);
            if (k == "dx1") dx1 = v; else if (k == "dx2") dx2 = v; else if (k == "dx3") dx3 = v;
            else if (k == "dt") dt = v; else if (k == "CFL") CFL = v; else if (k == "VIS") VIS = v;
            else if (k == "sound") sound = v; else if (k == "ro0") ro0_g = v; else if (k == "t0") t0 = v;
            else if (k == "TIME") TIME = v;
        }
    }
    fclose(f);
    return true;
}

void dump_local(const std::string& dir, int N1, int N2, int N3) {
    for (int a = 0; a < kNumArrays; ++a) {
        std::string fn = dir + "/" + g_arrays[a].name + ".f64";
        FILE* f = fopen(fn.c_str(), "wb");
        if (!f) { fprintf(stderr, "cannot write %s\n", fn.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
        fwrite((*g_arrays[a].arr)->ref(), sizeof(double), count_of(g_arrays[a]), f);
        fclose(f);
    }
    write_meta(dir, N1, N2, N3);
}

bool load_local(const std::string& dir) {
    for (int a = 0; a < kNumArrays; ++a) {
        std::string fn = dir + "/" + g_arrays[a].name + ".f64";
        FILE* f = fopen(fn.c_str(), "rb");
        if (!f) { fprintf(stderr, "missing %s\n", fn.c_str()); return false; }
        long n = count_of(g_arrays[a]);
        long got = (long)fread((*g_arrays[a].arr)->ref(), sizeof(double), n, f);
        fclose(f);
        if (got != n) { fprintf(stderr, "short %s: %ld of %ld\n", fn.c_str(), got, n); return false; }
    }
    return true;
}

// Global interior of the five level-n cell fields, reduced onto rank 0.
void dump_global(const std::string& dir, int N1, int N2, int N3) {
    Arr3d* F[5] = {ronCon, u1nCon, u2nCon, u3nCon, tnCon};
    const char* names[5] = {"ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon"};
    long tot = (long)N1 * N2 * N3;
    std::vector<double> loc(tot), glob(tot);
    int ox = coords[0] * (n1 - 1), oy = coords[1] * (n2 - 1), oz = coords[2] * (n3 - 1);
    for (int f = 0; f < 5; ++f) {
        for (long q = 0; q < tot; ++q) loc[q] = 0.0;
        for (int k = 1; k < n3; ++k)
            for (int j = 1; j < n2; ++j)
                for (int i = 1; i < n1; ++i)
                    loc[((long)(oz + k - 1) * N2 + (oy + j - 1)) * N1 + (ox + i - 1)] = F[f]->elem(i, j, k);
        MPI_Reduce(loc.data(), glob.data(), (int)tot, MPI_DOUBLE, MPI_SUM, 0, newComm);
        if (rank == 0) {
            std::string fn = dir + "/" + names[f] + ".interior.f64";
            FILE* fp = fopen(fn.c_str(), "wb");
            if (!fp) { fprintf(stderr, "cannot write %s\n", fn.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
            fwrite(glob.data(), sizeof(double), tot, fp);
            fclose(fp);
        }
    }
    if (rank == 0) write_meta(dir, N1, N2, N3);
}

// Kinetic-energy diagnostic with the reference's own definition and loop order
// (WriteEnergy, source/main.cpp:2851-2876), returned instead of appended to a file.
double energy_like_reference() {
    double energy = 0;
    for (int i = 1; i < n1; i++)
        for (int j = 1; j < n2; j++)
            for (int k = 1; k < n3; k++)
                energy += (u1nCon->elem(i, j, k) * u1nCon->elem(i, j, k) +
                           u2nCon->elem(i, j, k) * u2nCon->elem(i, j, k) +
                           u3nCon->elem(i, j, k) * u3nCon->elem(i, j, k)) / 2;
    double all = 0.;
    MPI_Allreduce(&energy, &all, 1, MPI_DOUBLE, MPI_SUM, newComm);
    return all * (x3_t - x3_b) * (x2_n - x2_s) * (x1_e - x1_w);
}

double mass_sum() {
    double m = 0;
    double dvc = dx1 * dx2 * dx3;
    for (int k = 1; k < n3; k++)
        for (int j = 1; j < n2; j++)
            for (int i = 1; i < n1; i++) m += ronCon->elem(i, j, k) * dvc;
    double all = 0.;
    MPI_Allreduce(&m, &all, 1, MPI_DOUBLE, MPI_SUM, newComm);
    return all;
}

// One time step in the reference's own order (source/main.cpp:80-104).
void one_step() {
    TimeStepSize();
    nStep++;
    TIME += 0.5 * dt;
    Phase1();
    StressTensor();
    UseForces();
    Phase2();
    Phase1();
    UseForces();
    TIME += 0.5 * dt;
}

void set_domain(int N1, int N2, int N3) {
    n1_g = N1 + 1; n2_g = N2 + 1; n3_g = N3 + 1;
    if (N1 % px || N2 % py || N3 % pz) {
        if (rank == 0) fprintf(stderr, "grid %dx%dx%d not divisible by ranks %dx%dx%d\n", N1, N2, N3, px, py, pz);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    n1 = (n1_g - 1) / px + 1;
    n2 = (n2_g - 1) / py + 1;
    n3 = (n3_g - 1) / pz + 1;
}

int usage() {
    fprintf(stderr,
            "usage: ark_ref_harness [-nu V] [-cfl V] [-c V] [-quiet] tgv N1 N2 N3 STEPS OUTDIR [DUMPSTEP...]\n"
            "       ark_ref_harness phase NAME INDIR OUTDIR\n"
            "       ark_ref_harness run INDIR STEPS OUTDIR\n"
            "       ark_ref_harness channel STEPS OUTDIR [DUMPSTEP...]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    double nu = 1.0 / 1600.0, cfl = 0.3, c = 10.0;
    bool want_vtk = false;
    int avg_from = -1;
    int a = 1;
    while (a < argc && argv[a][0] == '-') {
        std::string o(argv[a]);
        if (o == "-nu" && a + 1 < argc) nu = strtod(argv[++a], nullptr);
        else if (o == "-cfl" && a + 1 < argc) cfl = strtod(argv[++a], nullptr);
        else if (o == "-c" && a + 1 < argc) c = strtod(argv[++a], nullptr);
        else if (o == "-quiet") g_quiet = true;
        else if (o == "-vtk") want_vtk = true;
        else if (o == "-avg" && a + 1 < argc) avg_from = atoi(argv[++a]);
        else { MPI_Finalize(); return usage(); }
        ++a;
    }
    if (a >= argc) { MPI_Finalize(); return usage(); }
    std::string mode(argv[a++]);
    setup_comm(mode == "channel" ? 0 : 1);
    nStep = 0; TIME = 0.0; l = 1;
    needSwap = true; useTecplot = false; dirPath = "./";
    int rc = 0;

    if (mode == "tgv") {
        if (argc - a < 5) { MPI_Finalize(); return usage(); }
        int N1 = atoi(argv[a]), N2 = atoi(argv[a + 1]), N3 = atoi(argv[a + 2]);
        int steps = atoi(argv[a + 3]);
        std::string out(argv[a + 4]);
        std::vector<int> dumps;
        for (int q = a + 5; q < argc; ++q) dumps.push_back(atoi(argv[q]));
        set_domain(N1, N2, N3);
        sound = c; ro0_g = 1.0; VIS = nu; CFL = cfl; t0 = 1.0;
        x1_w = 0.; x1_e = 2.0 * M_PI; x2_s = 0.; x2_n = 2.0 * M_PI; x3_b = 0.; x3_t = 2.0 * M_PI;
        dx1 = (x1_e - x1_w) / (n1_g - 1);
        dx2 = (x2_n - x2_s) / (n2_g - 1);
        dx3 = (x3_t - x3_b) / (n3_g - 1);
        allocate_state();
        init_tgv(N1, N2, N3);
        coordX = coordY = coordZ = -1;  // periodic box: no physical-boundary branches
        bool single = (nproc == 1);
        FILE* series = nullptr;
        if (rank == 0 && out != "-") {
            std::string fn = out + "/series.txt";
            series = fopen(fn.c_str(), "w");
            if (!series) { fprintf(stderr, "cannot write %s\n", fn.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
            fprintf(series, "# step dt TIME ke_reference mass\n");
            fprintf(series, "0 %a %a %a %a\n", 0.0, 0.0, energy_like_reference(), mass_sum());
        } else {
            energy_like_reference(); mass_sum();
        }
        auto dump_if = [&](int s) {
            if (out == "-") return;
            for (int d : dumps)
                if (d == s) {
                    std::string dir = out + "/s" + std::to_string(s);
                    if (rank == 0) { std::string cmd = "mkdir -p " + dir; if (system(cmd.c_str()) != 0) MPI_Abort(MPI_COMM_WORLD, 2); }
                    MPI_Barrier(newComm);
                    if (single) dump_local(dir, N1, N2, N3);
                    else dump_global(dir, N1, N2, N3);
                }
        };
        dump_if(0);
        double loop_time = 0.0;
        for (int s = 1; s <= steps; ++s) {
            MPI_Barrier(newComm);
            double t_start = MPI_Wtime();
            one_step();
            loop_time += MPI_Wtime() - t_start;
            if (series) {
                double ke = energy_like_reference(), m = mass_sum();
                fprintf(series, "%d %a %a %a %a\n", nStep, dt, TIME, ke, m);
            } else if (out != "-") {
                energy_like_reference(); mass_sum();
            }
            dump_if(s);
        }
        if (series) fclose(series);
        double max_time = 0.0;
        MPI_Allreduce(&loop_time, &max_time, 1, MPI_DOUBLE, MPI_MAX, newComm);
        if (rank == 0) {
            double mcs = (double)N1 * N2 * N3 * steps / (max_time > 0 ? max_time : 1e-30) / 1e6;
            printf("{\"ranks\": %d, \"dims\": [%d, %d, %d], \"grid\": [%d, %d, %d], \"steps\": %d, "
                   "\"loop_seconds\": %.6f, \"mcell_steps_per_s\": %.6f}\n",
                   nproc, px, py, pz, N1, N2, N3, steps, max_time, mcs);
        }
    } else if (mode == "channel") {
        if (nproc != 1) { if (rank == 0) fprintf(stderr, "channel mode needs exactly one rank\n"); MPI_Abort(MPI_COMM_WORLD, 2); }
        if (argc - a < 2) { MPI_Finalize(); return usage(); }
        int steps = atoi(argv[a]);
        std::string out(argv[a + 1]);
        std::vector<int> dumps;
        for (int q = a + 2; q < argc; ++q) dumps.push_back(atoi(argv[q]));
        ark_ref_Input();
        for (int q = 0; q < kNumArrays; ++q) {
            double* p = (*g_arrays[q].arr)->ref();
            for (long e = 0; e < count_of(g_arrays[q]); ++e) p[e] = 0.0;
        }
        int nmax = (int)max3d(n1, n2, n3);
        for (int e = 0; e < nmax + 2; ++e) {
            rBuf[e] = tfBuf[e] = tbBuf[e] = u2fBuf[e] = u2bBuf[e] = u3fBuf[e] = u3bBuf[e] = 0.0;
            if (e < nmax + 1) qBuf[e] = 0.0;
        }
        ark_ref_InitializeData();
        for (int e = 0; e < (n1 + 1) * (n2 + 1) * (n3 + 1); ++e) {
            roCon->ref()[e] = ronCon->ref()[e];
            u1Con->ref()[e] = u1nCon->ref()[e];
            u2Con->ref()[e] = u2nCon->ref()[e];
            u3Con->ref()[e] = u3nCon->ref()[e];
            tCon->ref()[e] = tnCon->ref()[e];
        }
        const int N1 = n1 - 1, N2 = n2 - 1, N3 = n3 - 1;
        static std::string vtk_dir;
        vtk_dir = out + "/";
        dirPath = vtk_dir.c_str();
        auto dump_if = [&](int s) {
            for (int d : dumps)
                if (d == s) {
                    std::string dir = out + "/s" + std::to_string(s);
                    std::string cmd = "mkdir -p " + dir;
                    if (system(cmd.c_str()) != 0) MPI_Abort(MPI_COMM_WORLD, 2);
                    dump_local(dir, N1, N2, N3);
                    if (want_vtk) WriteDataParaView();
                }
        };
        for (int k = 0; k <= n3; ++k)
            for (int j = 0; j <= n2 + 1; ++j)
                for (int i = 0; i <= n1; ++i)
                    averageU1->elem(i, j, k) = averageU2->elem(i, j, k) = averageU3->elem(i, j, k) = 0.0;
        double start_time = 0.0;
        std::string fn = out + "/series.txt";
        std::string mk = "mkdir -p " + out;
        if (system(mk.c_str()) != 0) MPI_Abort(MPI_COMM_WORLD, 2);
        FILE* series = fopen(fn.c_str(), "w");
        if (!series) { fprintf(stderr, "cannot write %s\n", fn.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
        fprintf(series, "# step dt TIME ke_reference mass\n");
        fprintf(series, "0 %a %a %a %a\n", 0.0, 0.0, energy_like_reference(), mass_sum());
        dump_if(0);
        double loop_time = 0.0;
        for (int s = 1; s <= steps; ++s) {
            double t_start = MPI_Wtime();
            one_step();
            loop_time += MPI_Wtime() - t_start;
            if (avg_from >= 0 && s >= avg_from) {
                if (s == avg_from) start_time = TIME;
                AverageValues();
            }
            fprintf(series, "%d %a %a %a %a\n", nStep, dt, TIME, energy_like_reference(), mass_sum());
            dump_if(s);
        }
        fclose(series);
        if (avg_from >= 0) WriteAverageValues(start_time);
        printf("{\"ranks\": 1, \"grid\": [%d, %d, %d], \"steps\": %d, \"loop_seconds\": %.6f, "
               "\"mcell_steps_per_s\": %.6f}\n",
               N1, N2, N3, steps, loop_time, (double)N1 * N2 * N3 * steps / (loop_time > 0 ? loop_time : 1e-30) / 1e6);
    } else if (mode == "phase" || mode == "run") {
        if (nproc != 1) { if (rank == 0) fprintf(stderr, "%s mode needs exactly one rank\n", mode.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
        std::string name, in, out;
        int steps = 0;
        if (mode == "phase") {
            if (argc - a < 3) { MPI_Finalize(); return usage(); }
            name = argv[a]; in = argv[a + 1]; out = argv[a + 2];
        } else {
            if (argc - a < 3) { MPI_Finalize(); return usage(); }
            in = argv[a]; steps = atoi(argv[a + 1]); out = argv[a + 2];
        }
        int N1 = 0, N2 = 0, N3 = 0;
        if (!read_meta(in, &N1, &N2, &N3)) { fprintf(stderr, "cannot read %s/meta.txt\n", in.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
        set_domain(N1, N2, N3);
        x1_w = 0.; x1_e = N1 * dx1; x2_s = 0.; x2_n = N2 * dx2; x3_b = 0.; x3_t = N3 * dx3;
        allocate_state();
        if (!load_local(in)) MPI_Abort(MPI_COMM_WORLD, 2);
        coordX = coordY = coordZ = -1;
        if (mode == "phase") {
            if (name == "TimeStepSize") TimeStepSize();
            else if (name == "Phase1") Phase1();
            else if (name == "StressTensor") StressTensor();
            else if (name == "UseForces") UseForces();
            else if (name == "Phase2") Phase2();
            else if (name == "Step") one_step();
            else { fprintf(stderr, "unknown phase %s\n", name.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
        } else {
            for (int s = 0; s < steps; ++s) one_step();
        }
        dump_local(out, N1, N2, N3);
    } else {
        rc = usage();
    }
    MPI_Finalize();
    return rc;
}

// The reference's output path: WriteDataParaView, AverageValues and
// WriteAverageValues (source/main.cpp:2577-2891).  The fields are computed
// on the device (vorticity, running averages); the host writes the same
// bytes the reference's single rank writes: binary legacy VTK, rectilinear
// grid, big-endian when need_swap (the reference sets needSwap = true,
// source/main.cpp:120; its swap8, 2975-2988).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ark_internal.hpp"
#include "ark_kernels.hpp"

using namespace arkd;

namespace {

void put_double(std::vector<char>& out, double v, bool swap) {
    char b[8];
    std::memcpy(b, &v, 8);
    if (swap)
        for (int q = 0; q < 4; ++q) std::swap(b[q], b[7 - q]);
    out.insert(out.end(), b, b + 8);
}

void put_text(std::vector<char>& out, const char* s) { out.insert(out.end(), s, s + std::strlen(s)); }

// Interior values of a device field in the common index map: (N1*N2*N3),
// x fastest.
ark_status interior_of(ark_ctx* c, const double* dev, std::vector<double>& out) {
    std::vector<double> all(c->elems);   // the whole allocation, from kOriginPad before element (0,0,0)
    ark_status s = c->hip(hipMemcpyAsync(all.data(), dev - kOriginPad, c->elems * sizeof(double),
                                         hipMemcpyDeviceToHost, c->stream), "copy field");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s != ARK_OK) return s;
    const Geo& g = c->geo;
    out.resize((size_t)g.N1 * g.N2 * g.N3);
    size_t q = 0;
    for (int k = 1; k <= g.N3; ++k)
        for (int j = 1; j <= g.N2; ++j)
            for (int i = 1; i <= g.N1; ++i) out[q++] = all[kOriginPad + i + (size_t)j * g.px + (size_t)k * g.sxy];
    return ARK_OK;
}

// Header and coordinates shared by both files (source/main.cpp:2583-2641,
// 2753-2811).  One rank: n_g = N + 1 nodes per axis, node i at (i-1) dx.
void header_and_grid(const ark_ctx* c, std::vector<char>& out, bool swap) {
    const int n1g = c->geo.N1 + 1, n2g = c->geo.N2 + 1, n3g = c->geo.N3 + 1;
    char buf[256];
    put_text(out, "# vtk DataFile Version 3.0\nvtk output\nBINARY\n");
    std::snprintf(buf, sizeof buf, "DATASET RECTILINEAR_GRID\nDIMENSIONS %d %d %d", n1g, n2g, n3g);
    put_text(out, buf);
    const int n[3] = {n1g, n2g, n3g};
    const double dx[3] = {c->prm.dx1, c->prm.dx2, c->prm.dx3};
    const char* name[3] = {"X", "Y", "Z"};
    for (int d = 0; d < 3; ++d) {
        std::snprintf(buf, sizeof buf, "\n%s_COORDINATES %d double\n", name[d], n[d]);
        put_text(out, buf);
        for (int i = 1; i <= n[d]; ++i) put_double(out, (i - 1) * dx[d], swap);
    }
    std::snprintf(buf, sizeof buf, "\nCELL_DATA %d\nVECTORS U double\n", (n1g - 1) * (n2g - 1) * (n3g - 1));
    put_text(out, buf);
}

ark_status write_file(ark_ctx* c, const std::string& path, const std::vector<char>& bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return c->fail(ARK_ERR_INVALID, "cannot open " + path);
    const size_t w = std::fwrite(bytes.data(), 1, bytes.size(), f);
    const int e = std::fclose(f);
    if (w != bytes.size() || e != 0) return c->fail(ARK_ERR_INVALID, "cannot write " + path);
    return ARK_OK;
}

ark_status scalars_of(ark_ctx* c, Scalars& sc) {
    ark_status s = c->hip(hipMemcpyAsync(&sc, c->S, sizeof sc, hipMemcpyDeviceToHost, c->stream), "scalars");
    return s == ARK_OK ? c->hip(hipStreamSynchronize(c->stream), "sync") : s;
}

struct DevTemp {
    double* p = nullptr;
    ~DevTemp() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" {

ark_status ark_average_values(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    if (!c->average) {
        double* p = nullptr;
        ark_status s = c->hip(hipMalloc(&p, 3 * c->elems * sizeof(double)), "hipMalloc averages");
        if (s != ARK_OK) return s;
        c->allocs.push_back(p);
        c->bytes += 3 * c->elems * sizeof(double);
        s = c->hip(hipMemsetAsync(p, 0, 3 * c->elems * sizeof(double), c->stream), "zero averages");
        if (s != ARK_OK) return s;
        c->average = p + kOriginPad;
    }
    double* a = c->average;
    return c->hip(launch_average(c->stream, c->geo, c->fields(), c->S, a, a + c->elems, a + 2 * c->elems),
                  "average");
}

ark_status ark_write_paraview(ark_ctx* c, const char* dir_path, int need_swap) {
    if (!c || !dir_path) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    Scalars sc;
    ark_status s = scalars_of(c, sc);
    if (s != ARK_OK) return s;
    std::vector<double> u[3], ro, w[3];
    const int cells[3] = {F_u1nCon, F_u2nCon, F_u3nCon};
    for (int m = 0; m < 3 && s == ARK_OK; ++m) s = interior_of(c, c->cur[cells[m]], u[m]);
    if (s == ARK_OK) s = interior_of(c, c->cur[F_ronCon], ro);
    DevTemp t;
    if (s == ARK_OK) s = c->hip(hipMalloc(&t.p, 3 * c->elems * sizeof(double)), "hipMalloc vorticity");
    if (s == ARK_OK) {
        double* o = t.p + kOriginPad;
        if (c->prm.precision == ARK_PREC_MIXED)
            s = ark_mixed_vorticity(c, o, o + c->elems, o + 2 * c->elems);
        else
            s = c->hip(launch_vorticity(c->stream, c->geo, c->fields(), c->K, o, o + c->elems, o + 2 * c->elems),
                       "vorticity");
        for (int m = 0; m < 3 && s == ARK_OK; ++m) s = interior_of(c, o + m * c->elems, w[m]);
    }
    if (s != ARK_OK) return s;
    const bool swap = need_swap != 0;
    std::vector<char> out;
    header_and_grid(c, out, swap);
    const size_t n = ro.size();
    for (size_t q = 0; q < n; ++q)
        for (int m = 0; m < 3; ++m) put_double(out, u[m][q], swap);
    put_text(out, "\nscalars PC double\nLOOKUP_TABLE default\n");
    for (size_t q = 0; q < n; ++q) put_double(out, ro[q], swap);
    put_text(out, "\nVECTORS R double\n");
    for (size_t q = 0; q < n; ++q)
        for (int m = 0; m < 3; ++m) put_double(out, w[m][q], swap);
    char name[64];
    std::snprintf(name, sizeof name, "out_%d.vtk", (int)sc.nstep);
    return write_file(c, std::string(dir_path) + name, out);
}

ark_status ark_write_average_values(ark_ctx* c, const char* dir_path, double start_time, int need_swap) {
    if (!c || !dir_path) return ARK_ERR_INVALID;
    if (!c->average) return c->fail(ARK_ERR_STATE, "write_average_values: no ark_average_values call yet");
    Scalars sc;
    ark_status s = scalars_of(c, sc);
    std::vector<double> a[3];
    for (int m = 0; m < 3 && s == ARK_OK; ++m) s = interior_of(c, c->average + m * c->elems, a[m]);
    if (s != ARK_OK) return s;
    const bool swap = need_swap != 0;
    std::vector<char> out;
    header_and_grid(c, out, swap);
    const double span = sc.time - start_time;
    for (size_t q = 0; q < a[0].size(); ++q)
        for (int m = 0; m < 3; ++m) put_double(out, a[m][q] / span, swap);
    return write_file(c, std::string(dir_path) + "average.vtk", out);
}

}  // extern "C"

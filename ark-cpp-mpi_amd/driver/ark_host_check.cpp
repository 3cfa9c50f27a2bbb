// Exercises the C++ host interface (include/ark/ark_host.hpp) the way a
// reference-side program uses it; the GPU tests compare its outputs with the
// reference's.
//
//   ark_host_check bc N1 N2 N3 AXIS PERIODIC T0 IN OUT
//       the command line of the reference BoundaryConditions driver: IN holds
//       five cell arrays u1, u2, u3, ro, t (Arr3d order, float64); they are
//       bound to u1nCon, u2nCon, u3nCon, ronCon, tnCon of a Solver, uploaded,
//       BoundaryConditions(..., n_d = N_d + 1, t0).alongX{AXIS+1}() runs on the
//       device and the downloaded arrays go to OUT.
//   ark_host_check plane N1 N2 N3 FIELD AXIS Q IN OUT
//       IN holds one array of field FIELD (ark_field id); it is uploaded, plane
//       Q normal to AXIS is packed into an Arr2d and written to OUT, then the
//       Arr2d, negated, is unpacked into plane Q and the whole array is
//       appended to OUT.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "ark/ark_host.hpp"

namespace {

int usage() {
    fprintf(stderr,
            "usage: ark_host_check bc N1 N2 N3 AXIS PERIODIC T0 IN OUT\n"
            "       ark_host_check plane N1 N2 N3 FIELD AXIS Q IN OUT\n");
    return 2;
}

bool read_into(FILE* fh, ark::Arr3d& a) {
    const size_t n = (size_t)a.n1() * a.n2() * a.n3();
    return fread(a.ref(), sizeof(double), n, fh) == n;
}

void write_from(FILE* fh, ark::Arr3d& a) { fwrite(a.ref(), sizeof(double), (size_t)a.n1() * a.n2() * a.n3(), fh); }

int run_bc(int N1, int N2, int N3, int axis, bool periodic, double t0, const char* in, const char* out) {
    ark::Solver s(ark::Solver::periodic_box(N1, N2, N3, 1.0, 1.0, 1.0, 0.3, 1e-3, 10.0));
    const int fields[5] = {ARK_u1nCon, ARK_u2nCon, ARK_u3nCon, ARK_ronCon, ARK_tnCon};
    std::vector<ark::Arr3d*> a;
    FILE* fh = fopen(in, "rb");
    if (!fh) return 1;
    for (int q = 0; q < 5; ++q) {
        a.push_back(new ark::Arr3d(N1 + 2, N2 + 2, N3 + 2));
        if (!read_into(fh, *a[q])) return 1;
        a[q]->bind(&s, fields[q]);
        a[q]->upload();
    }
    fclose(fh);
    ark::BoundaryConditions bc(a[0], a[1], a[2], a[3], a[4], periodic, periodic, periodic, N1 + 1, N2 + 1, N3 + 1,
                               t0);
    if (axis == 0) bc.alongX1();
    else if (axis == 1) bc.alongX2();
    else bc.alongX3();
    fh = fopen(out, "wb");
    if (!fh) return 1;
    for (int q = 0; q < 5; ++q) {
        a[q]->download();
        write_from(fh, *a[q]);
        delete a[q];
    }
    fclose(fh);
    return 0;
}

int run_plane(int N1, int N2, int N3, int field, int axis, int q, const char* in, const char* out) {
    ark::Solver s(ark::Solver::periodic_box(N1, N2, N3, 1.0, 1.0, 1.0, 0.3, 1e-3, 10.0));
    ark::Arr3d a(s.extent(field, 0), s.extent(field, 1), s.extent(field, 2));
    FILE* fh = fopen(in, "rb");
    if (!fh || !read_into(fh, a)) return 1;
    fclose(fh);
    a.bind(&s, field);
    a.upload();
    const int na = axis == 0 ? N2 : N1, nb = axis == 2 ? N2 : N3;
    ark::Arr2d buf(na, nb);
    buf.pack(a, axis, q);
    fh = fopen(out, "wb");
    if (!fh) return 1;
    fwrite(buf.ref(), sizeof(double), (size_t)na * nb, fh);
    for (int j = 0; j < nb; ++j)
        for (int i = 0; i < na; ++i) buf.elem(i, j) = -buf.elem(i, j);
    buf.unpack(a, axis, q);
    a.download();
    write_from(fh, a);
    fclose(fh);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc == 10 && !strcmp(argv[1], "bc"))
            return run_bc(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]) != 0,
                          atof(argv[7]), argv[8], argv[9]);
        if (argc == 10 && !strcmp(argv[1], "plane"))
            return run_plane(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]),
                             atoi(argv[7]), argv[8], argv[9]);
    } catch (const ark::Error& e) {
        fprintf(stderr, "ark_host_check: %s (status %d)\n", e.what(), (int)e.status);
        return 1;
    }
    return usage();
}

// Test infrastructure only: drives the UNMODIFIED reference BoundaryConditions
// class (source/BoundaryConditions.cpp, with source/Arr3d.cpp), compiled from
// the reference checkout by oracle/Makefile.  Nothing in the product links or
// runs it.
//
//   ark_ref_bc N1 N2 N3 AXIS PERIODIC T0 IN OUT
//
// IN holds five cell arrays u1, u2, u3, ro, t (float64, little endian, each
// (N1+2)*(N2+2)*(N3+2) values in the reference's Arr3d order, x fastest).  The
// program constructs BoundaryConditions with node counts n_d = N_d + 1 and
// the period flag PERIODIC on every axis, calls alongX{AXIS+1}() and writes
// the five arrays to OUT.
#include <stdio.h>
#include <stdlib.h>

#include "Arr3d.h"
#include "BoundaryConditions.h"

int main(int argc, char** argv) {
    if (argc != 9) {
        fprintf(stderr, "usage: ark_ref_bc N1 N2 N3 AXIS PERIODIC T0 IN OUT\n");
        return 2;
    }
    const int N1 = atoi(argv[1]), N2 = atoi(argv[2]), N3 = atoi(argv[3]);
    const int axis = atoi(argv[4]);
    const bool periodic = atoi(argv[5]) != 0;
    const double t0 = atof(argv[6]);
    const long n = (long)(N1 + 2) * (N2 + 2) * (N3 + 2);
    Arr3d* a[5];
    FILE* in = fopen(argv[7], "rb");
    if (!in) return 1;
    for (int q = 0; q < 5; ++q) {
        a[q] = new Arr3d(N1 + 2, N2 + 2, N3 + 2);
        if (fread(a[q]->ref(), sizeof(double), n, in) != (size_t)n) return 1;
    }
    fclose(in);
    BoundaryConditions bc(a[0], a[1], a[2], a[3], a[4], periodic, periodic, periodic, N1 + 1, N2 + 1, N3 + 1, t0);
    if (axis == 0) bc.alongX1();
    else if (axis == 1) bc.alongX2();
    else bc.alongX3();
    FILE* out = fopen(argv[8], "wb");
    if (!out) return 1;
    for (int q = 0; q < 5; ++q) fwrite(a[q]->ref(), sizeof(double), n, out);
    fclose(out);
    for (int q = 0; q < 5; ++q) delete a[q];
    return 0;
}

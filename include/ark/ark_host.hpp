/* C++ host interface of the MI355X CABARET/ARK step.
 *
 * The reference's host objects, with their interfaces kept, over the C-ABI of
 * ark/ark.h:
 *   Arr3d              headers/Arr3d.h:9-18, source/Arr3d.cpp:7-25
 *   Arr2d              headers/Arr2d.h:9-18, source/Arr2d.cpp:7-24
 *   BoundaryConditions headers/BoundaryConditions.h:11-23
 * and the phase-by-phase step (TimeStepSize, Phase1, StressTensor, UseForces,
 * Phase2; source/main.cpp:12-17) as methods of Solver, which owns one rank's
 * subdomain on the GPU.
 *
 * The fields themselves live on the device.  An Arr3d keeps the reference's
 * host storage and index map; bound to a field of a Solver (Arr3d::bind) it is
 * that field's host copy, moved by upload() and download().  Arr2d is the
 * halo-plane buffer of the reference's MPI exchange; pack() and unpack() move
 * one plane of a bound Arr3d through it.  BoundaryConditions keeps the
 * reference's constructor; its alongX1/2/3() run the reference's boundary
 * loops on the device copies of the five bound arrays.
 *
 * Errors from the library are thrown as ark::Error (the reference has none:
 * MPI's handler aborts).
 */
#ifndef ARK_ARK_HOST_HPP
#define ARK_ARK_HOST_HPP

#include <stddef.h>

#include <stdexcept>
#include <string>

#include "ark/ark.h"

namespace ark {

class Solver;

class Error : public std::runtime_error {
public:
    Error(ark_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    ark_status status;
};

/* 3-D array of doubles with the reference's layout: elem(i,j,k) is
 * ref()[k*n1*n2 + j*n1 + i] (source/Arr3d.cpp:19-21). */
class Arr3d {
public:
    Arr3d(int n1, int n2, int n3);
    ~Arr3d();
    Arr3d(const Arr3d&) = delete;
    Arr3d& operator=(const Arr3d&) = delete;

    double& elem(int i, int j, int k) { return _arr[((size_t)k * _n2 + j) * _n1 + i]; }
    double* ref() { return _arr; }
    int n1() const { return _n1; }
    int n2() const { return _n2; }
    int n3() const { return _n3; }

    /* Make this array the host copy of `field` (ark_field) of `s`.  The
     * extents must be the reference's for that field. */
    void bind(Solver* s, int field);
    Solver* solver() const { return _s; }
    int field() const { return _field; }
    void upload();     /* host -> device (ark_upload) */
    void download();   /* device -> host (ark_download) */

private:
    int _n1, _n2, _n3;
    double* _arr;
    Solver* _s = nullptr;
    int _field = -1;
};

/* 2-D array of doubles, elem(i,j) = ref()[j*n1 + i] (source/Arr2d.cpp:14-16):
 * one halo plane. */
class Arr2d {
public:
    Arr2d(int n1, int n2);
    ~Arr2d();
    Arr2d(const Arr2d&) = delete;
    Arr2d& operator=(const Arr2d&) = delete;

    double& elem(int i, int j) { return _arr[(size_t)j * _n1 + i]; }
    double* ref() { return _arr; }
    int n1() const { return _n1; }
    int n2() const { return _n2; }

    /* Plane q normal to `axis` of the bound array a (interior of the other
     * two axes), as the reference packs its send buffers (e.g.
     * source/main.cpp:619-630) and unpacks its receive buffers (642-654). */
    void pack(const Arr3d& a, int axis, int q);
    void unpack(Arr3d& a, int axis, int q) const;

private:
    int _n1, _n2;
    double* _arr;
};

/* One rank's subdomain on one GPU. */
class Solver {
public:
    explicit Solver(const ark_params& p);
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    /* A periodic box of N1 x N2 x N3 cells and side lengths L1, L2, L3 with
     * the given physical constants (FP64, reference scheme unless set). */
    static ark_params periodic_box(int N1, int N2, int N3, double L1, double L2, double L3, double CFL,
                                   double VIS, double sound, double ro0 = 1.0, double t0 = 1.0);
    /* The reference's own problem as its Input() sets it up
     * (source/main.cpp:118-195): n1_g x n2_g x n3_g nodes on [0, x1_e] x
     * [0, x2_n] x [0, x3_t], no periodic axis.  Fill it with InitChannel. */
    static ark_params channel(int n1_g = 201, int n2_g = 61, int n3_g = 2, double x1_e = 20.0, double x2_n = 6.0,
                              double x3_t = 0.1, double CFL = 0.2, double VIS = 1.0 / 1000.0, double sound = 10.0);

    /* The reference's phase functions (source/main.cpp:374-2495). */
    void TimeStepSize();
    /* nStep++ and TIME += 0.5 * dt of the reference's loop (83-84, 91). */
    void NextStep();
    void HalfTime();
    void Phase1();
    void StressTensor();
    void UseForces();
    void Phase2();

    /* n full steps of the driver loop (source/main.cpp:80-104), fused. */
    void Step(int n = 1);
    /* n fused steps that keep only the persistent state current. */
    void Advance(int n);
    /* n full steps through the phase functions. */
    void StepPhases(int n);

    void InitTGV(double U0 = 1.0);
    /* InitializeData (source/main.cpp:280-372) of a channel() context. */
    void InitChannel(double u10 = 1.0);

    /* The output path (source/main.cpp:2577-2891): <dir>out_<nStep>.vtk,
     * the running averages, and <dir>average.vtk; need_swap as the
     * reference's needSwap (big-endian files). */
    void WriteDataParaView(const std::string& dir, bool need_swap = true);
    void AverageValues();
    void WriteAverageValues(const std::string& dir, double start_time, bool need_swap = true);
    ark_diag Diagnostics();
    double dt();
    void SetDt(double dt);
    void Synchronize();

    /* Reference extent of `field` along axis 0..2 (cells N+2, faces N+3). */
    int extent(int field, int axis) const;
    int N(int axis) const { return axis == 0 ? _p.N1 : axis == 1 ? _p.N2 : _p.N3; }
    const ark_params& params() const { return _p; }
    ark_ctx* ctx() { return _c; }
    /* Throws ark::Error for a status other than ARK_OK. */
    void check(ark_status s, const char* what) const;

private:
    ark_ctx* _c = nullptr;
    ark_params _p;
};

/* BoundaryConditions (headers/BoundaryConditions.h:11-23): the reference's
 * constructor, with n1..n3 the node counts (cells + 1).  The five arrays must
 * be bound to cell fields of one Solver whose box has those node counts. */
class BoundaryConditions {
public:
    BoundaryConditions(Arr3d* u1, Arr3d* u2, Arr3d* u3, Arr3d* ro, Arr3d* t, bool x1Period, bool x2Period,
                       bool x3Period, int n1, int n2, int n3, double t0);
    void alongX1();
    void alongX2();
    void alongX3();

private:
    void along(int axis, bool periodic);
    Arr3d* _a[5];
    bool _period[3];
    int _n[3];
    double _t0;
};

}  // namespace ark

#endif /* ARK_ARK_HOST_HPP */

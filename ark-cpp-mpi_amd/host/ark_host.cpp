// C++ host interface (include/ark/ark_host.hpp) over the C-ABI.
#include "ark/ark_host.hpp"

#include <string.h>

#include <string>

namespace ark {

// ---------------------------------------------------------------------------
Arr3d::Arr3d(int n1, int n2, int n3) : _n1(n1), _n2(n2), _n3(n3) {
    if (n1 <= 0 || n2 <= 0 || n3 <= 0) throw Error(ARK_ERR_INVALID, "Arr3d: extents must be positive");
    _arr = new double[(size_t)n1 * n2 * n3]();
}

Arr3d::~Arr3d() { delete[] _arr; }

void Arr3d::bind(Solver* s, int field) {
    if (!s || field < 0 || field >= ARK_NUM_FIELDS) throw Error(ARK_ERR_INVALID, "Arr3d::bind: bad solver or field");
    if (_n1 != s->extent(field, 0) || _n2 != s->extent(field, 1) || _n3 != s->extent(field, 2))
        throw Error(ARK_ERR_INVALID, "Arr3d::bind: extents differ from the field's");
    _s = s;
    _field = field;
}

void Arr3d::upload() {
    if (!_s) throw Error(ARK_ERR_STATE, "Arr3d::upload: not bound to a solver field");
    _s->check(ark_upload(_s->ctx(), _field, _arr, _n1, _n2, _n3), "ark_upload");
}

void Arr3d::download() {
    if (!_s) throw Error(ARK_ERR_STATE, "Arr3d::download: not bound to a solver field");
    _s->check(ark_download(_s->ctx(), _field, _arr, _n1, _n2, _n3), "ark_download");
}

// ---------------------------------------------------------------------------
Arr2d::Arr2d(int n1, int n2) : _n1(n1), _n2(n2) {
    if (n1 <= 0 || n2 <= 0) throw Error(ARK_ERR_INVALID, "Arr2d: extents must be positive");
    _arr = new double[(size_t)n1 * n2]();
}

Arr2d::~Arr2d() { delete[] _arr; }

void Arr2d::pack(const Arr3d& a, int axis, int q) {
    Solver* s = a.solver();
    if (!s) throw Error(ARK_ERR_STATE, "Arr2d::pack: array not bound to a solver field");
    s->check(ark_get_plane(s->ctx(), a.field(), axis, q, _arr, _n1, _n2), "ark_get_plane");
}

void Arr2d::unpack(Arr3d& a, int axis, int q) const {
    Solver* s = a.solver();
    if (!s) throw Error(ARK_ERR_STATE, "Arr2d::unpack: array not bound to a solver field");
    s->check(ark_set_plane(s->ctx(), a.field(), axis, q, _arr, _n1, _n2), "ark_set_plane");
}

// ---------------------------------------------------------------------------
Solver::Solver(const ark_params& p) : _p(p) {
    ark_status s = ark_create(&p, &_c);
    if (s != ARK_OK) {
        std::string msg = std::string("ark_create: ") + (_c ? ark_last_error(_c) : "failed");
        if (_c) ark_destroy(_c);
        _c = nullptr;
        throw Error(s, msg);
    }
}

Solver::~Solver() {
    if (_c) ark_destroy(_c);
}

ark_params Solver::periodic_box(int N1, int N2, int N3, double L1, double L2, double L3, double CFL, double VIS,
                                double sound, double ro0, double t0) {
    ark_params p;
    memset(&p, 0, sizeof p);
    p.N1 = N1;
    p.N2 = N2;
    p.N3 = N3;
    // node spacing of the reference: dx = (x_e - x_w)/(n - 1) (source/main.cpp:291-293)
    p.dx1 = L1 / N1;
    p.dx2 = L2 / N2;
    p.dx3 = L3 / N3;
    p.CFL = CFL;
    p.VIS = VIS;
    p.sound = sound;
    p.ro0 = ro0;
    p.t0 = t0;
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 1;
    p.precision = ARK_PREC_FP64;
    p.scheme = ARK_SCHEME_REFERENCE;
    p.device = 0;
    return p;
}

ark_params Solver::channel(int n1_g, int n2_g, int n3_g, double x1_e, double x2_n, double x3_t, double CFL,
                           double VIS, double sound) {
    // node spacing dx = (x_e - x_w)/(n - 1) with x_w = 0 (source/main.cpp:291-293)
    ark_params p = periodic_box(n1_g - 1, n2_g - 1, n3_g - 1, x1_e - 0., x2_n - 0., x3_t - 0., CFL, VIS, sound);
    p.periodic[0] = p.periodic[1] = p.periodic[2] = 0;
    return p;
}

void Solver::check(ark_status s, const char* what) const {
    if (s != ARK_OK) throw Error(s, std::string(what) + ": " + (_c ? ark_last_error(_c) : ""));
}

void Solver::TimeStepSize() { check(ark_time_step_size(_c, nullptr), "TimeStepSize"); }
void Solver::NextStep() { check(ark_next_step(_c), "NextStep"); }
void Solver::HalfTime() { check(ark_half_time(_c), "HalfTime"); }
void Solver::Phase1() { check(ark_phase1(_c), "Phase1"); }
void Solver::StressTensor() { check(ark_stress_tensor(_c), "StressTensor"); }
void Solver::UseForces() { check(ark_use_forces(_c), "UseForces"); }
void Solver::Phase2() { check(ark_phase2(_c), "Phase2"); }
void Solver::Step(int n) { check(ark_step(_c, n), "Step"); }
void Solver::Advance(int n) { check(ark_advance(_c, n), "Advance"); }
void Solver::StepPhases(int n) { check(ark_step_phases(_c, n), "StepPhases"); }
void Solver::InitTGV(double U0) { check(ark_init_tgv(_c, U0), "InitTGV"); }
void Solver::InitChannel(double u10) { check(ark_init_channel(_c, u10), "InitChannel"); }
void Solver::WriteDataParaView(const std::string& dir, bool need_swap) {
    check(ark_write_paraview(_c, dir.c_str(), need_swap ? 1 : 0), "WriteDataParaView");
}
void Solver::AverageValues() { check(ark_average_values(_c), "AverageValues"); }
void Solver::WriteAverageValues(const std::string& dir, double start_time, bool need_swap) {
    check(ark_write_average_values(_c, dir.c_str(), start_time, need_swap ? 1 : 0), "WriteAverageValues");
}
void Solver::Synchronize() { check(ark_synchronize(_c), "Synchronize"); }
void Solver::SetDt(double dt) { check(ark_set_dt(_c, dt), "SetDt"); }

double Solver::dt() {
    double v = 0.0;
    check(ark_get_dt(_c, &v), "dt");
    return v;
}

ark_diag Solver::Diagnostics() {
    ark_diag d;
    check(ark_diagnostics(_c, &d), "Diagnostics");
    return d;
}

int Solver::extent(int field, int axis) const {
    const int e = field >= ARK_ro1 ? 3 : 2;   // face and stress fields (N+3), cell fields (N+2)
    return N(axis) + e;
}

// ---------------------------------------------------------------------------
BoundaryConditions::BoundaryConditions(Arr3d* u1, Arr3d* u2, Arr3d* u3, Arr3d* ro, Arr3d* t, bool x1Period,
                                       bool x2Period, bool x3Period, int n1, int n2, int n3, double t0)
    : _a{u1, u2, u3, ro, t}, _period{x1Period, x2Period, x3Period}, _n{n1, n2, n3}, _t0(t0) {}

void BoundaryConditions::along(int axis, bool periodic) {
    Solver* s = _a[0] ? _a[0]->solver() : nullptr;
    int f[5];
    for (int q = 0; q < 5; ++q) {
        if (!_a[q] || !_a[q]->solver() || _a[q]->solver() != s)
            throw Error(ARK_ERR_STATE, "BoundaryConditions: the five arrays must be bound to one solver");
        f[q] = _a[q]->field();
    }
    for (int d = 0; d < 3; ++d)
        if (_n[d] != s->N(d) + 1)
            throw Error(ARK_ERR_INVALID, "BoundaryConditions: node counts differ from the solver's box");
    s->check(ark_boundary_conditions(s->ctx(), axis, f, periodic ? 1 : 0, _t0), "BoundaryConditions");
}

void BoundaryConditions::alongX1() { along(0, _period[0]); }
void BoundaryConditions::alongX2() { along(1, _period[1]); }
void BoundaryConditions::alongX3() { along(2, _period[2]); }

}  // namespace ark

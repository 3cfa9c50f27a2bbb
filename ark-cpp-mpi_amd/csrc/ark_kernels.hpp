// Kernel argument blocks and host-side launchers of the gfx950 kernels.
#pragma once

#include "ark_device.hpp"

namespace arkd {

// A list of arrays to ghost-fill: axis_mask bit d set = fill along axis d.
// The arrays of one ghost fill and, per array, the axes (bit d) that are
// copied periodically, set to zero, or given zero-gradient ends.  An axis is
// in at most one of the three masks.
struct ExchangePlan {
    double* p[24];
    int is_face[24];
    int copy_mask[24], zero_mask[24], grad_mask[24];
    int n;
};

// The next-buffer pointers of the six face fields along one axis, in the
// order normal velocity, tangential a, tangential b, p, rho, t.
struct NextFaces {
    double* p[6];
};

// 1-D Taylor-Green tables per axis, indices 0..N_d+2: sin/cos/cos2 of the
// cell-centre coordinate (c) and of the face coordinate (f).
struct TgvTables {
    const double *sc[3], *cc[3], *c2c[3];
    const double *sf[3], *cf[3], *c2f[3];
};

// Launchers (ark_kernels.hip).  Every one enqueues on `st` and returns the
// launch status.
hipError_t launch_cfl(hipStream_t st, const Geo& g, const Fields& F, double c1, double c2, double c3, double sound,
                      Scalars* S);
hipError_t launch_cfl_finalize(hipStream_t st, Scalars* S, int advance);
hipError_t launch_clock(hipStream_t st, Scalars* S, int what);   // 0: nStep++, 1: TIME += 0.5 dt
hipError_t launch_phase1(hipStream_t st, const Geo& g, const Fields& F, const Consts& K, const Scalars* S);
hipError_t launch_exchange(hipStream_t st, const Geo& g, const ExchangePlan& plan);
hipError_t launch_vorticity(hipStream_t st, const Geo& g, const Fields& F, const Consts& K,
                            double* o1, double* o2, double* o3);
hipError_t launch_average(hipStream_t st, const Geo& g, const Fields& F, const Scalars* S, double* a1, double* a2,
                          double* a3);
struct BcFields { double* f[5]; };   // u1, u2, u3, ro, t
hipError_t launch_boundary(hipStream_t st, const Geo& g, const BcFields& F, int axis, int periodic, double t0);
hipError_t launch_plane(hipStream_t st, const Geo& g, double* field, int axis, int q, double* buf, int to_buf);
hipError_t launch_copy_rows(hipStream_t st, const double* src, long long s_row, long long s_plane, double* dst,
                            long long d_row, long long d_plane, int e1, int e2, int e3);
hipError_t launch_stress_faces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K);
// Zero narr arrays of n elements each; dev_ptrs is a device array of pointers.
hipError_t launch_zero(hipStream_t st, double* const* dev_ptrs, int narr, long long n);
hipError_t launch_forces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K);
hipError_t launch_use_forces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K, const Scalars* S);
// The three Phase2 sweeps in one launch (nx[d]: the next faces of axis d).
hipError_t launch_phase2_xyz(hipStream_t st, int scheme, const Geo& g, const Fields& F, const NextFaces nx[3],
                             const Consts& K, const Scalars* S, int inflow);
// Force then UseForces, and Phase1 then UseForces, per cell in one launch.
hipError_t launch_forces_use_forces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K,
                                   const Scalars* S);
hipError_t launch_phase1_use_forces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K,
                                    const Scalars* S);
   // inflow: non-periodic X1 ends
hipError_t launch_rho_faces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K);
hipError_t launch_diag(hipStream_t st, const Geo& g, const Fields& F, const Consts& K, double* part, int nblocks,
                       double* out);
hipError_t launch_init_tgv(hipStream_t st, const Geo& g, const Fields& F, const Fields& F2, const TgvTables& T,
                           double U0, double amp, const Consts& K);

}  // namespace arkd

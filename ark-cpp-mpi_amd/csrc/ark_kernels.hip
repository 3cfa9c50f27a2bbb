// Phase-by-phase gfx950 kernels of the CABARET/ARK step: one kernel (or a
// short sequence) per reference phase function.  They define the semantics
// the fused step kernels are checked against, and they back the per-phase
// entry points of the C-ABI.
//
// Thread mapping: x-fastest everywhere.  A block is 64 x 4 threads over
// (i, j) of one k-plane, so a wavefront reads 64 consecutive doubles of each
// field it touches (coalesced 512-byte rows).
#include "ark_kernels.hpp"

namespace arkd {

// ---------------------------------------------------------------------------
// TimeStepSize (source/main.cpp:374-415).  Per cell: min3d of the three
// convective bounds; across cells a MIN that ignores NaN, which is what the
// reference's sequential min3d(dt, dtu, dtv) does.  Positive doubles order
// like their bit patterns as unsigned integers (NaN patterns sort above +Inf),
// so the MIN is an integer atomicMin on the bits.
__global__ void __launch_bounds__(256) k_cfl(Geo g, const double* __restrict__ u1, const double* __restrict__ u2,
                                              const double* __restrict__ u3, double c1, double c2, double c3,
                                              double sound, Scalars* S) {
    const int ncell = g.N1 * g.N2 * g.N3;
    unsigned long long bits = ~0ull;
    int bad = 0;
    for (int lin = blockIdx.x * blockDim.x + threadIdx.x; lin < ncell; lin += gridDim.x * blockDim.x) {
        int i = lin % g.N1 + 1, r = lin / g.N1;
        int j = r % g.N2 + 1, k = r / g.N2 + 1;
        int c = at(g, i, j, k);
        double a1 = u1[c], a2 = u2[c], a3 = u3[c];
        double dtu1 = c1 / (sound + fabs(a1));
        double dtu2 = c2 / (sound + fabs(a2));
        double dtu3 = c3 / (sound + fabs(a3));
        double dtu = min3d(dtu1, dtu2, dtu3);
        bits = umin64(bits, (unsigned long long)__double_as_longlong(dtu));
        bad += !(isfinite(a1) && isfinite(a2) && isfinite(a3));
    }
    block_min_count(bits, (unsigned)bad, &S->dt_bits, S->nf, S->nstep);
}

// mode 0: dt := MIN and reset the running MIN; mode 1: the same plus the
// driver's bookkeeping TIME += dt/2 twice and nStep++ (source/main.cpp:83-84,
// 91); mode 2: only reset the running MIN (discard it); mode 3: dt := MIN and
// keep the running MIN (a TimeStepSize between fused steps, whose next step
// finalizes the same MIN again).
__global__ void k_cfl_finalize(Scalars* S, int mode) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double dt = __longlong_as_double((long long)S->dt_bits);
        if (mode == 3) {
            S->dt = dt;
            return;
        }
        S->dt_bits = S->dt_init;
        if (mode == 2) return;
        S->dt = dt;
        if (mode == 1) {
            S->nstep += 1;
            double t = S->time + 0.5 * dt;
            S->time = t + 0.5 * dt;
        }
    }
}

// The bookkeeping statements of the reference's loop, one at a time
// (source/main.cpp:83-84, 91): what 0 is nStep++, what 1 is TIME += 0.5 * dt.
__global__ void k_clock(Scalars* S, int what) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (what == 0) S->nstep += 1;
        else S->time = S->time + 0.5 * S->dt;
    }
}

// ---------------------------------------------------------------------------
// Phase1 (source/main.cpp:417-575): Con <- nCon, then the flux-form half-step.
__device__ __forceinline__ void phase1_cell(const Fields& F, int c, int px, int sxy, double dt, const Consts& K) {
    double** a = (double**)F.a;
    const double ds1 = K.ds1, ds2 = K.ds2, ds3 = K.ds3, dvc = K.dvc;
    double u1_c = a[F_u1nCon][c], u2_c = a[F_u2nCon][c], u3_c = a[F_u3nCon][c];
    double ro_c = a[F_ronCon][c], t_c = a[F_tnCon][c];
    a[F_u1Con][c] = u1_c;
    a[F_u2Con][c] = u2_c;
    a[F_u3Con][c] = u3_c;
    a[F_roCon][c] = ro_c;
    a[F_tCon][c] = t_c;
    double u1_cn, u2_cn, u3_cn, ro_cn, t_cn;
    if (a[F_condition][c] < 0.5) {
        const int fe = c + 1, fw = c, fn = c + px, fs = c, ft = c + sxy, fb = c;
        double u1_e = a[F_u11][fe], u2_e = a[F_u21][fe], u3_e = a[F_u31][fe];
        double ro_e = a[F_ro1][fe], t_e = a[F_t1][fe], p_e = a[F_p1][fe];
        double u1_w = a[F_u11][fw], u2_w = a[F_u21][fw], u3_w = a[F_u31][fw];
        double ro_w = a[F_ro1][fw], t_w = a[F_t1][fw], p_w = a[F_p1][fw];
        double u1_n = a[F_u12][fn], u2_n = a[F_u22][fn], u3_n = a[F_u32][fn];
        double ro_n = a[F_ro2][fn], t_n = a[F_t2][fn], p_n = a[F_p2][fn];
        double u1_s = a[F_u12][fs], u2_s = a[F_u22][fs], u3_s = a[F_u32][fs];
        double ro_s = a[F_ro2][fs], t_s = a[F_t2][fs], p_s = a[F_p2][fs];
        double u1_t = a[F_u13][ft], u2_t = a[F_u23][ft], u3_t = a[F_u33][ft];
        double ro_t = a[F_ro3][ft], t_t = a[F_t3][ft], p_t = a[F_p3][ft];
        double u1_b = a[F_u13][fb], u2_b = a[F_u23][fb], u3_b = a[F_u33][fb];
        double ro_b = a[F_ro3][fb], t_b = a[F_t3][fb], p_b = a[F_p3][fb];

        ro_cn = (ro_c * dvc - 0.5 * dt * (
                     (ro_e * u1_e - ro_w * u1_w) * ds1 +
                     (ro_n * u2_n - ro_s * u2_s) * ds2 +
                     (ro_t * u3_t - ro_b * u3_b) * ds3)) / dvc;
        u1_cn = (ro_c * u1_c * dvc - 0.5 * dt * (
                     ((ro_e * u1_e * u1_e + p_e) - (ro_w * u1_w * u1_w + p_w)) * ds1 +
                     (ro_n * u1_n * u2_n - ro_s * u1_s * u2_s) * ds2 +
                     (ro_t * u1_t * u3_t - ro_b * u1_b * u3_b) * ds3)) / (ro_cn * dvc);
        u2_cn = (ro_c * u2_c * dvc - 0.5 * dt * (
                     (ro_e * u2_e * u1_e - ro_w * u2_w * u1_w) * ds1 +
                     ((ro_n * u2_n * u2_n + p_n) - (ro_s * u2_s * u2_s + p_s)) * ds2 +
                     (ro_t * u2_t * u3_t - ro_b * u2_b * u3_b) * ds3)) / (ro_cn * dvc);
        u3_cn = (ro_c * u3_c * dvc - 0.5 * dt * (
                     (ro_e * u3_e * u1_e - ro_w * u3_w * u1_w) * ds1 +
                     (ro_n * u3_n * u2_n - ro_s * u3_s * u2_s) * ds2 +
                     ((ro_t * u3_t * u3_t + p_t) - (ro_b * u3_b * u3_b + p_b)) * ds3)) / (ro_cn * dvc);
        t_cn = (ro_c * t_c * dvc - 0.5 * dt * (
                    (ro_e * t_e * u1_e - ro_w * t_w * u1_w) * ds1 +
                    (ro_n * t_n * u2_n - ro_s * t_s * u2_s) * ds2 +
                    (ro_t * t_t * u3_t - ro_b * t_b * u3_b) * ds3)) / (dvc * ro_cn);
    } else {
        u1_cn = 0.; u2_cn = 0.; u3_cn = 0.; ro_cn = K.ro0; t_cn = K.t0;
    }
    a[F_u1nCon][c] = u1_cn;
    a[F_u2nCon][c] = u2_cn;
    a[F_u3nCon][c] = u3_cn;
    a[F_ronCon][c] = ro_cn;
    a[F_tnCon][c] = t_cn;
}

__global__ void __launch_bounds__(256) k_phase1(Geo g, Fields F, Consts K, const Scalars* S) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > g.N1 || j > g.N2 || k > g.N3) return;
    phase1_cell(F, at(g, i, j, k), g.px, g.sxy, S->dt, K);
}

// ---------------------------------------------------------------------------
// Ghost fill of one exchange, one launch.  Per array and axis d, one of:
//
// copy: the single-GPU image of the reference's MPI halo exchange with
// periods={1,1,1} (source/main.cpp:619-653, 688-719, 754-785, 1101-1189,
// 1296-1384, 1491-1579).  For a cell array: plane 0 <- plane N_d, plane
// N_d+1 <- plane 1; for a face array along its own axis: face 0 <- face N_d,
// face N_d+2 <- face 2.
//
// zero: the same planes set to 0.  The exchange along a non-periodic axis
// of one rank, whose neighbour is MPI_PROC_NULL, copies a receive buffer that
// nothing has written (the reference's buffers are uninitialised,
// Arr2d.cpp:7-12; zero-filled buffers define the result).
//
// grad: the zero-gradient ends that the reference's physical-boundary blocks
// write after that exchange (source/main.cpp:803-845 and 1633-1698): cell
// planes 0 <- 1 and N+1 <- N, face planes 0 <- 1 and N+2 <- N+1.  They
// overwrite exactly the planes the zero fill writes, so the host drops the
// zero fill wherever they apply.
//
// Only the interior of the two other axes is written (no edges or corners),
// and every read is of an interior plane, which no mode writes: the modes
// and arrays are independent and run together.
__global__ void __launch_bounds__(256) k_exchange(Geo g, ExchangePlan plan) {
    const int a = blockIdx.y, d = blockIdx.z;
    if (a >= plan.n) return;
    const int bit = 1 << d;
    const int mode = (plan.copy_mask[a] & bit) ? 0 : (plan.zero_mask[a] & bit) ? 1 : (plan.grad_mask[a] & bit) ? 2 : -1;
    if (mode < 0) return;
    const int N[3] = {g.N1, g.N2, g.N3};
    const int st[3] = {1, g.px, g.sxy};
    const int o1 = d == 0 ? 1 : 0;   // x-fastest enumeration of the plane
    const int o2 = d == 2 ? 1 : 2;
    const int plane = N[o1] * N[o2];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * plane) return;
    const int side = t >= plane;
    const int r = t - side * plane;
    const int q = r % N[o1] + 1, p = r / N[o1] + 1;
    double* A = plan.p[a];
    const int f = plan.is_face[a];
    const int base = q * st[o1] + p * st[o2];
    const int hi = (N[d] + 1 + f) * st[d];   // the far ghost plane
    if (mode == 1) {
        A[base + (side == 0 ? 0 : hi)] = 0.0;
    } else if (mode == 0) {
        if (side == 0) A[base] = A[base + N[d] * st[d]];
        else A[base + hi] = A[base + (1 + f) * st[d]];
    } else {
        if (side == 0) A[base] = A[base + st[d]];
        else A[base + hi] = A[base + (N[d] + f) * st[d]];
    }
}

// ---------------------------------------------------------------------------
// StressTensor (source/main.cpp:856-982): sigma on the faces of every
// orientation; one thread per (i,j,k) in [1,N1+1]x[1,N2+1]x[1,N3+1] writes
// the faces of the orientations for which (i,j,k) is a computed face.
__device__ __forceinline__ void stress_one(double* const* a, int c, int cw, double dx, double VIS, double& v1,
                                           double& v2, double& v3) {
    double cond_c = a[F_condition][c], cond_cw = a[F_condition][cw];
    double u1_c = a[F_u1Con][c], u1_cw = a[F_u1Con][cw];
    double u2_c = a[F_u2Con][c], u2_cw = a[F_u2Con][cw];
    double u3_c = a[F_u3Con][c], u3_cw = a[F_u3Con][cw];
    if (cond_cw < 0.5 && cond_c < 0.5) {
        v1 = VIS * (u1_c - u1_cw) / dx;
        v2 = VIS * (u2_c - u2_cw) / dx;
        v3 = VIS * (u3_c - u3_cw) / dx;
    } else if (cond_cw < 0.5 && cond_c > 0.5) {
        v1 = 2 * VIS * (-u1_cw) / dx;
        v2 = 2 * VIS * (-u2_cw) / dx;
        v3 = 2 * VIS * (-u3_cw) / dx;
    } else if (cond_cw > 0.5 && cond_c < 0.5) {
        v1 = 2 * VIS * u1_c / dx;
        v2 = 2 * VIS * u2_c / dx;
        v3 = 2 * VIS * u3_c / dx;
    } else {
        v1 = v2 = v3 = 0.;
    }
}

__global__ void __launch_bounds__(256) k_stress_faces(Geo g, Fields F, Consts K) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > g.N1 + 1 || j > g.N2 + 1 || k > g.N3 + 1) return;
    double* const* a = F.a;
    int c = at(g, i, j, k);
    double v1, v2, v3;
    if (j <= g.N2 && k <= g.N3) {
        stress_one(a, c, c - 1, K.dx[0], K.VIS, v1, v2, v3);
        a[F_sigm11][c] = v1; a[F_sigm21][c] = v2; a[F_sigm31][c] = v3;
    }
    if (i <= g.N1 && k <= g.N3) {
        stress_one(a, c, c - g.px, K.dx[1], K.VIS, v1, v2, v3);
        a[F_sigm12][c] = v1; a[F_sigm22][c] = v2; a[F_sigm32][c] = v3;
    }
    if (i <= g.N1 && j <= g.N2) {
        stress_one(a, c, c - g.sxy, K.dx[2], K.VIS, v1, v2, v3);
        a[F_sigm13][c] = v1; a[F_sigm23][c] = v2; a[F_sigm33][c] = v3;
    }
}

// The reference zeroes the nine stress arrays over their full extent before
// every evaluation (source/main.cpp:579-595).  The face kernel above writes
// only computed faces and nothing else ever writes a stress array, so a
// context zeroes them once after any upload made them non-zero.
__global__ void __launch_bounds__(256) k_zero(double* const* arrs, int narr, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (int a = 0; a < narr; ++a)
        for (long long q = t; q < n; q += stride) arrs[a][q] = 0.0;
}

// Friction force (source/main.cpp:995-1023).
__device__ __forceinline__ void forces_cell(const Fields& F, int c, int px, int sxy, const Consts& K) {
    double* const* a = F.a;
    const int sx = 1, sy = px, sz = sxy;
    if (a[F_condition][c] == 0) {
        double ro_c = a[F_roCon][c];
        a[F_f1][c] = ((a[F_sigm11][c + sx] - a[F_sigm11][c]) * K.ds1 +
                      (a[F_sigm12][c + sy] - a[F_sigm12][c]) * K.ds2 +
                      (a[F_sigm13][c + sz] - a[F_sigm13][c]) * K.ds3) * ro_c;
        a[F_f2][c] = ((a[F_sigm21][c + sx] - a[F_sigm21][c]) * K.ds1 +
                      (a[F_sigm22][c + sy] - a[F_sigm22][c]) * K.ds2 +
                      (a[F_sigm23][c + sz] - a[F_sigm23][c]) * K.ds3) * ro_c;
        a[F_f3][c] = ((a[F_sigm31][c + sx] - a[F_sigm31][c]) * K.ds1 +
                      (a[F_sigm32][c + sy] - a[F_sigm32][c]) * K.ds2 +
                      (a[F_sigm33][c + sz] - a[F_sigm33][c]) * K.ds3) * ro_c;
    } else {
        a[F_f1][c] = 0.;
        a[F_f2][c] = 0.;
        a[F_f3][c] = 0.;
    }
}

__global__ void __launch_bounds__(256) k_forces(Geo g, Fields F, Consts K) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > g.N1 || j > g.N2 || k > g.N3) return;
    forces_cell(F, at(g, i, j, k), g.px, g.sxy, K);
}

// UseForces (source/main.cpp:1026-1045).
__device__ __forceinline__ void use_forces_cell(const Fields& F, int c, const Consts& K, const Scalars* S) {
    double* const* a = F.a;
    if (a[F_condition][c] == 0) {
        const double dt = S->dt, dvc = K.dvc;
        double ro_cn = a[F_ronCon][c];
        a[F_u1nCon][c] = (ro_cn * dvc * a[F_u1nCon][c] + 0.5 * dt * a[F_f1][c]) / (dvc * ro_cn);
        a[F_u2nCon][c] = (ro_cn * dvc * a[F_u2nCon][c] + 0.5 * dt * a[F_f2][c]) / (dvc * ro_cn);
        a[F_u3nCon][c] = (ro_cn * dvc * a[F_u3nCon][c] + 0.5 * dt * a[F_f3][c]) / (dvc * ro_cn);
    }
}

__global__ void __launch_bounds__(256) k_use_forces(Geo g, Fields F, Consts K, const Scalars* S) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > g.N1 || j > g.N2 || k > g.N3) return;
    use_forces_cell(F, at(g, i, j, k), K, S);
}

// The step's own sequences, one launch each where a phase only needs the
// previous phase's values of the same cell: the force and UseForces after
// the stresses, and Phase1 followed by UseForces (the forces of the step's
// StressTensor).  Each cell's arithmetic is the separate kernels'.
__global__ void __launch_bounds__(256) k_forces_use_forces(Geo g, Fields F, Consts K, const Scalars* S) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > g.N1 || j > g.N2 || k > g.N3) return;
    const int c = at(g, i, j, k);
    forces_cell(F, c, g.px, g.sxy, K);
    use_forces_cell(F, c, K, S);
}

__global__ void __launch_bounds__(256) k_phase1_use_forces(Geo g, Fields F, Consts K, const Scalars* S) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > g.N1 || j > g.N2 || k > g.N3) return;
    const int c = at(g, i, j, k);
    phase1_cell(F, c, g.px, g.sxy, S->dt, K);
    use_forces_cell(F, c, K, S);
}

// ---------------------------------------------------------------------------
// Phase2 sweep along axis D, one thread per face (source/main.cpp:1726-2011,
// 2023-2255, 2261-2494).  The reference fills 1-D line buffers cell by cell
// and then reads them face by face; the buffers are not a recurrence (face q
// reads the forward extrapolation of cell q-1 and the backward one of cell q),
// so a face thread evaluates the two cell halves it needs directly.  Faces are
// read from the current buffer and written to the next one (the reference's
// in-place update reads only old values).  Threads of the first and last face
// also carry the ghost faces 0 and N_D+2, set by the ghost fill, across.
// Backward extrapolation of t of one cell (the t part of cell_backward).
__device__ __forceinline__ double backward_t(double t_f, double t_b, double t_cn, double t_c, double un_cn, double dt,
                                            double dx) {
    double t = 2 * t_cn - t_f;
    double gt = 2 * (t_cn - t_c) / dt + un_cn * (t_f - t_b) / dx;
    clip(t, max3d(t_f, t_c, t_b) + dt * gt, min3d(t_f, t_c, t_b) + dt * gt);
    return t;
}

// The reference's X2 wall branch takes t from tbBuf[i] (source/main.cpp:2215),
// the line buffer at the position of the line's x index instead of the face's
// y index.  Position i <= N2+1 holds the backward t of cell (i, i, k) along y
// (this line's own cell loop, source/main.cpp:2023-2100); beyond, what the
// X1 sweep's last line (j = N2, k = N3) left: the backward t of cell
// (i, N2, N3) along x, from the x-faces before that sweep.
__device__ __forceinline__ double x2_wall_t(const Geo& g, const Fields& F, int i, int k, double dt, const Consts& K) {
    double* const* a = F.a;
    if (i <= g.N2 + 1) {
        const int c = at(g, i, i, k);
        return backward_t(a[F_t2][c + g.px], a[F_t2][c], a[F_tnCon][c], a[F_tCon][c], a[F_u2nCon][c], dt, K.dx[1]);
    }
    const int c = at(g, i, g.N2, g.N3);
    return backward_t(a[F_t1][c + 1], a[F_t1][c], a[F_tnCon][c], a[F_tCon][c], a[F_u1nCon][c], dt, K.dx[0]);
}

template <int D, bool GqUsesR>
__device__ __forceinline__ void phase2_face(const Geo& g, const Fields& F, const NextFaces& nx, const Consts& K,
                                            const Scalars* S, int inflow, int i, int j, int k) {
    const int E1 = g.N1 + (D == 0), E2 = g.N2 + (D == 1), E3 = g.N3 + (D == 2);
    if (i > E1 || j > E2 || k > E3) return;
    constexpr int TA = D == 0 ? 1 : 0, TB = D == 2 ? 1 : 2;
    const int sd = D == 0 ? 1 : D == 1 ? g.px : g.sxy;
    const int q = D == 0 ? i : D == 1 ? j : k;
    const int Nd = D == 0 ? g.N1 : D == 1 ? g.N2 : g.N3;
    double* const* a = F.a;
    const double* Un = a[face_u(D, D)];
    const double* Ua = a[face_u(D, TA)];
    const double* Ub = a[face_u(D, TB)];
    const double* P = a[face_p(D)];
    const double* T = a[face_t(D)];
    const double* unN = a[F_u1nCon + D];
    const double* unC = a[F_u1Con + D];
    const double* uaN = a[F_u1nCon + TA];
    const double* uaC = a[F_u1Con + TA];
    const double* ubN = a[F_u1nCon + TB];
    const double* ubC = a[F_u1Con + TB];
    const double* roN = a[F_ronCon];
    const double* roC = a[F_roCon];
    const double* tN = a[F_tnCon];
    const double* tC = a[F_tCon];
    const double dt = S->dt, dx = K.dx[D];

    const int c = at(g, i, j, k);   // cell q and face q
    const int cb = c - sd;          // cell q-1 and face q-1
    const int cf = c + sd;          // face q+1

    CellSide sb;   // cell q-1: forward face q, backward face q-1
    sb.un_f = Un[c]; sb.un_b = Un[cb]; sb.un_cn = unN[cb]; sb.un_c = unC[cb];
    sb.ua_f = Ua[c]; sb.ua_b = Ua[cb]; sb.ua_cn = uaN[cb]; sb.ua_c = uaC[cb];
    sb.ub_f = Ub[c]; sb.ub_b = Ub[cb]; sb.ub_cn = ubN[cb]; sb.ub_c = ubC[cb];
    sb.ro_cn = roN[cb]; sb.ro_c = roC[cb];
    sb.t_f = T[c]; sb.t_b = T[cb]; sb.t_cn = tN[cb]; sb.t_c = tC[cb];
    sb.p_f = P[c]; sb.p_b = P[cb];
    CellSide sf;   // cell q: forward face q+1, backward face q
    sf.un_f = Un[cf]; sf.un_b = Un[c]; sf.un_cn = unN[c]; sf.un_c = unC[c];
    sf.ua_f = Ua[cf]; sf.ua_b = Ua[c]; sf.ua_cn = uaN[c]; sf.ua_c = uaC[c];
    sf.ub_f = Ub[cf]; sf.ub_b = Ub[c]; sf.ub_cn = ubN[c]; sf.ub_c = ubC[c];
    sf.ro_cn = roN[c]; sf.ro_c = roC[c];
    sf.t_f = T[cf]; sf.t_b = T[c]; sf.t_cn = tN[c]; sf.t_c = tC[c];
    sf.p_f = P[cf]; sf.p_b = P[c];

    Fwd fw = cell_forward(sb, dt, dx, K);
    Bwd bw = cell_backward<GqUsesR>(sf, dt, dx, K);
    if (D == 0 && inflow) {
        // non-periodic X1 (source/main.cpp:1729-1745): the line's cell loop
        // skips both ghost cells; the inflow face takes R = 1 in front of a
        // fluid cell (0 in front of a solid one), t = t0 and no tangential
        // velocity, the outflow face Q = 5/6, t = t0, no tangential velocity
        if (q == 1) fw = Fwd{a[F_condition][c] < 0.5 ? 1. : 0., K.t0, 0., 0.};
        if (q == Nd + 1) bw = Bwd{5. / 6., K.t0, 0., 0.};
    }
    const double cond_b = a[F_condition][cb], cond_f = a[F_condition][c];
    FaceOut o;
    if (cond_b < 0.5 && cond_f < 0.5) {
        o = face_combine(fw, bw, sb.un_cn, sf.un_cn, sb.t_cn, sf.t_cn, sb.ua_cn, sf.ua_cn, sb.ub_cn, sf.ub_cn, T[c],
                         Ua[c], Ub[c], K);
    } else {
        // a face on a solid cell (source/main.cpp:1962-2001, 2206-2245,
        // 2445-2484): no normal flow; the fluid side's outgoing invariant or
        // its own state, by the sign of its normal velocity.  The X2 branch
        // takes the first tangential velocity from the normal one (u2nCon
        // where u1nCon belongs) and t from tbBuf[i].
        const double sound = K.sound, ro0_g = K.ro0;
        const double* ua_wall = D == 1 ? unN : uaN;
        o.un = 0.;
        if (cond_b > 0.5 && cond_f < 0.5) {
            if (unN[c] <= 0) {
                const double qn = bw.q;
                o.a = bw.a;
                o.b = bw.b;
                o.p = -qn * sound * ro0_g;
                o.ro = ro0_g + o.p / (sound * sound);
                o.t = D == 1 ? x2_wall_t(g, F, i, k, dt, K) : bw.t;
            } else {
                o.a = ua_wall[c];
                o.b = ubN[c];
                o.ro = roN[c];
                o.p = (o.ro - ro0_g) * (sound * sound);
                o.t = tN[c];
            }
        } else if (cond_b < 0.5 && cond_f > 0.5) {
            if (unN[cb] >= 0) {
                const double rn = fw.r;
                o.a = fw.a;
                o.b = fw.b;
                o.p = rn * sound * ro0_g;
                o.ro = ro0_g + o.p / (sound * sound);
                o.t = fw.t;
            } else {
                o.a = ua_wall[cb];
                o.b = ubN[cb];
                o.ro = roN[cb];
                o.p = (o.ro - ro0_g) * (sound * sound);
                o.t = tN[cb];
            }
        } else {
            o.a = 0.;
            o.b = 0.;
            o.p = 0.;
            o.ro = ro0_g;
            o.t = K.t0;
        }
    }
    nx.p[0][c] = o.un;
    nx.p[1][c] = o.a;
    nx.p[2][c] = o.b;
    nx.p[3][c] = o.p;
    nx.p[4][c] = o.ro;
    nx.p[5][c] = o.t;
    if (q == 1) {
        nx.p[0][cb] = Un[cb]; nx.p[1][cb] = Ua[cb]; nx.p[2][cb] = Ub[cb];
        nx.p[3][cb] = P[cb]; nx.p[4][cb] = a[face_ro(D)][cb];
    }
    if (q == Nd + 1) {
        nx.p[0][cf] = Un[cf]; nx.p[1][cf] = Ua[cf]; nx.p[2][cf] = Ub[cf];
        nx.p[3][cf] = P[cf]; nx.p[4][cf] = a[face_ro(D)][cf];
    }
}

// The three sweeps in one launch.  They are independent: each reads the
// current cell and face arrays only and writes the next faces of its own
// axis.  z-blocks [d (N3+1), (d+1)(N3+1)) run axis d.
// Bounded to 128 VGPRs, four waves per SIMD (131 unbounded, three): the
// sweeps' chains of IEEE divisions wait on latency, not on HBM, and the
// fourth wave shortens the launch by about 7% with nothing spilled.
template <bool Quirk>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_phase2_xyz(Geo g, Fields F, NextFaces nx0, NextFaces nx1, NextFaces nx2,
                                                    Consts K, const Scalars* S, int inflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int nz = g.N3 + 1, d = blockIdx.z / nz, k = blockIdx.z % nz + 1;
    if (d == 0) phase2_face<0, Quirk>(g, F, nx0, K, S, inflow, i, j, k);
    else if (d == 1) phase2_face<1, false>(g, F, nx1, K, S, 0, i, j, k);
    else phase2_face<2, Quirk>(g, F, nx2, K, S, 0, i, j, k);
}


// ---------------------------------------------------------------------------
// Vorticity at the cell centres from the face velocities, the expression of
// WriteDataParaView (source/main.cpp:2712-2733), interior cells; T is the
// face storage type (double, or float for mixed precision).  out[m] holds
// component m in the common index map.
__global__ void __launch_bounds__(256) k_vorticity(Geo g, Fields F, Consts K, double* o1, double* o2, double* o3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > g.N1) return;
    double* const* a = F.a;
    const double* u32 = a[F_u32];
    const double* u23 = a[F_u23];
    const double* u13 = a[F_u13];
    const double* u31 = a[F_u31];
    const double* u21 = a[F_u21];
    const double* u12 = a[F_u12];
    const int c = at(g, i, j, k);
    const double dx1 = K.dx[0], dx2 = K.dx[1], dx3 = K.dx[2];
    double d2u3 = ((double)u32[c + g.px] - (double)u32[c]) / dx2,
           d3u2 = ((double)u23[c + g.sxy] - (double)u23[c]) / dx3,
           d3u1 = ((double)u13[c + g.sxy] - (double)u13[c]) / dx3,
           d1u3 = ((double)u31[c + 1] - (double)u31[c]) / dx1,
           d1u2 = ((double)u21[c + 1] - (double)u21[c]) / dx1,
           d2u1 = ((double)u12[c + g.px] - (double)u12[c]) / dx2;
    o1[c] = d2u3 - d3u2;
    o2[c] = d3u1 - d1u3;
    o3[c] = d1u2 - d2u1;
}

// AverageValues (source/main.cpp:2878-2891): averageU_m += u_m * dt over the
// interior cells.
__global__ void __launch_bounds__(256) k_average(Geo g, Fields F, const Scalars* S, double* a1, double* a2,
                                                 double* a3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > g.N1) return;
    const int c = at(g, i, j, k);
    const double dt = S->dt;
    a1[c] += F.a[F_u1nCon][c] * dt;
    a2[c] += F.a[F_u2nCon][c] * dt;
    a3[c] += F.a[F_u3nCon][c] * dt;
}

// ---------------------------------------------------------------------------
// Face densities from face pressures, rho = ro0 + p/c^2 (source/main.cpp:1930),
// on the faces a Phase2 sweep writes (1..N_d+1 along d, interior across).
// The fused step does not store face densities (it derives them); this
// restores them before the phase-by-phase kernels read them.
__global__ void __launch_bounds__(256) k_rho_faces(Geo g, Fields F, Consts K) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > g.N1 + 1 || j > g.N2 + 1 || k > g.N3 + 1) return;
    const int c = at(g, i, j, k);
    const double s2 = K.sound * K.sound;
    if (j <= g.N2 && k <= g.N3) F.a[F_ro1][c] = K.ro0 + F.a[F_p1][c] / s2;
    if (i <= g.N1 && k <= g.N3) F.a[F_ro2][c] = K.ro0 + F.a[F_p2][c] / s2;
    if (i <= g.N1 && j <= g.N2) F.a[F_ro3][c] = K.ro0 + F.a[F_p3][c] / s2;
}

// ---------------------------------------------------------------------------
// Diagnostics (WriteEnergy, source/main.cpp:2851-2876, plus mass and a
// non-finite scan): fixed block->cell assignment and a fixed-order final
// reduction, so repeated calls return identical sums.
__global__ void __launch_bounds__(256) k_diag_partial(Geo g, Fields F, Consts K, double* part) {
    const int ncell = g.N1 * g.N2 * g.N3;
    double ke = 0, kp = 0, m = 0, umax = 0, bad = 0;
    for (int lin = blockIdx.x * blockDim.x + threadIdx.x; lin < ncell; lin += gridDim.x * blockDim.x) {
        int i = lin % g.N1 + 1, r = lin / g.N1;
        int j = r % g.N2 + 1, k = r / g.N2 + 1;
        int c = at(g, i, j, k);
        double u1 = F.a[F_u1nCon][c], u2 = F.a[F_u2nCon][c], u3 = F.a[F_u3nCon][c];
        double ro = F.a[F_ronCon][c], t = F.a[F_tnCon][c];
        double q = u1 * u1 + u2 * u2 + u3 * u3;
        ke += q / 2;
        kp += 0.5 * ro * q * K.dvc;
        m += ro * K.dvc;
        double um = fmax(fabs(u1), fmax(fabs(u2), fabs(u3)));
        umax = fmax(umax, um);
        if (!(isfinite(u1) && isfinite(u2) && isfinite(u3) && isfinite(ro) && isfinite(t))) bad += 1;
    }
    __shared__ double sh[5][256];
    sh[0][threadIdx.x] = ke; sh[1][threadIdx.x] = kp; sh[2][threadIdx.x] = m;
    sh[3][threadIdx.x] = umax; sh[4][threadIdx.x] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int v = 0; v < 5; ++v) {
                if (v == 3) sh[v][threadIdx.x] = fmax(sh[v][threadIdx.x], sh[v][threadIdx.x + s]);
                else sh[v][threadIdx.x] += sh[v][threadIdx.x + s];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < 5; ++v) part[v * gridDim.x + blockIdx.x] = sh[v][0];
}

__global__ void __launch_bounds__(256) k_diag_final(const double* part, int nblocks, double* out) {
    __shared__ double sh[5][256];
    double acc[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int v = 0; v < 5; ++v) {
            double x = part[v * nblocks + b];
            acc[v] = v == 3 ? fmax(acc[v], x) : acc[v] + x;
        }
    for (int v = 0; v < 5; ++v) sh[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int v = 0; v < 5; ++v)
                sh[v][threadIdx.x] = v == 3 ? fmax(sh[v][threadIdx.x], sh[v][threadIdx.x + s])
                                            : sh[v][threadIdx.x] + sh[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < 5; ++v) out[v] = sh[v][0];
}

// ---------------------------------------------------------------------------
// Taylor-Green initial state from 1-D tables (sin, cos, cos 2x of every cell
// centre and face coordinate, evaluated on the host): the products below
// are the host formula's, term for term.
__global__ void __launch_bounds__(256) k_init_tgv(Geo g, Fields F, Fields F2, TgvTables T, double U0, double amp,
                                                   Consts K) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > g.N1 + 2 || j > g.N2 + 2 || k > g.N3 + 2) return;
    int c = at(g, i, j, k);
    double* const* a = F.a;
    const double s2 = K.sound * K.sound;
    if (i <= g.N1 + 1 && j <= g.N2 + 1 && k <= g.N3 + 1) {
        double sx = T.sc[0][i], cx = T.cc[0][i], c2x = T.c2c[0][i];
        double sy = T.sc[1][j], cy = T.cc[1][j], c2y = T.c2c[1][j];
        double cz = T.cc[2][k], c2z = T.c2c[2][k];
        double u = U0 * sx * cy * cz;
        double v = -U0 * cx * sy * cz;
        double p = amp * (c2x + c2y) * (c2z + 2.0);
        double ro = K.ro0 + p / s2;
        a[F_u1nCon][c] = u; a[F_u2nCon][c] = v; a[F_u3nCon][c] = 0.0;
        a[F_ronCon][c] = ro; a[F_tnCon][c] = K.t0;
        if (F2.a[F_ronCon]) {   // second buffer of the level-n fields: identical, edges included
            F2.a[F_u1nCon][c] = u; F2.a[F_u2nCon][c] = v; F2.a[F_u3nCon][c] = 0.0;
            F2.a[F_ronCon][c] = ro; F2.a[F_tnCon][c] = K.t0;
        }
        a[F_u1Con][c] = u; a[F_u2Con][c] = v; a[F_u3Con][c] = 0.0;
        a[F_roCon][c] = ro; a[F_tCon][c] = K.t0;
        a[F_condition][c] = 0.0;
        a[F_f1][c] = 0.0; a[F_f2][c] = 0.0; a[F_f3][c] = 0.0;
    }
    for (int d = 0; d < 3; ++d) {
        const double* tx_s = d == 0 ? T.sf[0] : T.sc[0];
        const double* tx_c = d == 0 ? T.cf[0] : T.cc[0];
        const double* tx_2 = d == 0 ? T.c2f[0] : T.c2c[0];
        const double* ty_s = d == 1 ? T.sf[1] : T.sc[1];
        const double* ty_c = d == 1 ? T.cf[1] : T.cc[1];
        const double* ty_2 = d == 1 ? T.c2f[1] : T.c2c[1];
        const double* tz_c = d == 2 ? T.cf[2] : T.cc[2];
        const double* tz_2 = d == 2 ? T.c2f[2] : T.c2c[2];
        double u = U0 * tx_s[i] * ty_c[j] * tz_c[k];
        double v = -U0 * tx_c[i] * ty_s[j] * tz_c[k];
        double p = amp * (tx_2[i] + ty_2[j]) * (tz_2[k] + 2.0);
        double ro = K.ro0 + p / s2;
        double vals[6] = {u, v, 0.0, p, ro, K.t0};
        const int slots[6] = {face_u(d, 0), face_u(d, 1), face_u(d, 2), face_p(d), face_ro(d), face_t(d)};
        for (int s = 0; s < 6; ++s) {
            a[slots[s]][c] = vals[s];
            if (F2.a[slots[s]]) F2.a[slots[s]][c] = vals[s];
        }
    }
    for (int s = F_sigm11; s <= F_sigm33; ++s) a[s][c] = 0.0;
}

}  // namespace arkd

// ---------------------------------------------------------------------------
// Launchers.
namespace arkd {

static dim3 cell_grid(int e1, int e2, int e3) { return dim3((e1 + 63) / 64, (e2 + 3) / 4, e3); }
static const dim3 kBlock(64, 4, 1);

hipError_t launch_cfl(hipStream_t st, const Geo& g, const Fields& F, double c1, double c2, double c3, double sound,
                      Scalars* S) {
    hipLaunchKernelGGL(k_cfl, dim3(1024), dim3(256), 0, st, g, F.a[F_u1Con], F.a[F_u2Con],
                       F.a[F_u3Con], c1, c2, c3, sound, S);
    return hipGetLastError();
}

hipError_t launch_cfl_finalize(hipStream_t st, Scalars* S, int advance) {
    hipLaunchKernelGGL(k_cfl_finalize, dim3(1), dim3(64), 0, st, S, advance);
    return hipGetLastError();
}

hipError_t launch_clock(hipStream_t st, Scalars* S, int what) {
    hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, st, S, what);
    return hipGetLastError();
}

hipError_t launch_phase1(hipStream_t st, const Geo& g, const Fields& F, const Consts& K, const Scalars* S) {
    hipLaunchKernelGGL(k_phase1, cell_grid(g.N1, g.N2, g.N3), kBlock, 0, st, g, F, K, S);
    return hipGetLastError();
}

static dim3 plane_grid(const Geo& g, int n) {
    int m = g.N1 * g.N2;
    if (g.N1 * g.N3 > m) m = g.N1 * g.N3;
    if (g.N2 * g.N3 > m) m = g.N2 * g.N3;
    return dim3((2 * m + 255) / 256, n, 3);
}

hipError_t launch_exchange(hipStream_t st, const Geo& g, const ExchangePlan& plan) {
    if (plan.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_exchange, plane_grid(g, plan.n), dim3(256), 0, st, g, plan);
    return hipGetLastError();
}

hipError_t launch_stress_faces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K) {
    hipLaunchKernelGGL(k_stress_faces, cell_grid(g.N1 + 1, g.N2 + 1, g.N3 + 1), kBlock, 0, st, g, F, K);
    return hipGetLastError();
}

hipError_t launch_zero(hipStream_t st, double* const* dev_ptrs, int narr, long long n) {
    hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, dev_ptrs, narr, n);
    return hipGetLastError();
}

hipError_t launch_forces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K) {
    hipLaunchKernelGGL(k_forces, cell_grid(g.N1, g.N2, g.N3), kBlock, 0, st, g, F, K);
    return hipGetLastError();
}

hipError_t launch_use_forces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K, const Scalars* S) {
    hipLaunchKernelGGL(k_use_forces, cell_grid(g.N1, g.N2, g.N3), kBlock, 0, st, g, F, K, S);
    return hipGetLastError();
}

hipError_t launch_forces_use_forces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K,
                                   const Scalars* S) {
    hipLaunchKernelGGL(k_forces_use_forces, cell_grid(g.N1, g.N2, g.N3), kBlock, 0, st, g, F, K, S);
    return hipGetLastError();
}

hipError_t launch_phase1_use_forces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K,
                                    const Scalars* S) {
    hipLaunchKernelGGL(k_phase1_use_forces, cell_grid(g.N1, g.N2, g.N3), kBlock, 0, st, g, F, K, S);
    return hipGetLastError();
}

hipError_t launch_phase2_xyz(hipStream_t st, int scheme, const Geo& g, const Fields& F, const NextFaces nx[3],
                             const Consts& K, const Scalars* S, int inflow) {
    dim3 grid = cell_grid(g.N1 + 1, g.N2 + 1, 3 * (g.N3 + 1));
    if (scheme == 0)
        hipLaunchKernelGGL(k_phase2_xyz<true>, grid, kBlock, 0, st, g, F, nx[0], nx[1], nx[2], K, S, inflow);
    else
        hipLaunchKernelGGL(k_phase2_xyz<false>, grid, kBlock, 0, st, g, F, nx[0], nx[1], nx[2], K, S, inflow);
    return hipGetLastError();
}

hipError_t launch_vorticity(hipStream_t st, const Geo& g, const Fields& F, const Consts& K,
                            double* o1, double* o2, double* o3) {
    dim3 grid((g.N1 + 255) / 256, g.N2, g.N3);
    hipLaunchKernelGGL(k_vorticity, grid, dim3(256), 0, st, g, F, K, o1, o2, o3);
    return hipGetLastError();
}

hipError_t launch_average(hipStream_t st, const Geo& g, const Fields& F, const Scalars* S, double* a1, double* a2,
                          double* a3) {
    hipLaunchKernelGGL(k_average, dim3((g.N1 + 255) / 256, g.N2, g.N3), dim3(256), 0, st, g, F, S, a1, a2, a3);
    return hipGetLastError();
}

hipError_t launch_rho_faces(hipStream_t st, const Geo& g, const Fields& F, const Consts& K) {
    hipLaunchKernelGGL(k_rho_faces, cell_grid(g.N1 + 1, g.N2 + 1, g.N3 + 1), kBlock, 0, st, g, F, K);
    return hipGetLastError();
}

hipError_t launch_diag(hipStream_t st, const Geo& g, const Fields& F, const Consts& K, double* part, int nblocks,
                       double* out) {
    hipLaunchKernelGGL(k_diag_partial, dim3(nblocks), dim3(256), 0, st, g, F, K, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, st, part, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_init_tgv(hipStream_t st, const Geo& g, const Fields& F, const Fields& F2, const TgvTables& T,
                           double U0, double amp, const Consts& K) {
    hipLaunchKernelGGL(k_init_tgv, cell_grid(g.N1 + 3, g.N2 + 3, g.N3 + 3), kBlock, 0, st, g, F, F2, T, U0, amp, K);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// BoundaryConditions::alongX1/2/3 (source/BoundaryConditions.cpp:22-85) on
// five cell fields u1, u2, u3, ro, t.  With n = N_d + 1 (the reference's node
// count) the ghost planes are 0 and n and the loops run over the interior
// 1..N of the two other axes.  alongX1 copies the neighbouring interior plane
// whatever the period flag (22-42); alongX2/alongX3 wrap when periodic and
// otherwise set walls (u = 0, ro zero-gradient, t = t0); the periodic top
// plane of alongX3 takes t from ro(…, 1) as the reference does (73).
__global__ void __launch_bounds__(256) k_boundary(Geo g, BcFields F, int axis, int periodic, double t0) {
    const int Na = axis == 0 ? g.N2 : g.N1, Nb = axis == 2 ? g.N2 : g.N3;
    const int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int b = blockIdx.y + 1;
    if (a > Na || b > Nb) return;
    const int n = (axis == 0 ? g.N1 : axis == 1 ? g.N2 : g.N3) + 1;
    const int stride = axis == 0 ? 1 : axis == 1 ? g.px : g.sxy;
    const int c0 = axis == 0 ? at(g, 0, a, b) : axis == 1 ? at(g, a, 0, b) : at(g, a, b, 0);
    const int top = c0 + n * stride, bot = c0;
    if (axis == 0) {
        for (int q = 0; q < 5; ++q) {
            F.f[q][top] = F.f[q][top - stride];
            F.f[q][bot] = F.f[q][bot + stride];
        }
        return;
    }
    const int one = c0 + stride, last = c0 + (n - 1) * stride;
    if (periodic) {
        for (int q = 0; q < 3; ++q) F.f[q][top] = F.f[q][one];
        F.f[3][top] = F.f[3][one];
        F.f[4][top] = axis == 2 ? F.f[3][one] : F.f[4][one];
        for (int q = 0; q < 5; ++q) F.f[q][bot] = F.f[q][last];
    } else {
        for (int q = 0; q < 3; ++q) {
            F.f[q][top] = 0.;
            F.f[q][bot] = 0.;
        }
        F.f[3][top] = F.f[3][last];
        F.f[3][bot] = F.f[3][one];
        F.f[4][top] = t0;
        F.f[4][bot] = t0;
    }
}

// One plane normal to `axis` at index q, interior 1..N of the two other axes,
// to or from a buffer in the Arr2d layout buf[b*Na + a] (source/Arr2d.cpp:14-16),
// as the reference's halo packing loops fill it (e.g. source/main.cpp:619-630).
__global__ void __launch_bounds__(256) k_plane(Geo g, double* field, int axis, int q, double* buf, int to_buf) {
    const int Na = axis == 0 ? g.N2 : g.N1, Nb = axis == 2 ? g.N2 : g.N3;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (a >= Na || b >= Nb) return;
    const int c = axis == 0 ? at(g, q, a + 1, b + 1) : axis == 1 ? at(g, a + 1, q, b + 1) : at(g, a + 1, b + 1, q);
    if (to_buf) buf[(size_t)b * Na + a] = field[c];
    else field[c] = buf[(size_t)b * Na + a];
}

hipError_t launch_boundary(hipStream_t st, const Geo& g, const BcFields& F, int axis, int periodic, double t0) {
    const int Na = axis == 0 ? g.N2 : g.N1, Nb = axis == 2 ? g.N2 : g.N3;
    hipLaunchKernelGGL(k_boundary, dim3((Na + 255) / 256, Nb), dim3(256), 0, st, g, F, axis, periodic, t0);
    return hipGetLastError();
}

// Rows of e1 elements between two 3-D layouts: row (j, k) starts at
// src + k*s_plane + j*s_row and at dst + k*d_plane + j*d_row.  One workgroup
// per row: a host array in the reference's dense layout and a field in the
// padded device layout trade places in one contiguous transfer.
__global__ void k_copy_rows(const double* __restrict__ src, long long s_row, long long s_plane,
                            double* __restrict__ dst, long long d_row, long long d_plane, int e1, int e2) {
    const int r = blockIdx.x, j = r % e2, k = r / e2;
    const double* s = src + k * s_plane + j * s_row;
    double* d = dst + k * d_plane + j * d_row;
    for (int i = threadIdx.x; i < e1; i += blockDim.x) d[i] = s[i];
}

hipError_t launch_copy_rows(hipStream_t st, const double* src, long long s_row, long long s_plane, double* dst,
                            long long d_row, long long d_plane, int e1, int e2, int e3) {
    hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)e2 * (unsigned)e3), dim3(256), 0, st, src, s_row, s_plane, dst,
                       d_row, d_plane, e1, e2);
    return hipGetLastError();
}

hipError_t launch_plane(hipStream_t st, const Geo& g, double* field, int axis, int q, double* buf, int to_buf) {
    const int Na = axis == 0 ? g.N2 : g.N1, Nb = axis == 2 ? g.N2 : g.N3;
    hipLaunchKernelGGL(k_plane, dim3((Na + 255) / 256, Nb), dim3(256), 0, st, g, field, axis, q, buf, to_buf);
    return hipGetLastError();
}

}  // namespace arkd

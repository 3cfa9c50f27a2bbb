// Fused CABARET/ARK step: one kernel per time step.
//
// Reads the persistent state once (five level-n cell fields and the five
// face fields u1,u2,u3,p,t of each orientation) and writes it once (level
// n+1 cells into the other cell buffer, new faces into the other face
// buffer).  Everything the reference stores in between -- the half-level
// cells, the stresses, the forces, the face densities -- stays on chip.
//
// Blocking ("2.5-D"): a block owns TX x TY columns of cells and marches a
// chunk of planes in z (a host table sizes the chunks).  Its threads cover the owned columns plus a ring
// of one cell (RX x RY threads), because the new faces on the tile's edge
// need the half-level state of the neighbouring column, which the ring
// threads recompute.  Per plane the block computes, in order:
//   stress force f and the predictor (Phase1 + UseForces) for every thread's
//   cell, its Phase2 extrapolations along z, the new z-face below it (from
//   the previous plane's forward extrapolation) and the corrector (Phase1 +
//   UseForces) of the previous plane, whose top face has just been formed;
//   then its Phase2 extrapolations along x and y and the new x- and y-faces
//   (LDS exchange of the neighbour's forward extrapolation).
// Flux terms such as rho*u1*u2 are computed once per face and used by both
// cells that share it; the operands and their association are the
// reference's, so FP64 results are bitwise equal to the phase-by-phase path.
// Divisions by the step's constants (dx, dt, sound^2, ro0*sound, dvc) and by
// a cell's rho*dvc use div_rn with a correctly rounded reciprocal: the same
// bits as IEEE division for every finite quotient, at half the instructions.
//
// Periodic box: cell indices wrap (ghost cell 0 is cell N, N+1 is cell 1,
// which is what the reference's halo exchange delivers), face N+1 is a real
// face, u/p ghost faces 0 and N+2 map to faces N and 2, and t ghost faces are
// read from storage because the reference never refreshes them.
#include "ark_internal.hpp"
#include "ark_kernels.hpp"

namespace arkd {

// 32x13 owned columns plus the ring: 34x15 = 510 threads, eight wavefronts
// (the last one 62 lanes), two per SIMD (at most 256 VGPRs).  Eight wavefronts keep
// the four SIMDs of a CU equally loaded.  32x13 covers a 256^2 plane with 160
// tiles, 1.25 threads per owned column (30x14: 171 tiles, 1.34; 62x6: 1.68).
// Side by side on an earlier form of this kernel, the step at 256^3 took
// 2.21 ms (30x14: 2.27), at 320^3 4.18 ms (4.28); at 384^3 the two were
// within 0.5%.
constexpr int TX = 32, TY = 13;
constexpr int RX = TX + 2, RY = TY + 2;
constexpr int NTHR = RX * RY;

// Field groups: five fields `fb` bytes apart, element (0,0,0) of the first
// at the group pointer.  A field is read as group + v*fb + 8*idx with the
// group pointer in scalar registers and a 32-bit per-thread offset.
struct FusedArgs {
    Geo g;
    Consts K;
    double cfl_dx[3];          // CFL*dx_d, the numerators of the next step's CFL bound
    const char* A;             // level-n cells: rho, u1, u2, u3, t
    char* Ao;                  // level n+1 cells
    const char* F[3];          // old faces normal to d: u1, u2, u3, p, t
    char* Fo[3];               // new faces
    unsigned long long fb;     // bytes between the fields of a group
    const Scalars* S;          // S->dt is this step's dt, S->nstep its number
    unsigned long long* dt_next;
    long long* nf;             // non-finite count and first step (Scalars::nf)
};

__device__ __forceinline__ double gld(const char* grp, unsigned long long fb, int v, int idx) {
    const char* base = grp + v * fb;
    return *(const double*)(base + (unsigned)idx * 8u);
}
__device__ __forceinline__ void gst(char* grp, unsigned long long fb, int v, int idx, double x) {
    char* base = grp + v * fb;
    *(double*)(base + (unsigned)idx * 8u) = x;
}

__device__ __forceinline__ int wrapc(int q, int N) { return q < 1 ? q + N : (q > N ? q - N : q); }
__device__ __forceinline__ int mapf(int f, int N) { return f == 0 ? N : (f == N + 2 ? 2 : f); }

// Record layout of a group (ark_fused_layout; a long run of fused steps):
// fields 0-3 of index idx as one 32-byte record at grp + 8 + 32 idx (two
// 16-byte accesses instead of four 8-byte ones; 32-byte aligned, since
// grp + 8 starts a 128-byte line), field 4 where the per-field layout has it
// (the records stop short of it).  64-bit offset: 32 idx passes 2^32 above
// about 500^3 cells.
__device__ __forceinline__ const double2* grec(const char* grp, int idx) {
    return (const double2*)(grp + 8 + (unsigned long long)(unsigned)idx * 32ull);
}
struct Cell5 { double ro, u1, u2, u3, t; };
struct Face5 { double u1, u2, u3, p, t; };
struct Flux5 { double m[5]; };

template <bool Rec>
__device__ __forceinline__ Cell5 load_cell(const FusedArgs& a, int idx) {
    Cell5 c;
    if constexpr (Rec) {
        const double2* r = grec(a.A, idx);
        const double2 a01 = r[0], a23 = r[1];
        c.ro = a01.x; c.u1 = a01.y; c.u2 = a23.x; c.u3 = a23.y;
    } else {
        c.ro = gld(a.A, a.fb, 0, idx); c.u1 = gld(a.A, a.fb, 1, idx); c.u2 = gld(a.A, a.fb, 2, idx);
        c.u3 = gld(a.A, a.fb, 3, idx);
    }
    c.t = gld(a.A, a.fb, 4, idx);
    return c;
}
// The velocity (fields 1-3) of a cell, for the stress.
template <bool Rec>
__device__ __forceinline__ void load_velocity(const FusedArgs& a, int idx, double u[3]) {
    if constexpr (Rec) {
        const double2* r = grec(a.A, idx);
        u[0] = ((const double*)r)[1];
        const double2 b = r[1];
        u[1] = b.x; u[2] = b.y;
    } else {
        for (int m = 0; m < 3; ++m) u[m] = gld(a.A, a.fb, 1 + m, idx);
    }
}

template <bool Rec>
__device__ __forceinline__ Face5 load_face(const FusedArgs& a, int d, int idx) {
    Face5 f;
    if constexpr (Rec) {
        const double2* r = grec(a.F[d], idx);
        const double2 a01 = r[0], a23 = r[1];
        f.u1 = a01.x; f.u2 = a01.y; f.u3 = a23.x; f.p = a23.y;
    } else {
        f.u1 = gld(a.F[d], a.fb, 0, idx); f.u2 = gld(a.F[d], a.fb, 1, idx); f.u3 = gld(a.F[d], a.fb, 2, idx);
        f.p = gld(a.F[d], a.fb, 3, idx);
    }
    f.t = gld(a.F[d], a.fb, 4, idx);
    return f;
}

// The five flux terms of a face normal to D with density rho (the operands of
// source/main.cpp:530-557, one face at a time).
template <int D>
__device__ __forceinline__ Flux5 flux_terms(double ro, double u1, double u2, double u3, double p, double t) {
    Flux5 q;
    const double un = D == 0 ? u1 : D == 1 ? u2 : u3;
    q.m[0] = ro * un;
    q.m[1] = D == 0 ? ro * u1 * u1 + p : ro * u1 * un;
    q.m[2] = D == 1 ? ro * u2 * u2 + p : ro * u2 * un;
    q.m[3] = D == 2 ? ro * u3 * u3 + p : ro * u3 * un;
    q.m[4] = ro * t * un;
    return q;
}

template <int D>
__device__ __forceinline__ Flux5 old_flux(const Face5& f, const Consts& K) {
    double ro = K.ro0 + div_rn(f.p, K.sound * K.sound, K.rc2);   // the reference's face density, source/main.cpp:1930
    return flux_terms<D>(ro, f.u1, f.u2, f.u3, f.p, f.t);
}

// Conservative half-step from the flux differences (Phase1's update) followed
// by the force half-step (UseForces).  The seven divisions by ro_cn*dvc (the
// same product as dvc*ro_cn) share one reciprocal.
__device__ __forceinline__ Cell5 conservative_update(const Cell5& c, const double S[5], const double f[3], double dt,
                                                     const Consts& K) {
    const double dvc = K.dvc;
    Cell5 n;
    double ro_cn = div_rn(c.ro * dvc - 0.5 * dt * S[0], dvc, K.rdvc);
    const double m = ro_cn * dvc, rm = 1.0 / m;
    double u1_cn = div_rn(c.ro * c.u1 * dvc - 0.5 * dt * S[1], m, rm);
    double u2_cn = div_rn(c.ro * c.u2 * dvc - 0.5 * dt * S[2], m, rm);
    double u3_cn = div_rn(c.ro * c.u3 * dvc - 0.5 * dt * S[3], m, rm);
    double t_cn = div_rn(c.ro * c.t * dvc - 0.5 * dt * S[4], m, rm);
    n.ro = ro_cn;
    n.u1 = div_rn(ro_cn * dvc * u1_cn + 0.5 * dt * f[0], m, rm);
    n.u2 = div_rn(ro_cn * dvc * u2_cn + 0.5 * dt * f[1], m, rm);
    n.u3 = div_rn(ro_cn * dvc * u3_cn + 0.5 * dt * f[2], m, rm);
    n.t = t_cn;
    return n;
}

__device__ __forceinline__ double ucomp(const Cell5& c, int m) { return m == 0 ? c.u1 : m == 1 ? c.u2 : c.u3; }

// CellSide of a cell for the sweep along D from its level-n (a) and
// half-level (h) values and its two faces along D.
template <int D>
__device__ __forceinline__ CellSide make_side(const Cell5& a, const Cell5& h, const Face5& fb, const Face5& ff) {
    constexpr int TA = D == 0 ? 1 : 0, TB = D == 2 ? 1 : 2;
    auto fu = [](const Face5& f, int m) { return m == 0 ? f.u1 : m == 1 ? f.u2 : f.u3; };
    CellSide s;
    s.un_f = fu(ff, D); s.un_b = fu(fb, D); s.un_cn = ucomp(h, D); s.un_c = ucomp(a, D);
    s.ua_f = fu(ff, TA); s.ua_b = fu(fb, TA); s.ua_cn = ucomp(h, TA); s.ua_c = ucomp(a, TA);
    s.ub_f = fu(ff, TB); s.ub_b = fu(fb, TB); s.ub_cn = ucomp(h, TB); s.ub_c = ucomp(a, TB);
    s.ro_cn = h.ro; s.ro_c = a.ro;
    s.t_f = ff.t; s.t_b = fb.t; s.t_cn = h.t; s.t_c = a.t;
    s.p_f = ff.p; s.p_b = fb.p;
    return s;
}

// New face along D from the forward half of the cell behind (fw, hb) and the
// backward half of the cell ahead (bw, hf); old: the face's old values.
template <int D>
__device__ __forceinline__ FaceOut new_face(const Fwd& fw, const Cell5& hb, const Bwd& bw, const Cell5& hf,
                                            const Face5& old, const Consts& K) {
    constexpr int TA = D == 0 ? 1 : 0, TB = D == 2 ? 1 : 2;
    auto fu = [](const Face5& f, int m) { return m == 0 ? f.u1 : m == 1 ? f.u2 : f.u3; };
    return face_combine<true>(fw, bw, ucomp(hb, D), ucomp(hf, D), hb.t, hf.t, ucomp(hb, TA), ucomp(hf, TA),
                              ucomp(hb, TB), ucomp(hf, TB), old.t, fu(old, TA), fu(old, TB), K);
}

template <int D, bool Rec>
__device__ __forceinline__ void store_face(const FusedArgs& a, int idx, const FaceOut& o) {
    constexpr int TA = D == 0 ? 1 : 0, TB = D == 2 ? 1 : 2;
    if constexpr (Rec) {
        double u[3];
        u[D] = o.un; u[TA] = o.a; u[TB] = o.b;
        double2* w = (double2*)grec(a.Fo[D], idx);
        w[0] = make_double2(u[0], u[1]);
        w[1] = make_double2(u[2], o.p);
    } else {
        gst(a.Fo[D], a.fb, D, idx, o.un);
        gst(a.Fo[D], a.fb, TA, idx, o.a);
        gst(a.Fo[D], a.fb, TB, idx, o.b);
        gst(a.Fo[D], a.fb, 3, idx, o.p);
    }
    gst(a.Fo[D], a.fb, 4, idx, o.t);
}

template <int D>
__device__ __forceinline__ Flux5 new_flux(const FaceOut& o) {
    constexpr int TA = D == 0 ? 1 : 0, TB = D == 2 ? 1 : 2;
    double u[3];
    u[D] = o.un; u[TA] = o.a; u[TB] = o.b;
    return flux_terms<D>(o.ro, u[0], u[1], u[2], o.p, o.t);
}

// Per-thread LDS record of one plane.
struct LdsCell {
    double h[5];        // half-level rho, u1, u2, u3, t
    double fx[4];       // forward x-extrapolation (r, t, a, b)
    double fy[4];       // forward y-extrapolation
};
struct LdsFlux {
    double mx[5], my[5];
};
union LdsSlot {
    LdsCell c;
    LdsFlux f;
};

// One workgroup per chunk of a tile column.  units[u] = (tile column x
// fastest then y, first plane, end plane) lists the chunks in rounds (round q
// holds chunk q of every column), so the workgroups that run together work at
// the same height and share their rings through L2; die (XCD) b % 8 takes a
// contiguous run of chunks, so the columns in flight on one die are
// neighbours.
template <bool Quirk, bool Rec>
__global__ void __launch_bounds__(NTHR) k_fused_step(FusedArgs a, const int4* __restrict__ units, int nbx) {
    __shared__ LdsSlot lds[NTHR];
    const Geo& g = a.g;
    const Consts& K = a.K;
    const double dt = a.S->dt, rdt = 1.0 / dt;
    const int tid = threadIdx.x;
    const int tx = tid % RX, ty = tid / RX;
    const int4 w = units[die_major(gridDim.x)];
    const int i0 = (w.x % nbx) * TX + 1, j0 = (w.x / nbx) * TY + 1, k0 = w.y;
    const int iend = min(i0 + TX, g.N1 + 1), jend = min(j0 + TY, g.N2 + 1), kend = w.z;
    // logical cell of this thread (clamped into the region so every access stays in bounds)
    const int i = min(i0 - 1 + tx, iend), j = min(j0 - 1 + ty, jend);
    const bool in_region = (i0 - 1 + tx <= iend) && (j0 - 1 + ty <= jend);
    const bool own_x = i >= i0 && i < iend, own_y = j >= j0 && j < jend;
    const bool owned = in_region && own_x && own_y;
    const bool xface = in_region && own_y && i >= i0 && i <= iend;     // computes the x-face at i
    const bool yface = in_region && own_x && j >= j0 && j <= jend;     // computes the y-face at j
    const bool xface_store = xface && (i < iend || iend == g.N1 + 1);
    const bool yface_store = yface && (j < jend || jend == g.N2 + 1);
    const int iw = wrapc(i, g.N1), jw = wrapc(j, g.N2);
    const int iwm = wrapc(i - 1, g.N1), iwp = wrapc(i + 1, g.N1);
    const int jwm = wrapc(j - 1, g.N2), jwp = wrapc(j + 1, g.N2);
    const int px = g.px, sxy = g.sxy;
    const double VIS = K.VIS;
    FusedArgs A = a;

    // rolling state along z; plane p+1's cells are loaded during plane p-1
    Cell5 A_cur = load_cell<Rec>(A, at(g, iw, jw, wrapc(k0 - 1, g.N3)));
    Cell5 A_nx = load_cell<Rec>(A, at(g, iw, jw, wrapc(k0, g.N3)));
    double sz_lo[3];
    {
        const Cell5 A_below = load_cell<Rec>(A, at(g, iw, jw, wrapc(k0 - 2, g.N3)));
        sz_lo[0] = div_rn(VIS * (A_cur.u1 - A_below.u1), K.dx[2], K.rdx[2]);
        sz_lo[1] = div_rn(VIS * (A_cur.u2 - A_below.u2), K.dx[2], K.rdx[2]);
        sz_lo[2] = div_rn(VIS * (A_cur.u3 - A_below.u3), K.dx[2], K.rdx[2]);
    }
    // What plane p-1 hands to plane p -- its half level H, forward z
    // extrapolation, stress force, in-plane flux sums and z-face flux -- waits
    // in the lane's own LDS column (22 doubles) rather than in registers: only
    // the lane itself reads it, so no barrier guards it, and the 44 VGPRs it
    // frees keep the plane's work within 256 without spills.
    __shared__ double carry[22][NTHR];
    double* const cy = &carry[0][tid];
    auto cy_put = [&](int q, double v) { cy[q * NTHR] = v; };
    auto cy_get = [&](int q) { return cy[q * NTHR]; };
    // The top z-face of plane p-1 is the bottom z-face of plane p (the same
    // stored element) except where z wraps.
    Face5 zt_prev{};
    int zt_prev_at = -1;    // stored plane of zt_prev
    unsigned long long dtbits = ~0ull;
    unsigned bad = 0;

    for (int p = k0 - 1; p <= kend; ++p) {
        // keep the per-field addresses (group + v*stride) from being hoisted
        // out of the loop into scalar registers: they are cheap to recompute
        asm volatile("" : "+s"(A.fb));
        const int pw = wrapc(p, g.N3);
        const int c = at(g, iw, jw, pw);
        const Cell5 A_next = A_nx;
        if (p < kend) A_nx = load_cell<Rec>(A, at(g, iw, jw, wrapc(p + 2, g.N3)));

        // ---- stress force f (StressTensor, source/main.cpp:856-1023) -------
        double f[3];
        {
            const int cw = at(g, iwm, jw, pw), ce = at(g, iwp, jw, pw);
            const int cs = at(g, iw, jwm, pw), cn = at(g, iw, jwp, pw);
            double vW[3], vE[3], vS[3], vN[3];
            load_velocity<Rec>(A, cw, vW);
            load_velocity<Rec>(A, ce, vE);
            load_velocity<Rec>(A, cs, vS);
            load_velocity<Rec>(A, cn, vN);
            for (int m = 0; m < 3; ++m) {
                const double uc = ucomp(A_cur, m);
                const double uw = vW[m], ue = vE[m];
                const double us = vS[m], un = vN[m];
                double sx_lo = div_rn(VIS * (uc - uw), K.dx[0], K.rdx[0]);
                double sx_hi = div_rn(VIS * (ue - uc), K.dx[0], K.rdx[0]);
                double sy_lo = div_rn(VIS * (uc - us), K.dx[1], K.rdx[1]);
                double sy_hi = div_rn(VIS * (un - uc), K.dx[1], K.rdx[1]);
                double sz_hi = div_rn(VIS * (ucomp(A_next, m) - uc), K.dx[2], K.rdx[2]);
                f[m] = ((sx_hi - sx_lo) * K.ds1 + (sy_hi - sy_lo) * K.ds2 + (sz_hi - sz_lo[m]) * K.ds3) * A_cur.ro;
                sz_lo[m] = sz_hi;
            }
        }

        // ---- predictor (Phase1 + UseForces, old faces) and the Phase2 -------
        // ---- extrapolations of this cell along z -----------------------------
        Cell5 H;
        Face5 Pxb, Pxf, Pyb, Pyf;     // the faces of the x and y extrapolations
        Fwd fz;
        Bwd bz;
        double oz[3];                 // old t, tangential a, tangential b of the low z-face
        {
            const Face5 Xw = load_face<Rec>(A, 0, c), Xe = load_face<Rec>(A, 0, c + 1);
            const Face5 Ys = load_face<Rec>(A, 1, c), Yn = load_face<Rec>(A, 1, c + px);
            const Face5 Zb = zt_prev_at == pw ? zt_prev : load_face<Rec>(A, 2, c);
            const Face5 Zt = load_face<Rec>(A, 2, c + sxy);
            zt_prev = Zt;
            zt_prev_at = pw + 1;
            {
                const Flux5 fxw = old_flux<0>(Xw, K), fxe = old_flux<0>(Xe, K);
                const Flux5 fys = old_flux<1>(Ys, K), fyn = old_flux<1>(Yn, K);
                const Flux5 fzb = old_flux<2>(Zb, K), fzt = old_flux<2>(Zt, K);
                double S[5];
                for (int v = 0; v < 5; ++v)
                    S[v] = (fxe.m[v] - fxw.m[v]) * K.ds1 + (fyn.m[v] - fys.m[v]) * K.ds2 +
                           (fzt.m[v] - fzb.m[v]) * K.ds3;
                H = conservative_update(A_cur, S, f, dt, K);
            }
            // Phase2 uses the faces of the logical cell: the same as the
            // predictor's except for a ring cell just outside the box, whose
            // u/p ghost faces map into the box and whose t ghost faces are read
            // as stored
            Pxb = Xw; Pxf = Xe; Pyb = Ys; Pyf = Yn;
            Face5 Pzb = Zb, Pzf = Zt;
            if (i != iw) {
                const int cl = at(g, 0, jw, pw);
                Pxb = load_face<Rec>(A, 0, cl + mapf(i, g.N1));
                Pxf = load_face<Rec>(A, 0, cl + mapf(i + 1, g.N1));
                Pxb.t = gld(A.F[0], A.fb, 4, cl + i);
                Pxf.t = gld(A.F[0], A.fb, 4, cl + i + 1);
            }
            if (j != jw) {
                const int cl = at(g, iw, 0, pw);
                Pyb = load_face<Rec>(A, 1, cl + mapf(j, g.N2) * px);
                Pyf = load_face<Rec>(A, 1, cl + mapf(j + 1, g.N2) * px);
                Pyb.t = gld(A.F[1], A.fb, 4, cl + j * px);
                Pyf.t = gld(A.F[1], A.fb, 4, cl + (j + 1) * px);
            }
            if (p != pw) {
                const int cl = at(g, iw, jw, 0);
                Pzb = load_face<Rec>(A, 2, cl + mapf(p, g.N3) * sxy);
                Pzf = load_face<Rec>(A, 2, cl + mapf(p + 1, g.N3) * sxy);
                Pzb.t = gld(A.F[2], A.fb, 4, cl + p * sxy);
                Pzf.t = gld(A.F[2], A.fb, 4, cl + (p + 1) * sxy);
            }
            {
                const CellSide s = make_side<2>(A_cur, H, Pzb, Pzf);
                fz = cell_forward<true>(s, dt, K.dx[2], K, rdt, K.rdx[2]);
                bz = cell_backward<Quirk, true>(s, dt, K.dx[2], K, rdt, K.rdx[2]);
            }
            oz[0] = Pzb.t; oz[1] = Pzb.u1; oz[2] = Pzb.u2;
        }
        if (owned && p >= k0 && p < kend) {
            // the next step's CFL bound from the post-force half level (TimeStepSize, source/main.cpp:379-405)
            double d1 = A.cfl_dx[0] / (K.sound + fabs(H.u1));
            double d2 = A.cfl_dx[1] / (K.sound + fabs(H.u2));
            double d3 = A.cfl_dx[2] / (K.sound + fabs(H.u3));
            unsigned long long b = (unsigned long long)__double_as_longlong(min3d(d1, d2, d3));
            dtbits = b < dtbits ? b : dtbits;
            // a NaN or Inf anywhere in the half level makes this sum non-finite
            bad += !isfinite(H.ro + H.u1 + H.u2 + H.u3);
        }

        // ---- z-face p and the corrector of plane p-1 ------------------------
        Flux5 mz{};
        if (p >= k0) {
            Cell5 H_prev;
            H_prev.ro = cy_get(0); H_prev.u1 = cy_get(1); H_prev.u2 = cy_get(2);
            H_prev.u3 = cy_get(3); H_prev.t = cy_get(4);
            const Fwd fz_prev{cy_get(5), cy_get(6), cy_get(7), cy_get(8)};
            const FaceOut o = face_combine<true>(fz_prev, bz, H_prev.u3, H.u3, H_prev.t, H.t, H_prev.u1, H.u1,
                                                 H_prev.u2, H.u2, oz[0], oz[1], oz[2], K);
            mz = new_flux<2>(o);
            if (owned) {
                if (p < kend || kend == g.N3 + 1) store_face<2, Rec>(A, at(g, i, j, p), o);
                if (p > k0) {
                    double S[5], f_prev[3];
                    for (int v = 0; v < 5; ++v) S[v] = cy_get(12 + v) + (mz.m[v] - cy_get(17 + v)) * K.ds3;
                    for (int m = 0; m < 3; ++m) f_prev[m] = cy_get(9 + m);
                    const Cell5 An = conservative_update(H_prev, S, f_prev, dt, K);
                    const int co = at(g, i, j, p - 1);
                    if constexpr (Rec) {
                        double2* w = (double2*)grec(A.Ao, co);
                        w[0] = make_double2(An.ro, An.u1);
                        w[1] = make_double2(An.u2, An.u3);
                    } else {
                        gst(A.Ao, A.fb, 0, co, An.ro);
                        gst(A.Ao, A.fb, 1, co, An.u1);
                        gst(A.Ao, A.fb, 2, co, An.u2);
                        gst(A.Ao, A.fb, 3, co, An.u3);
                    }
                    gst(A.Ao, A.fb, 4, co, An.t);
                }
            }
        }

        // ---- Phase2 extrapolations along x and y ----------------------------
        Fwd fx, fy;
        Bwd bx, by;
        {
            const CellSide s = make_side<0>(A_cur, H, Pxb, Pxf);
            fx = cell_forward<true>(s, dt, K.dx[0], K, rdt, K.rdx[0]);
            bx = cell_backward<Quirk, true>(s, dt, K.dx[0], K, rdt, K.rdx[0]);
        }
        {
            const CellSide s = make_side<1>(A_cur, H, Pyb, Pyf);
            fy = cell_forward<true>(s, dt, K.dx[1], K, rdt, K.rdx[1]);
            by = cell_backward<false, true>(s, dt, K.dx[1], K, rdt, K.rdx[1]);
        }
        // old t, tangential a, tangential b of the low x- and y-faces
        const double ox[3] = {Pxb.t, Pxb.u2, Pxb.u3};
        const double oy[3] = {Pyb.t, Pyb.u1, Pyb.u3};

        // ---- x/y faces and the in-plane part of the corrector ---------------
        double S_xy[5] = {0, 0, 0, 0, 0};
        if (p >= k0 && p < kend) {
            LdsCell& me = lds[tid].c;
            me.h[0] = H.ro; me.h[1] = H.u1; me.h[2] = H.u2; me.h[3] = H.u3; me.h[4] = H.t;
            me.fx[0] = fx.r; me.fx[1] = fx.t; me.fx[2] = fx.a; me.fx[3] = fx.b;
            me.fy[0] = fy.r; me.fy[1] = fy.t; me.fy[2] = fy.a; me.fy[3] = fy.b;
            __syncthreads();
            Flux5 mx{}, my{};
            if (xface) {
                const LdsCell& nb = lds[tid - 1].c;
                const Fwd fwb{nb.fx[0], nb.fx[1], nb.fx[2], nb.fx[3]};
                const FaceOut o = face_combine<true>(fwb, bx, nb.h[1], H.u1, nb.h[4], H.t, nb.h[2], H.u2, nb.h[3],
                                                     H.u3, ox[0], ox[1], ox[2], K);
                if (xface_store) store_face<0, Rec>(A, at(g, i, j, p), o);
                mx = new_flux<0>(o);
            }
            if (yface) {
                const LdsCell& nb = lds[tid - RX].c;
                const Fwd fwb{nb.fy[0], nb.fy[1], nb.fy[2], nb.fy[3]};
                const FaceOut o = face_combine<true>(fwb, by, nb.h[2], H.u2, nb.h[4], H.t, nb.h[1], H.u1, nb.h[3],
                                                     H.u3, oy[0], oy[1], oy[2], K);
                if (yface_store) store_face<1, Rec>(A, at(g, i, j, p), o);
                my = new_flux<1>(o);
            }
            __syncthreads();
            LdsFlux& mf = lds[tid].f;
            for (int v = 0; v < 5; ++v) { mf.mx[v] = mx.m[v]; mf.my[v] = my.m[v]; }
            __syncthreads();
            if (owned) {
                const LdsFlux& e = lds[tid + 1].f;
                const LdsFlux& n = lds[tid + RX].f;
                for (int v = 0; v < 5; ++v) S_xy[v] = (e.mx[v] - mx.m[v]) * K.ds1 + (n.my[v] - my.m[v]) * K.ds2;
            }
            __syncthreads();
        }

        cy_put(0, H.ro); cy_put(1, H.u1); cy_put(2, H.u2); cy_put(3, H.u3); cy_put(4, H.t);
        cy_put(5, fz.r); cy_put(6, fz.t); cy_put(7, fz.a); cy_put(8, fz.b);
        for (int m = 0; m < 3; ++m) cy_put(9 + m, f[m]);
        for (int v = 0; v < 5; ++v) { cy_put(12 + v, S_xy[v]); cy_put(17 + v, mz.m[v]); }
        A_cur = A_next;
    }
    block_min_count(dtbits, bad, A.dt_next, A.nf, A.S->nstep);
}

}  // namespace arkd

// ---------------------------------------------------------------------------
namespace {

using namespace arkd;

hipError_t run_fused(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    FusedArgs a;
    a.g = c->geo;
    a.K = c->K;
    a.cfl_dx[0] = c->prm.CFL * c->prm.dx1;
    a.cfl_dx[1] = c->prm.CFL * c->prm.dx2;
    a.cfl_dx[2] = c->prm.CFL * c->prm.dx3;
    a.A = c->gbase[0][c->gpar[0]];
    a.Ao = c->gbase[0][c->gpar[0] ^ 1];
    for (int d = 0; d < 3; ++d) {
        a.F[d] = c->gbase[1 + d][c->gpar[1 + d]];
        a.Fo[d] = c->gbase[1 + d][c->gpar[1 + d] ^ 1];
    }
    a.fb = c->field_bytes;
    a.S = c->S;
    a.dt_next = &c->S->dt_bits;
    a.nf = c->S->nf;
    const dim3 grid(c->fused_nunits);
    const int4* units = (const int4*)c->fused_units;
    const int nbx = (c->geo.N1 + TX - 1) / TX;
    const bool q = c->prm.scheme == ARK_SCHEME_REFERENCE;
    auto kern = c->fused_rec ? (q ? k_fused_step<true, true> : k_fused_step<false, true>)
                             : (q ? k_fused_step<true, false> : k_fused_step<false, false>);
    hipLaunchKernelGGL(kern, grid, dim3(NTHR), 0, c->stream, a, units, nbx);
    return hipGetLastError();
}

// One buffer of a group between the per-field layout and the record layout
// (grec), index for index over the n indices of the common index map.
__global__ void k_group_layout(const char* __restrict__ src, char* __restrict__ dst, unsigned long long fb, long long n,
                               int to_records) {
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (to_records) {
        double v[4];
        for (int f = 0; f < 4; ++f) v[f] = *(const double*)(src + f * fb + q * 8);
        double2* w = (double2*)(dst + 8 + q * 32);
        w[0] = make_double2(v[0], v[1]);
        w[1] = make_double2(v[2], v[3]);
    } else {
        const double2* r = (const double2*)(src + 8 + q * 32);
        const double2 a = r[0], b = r[1];
        *(double*)(dst + 0 * fb + q * 8) = a.x;
        *(double*)(dst + 1 * fb + q * 8) = a.y;
        *(double*)(dst + 2 * fb + q * 8) = b.x;
        *(double*)(dst + 3 * fb + q * 8) = b.y;
    }
    *(double*)(dst + 4 * fb + q * 8) = *(const double*)(src + 4 * fb + q * 8);
}

// The chunk table of the FP64 fused step, built once per context for the
// workgroups the chip holds at once.
ark_status ensure_fused_units(ark_ctx* c) {
    if (c->fused_units) return ARK_OK;
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    const void* fn = c->prm.scheme == ARK_SCHEME_REFERENCE ? (const void*)k_fused_step<true, true>
                                                          : (const void*)k_fused_step<false, true>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NTHR, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return ark_upload_chunks(c, ark_chunk_table(c->geo, TX, TY, cus * per_cu), &c->fused_units, &c->fused_nunits);
}

}  // namespace

bool ark_fused_eligible(const ark_ctx* c) {
    return c->geo.N1 >= 2 && c->geo.N2 >= 2 && c->geo.N3 >= 2 && c->rho_consistent && !c->has_solid &&
           c->prm.precision == ARK_PREC_FP64 && c->prm.periodic[0] && c->prm.periodic[1] && c->prm.periodic[2];
}

// Every index of both buffers of every group is converted: current -> spare,
// other -> current, and the three buffers rotate.  Each buffer then holds in
// the new layout exactly what it held in the old one, ghosts and entries a
// step does not write included, so the fused step's results do not depend on
// the layout.  Cost at 256^3: two conversions of each of the four groups,
// about 2.1 ms.
ark_status ark_fused_layout(ark_ctx* c, bool records) {
    if (c->prm.precision != ARK_PREC_FP64 || c->fused_rec == records) return ARK_OK;
    if (!c->group_spare) {   // only ever missing on the way to records
        char* mem = nullptr;
        const size_t nb = 5 * c->field_bytes;
        if (hipMalloc(&mem, nb) != hipSuccess) {
            (void)hipGetLastError();
            return ARK_ERR_UNSUPPORTED;   // no room for the spare: the caller stays per field
        }
        c->allocs.push_back(mem);
        c->bytes += nb;
        ark_status s = c->hip(hipMemsetAsync(mem, 0, nb, c->stream), "hipMemset spare group");
        if (s != ARK_OK) return s;
        c->group_spare = mem + kOriginPad * sizeof(double);
    }
    const long long n = (long long)c->geo.sxy * (c->geo.N3 + 3);
    const dim3 grid((unsigned)((n + 255) / 256));
    for (int g = 0; g < 4; ++g) {
        char* cur = c->gbase[g][c->gpar[g]];
        char* oth = c->gbase[g][c->gpar[g] ^ 1];
        char* spare = c->group_spare;
        hipLaunchKernelGGL(k_group_layout, grid, dim3(256), 0, c->stream, (const char*)cur, spare, c->field_bytes, n,
                           records ? 1 : 0);
        hipLaunchKernelGGL(k_group_layout, grid, dim3(256), 0, c->stream, (const char*)oth, cur, c->field_bytes, n,
                           records ? 1 : 0);
        ark_status s = c->hip(hipGetLastError(), records ? "groups to records" : "groups to fields");
        if (s != ARK_OK) return s;
        ark_set_group_buffers(c, g, spare, cur);
        c->group_spare = oth;
    }
    c->fused_rec = records;
    return ARK_OK;
}

// A run of at least this many fused steps converts the groups to the record
// layout first (a switch there and back, about 4.2 ms at 256^3, costs what
// 15 record steps save); shorter runs use it only if a previous run left it.
constexpr int kRecordSteps = 32;

// One fused step per iteration: begin the step (dt), run the kernel, swap
// the double-buffered persistent fields.  The kernel leaves the next step's
// dt as the running MIN in S->dt_bits (dt_pending).
static ark_status fused_steps(ark_ctx* c, int nsteps) {
    ark_status s = ensure_fused_units(c);
    for (int n = 0; s == ARK_OK && n < nsteps; ++n) {
        s = ark_begin_step(c);
        if (s == ARK_OK && n == 0 && nsteps >= kRecordSteps) {
            // the record layout is only faster: without memory for its spare
            // buffer (ARK_ERR_UNSUPPORTED) the run stays on the per-field one
            const ark_status r = ark_fused_layout(c, true);
            if (r != ARK_ERR_UNSUPPORTED) s = r;
        }
        if (s == ARK_OK) s = c->launch("fused_step", run_fused, c);
        if (s != ARK_OK) break;
        for (int g = 0; g < 4; ++g) ark_swap_group(c, g);
        c->rho_stored = false;
        c->dt_pending = true;
    }
    return s;
}

// nsteps steps: nsteps-1 fused steps, then the last step through the
// phase-by-phase kernels, so that on return every array -- including the
// reference's scratch arrays Con, f, sigma and the ghost copies -- holds what
// the reference holds after the same steps.  A state the fused kernel does
// not cover (face densities not known to equal ro0 + p/c^2, a dimension
// below 2) takes its first step phase by phase, which makes it covered.
ark_status ark_step_fused(ark_ctx* c, int nsteps) {
    ark_status s = ARK_OK;
    int done = 0;
    if (nsteps > 1 && !ark_fused_eligible(c)) {
        s = ark_begin_step(c);
        if (s == ARK_OK) s = ark_phase_body(c);
        done = 1;
    }
    if (s == ARK_OK && done < nsteps - 1 && ark_fused_eligible(c)) {
        s = fused_steps(c, nsteps - 1 - done);
        done = nsteps - 1;
    }
    // a long run on the phase path (solid cells, physical boundaries) is
    // launch-bound: its steps go in pairs through one captured graph
    if (s == ARK_OK && nsteps - done >= 8 && !c->profiling) {
        const int pairs = (nsteps - done) / 2;
        s = ark_phase_pairs(c, pairs);
        done += 2 * pairs;
    }
    while (s == ARK_OK && done < nsteps) {
        s = ark_refresh_rho(c);
        if (s == ARK_OK) s = ark_begin_step(c);
        if (s == ARK_OK) s = ark_phase_body(c);
        ++done;
    }
    return s;
}

// ark_advance: fused steps only.
extern "C" ark_status ark_advance(ark_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return ARK_ERR_INVALID;
    if (c->has_solid || !(c->prm.periodic[0] && c->prm.periodic[1] && c->prm.periodic[2]))
        return c->fail(ARK_ERR_UNSUPPORTED,
                       "the fused steps need a fluid, triply periodic box; solid cells and non-periodic axes run on "
                       "the phase path (ark_step)");
    ark_status s = ark_nf_check(c, 0);
    if (s != ARK_OK) return s;
    if (c->prm.precision == ARK_PREC_MIXED) {
        s = ark_mixed_steps(c, nsteps);
    } else {
        if (!ark_fused_eligible(c))
            return c->fail(ARK_ERR_STATE,
                           "ark_advance needs face densities equal to ro0 + p/c^2 (after ark_init_tgv or any step) "
                           "and N1, N2, N3 >= 2 in FP64 mode; call ark_step first");
        s = fused_steps(c, nsteps);
    }
    return s == ARK_OK ? ark_nf_post(c) : s;
}

// Mixed-precision fused CABARET/ARK step.
//
// Storage: the level-n cell fields stay FP64 (rho, u1, u2, u3, t); the face
// fields are FP32: u1, u2, u3, p and t' = t - t0.  Each face has (u1, u2, u3,
// p) as one float4 record; the t' of the x-, y- and z-faces at one index share
// a three-float record (rec_of, trec_of below).  The face density is
// rho0 + p/c^2 and is never stored; p itself is the density perturbation
// c^2 (rho - rho0), so the weakly compressible equation of state
// p = c^2 (rho - rho0) (source/main.cpp:1774-1775) never subtracts two
// numbers near rho0 in FP32: cell densities are turned into FP32
// perturbations rho - rho0 after an FP64 subtraction.
//
// Arithmetic: fluxes, stresses, Riemann invariants, the limiter and the new
// face values are FP32 (divisions by step constants become multiplications by
// reciprocals).  The closing conservative update is FP64: with c = dt/(2 dvc)
// and the predictor's and corrector's flux sums S, S' (FP32), the
// reference's predictor-force-corrector-force sequence (source/main.cpp:
// 530-557, 1038-1040, applied twice) is exactly
//     rho'   = rho - c (S_rho + S'_rho)
//     (rho u)' = rho u - c (S_u + S'_u - 2 f)
//     (rho t)' = rho t - c (S_t + S'_t)
// so level n+1 is level n plus two FP32 increments accumulated in FP64.  The
// half-level state the Phase2 sweeps need is formed in FP32 from the
// predictor's increments.
//
// Blocking and the order of work per plane are those of the FP64 fused
// kernel (ark_fused.hip): owned columns plus a one-cell ring, z-march.
#include <algorithm>
#include <string>
#include <array>
#include <vector>

#include "ark_internal.hpp"
#include "ark_kernels.hpp"

// This file is not bound to the reference's rounding (its results are
// checked against the FP64 path with a tolerance), so a*b+c may become one
// fused multiply-add and divisions by cell quantities use the hardware
// reciprocal (1 ulp in FP32; in the closing update refined by one Newton
// step to about 2^-46, far below the FP32 flux arithmetic's rounding).
#pragma clang fp contract(fast)

namespace arkd {
namespace mixed {

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpd(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    return fma(y, fma(-x, y, 1.0), y);
}

// Tile shapes: TX x TY owned columns plus a one-cell ring, RX x RY = (TX+2)
// x (TY+2) threads, a multiple of four wavefronts so the four SIMDs of a CU
// are equally loaded.  The single-box step times 30x14 (512 threads) and
// 32x13 (510 threads), two waves per SIMD each, on its first step and keeps
// the faster; the batched step lets its autotuner choose among all four.
template <int TX_, int TY_>
struct Shape {
    static constexpr int TX = TX_, TY = TY_, RX = TX_ + 2, RY = TY_ + 2, NTHR = (TX_ + 2) * (TY_ + 2);
};
using Shape30x6 = Shape<30, 6>;
using Shape30x14 = Shape<30, 14>;
using Shape62x6 = Shape<62, 6>;
using Shape32x13 = Shape<32, 13>;   // 34 x 15 = 510 threads: whole 32-cell rows owned

struct Args {
    Geo g;
    // FP32 step constants
    float VIS, c, c2, inv_c2, inv_rc, half_rc, ro0;
    float inv_dx[3], ds[3];
    float cfl_dx[3];
    // FP64 constants of the closing update
    double ro0d, t0d, dvc;
    const char* A;      // FP64 level-n cells: (rho, u1, u2, u3) records (cell_rec), t in field 4
    char* Ao;
    const char* F[3];   // FP32 faces normal to d: u1, u2, u3, p, t'
    char* Fo[3];
    unsigned long long fb;
    const Scalars* S;
    unsigned long long* dt_next;
    long long* nf;      // non-finite count and first step (Scalars::nf)
    int LZ;
    // body acceleration (forcing, FP32, components force + m*fstride in the
    // field index map) or null; the cell force is rho dvc F
    const float* force;
    long long fstride;
    // or, with force null, the forcing evaluated per cell: this step's
    // 16-mode table (ark_turb: mode n, component d at 6n + 2d, re then im)
    // and the per-axis {cos, sin} of theta and 2 theta at the cell centres
    // (x entries, then y, then z); null when the box is not forced
    const float* fmodes;
    const float4* ftrig;
    float dvcf;
};

__device__ __forceinline__ double ldd(const char* grp, unsigned long long fb, int v, int idx) {
    return *(const double*)(grp + v * fb + (unsigned)idx * 8u);
}
__device__ __forceinline__ void std_(char* grp, unsigned long long fb, int v, int idx, double x) {
    *(double*)(grp + v * fb + (unsigned)idx * 8u) = x;
}
// Face records of one orientation, in the allocation of its five FP64 fields
// (group pointer grp = FP64 element (0,0,0), 120 bytes into the allocation):
// face idx has (u1, u2, u3, p) as the float4 at grp + 8 + 16 idx (16-byte
// aligned).  The t' of the three faces at idx (x, y, z) are one T3 record at
// xg + 2 fb + 12 idx, xg the x-faces' group pointer of the same buffer: that
// area lies past the float4 records, inside the group's 5 fb bytes.  A lane
// reads its x-, y- and z-faces' t' with one load and writes them with one
// store, where one float per face took three of each.
struct T3 { float x, y, z; };
__device__ __forceinline__ const float4* rec_of(const char* grp, unsigned idx) {
    return (const float4*)(grp + 8 + idx * 16u);
}
__device__ __forceinline__ const T3* trec_of(const char* xg, unsigned long long fb, unsigned idx) {
    return (const T3*)(xg + 2 * fb + idx * 12u);
}
// t' of the orientation-d face at idx (a component of its T3 record)
__device__ __forceinline__ const float* tp_of(const char* xg, unsigned long long fb, int d, unsigned idx) {
    return (const float*)(xg + 2 * fb + idx * 12u + 4u * (unsigned)d);
}

__device__ __forceinline__ int wrapc(int q, int N) { return q < 1 ? q + N : (q > N ? q - N : q); }
__device__ __forceinline__ int mapf(int f, int N) { return f == 0 ? N : (f == N + 2 ? 2 : f); }

// Level-n cells while a context steps in mixed precision: (rho, u1, u2, u3)
// of cell idx as one 32-byte record at grp + 8 + 32 idx (two 16-byte loads
// instead of four 8-byte ones; 32-byte aligned, since grp + 8 starts a
// 128-byte line), t in its own field (field 4 of the group, which the
// records stop short of).  The offset is 64-bit: 32 idx passes 2^32 above
// about 500^3 cells.
__device__ __forceinline__ const double2* cell_rec(const char* grp, int idx) {
    return (const double2*)(grp + 8 + (unsigned long long)(unsigned)idx * 32ull);
}
struct CellD { double ro, u[3], t; };                 // FP64 level-n cell
struct CellF { float dr, u[3], tp; };                 // FP32 view: rho - rho0, u, t - t0
struct FaceF { float u[3], p, tp; };

__device__ __forceinline__ CellD load_cell(const Args& a, int idx) {
    CellD c;
    const double2* r = cell_rec(a.A, idx);
    const double2 a01 = r[0], a23 = r[1];
    c.ro = a01.x; c.u[0] = a01.y; c.u[1] = a23.x; c.u[2] = a23.y;
    c.t = ldd(a.A, a.fb, 4, idx);
    return c;
}
__device__ __forceinline__ CellF to_f(const CellD& c, const Args& a) {
    CellF f;
    f.dr = (float)(c.ro - a.ro0d);
    f.u[0] = (float)c.u[0]; f.u[1] = (float)c.u[1]; f.u[2] = (float)c.u[2];
    f.tp = (float)(c.t - a.t0d);
    return f;
}
// The float4 record of an orientation-d face; t' is set by the caller.
__device__ __forceinline__ FaceF load_rec(const Args& a, int d, int idx) {
    FaceF f;
    const float4 q = *rec_of(a.F[d], (unsigned)idx);
    f.u[0] = q.x;
    f.u[1] = q.y;
    f.u[2] = q.z;
    f.p = q.w;
    f.tp = 0.f;
    return f;
}
// One orientation-d face with its t'.
__device__ __forceinline__ FaceF load_face(const Args& a, int d, int idx) {
    FaceF f = load_rec(a, d, idx);
    f.tp = *tp_of(a.F[0], a.fb, d, (unsigned)idx);
    return f;
}

struct Flux { float m[5]; };

// Flux terms of a face normal to D (momentum flux carries p on the normal
// component); rho = rho0 + p/c^2.
template <int D>
__device__ __forceinline__ Flux flux(const float u[3], float p, float tp, const Args& a) {
    const float ro = a.ro0 + p * a.inv_c2;
    const float mro = ro * u[D];
    Flux q;
    q.m[0] = mro;
    q.m[1] = mro * u[0] + (D == 0 ? p : 0.f);
    q.m[2] = mro * u[1] + (D == 1 ? p : 0.f);
    q.m[3] = mro * u[2] + (D == 2 ? p : 0.f);
    q.m[4] = mro * tp;
    return q;
}

struct Side {   // one cell along the sweep axis: level n (c), half level (cn), faces f and b
    float un_f, un_b, un_cn, un_c;
    float ua_f, ua_b, ua_cn, ua_c;
    float ub_f, ub_b, ub_cn, ub_c;
    float p_f, p_b, p_cn, p_c;
    float t_f, t_b, t_cn, t_c;
};
struct FwdF { float r, t, a, b; };
struct BwdF { float q, t, a, b; };

// The limiter's clamp.  Its bounds are min3 + dt g and max3 + dt g with the
// same dt g, so lo <= hi, and the median of (v, lo, hi) is min(max(v, lo), hi)
// exactly: one v_med3_f32.
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return __builtin_amdgcn_fmed3f(v, lo, hi); }
__device__ __forceinline__ float max3f(float x, float y, float z) { return fmaxf(fmaxf(x, y), z); }
__device__ __forceinline__ float min3f(float x, float y, float z) { return fminf(fminf(x, y), z); }

// Limited extrapolations (CABARET min-max limiter, source/main.cpp:1777-1873).
// The Riemann invariants r = u + p/(rho0 c) and q = u - p/(rho0 c) are held
// relative to the cell's half-level normal velocity un_cn: every operation on
// them (differences, 2 r_cn - r_b, min, max, clamp) commutes with a common
// shift, and the shifted values are of the size of the velocity differences
// and p/(rho0 c), so FP32 keeps the digits that the face pressure
// (r - q) rho0 c / 2 is made of.  combine() adds the shifts back.
__device__ __forceinline__ FwdF forward(const Side& s, float dt, float inv_dt2, float inv_dx, const Args& a) {
    const float r_f = (s.un_f - s.un_cn) + s.p_f * a.inv_rc, r_b = (s.un_b - s.un_cn) + s.p_b * a.inv_rc;
    const float r_cn = s.p_cn * a.inv_rc, r_c = (s.un_c - s.un_cn) + s.p_c * a.inv_rc;
    const float gr = (r_cn - r_c) * inv_dt2 + (s.un_cn + a.c) * (r_f - r_b) * inv_dx;
    const float gt = (s.t_cn - s.t_c) * inv_dt2 + s.un_cn * (s.t_f - s.t_b) * inv_dx;
    const float ga = (s.ua_cn - s.ua_c) * inv_dt2 + s.un_cn * (s.ua_f - s.ua_b) * inv_dx;
    const float gb = (s.ub_cn - s.ub_c) * inv_dt2 + s.un_cn * (s.ub_f - s.ub_b) * inv_dx;
    FwdF o;
    o.r = clampf(2.f * r_cn - r_b, min3f(r_f, r_c, r_b) + dt * gr, max3f(r_f, r_c, r_b) + dt * gr);
    o.t = clampf(2.f * s.t_cn - s.t_b, min3f(s.t_f, s.t_c, s.t_b) + dt * gt, max3f(s.t_f, s.t_c, s.t_b) + dt * gt);
    o.a = clampf(2.f * s.ua_cn - s.ua_b, min3f(s.ua_f, s.ua_c, s.ua_b) + dt * ga,
                 max3f(s.ua_f, s.ua_c, s.ua_b) + dt * ga);
    o.b = clampf(2.f * s.ub_cn - s.ub_b, min3f(s.ub_f, s.ub_c, s.ub_b) + dt * gb,
                 max3f(s.ub_f, s.ub_c, s.ub_b) + dt * gb);
    return o;
}

// GqUsesR: the X1 and X3 sweeps take gq's time derivative from r
// (source/main.cpp:1804, 2321), the X2 sweep from q (2083).
template <bool GqUsesR>
__device__ __forceinline__ BwdF backward(const Side& s, float dt, float inv_dt2, float inv_dx, const Args& a) {
    const float q_f = (s.un_f - s.un_cn) - s.p_f * a.inv_rc, q_b = (s.un_b - s.un_cn) - s.p_b * a.inv_rc;
    const float q_cn = -(s.p_cn * a.inv_rc), q_c = (s.un_c - s.un_cn) - s.p_c * a.inv_rc;
    float dq;
    if (GqUsesR) {
        const float r_cn = s.p_cn * a.inv_rc, r_c = (s.un_c - s.un_cn) + s.p_c * a.inv_rc;
        dq = (r_cn - r_c) * inv_dt2;
    } else {
        dq = (q_cn - q_c) * inv_dt2;
    }
    const float gq = dq + (s.un_cn - a.c) * (q_f - q_b) * inv_dx;
    const float gt = (s.t_cn - s.t_c) * inv_dt2 + s.un_cn * (s.t_f - s.t_b) * inv_dx;
    const float ga = (s.ua_cn - s.ua_c) * inv_dt2 + s.un_cn * (s.ua_f - s.ua_b) * inv_dx;
    const float gb = (s.ub_cn - s.ub_c) * inv_dt2 + s.un_cn * (s.ub_f - s.ub_b) * inv_dx;
    BwdF o;
    o.q = clampf(2.f * q_cn - q_f, min3f(q_f, q_c, q_b) + dt * gq, max3f(q_f, q_c, q_b) + dt * gq);
    o.t = clampf(2.f * s.t_cn - s.t_f, min3f(s.t_f, s.t_c, s.t_b) + dt * gt, max3f(s.t_f, s.t_c, s.t_b) + dt * gt);
    o.a = clampf(2.f * s.ua_cn - s.ua_f, min3f(s.ua_f, s.ua_c, s.ua_b) + dt * ga,
                 max3f(s.ua_f, s.ua_c, s.ua_b) + dt * ga);
    o.b = clampf(2.f * s.ub_cn - s.ub_f, min3f(s.ub_f, s.ub_c, s.ub_b) + dt * gb,
                 max3f(s.ub_f, s.ub_c, s.ub_b) + dt * gb);
    return o;
}

struct FaceOutF { float un, p, t, a, b; };

// New face from the cell behind (forward half fw, half-level hb: un, t, a, b)
// and the cell ahead (bw, hf); upwinding of the passive quantities as in
// source/main.cpp:1932-1961.  fw.r and bw.q are relative to the half-level
// normal velocities ucb and ucf of their cells.
__device__ __forceinline__ FaceOutF combine(const FwdF& fw, const BwdF& bw, float ucb, float ucf, float tb_,
                                            float tf_, float ab_, float af_, float bb_, float bf_, float t_old,
                                            float a_old, float b_old, const Args& a) {
    FaceOutF o;
    o.un = ((fw.r + bw.q) + (ucb + ucf)) * 0.5f;
    o.p = ((fw.r - bw.q) + (ucb - ucf)) * a.half_rc;
    if (ucf >= 0.f && ucb >= 0.f) {
        o.t = fw.t; o.a = fw.a; o.b = fw.b;
    } else if (ucf <= 0.f && ucb <= 0.f) {
        o.t = bw.t; o.a = bw.a; o.b = bw.b;
    } else if (ucb >= 0.f && ucf <= 0.f) {
        if (ucb > -ucf) { o.t = fw.t; o.a = fw.a; o.b = fw.b; }
        else { o.t = bw.t; o.a = bw.a; o.b = bw.b; }
    } else {
        o.t = tf_ + tb_ - t_old;
        o.a = af_ + ab_ - a_old;
        o.b = bf_ + bb_ - b_old;
    }
    return o;
}

template <int D>
__device__ __forceinline__ Flux out_flux(const FaceOutF& o, const Args& a) {
    constexpr int TA = D == 0 ? 1 : 0, TB = D == 2 ? 1 : 2;
    float u[3];
    u[D] = o.un; u[TA] = o.a; u[TB] = o.b;
    return flux<D>(u, o.p, o.t, a);
}

template <int D>
__device__ __forceinline__ void store_face(const Args& a, int idx, const FaceOutF& o) {
    constexpr int TA = D == 0 ? 1 : 0, TB = D == 2 ? 1 : 2;
    float u[3];
    u[D] = o.un;
    u[TA] = o.a;
    u[TB] = o.b;
    *(float4*)rec_of(a.Fo[D], (unsigned)idx) = make_float4(u[0], u[1], u[2], o.p);
}
// The t' of one orientation-d face alone (the periodic copies of face 1).
template <int D>
__device__ __forceinline__ void store_tp(const Args& a, int idx, float t) {
    *(float*)tp_of(a.Fo[0], a.fb, D, (unsigned)idx) = t;
}

template <int D>
__device__ __forceinline__ Side make_side(const CellF& ac, const CellF& h, const FaceF& fb, const FaceF& ff,
                                          const Args& a) {
    constexpr int TA = D == 0 ? 1 : 0, TB = D == 2 ? 1 : 2;
    Side s;
    s.un_f = ff.u[D]; s.un_b = fb.u[D]; s.un_cn = h.u[D]; s.un_c = ac.u[D];
    s.ua_f = ff.u[TA]; s.ua_b = fb.u[TA]; s.ua_cn = h.u[TA]; s.ua_c = ac.u[TA];
    s.ub_f = ff.u[TB]; s.ub_b = fb.u[TB]; s.ub_cn = h.u[TB]; s.ub_c = ac.u[TB];
    s.p_f = ff.p; s.p_b = fb.p; s.p_cn = a.c2 * h.dr; s.p_c = a.c2 * ac.dr;
    s.t_f = ff.tp; s.t_b = fb.tp; s.t_cn = h.tp; s.t_c = ac.tp;
    return s;
}

// One lane's inputs for one plane: its own x-, y- and top z-face (the
// neighbours' faces come through LDS), the face beyond the region for lanes
// on its last column or row, and the velocity of the stress neighbour outside
// the region for lanes on its edge.
struct PlaneIn {
    FaceF xw, ys, zt, ex;
    double ou[3];
    float fo[3];    // body acceleration at the lane's cell (forcing)
    float tn[2];    // t' of the x- and y-face one plane up (from zt's T3 record)
};

// One tile of one box: tile column (tbx, tby), owned planes k0 .. kend-1.
// A lane keeps in registers only what is not also in LDS: its own x- and
// y-faces and the FP64 velocities of its cells are read back from the
// staged planes, which keeps the tile within two waves per SIMD (about 225
// VGPRs, none spilled).
template <bool Quirk, class S>
__device__ __forceinline__ void mixed_tile(const Args& a, int tbx, int tby, int k0, int kend) {
    constexpr int TX = S::TX, TY = S::TY, RX = S::RX, RY = S::RY, NTHR = S::NTHR;
    __shared__ float4 s_h[NTHR];             // half level: u1, u2, u3, t'
    __shared__ float4 s_fx[NTHR], s_fy[NTHR];  // forward extrapolations along x, y
    __shared__ float s_m[10][NTHR];          // new x- and y-face fluxes
    // Level-n velocities of a plane (FP32: the stress is FP32 arithmetic) and
    // the x- and y-faces of each lane's cell, with a halo that the edge lanes
    // fill with the value beyond the region, so every neighbour read is one
    // unconditional LDS load.  The velocities of two planes are held (p and
    // p+1); the faces of one plane, read before the plane's first barrier and
    // rewritten after it.
    constexpr int UX = RX + 2, UN = (RX + 2) * (RY + 2);   // velocities: one-cell halo all round
    constexpr int FX = RX + 1, FN = (RX + 1) * (RY + 1);   // faces: one more column and row
    __shared__ float s_u[2][3][UN];
    __shared__ float s_f[10][FN];
    const Geo& g = a.g;
    const double dtd = a.S->dt;
    const float dt = (float)dtd, inv_dt2 = (float)(2.0 / dtd);
    const double cd = 0.5 * dtd / a.dvc;       // FP64 closing-update factor
    const float cf = (float)cd;
    const int tid = threadIdx.x;
    const int tx = tid % RX, ty = tid / RX;
    const int i0 = tbx * TX + 1, j0 = tby * TY + 1;
    const int iend = min(i0 + TX, g.N1 + 1), jend = min(j0 + TY, g.N2 + 1);
    const int i = min(i0 - 1 + tx, iend), j = min(j0 - 1 + ty, jend);
    const bool in_region = (i0 - 1 + tx <= iend) && (j0 - 1 + ty <= jend);
    const bool own_x = i >= i0 && i < iend, own_y = j >= j0 && j < jend;
    const bool owned = in_region && own_x && own_y;
    const bool xface = in_region && own_y && i >= i0;
    const bool yface = in_region && own_x && j >= j0;
    const bool xfirst = tx == 0, yfirst = ty == 0;
    const bool xlast = i == iend, ylast = j == jend;
    const int iw = wrapc(i, g.N1), jw = wrapc(j, g.N2);
    const int px = g.px, sxy = g.sxy;
    // LDS neighbours, clamped so every lane reads inside the arrays
    const int tW = xfirst ? tid : tid - 1, tE = xlast ? tid : tid + 1;
    const int tS = yfirst ? tid : tid - RX, tN = ylast ? tid : tid + RX;
    const int upos = (ty + 1) * UX + tx + 1, fpos = ty * FX + tx;
    // cell offset of this lane within a plane
    const int base = iw + jw * px;
    // What lies beyond the region, gathered onto dense lanes so that few
    // wavefronts issue its loads.  The region is wx x wy lanes (wx, wy >= 3).
    // (1) The velocity of each edge cell's outward stress neighbour, nh of
    // them: lane h < nh takes the west column (LDS u-grid x = -1, y = h), the
    // east column (x = wx), then the south (y = -1) and north (y = wy) rows
    // without their corners (x = 1 .. wx-2).
    // (2) The x-face beyond the last column (wy of them) and the y-face beyond
    // the last row (wx-1 of them; the corner lane's is its x-face): lane
    // EX0 + e, the last wavefront(s) of the workgroup.
    const int wx = iend - i0 + 2, wy = jend - j0 + 2;
    const int nh = 2 * wy + 2 * (wx - 2);
    const bool do_o = tid < nh;
    int o_off = 0, opos = 0;
    if (do_o) {
        int hx, hy;
        if (tid < wy) { hx = -1; hy = tid; }
        else if (tid < 2 * wy) { hx = wx; hy = tid - wy; }
        else if (tid < 2 * wy + wx - 2) { hx = 1 + tid - 2 * wy; hy = -1; }
        else { hx = 1 + tid - 2 * wy - (wx - 2); hy = wy; }
        o_off = wrapc(i0 - 1 + hx, g.N1) + wrapc(j0 - 1 + hy, g.N2) * px;
        opos = (hy + 1) * UX + hx + 1;
    }
    constexpr int EX0 = NTHR - (RX + RY - 1);
    const int e = tid - EX0;
    const bool do_ex = e >= 0 && e < wx + wy - 1;
    const int ex_d = e < wy ? 0 : 1;       // orientation of the face beyond
    int ex_off = 0, expos = 0;
    if (do_ex) {
        if (ex_d == 0) {
            ex_off = wrapc(iend + 1, g.N1) + wrapc(j0 - 1 + e, g.N2) * px;
            expos = e * FX + wx;
        } else {
            ex_off = wrapc(i0 - 1 + e - wy, g.N1) + wrapc(jend + 1, g.N2) * px;
            expos = wy * FX + e - wy;
        }
    }
    Args A = a;

    // Forcing evaluated in place.  F_d(x) = sum_n 2 Re(a_nd e^{i m_n.theta}),
    // theta the per-axis phases of the cell centre.  The x and y factors are
    // fixed along the lane's column, so the 16 modes collapse by their z
    // component m_z in {-1, 0, 1, 2} into five coefficients per component:
    //   F_d = R0 + R1 cos(tz) + R2 sin(tz) + R3 cos(2 tz) + R4 sin(2 tz).
    float fr[3][5] = {};
    if (A.fmodes) {
        const float4 px4 = A.ftrig[iw - 1], py4 = A.ftrig[g.N1 + jw - 1];
        auto phase = [](const float4& t, int m, float& r, float& i) {
            r = m == 0 ? 1.f : m == 2 ? t.z : t.x;
            i = m == 0 ? 0.f : m == 2 ? t.w : m == 1 ? t.y : -t.y;
        };
#pragma unroll
        for (int n = 0; n < 16; ++n) {
            float xr, xi, yr, yi;
            phase(px4, forced_mode(n, 0), xr, xi);
            phase(py4, forced_mode(n, 1), yr, yi);
            const float er = xr * yr - xi * yi, ei = xr * yi + xi * yr;
            const int mz = forced_mode(n, 2);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float tr = A.fmodes[6 * n + 2 * d], ti = A.fmodes[6 * n + 2 * d + 1];
                const float qr = 2.f * (tr * er - ti * ei), qi = 2.f * (tr * ei + ti * er);
                if (mz == 0) {
                    fr[d][0] += qr;
                } else if (mz == 1) {
                    fr[d][1] += qr;
                    fr[d][2] -= qi;
                } else if (mz == -1) {
                    fr[d][1] += qr;
                    fr[d][2] += qi;
                } else {
                    fr[d][3] += qr;
                    fr[d][4] -= qi;
                }
            }
        }
    }

    // Plane p's inputs.  The x- and y-faces' t' come from the T3 record that
    // plane p-1 read with its top z-face (prev.tn); this plane's top record
    // gives zt's t' and hands the next plane its x and y t'.
    auto load_plane = [&](int p, const PlaneIn& prev) {
        PlaneIn r;
        const int zo = wrapc(p, g.N3) * sxy, zt_idx = base + wrapc(p + 1, g.N3) * sxy;
        r.xw = load_rec(A, 0, base + zo);
        r.ys = load_rec(A, 1, base + zo);
        r.zt = load_rec(A, 2, zt_idx);
        const T3 t3 = *trec_of(A.F[0], A.fb, (unsigned)zt_idx);
        r.xw.tp = prev.tn[0];
        r.ys.tp = prev.tn[1];
        r.zt.tp = t3.z;
        r.tn[0] = t3.x;
        r.tn[1] = t3.y;
        r.ex = FaceF{{0.f, 0.f, 0.f}, 0.f, 0.f};
        if (do_ex) r.ex = load_face(A, ex_d, ex_off + zo);
        r.ou[0] = r.ou[1] = r.ou[2] = 0.0;
        if (do_o)
        {
            const double2* rc = cell_rec(A.A, o_off + zo);
            r.ou[0] = ((const double*)rc)[1];
            const double2 u23 = rc[1];
            r.ou[1] = u23.x;
            r.ou[2] = u23.y;
        }
        r.fo[0] = r.fo[1] = r.fo[2] = 0.f;
        if (A.force) {
            for (int m = 0; m < 3; ++m) r.fo[m] = A.force[m * A.fstride + base + zo];
        } else if (A.fmodes) {
            const float4 tz = A.ftrig[g.N1 + g.N2 + wrapc(p, g.N3) - 1];
            for (int m = 0; m < 3; ++m)
                r.fo[m] = fr[m][0] + fr[m][1] * tz.x + fr[m][2] * tz.y + fr[m][3] * tz.z + fr[m][4] * tz.w;
        }
        return r;
    };
    auto put_face = [&](float* f, int q, const FaceF& r) {
        f[0 * FN + q] = r.u[0]; f[1 * FN + q] = r.u[1]; f[2 * FN + q] = r.u[2];
        f[3 * FN + q] = r.p; f[4 * FN + q] = r.tp;
    };
    // Plane p's faces and outward velocities into LDS (the lane's own
    // velocity is written separately, from registers that are already there).
    auto stage = [&](int p, const PlaneIn& r) {
        float* f = &s_f[0][0];
        if (in_region) {
            put_face(f, fpos, r.xw);
            put_face(f + 5 * FN, fpos, r.ys);
        }
        if (do_ex) put_face(f + ex_d * 5 * FN, expos, r.ex);
        if (do_o)
            for (int m = 0; m < 3; ++m) s_u[p & 1][m][opos] = (float)r.ou[m];
    };
    auto put_u = [&](int p, const CellD& c) {
        if (in_region)
            for (int m = 0; m < 3; ++m) s_u[p & 1][m][upos] = (float)c.u[m];
    };
    auto staged_face = [&](int q, int off) {
        const float* f = &s_f[off][0];
        FaceF r;
        r.u[0] = f[0 * FN + q]; r.u[1] = f[1 * FN + q]; r.u[2] = f[2 * FN + q];
        r.p = f[3 * FN + q]; r.tp = f[4 * FN + q];
        return r;
    };

    // prologue: plane k0-1 (the ring plane below the chunk)
    CellD Acur = load_cell(A, base + wrapc(k0 - 1, g.N3) * sxy);
    CellD Anext = load_cell(A, base + wrapc(k0, g.N3) * sxy);
    float sz_lo[3];
    {
        const CellD Ab = load_cell(A, base + wrapc(k0 - 2, g.N3) * sxy);
        for (int m = 0; m < 3; ++m) sz_lo[m] = A.VIS * ((float)Acur.u[m] - (float)Ab.u[m]) * A.inv_dx[2];
    }
    PlaneIn cur;
    FaceF zb;
    {
        const int idx = base + wrapc(k0 - 1, g.N3) * sxy;
        const T3 t3 = *trec_of(A.F[0], A.fb, (unsigned)idx);
        PlaneIn below;
        below.tn[0] = t3.x;
        below.tn[1] = t3.y;
        cur = load_plane(k0 - 1, below);
        zb = load_rec(A, 2, idx);
        zb.tp = t3.z;
    }
    put_u(k0 - 1, Acur);
    stage(k0 - 1, cur);
    // Retire every prologue load here, so that the loop's waits only ever
    // cover the plane being consumed, never the plane being prefetched.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // carried from plane p-1 to plane p
    double cons_prev[5] = {0, 0, 0, 0, 0};
    CellF Hprev{};
    FwdF fz_prev{};
    unsigned long long dtbits = ~0ull;
    unsigned bad = 0;

    for (int p = k0 - 1; p <= kend; ++p) {
        asm volatile("" : "+s"(A.fb));
        const bool inner = p >= k0 && p < kend;
        // plane p+1's velocities into LDS (it is read there from now on), then
        // prefetch plane p+1 (and the cell two planes up, for its z-stress)
        put_u(p + 1, Anext);
        PlaneIn nxt = cur;
        CellD Anext2 = Anext;
        if (p < kend) {
            nxt = load_plane(p + 1, cur);
            Anext2 = load_cell(A, base + wrapc(p + 2, g.N3) * sxy);
        }
        const CellF Af = to_f(Acur, A);

        // stress force (FP32 differences of the FP32-rounded velocities)
        float f[3];
        {
            const float roc = (float)Acur.ro;
            for (int m = 0; m < 3; ++m) {
                const float* u = &s_u[p & 1][m][0];
                const float uc = Af.u[m];
                const float uW = u[upos - 1], uE = u[upos + 1];
                const float uS = u[upos - UX], uN = u[upos + UX];
                const float sxl = A.VIS * (uc - uW) * A.inv_dx[0];
                const float sxh = A.VIS * (uE - uc) * A.inv_dx[0];
                const float syl = A.VIS * (uc - uS) * A.inv_dx[1];
                const float syh = A.VIS * (uN - uc) * A.inv_dx[1];
                const float szh = A.VIS * (s_u[(p + 1) & 1][m][upos] - uc) * A.inv_dx[2];
                f[m] = ((sxh - sxl) * A.ds[0] + (syh - syl) * A.ds[1] + (szh - sz_lo[m]) * A.ds[2]) * roc +
                       roc * A.dvcf * cur.fo[m];
                sz_lo[m] = szh;
            }
        }

        // predictor increments, half level (FP32) and the Phase2 extrapolations
        float Sp[5];
        CellF H;
        FwdF fx, fy, fz;
        BwdF bx, by, bz;
        float oz[3];
        {
            const FaceF Xw = staged_face(fpos, 0);
            const FaceF Xe = staged_face(fpos + 1, 0);
            const FaceF Ys = staged_face(fpos, 5);
            const FaceF Yn = staged_face(fpos + FX, 5);
            const FaceF& Zt = cur.zt;
            {
                const Flux a0 = flux<0>(Xw.u, Xw.p, Xw.tp, A), a1 = flux<0>(Xe.u, Xe.p, Xe.tp, A);
                const Flux b0 = flux<1>(Ys.u, Ys.p, Ys.tp, A), b1 = flux<1>(Yn.u, Yn.p, Yn.tp, A);
                const Flux c0 = flux<2>(zb.u, zb.p, zb.tp, A), c1 = flux<2>(Zt.u, Zt.p, Zt.tp, A);
                for (int v = 0; v < 5; ++v)
                    Sp[v] = (a1.m[v] - a0.m[v]) * A.ds[0] + (b1.m[v] - b0.m[v]) * A.ds[1] +
                            (c1.m[v] - c0.m[v]) * A.ds[2];
            }
            {
                const float roA = A.ro0 + Af.dr;
                H.dr = Af.dr - cf * Sp[0];
                const float inv_roH = rcpf(A.ro0 + H.dr);
                for (int m = 0; m < 3; ++m) H.u[m] = (roA * Af.u[m] - cf * (Sp[1 + m] - f[m])) * inv_roH;
                H.tp = (roA * Af.tp - cf * Sp[4]) * inv_roH;
            }
            {
                const Side s = make_side<0>(Af, H, Xw, Xe, A);
                fx = forward(s, dt, inv_dt2, A.inv_dx[0], A);
                bx = backward<Quirk>(s, dt, inv_dt2, A.inv_dx[0], A);
            }
            {
                const Side s = make_side<1>(Af, H, Ys, Yn, A);
                fy = forward(s, dt, inv_dt2, A.inv_dx[1], A);
                by = backward<false>(s, dt, inv_dt2, A.inv_dx[1], A);
            }
            {
                const Side s = make_side<2>(Af, H, zb, Zt, A);
                fz = forward(s, dt, inv_dt2, A.inv_dx[2], A);
                bz = backward<Quirk>(s, dt, inv_dt2, A.inv_dx[2], A);
            }
            oz[0] = zb.tp; oz[1] = zb.u[0]; oz[2] = zb.u[1];
        }
        if (owned && inner) {
            const float d1 = A.cfl_dx[0] * rcpf(A.c + fabsf(H.u[0]));
            const float d2 = A.cfl_dx[1] * rcpf(A.c + fabsf(H.u[1]));
            const float d3 = A.cfl_dx[2] * rcpf(A.c + fabsf(H.u[2]));
            const double dmin = (double)fminf(fminf(d1, d2), d3);
            const unsigned long long b = (unsigned long long)__double_as_longlong(dmin);
            dtbits = b < dtbits ? b : dtbits;
            // a NaN or Inf anywhere in the half level makes this sum non-finite
            bad += !isfinite(H.dr + H.u[0] + H.u[1] + H.u[2]);
        }

        // z-face p, then the FP64 closing update of plane p-1.  The new t' of
        // the lane's z-, x- and y-face (tz_s, tx_s, ty_s) go out together in
        // one T3 record once the plane's x- and y-faces are done.
        Flux mz{};
        float tx_s = 0.f, ty_s = 0.f, tz_s = 0.f;
        if (p >= k0) {
            const FaceOutF o = combine(fz_prev, bz, Hprev.u[2], H.u[2], Hprev.tp, H.tp, Hprev.u[0], H.u[0],
                                       Hprev.u[1], H.u[1], oz[0], oz[1], oz[2], A);
            mz = out_flux<2>(o, A);
            tz_s = o.t;
            if (owned) {
                if (p < kend) {
                    store_face<2>(A, at(g, i, j, p), o);
                    if (p == 1) {
                        store_face<2>(A, at(g, i, j, g.N3 + 1), o);
                        store_tp<2>(A, at(g, i, j, g.N3 + 1), o.t);
                    }
                }
                if (p > k0) {
                    double Sc[5];
                    for (int v = 0; v < 5; ++v) Sc[v] = cons_prev[v] - cd * (double)(mz.m[v] * A.ds[2]);
                    const double ro1 = Sc[0];
                    const double inv = rcpd(ro1);
                    const int co = at(g, i, j, p - 1);
                    double2* w = (double2*)cell_rec(A.Ao, co);
                    w[0] = make_double2(ro1, Sc[1] * inv);
                    w[1] = make_double2(Sc[2] * inv, Sc[3] * inv);
                    std_(A.Ao, A.fb, 4, co, A.t0d + Sc[4] * inv);
                }
            }
        }

        // x/y faces and the in-plane corrector flux sums
        float Sxy[5] = {0, 0, 0, 0, 0};
        if (p < kend) {
            if (inner) {
                s_h[tid] = make_float4(H.u[0], H.u[1], H.u[2], H.tp);
                s_fx[tid] = make_float4(fx.r, fx.t, fx.a, fx.b);
                s_fy[tid] = make_float4(fy.r, fy.t, fy.a, fy.b);
            }
            __syncthreads();
            Flux mx{}, my{};
            if (inner) {
                if (xface) {
                    const float4 nb = s_h[tW], fw = s_fx[tW];
                    const FwdF fwb{fw.x, fw.y, fw.z, fw.w};
                    // the old tangential values of the lane's own x-face, from LDS
                    const FaceOutF o = combine(fwb, bx, nb.x, H.u[0], nb.w, H.tp, nb.y, H.u[1], nb.z, H.u[2],
                                               s_f[4][fpos], s_f[1][fpos], s_f[2][fpos], A);
                    if (own_x) {
                        store_face<0>(A, at(g, i, j, p), o);
                        if (i == 1) {
                            store_face<0>(A, at(g, g.N1 + 1, j, p), o);
                            store_tp<0>(A, at(g, g.N1 + 1, j, p), o.t);
                        }
                    }
                    tx_s = o.t;
                    mx = out_flux<0>(o, A);
                }
                if (yface) {
                    const float4 nb = s_h[tS], fw = s_fy[tS];
                    const FwdF fwb{fw.x, fw.y, fw.z, fw.w};
                    const FaceOutF o = combine(fwb, by, nb.y, H.u[1], nb.w, H.tp, nb.x, H.u[0], nb.z, H.u[2],
                                               s_f[9][fpos], s_f[5][fpos], s_f[7][fpos], A);
                    if (own_y) {
                        store_face<1>(A, at(g, i, j, p), o);
                        if (j == 1) {
                            store_face<1>(A, at(g, i, g.N2 + 1, p), o);
                            store_tp<1>(A, at(g, i, g.N2 + 1, p), o.t);
                        }
                    }
                    ty_s = o.t;
                    my = out_flux<1>(o, A);
                }
                // an owned lane has computed all three of its faces here
                if (owned) *(T3*)trec_of(A.Fo[0], A.fb, (unsigned)at(g, i, j, p)) = T3{tx_s, ty_s, tz_s};
                for (int v = 0; v < 5; ++v) { s_m[v][tid] = mx.m[v]; s_m[5 + v][tid] = my.m[v]; }
            }
            stage(p + 1, nxt);
            __syncthreads();
            if (owned && inner)
                for (int v = 0; v < 5; ++v)
                    Sxy[v] = (s_m[v][tE] - mx.m[v]) * A.ds[0] + (s_m[5 + v][tN] - my.m[v]) * A.ds[1];
        }

        Hprev = H;
        fz_prev = fz;
        for (int v = 0; v < 5; ++v) {
            // this step's FP32 increments summed in FP32, accumulated in FP64
            float inc = Sp[v] + (Sxy[v] - mz.m[v] * A.ds[2]);
            if (v >= 1 && v <= 3) inc -= 2.f * f[v - 1];
            const double q = v == 0 ? Acur.ro : Acur.ro * (v == 4 ? Acur.t - A.t0d : Acur.u[v - 1]);
            cons_prev[v] = q - cd * (double)inc;
        }
        zb = cur.zt;
        cur = nxt;
        Acur = Anext;
        Anext = Anext2;
    }
    block_min_count(dtbits, bad, A.dt_next, A.nf, A.S->nstep);
}

// One box, one workgroup per chunk of a tile column.  units[u] = (tile
// column x fastest then y, first plane, end plane) lists the chunks in
// rounds: round q holds chunk q of every tile column, so the workgroups that
// run together work at the same height and their rings come from L2.  The
// host sizes the chunks (mixed_units) so the count is a multiple of the
// workgroups the chip holds at once, with near-equal heights, and no round
// finishes with the chip part-empty.  Die (XCD) b % 8 takes a contiguous run
// of chunks, so the columns in flight on one die are neighbours.
template <bool Quirk, class S>
__global__ void __launch_bounds__(S::NTHR) k_fused_mixed(Args a, const int4* __restrict__ units, int nbx) {
    const int4 w = units[die_major(gridDim.x)];
    mixed_tile<Quirk, S>(a, w.x % nbx, w.x / nbx, w.y, w.z);
}

// Many boxes in one launch.  start[b] is the first global tile of box b
// (start[nboxes] the total); the global tile of a workgroup follows the same
// die-major order as a single box, so each die works through one run of
// tiles, box after box.  Every box is computed exactly as by k_fused_mixed:
// the values of a cell do not depend on the tiling.
template <bool Quirk, class S>
__global__ void __launch_bounds__(S::NTHR) k_fused_mixed_batch(const Args* __restrict__ boxes,
                                                               const int* __restrict__ start, int nboxes) {
    const int t = die_major(start[nboxes]);
    int lo = 0, hi = nboxes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const Args& a = boxes[lo];
    const int tl = t - start[lo];
    const int nbx = (a.g.N1 + S::TX - 1) / S::TX, nby = (a.g.N2 + S::TY - 1) / S::TY;
    const int k0 = (tl / (nbx * nby)) * a.LZ + 1;
    mixed_tile<Quirk, S>(a, tl % nbx, (tl / nbx) % nby, k0, min(k0 + a.LZ, a.g.N3 + 1));
}

// dt of a new step for every box of a batch (k_cfl_finalize, mode 1, per box).
// Level-n cells between the per-field layout (fields 0-4 of the group at
// stride fb) and the stepping layout (cell_rec, t in field 4), from one
// buffer of the group into the other; n indices of the common index map.
// Packing converts every index.  Unpacking writes every index too: interior
// cells from their records, ghost cells (and row padding) from the interior
// cell they are a periodic image of, since the mixed step reads no ghost.
__global__ void k_cells_convert(Geo g, const char* __restrict__ src, char* __restrict__ dst, unsigned long long fb,
                                long long n, int pack) {
    const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (q >= n) return;
    double v[5];
    if (pack) {
        for (int f = 0; f < 5; ++f) v[f] = *(const double*)(src + f * fb + q * 8);
        double2* w = (double2*)(dst + 8 + q * 32);
        w[0] = make_double2(v[0], v[1]);
        w[1] = make_double2(v[2], v[3]);
        *(double*)(dst + 4 * fb + q * 8) = v[4];
        return;
    }
    const int k = (int)(q / g.sxy), rem = (int)(q - (long long)k * g.sxy);
    const int j = rem / g.px, i = rem - j * g.px;
    auto img = [](int a, int N) { return (((a - 1) % N) + N) % N + 1; };
    const long long qs = img(i, g.N1) + (long long)img(j, g.N2) * g.px + (long long)img(k, g.N3) * g.sxy;
    const double2* r = (const double2*)(src + 8 + qs * 32);
    const double2 a = r[0], b = r[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    v[4] = *(const double*)(src + 4 * fb + qs * 8);
    for (int f = 0; f < 5; ++f) *(double*)(dst + f * fb + q * 8) = v[f];
}

__global__ void k_batch_begin_step(Scalars* const* __restrict__ S, int nboxes) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nboxes) return;
    Scalars* s = S[b];
    const double dt = __longlong_as_double((long long)s->dt_bits);
    s->dt_bits = s->dt_init;
    s->dt = dt;
    s->nstep += 1;
    const double tm = s->time + 0.5 * dt;
    s->time = tm + 0.5 * dt;
}

// ---------------------------------------------------------------------------
// Conversions between an FP64 field (element (0,0,0) at src/dst, elements
// 0..n-1) and component v of the records of the orientation-d faces (group
// pointer grp; xg the x-faces' group pointer of the same buffer, which holds
// the T3 records): v < 4: u1, u2, u3, p; v = 4: t'.  To the records:
// t' = t - t0 for v = 4.  From them: t0 + t' for v = 4, and with rho set,
// rho0 + p/c^2 from p.
__global__ void k_face_from_double(const double* __restrict__ src, char* grp, char* xg, unsigned long long fb, int d,
                                   int v, long long n, double t0) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        if (v == 4) *(float*)tp_of(xg, fb, d, (unsigned)q) = (float)(src[q] - t0);
        else ((float*)rec_of(grp, (unsigned)q))[v] = (float)src[q];
    }
}
__global__ void k_face_to_double(const char* grp, const char* xg, unsigned long long fb, int d, int v, int rho,
                                 long long n, double t0, double ro0, double c2, double* __restrict__ dst) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        if (v == 4) {
            dst[q] = t0 + (double)*tp_of(xg, fb, d, (unsigned)q);
        } else {
            const double x = (double)((const float*)rec_of(grp, (unsigned)q))[v];
            dst[q] = rho ? ro0 + x / c2 : x;
        }
    }
}

// The reference's vorticity from the face velocities (source/main.cpp:
// 2712-2733, as k_vorticity) on the face records of the three orientations.
__global__ void k_vorticity_records(Geo g, const char* fx, const char* fy, const char* fz, double dx1, double dx2,
                                    double dx3, double* o1, double* o2, double* o3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > g.N1) return;
    const unsigned c = (unsigned)at(g, i, j, k);
    auto u = [&](const char* grp, unsigned q, int m) { return (double)((const float*)rec_of(grp, q))[m]; };
    const double d2u3 = (u(fy, c + g.px, 2) - u(fy, c, 2)) / dx2, d3u2 = (u(fz, c + g.sxy, 1) - u(fz, c, 1)) / dx3,
                 d3u1 = (u(fz, c + g.sxy, 0) - u(fz, c, 0)) / dx3, d1u3 = (u(fx, c + 1, 2) - u(fx, c, 2)) / dx1,
                 d1u2 = (u(fx, c + 1, 1) - u(fx, c, 1)) / dx1, d2u1 = (u(fy, c + g.px, 0) - u(fy, c, 0)) / dx2;
    o1[c] = d2u3 - d3u2;
    o2[c] = d3u1 - d1u3;
    o3[c] = d1u2 - d2u1;
}

}  // namespace mixed
}  // namespace arkd

// ---------------------------------------------------------------------------
using namespace arkd;

namespace {

// Tile shapes a launch can use (mixed::Shape): owned 30x6, 30x14, 62x6, 32x13.
constexpr int kShapes = 4;
constexpr int kShapeTX[kShapes] = {30, 30, 62, 32}, kShapeTY[kShapes] = {6, 14, 6, 13};
constexpr int kShapeThreads[kShapes] = {256, 512, 512, 510};
constexpr int kDefaultShape = 1;

int tiles_of(const Geo& g, int shape, int LZ) {
    return ((g.N1 + kShapeTX[shape] - 1) / kShapeTX[shape]) * ((g.N2 + kShapeTY[shape] - 1) / kShapeTY[shape]) *
           ((g.N3 + LZ - 1) / LZ);
}

// Kernel arguments of one mixed step of context c: read the current buffers,
// write the other ones; dt_next is where the next step's dt MIN goes.
// nf: where the non-finite count goes (Scalars::nf of the box, or a scratch
// pair for the timing launches).
mixed::Args make_args(const ark_ctx* c, int LZ, unsigned long long* dt_next, long long* nf, int flip = 0) {
    const ark_params& P = c->prm;
    mixed::Args a;
    a.g = c->geo;
    a.VIS = (float)P.VIS;
    a.c = (float)P.sound;
    a.c2 = (float)(P.sound * P.sound);
    a.inv_c2 = (float)(1.0 / (P.sound * P.sound));
    a.inv_rc = (float)(1.0 / (P.ro0 * P.sound));
    a.half_rc = (float)(0.5 * P.sound * P.ro0);
    a.ro0 = (float)P.ro0;
    const double dx[3] = {P.dx1, P.dx2, P.dx3};
    for (int d = 0; d < 3; ++d) {
        a.inv_dx[d] = (float)(1.0 / dx[d]);
        a.cfl_dx[d] = (float)(P.CFL * dx[d]);
    }
    a.ds[0] = (float)c->K.ds1;
    a.ds[1] = (float)c->K.ds2;
    a.ds[2] = (float)c->K.ds3;
    a.ro0d = P.ro0;
    a.t0d = P.t0;
    a.dvc = c->K.dvc;
    a.A = c->gbase[0][c->gpar[0] ^ flip];
    a.Ao = c->gbase[0][c->gpar[0] ^ flip ^ 1];
    for (int d = 0; d < 3; ++d) {
        a.F[d] = c->gbase[1 + d][c->gpar[1 + d] ^ flip];
        a.Fo[d] = c->gbase[1 + d][c->gpar[1 + d] ^ flip ^ 1];
    }
    a.fb = c->field_bytes;
    a.S = c->S;
    a.dt_next = dt_next;
    a.nf = nf;
    a.LZ = LZ;
    // the force field when a closure adds to it, else the forcing in place
    a.force = c->closure ? c->force : nullptr;
    a.fstride = c->force_stride;
    a.fmodes = (!c->closure && c->force_eps > 0.0) ? c->force_table : nullptr;
    a.ftrig = (const float4*)c->force_trig;
    a.dvcf = (float)c->K.dvc;
    return a;
}

template <class S>
void launch_single(bool quirk, int units, const int4* table, hipStream_t st, const mixed::Args& a) {
    const int nbx = (a.g.N1 + S::TX - 1) / S::TX;
    if (quirk) hipLaunchKernelGGL((mixed::k_fused_mixed<true, S>), dim3(units), dim3(S::NTHR), 0, st, a, table, nbx);
    else hipLaunchKernelGGL((mixed::k_fused_mixed<false, S>), dim3(units), dim3(S::NTHR), 0, st, a, table, nbx);
}

template <class S>
std::vector<int4> mixed_units(const Geo& g, int W) {
    return ark_chunk_table(g, S::TX, S::TY, W);
}

template <class S>
void launch_batch(bool quirk, dim3 grid, hipStream_t st, const mixed::Args* boxes, const int* start, int n) {
    if (quirk) hipLaunchKernelGGL((mixed::k_fused_mixed_batch<true, S>), grid, dim3(S::NTHR), 0, st, boxes, start, n);
    else hipLaunchKernelGGL((mixed::k_fused_mixed_batch<false, S>), grid, dim3(S::NTHR), 0, st, boxes, start, n);
}

template <class S>
int resident_blocks(bool quirk) {
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    const void* fn = quirk ? (const void*)mixed::k_fused_mixed<true, S> : (const void*)mixed::k_fused_mixed<false, S>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, S::NTHR, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return cus * per_cu;
}

hipError_t launch_shape_batch(int shape, bool quirk, dim3 grid, hipStream_t st, const mixed::Args* boxes,
                              const int* start, int n) {
    if (shape == 0) launch_batch<mixed::Shape30x6>(quirk, grid, st, boxes, start, n);
    else if (shape == 1) launch_batch<mixed::Shape30x14>(quirk, grid, st, boxes, start, n);
    else if (shape == 3) launch_batch<mixed::Shape32x13>(quirk, grid, st, boxes, start, n);
    else launch_batch<mixed::Shape62x6>(quirk, grid, st, boxes, start, n);
    return hipGetLastError();
}

// The single-box step's two tile shapes: 30x14 owned (32 x 16 threads) and
// 32x13 owned (34 x 15).  Which is faster depends on the box: at 256^3 the
// 32-wide tiles cover each row with 8 tiles instead of 9 (1.06 against
// 1.12 ms), at 320^3 they are slower (2.09 against 1.88 ms).
using SingleShape = mixed::Shape30x14;
using SingleWide = mixed::Shape32x13;

void launch_single_shape(const ark_ctx* c, int shape, const mixed::Args& a) {
    const bool quirk = c->prm.scheme == ARK_SCHEME_REFERENCE;
    if (shape == 1)
        launch_single<SingleWide>(quirk, c->mixed_nunits[1], (const int4*)c->mixed_units[1], c->stream, a);
    else
        launch_single<SingleShape>(quirk, c->mixed_nunits[0], (const int4*)c->mixed_units[0], c->stream, a);
}

hipError_t run_mixed(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    launch_single_shape(c, c->mixed_shape, make_args(c, 0, &c->S->dt_bits, c->S->nf));
    return hipGetLastError();
}

// The chunk tables of the single-box step, built once per context.
ark_status ensure_units(ark_ctx* c) {
    if (c->mixed_units[0]) return ARK_OK;
    const bool quirk = c->prm.scheme == ARK_SCHEME_REFERENCE;
    ark_status s = ark_upload_chunks(c, mixed_units<SingleShape>(c->geo, resident_blocks<SingleShape>(quirk)),
                                     &c->mixed_units[0], &c->mixed_nunits[0]);
    if (s == ARK_OK)
        s = ark_upload_chunks(c, mixed_units<SingleWide>(c->geo, resident_blocks<SingleWide>(quirk)),
                              &c->mixed_units[1], &c->mixed_nunits[1]);
    return s;
}

// Times both single-box shapes on this context and keeps the faster: one
// untimed launch of each, then kRounds rounds that time one launch of each
// shape in turn, each shape scored by its fastest launch (alternating and the
// minimum keep a passing disturbance from deciding the choice).  The timing
// launches write only the buffers the step is about to overwrite and send
// their dt MIN to a scratch slot, so the state is untouched; every cell's
// arithmetic is the same in either tiling, so the choice changes only the
// speed.
ark_status tune_single_shape(ark_ctx* c) {
    constexpr int kRounds = 4;
    unsigned long long* scratch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // three slots: the dt MIN, and the non-finite count and first step
    ark_status s = c->hip(hipMalloc(&scratch, 3 * sizeof(unsigned long long)), "hipMalloc tuning slot");
    if (s == ARK_OK) s = c->hip(hipEventCreate(&e0), "event");
    if (s == ARK_OK) s = c->hip(hipEventCreate(&e1), "event");
    const mixed::Args a = make_args(c, 0, scratch, (long long*)(scratch + 1));
    for (int sh = 0; sh < 2 && s == ARK_OK; ++sh) {
        launch_single_shape(c, sh, a);
        s = c->hip(hipGetLastError(), "tuning launch");
    }
    float best[2] = {0.f, 0.f};
    for (int r = 0; r < kRounds && s == ARK_OK; ++r)
        for (int sh = 0; sh < 2 && s == ARK_OK; ++sh) {
            s = c->hip(hipEventRecord(e0, c->stream), "event");
            if (s == ARK_OK) {
                launch_single_shape(c, sh, a);
                s = c->hip(hipGetLastError(), "tuning launch");
            }
            if (s == ARK_OK) s = c->hip(hipEventRecord(e1, c->stream), "event");
            if (s == ARK_OK) s = c->hip(hipEventSynchronize(e1), "event sync");
            float ms = 0.f;
            if (s == ARK_OK) s = c->hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            if (s == ARK_OK && (r == 0 || ms < best[sh])) best[sh] = ms;
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (scratch) (void)hipFree(scratch);
    if (s == ARK_OK) c->mixed_shape = best[1] < best[0] ? 1 : 0;
    return s;
}

// Group (1..3) and record component (0..4) of a face field.
int face_slot(int f, int* v) {
    for (int g = 1; g < 4; ++g)
        for (int q = 0; q < 5; ++q)
            if (kGroupFields[g][q] == f) {
                *v = q;
                return g;
            }
    return -1;
}

// Elements of a field in the common index map (the records cover these).
long long face_elems(const ark_ctx* c) { return (long long)c->elems - kOriginPad - 16; }

}  // namespace

// Chunks of a z-marching step with TX x TY owned columns per tile.  With T
// tile columns, N3 planes and W workgroups resident at once (CUs x
// workgroups per CU), U chunks give each column U/T or U/T + 1 near-equal
// chunks (+2 ring planes each).  U is taken among the multiples of W (or all
// T*N3 planes, if fewer) by the finishing time of in-order dispatch onto each
// die's share of the workgroups, in units of plane-steps.  The table lists
// the chunks in rounds: round q holds chunk q of every tile column.
std::vector<int4> ark_chunk_table(const Geo& g, int TX, int TY, int W) {
    const int nbx = (g.N1 + TX - 1) / TX, nby = (g.N2 + TY - 1) / TY;
    const long long T = (long long)nbx * nby;
    auto chunks_of = [&](long long U, long long t) { return (int)(U / T + (t < U % T ? 1 : 0)); };
    auto build = [&](long long U) {
        std::vector<int4> u;
        u.reserve((size_t)U);
        const int cmax = chunks_of(U, 0);
        for (int c = 0; c < cmax; ++c)
            for (long long t = 0; t < T; ++t) {
                const int n = chunks_of(U, t);
                if (c >= n) continue;
                const int k0 = 1 + (int)((long long)g.N3 * c / n), k1 = 1 + (int)((long long)g.N3 * (c + 1) / n);
                u.push_back(make_int4((int)t, k0, k1, 0));
            }
        return u;
    };
    auto finish = [&](const std::vector<int4>& u) {
        const long long n = (long long)u.size();
        const int per_die = std::max(1, W / kDies);
        long long worst = 0;
        for (int d = 0; d < kDies; ++d) {
            const long long q = n / kDies, r = n % kDies;
            const long long lo = d * q + std::min<long long>(d, r), hi = lo + q + (d < r ? 1 : 0);
            std::vector<long long> free_at(per_die, 0);
            for (long long i = lo; i < hi; ++i) {
                auto m = std::min_element(free_at.begin(), free_at.end());
                *m += u[i].z - u[i].y + 2;
            }
            worst = std::max(worst, *std::max_element(free_at.begin(), free_at.end()));
        }
        return worst;
    };
    const long long planes = T * g.N3;
    if (planes <= W) return build(planes);
    std::vector<int4> best;
    long long best_t = -1;
    for (long long U = W; U <= planes && U <= 64LL * W; U += W) {
        if (U < T) continue;
        std::vector<int4> u = build(U);
        const long long f = finish(u);
        if (best_t < 0 || f < best_t) {
            best_t = f;
            best.swap(u);
        }
    }
    if (best_t < 0) best = build(std::min<long long>(planes, ((T + W - 1) / W) * W));
    return best;
}

ark_status ark_upload_chunks(ark_ctx* c, const std::vector<int4>& u, void** dst, int* n) {
    void* d = nullptr;
    ark_status s = c->hip(hipMalloc(&d, u.size() * sizeof(int4)), "hipMalloc chunk table");
    if (s != ARK_OK) return s;
    c->allocs.push_back(d);
    c->bytes += u.size() * sizeof(int4);
    s = c->hip(hipMemcpy(d, u.data(), u.size() * sizeof(int4), hipMemcpyHostToDevice), "chunk table");
    if (s != ARK_OK) return s;
    *dst = d;
    *n = (int)u.size();
    return ARK_OK;
}

ark_status ark_set_tile_shape(ark_ctx* c, int shape) {
    if (!c) return ARK_ERR_INVALID;
    if (shape < -1 || shape > 1) return c->fail(ARK_ERR_INVALID, "tile shape must be -1 (timed), 0 (30x14) or 1 (32x13)");
    c->mixed_shape = shape;
    return ARK_OK;
}

ark_status ark_get_tile_shape(const ark_ctx* c, int* shape) {
    if (!c || !shape) return ARK_ERR_INVALID;
    *shape = c->mixed_shape;
    return ARK_OK;
}

ark_status ark_cells_layout(ark_ctx* c, bool records) {
    if (c->prm.precision != ARK_PREC_MIXED || c->cells_rec == records) return ARK_OK;
    const long long n = (long long)c->geo.sxy * (c->geo.N3 + 3);
    hipLaunchKernelGGL(mixed::k_cells_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->geo,
                       (const char*)c->gbase[0][c->gpar[0]], c->gbase[0][c->gpar[0] ^ 1], c->field_bytes, n,
                       records ? 1 : 0);
    ark_status s = c->hip(hipGetLastError(), records ? "cells to records" : "cells to fields");
    if (s != ARK_OK) return s;
    ark_swap_group(c, 0);
    c->cells_rec = records;
    return ARK_OK;
}

ark_status ark_mixed_steps(ark_ctx* c, int nsteps) {
    if (c->geo.N1 < 2 || c->geo.N2 < 2 || c->geo.N3 < 2)
        return c->fail(ARK_ERR_UNSUPPORTED, "mixed precision needs N1, N2, N3 >= 2");
    ark_status s = ARK_OK;
    s = ensure_units(c);
    for (int n = 0; s == ARK_OK && n < nsteps; ++n) {
        s = ark_begin_step(c);
        if (s == ARK_OK) s = ark_cells_layout(c, true);
        // the first step of a context picks its tile shape, with this step's dt
        if (s == ARK_OK && c->mixed_shape < 0) s = tune_single_shape(c);
        if (s == ARK_OK) s = ark_turb_force_step(c, c->stream);
        if (s == ARK_OK) s = ark_closure_step(c, c->stream);
        if (s == ARK_OK) s = c->launch("fused_mixed", run_mixed, c);
        if (s != ARK_OK) break;
        for (int g = 0; g < 4; ++g) ark_swap_group(c, g);
        c->dt_pending = true;
    }
    return s;
}

// The FP32 storage of one face field, as seen through the FP64 API (the face
// densities ro1..ro3 from the pressures).
ark_status ark_mixed_face_to_host_layout(ark_ctx* c, int f, double* staging) {
    const ark_params& P = c->prm;
    const bool rho = f == F_ro1 || f == F_ro2 || f == F_ro3;
    const int fld = f == F_ro1 ? F_p1 : f == F_ro2 ? F_p2 : f == F_ro3 ? F_p3 : f;
    int v = 0;
    const int g = face_slot(fld, &v);
    if (g < 0) return c->fail(ARK_ERR_INVALID, "not a face field");
    hipLaunchKernelGGL(mixed::k_face_to_double, dim3(2048), dim3(256), 0, c->stream, c->gbase[g][c->gpar[g]],
                       c->gbase[1][c->gpar[1]], (unsigned long long)c->field_bytes, g - 1, v, rho ? 1 : 0,
                       face_elems(c), P.t0, P.ro0, P.sound * P.sound, staging);
    return c->hip(hipGetLastError(), "face to double");
}

// FP64 values (src, field layout) into the records of face field f in the
// group buffers named by `which` (1: current, 2: other, 3: both).
static ark_status face_from(ark_ctx* c, int f, const double* src, int which) {
    int v = 0;
    const int g = face_slot(f, &v);
    if (g < 0) return c->fail(ARK_ERR_INVALID, "not a face field");
    for (int b = 0; b < 2; ++b) {
        if (!(which & (1 << b))) continue;
        hipLaunchKernelGGL(mixed::k_face_from_double, dim3(2048), dim3(256), 0, c->stream, src,
                           c->gbase[g][c->gpar[g] ^ b], c->gbase[1][c->gpar[1] ^ b], (unsigned long long)c->field_bytes,
                           g - 1, v, face_elems(c), c->prm.t0);
        ark_status s = c->hip(hipGetLastError(), "face from double");
        if (s != ARK_OK) return s;
    }
    return ARK_OK;
}

ark_status ark_mixed_face_from_host_layout(ark_ctx* c, int f, const double* staging) {
    return face_from(c, f, staging, 3);
}

// After an initial state written as FP64 fields into the current buffer of
// each face group: the records of both buffers.  The records overlap the FP64
// fields of their own allocation, and the x group's also holds every
// orientation's t' (T3 records), so all three groups are first copied to
// their other buffer, then converted from that FP64 copy into the current
// buffer, then copied over it.
ark_status ark_mixed_faces_from_fp64(ark_ctx* c) {
    const size_t group_bytes = 5 * c->field_bytes;
    ark_status s = ARK_OK;
    auto copy_groups = [&]() {
        for (int g = 1; g < 4 && s == ARK_OK; ++g) {
            char* cur = c->gbase[g][c->gpar[g]] - kOriginPad * sizeof(double);
            char* oth = c->gbase[g][c->gpar[g] ^ 1] - kOriginPad * sizeof(double);
            s = c->hip(hipMemcpyAsync(oth, cur, group_bytes, hipMemcpyDeviceToDevice, c->stream), "copy faces");
        }
    };
    copy_groups();
    for (int g = 1; g < 4 && s == ARK_OK; ++g)
        for (int v = 0; v < 5 && s == ARK_OK; ++v) s = face_from(c, kGroupFields[g][v], c->alt[kGroupFields[g][v]], 1);
    copy_groups();
    return s;
}

ark_status ark_mixed_vorticity(ark_ctx* c, double* o1, double* o2, double* o3) {
    const Geo& g = c->geo;
    hipLaunchKernelGGL(mixed::k_vorticity_records, dim3((g.N1 + 255) / 256, g.N2, g.N3), dim3(256), 0, c->stream, g,
                       c->gbase[1][c->gpar[1]], c->gbase[2][c->gpar[2]], c->gbase[3][c->gpar[3]], c->prm.dx1,
                       c->prm.dx2, c->prm.dx3, o1, o2, o3);
    return c->hip(hipGetLastError(), "vorticity");
}

// ---------------------------------------------------------------------------
// Batched mixed-precision steps: many independent boxes in one launch per
// step, the launches replayed from two captured HIP graphs (one per buffer
// parity), tile shape and chunk length chosen by timing the candidates.
struct ark_batch {
    std::vector<ark_ctx*> box;
    hipStream_t stream = nullptr;
    int device = 0;
    bool quirk = true;
    int shape = kDefaultShape, lz_target = 32;
    int tiles = 0;
    mixed::Args* args[2] = {nullptr, nullptr};   // device: parity 0 (current -> other), parity 1
    int* start = nullptr;                         // device: first tile of each box, total last
    Scalars** sptr = nullptr;                     // device: each box's scalars
    unsigned long long* scratch_dt = nullptr;     // device: dt MIN and non-finite slots of the tuning launches,
                                                  // three per box
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    ark_force_batch* forcing = nullptr;           // the forced boxes' mode tables, one launch per step
    bool built = false;
    std::vector<std::array<int, 4>> gpar_at_build;   // box group parities the tables were built for
    int parity = 0;                               // which table the next step uses
    std::vector<double> tune_ms;                  // ms per step of each candidate (shape-major)
    std::string err;

    ark_status fail(ark_status s, const std::string& m) {
        err = m;
        return s;
    }
    ark_status hip(hipError_t e, const char* what) {
        if (e == hipSuccess) return ARK_OK;
        return fail(ARK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
};

namespace {

int box_lz(const ark_batch* b, const ark_ctx* c) { return std::min(b->lz_target, c->geo.N3); }

void drop_graphs(ark_batch* b) {
    for (int p = 0; p < 2; ++p) {
        if (b->exec[p]) (void)hipGraphExecDestroy(b->exec[p]);
        if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
        b->exec[p] = nullptr;
        b->graph[p] = nullptr;
    }
    b->built = false;
}

// Argument tables for the boxes' current buffers (and, for the tuning
// launches, dt MINs redirected to scratch slots), the tile table, and the two
// one-step graphs.
ark_status build_tables(ark_batch* b, bool tuning) {
    const int n = (int)b->box.size();
    std::vector<int> start(n + 1, 0);
    std::vector<mixed::Args> h[2];
    for (int p = 0; p < 2; ++p) h[p].resize(n);
    for (int i = 0; i < n; ++i) {
        ark_ctx* c = b->box[i];
        const int lz = box_lz(b, c);
        start[i + 1] = start[i] + tiles_of(c->geo, b->shape, lz);
        for (int p = 0; p < 2; ++p)
            h[p][i] = tuning ? make_args(c, lz, b->scratch_dt + 3 * i, (long long*)(b->scratch_dt + 3 * i + 1), p)
                             : make_args(c, lz, &c->S->dt_bits, c->S->nf, p);
    }
    b->tiles = start[n];
    ark_status s = ARK_OK;
    for (int p = 0; p < 2 && s == ARK_OK; ++p)
        s = b->hip(hipMemcpyAsync(b->args[p], h[p].data(), n * sizeof(mixed::Args), hipMemcpyHostToDevice,
                                  b->stream), "batch args");
    if (s == ARK_OK)
        s = b->hip(hipMemcpyAsync(b->start, start.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, b->stream),
                   "batch tiles");
    if (s == ARK_OK) s = b->hip(hipStreamSynchronize(b->stream), "batch tables");
    return s;
}

ark_status enqueue_step(ark_batch* b, int parity, bool begin) {
    const int n = (int)b->box.size();
    if (begin) {
        hipLaunchKernelGGL(mixed::k_batch_begin_step, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->sptr, n);
        ark_status s = b->hip(hipGetLastError(), "batch begin step");
        if (s == ARK_OK && ark_force_batch_enqueue(b->forcing, b->stream) != ARK_OK)
            s = b->fail(ARK_ERR_HIP, "batch forcing");
        if (s != ARK_OK) return s;
    }
    return b->hip(launch_shape_batch(b->shape, b->quirk, dim3(b->tiles), b->stream, b->args[parity], b->start, n),
                  "batch step");
}

ark_status build_graphs(ark_batch* b) {
    drop_graphs(b);
    ark_status s = build_tables(b, false);
    if (s == ARK_OK && ark_force_batch_build(b->box.data(), (int)b->box.size(), &b->forcing) != ARK_OK)
        s = b->fail(ARK_ERR_HIP, "batch forcing tables");
    for (int p = 0; p < 2 && s == ARK_OK; ++p) {
        s = b->hip(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
        if (s != ARK_OK) break;
        ark_status se = enqueue_step(b, p, true);
        s = b->hip(hipStreamEndCapture(b->stream, &b->graph[p]), "end capture");
        if (se != ARK_OK) s = se;
        if (s == ARK_OK) s = b->hip(hipGraphInstantiate(&b->exec[p], b->graph[p], nullptr, nullptr, 0), "instantiate");
    }
    if (s != ARK_OK) return s;
    b->gpar_at_build.clear();
    for (ark_ctx* c : b->box) b->gpar_at_build.push_back({c->gpar[0], c->gpar[1], c->gpar[2], c->gpar[3]});
    b->parity = 0;
    b->built = true;
    return ARK_OK;
}

// Every box's work so far is done, and every box has the next step's dt as a
// pending MIN (a box that has not stepped yet gets it from a CFL scan).
ark_status prepare_boxes(ark_batch* b) {
    for (ark_ctx* c : b->box) {
        if (!c->dt_pending) {
            ark_status s = ark_cfl_scan(c);
            if (s != ARK_OK) return b->fail(s, std::string("batch: CFL scan of a box: ") + c->err);
            c->dt_pending = true;
        }
        ark_status s = ark_cells_layout(c, true);
        if (s != ARK_OK) return b->fail(s, std::string("batch: cell layout of a box: ") + c->err);
        s = b->hip(hipStreamSynchronize(c->stream), "box stream");
        if (s != ARK_OK) return s;
    }
    return ARK_OK;
}

// The boxes' streams wait for the batch's work.
ark_status release_boxes(ark_batch* b) {
    hipEvent_t ev;
    ark_status s = b->hip(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
    if (s != ARK_OK) return s;
    s = b->hip(hipEventRecord(ev, b->stream), "event record");
    for (ark_ctx* c : b->box)
        if (s == ARK_OK) s = b->hip(hipStreamWaitEvent(c->stream, ev, 0), "box wait");
    (void)hipEventDestroy(ev);
    return s;
}

}  // namespace

extern "C" {

ark_status ark_batch_create(ark_ctx* const* boxes, int nboxes, ark_batch** out) {
    if (!boxes || nboxes < 1 || !out) return ARK_ERR_INVALID;
    *out = nullptr;
    ark_batch* b = new ark_batch();
    ark_status s = ARK_OK;
    for (int i = 0; i < nboxes && s == ARK_OK; ++i) {
        ark_ctx* c = boxes[i];
        if (!c) s = b->fail(ARK_ERR_INVALID, "batch: null box");
        else if (c->prm.precision != ARK_PREC_MIXED)
            s = b->fail(ARK_ERR_UNSUPPORTED, "batch: boxes must use mixed precision");
        else if (c->geo.N1 < 2 || c->geo.N2 < 2 || c->geo.N3 < 2)
            s = b->fail(ARK_ERR_UNSUPPORTED, "batch: boxes need N1, N2, N3 >= 2");
        else if (c->has_solid) s = b->fail(ARK_ERR_UNSUPPORTED, "batch: boxes must be all fluid");
        else if (c->closure) s = b->fail(ARK_ERR_UNSUPPORTED, "batch: boxes with a learned closure step alone");
        else if (i > 0 && (c->prm.device != boxes[0]->prm.device || c->prm.scheme != boxes[0]->prm.scheme))
            s = b->fail(ARK_ERR_UNSUPPORTED, "batch: boxes must share the device and the scheme");
        for (int j = 0; j < i && s == ARK_OK; ++j)
            if (boxes[j] == c) s = b->fail(ARK_ERR_INVALID, "batch: a box appears twice");
        if (s == ARK_OK) b->box.push_back(c);
    }
    if (s == ARK_OK) {
        b->device = boxes[0]->prm.device;
        b->quirk = boxes[0]->prm.scheme == ARK_SCHEME_REFERENCE;
        s = b->hip(hipSetDevice(b->device), "hipSetDevice");
    }
    if (s == ARK_OK) s = b->hip(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
    const size_t n = b->box.size();
    for (int p = 0; p < 2 && s == ARK_OK; ++p)
        s = b->hip(hipMalloc(&b->args[p], n * sizeof(mixed::Args)), "hipMalloc batch args");
    if (s == ARK_OK) s = b->hip(hipMalloc(&b->start, (n + 1) * sizeof(int)), "hipMalloc batch tiles");
    if (s == ARK_OK) s = b->hip(hipMalloc(&b->sptr, n * sizeof(Scalars*)), "hipMalloc batch scalars");
    if (s == ARK_OK) s = b->hip(hipMalloc(&b->scratch_dt, 3 * n * sizeof(unsigned long long)), "hipMalloc batch dt");
    if (s == ARK_OK) {
        std::vector<Scalars*> sp;
        for (ark_ctx* c : b->box) sp.push_back(c->S);
        s = b->hip(hipMemcpy(b->sptr, sp.data(), n * sizeof(Scalars*), hipMemcpyHostToDevice), "batch scalars");
    }
    if (s != ARK_OK) {
        ark_batch_destroy(b);
        return s;
    }
    *out = b;
    return ARK_OK;
}

void ark_batch_destroy(ark_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    ark_force_batch_free(b->forcing);
    for (int p = 0; p < 2; ++p)
        if (b->args[p]) (void)hipFree(b->args[p]);
    if (b->start) (void)hipFree(b->start);
    if (b->sptr) (void)hipFree(b->sptr);
    if (b->scratch_dt) (void)hipFree(b->scratch_dt);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* ark_batch_last_error(const ark_batch* b) { return b ? b->err.c_str() : "null batch"; }

ark_status ark_batch_set_config(ark_batch* b, int shape, int lz) {
    if (!b || shape < 0 || shape >= kShapes || lz < 1) return ARK_ERR_INVALID;
    b->shape = shape;
    b->lz_target = lz;
    drop_graphs(b);
    return ARK_OK;
}

ark_status ark_batch_get_config(const ark_batch* b, int* shape, int* lz, int* tiles, int* threads) {
    if (!b) return ARK_ERR_INVALID;
    if (shape) *shape = b->shape;
    if (lz) *lz = b->lz_target;
    if (tiles) *tiles = b->tiles;
    if (threads) *threads = kShapeThreads[b->shape];
    return ARK_OK;
}

ark_status ark_batch_autotune(ark_batch* b, int reps, double* ms_out, int max_out) {
    if (!b || reps < 1) return ARK_ERR_INVALID;
    ark_status s = b->hip(hipSetDevice(b->device), "hipSetDevice");
    if (s == ARK_OK) s = prepare_boxes(b);
    if (s != ARK_OK) return s;
    const int lzs[3] = {16, 32, 64};
    b->tune_ms.assign(kShapes * 3, 0.0);
    hipEvent_t e0, e1;
    s = b->hip(hipEventCreate(&e0), "event");
    if (s == ARK_OK) s = b->hip(hipEventCreate(&e1), "event");
    if (s != ARK_OK) return s;
    double best = 1e300;
    int bs = kDefaultShape, bl = 32;
    drop_graphs(b);
    for (int sh = 0; sh < kShapes && s == ARK_OK; ++sh) {
        for (int l = 0; l < 3 && s == ARK_OK; ++l) {
            b->shape = sh;
            b->lz_target = lzs[l];
            // Tuning launches write only the boxes' other buffers (never
            // swapped in) and their dt MINs go to scratch slots: the boxes'
            // state is left as it was.
            s = build_tables(b, true);
            if (s == ARK_OK) s = enqueue_step(b, 0, false);
            if (s == ARK_OK) s = b->hip(hipEventRecord(e0, b->stream), "event");
            for (int r = 0; r < reps && s == ARK_OK; ++r) s = enqueue_step(b, 0, false);
            if (s == ARK_OK) s = b->hip(hipEventRecord(e1, b->stream), "event");
            if (s == ARK_OK) s = b->hip(hipEventSynchronize(e1), "event sync");
            float ms = 0.f;
            if (s == ARK_OK) s = b->hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            const double per = (double)ms / reps;
            b->tune_ms[sh * 3 + l] = per;
            if (s == ARK_OK && per < best) {
                best = per;
                bs = sh;
                bl = lzs[l];
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->shape = bs;
    b->lz_target = bl;
    if (ms_out)
        for (int q = 0; q < max_out && q < (int)b->tune_ms.size(); ++q) ms_out[q] = b->tune_ms[q];
    return s;
}

ark_status ark_batch_step(ark_batch* b, int nsteps) {
    if (!b || nsteps < 0) return ARK_ERR_INVALID;
    if (nsteps == 0) return ARK_OK;
    ark_status s = b->hip(hipSetDevice(b->device), "hipSetDevice");
    if (s == ARK_OK) s = prepare_boxes(b);
    if (s != ARK_OK) return s;
    // The tables hold buffer addresses: rebuild them when a box's buffers
    // have swapped since other than by the batch's own steps (a step on its
    // own, a change of cell layout), or when the batch is new.
    bool same = b->built;
    for (size_t i = 0; same && i < b->box.size(); ++i)
        for (int g = 0; g < 4; ++g) same = same && b->box[i]->gpar[g] == (b->gpar_at_build[i][g] ^ b->parity);
    if (!same) s = build_graphs(b);
    for (int n = 0; n < nsteps && s == ARK_OK; ++n) {
        s = b->hip(hipGraphLaunch(b->exec[b->parity], b->stream), "graph launch");
        b->parity ^= 1;
        for (ark_ctx* c : b->box)
            for (int g = 0; g < 4; ++g) ark_swap_group(c, g);
    }
    if (s == ARK_OK) s = release_boxes(b);
    return s;
}

ark_status ark_batch_synchronize(ark_batch* b) {
    if (!b) return ARK_ERR_INVALID;
    return b->hip(hipStreamSynchronize(b->stream), "batch sync");
}

}  // extern "C"

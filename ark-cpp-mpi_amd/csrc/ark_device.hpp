// Device-side definitions shared by every kernel of the CABARET/ARK step.
//
// Layout in HBM: structure of arrays, one allocation per reference array,
// x fastest.  Cell and face arrays share ONE index map
//     idx(i,j,k) = i + j*px + k*sxy,  px = pitch along x, sxy = px*(N2+3)
// so a thread computes one index and uses it for every field it touches.
// Cell arrays use i in [0, N1+1], face arrays i in [0, N1+2] (the reference's
// extents, source/main.cpp:203-277); px >= N1+3 is padded so that the
// interior element i=1 of every row starts a 128-byte line.
//
// Arithmetic: every expression below keeps the reference's operand order and
// association.  The library is compiled with -ffp-contract=off, so nothing is
// fused into an FMA and FP64 results are bitwise equal to the reference's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace arkd {

struct Geo {
    int N1, N2, N3;  // interior cells per axis (the reference's n_d - 1)
    int px;          // x pitch in elements
    int sxy;         // elements per z-plane
};

__device__ __forceinline__ int at(const Geo& g, int i, int j, int k) { return i + j * g.px + k * g.sxy; }

// Scalars that live on the device so consecutive steps never wait on the host.
struct Scalars {
    double dt;                    // current dt (the reference global, headers/constants.h:20)
    double time;                  // TIME
    long long nstep;              // nStep
    unsigned long long dt_bits;   // running MIN of the CFL bound, as the bits of a positive double
    unsigned long long dt_init;   // bits of min(1e8, dtv): the value dt_bits is reset to
    // Non-finite detection, folded into the CFL reduction of every step
    // (TimeStepSize, source/main.cpp:374-415, has none): nf[0] counts cells
    // seen with a non-finite velocity (or, in the fused steps, density), nf[1]
    // is the first step number nStep at which one was seen (LLONG_MAX: none).
    long long nf[2];
};

// Field slots, same order as ark_field in include/ark/ark.h.
enum {
    F_roCon, F_u1Con, F_u2Con, F_u3Con, F_tCon,
    F_ronCon, F_u1nCon, F_u2nCon, F_u3nCon, F_tnCon,
    F_f1, F_f2, F_f3,
    F_condition,
    F_ro1, F_t1, F_u11, F_u21, F_u31, F_p1,
    F_ro2, F_t2, F_u12, F_u22, F_u32, F_p2,
    F_ro3, F_t3, F_u13, F_u23, F_u33, F_p3,
    F_sigm11, F_sigm21, F_sigm31,
    F_sigm12, F_sigm22, F_sigm32,
    F_sigm13, F_sigm23, F_sigm33,
    F_NUM
};

struct Fields {
    double* a[F_NUM];
};

// Face slots along axis d: velocity component m (0..2) on faces normal to d,
// and the pressure, density and t of those faces.
__host__ __device__ constexpr int face_u(int d, int m) { return (d == 0 ? F_u11 : d == 1 ? F_u12 : F_u13) + m; }
__host__ __device__ constexpr int face_p(int d) { return d == 0 ? F_p1 : d == 1 ? F_p2 : F_p3; }
__host__ __device__ constexpr int face_ro(int d) { return d == 0 ? F_ro1 : d == 1 ? F_ro2 : F_ro3; }
__host__ __device__ constexpr int face_t(int d) { return d == 0 ? F_t1 : d == 1 ? F_t2 : F_t3; }

// min3d / max3d, source/main.cpp:2963-2973: the ternaries are kept because
// their NaN and signed-zero behaviour differs from fmin/fmax.
__device__ __forceinline__ double min3d(double x1, double x2, double x3) {
    x1 = x1 > x2 ? x2 : x1;
    x1 = x1 > x3 ? x3 : x1;
    return x1;
}
__device__ __forceinline__ double max3d(double x1, double x2, double x3) {
    x1 = x1 < x2 ? x2 : x1;
    x1 = x1 < x3 ? x3 : x1;
    return x1;
}

__device__ __forceinline__ unsigned long long umin64(unsigned long long a, unsigned long long b) {
    return a < b ? a : b;
}

// Block-wide MIN of `bits` into *slot: an LDS atomic per thread (valid for
// any block size, including a partial last wavefront, whose inactive lanes a
// cross-lane shuffle would read), then one global atomic per block.
__device__ __forceinline__ void block_min_atomic(unsigned long long bits, unsigned long long* slot) {
    __shared__ unsigned long long block_min;
    if (threadIdx.x == 0) block_min = ~0ull;
    __syncthreads();
    if (bits != ~0ull) atomicMin(&block_min, bits);
    __syncthreads();
    if (threadIdx.x == 0 && block_min != ~0ull) atomicMin(slot, block_min);
}

// The same MIN, plus the count of the block's non-finite cells (`bad` per
// thread) added to nf[0], with nf[1] lowered to `nstep` when there are any.
// The count costs one LDS atomic per thread that saw a bad cell, which a
// finite state never does.
__device__ __forceinline__ void block_min_count(unsigned long long bits, unsigned bad, unsigned long long* slot,
                                                long long* nf, long long nstep) {
    __shared__ unsigned long long block_min;
    __shared__ unsigned block_bad;
    if (threadIdx.x == 0) {
        block_min = ~0ull;
        block_bad = 0u;
    }
    __syncthreads();
    if (bits != ~0ull) atomicMin(&block_min, bits);
    if (bad) atomicAdd(&block_bad, bad);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (block_min != ~0ull) atomicMin(slot, block_min);
        if (block_bad) {
            atomicAdd((unsigned long long*)&nf[0], (unsigned long long)block_bad);
            atomicMin(&nf[1], nstep);
        }
    }
}

// Work order of the z-marching kernels.  The launch is one-dimensional, and
// the hardware hands workgroup b to die (XCD) b % 8, each with its own L2.
// die_major(T) maps b to a unit in [0, T) so that each die takes one
// contiguous run of units: units numbered with neighbouring tiles adjacent
// are then in flight on the same die, and the one-cell rings they share come
// from that die's L2 instead of HBM.  It is a bijection on [0, T) for any T.
constexpr int kDies = 8;
__device__ __forceinline__ int die_major(int T) {
    const int b = blockIdx.x;
    const int die = b % kDies, local = b / kDies;
    return die * (T / kDies) + min(die, T % kDies) + local;
}

// Numeric constants of a context.
struct Consts {
    double dx[3];
    double ds1, ds2, ds3, dvc;   // face areas and cell volume (source/main.cpp:455-460)
    double VIS, sound, ro0, t0;
    // correctly rounded reciprocals of the divisors the fused step uses:
    // 1/dx[d], 1/(sound*sound), 1/(ro0*sound), 1/dvc
    double rdx[3], rc2, rrs, rdvc;
};

// x / b from rb = 1/b (correctly rounded): the IEEE quotient, bit for bit,
// whenever that is finite and not subnormal.  q0 = x*rb lies within 1.5 ulp
// of x/b; one residual correction fma(fma(-q, b, x), rb, q) brings it within
// one ulp, and a second one then yields the correctly rounded quotient
// (Markstein, IBM J. Res. Develop. 34 (1990) 111-119).  The sign of a zero
// quotient is that of x*rb, which is the IEEE sign.  An overflowing quotient
// comes out NaN where IEEE division gives inf.  Five FP64 operations at full
// rate, against eleven, one of them a quarter-rate reciprocal, for a division.
__device__ __forceinline__ double div_rn(double x, double b, double rb) {
    const double q0 = x * rb;
    const double q1 = fma(fma(-q0, b, x), rb, q0);
    return copysign(fma(fma(-q1, b, x), rb, q1), q0);
}

// ---------------------------------------------------------------------------
// Phase2 cell extrapolation (source/main.cpp:1747-1874 and the X2/X3
// analogues).  The sweep along axis d computes, for every cell, limited
// extrapolations of the Riemann invariants and of the passive quantities to
// its forward face (r, t, tangential a, tangential b) and to its backward face
// (q, t, a, b).  "a"/"b" are the two tangential velocity components in
// increasing axis order, which is the order of the reference's line buffers.
struct CellSide {
    double un_f, un_b, un_cn, un_c;   // normal velocity: faces f, b; centre n+1/2, n
    double ua_f, ua_b, ua_cn, ua_c;   // tangential a
    double ub_f, ub_b, ub_cn, ub_c;   // tangential b
    double ro_cn, ro_c;
    double t_f, t_b, t_cn, t_c;
    double p_f, p_b;
};

struct Fwd { double r, t, a, b; };   // rBuf[i+1], tfBuf[i+1], u2fBuf[i+1], u3fBuf[i+1]
struct Bwd { double q, t, a, b; };   // qBuf[i],  tbBuf[i],  u2bBuf[i],  u3bBuf[i]

__device__ __forceinline__ void clip(double& v, double vmax, double vmin) {
    if (v > vmax) v = vmax;
    if (v < vmin) v = vmin;
}

// The divisions of the shared expressions below: IEEE division (Rcp false),
// or div_rn with the correctly rounded reciprocals (the same bits for every
// finite quotient).  rdt and rdx are 1/dt and 1/dx; 1/(ro0*sound) and
// 1/(sound*sound) come from K.
template <bool Rcp>
struct Divide {
    double dt, rdt, dx, rdx;
    const Consts& K;
    __device__ __forceinline__ double by_dt(double x) const { return Rcp ? div_rn(x, dt, rdt) : x / dt; }
    __device__ __forceinline__ double by_dx(double x) const { return Rcp ? div_rn(x, dx, rdx) : x / dx; }
    __device__ __forceinline__ double by_rs(double x) const {
        return Rcp ? div_rn(x, K.ro0 * K.sound, K.rrs) : x / (K.ro0 * K.sound);
    }
    __device__ __forceinline__ double by_c2(double x) const {
        return Rcp ? div_rn(x, K.sound * K.sound, K.rc2) : x / (K.sound * K.sound);
    }
};

// Forward extrapolations of one cell.
template <bool Rcp = false>
__device__ __forceinline__ Fwd cell_forward(const CellSide& s, double dt, double dx, const Consts& K, double rdt = 0.0,
                                            double rdx = 0.0) {
    const Divide<Rcp> D{dt, rdt, dx, rdx, K};
    const double sound = K.sound, ro0_g = K.ro0;
    double p_cn = sound * sound * (s.ro_cn - ro0_g);
    double p_c = sound * sound * (s.ro_c - ro0_g);
    double r_f = s.un_f + D.by_rs(s.p_f);
    double r_b = s.un_b + D.by_rs(s.p_b);
    double r_cn = s.un_cn + D.by_rs(p_cn);
    double r_c = s.un_c + D.by_rs(p_c);
    Fwd o;
    o.r = 2 * r_cn - r_b;
    o.t = 2 * s.t_cn - s.t_b;
    o.a = 2 * s.ua_cn - s.ua_b;
    o.b = 2 * s.ub_cn - s.ub_b;
    double gr = D.by_dt(2 * (r_cn - r_c)) + D.by_dx((s.un_cn + sound) * (r_f - r_b));
    double gt = D.by_dt(2 * (s.t_cn - s.t_c)) + D.by_dx(s.un_cn * (s.t_f - s.t_b));
    double ga = D.by_dt(2 * (s.ua_cn - s.ua_c)) + D.by_dx(s.un_cn * (s.ua_f - s.ua_b));
    double gb = D.by_dt(2 * (s.ub_cn - s.ub_c)) + D.by_dx(s.un_cn * (s.ub_f - s.ub_b));
    clip(o.r, max3d(r_f, r_c, r_b) + dt * gr, min3d(r_f, r_c, r_b) + dt * gr);
    clip(o.t, max3d(s.t_f, s.t_c, s.t_b) + dt * gt, min3d(s.t_f, s.t_c, s.t_b) + dt * gt);
    clip(o.a, max3d(s.ua_f, s.ua_c, s.ua_b) + dt * ga, min3d(s.ua_f, s.ua_c, s.ua_b) + dt * ga);
    clip(o.b, max3d(s.ub_f, s.ub_c, s.ub_b) + dt * gb, min3d(s.ub_f, s.ub_c, s.ub_b) + dt * gb);
    return o;
}

// Backward extrapolations of one cell.  GqUsesR reproduces the X1/X3 sweeps,
// whose gq takes its time derivative from r (source/main.cpp:1804, 2321); the
// X2 sweep uses q (source/main.cpp:2083).
template <bool GqUsesR, bool Rcp = false>
__device__ __forceinline__ Bwd cell_backward(const CellSide& s, double dt, double dx, const Consts& K,
                                             double rdt = 0.0, double rdx = 0.0) {
    const Divide<Rcp> D{dt, rdt, dx, rdx, K};
    const double sound = K.sound, ro0_g = K.ro0;
    double p_cn = sound * sound * (s.ro_cn - ro0_g);
    double p_c = sound * sound * (s.ro_c - ro0_g);
    double q_f = s.un_f - D.by_rs(s.p_f);
    double q_b = s.un_b - D.by_rs(s.p_b);
    double q_cn = s.un_cn - D.by_rs(p_cn);
    double q_c = s.un_c - D.by_rs(p_c);
    Bwd o;
    o.q = 2 * q_cn - q_f;
    o.t = 2 * s.t_cn - s.t_f;
    o.a = 2 * s.ua_cn - s.ua_f;
    o.b = 2 * s.ub_cn - s.ub_f;
    double dq;
    if (GqUsesR) {
        double r_cn = s.un_cn + D.by_rs(p_cn);
        double r_c = s.un_c + D.by_rs(p_c);
        dq = D.by_dt(2 * (r_cn - r_c));
    } else {
        dq = D.by_dt(2 * (q_cn - q_c));
    }
    double gq = dq + D.by_dx((s.un_cn - sound) * (q_f - q_b));
    double gt = D.by_dt(2 * (s.t_cn - s.t_c)) + D.by_dx(s.un_cn * (s.t_f - s.t_b));
    double ga = D.by_dt(2 * (s.ua_cn - s.ua_c)) + D.by_dx(s.un_cn * (s.ua_f - s.ua_b));
    double gb = D.by_dt(2 * (s.ub_cn - s.ub_c)) + D.by_dx(s.un_cn * (s.ub_f - s.ub_b));
    clip(o.q, max3d(q_f, q_c, q_b) + dt * gq, min3d(q_f, q_c, q_b) + dt * gq);
    clip(o.t, max3d(s.t_f, s.t_c, s.t_b) + dt * gt, min3d(s.t_f, s.t_c, s.t_b) + dt * gt);
    clip(o.a, max3d(s.ua_f, s.ua_c, s.ua_b) + dt * ga, min3d(s.ua_f, s.ua_c, s.ua_b) + dt * ga);
    clip(o.b, max3d(s.ub_f, s.ub_c, s.ub_b) + dt * gb, min3d(s.ub_f, s.ub_c, s.ub_b) + dt * gb);
    return o;
}

// New face state from the forward extrapolation of the cell behind it and
// the backward extrapolation of the cell ahead (fluid-fluid face,
// source/main.cpp:1924-1961).  ucb/ucf: normal velocity at n+1/2 in the two
// cells; the divergent case needs the cells' n+1/2 values and the old face.
struct FaceOut { double un, p, ro, t, a, b; };

template <bool Rcp = false>
__device__ __forceinline__ FaceOut face_combine(const Fwd& fb, const Bwd& bf, double ucb, double ucf,
                                                double tn_b, double tn_f, double ta_b, double ta_f,
                                                double tb_b, double tb_f, double t_old, double a_old,
                                                double b_old, const Consts& K) {
    const double sound = K.sound, ro0_g = K.ro0;
    FaceOut o;
    double rn = fb.r, qn = bf.q;
    o.un = (rn + qn) / 2;
    o.p = (rn - qn) * sound * ro0_g / 2;
    o.ro = ro0_g + (Rcp ? div_rn(o.p, sound * sound, K.rc2) : o.p / (sound * sound));
    if (ucf >= 0 && ucb >= 0) {
        o.t = fb.t; o.a = fb.a; o.b = fb.b;
    } else if (ucf <= 0 && ucb <= 0) {
        o.t = bf.t; o.a = bf.a; o.b = bf.b;
    } else if (ucb >= 0 && ucf <= 0) {
        if (ucb > -ucf) { o.t = fb.t; o.a = fb.a; o.b = fb.b; }
        else { o.t = bf.t; o.a = bf.a; o.b = bf.b; }
    } else {
        o.t = tn_f + tn_b - t_old;
        o.a = ta_f + ta_b - a_old;
        o.b = tb_f + tb_b - b_old;
    }
    return o;
}

}  // namespace arkd

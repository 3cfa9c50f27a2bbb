// Homogeneous turbulence on the periodic box (no counterpart in the
// reference, which solves a channel; these build and measure the forced
// turbulence runs the learned closure is trained on):
//   * a seeded solenoidal initial velocity with a prescribed spectrum,
//   * stochastic random-phase forcing on the lowest wavenumber shells,
//   * the shell-summed kinetic-energy spectrum and the dissipation rate,
//   * a w-cell box filter with the subgrid stresses and the filtered
//     velocity gradients (closure training pairs).
// FFTs are hipFFT double-precision real-to-complex transforms of the
// interior N3 x N2 x N1 cells (x fastest).  Random numbers are Philox4x32-10,
// a pure function of (seed, counter): results do not depend on the launch.
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ark/ark.h"
#include "ark_internal.hpp"
#include "ark_kernels.hpp"
#include "ark_rng.hpp"

using namespace arkd;

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

struct Spec {
    int N1, N2, N3;     // interior cells
    int H1;             // N1/2 + 1: complex extent of the fastest axis
    double L[3];        // box lengths
};

__device__ __forceinline__ int wave_index(int q, int N) { return q <= N / 2 ? q : q - N; }

// Gaussian white noise, three components, contiguous N3*N2*N1 each.
__global__ void k_noise(Spec s, double* out, unsigned long long seed) {
    const long long n = (long long)s.N1 * s.N2 * s.N3;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        double g[4];
        gauss4(seed, (uint32_t)q, (uint32_t)(q >> 32), 0x7a3c51u, 0u, g);
        out[q] = g[0];
        out[n + q] = g[1];
        out[2 * n + q] = g[2];
    }
}

// Spectral shaping of the three transformed components in place: amplitude
// |k| exp(-(|k|/k0)^2) per mode, so that the shell energy of white noise
// (proportional to the 4 pi k^2 modes of a shell) becomes
// E(k) ~ k^4 exp(-2 (k/k0)^2); projection on the plane normal to k (a
// solenoidal field); the mean and the Nyquist modes are zeroed.
__global__ void k_shape(Spec s, hipfftDoubleComplex* u, double k0) {
    const long long nc = (long long)s.H1 * s.N2 * s.N3;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nc; q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q % s.H1), j = (int)((q / s.H1) % s.N2), k = (int)(q / ((long long)s.H1 * s.N2));
        const int m1 = i, m2 = wave_index(j, s.N2), m3 = wave_index(k, s.N3);
        const double kv[3] = {kTwoPi / s.L[0] * m1, kTwoPi / s.L[1] * m2, kTwoPi / s.L[2] * m3};
        const double kk = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        const bool nyq = (2 * m1 == s.N1) || (2 * j == s.N2) || (2 * k == s.N3);
        hipfftDoubleComplex* c[3] = {u + q, u + nc + q, u + 2 * nc + q};
        if (kk == 0.0 || nyq) {
            for (int d = 0; d < 3; ++d) c[d]->x = c[d]->y = 0.0;
            continue;
        }
        const double km = sqrt(kk);
        const double amp = km * exp(-(km / k0) * (km / k0));
        double re[3], im[3];
        for (int d = 0; d < 3; ++d) {
            re[d] = c[d]->x;
            im[d] = c[d]->y;
        }
        const double pr = (kv[0] * re[0] + kv[1] * re[1] + kv[2] * re[2]) / kk;
        const double pi = (kv[0] * im[0] + kv[1] * im[1] + kv[2] * im[2]) / kk;
        for (int d = 0; d < 3; ++d) {
            c[d]->x = amp * (re[d] - pr * kv[d]);
            c[d]->y = amp * (im[d] - pi * kv[d]);
        }
    }
}

// Sums in a fixed order, so every result is the same bit for bit from run to
// run: a grid-stride loop over a fixed grid of kSumBlocks x 256 threads, a
// tree over each block (block_sum), one partial per block, and the partials
// summed by one block (k_sum_parts).
constexpr int kSumBlocks = 1024;

__device__ double block_sum(double v) {
    __shared__ double red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    return red[0];
}

__global__ void k_sum_parts(const double* part, int n, double* out) {
    double acc = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) acc += part[q];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) *out = t;
}

// sum over the full spectrum of |u_hat|^2 (three components), counting the
// modes the half spectrum stands for twice: N^3 * sum_x |u|^2 (Parseval).
// One partial per block into part[blockIdx.x].
__global__ void k_spec_energy(Spec s, const hipfftDoubleComplex* u, double* part) {
    const long long nc = (long long)s.H1 * s.N2 * s.N3;
    double acc = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nc; q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q % s.H1);
        const double w = (i == 0 || 2 * i == s.N1) ? 1.0 : 2.0;
        for (int d = 0; d < 3; ++d) {
            const hipfftDoubleComplex v = u[d * nc + q];
            acc += w * (v.x * v.x + v.y * v.y);
        }
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// dst = src * exp(-i k_d dx_d / 2): the field sampled half a cell back along
// axis d, i.e. on the faces normal to d (face q lies dx_d/2 below cell q).
__global__ void k_shift(Spec s, const hipfftDoubleComplex* src, hipfftDoubleComplex* dst, int d) {
    const long long nc = (long long)s.H1 * s.N2 * s.N3;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nc; q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q % s.H1), j = (int)((q / s.H1) % s.N2), k = (int)(q / ((long long)s.H1 * s.N2));
        const int m = d == 0 ? i : d == 1 ? wave_index(j, s.N2) : wave_index(k, s.N3);
        const int Nd = d == 0 ? s.N1 : d == 1 ? s.N2 : s.N3;
        const double ph = -0.5 * kTwoPi * m / Nd;
        const double cr = cos(ph), ci = sin(ph);
        const hipfftDoubleComplex v = src[q];
        dst[q].x = v.x * cr - v.y * ci;
        dst[q].y = v.x * ci + v.y * cr;
    }
}

// Field (cell or face extents, periodic ghosts) from an interior array
// src[(k-1)*N1*N2 + (j-1)*N1 + (i-1)] times scale plus offset; src may be
// null (constant).  Every index wraps to 1..N: face N+1 is face 1.
__global__ void k_scatter(Geo g, double* dst, const double* src, double scale, double offset, int ext) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= g.N1 + ext || j >= g.N2 + ext || k >= g.N3 + ext) return;
    const int iw = ((i - 1) % g.N1 + g.N1) % g.N1, jw = ((j - 1) % g.N2 + g.N2) % g.N2,
              kw = ((k - 1) % g.N3 + g.N3) % g.N3;
    const double v = src ? src[((size_t)kw * g.N2 + jw) * g.N1 + iw] : 0.0;
    dst[at(g, i, j, k)] = v * scale + offset;
}

// Interior of a cell field into a contiguous array.
__global__ void k_extract(Geo g, const double* f, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= g.N1) return;
    out[((size_t)k * g.N2 + j) * g.N1 + i] = f[at(g, i + 1, j + 1, k + 1)];
}

// Shell sums of 1/2 |u_hat|^2 / (N^3)^2 over integer-index shells
// b = round(|m|): E[b] summed over b equals 1/2 <u.u> (mean over cells).
// One block per shell, so no two blocks add into one bin: the block's threads
// take the spectrum's rows (j, k) in turn and, on each, only the i where the
// shell can lie (b - 1/2 <= |m| < b + 1/2, widened by one either side and
// each i tested with the exact bin formula), then the block's tree.
__global__ void k_shell_sum(Spec s, const hipfftDoubleComplex* u, double* E) {
    const int b = blockIdx.x;
    const long long nc = (long long)s.H1 * s.N2 * s.N3;
    const double n3 = (double)s.N1 * s.N2 * s.N3;
    const double norm = 0.5 / (n3 * n3);
    const double lo2 = (b - 0.5) * (b - 0.5), hi2 = (b + 0.5) * (b + 0.5);
    double acc = 0.0;
    for (int r = threadIdx.x; r < s.N2 * s.N3; r += blockDim.x) {
        const int j = r % s.N2, k = r / s.N2;
        const int m2 = wave_index(j, s.N2), m3 = wave_index(k, s.N3);
        const int q2 = m2 * m2 + m3 * m3;
        if ((double)q2 >= hi2) continue;   // |m| >= b + 1/2 for every i
        const int ilo = max(0, (int)sqrt(fmax(0.0, lo2 - q2)) - 1);
        const int ihi = min(s.H1 - 1, (int)sqrt(fmax(0.0, hi2 - q2)) + 1);
        const long long row = ((long long)k * s.N2 + j) * s.H1;
        for (int i = ilo; i <= ihi; ++i) {
            if ((int)(sqrt((double)(i * i + q2)) + 0.5) != b) continue;
            const double w = (i == 0 || 2 * i == s.N1) ? 1.0 : 2.0;
            double e = 0.0;
            for (int d = 0; d < 3; ++d) {
                const hipfftDoubleComplex v = u[d * nc + row + i];
                e += v.x * v.x + v.y * v.y;
            }
            acc += w * norm * e;
        }
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) E[b] = t;
}

// nu * <sum_ij (d_j u_i)^2> with central differences of the level-n cells:
// the sum of the squares, one partial per block into part[blockIdx.x].
__global__ void k_dissipation(Geo g, const double* u1, const double* u2, const double* u3, double inv2dx0,
                              double inv2dx1, double inv2dx2, double* part) {
    const long long n = (long long)g.N1 * g.N2 * g.N3;
    const double* u[3] = {u1, u2, u3};
    double acc = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.N1) + 1, j = (int)((q / g.N1) % g.N2) + 1, k = (int)(q / ((long long)g.N1 * g.N2)) + 1;
        const int ip = i == g.N1 ? 1 : i + 1, im = i == 1 ? g.N1 : i - 1;
        const int jp = j == g.N2 ? 1 : j + 1, jm = j == 1 ? g.N2 : j - 1;
        const int kp = k == g.N3 ? 1 : k + 1, km = k == 1 ? g.N3 : k - 1;
        for (int m = 0; m < 3; ++m) {
            const double dx = (u[m][at(g, ip, j, k)] - u[m][at(g, im, j, k)]) * inv2dx0;
            const double dy = (u[m][at(g, i, jp, k)] - u[m][at(g, i, jm, k)]) * inv2dx1;
            const double dz = (u[m][at(g, i, j, kp)] - u[m][at(g, i, j, km)]) * inv2dx2;
            acc += dx * dx + dy * dy + dz * dz;
        }
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// Forcing.  The forced wavevectors are the integer m with 1 <= |m|^2 <= 4
// whose first non-zero component is positive (16 of the 32; the other half
// are their conjugates).  Each step, mode n gets a random complex amplitude
// vector a_n normal to its k, of norm A = sqrt(eps / (dt * 16)), drawn from
// (seed, step, n); the acceleration field F(x) = sum_n 2 Re(a_n e^{ik.x}) then
// has <|F|^2> = 32 A^2, and a white-in-time force injects on average
// <|F|^2> dt / 2 = eps per unit mass (the force is independent of the
// velocity, so <u.F> averages to zero).
struct ForceMode { int m[3]; };

__constant__ ForceMode c_modes[16];

// 16 modes x (re, im) x 3 components.
// With tau > 0 the amplitudes are redrawn once per time interval of length
// tau (counter = the interval of the step's mid time) instead of every step,
// with norm sqrt(eps/(16 tau)): a force constant over tau injects
// <|F|^2> tau / 2 = eps on average, and runs of any resolution see the same
// force history.
__device__ void force_mode(const Scalars* S, unsigned long long seed, double eps, double tau, double L0, double L1,
                           double L2, float* table, int n) {
    const double dt = tau > 0.0 ? tau : S->dt;
    const uint32_t step = tau > 0.0 ? (uint32_t)floor((S->time - 0.5 * S->dt) / tau) : (uint32_t)S->nstep;
    double g[8];
    gauss4(seed, (uint32_t)n, step, 0xf0c3u, 0u, g);
    gauss4(seed, (uint32_t)n, step, 0xf0c3u, 1u, g + 4);
    const double kv[3] = {kTwoPi / L0 * c_modes[n].m[0], kTwoPi / L1 * c_modes[n].m[1], kTwoPi / L2 * c_modes[n].m[2]};
    const double kk = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    double re[3] = {g[0], g[1], g[2]}, im[3] = {g[3], g[4], g[5]};
    const double pr = (kv[0] * re[0] + kv[1] * re[1] + kv[2] * re[2]) / kk;
    const double pi = (kv[0] * im[0] + kv[1] * im[1] + kv[2] * im[2]) / kk;
    double nrm = 0.0;
    for (int d = 0; d < 3; ++d) {
        re[d] -= pr * kv[d];
        im[d] -= pi * kv[d];
        nrm += re[d] * re[d] + im[d] * im[d];
    }
    const double A = (dt > 0.0 && nrm > 0.0) ? sqrt(eps / (dt * 16.0)) / sqrt(nrm) : 0.0;
    for (int d = 0; d < 3; ++d) {
        table[n * 6 + 2 * d] = (float)(A * re[d]);
        table[n * 6 + 2 * d + 1] = (float)(A * im[d]);
    }
}

__global__ void k_force_modes(const Scalars* S, unsigned long long seed, double eps, double tau, double L0, double L1,
                              double L2, float* table) {
    if (threadIdx.x < 16) force_mode(S, seed, eps, tau, L0, L1, L2, table, threadIdx.x);
}

// Forcing of many boxes at once (batched steps): one record per forced box.
struct ForceBox {
    const Scalars* S;
    unsigned long long seed;
    double eps, tau, L0, L1, L2;
    float* table;
};

// Block b: the 16 modes of box b.
__global__ void k_force_modes_batch(const ForceBox* __restrict__ fb) {
    const ForceBox& f = fb[blockIdx.x];
    if (threadIdx.x < 16) force_mode(f.S, f.seed, f.eps, f.tau, f.L0, f.L1, f.L2, f.table, threadIdx.x);
}

// F at the cell centres x = (i - 1/2) dx, three FP32 components at force +
// m*stride in the field index map.  trig holds, per axis d and cell index
// 1..N_d, {cos, sin} of (2 pi / L_d) m x for m = 1, 2 (axes one after the
// other); each mode's phase is the product of three of them.
__device__ __forceinline__ void force_cell(const Geo& g, const float* tab, const float4* trig, float* force,
                                           long long stride, int i, int j, int k) {
    const float4 t[3] = {trig[i - 1], trig[g.N1 + j - 1], trig[g.N1 + g.N2 + k - 1]};
    float cr[3][3], ci[3][3];
    for (int d = 0; d < 3; ++d) {
        cr[d][0] = 1.f;
        ci[d][0] = 0.f;
        cr[d][1] = t[d].x;
        ci[d][1] = t[d].y;
        cr[d][2] = t[d].z;
        ci[d][2] = t[d].w;
    }
    // the 16 forced wavevectors (forced_mode), compile-time constants once unrolled
    float F[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        float er = 1.f, ei = 0.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int m = forced_mode(n, d);
            const float r = cr[d][m < 0 ? -m : m], s = m < 0 ? -ci[d][-m] : ci[d][m];
            const float nr = er * r - ei * s, ni = er * s + ei * r;
            er = nr;
            ei = ni;
        }
        for (int d = 0; d < 3; ++d) F[d] += 2.f * (tab[n * 6 + 2 * d] * er - tab[n * 6 + 2 * d + 1] * ei);
    }
    const int c = at(g, i, j, k);
    force[c] = F[0];
    force[stride + c] = F[1];
    force[2 * stride + c] = F[2];
}

__global__ void k_force_field(Geo g, const float* table, const float4* trig, float* force, long long stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    __shared__ float tab[96];
    for (int q = threadIdx.x; q < 96; q += blockDim.x) tab[q] = table[q];
    __syncthreads();
    if (i > g.N1) return;
    force_cell(g, tab, trig, force, stride, i, j, k);
}

// ---------------------------------------------------------------------------
// Box filter of width w cells (w divides N): per coarse cell the mean of u_i
// and of u_i u_j over its w^3 fine cells (FP64 sums), giving the filtered
// velocity and the subgrid stress tau_ij = <u_i u_j> - <u_i><u_j>
// (ij = 11, 22, 33, 12, 13, 23), coarse arrays contiguous (x fastest).
__global__ void k_box_filter(Geo g, const double* u1, const double* u2, const double* u3, int w, double* ubar,
                             double* tau) {
    const int Nc1 = g.N1 / w, Nc2 = g.N2 / w, Nc3 = g.N3 / w;
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y, K = blockIdx.z;
    if (I >= Nc1) return;
    double s[3] = {0, 0, 0}, p[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < w; ++c)
        for (int b = 0; b < w; ++b)
            for (int a = 0; a < w; ++a) {
                const int q = at(g, I * w + a + 1, J * w + b + 1, K * w + c + 1);
                const double v[3] = {u1[q], u2[q], u3[q]};
                for (int m = 0; m < 3; ++m) s[m] += v[m];
                p[0] += v[0] * v[0];
                p[1] += v[1] * v[1];
                p[2] += v[2] * v[2];
                p[3] += v[0] * v[1];
                p[4] += v[0] * v[2];
                p[5] += v[1] * v[2];
            }
    const double inv = 1.0 / ((double)w * w * w);
    const long long nc = (long long)Nc1 * Nc2 * Nc3, o = ((long long)K * Nc2 + J) * Nc1 + I;
    double m[3];
    for (int d = 0; d < 3; ++d) {
        m[d] = s[d] * inv;
        ubar[d * nc + o] = m[d];
    }
    const int pi[6] = {0, 1, 2, 0, 0, 1}, pj[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) tau[q * nc + o] = p[q] * inv - m[pi[q]] * m[pj[q]];
}

// Block means of one cell field over w^3 cells (the box filter of a scalar).
__global__ void k_box_mean(Geo g, const double* f, int w, double* out) {
    const int Nc1 = g.N1 / w, Nc2 = g.N2 / w;
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y, K = blockIdx.z;
    if (I >= Nc1) return;
    double s = 0.0;
    for (int c = 0; c < w; ++c)
        for (int b = 0; b < w; ++b)
            for (int a = 0; a < w; ++a) s += f[at(g, I * w + a + 1, J * w + b + 1, K * w + c + 1)];
    out[((long long)K * Nc2 + J) * Nc1 + I] = s * (1.0 / ((double)w * w * w));
}

// Gradient d_j ubar_i of the filtered velocity on the coarse grid (central
// differences, periodic), component 3*i + j.
__global__ void k_coarse_gradient(int Nc1, int Nc2, int Nc3, const double* ubar, double h0, double h1, double h2,
                                  double* grad) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y, K = blockIdx.z;
    if (I >= Nc1) return;
    const long long nc = (long long)Nc1 * Nc2 * Nc3;
    auto id = [&](int a, int b, int c) {
        a = (a + Nc1) % Nc1;
        b = (b + Nc2) % Nc2;
        c = (c + Nc3) % Nc3;
        return ((long long)c * Nc2 + b) * Nc1 + a;
    };
    const long long o = id(I, J, K);
    for (int m = 0; m < 3; ++m) {
        const double* u = ubar + m * nc;
        grad[(3 * m + 0) * nc + o] = (u[id(I + 1, J, K)] - u[id(I - 1, J, K)]) / (2.0 * h0);
        grad[(3 * m + 1) * nc + o] = (u[id(I, J + 1, K)] - u[id(I, J - 1, K)]) / (2.0 * h1);
        grad[(3 * m + 2) * nc + o] = (u[id(I, J, K + 1)] - u[id(I, J, K - 1)]) / (2.0 * h2);
    }
}

// ---------------------------------------------------------------------------
// Learned closure on the solver's grid: from the network's output, the
// subgrid stresses tau_ij of every cell (the network and its inputs, the
// central-difference velocity gradient, run in ark_mlp.hip), the
// acceleration -d_j tau_ij added to (or written into) the body force.
__global__ void k_les_force(Geo g, const float* Y, float inv0, float inv1, float inv2, float* force, long long stride,
                            int add) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > g.N1) return;
    const int ip = i == g.N1 ? 1 : i + 1, im = i == 1 ? g.N1 : i - 1;
    const int jp = j == g.N2 ? 1 : j + 1, jm = j == 1 ? g.N2 : j - 1;
    const int kp = k == g.N3 ? 1 : k + 1, km = k == 1 ? g.N3 : k - 1;
    auto tau = [&](int a, int b, int c, int q) { return Y[(((size_t)(c - 1) * g.N2 + (b - 1)) * g.N1 + (a - 1)) * 6 + q]; };
    // tau components 11 22 33 12 13 23
    const int cmp[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
    const int c = at(g, i, j, k);
    for (int m = 0; m < 3; ++m) {
        const float div = (tau(ip, j, k, cmp[m][0]) - tau(im, j, k, cmp[m][0])) * inv0 +
                          (tau(i, jp, k, cmp[m][1]) - tau(i, jm, k, cmp[m][1])) * inv1 +
                          (tau(i, j, kp, cmp[m][2]) - tau(i, j, km, cmp[m][2])) * inv2;
        float* f = force + m * stride + c;
        *f = (add ? *f : 0.f) - div;
    }
}

// The closure's acceleration -d_j tau_ij into three contiguous N1 x N2 x N3
// arrays (i fastest), the layout of a 3-D real transform.
__global__ void k_les_force_dense(Geo g, const float* Y, float inv0, float inv1, float inv2, float* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > g.N1) return;
    const int ip = i == g.N1 ? 1 : i + 1, im = i == 1 ? g.N1 : i - 1;
    const int jp = j == g.N2 ? 1 : j + 1, jm = j == 1 ? g.N2 : j - 1;
    const int kp = k == g.N3 ? 1 : k + 1, km = k == 1 ? g.N3 : k - 1;
    auto tau = [&](int a, int b, int c, int q) { return Y[(((size_t)(c - 1) * g.N2 + (b - 1)) * g.N1 + (a - 1)) * 6 + q]; };
    const int cmp[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
    const size_t n = (size_t)g.N1 * g.N2 * g.N3, o = ((size_t)(k - 1) * g.N2 + (j - 1)) * g.N1 + (i - 1);
    for (int m = 0; m < 3; ++m) {
        const float div = (tau(ip, j, k, cmp[m][0]) - tau(im, j, k, cmp[m][0])) * inv0 +
                          (tau(i, jp, k, cmp[m][1]) - tau(i, jm, k, cmp[m][1])) * inv1 +
                          (tau(i, j, kp, cmp[m][2]) - tau(i, j, km, cmp[m][2])) * inv2;
        dst[m * n + o] = -div;
    }
}

// Keeps the Fourier modes of shells round(|m|) >= kcut (m the integer
// wavevector) and scales them by `scale` (the inverse transform's 1/n); the
// others are zeroed.  Three half spectra of H1 x N2 x N3, H1 fastest.
__global__ void k_keep_high_shells(float2* spec, int N1, int N2, int N3, int H1, int kcut, float scale) {
    const size_t nc = (size_t)H1 * N2 * N3;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < 3 * nc; q += (size_t)gridDim.x * blockDim.x) {
        const size_t r = q % nc;
        const int a = (int)(r % H1), b = (int)((r / H1) % N2), c = (int)(r / ((size_t)H1 * N2));
        const int m1 = a, m2 = wave_index(b, N2), m3 = wave_index(c, N3);
        const int shell = (int)lrint(sqrt((double)m1 * m1 + (double)m2 * m2 + (double)m3 * m3));
        const float w = shell >= kcut ? scale : 0.f;
        spec[q] = make_float2(spec[q].x * w, spec[q].y * w);
    }
    (void)N1;
}

// force (common index map) := force + src (add) or src, for the three
// components of a dense field.
__global__ void k_put_force(Geo g, const float* src, float* force, long long stride, int add) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > g.N1) return;
    const size_t n = (size_t)g.N1 * g.N2 * g.N3, o = ((size_t)(k - 1) * g.N2 + (j - 1)) * g.N1 + (i - 1);
    const int c = at(g, i, j, k);
    for (int m = 0; m < 3; ++m) {
        float* f = force + m * stride + c;
        *f = (add ? *f : 0.f) + src[m * n + o];
    }
}

// ---------------------------------------------------------------------------
struct Fft {
    hipfftHandle plan_f = 0, plan_b = 0;
    Spec s{};
    ~Fft() {
        if (plan_f) hipfftDestroy(plan_f);
        if (plan_b) hipfftDestroy(plan_b);
    }
};

ark_status make_fft(ark_ctx* c, Fft& f, bool backward) {
    f.s.N1 = c->geo.N1;
    f.s.N2 = c->geo.N2;
    f.s.N3 = c->geo.N3;
    f.s.H1 = c->geo.N1 / 2 + 1;
    f.s.L[0] = c->prm.dx1 * c->geo.N1;
    f.s.L[1] = c->prm.dx2 * c->geo.N2;
    f.s.L[2] = c->prm.dx3 * c->geo.N3;
    if (hipfftPlan3d(&f.plan_f, f.s.N3, f.s.N2, f.s.N1, HIPFFT_D2Z) != HIPFFT_SUCCESS)
        return c->fail(ARK_ERR_HIP, "hipfftPlan3d D2Z failed");
    hipfftSetStream(f.plan_f, c->stream);
    if (backward) {
        if (hipfftPlan3d(&f.plan_b, f.s.N3, f.s.N2, f.s.N1, HIPFFT_Z2D) != HIPFFT_SUCCESS)
            return c->fail(ARK_ERR_HIP, "hipfftPlan3d Z2D failed");
        hipfftSetStream(f.plan_b, c->stream);
    }
    return ARK_OK;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

dim3 interior_grid(const Geo& g, int ext) {
    return dim3((g.N1 + ext + 255) / 256, g.N2 + ext, g.N3 + ext);
}

// A velocity field given by its half spectrum U (three components, the
// unnormalised forward transforms of the cell values) times `scale`: every
// array zeroed, then the velocities on the cells (level n, both buffers, and
// Con) and, by the half-cell shift, on the three face families; t0.  The
// density is rho0 with zero pressure, or, given the half spectrum P of
// rho - rho0 (same scaling), rho on the cells and, by the same shift, on the
// faces with their pressures c^2 (rho - rho0).  W (one component) and R (one
// real field) are work.
ark_status install_velocity(ark_ctx* c, Fft& f, const hipfftDoubleComplex* U, hipfftDoubleComplex* W, double* R,
                            double scale, const hipfftDoubleComplex* P = nullptr) {
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    const Geo& g = c->geo;
    const long long nc = (long long)f.s.H1 * g.N2 * g.N3;
    ark_status s = ARK_OK;
    // every array zero, then velocities on cells and the three face families
    for (int q = 0; q < F_NUM && s == ARK_OK; ++q) {
        s = c->hip(hipMemsetAsync(c->cur[q] - 15, 0, c->elems * sizeof(double), c->stream), "zero");
        if (s == ARK_OK && c->alt[q]) s = c->hip(hipMemsetAsync(c->alt[q] - 15, 0, c->elems * sizeof(double), c->stream), "zero");
    }
    const int cellu[3] = {F_u1nCon, F_u2nCon, F_u3nCon}, conu[3] = {F_u1Con, F_u2Con, F_u3Con};
    for (int m = 0; m < 3 && s == ARK_OK; ++m) {
        s = c->hip(hipMemcpyAsync(W, U + m * nc, nc * sizeof(hipfftDoubleComplex), hipMemcpyDeviceToDevice, c->stream),
                   "copy spectrum");
        if (s == ARK_OK && hipfftExecZ2D(f.plan_b, W, R) != HIPFFT_SUCCESS) s = c->fail(ARK_ERR_HIP, "Z2D");
        for (double* dst : {c->cur[cellu[m]], c->alt[cellu[m]], c->cur[conu[m]]}) {
            if (s != ARK_OK || !dst) continue;
            hipLaunchKernelGGL(k_scatter, interior_grid(g, 2), dim3(256), 0, c->stream, g, dst, R, scale, 0.0, 2);
            s = c->hip(hipGetLastError(), "scatter");
        }
        for (int d = 0; d < 3 && s == ARK_OK; ++d) {
            hipLaunchKernelGGL(k_shift, dim3(4096), dim3(256), 0, c->stream, f.s, U + m * nc, W, d);
            s = c->hip(hipGetLastError(), "shift");
            if (s == ARK_OK && hipfftExecZ2D(f.plan_b, W, R) != HIPFFT_SUCCESS) s = c->fail(ARK_ERR_HIP, "Z2D");
            const int fu = face_u(d, m);
            for (double* dst : {c->cur[fu], c->alt[fu]}) {
                if (s != ARK_OK || !dst) continue;
                hipLaunchKernelGGL(k_scatter, interior_grid(g, 3), dim3(256), 0, c->stream, g, dst, R, scale, 0.0, 3);
                s = c->hip(hipGetLastError(), "scatter");
            }
        }
    }
    // t0; density rho0 and zero pressure, or the given density perturbation
    const double ro0 = c->prm.ro0, c2 = c->prm.sound * c->prm.sound;
    if (P && s == ARK_OK) {
        s = c->hip(hipMemcpyAsync(W, P, nc * sizeof(hipfftDoubleComplex), hipMemcpyDeviceToDevice, c->stream),
                   "copy spectrum");
        if (s == ARK_OK && hipfftExecZ2D(f.plan_b, W, R) != HIPFFT_SUCCESS) s = c->fail(ARK_ERR_HIP, "Z2D");
        for (int q : {F_ronCon, F_roCon})
            for (double* dst : {c->cur[q], c->alt[q]}) {
                if (s != ARK_OK || !dst) continue;
                hipLaunchKernelGGL(k_scatter, interior_grid(g, 2), dim3(256), 0, c->stream, g, dst, R, scale, ro0, 2);
                s = c->hip(hipGetLastError(), "scatter");
            }
        for (int d = 0; d < 3 && s == ARK_OK; ++d) {
            hipLaunchKernelGGL(k_shift, dim3(4096), dim3(256), 0, c->stream, f.s, P, W, d);
            s = c->hip(hipGetLastError(), "shift");
            if (s == ARK_OK && hipfftExecZ2D(f.plan_b, W, R) != HIPFFT_SUCCESS) s = c->fail(ARK_ERR_HIP, "Z2D");
            for (int b = 0; b < 2 && s == ARK_OK; ++b) {
                double* rf = b ? c->alt[face_ro(d)] : c->cur[face_ro(d)];
                double* pf = b ? c->alt[face_p(d)] : c->cur[face_p(d)];
                if (rf) hipLaunchKernelGGL(k_scatter, interior_grid(g, 3), dim3(256), 0, c->stream, g, rf, R, scale,
                                           ro0, 3);
                if (pf) hipLaunchKernelGGL(k_scatter, interior_grid(g, 3), dim3(256), 0, c->stream, g, pf, R,
                                           scale * c2, 0.0, 3);
                s = c->hip(hipGetLastError(), "scatter");
            }
        }
    }
    const int cellc[4] = {F_ronCon, F_roCon, F_tnCon, F_tCon};
    const double cellv[4] = {ro0, ro0, c->prm.t0, c->prm.t0};
    for (int q = P ? 2 : 0; q < 4 && s == ARK_OK; ++q)
        for (double* dst : {c->cur[cellc[q]], c->alt[cellc[q]]}) {
            if (s != ARK_OK || !dst) continue;
            hipLaunchKernelGGL(k_scatter, interior_grid(g, 2), dim3(256), 0, c->stream, g, dst, (const double*)nullptr,
                               0.0, cellv[q], 2);
            s = c->hip(hipGetLastError(), "fill");
        }
    for (int d = 0; d < 3 && s == ARK_OK; ++d) {
        const int fq[2] = {face_ro(d), face_t(d)};
        const double fv[2] = {ro0, c->prm.t0};
        for (int q = P ? 1 : 0; q < 2; ++q)
            for (double* dst : {c->cur[fq[q]], c->alt[fq[q]]}) {
                if (s != ARK_OK || !dst) continue;
                hipLaunchKernelGGL(k_scatter, interior_grid(g, 3), dim3(256), 0, c->stream, g, dst,
                                   (const double*)nullptr, 0.0, fv[q], 3);
                s = c->hip(hipGetLastError(), "fill");
            }
    }
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s == ARK_OK) s = ark_finish_init(c);
    return s;
}

}  // namespace

// Mixed precision: this step's forcing (nothing when off).  The 16-mode
// table always; the acceleration field only when a learned closure will add
// its own acceleration to it (ark_closure_step).  Without a closure the fused
// step evaluates the forcing per cell from the table itself (mixed::Args::
// fmodes), which keeps the field's 12 bytes per cell out of HBM.
ark_status ark_turb_force_step(ark_ctx* c, hipStream_t st) {
    if (!c->force || c->force_eps <= 0.0) return ARK_OK;
    const double L0 = c->prm.dx1 * c->geo.N1, L1 = c->prm.dx2 * c->geo.N2, L2 = c->prm.dx3 * c->geo.N3;
    hipLaunchKernelGGL(k_force_modes, dim3(1), dim3(64), 0, st, c->S, (unsigned long long)c->force_seed,
                       c->force_eps, c->force_tau, L0, L1, L2, c->force_table);
    ark_status s = c->hip(hipGetLastError(), "force modes");
    if (s != ARK_OK || !c->closure) return s;
    hipLaunchKernelGGL(k_force_field, interior_grid(c->geo, 0), dim3(256), 0, st, c->geo, c->force_table,
                       (const float4*)c->force_trig, c->force, (long long)c->force_stride);
    return c->hip(hipGetLastError(), "force field");
}

// Forcing of a batch: one record per forced box, in device memory, built
// once per set of boxes.  Batched boxes carry no closure, so the batch kernel
// evaluates each box's forcing from its mode table.
struct ark_force_batch {
    ForceBox* fb = nullptr;
    int n = 0;
};

void ark_force_batch_free(ark_force_batch* f) {
    if (!f) return;
    if (f->fb) (void)hipFree(f->fb);
    delete f;
}

ark_status ark_force_batch_build(ark_ctx* const* boxes, int nbox, ark_force_batch** out) {
    ark_force_batch_free(*out);
    *out = nullptr;
    std::vector<ForceBox> h;
    for (int b = 0; b < nbox; ++b) {
        const ark_ctx* c = boxes[b];
        if (!c->force || c->force_eps <= 0.0) continue;
        ForceBox f;
        f.S = c->S;
        f.seed = c->force_seed;
        f.eps = c->force_eps;
        f.tau = c->force_tau;
        f.L0 = c->prm.dx1 * c->geo.N1;
        f.L1 = c->prm.dx2 * c->geo.N2;
        f.L2 = c->prm.dx3 * c->geo.N3;
        f.table = c->force_table;
        h.push_back(f);
    }
    ark_force_batch* f = new ark_force_batch();
    f->n = (int)h.size();
    if (f->n > 0) {
        if (hipMalloc(&f->fb, h.size() * sizeof(ForceBox)) != hipSuccess ||
            hipMemcpy(f->fb, h.data(), h.size() * sizeof(ForceBox), hipMemcpyHostToDevice) != hipSuccess) {
            ark_force_batch_free(f);
            return ARK_ERR_HIP;
        }
    }
    *out = f;
    return ARK_OK;
}

// This step's mode tables of every forced box: one launch, whatever the count.
ark_status ark_force_batch_enqueue(const ark_force_batch* f, hipStream_t st) {
    if (!f || f->n == 0) return ARK_OK;
    hipLaunchKernelGGL(k_force_modes_batch, dim3(f->n), dim3(64), 0, st, f->fb);
    return hipGetLastError() == hipSuccess ? ARK_OK : ARK_ERR_HIP;
}

// The force buffer (acceleration, 3 x FP32 fields, and the 16-mode table).
static ark_status ensure_force(ark_ctx* c) {
    if (c->force) return ARK_OK;
    float* p = nullptr;
    ark_status s = c->hip(hipMalloc(&p, (3 * c->elems + 96) * sizeof(float)), "hipMalloc force");
    if (s != ARK_OK) return s;
    c->allocs.push_back(p);
    c->bytes += (3 * c->elems + 96) * sizeof(float);
    s = c->hip(hipMemsetAsync(p, 0, (3 * c->elems + 96) * sizeof(float), c->stream), "memset force");
    c->force_table = p;
    c->force = p + 96 + 15;   // element (0,0,0) of component 0
    c->force_stride = (long long)c->elems;
    // per-axis phase tables of the forced wavenumbers (C library cos/sin in FP64)
    const int N[3] = {c->geo.N1, c->geo.N2, c->geo.N3};
    const double dx[3] = {c->prm.dx1, c->prm.dx2, c->prm.dx3};
    std::vector<float> h;
    for (int d = 0; d < 3; ++d)
        for (int i = 1; i <= N[d]; ++i) {
            const double ph = kTwoPi / (dx[d] * N[d]) * (i - 0.5) * dx[d];
            h.push_back((float)cos(ph));
            h.push_back((float)sin(ph));
            h.push_back((float)cos(2.0 * ph));
            h.push_back((float)sin(2.0 * ph));
        }
    float* tr = nullptr;
    if (s == ARK_OK) s = c->hip(hipMalloc(&tr, h.size() * sizeof(float)), "hipMalloc force tables");
    if (s == ARK_OK) {
        c->allocs.push_back(tr);
        c->bytes += h.size() * sizeof(float);
        c->force_trig = tr;
        s = c->hip(hipMemcpy(tr, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice), "force tables");
    }
    return s;
}

// Scale-selective application: FP32 transforms of the closure's
// acceleration, in the context's stream.
struct ark_closure_hp {
    hipfftHandle plan_f = 0, plan_b = 0;
    float* real = nullptr;     // 3 x N1 x N2 x N3
    float2* spec = nullptr;    // 3 x H1 x N2 x N3
    int kcut = 0;
};

void ark_closure_hp_free(ark_ctx* c) {
    ark_closure_hp* h = c->closure_hp;
    if (!h) return;
    if (h->plan_f) (void)hipfftDestroy(h->plan_f);
    if (h->plan_b) (void)hipfftDestroy(h->plan_b);
    if (h->real) (void)hipFree(h->real);
    if (h->spec) (void)hipFree(h->spec);
    delete h;
    c->closure_hp = nullptr;
}

// The closure's contribution to this step's body force (after the forcing).
ark_status ark_closure_step(ark_ctx* c, hipStream_t st) {
    if (!c->closure) return ARK_OK;
    const Geo& g = c->geo;
    const double inv[3] = {0.5 / c->prm.dx1, 0.5 / c->prm.dx2, 0.5 / c->prm.dx3};
    // velocities per field, or (while stepping) doubles 1-3 of each cell's
    // 32-byte record at group + 8 + 32 idx
    const double* u[3] = {c->cur[F_u1nCon], c->cur[F_u2nCon], c->cur[F_u3nCon]};
    int stride = 1;
    if (c->cells_rec) {
        const double* rec = (const double*)(c->gbase[0][c->gpar[0]] + 8);
        for (int m = 0; m < 3; ++m) u[m] = rec + 1 + m;
        stride = 4;
    }
    ark_status s = ark_mlp_closure_device(c->closure, g, u[0], u[1], u[2], stride, inv, (float)c->closure_delta,
                                          c->les_Y, st);
    if (s != ARK_OK) return s;
    const int add = c->force_eps > 0.0 ? 1 : 0;
    const float inv0 = (float)(0.5 / c->prm.dx1), inv1 = (float)(0.5 / c->prm.dx2), inv2 = (float)(0.5 / c->prm.dx3);
    ark_closure_hp* h = c->closure_hp;
    if (!h) {
        hipLaunchKernelGGL(k_les_force, interior_grid(g, 0), dim3(256), 0, st, g, c->les_Y, inv0, inv1, inv2, c->force,
                           c->force_stride, add);
        return c->hip(hipGetLastError(), "closure force");
    }
    // only the shells >= kcut of the acceleration
    const size_t n = (size_t)g.N1 * g.N2 * g.N3, nc = (size_t)(g.N1 / 2 + 1) * g.N2 * g.N3;
    hipLaunchKernelGGL(k_les_force_dense, interior_grid(g, 0), dim3(256), 0, st, g, c->les_Y, inv0, inv1, inv2,
                       h->real);
    s = c->hip(hipGetLastError(), "closure force");
    for (int m = 0; m < 3 && s == ARK_OK; ++m)
        if (hipfftExecR2C(h->plan_f, h->real + m * n, (hipfftComplex*)(h->spec + m * nc)) != HIPFFT_SUCCESS)
            s = c->fail(ARK_ERR_HIP, "closure R2C");
    if (s == ARK_OK) {
        hipLaunchKernelGGL(k_keep_high_shells, dim3(1024), dim3(256), 0, st, h->spec, g.N1, g.N2, g.N3, g.N1 / 2 + 1,
                           h->kcut, (float)(1.0 / (double)n));
        s = c->hip(hipGetLastError(), "closure projection");
    }
    for (int m = 0; m < 3 && s == ARK_OK; ++m)
        if (hipfftExecC2R(h->plan_b, (hipfftComplex*)(h->spec + m * nc), h->real + m * n) != HIPFFT_SUCCESS)
            s = c->fail(ARK_ERR_HIP, "closure C2R");
    if (s != ARK_OK) return s;
    hipLaunchKernelGGL(k_put_force, interior_grid(g, 0), dim3(256), 0, st, g, h->real, c->force, c->force_stride, add);
    return c->hip(hipGetLastError(), "closure force add");
}

extern "C" {

ark_status ark_set_closure(ark_ctx* c, ark_mlp* mlp, double delta) {
    if (!c) return ARK_ERR_INVALID;
    if (c->prm.precision != ARK_PREC_MIXED)
        return c->fail(ARK_ERR_UNSUPPORTED, "closure: mixed-precision contexts only");
    if (!mlp) {
        c->closure = nullptr;
        return ARK_OK;
    }
    ark_status s = ensure_force(c);
    const size_t n = (size_t)c->geo.N1 * c->geo.N2 * c->geo.N3;
    if (s == ARK_OK && !c->les_Y) {
        s = c->hip(hipMalloc(&c->les_Y, n * 6 * sizeof(float)), "hipMalloc closure");
        if (s == ARK_OK) {
            c->allocs.push_back(c->les_Y);
            c->bytes += n * 6 * sizeof(float);
        }
    }
    if (s == ARK_OK) {
        c->closure = mlp;
        c->closure_delta = delta;
    }
    return s;
}

ark_status ark_set_closure_cutoff(ark_ctx* c, int kcut) {
    if (!c || kcut < 0) return ARK_ERR_INVALID;
    if (c->prm.precision != ARK_PREC_MIXED)
        return c->fail(ARK_ERR_UNSUPPORTED, "closure: mixed-precision contexts only");
    if (kcut == 0) {
        ark_status s = c->hip(hipStreamSynchronize(c->stream), "sync");
        ark_closure_hp_free(c);
        return s;
    }
    if (c->closure_hp) {
        c->closure_hp->kcut = kcut;
        return ARK_OK;
    }
    const Geo& g = c->geo;
    ark_closure_hp* h = new ark_closure_hp;
    h->kcut = kcut;
    c->closure_hp = h;
    const size_t n = (size_t)g.N1 * g.N2 * g.N3, nc = (size_t)(g.N1 / 2 + 1) * g.N2 * g.N3;
    ark_status s = c->hip(hipMalloc(&h->real, 3 * n * sizeof(float)), "hipMalloc closure projection");
    if (s == ARK_OK) s = c->hip(hipMalloc(&h->spec, 3 * nc * sizeof(float2)), "hipMalloc closure projection");
    if (s == ARK_OK && (hipfftPlan3d(&h->plan_f, g.N3, g.N2, g.N1, HIPFFT_R2C) != HIPFFT_SUCCESS ||
                        hipfftPlan3d(&h->plan_b, g.N3, g.N2, g.N1, HIPFFT_C2R) != HIPFFT_SUCCESS))
        s = c->fail(ARK_ERR_HIP, "hipfftPlan3d for the closure projection failed");
    if (s == ARK_OK && (hipfftSetStream(h->plan_f, c->stream) != HIPFFT_SUCCESS ||
                        hipfftSetStream(h->plan_b, c->stream) != HIPFFT_SUCCESS))
        s = c->fail(ARK_ERR_HIP, "hipfftSetStream");
    if (s != ARK_OK) ark_closure_hp_free(c);
    return s;
}

ark_status ark_init_state(ark_ctx* c, const double* u, const double* rho) {
    if (!c || !u) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    const Geo& g = c->geo;
    if ((g.N1 & 1) || (g.N2 & 1) || (g.N3 & 1)) return c->fail(ARK_ERR_UNSUPPORTED, "init_velocity: N must be even");
    Fft f;
    ark_status s = make_fft(c, f, true);
    if (s != ARK_OK) return s;
    const long long n = (long long)g.N1 * g.N2 * g.N3, nc = (long long)f.s.H1 * g.N2 * g.N3;
    const int nf = rho ? 4 : 3;
    DevBuf real, spec, work;
    s = c->hip(hipMalloc(&real.p, nf * n * sizeof(double)), "hipMalloc");
    if (s == ARK_OK) s = c->hip(hipMalloc(&spec.p, nf * nc * sizeof(hipfftDoubleComplex)), "hipMalloc");
    if (s == ARK_OK) s = c->hip(hipMalloc(&work.p, nc * sizeof(hipfftDoubleComplex)), "hipMalloc");
    if (s == ARK_OK) s = c->hip(hipMemcpyAsync(real.p, u, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    double* R = (double*)real.p;
    if (s == ARK_OK && rho) {
        // the density perturbation rho - rho0, formed on the host in FP64
        std::vector<double> dr(rho, rho + n);
        for (double& v : dr) v -= c->prm.ro0;
        s = c->hip(hipMemcpyAsync(R + 3 * n, dr.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
        if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    }
    hipfftDoubleComplex* U = (hipfftDoubleComplex*)spec.p;
    for (int d = 0; d < nf && s == ARK_OK; ++d)
        if (hipfftExecD2Z(f.plan_f, R + d * n, U + d * nc) != HIPFFT_SUCCESS) s = c->fail(ARK_ERR_HIP, "D2Z");
    if (s != ARK_OK) return s;
    s = install_velocity(c, f, U, (hipfftDoubleComplex*)work.p, R, 1.0 / (double)n, rho ? U + 3 * nc : nullptr);
    // face densities are ro0 + p/c^2 of the installed pressures, bit for bit
    if (s == ARK_OK && rho) c->rho_stored = false;
    return s;
}

ark_status ark_init_velocity(ark_ctx* c, const double* u) { return ark_init_state(c, u, nullptr); }

ark_status ark_filter_density(ark_ctx* c, int width, double* rhobar) {
    if (!c || width < 1 || !rhobar) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    const Geo& g = c->geo;
    if (g.N1 % width || g.N2 % width || g.N3 % width)
        return c->fail(ARK_ERR_INVALID, "filter: the width must divide N1, N2, N3");
    const int Nc1 = g.N1 / width, Nc2 = g.N2 / width, Nc3 = g.N3 / width;
    const long long nc = (long long)Nc1 * Nc2 * Nc3;
    DevBuf db;
    ark_status s = c->hip(hipMalloc(&db.p, nc * sizeof(double)), "hipMalloc filter");
    if (s != ARK_OK) return s;
    hipLaunchKernelGGL(k_box_mean, dim3((Nc1 + 255) / 256, Nc2, Nc3), dim3(256), 0, c->stream, g, c->cur[F_ronCon],
                       width, (double*)db.p);
    s = c->hip(hipGetLastError(), "box mean");
    if (s == ARK_OK)
        s = c->hip(hipMemcpyAsync(rhobar, db.p, nc * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    return s == ARK_OK ? c->hip(hipStreamSynchronize(c->stream), "sync") : s;
}

ark_status ark_init_turbulence(ark_ctx* c, unsigned long long seed, double k0, double urms) {
    if (!c || !(k0 > 0.0) || !(urms >= 0.0)) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    const Geo& g = c->geo;
    if ((g.N1 & 1) || (g.N2 & 1) || (g.N3 & 1)) return c->fail(ARK_ERR_UNSUPPORTED, "turbulence: N must be even");
    Fft f;
    ark_status s = make_fft(c, f, true);
    if (s != ARK_OK) return s;
    const long long n = (long long)g.N1 * g.N2 * g.N3, nc = (long long)f.s.H1 * g.N2 * g.N3;
    DevBuf real, spec, work, acc;
    s = c->hip(hipMalloc(&real.p, 3 * n * sizeof(double)), "hipMalloc turbulence");
    if (s == ARK_OK) s = c->hip(hipMalloc(&spec.p, 3 * nc * sizeof(hipfftDoubleComplex)), "hipMalloc turbulence");
    if (s == ARK_OK) s = c->hip(hipMalloc(&work.p, nc * sizeof(hipfftDoubleComplex)), "hipMalloc turbulence");
    if (s == ARK_OK) s = c->hip(hipMalloc(&acc.p, (kSumBlocks + 1) * sizeof(double)), "hipMalloc turbulence");
    if (s != ARK_OK) return s;
    double* R = (double*)real.p;
    hipfftDoubleComplex* U = (hipfftDoubleComplex*)spec.p;
    hipfftDoubleComplex* W = (hipfftDoubleComplex*)work.p;
    hipLaunchKernelGGL(k_noise, dim3(4096), dim3(256), 0, c->stream, f.s, R, seed);
    s = c->hip(hipGetLastError(), "noise");
    for (int d = 0; d < 3 && s == ARK_OK; ++d)
        if (hipfftExecD2Z(f.plan_f, R + d * n, U + d * nc) != HIPFFT_SUCCESS) s = c->fail(ARK_ERR_HIP, "D2Z");
    if (s != ARK_OK) return s;
    hipLaunchKernelGGL(k_shape, dim3(4096), dim3(256), 0, c->stream, f.s, U, k0);
    double* part = (double*)acc.p;
    hipLaunchKernelGGL(k_spec_energy, dim3(kSumBlocks), dim3(256), 0, c->stream, f.s, U, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, part, kSumBlocks, part + kSumBlocks);
    s = c->hip(hipGetLastError(), "spectral energy");
    double e = 0.0;
    if (s == ARK_OK)
        s = c->hip(hipMemcpyAsync(&e, part + kSumBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s != ARK_OK) return s;
    // The unnormalised inverse R(x) = sum_k U_k e^{ikx} has (Parseval)
    // sum_x |R|^2 = N^3 sum_k |U_k|^2 = N^3 e, so its per-component rms is
    // sqrt(e / 3); R is scaled to the requested u_rms.
    const double rms_raw = sqrt(e / 3.0);
    const double scale = rms_raw > 0.0 ? urms / rms_raw : 0.0;

    return install_velocity(c, f, U, W, R, scale);
}

ark_status ark_spectrum(ark_ctx* c, int nbins, double* E) {
    if (!c || nbins < 1 || !E) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    const Geo& g = c->geo;
    Fft f;
    ark_status s = make_fft(c, f, false);
    if (s != ARK_OK) return s;
    const long long n = (long long)g.N1 * g.N2 * g.N3, nc = (long long)f.s.H1 * g.N2 * g.N3;
    DevBuf real, spec, bins;
    s = c->hip(hipMalloc(&real.p, n * sizeof(double)), "hipMalloc spectrum");
    if (s == ARK_OK) s = c->hip(hipMalloc(&spec.p, 3 * nc * sizeof(hipfftDoubleComplex)), "hipMalloc spectrum");
    if (s == ARK_OK) s = c->hip(hipMalloc(&bins.p, nbins * sizeof(double)), "hipMalloc spectrum");
    const int cellu[3] = {F_u1nCon, F_u2nCon, F_u3nCon};
    hipfftDoubleComplex* U = (hipfftDoubleComplex*)spec.p;
    for (int m = 0; m < 3 && s == ARK_OK; ++m) {
        hipLaunchKernelGGL(k_extract, interior_grid(g, 0), dim3(256), 0, c->stream, g, c->cur[cellu[m]], (double*)real.p);
        s = c->hip(hipGetLastError(), "extract");
        if (s == ARK_OK && hipfftExecD2Z(f.plan_f, (double*)real.p, U + m * nc) != HIPFFT_SUCCESS)
            s = c->fail(ARK_ERR_HIP, "D2Z");
    }
    if (s == ARK_OK) {
        hipLaunchKernelGGL(k_shell_sum, dim3(nbins), dim3(256), 0, c->stream, f.s, U, (double*)bins.p);
        s = c->hip(hipGetLastError(), "shell sum");
    }
    if (s == ARK_OK)
        s = c->hip(hipMemcpyAsync(E, bins.p, nbins * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    return s;
}

ark_status ark_dissipation(ark_ctx* c, double* eps) {
    if (!c || !eps) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    DevBuf acc;
    ark_status s = c->hip(hipMalloc(&acc.p, (kSumBlocks + 1) * sizeof(double)), "hipMalloc");
    double* part = (double*)acc.p;
    if (s == ARK_OK) {
        hipLaunchKernelGGL(k_dissipation, dim3(kSumBlocks), dim3(256), 0, c->stream, c->geo, c->cur[F_u1nCon],
                           c->cur[F_u2nCon], c->cur[F_u3nCon], 0.5 / c->prm.dx1, 0.5 / c->prm.dx2, 0.5 / c->prm.dx3,
                           part);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, part, kSumBlocks, part + kSumBlocks);
        s = c->hip(hipGetLastError(), "dissipation");
    }
    double v = 0.0;
    if (s == ARK_OK)
        s = c->hip(hipMemcpyAsync(&v, part + kSumBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    *eps = c->prm.VIS * v / ((double)c->geo.N1 * c->geo.N2 * c->geo.N3);
    return s;
}

ark_status ark_set_forcing(ark_ctx* c, unsigned long long seed, double eps, double tau) {
    if (!c || !(eps >= 0.0) || !(tau >= 0.0)) return ARK_ERR_INVALID;
    if (c->prm.precision != ARK_PREC_MIXED)
        return c->fail(ARK_ERR_UNSUPPORTED, "forcing: mixed-precision contexts only");
    ark_status s = ARK_OK;
    if (eps > 0.0) {
        ForceMode h[16];
        int q = 0;
        for (int a = -2; a <= 2; ++a)
            for (int b = -2; b <= 2; ++b)
                for (int d = -2; d <= 2; ++d) {
                    const int k2 = a * a + b * b + d * d;
                    const int first = a != 0 ? a : b != 0 ? b : d;
                    if (k2 >= 1 && k2 <= 4 && first > 0) h[q++] = ForceMode{{a, b, d}};
                }
        // the force kernels hold the same list as compile-time constants
        for (int n = 0; n < 16; ++n)
            for (int d = 0; d < 3; ++d)
                if (q != 16 || h[n].m[d] != forced_mode(n, d))
                    return c->fail(ARK_ERR_STATE, "forcing: mode table mismatch");
        s = c->hip(hipMemcpyToSymbol(HIP_SYMBOL(c_modes), h, sizeof h), "force modes");
    }
    if (s == ARK_OK && eps > 0.0) s = ensure_force(c);
    c->force_seed = seed;
    c->force_eps = eps;
    c->force_tau = tau;
    return s;
}

ark_status ark_filter_box(ark_ctx* c, int width, double* ubar, double* tau, double* grad) {
    if (!c || width < 1 || !ubar || !tau || !grad) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    const Geo& g = c->geo;
    if (g.N1 % width || g.N2 % width || g.N3 % width)
        return c->fail(ARK_ERR_INVALID, "filter: the width must divide N1, N2, N3");
    const int Nc1 = g.N1 / width, Nc2 = g.N2 / width, Nc3 = g.N3 / width;
    const long long nc = (long long)Nc1 * Nc2 * Nc3;
    DevBuf db;
    ark_status s = c->hip(hipMalloc(&db.p, 18 * nc * sizeof(double)), "hipMalloc filter");
    if (s != ARK_OK) return s;
    double* u = (double*)db.p;
    double* t = u + 3 * nc;
    double* gr = u + 9 * nc;
    dim3 grid((Nc1 + 255) / 256, Nc2, Nc3);
    hipLaunchKernelGGL(k_box_filter, grid, dim3(256), 0, c->stream, g, c->cur[F_u1nCon], c->cur[F_u2nCon],
                       c->cur[F_u3nCon], width, u, t);
    s = c->hip(hipGetLastError(), "box filter");
    if (s == ARK_OK) {
        hipLaunchKernelGGL(k_coarse_gradient, grid, dim3(256), 0, c->stream, Nc1, Nc2, Nc3, u, width * c->prm.dx1,
                           width * c->prm.dx2, width * c->prm.dx3, gr);
        s = c->hip(hipGetLastError(), "coarse gradient");
    }
    if (s == ARK_OK) s = c->hip(hipMemcpyAsync(ubar, u, 3 * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    if (s == ARK_OK) s = c->hip(hipMemcpyAsync(tau, t, 6 * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    if (s == ARK_OK) s = c->hip(hipMemcpyAsync(grad, gr, 9 * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    return s;
}

}  // extern "C"

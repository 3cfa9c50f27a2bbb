// Learned subgrid closure: a per-cell MLP 10 -> 64 -> 64 -> 6 (SiLU) mapping
// the filtered velocity-gradient tensor and the grid scale to the six subgrid
// stresses.  No counterpart in the reference.
//
// Arithmetic: BF16 weights and activations, FP32 accumulation, on the matrix
// cores (v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)+v],
// B[8(l/16)+v][l%16] and D[4(l/16)+v][l%16]).  Every layer of a wave's chunk
// of 32 samples is a product of 16 x 32 by 32 x 16 tiles.  Activations pass
// through the wave's LDS twice laid out: [sample][feature] (the B operand of
// the forward and input-gradient products) and [feature][sample] (both
// operands of the weight-gradient products, which sum over samples).
// Training keeps FP32 master weights and Adam moments; each wave writes its
// gradient sums to its own slot and one kernel adds the slots in a fixed
// order, applies Adam and refreshes the BF16 copies, so training is
// deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ark/ark.h"
#include "ark_internal.hpp"
#include "ark_rng.hpp"

using namespace arkd;

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;

constexpr int NI = 10, NH = 64, NO = 6;
// FP32 parameter vector: W1[64][10] b1[64] W2[64][64] b2[64] W3[6][64] b3[6]
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NI, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NO * NH, NPARAM = OFF_B3 + NO;
constexpr int SLOT = NPARAM + 2;        // per-wave partial: gradient, loss sum, sample count
// BF16 copies: W1b[64][32] (inputs padded to 32), W2b[64][64], W2t[64][64] (transposed),
// W3b[16][64] (outputs padded to 16), W3t[64][32] (transposed, padded); and
// the inference kernel's copies W1i[64][32], W2i[64][64], W3i[16][64] with
// their columns in the operand order of that kernel (in_pos, hid_pos below).
constexpr int BOFF_W1 = 0, BOFF_W2 = BOFF_W1 + 64 * 32, BOFF_W2T = BOFF_W2 + 64 * 64, BOFF_W3 = BOFF_W2T + 64 * 64,
              BOFF_W3T = BOFF_W3 + 16 * 64, BOFF_W1I = BOFF_W3T + 64 * 32, BOFF_W2I = BOFF_W1I + 64 * 32,
              BOFF_W3I = BOFF_W2I + 64 * 64, NBF = BOFF_W3I + 16 * 64;

// Operand order of the inference kernel.  Input feature f (d_j u_i at 3i + j,
// then the filter width) sits at K position in_pos(f): lane group g of the
// B operand holds the three derivatives of u_g, group 3 the width.  Hidden
// unit k sits at K position hid_pos(k) of the next layer's B operand: with
// that order a lane's four accumulator rows of output tiles 2c and 2c+1
// (units 16t + 4g + v) are exactly its eight B rows of K chunk c, so the
// activations never leave the lane's registers.
__host__ __device__ constexpr int in_pos(int f) { return f < 9 ? 8 * (f / 3) + f % 3 : 24; }
__host__ __device__ constexpr int hid_pos(int k) {
    return 32 * (k / 32) + 8 * ((k % 16) / 4) + 4 * ((k % 32) / 16) + k % 4;
}
constexpr int WAVES = 2, THREADS = 64 * WAVES;

struct Norm { float in_mean[NI], in_inv[NI], out_mean[NO], out_std[NO], out_inv[NO]; };

__device__ __forceinline__ f4 mma(bf16x8 a, bf16x8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *(const bf16x8*)p; }
// Logistic function from the hardware exp2 and reciprocal (about 1 ulp each;
// the activations are rounded to BF16 next, 2^-9 relative), the same in
// training and inference.  exp2 overflows to +inf for z < -88 (then 0) and
// underflows to 0 for large z (then 1), the logistic's own limits.
__device__ __forceinline__ float sigm(float z) {
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}

// Per-wave LDS: bf16 activations of one chunk of 32 samples.
struct WaveLds {
    __bf16 x_sf[32 * 32], x_fs[16 * 32];
    __bf16 h1_sf[32 * 64], h1_fs[64 * 32];
    __bf16 h2_sf[32 * 64], h2_fs[64 * 32];
    __bf16 dy_sf[32 * 32], dy_fs[16 * 32];
    __bf16 dz2_sf[32 * 64], dz2_fs[64 * 32];
    __bf16 dz1_fs[64 * 32];
};

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); }

// Inputs of the chunk: sample s (0..31) of lane s < 32, normalised, as BF16
// in both layouts; `idx` gives the sample's row in X.
__device__ void stage_inputs(WaveLds& L, const float* X, long long row, bool valid, const Norm& nm, int lane) {
    if (lane < 32) {
        float v[NI];
        for (int f = 0; f < NI; ++f) v[f] = valid ? (X[row * NI + f] - nm.in_mean[f]) * nm.in_inv[f] : 0.f;
        for (int f = 0; f < 32; ++f) L.x_sf[lane * 32 + f] = (__bf16)(f < NI ? v[f] : 0.f);
        for (int f = 0; f < 16; ++f) L.x_fs[f * 32 + lane] = (__bf16)(f < NI ? v[f] : 0.f);
    }
    wave_sync();
}

// Forward of one chunk: Z1, Z2 (FP32, lane layout) and Y; H1, H2 into LDS.
struct Fwd {
    f4 z1[4][2], z2[4][2], y[2];
};

__device__ void forward(WaveLds& L, const __bf16* Wb, const float* P, Fwd& F, int lane) {
    const int r = lane % 16, g = lane / 16;
    for (int t = 0; t < 4; ++t) {
        const bf16x8 a = ld8(Wb + BOFF_W1 + (16 * t + r) * 32 + 8 * g);
        for (int n = 0; n < 2; ++n) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mma(a, ld8(L.x_sf + (16 * n + r) * 32 + 8 * g), acc);
            const int s = 16 * n + r;
            for (int v = 0; v < 4; ++v) {
                const int h = 16 * t + 4 * g + v;
                acc[v] += P[OFF_B1 + h];
                const __bf16 hb = (__bf16)(acc[v] * sigm(acc[v]));
                L.h1_sf[s * 64 + h] = hb;
                L.h1_fs[h * 32 + s] = hb;
            }
            F.z1[t][n] = acc;
        }
    }
    wave_sync();
    for (int t = 0; t < 4; ++t)
        for (int n = 0; n < 2; ++n) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < 2; ++c)
                acc = mma(ld8(Wb + BOFF_W2 + (16 * t + r) * 64 + 32 * c + 8 * g),
                          ld8(L.h1_sf + (16 * n + r) * 64 + 32 * c + 8 * g), acc);
            const int s = 16 * n + r;
            for (int v = 0; v < 4; ++v) {
                const int h = 16 * t + 4 * g + v;
                acc[v] += P[OFF_B2 + h];
                const __bf16 hb = (__bf16)(acc[v] * sigm(acc[v]));
                L.h2_sf[s * 64 + h] = hb;
                L.h2_fs[h * 32 + s] = hb;
            }
            F.z2[t][n] = acc;
        }
    wave_sync();
    for (int n = 0; n < 2; ++n) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < 2; ++c)
            acc = mma(ld8(Wb + BOFF_W3 + r * 64 + 32 * c + 8 * g), ld8(L.h2_sf + (16 * n + r) * 64 + 32 * c + 8 * g),
                      acc);
        for (int v = 0; v < 4; ++v) {
            const int o = 4 * g + v;
            if (o < NO) acc[v] += P[OFF_B3 + o];
        }
        F.y[n] = acc;
    }
}

// Gradient of the mean squared error over `count` samples and 6 outputs for
// chunks drawn from the data set (rows from Philox when `sample`, else
// consecutive), one partial slot per wave.
__global__ void __launch_bounds__(THREADS) k_mlp_grad(const float* X, const float* T, long long ndata, long long nbatch,
                                                      int sample, unsigned long long seed, unsigned step,
                                                      const float* P, const __bf16* Wb, Norm nm, float* partial) {
    __shared__ WaveLds lds[WAVES];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    WaveLds& L = lds[wave];
    const int r = lane % 16, g = lane / 16;
    const long long gw = (long long)blockIdx.x * WAVES + wave, nw = (long long)gridDim.x * WAVES;
    const long long nchunks = (nbatch + 31) / 32;
    f4 dw3[4], dw2[4][4], dw1[4];
    for (int u = 0; u < 4; ++u) {
        dw3[u] = f4{0.f, 0.f, 0.f, 0.f};
        dw1[u] = f4{0.f, 0.f, 0.f, 0.f};
        for (int t = 0; t < 4; ++t) dw2[t][u] = f4{0.f, 0.f, 0.f, 0.f};
    }
    float db1[4][4] = {}, db2[4][4] = {}, db3[4] = {};
    float loss = 0.f;
    const float scale = 2.f / (float)(NO * nbatch);
    for (long long ch = gw; ch < nchunks; ch += nw) {
        // rows of this chunk's 32 samples
        long long row = 0;
        const long long pos = ch * 32 + (lane % 32);
        const bool valid = pos < nbatch;
        if (sample) {
            const U4 rr = philox4x32_10(U4{{(uint32_t)pos, (uint32_t)(pos >> 32), step, 0x6d6c70u}}, (uint32_t)seed,
                                        (uint32_t)(seed >> 32));
            row = (long long)((((uint64_t)rr.v[0] << 32) | rr.v[1]) % (uint64_t)ndata);
        } else {
            row = valid ? pos : 0;
        }
        stage_inputs(L, X, row, valid, nm, lane);
        Fwd F;
        forward(L, Wb, P, F, lane);
        // output error: lane holds outputs 4g+v of sample 16n+r; the target row
        // of that sample comes from the lane that staged it
        for (int n = 0; n < 2; ++n) {
            const int s = 16 * n + r;
            const long long srow = __shfl(row, s);
            const bool sval = ch * 32 + s < nbatch;
            for (int v = 0; v < 4; ++v) {
                const int o = 4 * g + v;
                float d = 0.f;
                if (o < NO && sval) {
                    const float t = (T[srow * NO + o] - nm.out_mean[o]) * nm.out_inv[o];
                    const float e = F.y[n][v] - t;
                    loss += e * e;
                    d = scale * e;
                }
                db3[v] += d;
                const __bf16 db = (__bf16)d;
                L.dy_sf[s * 32 + o] = db;
                if (o < 16) L.dy_fs[o * 32 + s] = db;
            }
            for (int o = 16 + 4 * g; o < 16 + 4 * g + 4; ++o) L.dy_sf[s * 32 + o] = (__bf16)0.f;
        }
        wave_sync();
        // dH2 = W3^T dY, dZ2 = dH2 * silu'(Z2)
        for (int t = 0; t < 4; ++t)
            for (int n = 0; n < 2; ++n) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mma(ld8(Wb + BOFF_W3T + (16 * t + r) * 32 + 8 * g), ld8(L.dy_sf + (16 * n + r) * 32 + 8 * g), acc);
                const int s = 16 * n + r;
                for (int v = 0; v < 4; ++v) {
                    const int h = 16 * t + 4 * g + v;
                    const float z = F.z2[t][n][v], sg = sigm(z);
                    const float d = acc[v] * sg * (1.f + z * (1.f - sg));
                    db2[t][v] += d;
                    const __bf16 db = (__bf16)d;
                    L.dz2_sf[s * 64 + h] = db;
                    L.dz2_fs[h * 32 + s] = db;
                }
            }
        wave_sync();
        // dH1 = W2^T dZ2, dZ1 = dH1 * silu'(Z1)
        for (int t = 0; t < 4; ++t)
            for (int n = 0; n < 2; ++n) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                for (int c = 0; c < 2; ++c)
                    acc = mma(ld8(Wb + BOFF_W2T + (16 * t + r) * 64 + 32 * c + 8 * g),
                              ld8(L.dz2_sf + (16 * n + r) * 64 + 32 * c + 8 * g), acc);
                const int s = 16 * n + r;
                for (int v = 0; v < 4; ++v) {
                    const int h = 16 * t + 4 * g + v;
                    const float z = F.z1[t][n][v], sg = sigm(z);
                    const float d = acc[v] * sg * (1.f + z * (1.f - sg));
                    db1[t][v] += d;
                    L.dz1_fs[h * 32 + s] = (__bf16)d;
                }
            }
        wave_sync();
        // weight gradients, summed over the chunk's 32 samples
        for (int u = 0; u < 4; ++u)
            dw3[u] = mma(ld8(L.dy_fs + r * 32 + 8 * g), ld8(L.h2_fs + (16 * u + r) * 32 + 8 * g), dw3[u]);
        for (int t = 0; t < 4; ++t)
            for (int u = 0; u < 4; ++u)
                dw2[t][u] = mma(ld8(L.dz2_fs + (16 * t + r) * 32 + 8 * g), ld8(L.h1_fs + (16 * u + r) * 32 + 8 * g),
                                dw2[t][u]);
        for (int t = 0; t < 4; ++t)
            dw1[t] = mma(ld8(L.dz1_fs + (16 * t + r) * 32 + 8 * g), ld8(L.x_fs + r * 32 + 8 * g), dw1[t]);
        wave_sync();
    }
    // this wave's slot: weights from the lane layout, biases summed over the
    // 16 lanes (samples) that share an output group
    float* out = partial + gw * SLOT;
    for (int v = 0; v < 4; ++v) {
        const int o = 4 * g + v;
        for (int u = 0; u < 4; ++u)
            if (o < NO) out[OFF_W3 + o * 64 + 16 * u + r] = dw3[u][v];
        for (int t = 0; t < 4; ++t) {
            const int h = 16 * t + 4 * g + v;
            for (int u = 0; u < 4; ++u) out[OFF_W2 + h * 64 + 16 * u + r] = dw2[t][u][v];
            if (r < NI) out[OFF_W1 + h * NI + r] = dw1[t][v];
        }
    }
    for (int m = 8; m >= 1; m >>= 1)
        for (int v = 0; v < 4; ++v) {
            db3[v] += __shfl_xor(db3[v], m);
            for (int t = 0; t < 4; ++t) {
                db1[t][v] += __shfl_xor(db1[t][v], m);
                db2[t][v] += __shfl_xor(db2[t][v], m);
            }
        }
    for (int m = 32; m >= 1; m >>= 1) loss += __shfl_xor(loss, m);
    if (r == 0)
        for (int v = 0; v < 4; ++v) {
            const int o = 4 * g + v;
            if (o < NO) out[OFF_B3 + o] = db3[v];
            for (int t = 0; t < 4; ++t) {
                out[OFF_B1 + 16 * t + 4 * g + v] = db1[t][v];
                out[OFF_B2 + 16 * t + 4 * g + v] = db2[t][v];
            }
        }
    if (lane == 0) {
        out[NPARAM] = loss;
        out[NPARAM + 1] = 0.f;
    }
}

// BF16 copies of the FP32 weights (padded and transposed where the products
// need it).
__device__ void refresh_one(const float* P, __bf16* Wb, int p) {
    if (p >= OFF_W1 && p < OFF_B1) {
        const int h = (p - OFF_W1) / NI, i = (p - OFF_W1) % NI;
        Wb[BOFF_W1 + h * 32 + i] = (__bf16)P[p];
        Wb[BOFF_W1I + h * 32 + in_pos(i)] = (__bf16)P[p];
    } else if (p >= OFF_W2 && p < OFF_B2) {
        const int h = (p - OFF_W2) / NH, k = (p - OFF_W2) % NH;
        Wb[BOFF_W2 + h * 64 + k] = (__bf16)P[p];
        Wb[BOFF_W2T + k * 64 + h] = (__bf16)P[p];
        Wb[BOFF_W2I + h * 64 + hid_pos(k)] = (__bf16)P[p];
    } else if (p >= OFF_W3 && p < OFF_B3) {
        const int o = (p - OFF_W3) / NH, h = (p - OFF_W3) % NH;
        Wb[BOFF_W3 + o * 64 + h] = (__bf16)P[p];
        Wb[BOFF_W3T + h * 32 + o] = (__bf16)P[p];
        Wb[BOFF_W3I + o * 64 + hid_pos(h)] = (__bf16)P[p];
    }
}

__global__ void k_refresh(const float* P, __bf16* Wb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPARAM) refresh_one(P, Wb, p);
}

// Sum of the wave slots (fixed order) into G (FP64), then optionally Adam.
__global__ void k_reduce_adam(const float* partial, int nslots, double* G, float* P, float* M, float* V, __bf16* Wb,
                              int adam, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM + 1) return;
    double s = 0.0;
    for (int w = 0; w < nslots; ++w) s += (double)partial[(long long)w * SLOT + p];
    G[p] = s;
    if (!adam || p >= NPARAM) return;
    const float gr = (float)s;
    const float m = b1 * M[p] + (1.f - b1) * gr;
    const float v = b2 * V[p] + (1.f - b2) * gr * gr;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    refresh_one(P, Wb, p);
}

// Inference of 16 samples per wave and step of the loop, everything in
// registers: the wave's weights (BF16 operands, in_pos/hid_pos order) and
// biases are loaded once; lane (g, r) works on sample r of the step.  Layer 1
// is four 16x16x32 products (output tile t = units 16t..16t+15), layers 2 and
// 3 take the SiLU activations of tiles 2c and 2c+1 as their K chunk c.  The
// samples come from X (n x 10, raw units) or, on the solver's grid, from the
// central-difference velocity gradient of interior cell s (x fastest) and
// the filter width; Y receives n x 6 stresses (raw units).
struct GridIn {
    Geo g;
    const double* u[3];
    int cs;             // element stride of u: 1 per field, 4 in the 32-byte cell records
    double inv[3];      // 1 / (2 dx_d)
    float delta;
};

// Bounded to three waves per SIMD (168 registers; unbounded it took 150 VGPRs
// and 32 AGPRs, two waves): its SiLU chains hide better with the third wave,
// and a 128^3 closure step went from 0.371 to 0.347 ms, bitwise the same.
template <bool Grid>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_mlp_infer(const float* X, GridIn gi, long long n, const float* P,
                                                   const __bf16* Wb, Norm nm, float* Y) {
    const int lane = threadIdx.x % 64, r = lane % 16, g = lane / 16;
    bf16x8 w1[4], w2[4][2], w3[2];
    float b1[4][4], b2[4][4], b3[4];
    for (int t = 0; t < 4; ++t) {
        w1[t] = ld8(Wb + BOFF_W1I + (16 * t + r) * 32 + 8 * g);
        for (int c = 0; c < 2; ++c) w2[t][c] = ld8(Wb + BOFF_W2I + (16 * t + r) * 64 + 32 * c + 8 * g);
        for (int v = 0; v < 4; ++v) {
            b1[t][v] = P[OFF_B1 + 16 * t + 4 * g + v];
            b2[t][v] = P[OFF_B2 + 16 * t + 4 * g + v];
        }
    }
    for (int c = 0; c < 2; ++c) w3[c] = ld8(Wb + BOFF_W3I + r * 64 + 32 * c + 8 * g);
    for (int v = 0; v < 4; ++v) b3[v] = 4 * g + v < NO ? P[OFF_B3 + 4 * g + v] : 0.f;
    // this lane's three input features (group 3: the filter width only)
    const int nf = g < 3 ? 3 : 1;
    float im[3], ii[3], om[4], os[4];
    for (int j = 0; j < 3; ++j) {
        const int f = g < 3 ? 3 * g + j : 9;
        im[j] = nm.in_mean[f];
        ii[j] = nm.in_inv[f];
    }
    for (int v = 0; v < 4; ++v) {
        const int o = 4 * g + v < NO ? 4 * g + v : 0;
        om[v] = nm.out_mean[o];
        os[v] = nm.out_std[o];
    }
    const long long nblk = (n + 15) / 16;
    const long long w0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const long long nw = (long long)gridDim.x * blockDim.x / 64;
    // two independent blocks of 16 samples per loop step (blk, blk + nw), so
    // one block's products overlap the other's activations
    for (long long b0 = w0; b0 < nblk; b0 += 2 * nw) {
        long long sq[2];
        bool valid[2];
        bf16x8 bx[2];
        for (int q = 0; q < 2; ++q) {
            sq[q] = (b0 + q * nw) * 16 + r;
            valid[q] = sq[q] < n;
            const long long s = sq[q];
            float x[3] = {0.f, 0.f, 0.f};
            if (valid[q]) {
                if (Grid) {
                    if (g < 3) {
                        // n < 2^31 cells on the solver's grid (ark_create's limit)
                        const Geo& G = gi.g;
                        const int si = (int)s, row = si / G.N1;
                        const int i = si - row * G.N1 + 1, j = row % G.N2 + 1, k = row / G.N2 + 1;
                        const int ip = i == G.N1 ? 1 : i + 1, im_ = i == 1 ? G.N1 : i - 1;
                        const int jp = j == G.N2 ? 1 : j + 1, jm = j == 1 ? G.N2 : j - 1;
                        const int kp = k == G.N3 ? 1 : k + 1, km = k == 1 ? G.N3 : k - 1;
                        const double* u = gi.u[g];
                        const long long cs = gi.cs;
                        x[0] = (float)((u[cs * at(G, ip, j, k)] - u[cs * at(G, im_, j, k)]) * gi.inv[0]);
                        x[1] = (float)((u[cs * at(G, i, jp, k)] - u[cs * at(G, i, jm, k)]) * gi.inv[1]);
                        x[2] = (float)((u[cs * at(G, i, j, kp)] - u[cs * at(G, i, j, km)]) * gi.inv[2]);
                    } else {
                        x[0] = gi.delta;
                    }
                } else {
                    for (int j = 0; j < nf; ++j) x[j] = X[s * NI + (g < 3 ? 3 * g + j : 9)];
                }
            }
            for (int v = 0; v < 8; ++v) bx[q][v] = (__bf16)0.f;
            if (valid[q])
                for (int j = 0; j < nf; ++j) bx[q][j] = (__bf16)((x[j] - im[j]) * ii[j]);
        }
        f4 z[2][4];
        bf16x8 h[2][2];
        for (int q = 0; q < 2; ++q)
            for (int t = 0; t < 4; ++t) z[q][t] = mma(w1[t], bx[q], f4{0.f, 0.f, 0.f, 0.f});
        for (int q = 0; q < 2; ++q)
            for (int c = 0; c < 2; ++c)
                for (int v = 0; v < 4; ++v) {
                    const float za = z[q][2 * c][v] + b1[2 * c][v], zb = z[q][2 * c + 1][v] + b1[2 * c + 1][v];
                    h[q][c][v] = (__bf16)(za * sigm(za));
                    h[q][c][4 + v] = (__bf16)(zb * sigm(zb));
                }
        for (int q = 0; q < 2; ++q)
            for (int t = 0; t < 4; ++t)
                z[q][t] = mma(w2[t][1], h[q][1], mma(w2[t][0], h[q][0], f4{0.f, 0.f, 0.f, 0.f}));
        for (int q = 0; q < 2; ++q)
            for (int c = 0; c < 2; ++c)
                for (int v = 0; v < 4; ++v) {
                    const float za = z[q][2 * c][v] + b2[2 * c][v], zb = z[q][2 * c + 1][v] + b2[2 * c + 1][v];
                    h[q][c][v] = (__bf16)(za * sigm(za));
                    h[q][c][4 + v] = (__bf16)(zb * sigm(zb));
                }
        for (int q = 0; q < 2; ++q) {
            const f4 y = mma(w3[1], h[q][1], mma(w3[0], h[q][0], f4{0.f, 0.f, 0.f, 0.f}));
            if (valid[q] && g < 2) {
                float* o = Y + sq[q] * NO + 4 * g;
                for (int v = 0; v < 4; ++v)
                    if (4 * g + v < NO) o[v] = (y[v] + b3[v]) * os[v] + om[v];
            }
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
struct ark_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;                          // recorded on `stream` before another stream reads P, Wb
    float *P = nullptr, *M = nullptr, *V = nullptr;   // FP32 weights, Adam moments
    __bf16* Wb = nullptr;
    double* G = nullptr;                                // last gradient (and loss sum)
    float* partial = nullptr;
    int slots = 0;
    float *X = nullptr, *T = nullptr;                   // training data (raw units)
    long long ndata = 0;
    Norm nm{};
    long long adam_t = 0;
    unsigned long long seed = 0;
    std::string err;
    ark_status fail(ark_status s, const std::string& m) {
        err = m;
        return s;
    }
    ark_status hip(hipError_t e, const char* what) {
        return e == hipSuccess ? ARK_OK : fail(ARK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
};

namespace {

constexpr int kGradBlocks = 256;

// Inference grid: 4 waves per block, about eight blocks per CU of a 256-CU
// chip, or fewer for small n (16 samples per wave and loop step).
int infer_blocks(long long n) {
    const long long waves = (n + 15) / 16;
    return (int)std::max<long long>(1, std::min<long long>(2048, (waves + 3) / 4));
}

ark_status ensure_partial(ark_mlp* m) {
    if (m->partial) return ARK_OK;
    m->slots = kGradBlocks * WAVES;
    return m->hip(hipMalloc(&m->partial, (size_t)m->slots * SLOT * sizeof(float)), "hipMalloc mlp partial");
}

ark_status run_grad(ark_mlp* m, const float* X, const float* T, long long ndata, long long nbatch, int sample,
                    unsigned step, int adam, float lr) {
    ark_status s = ensure_partial(m);
    if (s != ARK_OK) return s;
    hipLaunchKernelGGL(k_mlp_grad, dim3(kGradBlocks), dim3(THREADS), 0, m->stream, X, T, ndata, nbatch, sample,
                       m->seed, step, m->P, m->Wb, m->nm, m->partial);
    s = m->hip(hipGetLastError(), "mlp grad");
    if (s != ARK_OK) return s;
    float c1 = 1.f, c2 = 1.f;
    if (adam) {
        ++m->adam_t;
        c1 = 1.f - (float)std::pow(0.9, (double)m->adam_t);
        c2 = 1.f - (float)std::pow(0.999, (double)m->adam_t);
    }
    hipLaunchKernelGGL(k_reduce_adam, dim3((NPARAM + 256) / 256), dim3(256), 0, m->stream, m->partial, m->slots, m->G,
                       m->P, m->M, m->V, m->Wb, adam, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
    return m->hip(hipGetLastError(), "mlp reduce");
}

}  // namespace

extern "C" {

int ark_mlp_param_count(void) { return NPARAM; }

ark_status ark_mlp_create(int device, unsigned long long seed, ark_mlp** out) {
    if (!out) return ARK_ERR_INVALID;
    *out = nullptr;
    ark_mlp* m = new ark_mlp();
    m->device = device;
    m->seed = seed;
    ark_status s = m->hip(hipSetDevice(device), "hipSetDevice");
    if (s == ARK_OK) s = m->hip(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream");
    if (s == ARK_OK) s = m->hip(hipEventCreateWithFlags(&m->done, hipEventDisableTiming), "event");
    if (s == ARK_OK) s = m->hip(hipMalloc(&m->P, 3 * NPARAM * sizeof(float)), "hipMalloc mlp");
    if (s == ARK_OK) s = m->hip(hipMalloc(&m->Wb, NBF * sizeof(__bf16)), "hipMalloc mlp");
    if (s == ARK_OK) s = m->hip(hipMalloc(&m->G, (NPARAM + 1) * sizeof(double)), "hipMalloc mlp");
    if (s == ARK_OK) {
        m->M = m->P + NPARAM;
        m->V = m->P + 2 * NPARAM;
        // He-style init from Philox: W ~ N(0, 1/fan_in) (SiLU), biases 0
        std::vector<float> h(3 * NPARAM, 0.f);
        auto fill = [&](int off, int rows, int cols, uint32_t tag) {
            for (int q = 0; q < rows * cols; q += 4) {
                double gz[4];
                gauss4(seed, (uint32_t)q, tag, 0x494e4954u, 0u, gz);
                for (int e = 0; e < 4 && q + e < rows * cols; ++e) h[off + q + e] = (float)(gz[e] / std::sqrt((double)cols));
            }
        };
        fill(OFF_W1, NH, NI, 1);
        fill(OFF_W2, NH, NH, 2);
        fill(OFF_W3, NO, NH, 3);
        s = m->hip(hipMemcpyAsync(m->P, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, m->stream),
                   "mlp init");
        if (s == ARK_OK) s = m->hip(hipStreamSynchronize(m->stream), "sync");   // h is freed at the end of this block
    }
    // Every fill, copy and kernel of the network is ordered on its own stream
    // (a non-blocking stream, unordered with the null stream's hipMemset).
    if (s == ARK_OK) s = m->hip(hipMemsetAsync(m->Wb, 0, NBF * sizeof(__bf16), m->stream), "mlp init");
    if (s == ARK_OK) {
        hipLaunchKernelGGL(k_refresh, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->P, m->Wb);
        s = m->hip(hipGetLastError(), "refresh");
    }
    if (s == ARK_OK) s = m->hip(hipStreamSynchronize(m->stream), "sync");
    for (int i = 0; i < NI; ++i) m->nm.in_inv[i] = 1.f;
    for (int o = 0; o < NO; ++o) m->nm.out_std[o] = m->nm.out_inv[o] = 1.f;
    if (s != ARK_OK) {
        ark_mlp_destroy(m);
        return s;
    }
    *out = m;
    return ARK_OK;
}

void ark_mlp_destroy(ark_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : {(void*)m->P, (void*)m->Wb, (void*)m->G, (void*)m->partial, (void*)m->X, (void*)m->T})
        if (p) (void)hipFree(p);
    if (m->done) (void)hipEventDestroy(m->done);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char* ark_mlp_last_error(const ark_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

ark_status ark_mlp_set_params(ark_mlp* m, const float* params) {
    if (!m || !params) return ARK_ERR_INVALID;
    ark_status s = m->hip(hipMemcpyAsync(m->P, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream),
                          "set params");
    if (s == ARK_OK) s = m->hip(hipMemsetAsync(m->M, 0, 2 * NPARAM * sizeof(float), m->stream), "reset moments");
    m->adam_t = 0;
    if (s == ARK_OK) {
        hipLaunchKernelGGL(k_refresh, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->P, m->Wb);
        s = m->hip(hipGetLastError(), "refresh");
    }
    return s == ARK_OK ? m->hip(hipStreamSynchronize(m->stream), "sync") : s;
}

ark_status ark_mlp_get_params(ark_mlp* m, float* params) {
    if (!m || !params) return ARK_ERR_INVALID;
    ark_status s = m->hip(hipMemcpyAsync(params, m->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream),
                          "get params");
    return s == ARK_OK ? m->hip(hipStreamSynchronize(m->stream), "sync") : s;
}

ark_status ark_mlp_set_normalization(ark_mlp* m, const float* in_mean, const float* in_std, const float* out_mean,
                                     const float* out_std) {
    if (!m || !in_mean || !in_std || !out_mean || !out_std) return ARK_ERR_INVALID;
    for (int i = 0; i < NI; ++i) {
        if (!(in_std[i] > 0.f)) return m->fail(ARK_ERR_INVALID, "normalization: std must be positive");
        m->nm.in_mean[i] = in_mean[i];
        m->nm.in_inv[i] = 1.f / in_std[i];
    }
    for (int o = 0; o < NO; ++o) {
        if (!(out_std[o] > 0.f)) return m->fail(ARK_ERR_INVALID, "normalization: std must be positive");
        m->nm.out_mean[o] = out_mean[o];
        m->nm.out_std[o] = out_std[o];
        m->nm.out_inv[o] = 1.f / out_std[o];
    }
    return ARK_OK;
}

ark_status ark_mlp_set_data(ark_mlp* m, const float* X, const float* T, long long n) {
    if (!m || !X || !T || n < 1) return ARK_ERR_INVALID;
    ark_status s = m->hip(hipStreamSynchronize(m->stream), "sync");
    for (float* p : {m->X, m->T})
        if (p) (void)hipFree(p);
    m->X = m->T = nullptr;
    if (s == ARK_OK) s = m->hip(hipMalloc(&m->X, n * NI * sizeof(float)), "hipMalloc data");
    if (s == ARK_OK) s = m->hip(hipMalloc(&m->T, n * NO * sizeof(float)), "hipMalloc data");
    if (s == ARK_OK)
        s = m->hip(hipMemcpyAsync(m->X, X, n * NI * sizeof(float), hipMemcpyHostToDevice, m->stream), "data");
    if (s == ARK_OK)
        s = m->hip(hipMemcpyAsync(m->T, T, n * NO * sizeof(float), hipMemcpyHostToDevice, m->stream), "data");
    if (s == ARK_OK) s = m->hip(hipStreamSynchronize(m->stream), "sync");
    m->ndata = s == ARK_OK ? n : 0;
    return s;
}

ark_status ark_mlp_gradient(ark_mlp* m, const float* X, const float* T, long long n, double* grad, double* loss) {
    if (!m || !X || !T || n < 1 || !grad) return ARK_ERR_INVALID;
    float *dX = nullptr, *dT = nullptr;
    ark_status s = m->hip(hipMalloc(&dX, n * NI * sizeof(float)), "hipMalloc");
    if (s == ARK_OK) s = m->hip(hipMalloc(&dT, n * NO * sizeof(float)), "hipMalloc");
    if (s == ARK_OK)
        s = m->hip(hipMemcpyAsync(dX, X, n * NI * sizeof(float), hipMemcpyHostToDevice, m->stream), "copy");
    if (s == ARK_OK)
        s = m->hip(hipMemcpyAsync(dT, T, n * NO * sizeof(float), hipMemcpyHostToDevice, m->stream), "copy");
    if (s == ARK_OK) s = run_grad(m, dX, dT, n, n, 0, 0, 0, 0.f);
    std::vector<double> h(NPARAM + 1);
    if (s == ARK_OK) s = m->hip(hipMemcpyAsync(h.data(), m->G, h.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream), "copy");
    if (s == ARK_OK) s = m->hip(hipStreamSynchronize(m->stream), "sync");
    if (dX) (void)hipFree(dX);
    if (dT) (void)hipFree(dT);
    if (s == ARK_OK) {
        std::memcpy(grad, h.data(), NPARAM * sizeof(double));
        if (loss) *loss = h[NPARAM] / (double)(NO * n);
    }
    return s;
}

ark_status ark_mlp_train(ark_mlp* m, int nsteps, long long batch, float lr, double* last_loss) {
    if (!m || nsteps < 0 || batch < 1 || !(lr > 0.f)) return ARK_ERR_INVALID;
    if (!m->X) return m->fail(ARK_ERR_STATE, "train: no data (ark_mlp_set_data)");
    ark_status s = ARK_OK;
    for (int k = 0; k < nsteps && s == ARK_OK; ++k)
        s = run_grad(m, m->X, m->T, m->ndata, batch, 1, (unsigned)m->adam_t, 1, lr);
    if (s == ARK_OK && last_loss) {
        double l = 0.0;
        s = m->hip(hipMemcpyAsync(&l, m->G + NPARAM, sizeof(double), hipMemcpyDeviceToHost, m->stream), "copy");
        if (s == ARK_OK) s = m->hip(hipStreamSynchronize(m->stream), "sync");
        *last_loss = l / (double)(NO * batch);
    }
    return s;
}

ark_status ark_mlp_predict(ark_mlp* m, const float* X, float* Y, long long n) {
    if (!m || !X || !Y || n < 1) return ARK_ERR_INVALID;
    float *dX = nullptr, *dY = nullptr;
    ark_status s = m->hip(hipMalloc(&dX, n * NI * sizeof(float)), "hipMalloc");
    if (s == ARK_OK) s = m->hip(hipMalloc(&dY, n * NO * sizeof(float)), "hipMalloc");
    if (s == ARK_OK)
        s = m->hip(hipMemcpyAsync(dX, X, n * NI * sizeof(float), hipMemcpyHostToDevice, m->stream), "copy");
    if (s == ARK_OK) {
        hipLaunchKernelGGL(k_mlp_infer<false>, dim3(infer_blocks(n)), dim3(256), 0, m->stream, dX, GridIn{}, n, m->P,
                           m->Wb, m->nm, dY);
        s = m->hip(hipGetLastError(), "mlp forward");
    }
    if (s == ARK_OK)
        s = m->hip(hipMemcpyAsync(Y, dY, n * NO * sizeof(float), hipMemcpyDeviceToHost, m->stream), "copy");
    if (s == ARK_OK) s = m->hip(hipStreamSynchronize(m->stream), "sync");
    if (dX) (void)hipFree(dX);
    if (dY) (void)hipFree(dY);
    return s;
}

}  // extern "C"

// The closure's stresses on the solver's grid (its interior cells, x fastest,
// 6 per cell into Y), on the solver's stream, ordered after everything
// already enqueued on the network's own stream.
ark_status ark_mlp_closure_device(ark_mlp* m, const arkd::Geo& g, const double* u1, const double* u2, const double* u3,
                                  int stride, const double inv[3], float delta, float* Y, hipStream_t st) {
    ark_status s = m->hip(hipEventRecord(m->done, m->stream), "event");
    if (s == ARK_OK) s = m->hip(hipStreamWaitEvent(st, m->done, 0), "wait");
    if (s != ARK_OK) return s;
    GridIn gi;
    gi.g = g;
    gi.u[0] = u1;
    gi.u[1] = u2;
    gi.u[2] = u3;
    gi.cs = stride;
    for (int d = 0; d < 3; ++d) gi.inv[d] = inv[d];
    gi.delta = delta;
    const long long n = (long long)g.N1 * g.N2 * g.N3;
    hipLaunchKernelGGL(k_mlp_infer<true>, dim3(infer_blocks(n)), dim3(256), 0, st, (const float*)nullptr, gi, n, m->P,
                       m->Wb, m->nm, Y);
    return m->hip(hipGetLastError(), "mlp closure");
}

// ark_diagnostics_noalloc: the diagnostics of ark_diagnostics (WriteEnergy,
// source/main.cpp:2851-2876, plus the physical kinetic energy, the mass, the
// largest velocity component and a non-finite count) over five cell arrays
// that the caller holds in device memory in the reference's Arr3d layout,
// with no context and no allocation: for callers that cannot call hipMalloc
// (work under HIP graph capture, tight serving loops).
//
// The per-block partial sums live in `scratch`, a file-scope device array of
// fixed size.  Host code never names the array in a launch or copy: each
// device's address of it comes from hipGetSymbolAddress, taken once per
// device with that device current, and kept in g_scratch.  Every call on a
// device shares that device's array: calls that overlap on one GPU race.
// Nothing in it needs clearing between calls (every partial a call reads is
// written by that call's first kernel).
#include <hip/hip_runtime.h>

#include <mutex>

#include "ark/ark.h"
#include "ark_kernels.hpp"

using namespace arkd;

namespace {

// The same partial count as a context's diagnostics, so the fixed-order
// reduction gives the same sums bit for bit.
constexpr int kBlocks = 1024;
constexpr int kScratchDoubles = 5 * kBlocks;   // 40 KiB per device

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
double* g_scratch[kMaxDevices] = {};

// This device's address of `scratch`, taken on the first call with the
// device current.
hipError_t scratch_of(int dev, double** out) {
    std::lock_guard<std::mutex> hold(g_lock);
    if (!g_scratch[dev]) {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e == hipSuccess && prev != dev) e = hipSetDevice(dev);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != dev) {
            const hipError_t r = hipSetDevice(prev);
            if (e == hipSuccess) e = r;
        }
        if (e != hipSuccess) return e;
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

extern "C" ark_status ark_diagnostics_noalloc(const ark_params* p, const double* const cells[5], double* out,
                                              void* stream) {
    if (!p || !cells || !out) return ARK_ERR_INVALID;
    for (int q = 0; q < 5; ++q)
        if (!cells[q]) return ARK_ERR_INVALID;
    if (p->N1 < 1 || p->N2 < 1 || p->N3 < 1 || p->device < 0 || p->device >= kMaxDevices) return ARK_ERR_INVALID;
    // The kernels index a cell with 32-bit integers: the arrays, (N+2) per
    // axis, must hold fewer than 2^31 elements.
    const long long e1 = p->N1 + 2LL, e2 = p->N2 + 2LL, e3 = p->N3 + 2LL;
    if (e1 * e2 * e3 >= (1LL << 31)) return ARK_ERR_UNSUPPORTED;
    double* part = nullptr;
    if (scratch_of(p->device, &part) != hipSuccess) return ARK_ERR_HIP;
    Geo g;
    g.N1 = p->N1; g.N2 = p->N2; g.N3 = p->N3;
    g.px = (int)e1;
    g.sxy = (int)(e1 * e2);
    Fields F{};
    F.a[F_ronCon] = (double*)cells[0];
    F.a[F_u1nCon] = (double*)cells[1];
    F.a[F_u2nCon] = (double*)cells[2];
    F.a[F_u3nCon] = (double*)cells[3];
    F.a[F_tnCon] = (double*)cells[4];
    Consts K{};
    K.dvc = p->dx1 * p->dx2 * p->dx3;   // source/main.cpp:455-460
    const hipError_t e = launch_diag((hipStream_t)stream, g, F, K, part, kBlocks, out);
    return e == hipSuccess ? ARK_OK : ARK_ERR_HIP;
}

// C-ABI of the MI355X CABARET/ARK step (include/ark/ark.h): context, device
// memory in the padded SoA layout, per-phase entry points and the step
// drivers.  Everything runs on the context's one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "ark/ark.h"
#include "ark_internal.hpp"
#include "ark_kernels.hpp"

using namespace arkd;

namespace {


bool is_face_field(int f) { return f >= F_ro1; }
// The eighteen persistent face fields are double-buffered (Phase2 reads the
// old faces and writes the new ones).
bool is_face_state(int f) { return f >= F_ro1 && f <= F_p3; }
// The level-n cell fields are double-buffered too: the fused step reads
// them around each cell while it writes level n+1.
bool is_double_buffered(int f) { return is_face_state(f) || (f >= F_ronCon && f <= F_tnCon); }

const char* kFieldNames[F_NUM] = {
    "roCon", "u1Con", "u2Con", "u3Con", "tCon", "ronCon", "u1nCon", "u2nCon", "u3nCon", "tnCon",
    "f1", "f2", "f3", "condition",
    "ro1", "t1", "u11", "u21", "u31", "p1", "ro2", "t2", "u12", "u22", "u32", "p2",
    "ro3", "t3", "u13", "u23", "u33", "p3",
    "sigm11", "sigm21", "sigm31", "sigm12", "sigm22", "sigm32", "sigm13", "sigm23", "sigm33"};

}  // namespace

const int kGroupFields[4][5] = {
    {F_ronCon, F_u1nCon, F_u2nCon, F_u3nCon, F_tnCon},
    {F_u11, F_u21, F_u31, F_p1, F_t1},
    {F_u12, F_u22, F_u32, F_p2, F_t2},
    {F_u13, F_u23, F_u33, F_p3, F_t3},
};

void ark_set_group_buffers(ark_ctx* c, int g, char* cur, char* alt) {
    c->gbase[g][c->gpar[g]] = cur;
    c->gbase[g][c->gpar[g] ^ 1] = alt;
    for (int v = 0; v < 5; ++v) {
        c->cur[kGroupFields[g][v]] = (double*)(cur + v * c->field_bytes);
        c->alt[kGroupFields[g][v]] = (double*)(alt + v * c->field_bytes);
    }
}

ark_status ark_per_field(ark_ctx* c) {
    ark_status s = ark_cells_layout(c, false);
    return s == ARK_OK ? ark_fused_layout(c, false) : s;
}

void ark_swap_group(ark_ctx* c, int g) {
    for (int v = 0; v < 5; ++v) std::swap(c->cur[kGroupFields[g][v]], c->alt[kGroupFields[g][v]]);
    c->gpar[g] ^= 1;
}

static int group_of(int f, int* slot) {
    for (int g = 0; g < 4; ++g)
        for (int v = 0; v < 5; ++v)
            if (kGroupFields[g][v] == f) {
                *slot = v;
                return g;
            }
    return -1;
}

// ---------------------------------------------------------------------------
ark_status ark_ctx::fail(ark_status s, const std::string& msg) {
    err = msg;
    return s;
}

ark_status ark_ctx::hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return ARK_OK;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return ARK_ERR_HIP;
}

ark_status ark_ctx::launch(const char* name, hipError_t (*fn)(void*), void* arg) {
    if (!profiling) return hip(fn(arg), name);
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return fail(ARK_ERR_HIP, "hipEventCreate");
    ark_status s = hip(hipEventRecord(a, stream), "hipEventRecord");
    const hipError_t e = fn(arg);
    if (s == ARK_OK) s = hip(hipEventRecord(b, stream), "hipEventRecord");
    prof_events.push_back({name, a, b});
    const ark_status sk = hip(e, name);
    return sk != ARK_OK ? sk : s;
}

Fields ark_ctx::fields() const {
    Fields F;
    for (int f = 0; f < F_NUM; ++f) F.a[f] = cur[f];
    return F;
}

// ---------------------------------------------------------------------------
extern "C" {

ark_status ark_create(const ark_params* p, ark_ctx** out) {
    if (!p || !out) return ARK_ERR_INVALID;
    *out = nullptr;
    ark_ctx* c = new ark_ctx();
    c->prm = *p;
    if (p->N1 < 1 || p->N2 < 1 || p->N3 < 1) {
        c->err = "N1, N2, N3 must be >= 1";
        *out = c;
        return ARK_ERR_INVALID;
    }
    const bool all_periodic = p->periodic[0] && p->periodic[1] && p->periodic[2];
    if (!all_periodic && p->precision == ARK_PREC_MIXED) {
        c->err = "mixed precision needs a triply periodic box (non-periodic axes run the FP64 phase path)";
        *out = c;
        return ARK_ERR_UNSUPPORTED;
    }
    if (p->scheme != ARK_SCHEME_REFERENCE && p->scheme != ARK_SCHEME_CORRECTED) {
        c->err = "scheme must be ARK_SCHEME_REFERENCE or ARK_SCHEME_CORRECTED";
        *out = c;
        return ARK_ERR_INVALID;
    }
    if (p->precision != ARK_PREC_FP64 && p->precision != ARK_PREC_MIXED) {
        c->err = "precision must be ARK_PREC_FP64 or ARK_PREC_MIXED";
        *out = c;
        return ARK_ERR_INVALID;
    }
    if (!(p->dx1 > 0 && p->dx2 > 0 && p->dx3 > 0 && p->sound > 0 && p->CFL > 0 && p->VIS >= 0)) {
        c->err = "dx, sound and CFL must be > 0, VIS >= 0";
        *out = c;
        return ARK_ERR_INVALID;
    }
    Geo& g = c->geo;
    g.N1 = p->N1; g.N2 = p->N2; g.N3 = p->N3;
    // Row pitch: N1+3 elements rounded up to a multiple of 16 for FP64, of 4
    // for mixed precision.  The mixed step slows down with longer rows
    // (measured at 256^3: 1.151 ms at pitch 260, 1.205 ms at 272; at 384^3:
    // 3.59 ms at 388, 4.02 ms at 392); the FP64 step is within 1% either way.
    const long long pq = p->precision == ARK_PREC_MIXED ? 4 : 16;
    long long px = ((long long)p->N1 + 3 + pq - 1) / pq * pq;
    long long elems = px * (p->N2 + 3) * (long long)(p->N3 + 3);
    // The fused kernels address a field with a 32-bit byte offset from its
    // base (8 bytes per element): at most 2^29 elements, about 800^3 cells.
    if ((elems + kOriginPad) * (long long)sizeof(double) >= (1ll << 32)) {
        c->err = "grid too large: a field's byte offsets must fit in 32 bits (at most 2^29 elements per field, "
                 "about 800^3 cells)";
        *out = c;
        return ARK_ERR_UNSUPPORTED;
    }
    // Mixed precision: a face record is 16 bytes, also at a 32-bit offset.
    if (p->precision == ARK_PREC_MIXED && elems * 16 >= (1ll << 32)) {
        c->err = "grid too large for mixed precision: a face record's byte offset must fit in 32 bits (at most "
                 "2^28 elements per field, about 640^3 cells)";
        *out = c;
        return ARK_ERR_UNSUPPORTED;
    }
    g.px = (int)px;
    g.sxy = (int)(px * (p->N2 + 3));
    c->elems = elems + kOriginPad + 16;

    Consts& K = c->K;
    K.dx[0] = p->dx1; K.dx[1] = p->dx2; K.dx[2] = p->dx3;
    K.ds1 = p->dx2 * p->dx3;            // source/main.cpp:455-460
    K.ds2 = p->dx1 * p->dx3;
    K.ds3 = p->dx1 * p->dx2;
    K.dvc = p->dx1 * p->dx2 * p->dx3;
    K.VIS = p->VIS; K.sound = p->sound; K.ro0 = p->ro0; K.t0 = p->t0;
    // IEEE double division on the host: the correctly rounded reciprocals
    for (int d = 0; d < 3; ++d) K.rdx[d] = 1.0 / K.dx[d];
    K.rc2 = 1.0 / (K.sound * K.sound);
    K.rrs = 1.0 / (K.ro0 * K.sound);
    K.rdvc = 1.0 / K.dvc;

    ark_status s = c->hip(hipSetDevice(p->device), "hipSetDevice");
    if (s == ARK_OK) s = c->hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    c->field_bytes = (c->elems * sizeof(double) + 4095) / 4096 * 4096;
    for (int g = 0; g < 4 && s == ARK_OK; ++g)
        for (int b = 0; b < 2 && s == ARK_OK; ++b) {
            char* mem = nullptr;
            size_t nb = 5 * c->field_bytes;
            s = c->hip(hipMalloc(&mem, nb), "hipMalloc field group");
            if (s != ARK_OK) break;
            c->allocs.push_back(mem);
            c->bytes += nb;
            s = c->hip(hipMemsetAsync(mem, 0, nb, c->stream), "hipMemset field group");
            c->gbase[g][b] = mem + kOriginPad * sizeof(double);
        }
    for (int f = 0; f < F_NUM && s == ARK_OK; ++f) {
        int slot = 0, g = group_of(f, &slot);
        if (g >= 0) {
            c->cur[f] = (double*)(c->gbase[g][0] + slot * c->field_bytes);
            c->alt[f] = (double*)(c->gbase[g][1] + slot * c->field_bytes);
            continue;
        }
        int nbuf = is_double_buffered(f) ? 2 : 1;
        for (int b = 0; b < nbuf && s == ARK_OK; ++b) {
            double* mem = nullptr;
            s = c->hip(hipMalloc(&mem, c->elems * sizeof(double)), "hipMalloc field");
            if (s != ARK_OK) break;
            c->allocs.push_back(mem);
            c->bytes += c->elems * sizeof(double);
            s = c->hip(hipMemsetAsync(mem, 0, c->elems * sizeof(double), c->stream), "hipMemset field");
            if (b == 0) c->cur[f] = mem + kOriginPad;
            else c->alt[f] = mem + kOriginPad;
        }
    }
    if (s == ARK_OK) {
        s = c->hip(hipMalloc(&c->S, sizeof(Scalars)), "hipMalloc scalars");
        if (s == ARK_OK) c->allocs.push_back(c->S);
    }
    if (s == ARK_OK) {
        double* bases[9];
        for (int q = 0; q < 9; ++q) bases[q] = c->cur[F_sigm11 + q] - kOriginPad;
        s = c->hip(hipMalloc(&c->sigma_ptrs, sizeof bases), "hipMalloc sigma table");
        if (s == ARK_OK) {
            c->allocs.push_back(c->sigma_ptrs);
            s = c->hip(hipMemcpy(c->sigma_ptrs, bases, sizeof bases, hipMemcpyHostToDevice), "sigma table");
        }
    }
    if (s == ARK_OK && p->precision == ARK_PREC_MIXED) {
        s = c->hip(hipMalloc(&c->staging, c->elems * sizeof(double)), "hipMalloc staging");
        if (s == ARK_OK) {
            c->allocs.push_back(c->staging);
            c->bytes += c->elems * sizeof(double);
            c->staging += kOriginPad;
        }
    }
    if (s == ARK_OK) {
        c->diag_blocks = 1024;
        s = c->hip(hipMalloc(&c->diag_part, (5 * c->diag_blocks + 8) * sizeof(double)), "hipMalloc diag");
        if (s == ARK_OK) c->allocs.push_back(c->diag_part);
    }
    if (s == ARK_OK) {
        const long long N1 = p->N1, N2 = p->N2, N3 = p->N3;
        const long long plane = std::max(N2 * N3, std::max(N1 * N3, N1 * N2));
        s = c->hip(hipMalloc(&c->plane_buf, plane * sizeof(double)), "hipMalloc plane buffer");
        if (s == ARK_OK) {
            c->allocs.push_back(c->plane_buf);
            c->bytes += plane * sizeof(double);
        }
    }
    if (s == ARK_OK) {
        // dt starts from min(1e8, dtv), the constant part of the reference's
        // running min3d(dt, dtu, dtv) (source/main.cpp:377, 393-405).
        double dtv;
        if (p->VIS > pow(10, -16)) {
            double dtv1 = p->CFL * p->dx1 * p->dx1 / (2. * p->VIS);
            double dtv2 = p->CFL * p->dx2 * p->dx2 / (2. * p->VIS);
            double dtv3 = p->CFL * p->dx3 * p->dx3 / (2. * p->VIS);
            dtv = dtv1;
            dtv = dtv > dtv2 ? dtv2 : dtv;
            dtv = dtv > dtv3 ? dtv3 : dtv;
        } else {
            dtv = pow(10, 16);
        }
        double d0 = pow(10, 8);
        d0 = d0 > dtv ? dtv : d0;
        Scalars h{};
        h.dt = 0.0;
        h.time = 0.0;
        h.nstep = 0;
        memcpy(&h.dt_init, &d0, 8);
        h.dt_bits = h.dt_init;
        h.nf[0] = 0;
        h.nf[1] = LLONG_MAX;
        s = c->hip(hipMemcpyAsync(c->S, &h, sizeof(h), hipMemcpyHostToDevice, c->stream), "upload scalars");
        if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    }
    if (s == ARK_OK) s = c->hip(hipHostMalloc((void**)&c->nf_host, 2 * sizeof(long long)), "hipHostMalloc");
    if (s == ARK_OK) s = c->hip(hipEventCreateWithFlags(&c->nf_event, hipEventDisableTiming), "hipEventCreate");
    *out = c;
    return s;
}

void ark_destroy(ark_ctx* c) {
    if (!c) return;
    // best effort: nothing can be reported from here
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    ark_closure_hp_free(c);
    for (auto& e : c->prof_events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (void* m : c->allocs) (void)hipFree(m);
    if (c->dense) (void)hipFree(c->dense);
    if (c->nf_host) (void)hipHostFree(c->nf_host);
    if (c->nf_event) (void)hipEventDestroy(c->nf_event);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* ark_last_error(const ark_ctx* c) { return c ? c->err.c_str() : "null context"; }

static ark_status check_extents(ark_ctx* c, int field, int e1, int e2, int e3) {
    if (field < 0 || field >= F_NUM) return c->fail(ARK_ERR_INVALID, "field id out of range");
    int add = is_face_field(field) ? 3 : 2;
    if (e1 != c->geo.N1 + add || e2 != c->geo.N2 + add || e3 != c->geo.N3 + add) {
        char buf[160];
        snprintf(buf, sizeof buf, "%s: extents %dx%dx%d, expected %dx%dx%d", kFieldNames[field], e1, e2, e3,
                 c->geo.N1 + add, c->geo.N2 + add, c->geo.N3 + add);
        return c->fail(ARK_ERR_INVALID, buf);
    }
    return ARK_OK;
}

// A host array (dense, e1 x e2 x e3, the reference's Arr3d layout) to or from
// a field (padded rows and planes): one contiguous transfer through a dense
// device buffer, and the repack on the device.  Strided copies from host
// memory, one per plane, ran at 0.2-0.3 GB/s.
static ark_status copy_planes(ark_ctx* c, double* dev, const double* host_src, double* host_dst, int e1, int e2,
                              int e3) {
    const size_t n = (size_t)e1 * e2 * e3;
    if (c->dense_n < n) {
        if (c->dense) {
            ark_status s = c->hip(hipStreamSynchronize(c->stream), "sync");
            if (s == ARK_OK) s = c->hip(hipFree(c->dense), "hipFree dense buffer");
            if (s != ARK_OK) return s;
            c->bytes -= c->dense_n * sizeof(double);
            c->dense = nullptr;
            c->dense_n = 0;
        }
        ark_status s = c->hip(hipMalloc(&c->dense, n * sizeof(double)), "hipMalloc dense buffer");
        if (s != ARK_OK) return s;
        c->dense_n = n;
        c->bytes += n * sizeof(double);
    }
    const long long px = c->geo.px, sxy = c->geo.sxy, r = e1, pl = (long long)e1 * e2;
    ark_status s;
    if (host_src) {
        s = c->hip(hipMemcpyAsync(c->dense, host_src, n * sizeof(double), hipMemcpyHostToDevice, c->stream),
                   "hipMemcpyAsync upload");
        if (s == ARK_OK)
            s = c->hip(launch_copy_rows(c->stream, c->dense, r, pl, dev, px, sxy, e1, e2, e3), "copy rows");
    } else {
        s = c->hip(launch_copy_rows(c->stream, dev, px, sxy, c->dense, r, pl, e1, e2, e3), "copy rows");
        if (s == ARK_OK)
            s = c->hip(hipMemcpyAsync(host_dst, c->dense, n * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                       "hipMemcpyAsync download");
    }
    return s == ARK_OK ? c->hip(hipStreamSynchronize(c->stream), "sync after copy") : s;
}

static ark_status drop_pending(ark_ctx* c);

ark_status ark_upload(ark_ctx* c, int field, const double* host, int e1, int e2, int e3) {
    if (!c || !host) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    ark_status s = check_extents(c, field, e1, e2, e3);
    if (s != ARK_OK) return s;
    // TimeStepSize reads these: a dt left pending by fused steps no longer holds
    if (field == F_u1Con || field == F_u2Con || field == F_u3Con) s = drop_pending(c);
    if (s != ARK_OK) return s;
    if (field == F_condition) {
        long long n = (long long)e1 * e2 * e3;
        c->has_solid = false;
        for (long long q = 0; q < n; ++q)
            if (host[q] != 0.0) { c->has_solid = true; break; }
    }
    if (c->prm.precision == ARK_PREC_MIXED && is_face_state(field)) {
        // FP32 storage: face densities are derived from p and not stored
        if (field == F_ro1 || field == F_ro2 || field == F_ro3) return ARK_OK;
        s = copy_planes(c, c->staging, host, nullptr, e1, e2, e3);
        if (s == ARK_OK) s = ark_mixed_face_from_host_layout(c, field, c->staging);
        if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
        return s;
    }
    if (field >= F_sigm11) c->sigma_dirty = true;
    if (field == F_ro1 || field == F_ro2 || field == F_ro3 || field == F_p1 || field == F_p2 || field == F_p3) {
        // face densities left derived by fused steps are stored first, from
        // the pressures they belong to, so that the arrays not uploaded keep
        // the reference's values
        s = ark_refresh_rho(c);
        if (s != ARK_OK) return s;
        c->rho_consistent = false;
    }
    s = copy_planes(c, c->cur[field], host, nullptr, e1, e2, e3);
    // the second buffer from the dense copy still on the device
    if (s == ARK_OK && c->alt[field])
        s = c->hip(launch_copy_rows(c->stream, c->dense, e1, (long long)e1 * e2, c->alt[field], c->geo.px,
                                    c->geo.sxy, e1, e2, e3),
                   "copy rows");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync after copy");
    return s;
}

ark_status ark_download(ark_ctx* c, int field, double* host, int e1, int e2, int e3) {
    if (!c || !host) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    ark_status s = check_extents(c, field, e1, e2, e3);
    if (s != ARK_OK) return s;
    if (c->prm.precision == ARK_PREC_MIXED && is_face_state(field)) {
        s = ark_mixed_face_to_host_layout(c, field, c->staging);
        if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync before download");
        if (s != ARK_OK) return s;
        return copy_planes(c, c->staging, nullptr, host, e1, e2, e3);
    }
    if (field == F_ro1 || field == F_ro2 || field == F_ro3) s = ark_refresh_rho(c);
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync before download");
    if (s != ARK_OK) return s;
    return copy_planes(c, c->cur[field], nullptr, host, e1, e2, e3);
}

ark_status ark_set_dt(ark_ctx* c, double dt) {
    if (!c) return ARK_ERR_INVALID;
    c->host_dt = dt;
    ark_status s = c->hip(hipMemcpyAsync(&c->S->dt, &c->host_dt, sizeof(double), hipMemcpyHostToDevice, c->stream),
                          "set dt");
    return s == ARK_OK ? c->hip(hipStreamSynchronize(c->stream), "sync") : s;
}

ark_status ark_get_clock(ark_ctx* c, ark_clock* out) {
    if (!c || !out) return ARK_ERR_INVALID;
    Scalars sc;
    ark_status s = c->hip(hipMemcpyAsync(&sc, c->S, sizeof sc, hipMemcpyDeviceToHost, c->stream), "scalars");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s != ARK_OK) return s;
    out->time = sc.time;
    out->nstep = sc.nstep;
    out->dt = sc.dt;
    out->dt_next_known = c->dt_pending ? 1 : 0;
    out->dt_next = 0.0;
    if (c->dt_pending) memcpy(&out->dt_next, &sc.dt_bits, sizeof(double));
    return ARK_OK;
}

ark_status ark_restart(ark_ctx* c, const ark_clock* k) {
    if (!c || !k) return ARK_ERR_INVALID;
    if (k->dt_next_known && !(k->dt_next > 0.0 && std::isfinite(k->dt_next)))
        return c->fail(ARK_ERR_INVALID, "restart: dt_next must be positive and finite");
    Scalars hs;
    ark_status s = c->hip(hipMemcpyAsync(&hs, c->S, sizeof hs, hipMemcpyDeviceToHost, c->stream), "scalars");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s != ARK_OK) return s;
    hs.time = k->time;
    hs.nstep = k->nstep;
    hs.dt = k->dt;
    if (k->dt_next_known) memcpy(&hs.dt_bits, &k->dt_next, sizeof(double));
    else hs.dt_bits = hs.dt_init;
    hs.nf[0] = 0;
    hs.nf[1] = LLONG_MAX;
    s = c->hip(hipMemcpyAsync(c->S, &hs, sizeof hs, hipMemcpyHostToDevice, c->stream), "scalars");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s != ARK_OK) return s;
    c->host_dt = k->dt;
    c->dt_pending = k->dt_next_known != 0;
    c->nf_seen = false;
    c->nf_posted = false;
    const bool periodic = c->prm.periodic[0] && c->prm.periodic[1] && c->prm.periodic[2];
    if (c->prm.precision == ARK_PREC_FP64 && periodic && !c->has_solid) {
        c->rho_consistent = true;
        c->rho_stored = false;   // derived from p when next needed
    }
    return ARK_OK;
}

struct ClockArgs { ark_ctx* c; int what; };
hipError_t run_clock(void* p) {
    ClockArgs* a = (ClockArgs*)p;
    return launch_clock(a->c->stream, a->c->S, a->what);
}

ark_status ark_next_step(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ClockArgs a{c, 0};
    return c->launch("clock", run_clock, &a);
}

ark_status ark_half_time(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ClockArgs a{c, 1};
    return c->launch("clock", run_clock, &a);
}

ark_status ark_get_dt(ark_ctx* c, double* dt) {
    if (!c || !dt) return ARK_ERR_INVALID;
    ark_status s = c->hip(hipMemcpyAsync(&c->host_dt, &c->S->dt, sizeof(double), hipMemcpyDeviceToHost, c->stream),
                          "get dt");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s == ARK_OK) *dt = c->host_dt;
    return s;
}

// ---- phases ----------------------------------------------------------------
namespace {

struct CflArgs { ark_ctx* c; };
hipError_t run_cfl(void* p) {
    ark_ctx* c = ((CflArgs*)p)->c;
    const ark_params& P = c->prm;
    return launch_cfl(c->stream, c->geo, c->fields(), P.CFL * P.dx1, P.CFL * P.dx2, P.CFL * P.dx3, P.sound, c->S);
}
struct FinArgs { ark_ctx* c; int advance; };
hipError_t run_cfl_fin(void* p) {
    FinArgs* a = (FinArgs*)p;
    return launch_cfl_finalize(a->c->stream, a->c->S, a->advance);
}
hipError_t run_phase1(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    return launch_phase1(c->stream, c->geo, c->fields(), c->K, c->S);
}
// Axes of a context: periodic ones exchange with themselves; a non-periodic
// axis of one rank has MPI_PROC_NULL neighbours (its planes get zeros) and
// runs the reference's physical-boundary blocks: zero-gradient ends on X1 and
// X3 (source/main.cpp:803-845, 1633-1698), inflow and outflow in the X1 sweep.
int periodic_mask(const ark_ctx* c) {
    return (c->prm.periodic[0] ? 1 : 0) | (c->prm.periodic[1] ? 2 : 0) | (c->prm.periodic[2] ? 4 : 0);
}

// The entry of array p in the plan (added on first use).  The exchanges
// below name at most 22 distinct arrays; a full plan reuses its last entry
// and is refused by exchange().
constexpr int kPlanMax = (int)(sizeof(ExchangePlan::p) / sizeof(double*));
static int plan_entry(ExchangePlan& plan, double* p, int is_face) {
    for (int q = 0; q < plan.n; ++q)
        if (plan.p[q] == p) return q;
    if (plan.n == kPlanMax) return kPlanMax - 1;
    const int q = plan.n++;
    plan.p[q] = p;
    plan.is_face[q] = is_face;
    plan.copy_mask[q] = plan.zero_mask[q] = plan.grad_mask[q] = 0;
    return q;
}

// The exchange of the given cell arrays (every axis) and face arrays (pairs
// of field and axis: along their own axis), then, on non-periodic X1/X3, the
// zero-gradient ends of the arrays the reference sets there (x_ends on X1;
// the cell arrays and z_faces on X3), as one launch.  A zero-gradient end
// replaces the zero fill of the same planes.
hipError_t exchange(ark_ctx* c, const int* cells, int ncells, const int* faces, int nfaces, const int* x_ends,
                    int nx_ends, const int* z_faces, int nz_faces) {
    const int per = periodic_mask(c), non = 7 & ~per;
    ExchangePlan plan{};
    for (int q = 0; q < ncells; ++q) {
        const int e = plan_entry(plan, c->cur[cells[q]], 0);
        plan.copy_mask[e] |= per;
        plan.zero_mask[e] |= non;
    }
    for (int q = 0; q < nfaces; ++q) {
        const int d = faces[2 * q + 1];
        const int e = plan_entry(plan, c->cur[faces[2 * q]], 1);
        plan.copy_mask[e] |= per & (1 << d);
        plan.zero_mask[e] |= non & (1 << d);
    }
    for (int q = 0; q < nx_ends; ++q) plan.grad_mask[plan_entry(plan, c->cur[x_ends[q]], 0)] |= non & 1;
    for (int q = 0; q < ncells; ++q) plan.grad_mask[plan_entry(plan, c->cur[cells[q]], 0)] |= non & 4;
    for (int q = 0; q < nz_faces; ++q) plan.grad_mask[plan_entry(plan, c->cur[z_faces[q]], 1)] |= non & 4;
    if (ncells + nfaces > kPlanMax) return hipErrorInvalidValue;
    for (int q = 0; q < plan.n; ++q) plan.zero_mask[q] &= ~plan.grad_mask[q];
    return launch_exchange(c->stream, c->geo, plan);
}

hipError_t run_stress_ghost(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    const int cells[3] = {F_u1Con, F_u2Con, F_u3Con};
    return exchange(c, cells, 3, nullptr, 0, cells, 3, nullptr, 0);
}
hipError_t run_sigma_zero(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    return launch_zero(c->stream, c->sigma_ptrs, 9, (long long)c->elems);
}
hipError_t run_stress_faces(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    return launch_stress_faces(c->stream, c->geo, c->fields(), c->K);
}
hipError_t run_forces(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    return launch_forces(c->stream, c->geo, c->fields(), c->K);
}
hipError_t run_use_forces(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    return launch_use_forces(c->stream, c->geo, c->fields(), c->K, c->S);
}
hipError_t run_phase2_ghost(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    const int cells[7] = {F_u1nCon, F_u2nCon, F_u3nCon, F_roCon, F_ronCon, F_tCon, F_tnCon};
    int faces[30];
    for (int d = 0; d < 3; ++d) {
        const int f[5] = {face_u(d, 0), face_u(d, 1), face_u(d, 2), face_ro(d), face_p(d)};
        for (int q = 0; q < 5; ++q) {
            faces[2 * (5 * d + q)] = f[q];
            faces[2 * (5 * d + q) + 1] = d;
        }
    }
    // X1 ends: the level-n velocities, density and t (source/main.cpp:1633-1658);
    // X3 ends: all seven cell arrays and the z-faces (1660-1698)
    const int x_ends[5] = {F_u1nCon, F_u2nCon, F_u3nCon, F_ronCon, F_tnCon};
    const int z_faces[5] = {face_u(2, 0), face_u(2, 1), face_u(2, 2), face_ro(2), face_p(2)};
    return exchange(c, cells, 7, faces, 15, x_ends, 5, z_faces, 5);
}
hipError_t run_phase2_xyz(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    NextFaces nx[3];
    for (int d = 0; d < 3; ++d) {
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        nx[d].p[0] = c->alt[face_u(d, d)];
        nx[d].p[1] = c->alt[face_u(d, ta)];
        nx[d].p[2] = c->alt[face_u(d, tb)];
        nx[d].p[3] = c->alt[face_p(d)];
        nx[d].p[4] = c->alt[face_ro(d)];
        nx[d].p[5] = c->alt[face_t(d)];
    }
    return launch_phase2_xyz(c->stream, c->prm.scheme, c->geo, c->fields(), nx, c->K, c->S, !c->prm.periodic[0]);
}
hipError_t run_forces_use_forces(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    return launch_forces_use_forces(c->stream, c->geo, c->fields(), c->K, c->S);
}
hipError_t run_phase1_use_forces(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    return launch_phase1_use_forces(c->stream, c->geo, c->fields(), c->K, c->S);
}

}  // namespace

// The phase-by-phase entry points are FP64 and take solid cells and
// non-periodic axes; the fused and mixed-precision steps need a fluid,
// triply periodic box.
static ark_status require_fluid(ark_ctx* c) {
    if (c->prm.precision == ARK_PREC_MIXED)
        return c->fail(ARK_ERR_UNSUPPORTED,
                       "the phase-by-phase entry points are FP64; a mixed-precision context advances with ark_step "
                       "or ark_advance");
    return ARK_OK;
}

ark_status ark_cfl_finalize(ark_ctx* c, int advance) {
    FinArgs f{c, advance};
    return c->launch("cfl_finalize", run_cfl_fin, &f);
}

ark_status ark_cfl_scan(ark_ctx* c) {
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    CflArgs a{c};
    return c->launch("cfl", run_cfl, &a);
}

ark_status ark_cfl_step(ark_ctx* c, int advance) {
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    CflArgs a{c};
    ark_status s = c->launch("cfl", run_cfl, &a);
    return s == ARK_OK ? ark_cfl_finalize(c, advance) : s;
}

ark_status ark_begin_step(ark_ctx* c) {
    if (c->dt_pending) {
        c->dt_pending = false;
        return ark_cfl_finalize(c, 1);
    }
    return ark_cfl_step(c, 1);
}

// A phase call after fused steps: the pending dt no longer describes Con.
static ark_status drop_pending(ark_ctx* c) {
    if (!c->dt_pending) return ARK_OK;
    c->dt_pending = false;
    return ark_cfl_finalize(c, 2);
}

hipError_t run_rho(void* p) {
    ark_ctx* c = (ark_ctx*)p;
    return launch_rho_faces(c->stream, c->geo, c->fields(), c->K);
}

ark_status ark_refresh_rho(ark_ctx* c) {
    if (c->rho_stored) return ARK_OK;
    ark_status s = ark_per_field(c);
    if (s == ARK_OK) s = c->launch("rho_faces", run_rho, c);
    if (s == ARK_OK) c->rho_stored = true;
    return s;
}

// One step after TimeStepSize, in the reference's order (source/main.cpp:
// 85-101): Phase1, StressTensor, UseForces, Phase2, Phase1, UseForces.  The
// force and the first UseForces, and the second Phase1 and UseForces, each
// run as one launch: those phases use only the same cell's values of the
// phase before.
static ark_status stress_tensor(ark_ctx* c, bool use_forces);
static ark_status phase1(ark_ctx* c, bool use_forces);
ark_status ark_phase_body(ark_ctx* c) {
    ark_status s = ark_per_field(c);
    if (s == ARK_OK) s = phase1(c, false);
    if (s == ARK_OK) s = stress_tensor(c, true);
    if (s == ARK_OK) s = ark_phase2(c);
    if (s == ARK_OK) s = phase1(c, true);
    return s;
}

ark_status ark_time_step_size(ark_ctx* c, double* dt_out) {
    if (!c) return ARK_ERR_INVALID;
    ark_status s = c->has_solid ? require_fluid(c) : ARK_OK;
    if (s == ARK_OK) {
        if (c->dt_pending) {
            // after fused steps Con is not stored; their MIN is TimeStepSize's
            // result, and it stays pending: the next step takes the same dt,
            // as a second TimeStepSize call in the reference would
            s = ark_cfl_finalize(c, 3);
        } else {
            s = ark_cfl_step(c, 0);
        }
    }
    if (s == ARK_OK && dt_out) s = ark_get_dt(c, dt_out);
    return s;
}

// Phase1; with use_forces, also the UseForces that follows the second
// Phase1 of a step, in the same launch.
static ark_status phase1(ark_ctx* c, bool use_forces) {
    ark_status s = require_fluid(c);
    if (s == ARK_OK) s = drop_pending(c);
    if (s == ARK_OK) s = ark_refresh_rho(c);
    if (s != ARK_OK) return s;
    return use_forces ? c->launch("phase1_use_forces", run_phase1_use_forces, c) : c->launch("phase1", run_phase1, c);
}

ark_status ark_phase1(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    return phase1(c, false);
}

// StressTensor; with use_forces, also the UseForces that follows it in a
// step, in the force's launch.
static ark_status stress_tensor(ark_ctx* c, bool use_forces) {
    ark_status s = require_fluid(c);
    if (s == ARK_OK) s = drop_pending(c);
    if (s == ARK_OK && c->sigma_dirty) {
        s = c->launch("sigma_zero", run_sigma_zero, c);
        if (s == ARK_OK) c->sigma_dirty = false;
    }
    if (s == ARK_OK) s = c->launch("ghost_stress", run_stress_ghost, c);
    if (s == ARK_OK) s = c->launch("stress_faces", run_stress_faces, c);
    if (s == ARK_OK)
        s = use_forces ? c->launch("forces_use_forces", run_forces_use_forces, c) : c->launch("forces", run_forces, c);
    return s;
}

ark_status ark_stress_tensor(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    return stress_tensor(c, false);
}

ark_status ark_use_forces(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    ark_status s = require_fluid(c);
    if (s == ARK_OK) s = drop_pending(c);
    return s == ARK_OK ? c->launch("use_forces", run_use_forces, c) : s;
}

ark_status ark_phase2(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    ark_status s = require_fluid(c);
    if (s == ARK_OK) s = drop_pending(c);
    if (s == ARK_OK) s = c->launch("ghost_phase2", run_phase2_ghost, c);
    if (s == ARK_OK) s = c->launch("phase2", run_phase2_xyz, c);
    if (s == ARK_OK) {
        for (int g = 1; g < 4; ++g) ark_swap_group(c, g);
        std::swap(c->cur[F_ro1], c->alt[F_ro1]);
        std::swap(c->cur[F_ro2], c->alt[F_ro2]);
        std::swap(c->cur[F_ro3], c->alt[F_ro3]);
        c->rho_consistent = !c->has_solid;
        c->rho_stored = true;
    }
    return s;
}

// npairs x two phase-path steps from one captured graph.  A step is only
// kernel launches on c->stream, its dt in device memory and no memset or
// host synchronisation; the host side of a step swaps the face buffers, so
// two steps bring every buffer back to the role it had, and the host state
// after the capture (which ran that host side once) is the state after any
// number of pairs.  The graph lives for this call only.
ark_status ark_phase_pairs(ark_ctx* c, int npairs) {
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ark_status s = c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (s != ARK_OK) return s;
    ark_status body = ARK_OK;
    for (int k = 0; k < 2 && body == ARK_OK; ++k) {
        body = ark_refresh_rho(c);
        if (body == ARK_OK) body = ark_begin_step(c);
        if (body == ARK_OK) body = ark_phase_body(c);
    }
    s = c->hip(hipStreamEndCapture(c->stream, &graph), "end capture");
    if (body != ARK_OK) s = body;
    if (s == ARK_OK) s = c->hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    for (int k = 0; k < npairs && s == ARK_OK; ++k) s = c->hip(hipGraphLaunch(exec, c->stream), "graph launch");
    const ark_status sync = c->hip(hipStreamSynchronize(c->stream), "sync after graph");
    if (s == ARK_OK) s = sync;
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    return s;
}

ark_status ark_step_phases(ark_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return ARK_ERR_INVALID;
    ark_status s = require_fluid(c);
    if (s == ARK_OK) s = ark_nf_check(c, 0);
    for (int n = 0; n < nsteps && s == ARK_OK; ++n) {
        s = ark_begin_step(c);
        if (s == ARK_OK) s = ark_phase_body(c);
    }
    return s == ARK_OK ? ark_nf_post(c) : s;
}

// A stepping call first looks at the non-finite count its predecessor posted
// (without waiting for it), and posts its own behind its launches.
ark_status ark_step(ark_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return ARK_ERR_INVALID;
    ark_status s = ark_nf_check(c, 0);
    if (s != ARK_OK) return s;
    if (c->prm.precision == ARK_PREC_MIXED) {
        if (c->has_solid)
            return c->fail(ARK_ERR_UNSUPPORTED, "solid cells run on the FP64 phase path, not in mixed precision");
        s = ark_mixed_steps(c, nsteps);
    } else {
        s = require_fluid(c);
        if (s == ARK_OK) s = ark_step_fused(c, nsteps);
    }
    return s == ARK_OK ? ark_nf_post(c) : s;
}

ark_status ark_diagnostics(ark_ctx* c, ark_diag* out) {
    if (!c || !out) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    double* part = c->diag_part;
    double* res = c->diag_part + 5 * c->diag_blocks;
    ark_status s = c->hip(launch_diag(c->stream, c->geo, c->fields(), c->K, part, c->diag_blocks, res), "diag");
    double h[5];
    Scalars sc;
    if (s == ARK_OK) s = c->hip(hipMemcpyAsync(h, res, sizeof h, hipMemcpyDeviceToHost, c->stream), "diag copy");
    if (s == ARK_OK) s = c->hip(hipMemcpyAsync(&sc, c->S, sizeof sc, hipMemcpyDeviceToHost, c->stream), "scalars");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s != ARK_OK) return s;
    out->ke_reference = h[0];
    out->ke_physical = h[1];
    out->mass = h[2];
    out->max_abs_u = h[3];
    out->nonfinite = (int64_t)h[4];
    out->dt = sc.dt;
    out->time = sc.time;
    out->nstep = sc.nstep;
    if (out->nonfinite > 0) return c->fail(ARK_ERR_NONFINITE, "non-finite values in the cell fields");
    return ARK_OK;
}

// The reference's own problem, InitializeData (source/main.cpp:280-372) on
// one rank (coords 0) of Input's set-up (118-195): node coordinates by running
// sums from -dx, u1 = u10 above y = 1 (cells, x- and z-faces at the cell
// centre's y, y-faces at the node's y) and 0 below, u2 = u3 = p = 0, rho = ro0,
// t = t0 in the cells (the faces' t stays 0: the reference never sets it),
// solid cells where x <= 5 and y <= 1 and the two ghost rows in y (walls),
// Con := nCon; every other element 0.
ark_status ark_init_channel(ark_ctx* c, double u10) {
    if (!c) return ARK_ERR_INVALID;
    if (c->prm.precision == ARK_PREC_MIXED)
        return c->fail(ARK_ERR_UNSUPPORTED, "the channel has solid cells and physical boundaries: FP64 only");
    const int n1 = c->geo.N1 + 1, n2 = c->geo.N2 + 1, n3 = c->geo.N3 + 1;
    const double dx1 = c->prm.dx1, dx2 = c->prm.dx2;
    std::vector<double> x1(n1 + 2), x2(n2 + 2);
    x1[0] = 0. + ((n1 - 1) * dx1 * 0) - dx1;
    x2[0] = 0. + ((n2 - 1) * dx2 * 0) - dx2;
    for (int i = 0; i < n1 + 1; ++i) x1[i + 1] = x1[i] + dx1;
    for (int j = 0; j < n2 + 1; ++j) x2[j + 1] = x2[j] + dx2;
    const size_t e1c = n1 + 1, e2c = n2 + 1, e3c = n3 + 1, e1f = n1 + 2, e2f = n2 + 2, e3f = n3 + 2;
    auto cidx = [&](int i, int j, int k) { return (size_t)k * e1c * e2c + (size_t)j * e1c + i; };
    auto fidx = [&](int i, int j, int k) { return (size_t)k * e1f * e2f + (size_t)j * e1f + i; };
    std::vector<std::vector<double>> h(F_NUM);
    for (int f = 0; f < F_NUM; ++f) h[f].assign(f >= F_ro1 ? e1f * e2f * e3f : e1c * e2c * e3c, 0.0);
    const double ro0_g = c->prm.ro0, t0 = c->prm.t0;
    for (int i = 0; i <= n1; ++i)
        for (int j = 0; j <= n2; ++j)
            for (int k = 0; k <= n3; ++k) {
                const double y = 0.5 * (x2[j] + x2[j + 1]);
                const size_t q = cidx(i, j, k);
                h[F_u1nCon][q] = y <= 1 ? 0. : u10;
                h[F_ronCon][q] = ro0_g;
                h[F_tnCon][q] = t0;
            }
    for (int i = 1; i <= n1; ++i)
        for (int j = 1; j <= n2; ++j)
            for (int k = 1; k <= n3; ++k) {
                const double y = x2[j], y_c = 0.5 * (x2[j] + x2[j + 1]);
                const size_t q = fidx(i, j, k);
                h[F_u11][q] = y_c <= 1 ? 0. : u10;
                h[F_u12][q] = y <= 1 ? 0. : u10;
                h[F_u13][q] = y_c <= 1 ? 0. : u10;
                h[F_ro1][q] = h[F_ro2][q] = h[F_ro3][q] = ro0_g;
            }
    for (int i = 0; i < n1 + 1; ++i)
        for (int j = 0; j < n2 + 1; ++j)
            for (int k = 0; k < n3 + 1; ++k) {
                const double x = 0.5 * (x1[i] + x1[i + 1]), y = 0.5 * (x2[j] + x2[j + 1]);
                h[F_condition][cidx(i, j, k)] = (x <= 5 && y <= 1) ? 1 : 0;
            }
    for (int i = 0; i < n1 + 1; ++i)
        for (int k = 0; k < n3 + 1; ++k) h[F_condition][cidx(i, 0, k)] = h[F_condition][cidx(i, n2, k)] = 1;
    h[F_roCon] = h[F_ronCon];
    h[F_u1Con] = h[F_u1nCon];
    h[F_u2Con] = h[F_u2nCon];
    h[F_u3Con] = h[F_u3nCon];
    h[F_tCon] = h[F_tnCon];
    ark_status s = ARK_OK;
    for (int f = 0; f < F_NUM && s == ARK_OK; ++f) {
        const int e = f >= F_ro1 ? 3 : 2;
        s = ark_upload(c, f, h[f].data(), c->geo.N1 + e, c->geo.N2 + e, c->geo.N3 + e);
    }
    if (s == ARK_OK) s = ark_finish_init(c);
    c->has_solid = true;
    c->rho_consistent = false;
    return s;
}

ark_status ark_init_tgv(ark_ctx* c, double U0) {
    if (!c) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    const int N[3] = {c->geo.N1, c->geo.N2, c->geo.N3};
    const double dx[3] = {c->prm.dx1, c->prm.dx2, c->prm.dx3};
    // 18 tables of up to max(N)+3 entries
    int len = std::max(N[0], std::max(N[1], N[2])) + 3;
    std::vector<double> h(18 * (size_t)len, 0.0);
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < N[d] + 3; ++i) {
            int w = ((i - 1) % N[d] + N[d]) % N[d] + 1;   // periodic image, 1..N
            double xc = (w - 0.5) * dx[d];
            double xf = (w - 1.0) * dx[d];
            h[(0 * 3 + d) * (size_t)len + i] = sin(xc);
            h[(1 * 3 + d) * (size_t)len + i] = cos(xc);
            h[(2 * 3 + d) * (size_t)len + i] = cos(2.0 * xc);
            h[(3 * 3 + d) * (size_t)len + i] = sin(xf);
            h[(4 * 3 + d) * (size_t)len + i] = cos(xf);
            h[(5 * 3 + d) * (size_t)len + i] = cos(2.0 * xf);
        }
    double* dtab = nullptr;
    ark_status s = c->hip(hipMalloc(&dtab, h.size() * sizeof(double)), "hipMalloc tables");
    if (s != ARK_OK) return s;
    s = c->hip(hipMemcpyAsync(dtab, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream),
               "upload tables");
    TgvTables T;
    for (int d = 0; d < 3; ++d) {
        T.sc[d] = dtab + (0 * 3 + d) * (size_t)len;
        T.cc[d] = dtab + (1 * 3 + d) * (size_t)len;
        T.c2c[d] = dtab + (2 * 3 + d) * (size_t)len;
        T.sf[d] = dtab + (3 * 3 + d) * (size_t)len;
        T.cf[d] = dtab + (4 * 3 + d) * (size_t)len;
        T.c2f[d] = dtab + (5 * 3 + d) * (size_t)len;
    }
    Fields F2;
    for (int f = 0; f < F_NUM; ++f) F2.a[f] = c->alt[f];
    double amp = c->prm.ro0 * U0 * U0 / 16.0;
    if (s == ARK_OK) s = c->hip(launch_init_tgv(c->stream, c->geo, c->fields(), F2, T, U0, amp, c->K), "init_tgv");
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    (void)hipFree(dtab);
    return s == ARK_OK ? ark_finish_init(c) : s;
}

// After a new initial state in the FP64 layout: FP32 face storage in mixed
// precision, state flags, TIME = 0, nStep = 0, running MIN reset.
ark_status ark_finish_init(ark_ctx* c) {
    ark_status s = ARK_OK;
    if (c->prm.precision == ARK_PREC_MIXED) {
        // FP32 face records from the FP64 face fields just written
        s = ark_mixed_faces_from_fp64(c);
        if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    }
    c->has_solid = false;
    c->rho_consistent = true;
    c->rho_stored = true;
    c->sigma_dirty = false;
    c->dt_pending = false;   // the reset below discards a MIN left by fused steps
    c->nf_seen = false;
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    c->nf_posted = false;
    if (s == ARK_OK) {
        // fresh run: TIME = 0, nStep = 0, running MIN reset
        // (ordered on the context's stream, after the kernels that set dt_init)
        Scalars hs{};
        s = c->hip(hipMemcpyAsync(&hs, c->S, sizeof hs, hipMemcpyDeviceToHost, c->stream), "scalars");
        if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
        hs.time = 0.0;
        hs.nstep = 0;
        hs.dt = 0.0;
        hs.dt_bits = hs.dt_init;
        hs.nf[0] = 0;
        hs.nf[1] = LLONG_MAX;
        if (s == ARK_OK) s = c->hip(hipMemcpyAsync(c->S, &hs, sizeof hs, hipMemcpyHostToDevice, c->stream), "scalars");
        if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "sync");
    }
    return s;
}

ark_status ark_synchronize(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ark_status s = ark_nf_post(c);
    if (s == ARK_OK) s = c->hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return s == ARK_OK ? ark_nf_check(c, 1) : s;
}

ark_status ark_nf_post(ark_ctx* c) {
    ark_status s = c->hip(hipMemcpyAsync(c->nf_host, c->S->nf, 2 * sizeof(long long), hipMemcpyDeviceToHost,
                                         c->stream), "non-finite readback");
    if (s == ARK_OK) s = c->hip(hipEventRecord(c->nf_event, c->stream), "hipEventRecord");
    if (s == ARK_OK) c->nf_posted = true;
    return s;
}

ark_status ark_nf_check(ark_ctx* c, int wait) {
    if (c->nf_posted) {
        const hipError_t e = wait ? hipEventSynchronize(c->nf_event) : hipEventQuery(c->nf_event);
        if (e == hipSuccess) {
            c->nf_posted = false;
            if (c->nf_host[0] > 0 && !c->nf_seen) {
                c->nf_seen = true;
                c->nf_count = c->nf_host[0];
                c->nf_step = c->nf_host[1];
            }
        } else if (e != hipErrorNotReady) {
            return c->hip(e, "non-finite readback");
        }
    }
    if (!c->nf_seen) return ARK_OK;
    char buf[200];
    snprintf(buf, sizeof buf,
             "non-finite velocity or density: %lld cell(s) by the time it was read, first seen in step %lld; "
             "stepping is refused until a new initial state (ark_init_*, ark_restart)",
             c->nf_count, c->nf_step);
    return c->fail(ARK_ERR_NONFINITE, buf);
}

ark_status ark_get_stream(ark_ctx* c, void** stream) {
    if (!c || !stream) return ARK_ERR_INVALID;
    *stream = (void*)c->stream;
    return ARK_OK;
}

ark_status ark_profile_enable(ark_ctx* c, int on) {
    if (!c) return ARK_ERR_INVALID;
    c->profiling = on != 0;
    return ARK_OK;
}

ark_status ark_profile_reset(ark_ctx* c) {
    if (!c) return ARK_ERR_INVALID;
    ark_status s = c->hip(hipStreamSynchronize(c->stream), "sync");
    for (auto& e : c->prof_events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->prof_events.clear();
    return s;
}

ark_status ark_profile_read(ark_ctx* c, int max_entries, char (*names)[64], int64_t* launches, double* total_ms,
                            int* n_entries) {
    if (!c || !n_entries) return ARK_ERR_INVALID;
    ark_status s = c->hip(hipStreamSynchronize(c->stream), "sync");
    if (s != ARK_OK) return s;
    std::vector<std::string> keys;
    std::vector<int64_t> cnt;
    std::vector<double> tot;
    for (auto& e : c->prof_events) {
        float ms = 0.f;
        s = c->hip(hipEventElapsedTime(&ms, e.a, e.b), "hipEventElapsedTime");
        if (s != ARK_OK) return s;
        size_t q = std::find(keys.begin(), keys.end(), e.name) - keys.begin();
        if (q == keys.size()) {
            keys.push_back(e.name);
            cnt.push_back(0);
            tot.push_back(0.0);
        }
        cnt[q] += 1;
        tot[q] += ms;
    }
    int n = (int)std::min<size_t>(keys.size(), (size_t)std::max(max_entries, 0));
    for (int q = 0; q < n; ++q) {
        if (names) {
            strncpy(names[q], keys[q].c_str(), 63);
            names[q][63] = 0;
        }
        if (launches) launches[q] = cnt[q];
        if (total_ms) total_ms[q] = tot[q];
    }
    *n_entries = (int)keys.size();
    return ARK_OK;
}

// Field planes and boundary conditions (the host-side halo and BC interface).
static bool is_cell_field(int f) { return f >= F_roCon && f <= F_condition; }

static ark_status plane_args(ark_ctx* c, int field, int axis, int q, int na, int nb) {
    if (field < 0 || field >= F_NUM || axis < 0 || axis > 2) return c->fail(ARK_ERR_INVALID, "plane: bad field or axis");
    const int Nd = axis == 0 ? c->geo.N1 : axis == 1 ? c->geo.N2 : c->geo.N3;
    const int Na = axis == 0 ? c->geo.N2 : c->geo.N1, Nb = axis == 2 ? c->geo.N2 : c->geo.N3;
    const int qmax = Nd + (is_cell_field(field) ? 1 : 2);
    if (q < 0 || q > qmax || na != Na || nb != Nb) {
        char buf[160];
        snprintf(buf, sizeof buf, "plane: index %d (0..%d) or extents %dx%d (expected %dx%d) out of range", q, qmax,
                 na, nb, Na, Nb);
        return c->fail(ARK_ERR_INVALID, buf);
    }
    if (c->prm.precision == ARK_PREC_MIXED && !is_cell_field(field))
        return c->fail(ARK_ERR_UNSUPPORTED, "plane: face fields are FP32 in mixed precision");
    return ARK_OK;
}

ark_status ark_get_plane(ark_ctx* c, int field, int axis, int q, double* host, int na, int nb) {
    if (!c || !host) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    ark_status s = plane_args(c, field, axis, q, na, nb);
    if (s == ARK_OK && (field == F_ro1 || field == F_ro2 || field == F_ro3)) s = ark_refresh_rho(c);
    if (s == ARK_OK) s = c->hip(launch_plane(c->stream, c->geo, c->cur[field], axis, q, c->plane_buf, 1), "plane");
    if (s == ARK_OK)
        s = c->hip(hipMemcpyAsync(host, c->plane_buf, (size_t)na * nb * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream), "plane to host");
    return s == ARK_OK ? c->hip(hipStreamSynchronize(c->stream), "sync") : s;
}

ark_status ark_set_plane(ark_ctx* c, int field, int axis, int q, const double* host, int na, int nb) {
    if (!c || !host) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    ark_status s = plane_args(c, field, axis, q, na, nb);
    if (s != ARK_OK) return s;
    if (field >= F_sigm11) c->sigma_dirty = true;
    if (field == F_ro1 || field == F_ro2 || field == F_ro3 || field == F_p1 || field == F_p2 || field == F_p3) {
        s = ark_refresh_rho(c);   // as in ark_upload
        if (s != ARK_OK) return s;
        c->rho_consistent = false;
    }
    if (field == F_condition) {
        for (long long k = 0; k < (long long)na * nb; ++k)
            if (host[k] != 0.0) { c->has_solid = true; break; }
    }
    s = c->hip(hipMemcpyAsync(c->plane_buf, host, (size_t)na * nb * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream), "plane to device");
    if (s == ARK_OK) s = c->hip(launch_plane(c->stream, c->geo, c->cur[field], axis, q, c->plane_buf, 0), "plane");
    return s == ARK_OK ? c->hip(hipStreamSynchronize(c->stream), "sync") : s;
}

ark_status ark_boundary_conditions(ark_ctx* c, int axis, const int fields[5], int periodic, double t0) {
    if (!c || !fields || axis < 0 || axis > 2) return ARK_ERR_INVALID;
    ark_status lay = ark_per_field(c);   // every field in the per-field layout
    if (lay != ARK_OK) return lay;
    BcFields F;
    for (int q = 0; q < 5; ++q) {
        if (fields[q] < 0 || !is_cell_field(fields[q]))
            return c->fail(ARK_ERR_INVALID, "boundary conditions: fields must be cell fields");
        F.f[q] = c->cur[fields[q]];
    }
    return c->hip(launch_boundary(c->stream, c->geo, F, axis, periodic, t0), "boundary conditions");
}

ark_status ark_device_bytes(const ark_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return ARK_ERR_INVALID;
    *bytes = (int64_t)c->bytes;
    return ARK_OK;
}

}  // extern "C"

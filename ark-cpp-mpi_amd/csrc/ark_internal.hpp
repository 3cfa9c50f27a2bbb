// Internal definition of the context behind the C-ABI handle.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ark/ark.h"
#include "ark_device.hpp"

struct ark_prof_event {
    const char* name;
    hipEvent_t a, b;
};

// Element (0,0,0) of every field sits 15 elements into its allocation, so
// interior element i = 1 of every row starts a 128-byte line.
constexpr int kOriginPad = 15;

// The 16 forced wavevectors m (k = 2 pi m / L) of ark_set_forcing: every
// 1 <= |m|^2 <= 4 with its first non-zero component positive (the conjugates
// complete them), in the order the mode table holds them.  Component d of
// mode n, as a compile-time constant wherever n and d are.
__host__ __device__ constexpr int forced_mode(int n, int d) {
    constexpr int m[16][3] = {{0, 0, 1},  {0, 0, 2},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1},  {0, 2, 0},
                              {1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0},  {1, 0, 1},
                              {1, 1, -1},  {1, 1, 0},  {1, 1, 1},  {2, 0, 0}};
    return m[n][d];
}

struct ark_ctx {
    ark_params prm{};
    arkd::Geo geo{};
    arkd::Consts K{};
    hipStream_t stream = nullptr;
    size_t elems = 0;                  // elements per field allocation
    size_t field_bytes = 0;            // bytes between consecutive fields of a group
    double* cur[arkd::F_NUM] = {};     // element (0,0,0) of each field
    double* alt[arkd::F_NUM] = {};     // second buffer of the double-buffered fields
    // The persistent fields live in four groups of five consecutive fields
    // (level-n cells rho,u1,u2,u3,t; per orientation the faces u1,u2,u3,p,t),
    // each group twice.  gbase[g][b] is element (0,0,0) of the group's first
    // field in buffer b; gpar[g] is the buffer that is current.
    char* gbase[4][2] = {};
    int gpar[4] = {0, 0, 0, 0};
    arkd::Scalars* S = nullptr;        // device scalars
    double* diag_part = nullptr;
    double* staging = nullptr;         // one FP64 field (mixed precision conversions)
    double* plane_buf = nullptr;       // one interior plane (ark_get_plane / ark_set_plane)
    double* dense = nullptr;           // one host array in its dense layout (upload / download)
    size_t dense_n = 0;
    // body force of the mixed-precision step (ark_set_forcing): acceleration
    // components at force + m*force_stride in the field index map, FP32
    float* force = nullptr;
    float* force_table = nullptr;      // this step's 16 forced modes
    float* force_trig = nullptr;       // per-axis cos/sin of the forced wavenumbers at the cell centres
    long long force_stride = 0;
    unsigned long long force_seed = 0;
    double force_eps = 0.0;
    double force_tau = 0.0;           // forcing correlation time (0: new amplitudes every step)
    // learned closure (ark_set_closure): network, grid scale, its stresses (6 per cell)
    struct ark_mlp* closure = nullptr;
    double closure_delta = 0.0;
    float* les_Y = nullptr;
    struct ark_closure_hp* closure_hp = nullptr;   // scale-selective application (ark_set_closure_cutoff)
    double* average = nullptr;         // AverageValues' three accumulators (cell fields, common index map)
    // chunk tables of the mixed-precision step, one per single-box tile shape
    // (int4 per chunk), and the shape in use (-1: not chosen yet)
    void* mixed_units[2] = {nullptr, nullptr};
    int mixed_nunits[2] = {0, 0};
    int mixed_shape = -1;
    void* fused_units = nullptr;       // chunk table of the FP64 fused step
    int fused_nunits = 0;
    int diag_blocks = 0;
    double host_dt = 0.0;
    bool has_solid = false;
    bool rho_consistent = false;       // face densities 1..N+1 equal ro0 + p/c^2 bit for bit
    bool rho_stored = true;            // the stored face-density arrays are current
    bool dt_pending = false;           // S->dt_bits holds the next step's dt from a fused step
    bool cells_rec = false;            // mixed: the level-n cells are in the stepping layout (ark_cells_layout)
    bool fused_rec = false;            // FP64: the four groups are in the record layout (ark_fused_layout)
    char* group_spare = nullptr;       // FP64: a spare group buffer for the layout conversions (lazily allocated)
    bool sigma_dirty = false;          // an upload left non-zero values in a stress array
    double** sigma_ptrs = nullptr;     // device array: allocation bases of the nine stress arrays
    // Non-finite detection: the stepping calls copy Scalars::nf into nf_host
    // (pinned) behind their launches and record nf_event; the next stepping
    // call reads it if the copy has landed, a synchronising call always.
    // nf_seen is sticky until a new initial state.
    long long* nf_host = nullptr;
    hipEvent_t nf_event = nullptr;
    bool nf_posted = false;
    bool nf_seen = false;
    long long nf_count = 0, nf_step = 0;
    bool profiling = false;
    std::vector<ark_prof_event> prof_events;
    std::vector<void*> allocs;
    size_t bytes = 0;
    std::string err;

    ark_status fail(ark_status s, const std::string& msg);
    ark_status hip(hipError_t e, const char* what);
    // Runs fn(arg) (which enqueues kernels on `stream`), bracketed by events when profiling.
    ark_status launch(const char* name, hipError_t (*fn)(void*), void* arg);
    arkd::Fields fields() const;
};

// Step building blocks shared by the drivers (ark_api.hip; internal, not in ark.h).
extern "C" {
ark_status ark_cfl_scan(ark_ctx* c);                    // running MIN of the CFL bound over Con
ark_status ark_finish_init(ark_ctx* c);                 // after an initial state: FP32 faces, flags, TIME = 0
ark_status ark_turb_force_step(ark_ctx* c, hipStream_t st);   // this step's forcing field (mixed precision)
ark_status ark_closure_step(ark_ctx* c, hipStream_t st);      // the closure's body force (after the forcing)
// the learned closure's stresses on the solver's grid (ark_mlp.hip)
ark_status ark_mlp_closure_device(struct ark_mlp* m, const arkd::Geo& g, const double* u1, const double* u2,
                                  const double* u3, int stride, const double inv[3], float delta, float* Y,
                                  hipStream_t st);
ark_status ark_cfl_step(ark_ctx* c, int advance);       // TimeStepSize on Con, then dt := MIN
ark_status ark_cfl_finalize(ark_ctx* c, int mode);      // see k_cfl_finalize
ark_status ark_begin_step(ark_ctx* c);                  // dt of a new step (pending MIN or Con), TIME, nStep
ark_status ark_refresh_rho(ark_ctx* c);                 // store face densities if they are derived
ark_status ark_phase_body(ark_ctx* c);                  // Phase1 .. UseForces of one step, dt already set
ark_status ark_phase_pairs(ark_ctx* c, int npairs);     // 2*npairs phase-path steps from one captured graph
ark_status ark_nf_post(ark_ctx* c);                     // copy the non-finite count behind the queued work
ark_status ark_nf_check(ark_ctx* c, int wait);          // ARK_ERR_NONFINITE once a posted count is non-zero
}

// The closure's spectral projection state (ark_turb.hip); freed with the context.
void ark_closure_hp_free(ark_ctx* c);

// Forcing of a batch of contexts in two launches per step (ark_turb.hip).
struct ark_force_batch;
ark_status ark_force_batch_build(ark_ctx* const* boxes, int nbox, ark_force_batch** f);   // replaces *f
ark_status ark_force_batch_enqueue(const ark_force_batch* f, hipStream_t st);
void ark_force_batch_free(ark_force_batch* f);

// Field slots of the four groups, in group order.
extern const int kGroupFields[4][5];
// Exchange the current and the other buffer of group g.
void ark_swap_group(ark_ctx* c, int g);
// Make cur the current and alt the other buffer of group g (its field pointers too).
void ark_set_group_buffers(ark_ctx* c, int g, char* cur, char* alt);
// Every field in the per-field layout, whatever a stepping call left them in:
// call before reading or writing any field per field.
ark_status ark_per_field(ark_ctx* c);

// Mixed precision (ark_mixed.hip): fused steps, and the conversion of one
// face field between its FP32 record storage and an FP64 field in the common
// layout.
ark_status ark_mixed_steps(ark_ctx* c, int nsteps);
// Mixed precision steps the level-n cells as 32-byte records of (rho, u1, u2,
// u3) with t in its own field; everything else reads and writes them per
// field.  Converts the current buffer of the cell group into the other one
// and swaps the two when the layout differs; a no-op for FP64 contexts.
ark_status ark_cells_layout(ark_ctx* c, bool records);
ark_status ark_mixed_face_to_host_layout(ark_ctx* c, int f, double* staging);
ark_status ark_mixed_face_from_host_layout(ark_ctx* c, int f, const double* staging);
// After an initial state in FP64 face fields (both buffers): the FP32 records.
ark_status ark_mixed_faces_from_fp64(ark_ctx* c);
// The reference's vorticity from the face records (ark_write_paraview).
ark_status ark_mixed_vorticity(ark_ctx* c, double* o1, double* o2, double* o3);

// Chunk tables of the z-marching fused steps (ark_mixed.hip): (tile column
// x fastest then y, first plane, end plane) per chunk, for tiles of TX x TY
// owned columns and W workgroups resident at once; and their upload into a
// device allocation owned by c.
std::vector<int4> ark_chunk_table(const arkd::Geo& g, int TX, int TY, int W);
ark_status ark_upload_chunks(ark_ctx* c, const std::vector<int4>& u, void** dst, int* n);

// The fused step (ark_fused.hip).
ark_status ark_step_fused(ark_ctx* c, int nsteps);
// The FP64 fused step's record layout of all four groups (fields 0-3 of an
// index as one 32-byte record, field 4 per field), or back; both buffers of
// every group are converted index for index, through group_spare.
// ARK_ERR_UNSUPPORTED, with nothing changed, when group_spare cannot be
// allocated on the way to records.
ark_status ark_fused_layout(ark_ctx* c, bool records);
bool ark_fused_eligible(const ark_ctx* c);

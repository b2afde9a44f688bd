// Host-side context of the particle hot path, shared by the translation
// units that implement the C ABI of include/ipm_hip.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <utility>
#include <vector>
#include "ipm_device.hpp"

using namespace ipm;

struct ipm_window;
struct ipm_recomb;
// Frees the differentiable-window buffers of a context (ipm_window.hip).
int window_release(struct ipm_ctx *c);
// Frees the recombination work buffers of a context (ipm_recombine.hip).
int recomb_release(struct ipm_ctx *c);

inline thread_local char g_err[512] = "";

inline int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail(IPM_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                           \
    } while (0)

struct ipm_ctx {
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    int nz = 0, nr = 0, ns = 0, ng = 0;
    double dx = 0, dt = 0;
    DevState h{};
    DevState *d_state = nullptr;
    DevCounters *d_ctr = nullptr;
    double *buf[2][IPM_NS_MAX] = {};
    int cur[IPM_NS_MAX] = {};
    bool species_set[IPM_NS_MAX] = {};
    double chifac[IPM_NS_MAX] = {};
    double *d_chifac = nullptr;
    double *raw_all = nullptr, *rho_one = nullptr, *chi_one = nullptr;
    double *rho_tot = nullptr, *chi_tot = nullptr;
    double *inv_vol = nullptr;
    int *cell_material = nullptr, *hedge = nullptr, *vedge = nullptr, *mat_type = nullptr;
    int2 *row_vac = nullptr;               // nr + 2 vacuum runs of cell rows (DevState::row_vac)
    double *metal_charge = nullptr, *sigma = nullptr;
    double *Ez = nullptr, *Er = nullptr;
    unsigned long long *mcc_counts = nullptr;
    std::vector<void *> tab_mem[IPM_NS_MAX * IPM_NG_MAX];
    unsigned long long *hist = nullptr, *offs = nullptr, *cursor = nullptr, *partial = nullptr,
                       *partial2 = nullptr;
    double *qbuf = nullptr;                // q_cap records of 8 doubles
    double *p_stage = nullptr;             // per collision block: staged products, meta, counts
    unsigned *p_meta = nullptr, *p_cnt = nullptr;
    unsigned *q_tile_count = nullptr;      // q_ntiles
    int *tanc = nullptr;                   // q_ntiles deposit-window anchors (z, r)
    long long q_cap = 0, q_ntiles = 0;
    double *mom = nullptr;                 // moment sums, 5 * nz * nr (ipm_moments)
    double *scratch = nullptr;             // this device's scratch array (ipm_moments_noalloc)
    double *phi = nullptr;                 // potential (ipm_set_potential), nz * nr
    double *eps_cell = nullptr;            // cell_epsilon, (nz-1) * (nr-1) (ipm_set_solver_geometry)
    int *ptype = nullptr;                  // point_type, nz * nr
    double *sol = nullptr;                 // ipm_solver_system: 6 nz*nr outputs, 2 nz*nr inputs, nz*nr ints
    unsigned long long *rezone = nullptr;  // split/merge tallies (ipm_split_merge), 3
    hipGraphExec_t graph = nullptr;
    hipGraphExec_t sort_graph = nullptr;   // cell sort of every species (ipm_step_graph's interval sorts)
    int sort_every = 0, since_sort = 0;    // ipm_set_sort_interval
    int block = 256, blocks_per_cu = 0;
    int pk_block = 256, pk_winr = 24;      // fused particle pass: threads per block, deposit window rows
    int collide_bpc = 6;                   // collision-pass blocks per compute unit
    int sort_block = 256;                  // cell sort: threads per block (8 slots each)
    // optional HIP-event timing of the kernels ipm_step enqueues
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pairs[3];   // (start, stop) per kernel class
    size_t ev_used = 0;
    bool geometry_set = false, fields_set = false;
    // differentiable-window state (ipm_window_* entry points), or null
    ipm_window *win = nullptr;
    // recombination work buffers (ipm_recombine), or null
    ipm_recomb *recomb = nullptr;
};

inline int sync_state(ipm_ctx *c)
{
    HIPCHK(hipMemcpyAsync(c->d_state, &c->h, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IPM_OK;
}


inline int read_counters(ipm_ctx *c, DevCounters *out)
{
    HIPCHK(hipMemcpyAsync(out, c->d_ctr, sizeof(DevCounters), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (out->overflow & 1)
        return fail(IPM_ECAPACITY, "The size of particle bundle reaches the maximum (MCC products dropped).");
    if (out->overflow & 2) return fail(IPM_EINVAL, "MCC product with an invalid species index.");
    if (out->overflow & 4) return fail(IPM_ECAPACITY, "collision queue overflow (selections were not collided)");
    if (out->overflow & 8) return fail(IPM_EINVAL, "dead-slot count disagrees with the particle arrays");
    return IPM_OK;
}

// Species s's current buffer set becomes set `cur` of its two; f[] and alt[]
// of the host copy of DevState follow.  The device copy is updated by
// sync_state or, after a cell sort, by the sort's own swap kernel.
// Elements per field array of a species of the given capacity: a multiple of
// 8192 (64 KiB-aligned arrays: of the strides tried, 256 B-, 4 KiB- and
// 64 KiB-aligned, the particle pass streams its fifteen arrays fastest at
// this one) with at least one slot past the capacity, which the cell sort's
// scatter writes for sorted positions that hold no particle (so its stores
// need no branch).
inline size_t field_stride(long long capacity)
{
    return ((size_t)capacity + 8192) & ~(size_t)8191;
}

inline void set_buffers(ipm_ctx *c, int s, int cur)
{
    const size_t cap = field_stride(c->h.cap[s]);
    c->cur[s] = cur;
    for (int f = 0; f < 8; ++f) {
        c->h.f[s][f] = c->buf[cur][s] + (size_t)f * cap;
        c->h.alt[s][f] = c->buf[1 - cur][s] + (size_t)f * cap;
    }
}

// Captured graphs bake in grid sizes derived from the capacities and the
// launch shape: any change to those drops them (recaptured on next use).
inline void drop_graphs(ipm_ctx *c)
{
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->sort_graph) (void)hipGraphExecDestroy(c->sort_graph);
    c->graph = nullptr;
    c->sort_graph = nullptr;
}

inline int check_ready(ipm_ctx *c, bool need_fields, bool need_geom)
{
    if (!c) return fail(IPM_EINVAL, "null context");
    for (int s = 0; s < c->ns; ++s)
        if (!c->species_set[s]) return fail(IPM_EINVAL, "species %d has not been set", s);
    if (need_fields && !c->fields_set) return fail(IPM_EINVAL, "fields have not been set");
    if (need_geom && !c->geometry_set) return fail(IPM_EINVAL, "geometry has not been set");
    return IPM_OK;
}

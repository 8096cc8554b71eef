// api.hip -- the C-ABI: context lifetime, field transfer, stage entry points and the
// RHS sequence of Miluphpc::rhs (miluphpc.cpp:387-562) for one subdomain on one GPU.
//
// Asynchronous launches on the context stream; one synchronisation at the end of each
// entry point, where the sticky device error word is read and mapped to mhpc_status
// (replacing the reference's exits and cudaTerminate).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "mhpc_internal.h"

#define HIPC(expr)                                                            \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            c->last_error = std::string(#expr) + ": " + hipGetErrorString(e_); \
            return MHPC_E_HIP;                                                \
        }                                                                     \
    } while (0)

extern "C" void mhpc_default_params(mhpc_params *p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->theta = 0.5;
    p->smoothing = 0.0;
    p->mac = MHPC_MAC_CLASSIC;
    p->quadrupole = 0;
    p->G = 1.0;
    p->curve = MHPC_CURVE_LEBESGUE;
    p->eos = MHPC_EOS_IDEAL_GAS;
    p->gamma = 5.0 / 3.0;
    p->K = 0.0;
    p->cs_const = 0.0;
    p->alpha = 1.0;
    p->beta = 2.0;
    p->n_target = 64;
    p->band_lo = 50;
    p->band_hi = 80;
    p->max_counts = 7;
    p->max_nb = 180;
    p->sml_avg = 1;
    p->precision = MHPC_PREC_FP64;
    p->cs_before_p = 1;
    p->bh_accept = MHPC_BH_VOTE;
    p->smoothing_kernel = MHPC_KERNEL_CUBIC_SPLINE;
}

extern "C" const char *mhpc_status_string(int s) {
    switch (s) {
        case MHPC_OK: return "ok";
        case MHPC_E_CAP_NODES: return "node capacity exceeded";
        case MHPC_E_CAP_NEIGHBOURS: return "neighbour capacity exceeded (or no neighbours after h-iteration)";
        case MHPC_E_DEPTH: return "particles do not separate within 22 bisections (tree depth)";
        case MHPC_E_RHO_NONPOSITIVE: return "non-positive density";
        case MHPC_E_STACK: return "traversal stack overflow";
        case MHPC_E_ARG: return "bad argument";
        case MHPC_E_HIP: return "HIP runtime error";
        case MHPC_E_STATE: return "call order";
        case MHPC_E_SCRATCH: return "more elements than the no-allocation scratch holds";
        default: return "unknown";
    }
}

static int check_params(const mhpc_params *p) {
    if (!p) return 0;
    if (!(p->theta > 0.0) || p->smoothing < 0.0) return 0;
    if (p->curve != MHPC_CURVE_LEBESGUE && p->curve != MHPC_CURVE_HILBERT) return 0;
    if (p->mac != MHPC_MAC_CLASSIC && p->mac != MHPC_MAC_SQRT_THETA) return 0;
    if (p->max_nb < 1 || p->max_counts < 1 || p->max_counts > 64) return 0;
    if (p->n_target < 1 || p->band_lo > p->band_hi) return 0;
    if (p->precision != MHPC_PREC_FP64 && p->precision != MHPC_PREC_FP32_KAHAN) return 0;
    if (!(p->sml_max >= 0.0)) return 0;
    if (p->bh_accept < MHPC_BH_BOX || p->bh_accept > MHPC_BH_PER_PARTICLE) return 0;
    if (p->bh_accept == MHPC_BH_PER_PARTICLE && (p->mac != MHPC_MAC_CLASSIC || p->quadrupole)) return 0;
    if (p->smoothing_kernel < MHPC_KERNEL_SPIKY || p->smoothing_kernel > MHPC_KERNEL_WENDLAND_C6) return 0;
    return 1;
}

template <typename T>
static hipError_t dalloc(T **p, int64_t count) {
    if (count <= 0) count = 1;
    return hipMalloc((void **)p, sizeof(T) * (size_t)count);
}

static void upload_htable(mhpc_ctx *c) {
    std::vector<double> t(MHPC_HTABLE_MAX + 1);
    t[0] = 0.0;
    for (int k = 1; k <= MHPC_HTABLE_MAX; k++)
        t[k] = 0.5 * (1.0 + pow((double)c->p.n_target / (double)k, 1.0 / 3.0));
    (void)hipMemcpy(c->htable, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice);
}

static int ensure_nbr(mhpc_ctx *c) {
    if (c->nbr && c->nbr_rows >= c->p.max_nb) return MHPC_OK;
    if (c->nbr) (void)hipFree(c->nbr);
    c->nbr = nullptr;
    HIPC(dalloc(&c->nbr, ceil_div(c->p.max_nb, 4) * 4 * c->cap_pad));
    c->nbr_rows = c->p.max_nb;
    return MHPC_OK;
}

static void free_all(mhpc_ctx *c) {
    void *ptrs[] = {c->noi, c->level, c->h_iters, c->key, c->lkey, c->skey, c->keyB, c->perm,
                    c->valB, c->tile_hist, c->tile_off, c->digit_hist, c->slkey, c->sxm, c->sv4,
                    c->sr4, c->sh, c->sh2, c->se, c->srho, c->sp, c->scs, c->snoi, c->siters, c->sdone,
                    c->lcp, c->tcnt, c->toff,
                    c->level_count, c->cstart, c->cend, c->clevel, c->cchild,
                    c->cmass, c->cx, c->cy, c->cz, c->cquad, c->cquadf, c->cbmin, c->cbmax, c->chmax,
                    c->nbr, c->htable, c->bbox_partial, c->box, c->dstatus,
                    c->diag_partial, c->diag_out, c->pc_backup, c->mlp_params, c->mlp_m1, c->mlp_m2,
                    c->mlp_grad, c->mlp_norm, c->mlp_x, c->mlp_t, c->mlp_hprev, c->mlp_idx,
                    c->mlp_partial, c->mlp_loss_partial, c->mlp_loss, c->mlp_io_x, c->mlp_io_y,
                    c->hlist[0], c->hlist[1], c->hlist[2], c->sdirty,
                    c->h0slot, c->olist, c->ocount, c->hcnt, c->nn_start, c->nn_start_rng, c->hhard,
                    c->xyzm, c->ppar, c->cpar};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int f = 0; f < MHPC_F_NOI; f++)
        if (c->f_own[f]) (void)hipFree(c->f_own[f]);
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->h_status1) (void)hipHostFree(c->h_status1);
    for (int e = 0; e < MHPC_NUM_EVENTS; e++)
        if (c->ev[e]) (void)hipEventDestroy(c->ev[e]);
    if (c->sync_ev) (void)hipEventDestroy(c->sync_ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

static int create_impl(mhpc_ctx *c, int64_t cap, int64_t cap_cells) {
    HIPC(hipGetDevice(&c->device));
    HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int e = 0; e < MHPC_NUM_EVENTS; e++) HIPC(hipEventCreate(&c->ev[e]));
    HIPC(hipEventCreateWithFlags(&c->sync_ev, hipEventDisableTiming));
    HIPC(hipMalloc((void **)&c->dstatus, sizeof(DevStatus)));
    HIPC(hipMemset(c->dstatus, 0, sizeof(DevStatus)));
    HIPC(hipHostMalloc((void **)&c->h_status, sizeof(DevStatus), hipHostMallocDefault));
    memset(c->h_status, 0, sizeof(DevStatus));
    HIPC(hipHostMalloc((void **)&c->h_status1, sizeof(DevStatus), hipHostMallocDefault));
    memset(c->h_status1, 0, sizeof(DevStatus));
    c->err = &c->dstatus->err;
    c->counters = c->dstatus->counters;
    c->sort_check = &c->dstatus->sort_check;
    c->cell_base = c->dstatus->cell_base;
    c->hlist_count = c->dstatus->hlist_count;
    c->cap = cap;
    c->cap_pad = ceil_div(cap, 64) * 64;
    c->cap_cells = cap_cells;
    for (int f = 0; f < MHPC_F_NOI; f++) {
        HIPC(dalloc(&c->f_own[f], cap));
        HIPC(hipMemset(c->f_own[f], 0, sizeof(double) * (size_t)cap));
        c->f[f] = c->f_own[f];
    }
    HIPC(dalloc(&c->noi, cap));
    HIPC(dalloc(&c->level, cap));
    HIPC(dalloc(&c->h_iters, cap));
    HIPC(hipMemset(c->noi, 0, sizeof(int32_t) * (size_t)cap));
    HIPC(hipMemset(c->level, 0, sizeof(int32_t) * (size_t)cap));
    HIPC(hipMemset(c->h_iters, 0, sizeof(int32_t) * (size_t)cap));
    HIPC(dalloc(&c->key, cap));
    HIPC(dalloc(&c->lkey, cap));
    HIPC(dalloc(&c->skey, cap));
    HIPC(dalloc(&c->keyB, cap));
    HIPC(dalloc(&c->slkey, cap));
    HIPC(dalloc(&c->perm, cap));
    HIPC(dalloc(&c->valB, cap));
    HIPC(dalloc(&c->xyzm, cap));
    c->sort_tiles_cap = ceil_div(cap, MHPC_SORT_TILE);
    HIPC(dalloc(&c->tile_hist, MHPC_SORT_BUCKETS * c->sort_tiles_cap));
    HIPC(dalloc(&c->tile_off, MHPC_SORT_BUCKETS * c->sort_tiles_cap));
    HIPC(dalloc(&c->digit_hist, MHPC_SORT_PASSES * MHPC_SORT_BUCKETS));
    double **sorted[] = {&c->sh, &c->sh2, &c->se, &c->srho, &c->sp, &c->scs};
    for (double **p : sorted) HIPC(dalloc(p, cap));
    HIPC(dalloc(&c->sxm, cap));
    HIPC(dalloc(&c->sv4, cap));
    HIPC(dalloc(&c->sr4, cap));
    HIPC(dalloc(&c->snoi, cap));
    HIPC(dalloc(&c->siters, cap));
    HIPC(hipMemset(c->siters, 0, sizeof(int32_t) * (size_t)cap));
    HIPC(dalloc(&c->sdone, cap));
    HIPC(dalloc(&c->hlist[0], cap));
    HIPC(dalloc(&c->hlist[1], cap));
    HIPC(dalloc(&c->hlist[2], cap));
    HIPC(dalloc(&c->sdirty, cap));
    HIPC(dalloc(&c->h0slot, cap));
    HIPC(dalloc(&c->hhard, cap));
    HIPC(dalloc(&c->olist, cap));
    HIPC(dalloc(&c->ocount, 4));
    HIPC(dalloc(&c->hcnt, cap));
    // start table levels: the smallest K with 8^K >= cap (one geometric cell per particle on
    // average at the deepest level), 1 <= K <= MHPC_NN_KMAX
    c->nn_kmax = 1;
    while (c->nn_kmax < MHPC_NN_KMAX && (1ll << (3 * c->nn_kmax)) < cap) c->nn_kmax++;
    HIPC(dalloc(&c->nn_start, nn_tab_off(c->nn_kmax + 1)));
    HIPC(dalloc(&c->nn_start_rng, nn_tab_off(c->nn_kmax + 1)));
    HIPC(dalloc(&c->lcp, cap));
    c->tree_tiles_cap = ceil_div(cap, MHPC_TREE_TILE);
    HIPC(dalloc(&c->tcnt, (MHPC_LEVELS + 1) * c->tree_tiles_cap));
    HIPC(dalloc(&c->toff, (MHPC_LEVELS + 1) * c->tree_tiles_cap));
    HIPC(dalloc(&c->level_count, MHPC_LEVELS + 2));
    HIPC(hipMemset(c->level_count, 0, sizeof(int32_t) * (MHPC_LEVELS + 2)));
    HIPC(dalloc(&c->cstart, cap_cells));
    HIPC(dalloc(&c->cend, cap_cells));
    HIPC(dalloc(&c->clevel, cap_cells));
    HIPC(dalloc(&c->cpar, cap_cells));
    HIPC(dalloc(&c->ppar, cap));
    HIPC(dalloc(&c->cchild, 8 * cap_cells));
    HIPC(dalloc(&c->cmass, cap_cells));
    HIPC(dalloc(&c->cx, cap_cells));
    HIPC(dalloc(&c->cy, cap_cells));
    HIPC(dalloc(&c->cz, cap_cells));
    HIPC(dalloc(&c->cquad, 6 * cap_cells));
    HIPC(dalloc(&c->cquadf, 2 * cap_cells));
    HIPC(dalloc(&c->cbmin, 3 * cap_cells));
    HIPC(dalloc(&c->cbmax, 3 * cap_cells));
    HIPC(dalloc(&c->chmax, cap_cells));
    HIPC(dalloc(&c->htable, MHPC_HTABLE_MAX + 1));
    upload_htable(c);
    HIPC(dalloc(&c->bbox_partial, 2 * 1024));
    HIPC(dalloc(&c->diag_partial, diag_partial_size()));
    HIPC(dalloc(&c->diag_out, diag_num_quantities()));
    HIPC(dalloc(&c->box, 6));
    HIPC(hipMemset(c->box, 0, sizeof(double) * 6));
    return MHPC_OK;
}

extern "C" int mhpc_create(const mhpc_params *p, int64_t cap_particles, int64_t cap_cells,
                           mhpc_ctx **out) {
    if (!out) return MHPC_E_ARG;
    *out = nullptr;
    if (cap_particles < 1 || cap_particles > (int64_t)INT32_MAX / 2) return MHPC_E_ARG;
    if (cap_cells <= 0) cap_cells = cap_particles + cap_particles / 2 + 65536;
    if (cap_cells > (int64_t)INT32_MAX - cap_particles) return MHPC_E_ARG;
    mhpc_params def;
    mhpc_default_params(&def);
    if (!p) p = &def;
    if (!check_params(p)) return MHPC_E_ARG;
    mhpc_ctx *c = new mhpc_ctx();
    c->p = *p;
    int st = create_impl(c, cap_particles, cap_cells);
    if (st != MHPC_OK) {
        fprintf(stderr, "mhpc_create: %s\n", c->last_error.c_str());
        free_all(c);
        delete c;
        return st;
    }
    *out = c;
    return MHPC_OK;
}

extern "C" int mhpc_destroy(mhpc_ctx *c) {
    if (!c) return MHPC_E_ARG;
    (void)ctx_sync(c);
    free_all(c);
    delete c;
    return MHPC_OK;
}

extern "C" int mhpc_set_params(mhpc_ctx *c, const mhpc_params *p) {
    if (!c || !check_params(p)) return MHPC_E_ARG;
    const int old_target = c->p.n_target;
    const bool changed = memcmp(&c->p, p, sizeof(*p)) != 0;
    c->p = *p;
    if (p->n_target != old_target) upload_htable(c);
    if (changed) {
        // keys depend on the curve, lists on max_nb, sums on the precision and the EOS: every
        // stage after a parameter change recomputes from the keys (stage calls in between
        // report MHPC_E_STATE)
        c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
        c->nn_valid = c->rho_valid = c->eos_valid = false;
        c->nn_start_valid = false;
    }
    return MHPC_OK;
}

extern "C" int mhpc_get_params(mhpc_ctx *c, mhpc_params *p) {
    if (!c || !p) return MHPC_E_ARG;
    *p = c->p;
    return MHPC_OK;
}

extern "C" int mhpc_set_num_particles(mhpc_ctx *c, int64_t n) {
    if (!c || n < 0 || n > c->cap) return MHPC_E_ARG;
    c->n = n;
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
    c->nn_valid = c->rho_valid = c->eos_valid = false;
    return MHPC_OK;
}

extern "C" int64_t mhpc_num_particles(mhpc_ctx *c) { return c ? c->n : -1; }

extern "C" const char *mhpc_last_error(mhpc_ctx *c) { return c ? c->last_error.c_str() : "null context"; }

extern "C" int mhpc_synchronize(mhpc_ctx *c) {
    if (!c) return MHPC_E_ARG;
    HIPC(ctx_sync(c));
    return MHPC_OK;
}

struct FieldDesc {
    void *ptr;
    size_t elem;
};

static FieldDesc field_desc(mhpc_ctx *c, int field) {
    if (field >= 0 && field < MHPC_F_NOI) return {c->f[field], sizeof(double)};
    switch (field) {
        case MHPC_F_NOI: return {c->noi, sizeof(int32_t)};
        case MHPC_F_KEY: return {c->key, sizeof(uint64_t)};
        case MHPC_F_SORTED: return {c->perm, sizeof(int32_t)};
        case MHPC_F_LEVEL: return {c->level, sizeof(int32_t)};
        case MHPC_F_H_ITERS: return {c->h_iters, sizeof(int32_t)};
        default: return {nullptr, 0};
    }
}

extern "C" int mhpc_upload(mhpc_ctx *c, int field, const void *host, int64_t n) {
    if (!c || !host || n < 0 || n > c->n) return MHPC_E_ARG;
    FieldDesc d = field_desc(c, field);
    if (!d.ptr) return MHPC_E_ARG;
    HIPC(hipMemcpyAsync(d.ptr, host, d.elem * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPC(ctx_sync(c));
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
    c->nn_valid = c->rho_valid = c->eos_valid = false;
    return MHPC_OK;
}

extern "C" int mhpc_download(mhpc_ctx *c, int field, void *host, int64_t n) {
    if (!c || !host || n < 0 || n > c->n) return MHPC_E_ARG;
    FieldDesc d = field_desc(c, field);
    if (!d.ptr) return MHPC_E_ARG;
    HIPC(hipMemcpyAsync(host, d.ptr, d.elem * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    return MHPC_OK;
}

extern "C" int mhpc_upload_device(mhpc_ctx *c, int field, const void *device_src, int64_t n) {
    if (!c || !device_src || n < 0 || n > c->n) return MHPC_E_ARG;
    FieldDesc d = field_desc(c, field);
    if (!d.ptr) return MHPC_E_ARG;
    HIPC(hipMemcpyAsync(d.ptr, device_src, d.elem * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
    c->nn_valid = c->rho_valid = c->eos_valid = false;
    return MHPC_OK;
}

extern "C" int mhpc_field_ptr(mhpc_ctx *c, int field, void **device_ptr) {
    if (!c || !device_ptr) return MHPC_E_ARG;
    FieldDesc d = field_desc(c, field);
    if (!d.ptr) return MHPC_E_ARG;
    *device_ptr = d.ptr;
    return MHPC_OK;
}

static void invalidate(mhpc_ctx *c) {
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
    c->nn_valid = c->rho_valid = c->eos_valid = false;
}

extern "C" int mhpc_bind_field(mhpc_ctx *c, int field, void *device_ptr) {
    if (!c || field < 0 || field >= MHPC_F_NOI) return MHPC_E_ARG;
    HIPC(ctx_sync(c));   // work queued on the previous array finishes first
    c->f[field] = device_ptr ? static_cast<double *>(device_ptr) : c->f_own[field];
    invalidate(c);
    return MHPC_OK;
}

extern "C" int mhpc_device_alloc(mhpc_ctx *c, int64_t bytes, void **device_ptr) {
    if (!c || !device_ptr || bytes < 1) return MHPC_E_ARG;
    *device_ptr = nullptr;
    HIPC(hipMalloc(device_ptr, (size_t)bytes));
    HIPC(hipMemset(*device_ptr, 0, (size_t)bytes));
    return MHPC_OK;
}

extern "C" int mhpc_device_upload(mhpc_ctx *c, void *device_ptr, const void *host, int64_t bytes) {
    if (!c || !device_ptr || !host || bytes < 0) return MHPC_E_ARG;
    HIPC(ctx_sync(c));
    HIPC(hipMemcpy(device_ptr, host, (size_t)bytes, hipMemcpyHostToDevice));
    return MHPC_OK;
}

extern "C" int mhpc_device_download(mhpc_ctx *c, void *host, const void *device_ptr, int64_t bytes) {
    if (!c || !device_ptr || !host || bytes < 0) return MHPC_E_ARG;
    HIPC(ctx_sync(c));
    HIPC(hipMemcpy(host, device_ptr, (size_t)bytes, hipMemcpyDeviceToHost));
    return MHPC_OK;
}

extern "C" int mhpc_device_free(mhpc_ctx *c, void *device_ptr) {
    if (!c) return MHPC_E_ARG;
    if (!device_ptr) return MHPC_OK;
    for (int f = 0; f < MHPC_F_NOI; f++)
        if (c->f[f] == device_ptr) return MHPC_E_STATE;   // still bound to a field
    HIPC(ctx_sync(c));
    HIPC(hipFree(device_ptr));
    return MHPC_OK;
}

extern "C" int mhpc_get_box(mhpc_ctx *c, double box[6]) {
    if (!c || !box) return MHPC_E_ARG;
    HIPC(hipMemcpyAsync(box, c->box, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    return MHPC_OK;
}

// ------------------------------------------------------------------------------------
static void reset_flags(mhpc_ctx *c) {
    launch_fill_u32(c, c->err, 0u, 1);
    launch_fill_u64(c, c->counters, 0ull, CNT_SLOTS * CNT_LINE);
}

hipError_t ctx_sync(mhpc_ctx *c) {
    hipError_t e = hipEventRecord(c->sync_ev, c->stream);
    if (e != hipSuccess) return e;
    while ((e = hipEventQuery(c->sync_ev)) == hipErrorNotReady) __builtin_ia32_pause();
    return e;
}

hipError_t ctx_fetch_status(mhpc_ctx *c) {
    hipError_t e = hipMemcpyAsync(c->h_status, c->dstatus, sizeof(DevStatus), hipMemcpyDeviceToHost, c->stream);
    return e != hipSuccess ? e : ctx_sync(c);
}

// the status of a call from a fetched copy of the device status block (and the tree depth
// the next sort starts from)
static int status_of(mhpc_ctx *c, const DevStatus &hs) {
    const uint32_t e = hs.err;
    const unsigned long long depth = cnt_max(hs, CNT_TREE_DEPTH);
    if (depth > 0) c->com_depth_hint = (int)depth;
    if (e == 0) return MHPC_OK;
    int st = MHPC_OK;
    if (e & ERRF_DEPTH) st = MHPC_E_DEPTH;
    else if (e & ERRF_CAP_NODES) st = MHPC_E_CAP_NODES;
    else if (e & ERRF_STACK) st = MHPC_E_STACK;
    else if (e & ERRF_CAP_NEIGHBOURS) st = MHPC_E_CAP_NEIGHBOURS;
    else if (e & ERRF_RHO) st = MHPC_E_RHO_NONPOSITIVE;
    c->last_error = mhpc_status_string(st);
    return st;
}

static int finish(mhpc_ctx *c) {
    HIPC(hipGetLastError());
    HIPC(ctx_fetch_status(c));
    return status_of(c, *c->h_status);
}

// ms between two events of a finished call; 0 for an event the call did not record (a stage
// with nothing to do, e.g. no particles), whose query error is cleared here so that the next
// call's error check does not report it
static float elapsed_ev(const hipEvent_t *ev, int a, int b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) {
        ms = 0.f;
        (void)hipGetLastError();
    }
    return ms;
}

static float elapsed(mhpc_ctx *c, int a, int b) { return elapsed_ev(c->ev, a, b); }

// one stage: reset flags, launch, time with events, sync, map errors
template <typename F>
static int run_stage(mhpc_ctx *c, float *ms, F &&fn) {
    reset_flags(c);
    HIPC(hipEventRecord(c->ev[0], c->stream));
    int st = fn();
    if (st != MHPC_OK) return st;
    HIPC(hipEventRecord(c->ev[1], c->stream));
    st = finish(c);
    if (ms) *ms = elapsed(c, 0, 1);
    return st;
}

extern "C" int mhpc_bbox(mhpc_ctx *c, double box[6], float *ms) {
    if (!c) return MHPC_E_ARG;
    int st = run_stage(c, ms, [&] {
        launch_bbox(c);
        return (int)MHPC_OK;
    });
    if (st == MHPC_OK && box) st = mhpc_get_box(c, box);
    return st;
}

extern "C" int mhpc_keys(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    int st = run_stage(c, ms, [&] {
        launch_keys(c);
        return (int)MHPC_OK;
    });
    c->keys_valid = st == MHPC_OK;
    return st;
}

extern "C" int mhpc_sort(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->keys_valid) return MHPC_E_STATE;
    int st = run_stage(c, ms, [&] {
        launch_sort(c);
        return (int)MHPC_OK;
    });
    c->sorted_valid = st == MHPC_OK;
    return st;
}

extern "C" int mhpc_build_tree(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->sorted_valid) return MHPC_E_STATE;
    int st = run_stage(c, ms, [&] {
        launch_gather(c, true, true, false);
        launch_tree(c);
        launch_bottom_up(c, false, true);
        return (int)MHPC_OK;
    });
    c->tree_valid = st == MHPC_OK;
    return st;
}

extern "C" int mhpc_centres_of_mass(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->tree_valid) return MHPC_E_STATE;
    int st = run_stage(c, ms, [&] {
        launch_bottom_up(c, true, true);
        return (int)MHPC_OK;
    });
    c->com_valid = st == MHPC_OK;
    return st;
}

extern "C" int mhpc_gravity(mhpc_ctx *c, int potential, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->com_valid) return MHPC_E_STATE;
    return run_stage(c, ms, [&] {
        launch_gravity(c, potential != 0);
        return (int)MHPC_OK;
    });
}

extern "C" int mhpc_h_iterate(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->tree_valid) return MHPC_E_STATE;
    int st = run_stage(c, ms, [&] {
        launch_h_iterate(c, c->p.max_counts, false);
        launch_scatter_h(c);
        return (int)MHPC_OK;
    });
    c->nn_valid = false;
    return st;
}

extern "C" int mhpc_neighbours(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->tree_valid) return MHPC_E_STATE;
    int st = ensure_nbr(c);
    if (st != MHPC_OK) return st;
    st = run_stage(c, ms, [&] {
        launch_neighbours(c);
        return (int)MHPC_OK;
    });
    c->nn_valid = st == MHPC_OK;
    return st;
}

extern "C" int mhpc_density(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->nn_valid) return MHPC_E_STATE;
    int st = run_stage(c, ms, [&] {
        launch_density(c);
        return (int)MHPC_OK;
    });
    c->rho_valid = st == MHPC_OK;
    return st;
}

extern "C" int mhpc_eos(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->rho_valid) return MHPC_E_STATE;
    int st = run_stage(c, ms, [&] {
        launch_eos(c);
        return (int)MHPC_OK;
    });
    c->eos_valid = st == MHPC_OK;
    return st;
}

extern "C" int mhpc_internal_forces(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->eos_valid) return MHPC_E_STATE;
    return run_stage(c, ms, [&] {
        launch_forces(c, false);
        return (int)MHPC_OK;
    });
}

extern "C" int mhpc_time_step(mhpc_ctx *c, double courant_fact, double forces_fact, double dt_max,
                              double *dt, double parts[5]) {
    if (!c || !dt) return MHPC_E_ARG;
    double q[5];
    const DtRule rule = {courant_fact, forces_fact, dt_max, -1.0, nullptr};
    launch_time_step(c, c->diag_partial, c->diag_out, rule);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(q, c->diag_out, sizeof(q), hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    const double t = dt_rule(rule, q[0], q[1], q[2], q[3], q[4]);
    *dt = t;
    c->last_dt = t;
    if (parts)
        for (int k = 0; k < 5; k++) parts[k] = q[k];
    return MHPC_OK;
}

// ------------------------------------------------------------------------------------
// learned smoothing length (no reference counterpart)
#define MLP_MAX_SAMPLES 262144

static int ensure_mlp(mhpc_ctx *c) {
    if (c->mlp_params) return MHPC_OK;
    const int np = mlp_num_params();
    const int64_t max_waves = ceil_div(MLP_MAX_SAMPLES, 64);
    const int64_t max_chunks = ceil_div(max_waves, 32);   // stage-1 reduction sums
    HIPC(dalloc(&c->mlp_params, np));
    HIPC(dalloc(&c->mlp_m1, np));
    HIPC(dalloc(&c->mlp_m2, np));
    HIPC(dalloc(&c->mlp_grad, np));
    HIPC(dalloc(&c->mlp_norm, 18));
    HIPC(dalloc(&c->mlp_x, c->cap * 8));
    HIPC(dalloc(&c->mlp_t, c->cap));
    HIPC(dalloc(&c->mlp_hprev, c->cap));
    HIPC(dalloc(&c->mlp_idx, MLP_MAX_SAMPLES));
    HIPC(dalloc(&c->mlp_partial, (max_waves + max_chunks) * np));
    HIPC(dalloc(&c->mlp_loss_partial, max_waves + max_chunks));
    c->mlp_max_waves = max_waves;
    HIPC(dalloc(&c->mlp_loss, 1));
    float norm[18];
    for (int k = 0; k < 18; k++) norm[k] = (k >= 8 && k < 15) || k == 17 ? 1.f : 0.f;
    HIPC(hipMemcpy(c->mlp_norm, norm, sizeof(norm), hipMemcpyHostToDevice));
    c->mlp_norm_set = false;
    c->mlp_x_valid = false;
    return MHPC_OK;
}

static uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int upload_mlp_params(mhpc_ctx *c, const float *params) {
    const size_t bytes = sizeof(float) * (size_t)mlp_num_params();
    HIPC(hipMemcpy(c->mlp_params, params, bytes, hipMemcpyHostToDevice));
    HIPC(hipMemset(c->mlp_m1, 0, bytes));
    HIPC(hipMemset(c->mlp_m2, 0, bytes));
    c->mlp_adam_step = 0;
    return MHPC_OK;
}

// uniform(-a, a) weights with a = sqrt(6 / (fan_in + fan_out)) (Glorot), zero biases
extern "C" int mhpc_mlp_init(mhpc_ctx *c, uint64_t seed) {
    if (!c) return MHPC_E_ARG;
    int st = ensure_mlp(c);
    if (st != MHPC_OK) return st;
    const int H = 64, I = 8;
    std::vector<float> p((size_t)mlp_num_params(), 0.f);
    uint64_t s = seed;
    auto uni = [&](float a) { return (float)(((double)(splitmix64(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * a); };
    const float a1 = sqrtf(6.f / (7 + H)), a2 = sqrtf(6.f / (H + H)), a3 = sqrtf(6.f / (H + 1));
    for (int o = 0; o < H; o++)
        for (int k = 0; k < I; k++) p[(size_t)(o * I + k)] = k < 7 ? uni(a1) : 0.f;
    float *w2 = p.data() + H * I + H;
    for (int t = 0; t < H * H; t++) w2[t] = uni(a2);
    float *w3 = w2 + H * H + H;
    for (int t = 0; t < H; t++) w3[t] = uni(a3);
    return upload_mlp_params(c, p.data());
}

extern "C" int mhpc_mlp_set_params(mhpc_ctx *c, const float *params, int nparam) {
    if (!c || !params || nparam != mlp_num_params()) return MHPC_E_ARG;
    int st = ensure_mlp(c);
    if (st != MHPC_OK) return st;
    return upload_mlp_params(c, params);
}

extern "C" int mhpc_mlp_get_params(mhpc_ctx *c, float *params, int nparam) {
    if (!c || !params || nparam != mlp_num_params()) return MHPC_E_ARG;
    if (!c->mlp_params) return MHPC_E_STATE;
    HIPC(hipMemcpy(params, c->mlp_params, sizeof(float) * (size_t)nparam, hipMemcpyDeviceToHost));
    return MHPC_OK;
}

extern "C" int mhpc_mlp_set_norm(mhpc_ctx *c, const float norm[18]) {
    if (!c || !norm) return MHPC_E_ARG;
    for (int k = 0; k < 7; k++)
        if (!(norm[8 + k] > 0.f)) return MHPC_E_ARG;
    if (!(norm[17] > 0.f)) return MHPC_E_ARG;
    int st = ensure_mlp(c);
    if (st != MHPC_OK) return st;
    HIPC(hipMemcpy(c->mlp_norm, norm, sizeof(float) * 18, hipMemcpyHostToDevice));
    c->mlp_norm_set = true;
    c->mlp_x_valid = false;   // features must be recomputed with the new statistics
    return MHPC_OK;
}

extern "C" int mhpc_mlp_get_norm(mhpc_ctx *c, float norm[18]) {
    if (!c || !norm) return MHPC_E_ARG;
    if (!c->mlp_norm) return MHPC_E_STATE;
    HIPC(hipMemcpy(norm, c->mlp_norm, sizeof(float) * 18, hipMemcpyDeviceToHost));
    return MHPC_OK;
}

extern "C" int mhpc_mlp_features(mhpc_ctx *c, double dt) {
    if (!c || !(dt >= 0.0)) return MHPC_E_ARG;
    if (!c->mlp_params) return MHPC_E_STATE;
    launch_mlp_features(c, dt);
    HIPC(hipGetLastError());
    HIPC(ctx_sync(c));
    c->mlp_x_valid = true;
    return MHPC_OK;
}

extern "C" int mhpc_mlp_download_features(mhpc_ctx *c, float *X, int64_t n) {
    if (!c || !X || n != c->n) return MHPC_E_ARG;
    if (!c->mlp_x_valid) return MHPC_E_STATE;
    HIPC(hipMemcpy(X, c->mlp_x, sizeof(float) * 8 * (size_t)n, hipMemcpyDeviceToHost));
    return MHPC_OK;
}

extern "C" int mhpc_mlp_predict(mhpc_ctx *c, float *ms) {
    if (!c) return MHPC_E_ARG;
    if (!c->mlp_x_valid || !c->mlp_norm_set) return MHPC_E_STATE;
    return run_stage(c, ms, [&] {
        launch_mlp_predict(c);
        return (int)MHPC_OK;
    });
}

// feature / target statistics from the current raw features and the converged h
static int mlp_take_norm(mhpc_ctx *c) {
    const int64_t n = c->n;
    std::vector<float> X((size_t)n * 8);
    std::vector<double> h((size_t)n), hp((size_t)n);
    HIPC(hipMemcpy(X.data(), c->mlp_x, sizeof(float) * X.size(), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(h.data(), c->f[MHPC_F_SML], sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(hp.data(), c->mlp_hprev, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    float norm[18] = {};
    for (int k = 0; k < 8; k++) {
        double m = 0.0, q = 0.0;
        for (int64_t i = 0; i < n; i++) m += X[(size_t)(i * 8 + k)];
        m /= (double)n;
        for (int64_t i = 0; i < n; i++) {
            const double d = X[(size_t)(i * 8 + k)] - m;
            q += d * d;
        }
        const double sd = sqrt(q / (double)n);
        norm[k] = k < 7 ? (float)m : 0.f;
        norm[8 + k] = k < 7 && sd > 1e-6 * (fabs(m) + 1e-30) && sd > 0.0 ? (float)sd : 1.f;
    }
    double m = 0.0, q = 0.0;
    for (int64_t i = 0; i < n; i++) m += log(h[(size_t)i] / hp[(size_t)i]);
    m /= (double)n;
    for (int64_t i = 0; i < n; i++) {
        const double d = log(h[(size_t)i] / hp[(size_t)i]) - m;
        q += d * d;
    }
    const double sd = sqrt(q / (double)n);
    norm[16] = (float)m;
    norm[17] = sd > 1e-4 ? (float)sd : 1e-4f;
    HIPC(hipMemcpy(c->mlp_norm, norm, sizeof(norm), hipMemcpyHostToDevice));
    launch_mlp_renorm(c);
    HIPC(hipGetLastError());
    c->mlp_norm_set = true;
    return MHPC_OK;
}

extern "C" int mhpc_mlp_train(mhpc_ctx *c, int nsamples, int iters, float lr, uint64_t seed, float *loss) {
    if (!c || nsamples < 1 || nsamples > MLP_MAX_SAMPLES || iters < 0 || !(lr > 0.f)) return MHPC_E_ARG;
    if (!c->mlp_x_valid || c->n < 1) return MHPC_E_STATE;
    if (!c->mlp_norm_set) {
        int st = mlp_take_norm(c);
        if (st != MHPC_OK) return st;
    }
    launch_mlp_targets(c);
    for (int it = 0; it < iters; it++) {
        launch_mlp_sample(c, nsamples, seed + (uint64_t)it);
        launch_mlp_gradient(c, nsamples, c->mlp_x, c->mlp_t, c->mlp_idx, c->mlp_grad, c->mlp_loss);
        launch_mlp_adam(c, lr, ++c->mlp_adam_step);
    }
    HIPC(hipGetLastError());
    float l = 0.f;
    if (iters > 0) HIPC(hipMemcpyAsync(&l, c->mlp_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    if (loss) *loss = l;
    return MHPC_OK;
}

static int ensure_mlp_io(mhpc_ctx *c, int64_t n) {
    if (c->mlp_io_cap >= n) return MHPC_OK;
    if (c->mlp_io_x) (void)hipFree(c->mlp_io_x);
    if (c->mlp_io_y) (void)hipFree(c->mlp_io_y);
    c->mlp_io_x = c->mlp_io_y = nullptr;
    c->mlp_io_cap = 0;
    HIPC(dalloc(&c->mlp_io_x, n * 8));
    HIPC(dalloc(&c->mlp_io_y, n));
    c->mlp_io_cap = n;
    return MHPC_OK;
}

extern "C" int mhpc_mlp_forward(mhpc_ctx *c, const float *X, int64_t n, float *Y) {
    if (!c || !X || !Y || n < 1) return MHPC_E_ARG;
    if (!c->mlp_params) return MHPC_E_STATE;
    int st = ensure_mlp_io(c, n);
    if (st != MHPC_OK) return st;
    HIPC(hipMemcpyAsync(c->mlp_io_x, X, sizeof(float) * 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    launch_mlp_forward_raw(c, n, c->mlp_io_x, c->mlp_io_y);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(Y, c->mlp_io_y, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    return MHPC_OK;
}

extern "C" int mhpc_mlp_gradient(mhpc_ctx *c, const float *X, const float *T, int64_t n, float *grad,
                                 float *loss) {
    if (!c || !X || !T || !grad || n < 1 || n > MLP_MAX_SAMPLES) return MHPC_E_ARG;
    if (!c->mlp_params) return MHPC_E_STATE;
    int st = ensure_mlp_io(c, n);
    if (st != MHPC_OK) return st;
    HIPC(hipMemcpyAsync(c->mlp_io_x, X, sizeof(float) * 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(c->mlp_io_y, T, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    launch_mlp_gradient(c, n, c->mlp_io_x, c->mlp_io_y, nullptr, c->mlp_grad, c->mlp_loss);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(grad, c->mlp_grad, sizeof(float) * (size_t)mlp_num_params(), hipMemcpyDeviceToHost,
                        c->stream));
    float l = 0.f;
    HIPC(hipMemcpyAsync(&l, c->mlp_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    if (loss) *loss = l;
    return MHPC_OK;
}

static int ensure_pc_backup(mhpc_ctx *c) {
    if (c->pc_backup) return MHPC_OK;
    HIPC(dalloc(&c->pc_backup, (int64_t)pc_backup_fields() * c->cap));
    return MHPC_OK;
}

extern "C" int mhpc_pc_predictor(mhpc_ctx *c, double dt) {
    if (!c) return MHPC_E_ARG;
    int st = ensure_pc_backup(c);
    if (st != MHPC_OK) return st;
    launch_pc_predictor(c, dt);
    HIPC(hipGetLastError());
    HIPC(ctx_sync(c));
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
    c->nn_valid = c->rho_valid = c->eos_valid = false;
    return MHPC_OK;
}

extern "C" int mhpc_pc_corrector(mhpc_ctx *c, double dt) {
    if (!c || !c->pc_backup) return c ? MHPC_E_STATE : MHPC_E_ARG;
    launch_pc_corrector(c, dt);
    HIPC(hipGetLastError());
    HIPC(ctx_sync(c));
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
    c->nn_valid = c->rho_valid = c->eos_valid = false;
    return MHPC_OK;
}

extern "C" int mhpc_leapfrog_update_x(mhpc_ctx *c, double dt) {
    if (!c) return MHPC_E_ARG;
    launch_leapfrog_x(c, dt);
    HIPC(hipGetLastError());
    HIPC(ctx_sync(c));
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
    c->nn_valid = c->rho_valid = c->eos_valid = false;
    return MHPC_OK;
}

extern "C" int mhpc_leapfrog_update_v(mhpc_ctx *c, double dt) {
    if (!c) return MHPC_E_ARG;
    launch_leapfrog_v(c, dt);
    HIPC(hipGetLastError());
    HIPC(ctx_sync(c));
    return MHPC_OK;
}

extern "C" int mhpc_compute_diagnostics(mhpc_ctx *c, mhpc_diagnostics *out) {
    if (!c || !out) return MHPC_E_ARG;
    mhpc_diagnostics *d = reinterpret_cast<mhpc_diagnostics *>(c->diag_out);
    launch_diagnostics(c, c->diag_partial, d);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(out, d, sizeof(*out), hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    return MHPC_OK;
}

extern "C" int mhpc_download_neighbours(mhpc_ctx *c, int32_t *nnl, int32_t *noi, int64_t n, int max_nb) {
    if (!c || !nnl || !noi || n != c->n || max_nb < 1) return MHPC_E_ARG;
    if (!c->nn_valid) return MHPC_E_STATE;
    int32_t *d_nnl = nullptr, *d_noi = nullptr;
    HIPC(dalloc(&d_nnl, n * (int64_t)max_nb));
    HIPC(dalloc(&d_noi, n));
    HIPC(hipMemsetAsync(d_nnl, 0xff, sizeof(int32_t) * (size_t)(n * max_nb), c->stream));
    launch_neighbours_to_caller(c, d_nnl, d_noi, max_nb);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(nnl, d_nnl, sizeof(int32_t) * (size_t)(n * max_nb), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipMemcpyAsync(noi, d_noi, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    (void)hipFree(d_nnl);
    (void)hipFree(d_noi);
    return MHPC_OK;
}

extern "C" int mhpc_tree_size(mhpc_ctx *c, int64_t *num_cells) {
    if (!c || !num_cells) return MHPC_E_ARG;
    if (!c->tree_valid) return MHPC_E_STATE;
    int32_t cb[MHPC_LEVELS + 3];
    HIPC(hipMemcpyAsync(cb, c->cell_base, sizeof(cb), hipMemcpyDeviceToHost, c->stream));
    HIPC(ctx_sync(c));
    *num_cells = cb[MHPC_LEVELS + 1];
    return MHPC_OK;
}

extern "C" int mhpc_download_tree(mhpc_ctx *c, int32_t *cell_level, int32_t *cell_start,
                                  int32_t *cell_end, int32_t *cell_child, double *cell_mass,
                                  double *cell_x, double *cell_y, double *cell_z,
                                  double *cell_quad) {
    int64_t nc = 0;
    int st = mhpc_tree_size(c, &nc);
    if (st != MHPC_OK) return st;
    if (nc > c->cap_cells) return MHPC_E_CAP_NODES;
    const size_t i4 = sizeof(int32_t) * (size_t)nc, d8 = sizeof(double) * (size_t)nc;
    if (cell_level) HIPC(hipMemcpyAsync(cell_level, c->clevel, i4, hipMemcpyDeviceToHost, c->stream));
    if (cell_start) HIPC(hipMemcpyAsync(cell_start, c->cstart, i4, hipMemcpyDeviceToHost, c->stream));
    if (cell_end) HIPC(hipMemcpyAsync(cell_end, c->cend, i4, hipMemcpyDeviceToHost, c->stream));
    if (cell_child) HIPC(hipMemcpyAsync(cell_child, c->cchild, 8 * i4, hipMemcpyDeviceToHost, c->stream));
    if (cell_mass) HIPC(hipMemcpyAsync(cell_mass, c->cmass, d8, hipMemcpyDeviceToHost, c->stream));
    if (cell_x) HIPC(hipMemcpyAsync(cell_x, c->cx, d8, hipMemcpyDeviceToHost, c->stream));
    if (cell_y) HIPC(hipMemcpyAsync(cell_y, c->cy, d8, hipMemcpyDeviceToHost, c->stream));
    if (cell_z) HIPC(hipMemcpyAsync(cell_z, c->cz, d8, hipMemcpyDeviceToHost, c->stream));
    if (cell_quad)
        for (int e = 0; e < 6; e++)
            HIPC(hipMemcpyAsync(cell_quad + e * nc, c->cquad + e * c->cap_cells, d8, hipMemcpyDeviceToHost,
                                c->stream));
    HIPC(ctx_sync(c));
    return MHPC_OK;
}

// ------------------------------------------------------------------------------------
// Miluphpc::rhs (miluphpc.cpp:387-562) for one subdomain: reset -> bounding box -> keys ->
// sort -> tree -> [COM -> Barnes-Hut] -> [h-iteration -> neighbours -> density -> sound
// speed / pressure -> internal forces].  No host synchronisation until the end.
// what one RHS call does, from its flags
struct RhsPlan {
    bool grav, sph, hiter, pot, learned, feats, spec, predict, timing;
};

static int rhs_plan(mhpc_ctx *c, uint32_t flags, RhsPlan &P) {
    P.grav = flags & MHPC_RHS_GRAVITY;
    P.sph = flags & MHPC_RHS_SPH;
    P.hiter = P.sph && (flags & MHPC_RHS_VARIABLE_SML);
    P.pot = P.grav && (flags & MHPC_RHS_POTENTIAL);
    P.learned = P.hiter && (flags & MHPC_RHS_LEARNED_H);
    // learned-h features (MLP inputs and curl v / grad rho) without the prediction:
    // the plain steps that produce training pairs
    P.feats = P.hiter && (flags & (MHPC_RHS_LEARNED_H | MHPC_RHS_FEATURES));
    // speculative lists: pass 0 of the h-iteration writes the lists, and only particles
    // whose list can differ from the final one are re-listed (same neighbour sets)
    P.spec = P.hiter && (P.learned || (flags & MHPC_RHS_SPECULATIVE_NN));
    P.timing = flags & MHPC_RHS_TIMING;
    if (P.sph) {
        int st = ensure_nbr(c);
        if (st != MHPC_OK) return st;
    }
    if (P.feats) {
        int st = ensure_mlp(c);
        if (st != MHPC_OK) return st;
    }
    // the learned start applies once the model has feature statistics; before that the
    // features are still taken (raw) so that a following mhpc_mlp_train can use them
    P.predict = P.learned && c->mlp_norm_set;
    return MHPC_OK;
}

// Queue one RHS on the context stream (no host synchronisation besides the sort's check).
// ev: 16 events; per-phase events only with MHPC_RHS_TIMING; the total (13 -> 12), the pass-0
// walk (14 -> 15), the density (9 -> 10) and the forces (11 -> 12) are always timed.
static void rhs_launch(mhpc_ctx *c, const RhsPlan &P, hipEvent_t *ev) {
    auto mark = [&](int idx) {
        if (P.timing || idx >= 9) (void)hipEventRecord(ev[idx], c->stream);
    };
    reset_flags(c);
    mark(13);
    if (P.feats) {
        launch_mlp_features(c, c->last_dt);
        if (P.predict) launch_mlp_predict(c);
    }
    mark(0);
    launch_bbox(c);
    mark(1);
    launch_keys(c);
    mark(2);
    launch_sort(c);
    mark(3);
    launch_gather(c, P.grav, P.sph, true);
    mark(4);
    launch_tree(c);
    mark(5);
    launch_bottom_up(c, P.grav, P.sph);
    mark(6);
    if (P.grav) launch_gravity(c, P.pot);
    mark(7);
    if (P.hiter) {
        launch_h_iterate(c, c->p.max_counts, P.spec, ev + 14);
        if (!P.spec) launch_scatter_h(c);   // with lists from pass 0, the list fix-up publishes h
    }
    mark(8);
    if (P.spec) launch_neighbours_fixup(c);
    else if (P.sph) launch_neighbours(c);
    mark(9);
    // sound speed and pressure fused (ms_eos stays ~0); after the list fix-up it also publishes
    // h and the neighbour counts
    if (P.sph) launch_density_eos(c, P.spec);
    mark(10);
    mark(11);
    if (P.sph) launch_forces(c, P.feats);
    mark(12);
}

// the stage flags after an RHS that ended with status st
static void rhs_valid(mhpc_ctx *c, const RhsPlan &P, int st) {
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = (st == MHPC_OK);
    c->com_valid = c->com_valid && P.grav;
    c->nn_valid = c->rho_valid = c->eos_valid = (st == MHPC_OK) && P.sph;
    if (P.feats) c->mlp_x_valid = st == MHPC_OK;
}

// statistics of a finished RHS from its events and its copy of the device status block
static void rhs_stats(mhpc_ctx *c, const RhsPlan &P, const hipEvent_t *ev, const DevStatus &hs,
                      mhpc_rhs_stats *stats) {
    if (!stats) return;
    memset(stats, 0, sizeof(*stats));
    if (P.timing) {
        stats->ms_bbox = elapsed_ev(ev, 0, 1);
        stats->ms_keys = elapsed_ev(ev, 1, 2);
        stats->ms_sort = elapsed_ev(ev, 2, 3);
        stats->ms_gather = elapsed_ev(ev, 3, 4);
        stats->ms_tree = elapsed_ev(ev, 4, 5);
        stats->ms_com = elapsed_ev(ev, 5, 6);
        stats->ms_gravity = elapsed_ev(ev, 6, 7);
        stats->ms_hiter = elapsed_ev(ev, 7, 8);
        stats->ms_nn = elapsed_ev(ev, 8, 9);
        stats->ms_eos = elapsed_ev(ev, 10, 11);
        stats->ms_mlp = elapsed_ev(ev, 13, 0);
    }
    stats->ms_density = P.sph ? elapsed_ev(ev, 9, 10) : 0.f;
    stats->ms_forces = P.sph ? elapsed_ev(ev, 11, 12) : 0.f;
    stats->ms_total = elapsed_ev(ev, 13, 12);
    stats->num_cells = hs.cell_base[MHPC_LEVELS + 1];
    stats->n_cell_interactions = (int64_t)cnt_sum(hs, CNT_CELL_INT);
    stats->n_part_interactions = (int64_t)cnt_sum(hs, CNT_PART_INT);
    stats->n_neighbours = (int64_t)cnt_sum(hs, CNT_NEIGHBOURS);
    stats->max_noi = (int32_t)cnt_max(hs, CNT_MAX_NOI);
    stats->n_h_counts = (int64_t)cnt_sum(hs, CNT_H_COUNTS);
    stats->ms_nn_walk = P.hiter ? elapsed_ev(ev, 14, 15) : 0.f;
    stats->h_passes = 0;
    if (P.hiter) {
        // pass 0, then every pass whose input list (the previous pass's unconverged
        // particles) was not empty
        stats->h_passes = 1;
        for (int pass = 1; pass < c->p.max_counts; pass++) stats->h_passes += hs.hlist_count[pass - 1] > 0 ? 1 : 0;
    }
}

extern "C" int mhpc_rhs(mhpc_ctx *c, uint32_t flags, mhpc_rhs_stats *stats) {
    if (!c) return MHPC_E_ARG;
    RhsPlan P;
    int st = rhs_plan(c, flags, P);
    if (st != MHPC_OK) return st;
    rhs_launch(c, P, c->ev);
    st = finish(c);
    rhs_valid(c, P, st);
    rhs_stats(c, P, c->ev, *c->h_status, stats);
    return st;
}

// One sub-step of PredictorCorrectorEuler::integrate (predictor_corrector_euler.cpp): rhs with
// `flags`, the CFL time step, the predictor, rhs without self-gravity on the predicted state,
// the corrector -- queued back to back on the context stream.  The time step is computed and
// consumed on the device (k_time_step_final writes it, the predictor and corrector read it),
// the first RHS's status block is copied aside without waiting, and the host waits once at the
// end (besides each sort's check): the same kernels and results as mhpc_rhs, mhpc_time_step,
// mhpc_pc_predictor, mhpc_rhs and mhpc_pc_corrector called in turn, without their four
// round trips.  Returns the first RHS's error, else the second's; on an error the state is the
// caller's to discard (the reference stops).
extern "C" int mhpc_pc_euler_step(mhpc_ctx *c, uint32_t flags, double courant_fact, double forces_fact,
                                  double dt_max, double t_left, double *dt, mhpc_rhs_stats *stats0,
                                  mhpc_rhs_stats *stats1) {
    if (!c || !dt) return MHPC_E_ARG;
    // the learned-h features use the previous time step on the host side: the separate calls
    if (flags & (MHPC_RHS_LEARNED_H | MHPC_RHS_FEATURES)) return MHPC_E_ARG;
    int st = ensure_pc_backup(c);
    if (st != MHPC_OK) return st;
    RhsPlan P0, P1;
    st = rhs_plan(c, flags, P0);
    if (st == MHPC_OK) st = rhs_plan(c, flags & ~(uint32_t)MHPC_RHS_GRAVITY, P1);
    if (st != MHPC_OK) return st;
    double *dt_dev = &c->dstatus->dt;
    rhs_launch(c, P0, c->ev);
    HIPC(hipMemcpyAsync(c->h_status1, c->dstatus, sizeof(DevStatus), hipMemcpyDeviceToHost, c->stream));
    launch_time_step(c, c->diag_partial, c->diag_out, DtRule{courant_fact, forces_fact, dt_max, t_left, dt_dev});
    launch_pc_predictor(c, 0.0, dt_dev);
    rhs_launch(c, P1, c->ev + 16);
    launch_pc_corrector(c, 0.0, dt_dev);
    HIPC(hipGetLastError());
    HIPC(ctx_fetch_status(c));
    const int st0 = status_of(c, *c->h_status1);
    const int st1 = status_of(c, *c->h_status);   // last: the next sort starts from this tree's depth
    *dt = c->h_status->dt;
    c->last_dt = *dt;
    rhs_stats(c, P0, c->ev, *c->h_status1, stats0);
    rhs_stats(c, P1, c->ev + 16, *c->h_status, stats1);
    // the corrector moved the particles
    c->keys_valid = c->sorted_valid = c->tree_valid = c->com_valid = false;
    c->nn_valid = c->rho_valid = c->eos_valid = false;
    if (st0 != MHPC_OK) {
        c->last_error = mhpc_status_string(st0);
        return st0;
    }
    return st1;
}

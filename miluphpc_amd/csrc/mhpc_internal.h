// mhpc_internal.h -- context layout and wave64 device helpers shared by the kernels.
//
// Data layout in HBM (one context = one subdomain):
//  * caller-order SoA fields (FP64, plus int32 noi / level / h_iters, uint64 key), the
//    arrays the reference's Particles class holds (particles.cuh:50-264);
//  * key-sorted working copies s* of the fields the tree, gravity and SPH kernels read,
//    so every group of 64 consecutive particles is spatially compact and each octree
//    cell is a contiguous range [cstart, cend] of sorted positions;
//  * the octree in level-major cell order (cell 0 = root), child slots in Lebesgue
//    octant order (tree.cu:768-1082), node ids: particles [0,n), cells n + c;
//  * neighbour lists in chunks of four, entry s of sorted particle i at
//    ((s/4) * cap_pad + i) * 4 + s%4: a wave reading its next four neighbours per lane
//    touches one contiguous 1 KiB row, and the search writes whole 16-byte chunks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/mhpc.h"

#ifndef MHPC_WAVE
#define MHPC_WAVE 64
#endif

#define MHPC_LEVELS 21          // key levels (MAX_LEVEL, parameter.h)
#define MHPC_SORT_BITS 8
#define MHPC_SORT_BUCKETS 256
#define MHPC_SORT_PASSES 8
#define MHPC_SORT_TILE 2048    // keys per scatter block (256 threads x 8)
#define MHPC_TREE_TILE 1024     // positions per tree block (256 threads x 4)
#define MHPC_HTABLE_MAX 4096    // h-iteration factor table size
#define MHPC_MAX_PASSES 64       // upper bound of max_counts
#define MHPC_NUM_EVENTS 32   // two sets of 16: an RHS, and the second RHS of a fused PC-Euler step

// neighbour-walk start table: entries of levels 1 .. K-1 come before level K
#define MHPC_NN_KMAX 8
__host__ __device__ __forceinline__ int64_t nn_tab_off(int K) { return ((1ll << (3 * K)) - 8) / 7; }

// device error flags (sticky word, mapped onto mhpc_status by the host)
enum : uint32_t {
    ERRF_CAP_NODES = 1u,
    ERRF_CAP_NEIGHBOURS = 2u,
    ERRF_DEPTH = 4u,
    ERRF_RHO = 8u,
    ERRF_STACK = 16u,
};

// device counters (uint64) for statistics
enum {
    CNT_CELL_INT = 0,
    CNT_PART_INT = 1,
    CNT_NEIGHBOURS = 2,
    CNT_MAX_NOI = 3,
    CNT_H_COUNTS = 4,
    CNT_TREE_DEPTH = 5,     // deepest level with cells (written by the bottom-up pass)
    CNT_COUNT_ = 8
};
// The counters are kept in CNT_SLOTS copies, one 128-byte line each, and a workgroup adds to
// the copy of its index mod CNT_SLOTS: atomics on one address serialise (about 12 ns each on
// MI355X), so a kernel whose 32k wavefronts each add once spent ~0.4 ms draining them.  The
// host sums (or takes the maximum of) the copies.
#define CNT_SLOTS 16
#define CNT_LINE 16
static_assert(CNT_COUNT_ <= CNT_LINE, "one line per copy");

// The words the host reads after an RHS or a stage, in one device block so that one copy
// brings them all to the pinned host mirror (one round trip per call).
struct DevStatus {
    unsigned long long counters[CNT_SLOTS * CNT_LINE];   // copy s at [s * CNT_LINE, + CNT_COUNT_)
    uint32_t err;
    uint32_t sort_check;                         // set when a partial sort left equal prefixes
    int32_t cell_base[MHPC_LEVELS + 3];          // level-major cell offsets, [23] = total
    uint32_t hlist_count[2 * (MHPC_MAX_PASSES + 1)];
    uint32_t hchanged[MHPC_MAX_PASSES + 1];      // set when a pass changed some h (follows
                                                 // hlist_count: k_h_init zeroes both)
    double dt;                                   // the fused PC-Euler step's time step
};

struct mhpc_ctx {
    mhpc_params p;
    int64_t cap = 0, cap_pad = 0, cap_cells = 0, n = 0;
    int device = 0;
    hipStream_t stream = nullptr;

    // caller-order fields
    double *f[MHPC_F_COUNT_] = {};   // FP64 fields (ids < MHPC_F_NOI) the kernels use
    double *f_own[MHPC_F_COUNT_] = {};   // the context's own arrays (f differs while a field is bound
                                         // to a caller's array, mhpc_bind_field)
    int32_t *noi = nullptr, *level = nullptr, *h_iters = nullptr;
    uint64_t *key = nullptr;         // sort key per particle (caller order)
    uint64_t *lkey = nullptr;        // Lebesgue key per particle (caller order)

    // sort
    uint64_t *skey = nullptr, *keyB = nullptr;   // sorted keys / ping-pong buffer
    int32_t *perm = nullptr, *valB = nullptr;    // sorted position -> particle index
    double4 *xyzm = nullptr;                     // {x,y,z,m} in caller order (written with the keys)
    int32_t *ppar = nullptr;                     // sorted particle -> parent cell (tree build)
    uint32_t *sort_check = nullptr;              // = &dstatus->sort_check
    int32_t *cpar = nullptr;                     // cell -> parent cell (tree build)
    uint32_t *tile_hist = nullptr, *tile_off = nullptr;
    uint32_t *digit_hist = nullptr;              // [passes][256]
    int64_t sort_tiles_cap = 0;
    uint64_t *slkey = nullptr;                   // Lebesgue keys in sorted order

    // sorted working copies; 32-byte records so a neighbour costs two 16-byte loads
    double4 *sxm = nullptr;                      // {x, y, z, m}
    double4 *sv4 = nullptr;                      // {vx, vy, vz, cs}
    double4 *sr4 = nullptr;                      // {rho, p / rho^2, m / rho, p}
    double *sh = nullptr, *sh2 = nullptr;
    double *se = nullptr, *srho = nullptr, *sp = nullptr, *scs = nullptr;
    int32_t *snoi = nullptr, *siters = nullptr;
    uint8_t *sdone = nullptr;

    // tree
    uint8_t *lcp = nullptr;
    uint32_t *tcnt = nullptr, *toff = nullptr;   // [22][tree tiles]
    int64_t tree_tiles_cap = 0;
    int32_t *level_count = nullptr;              // [22]
    int32_t *cell_base = nullptr;                // = dstatus->cell_base
    int32_t *cstart = nullptr, *cend = nullptr, *clevel = nullptr, *cchild = nullptr;
    double *cmass = nullptr, *cx = nullptr, *cy = nullptr, *cz = nullptr;
    double *cquad = nullptr;                     // [6][cap_cells]
    float4 *cquadf = nullptr;                    // packed FP32 quadrupole (xx,xy,xz,yy),(yz,zz,-,-)
    double *cbmin = nullptr, *cbmax = nullptr;   // [3][cap_cells] particle AABB per cell
    double *chmax = nullptr;                     // [cap_cells] an upper bound of the particles' h
                                                 // per cell (neighbour searches prune with it)

    // neighbours
    int32_t *nbr = nullptr;                      // [max_nb][cap_pad]
    int nbr_rows = 0;

    // h-iteration
    double *htable = nullptr;

    // misc
    double *bbox_partial = nullptr;              // [2][bbox blocks]
    double *diag_partial = nullptr, *diag_out = nullptr;
    double *pc_backup = nullptr;                 // predictor-corrector saved state, lazily
    double last_dt = 0.0;                        // last CFL time step (learned-h feature)
    int32_t *hlist[3] = {nullptr, nullptr, nullptr};   // h-iteration: unconverged particles per pass
    uint8_t *sdirty = nullptr;                   // speculative lists: particles to re-list
    int32_t *olist = nullptr;                    // neighbour search: group-mode outliers
    uint32_t *ocount = nullptr;
    double *h0slot = nullptr;                    // pass-0 h of the unconverged particles (list slot)
    int32_t *hhard = nullptr;                    // tail passes: slots left to the split counts
    uint32_t *hlist_count = nullptr;             // = dstatus->hlist_count: list lengths of the
                                                 // passes, then the hard-list lengths of the passes
    uint32_t *hcnt = nullptr;                    // neighbour count per list slot

    // learned smoothing length
    float *mlp_params = nullptr, *mlp_m1 = nullptr, *mlp_m2 = nullptr, *mlp_grad = nullptr;
    float *mlp_norm = nullptr;                   // [18]
    float *mlp_x = nullptr;                      // [cap][8]
    float *mlp_t = nullptr;                      // [cap] targets
    double *mlp_hprev = nullptr;                 // [cap]
    int32_t *mlp_idx = nullptr;                  // [max samples]
    float *mlp_partial = nullptr, *mlp_loss_partial = nullptr, *mlp_loss = nullptr;
    float *mlp_io_x = nullptr, *mlp_io_y = nullptr;   // test-hook staging
    int64_t mlp_io_cap = 0;
    int64_t mlp_max_waves = 0;                   // per-wave gradient partials allocated
    int mlp_adam_step = 0;
    bool mlp_x_valid = false, mlp_norm_set = false;
    double *box = nullptr;                       // [6]
    DevStatus *dstatus = nullptr;                // device block: err, counters, cell_base, list lengths
    DevStatus *h_status = nullptr;               // its pinned host mirror
    DevStatus *h_status1 = nullptr;              // a second pinned copy: the first RHS of a fused
                                                 // PC-Euler step
    uint32_t *err = nullptr;                     // = &dstatus->err
    unsigned long long *counters = nullptr;      // = dstatus->counters
    hipEvent_t ev[MHPC_NUM_EVENTS] = {};
    hipEvent_t sync_ev = nullptr;                // ctx_sync's marker (no timing)

    bool keys_valid = false, sorted_valid = false, tree_valid = false, com_valid = false;
    // neighbour-walk start table: for levels 1..nn_kmax, the tree node covering each
    // geometric cell of that level (Lebesgue prefix order), -1 where there is no particle
    int32_t *nn_start = nullptr;
    int2 *nn_start_rng = nullptr;    // {first sorted particle, count} of each entry's node
    int nn_kmax = 0;
    bool nn_start_valid = false;
    // deepest level that had cells in the last finished call: the bottom-up pass launches
    // the levels below it as one single-workgroup sweep
    int com_depth_hint = MHPC_LEVELS;
    bool nn_valid = false, rho_valid = false, eos_valid = false;
    std::string last_error;
};

// ------------------------------------------------------------------------------------
// host-side launchers (defined in the k_*.hip files)
// waits for the context stream's queued work by spinning on an event (no blocking wait: the
// host thread sees completion within microseconds)
hipError_t ctx_sync(mhpc_ctx *c);
// copies the status block to its host mirror and waits for it (ctx_sync)
hipError_t ctx_fetch_status(mhpc_ctx *c);
void launch_fill_u32(mhpc_ctx *c, uint32_t *p, uint32_t v, int64_t n);
void fill_u32_on(hipStream_t stream, uint32_t *p, uint32_t v, int64_t n);
void launch_fill_u64(mhpc_ctx *c, unsigned long long *p, unsigned long long v, int64_t n);
void launch_bbox(mhpc_ctx *c);
void launch_keys(mhpc_ctx *c);
void launch_sort(mhpc_ctx *c);
void launch_gather(mhpc_ctx *c, bool gravity, bool sph, bool rhs);
void launch_tree(mhpc_ctx *c);
// sph: also the cells' largest h from the key-sorted h (sh must hold the gathered h)
void launch_bottom_up(mhpc_ctx *c, bool gravity, bool sph);
void launch_gravity(mhpc_ctx *c, bool potential);
// walk_events (may be null): two events recorded around the pass-0 group walk kernel
void launch_h_iterate(mhpc_ctx *c, int passes, bool speculative, hipEvent_t *walk_events = nullptr);
void launch_neighbours(mhpc_ctx *c);
void launch_neighbours_fixup(mhpc_ctx *c);
void launch_density(mhpc_ctx *c);
void launch_eos(mhpc_ctx *c);
void launch_density_eos(mhpc_ctx *c, bool publish);   // density with the EOS fused (the RHS path)
void launch_forces(mhpc_ctx *c, bool features);
void launch_scatter_h(mhpc_ctx *c);
void launch_leapfrog_x(mhpc_ctx *c, double dt);
void launch_leapfrog_v(mhpc_ctx *c, double dt);
void launch_neighbours_to_caller(mhpc_ctx *c, int32_t *dst_nnl, int32_t *dst_noi, int max_nb);
int mlp_num_params();
void launch_mlp_features(mhpc_ctx *c, double dt);
void launch_mlp_predict(mhpc_ctx *c);
void launch_mlp_forward_raw(mhpc_ctx *c, int64_t n, const float *X, float *Y);
void launch_mlp_gradient(mhpc_ctx *c, int64_t nb, const float *X, const float *T, const int32_t *idx,
                         float *grad, float *loss);
void launch_mlp_adam(mhpc_ctx *c, float lr, int step);
void launch_mlp_targets(mhpc_ctx *c);
void launch_mlp_sample(mhpc_ctx *c, int64_t nsamples, uint64_t seed);
void launch_mlp_renorm(mhpc_ctx *c);
int pc_backup_fields();
// The CFL time-step rule of setTimeStep (device_predictor_corrector_euler.cu:315-540) from the
// five folded terms q = {min sqrt(h/|a+g|), min h/cs, min h/(cs + 1.2 (alpha cs + beta mu)),
// max |v|, max h}: min(courant q1, forces q0, courant q2, q4 / q3 (when both > 0), dt_max (when
// > 0), t_left (when > 0)).  Host (mhpc_time_step) and device (the fused step) use this one
// function, so both give the same dt.
struct DtRule {
    double courant, forces, dt_max, t_left;
    double *dt_out;   // device: where k_time_step_final writes dt (nullptr: not computed)
};
__host__ __device__ inline double dt_rule(const DtRule &R, double q0, double q1, double q2, double q3, double q4) {
    double t = fmin(R.courant * q1, R.forces * q0);
    t = fmin(t, R.courant * q2);
    if (q3 > 0.0 && q4 > 0.0) t = fmin(t, q4 / q3);
    if (R.dt_max > 0.0) t = fmin(t, R.dt_max);
    if (R.t_left > 0.0) t = fmin(t, R.t_left);
    return t;
}
void launch_pc_predictor(mhpc_ctx *c, double dt, const double *dt_dev = nullptr);
void launch_pc_corrector(mhpc_ctx *c, double dt, const double *dt_dev = nullptr);
void launch_time_step(mhpc_ctx *c, double *partial, double *out, const DtRule &rule);
int diag_num_quantities();
int diag_partial_size();
void launch_diagnostics(mhpc_ctx *c, double *partial, mhpc_diagnostics *out);

// ------------------------------------------------------------------------------------
// device helpers
#if defined(__HIP_DEVICE_COMPILE__) && !defined(__gfx950__)
#error "these kernels are written for gfx950 (CDNA4, wave64)"
#endif

static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ int wave_lane() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return (uint64_t)__ballot(p); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = wave_lane();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }

// orders LDS traffic of one wavefront against the compiler (the hardware keeps one
// wave's LDS operations in order)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// inclusive prefix sum over the wave
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int l = wave_lane();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int w = __shfl_up(v, o);
        if (l >= o) v += w;
    }
    return v;
}

// the calling workgroup's copy of the counters
__device__ __forceinline__ unsigned long long *cnt_slot(unsigned long long *counters) {
    return counters + (blockIdx.x % CNT_SLOTS) * CNT_LINE;
}
// a counter from a fetched status block: the sum of the copies (the maximum for max-counters)
inline unsigned long long cnt_sum(const DevStatus &hs, int q) {
    unsigned long long v = 0;
    for (int s = 0; s < CNT_SLOTS; s++) v += hs.counters[s * CNT_LINE + q];
    return v;
}
inline unsigned long long cnt_max(const DevStatus &hs, int q) {
    unsigned long long v = 0;
    for (int s = 0; s < CNT_SLOTS; s++) v = hs.counters[s * CNT_LINE + q] > v ? hs.counters[s * CNT_LINE + q] : v;
    return v;
}

__device__ __forceinline__ void set_err(uint32_t *err, uint32_t f) { atomicOr(err, f); }

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each with its own
// L2.  Kernels whose neighbouring workgroups share data (tree walks, neighbour gathers over
// key-sorted particles) remap the hardware block id so every XCD works on one contiguous
// chunk of the key order and the shared data stays in one L2.  A bijection on [0, nb).
#define MHPC_NUM_XCD 8
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nb) {
    const int64_t x = bid % MHPC_NUM_XCD, i = bid / MHPC_NUM_XCD;
    const int64_t q = nb / MHPC_NUM_XCD, r = nb % MHPC_NUM_XCD;
    return x * q + (x < r ? x : r) + i;
}

// block-wide exclusive scan of one uint32 per thread (blockDim.x = 256); returns the total.
// Contains two __syncthreads: every thread of the block must call it.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *excl, uint32_t *lds4) {
    const int w = threadIdx.x / 64;
    int inc = wave_incl_scan((int)v);
    if (wave_lane() == 63) lds4[w] = (uint32_t)inc;
    __syncthreads();
    uint32_t off = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k < w) off += lds4[k];
        total += lds4[k];
    }
    *excl = off + (uint32_t)inc - v;
    __syncthreads();
    return total;
}

// exclusive scan of row[0..len) into out[0..len) plus `carry`, by one 256-thread block
__device__ __forceinline__ uint32_t block_scan_row(const uint32_t *row, uint32_t *out, int64_t len,
                                                   uint32_t carry, uint32_t *lds4) {
    const int64_t chunk = 256 * 16;
    for (int64_t c0 = 0; c0 < len; c0 += chunk) {
        uint32_t v[16];
        uint32_t s = 0;
        const int64_t b = c0 + (int64_t)threadIdx.x * 16;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            v[k] = (b + k < len) ? row[b + k] : 0u;
            s += v[k];
        }
        uint32_t excl;
        uint32_t tot = block_excl_scan256(s, &excl, lds4);
        uint32_t run = carry + excl;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (b + k < len) out[b + k] = run;
            run += v[k];
        }
        carry += tot;
    }
    return carry;
}

// exact FP64 arithmetic helpers: the .hip files are compiled with -ffp-contract=off, so
// a*b+c is two roundings, as in the reference build (nvcc -fmad=false, Makefile:33)

// Kahan-compensated FP32 accumulator
struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

// plain FP64 accumulator with the same interface
struct Plain64 {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
};

#define HIP_CHECK_LAUNCH() (void)hipGetLastError()

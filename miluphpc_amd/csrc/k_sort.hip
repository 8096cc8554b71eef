// k_sort.hip -- stable LSD radix sort of (64-bit key, 32-bit index) pairs.
//
// Replaces TreeNS::prepareSorting + HelperNS::sortArray = cub::DeviceRadixSort::SortPairs
// (tree.cu:1404-1413, helper.cu:150-167): same result, a stable sort by key in which
// ties keep ascending index order.  Eight passes of 8-bit digits, one kernel per pass:
//   - one up-front kernel reads the keys once and builds the digit histograms of all
//     eight passes (global digit offsets);
//   - per pass, each workgroup takes the next 4096-key tile by ticket (tickets are
//     handed out in launch order, so every tile a workgroup waits on is already
//     resident), ranks its keys with a ballot multisplit (peers = lanes holding the
//     same digit), publishes its per-digit counts, then walks back over the
//     preceding tiles' published words until it meets an inclusive prefix
//     (decoupled look-back), publishes its own inclusive prefix, stages the tile in
//     LDS in digit order and writes each digit run contiguously.
// A look-back word is [63:60] pass + 1, [59:58] kind (1 tile count, 2 inclusive
// prefix), [57:0] count; the status array is zeroed once per sort, so a word left by an
// earlier pass reads as "not yet published".  Every wait is bounded (SORT_SPIN_MAX
// polls); running out sets ERRF_SPIN and the call reports an error.
// HBM bytes per key: 8 for the histograms, then per pass 12 read (8 on the first pass,
// whose values are the indices) and 12 written.
#include "mhpc_internal.h"

#define SORT_BLOCK 256
#define SORT_ITEMS (MHPC_SORT_TILE / SORT_BLOCK)   // 16
#define SORT_WAVES (SORT_BLOCK / 64)

__global__ __launch_bounds__(SORT_BLOCK) void k_sort_hist_all(const uint64_t *__restrict__ keys,
                                                              int64_t n,
                                                              uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[MHPC_SORT_PASSES * MHPC_SORT_BUCKETS];
    for (int t = threadIdx.x; t < MHPC_SORT_PASSES * MHPC_SORT_BUCKETS; t += SORT_BLOCK) h[t] = 0;
    __syncthreads();
    int64_t i = blockIdx.x * (int64_t)SORT_BLOCK + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * SORT_BLOCK;
    for (; i < n; i += stride) {
        const uint64_t k = keys[i];
#pragma unroll
        for (int p = 0; p < MHPC_SORT_PASSES; p++)
            atomicAdd(&h[p * MHPC_SORT_BUCKETS + (int)((k >> (MHPC_SORT_BITS * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < MHPC_SORT_PASSES * MHPC_SORT_BUCKETS; t += SORT_BLOCK)
        if (h[t]) atomicAdd(&hist[t], h[t]);
}

#define SORT_SPIN_MAX (1 << 22)
#define SORT_CNT_MASK ((1ull << 58) - 1ull)

__device__ __forceinline__ uint64_t sort_word(int pass, uint64_t kind, uint64_t cnt) {
    return ((uint64_t)(pass + 1) << 60) | (kind << 58) | cnt;
}

__global__ __launch_bounds__(SORT_BLOCK) void k_sort_onesweep(const uint64_t *__restrict__ kin,
                                                              const int32_t *__restrict__ vin,
                                                              uint64_t *__restrict__ kout,
                                                              int32_t *__restrict__ vout, int64_t n,
                                                              int pass,
                                                              const uint32_t *__restrict__ digit_hist,
                                                              uint64_t *status, uint32_t *ticket,
                                                              uint32_t *err) {
    __shared__ uint64_t skeys[MHPC_SORT_TILE];
    __shared__ int32_t svals[MHPC_SORT_TILE];
    __shared__ uint32_t wcnt[SORT_WAVES][MHPC_SORT_BUCKETS];
    __shared__ uint32_t dstart[MHPC_SORT_BUCKETS];
    __shared__ uint64_t goff[MHPC_SORT_BUCKETS];
    __shared__ uint32_t lds4[SORT_WAVES];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x, w = tid / 64, lane = wave_lane();
    const int shift = MHPC_SORT_BITS * pass;
    if (tid == 0) s_tile = atomicAdd(&ticket[pass], 1u);
#pragma unroll
    for (int k = 0; k < SORT_WAVES; k++) wcnt[k][tid] = 0;
    __syncthreads();
    const int64_t tile = (int64_t)s_tile;
    const int64_t base = tile * MHPC_SORT_TILE;

    uint64_t kr[SORT_ITEMS];
    int32_t vr[SORT_ITEMS];
    uint32_t rank[SORT_ITEMS];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int s = 0; s < SORT_ITEMS; s++) {
        const int64_t idx = base + (int64_t)w * (64 * SORT_ITEMS) + s * 64 + lane;
        const bool valid = idx < n;
        const uint64_t key = valid ? kin[idx] : 0ull;
        const int32_t val = valid ? (vin ? vin[idx] : (int32_t)idx) : 0;
        const uint32_t d = (uint32_t)((key >> shift) & 255u);
        uint64_t peers = wave_ballot(valid);
#pragma unroll
        for (int b = 0; b < MHPC_SORT_BITS; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = wave_ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t r = (uint32_t)popc64(peers & lt);
        uint32_t prior = 0;
        if (valid) prior = wcnt[w][d];
        wave_sync();
        if (valid && r == 0) wcnt[w][d] = prior + (uint32_t)popc64(peers);
        wave_sync();
        kr[s] = key;
        vr[s] = val;
        rank[s] = prior + r;
    }
    __syncthreads();
    {
        // thread tid owns digit tid
        const uint32_t c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
        const uint32_t tot = c0 + c1 + c2 + c3;
        uint64_t *my = status + tile * MHPC_SORT_BUCKETS + tid;
        __hip_atomic_store(my, sort_word(pass, tile == 0 ? 2u : 1u, tot), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        wcnt[0][tid] = 0;
        wcnt[1][tid] = c0;
        wcnt[2][tid] = c0 + c1;
        wcnt[3][tid] = c0 + c1 + c2;
        uint32_t excl;
        block_excl_scan256(tot, &excl, lds4);
        dstart[tid] = excl;
        uint32_t dex;
        block_excl_scan256(digit_hist[pass * MHPC_SORT_BUCKETS + tid], &dex, lds4);
        // decoupled look-back over the preceding tiles for digit tid
        uint64_t before = 0;
        int64_t j = tile - 1;
        int polls = 0;
        while (j >= 0) {
            const uint64_t v = __hip_atomic_load(status + j * MHPC_SORT_BUCKETS + tid, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            if ((v >> 60) != (uint64_t)(pass + 1)) {
                if (++polls > SORT_SPIN_MAX) {
                    set_err(err, ERRF_SPIN);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            before += v & SORT_CNT_MASK;
            if (((v >> 58) & 3u) == 2u) break;
            j--;
        }
        if (tile > 0)
            __hip_atomic_store(my, sort_word(pass, 2u, before + tot), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        goff[tid] = (uint64_t)dex + before;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SORT_ITEMS; s++) {
        const int64_t idx = base + (int64_t)w * (64 * SORT_ITEMS) + s * 64 + lane;
        if (idx < n) {
            const uint32_t d = (uint32_t)((kr[s] >> shift) & 255u);
            const uint32_t pos = dstart[d] + wcnt[w][d] + rank[s];
            skeys[pos] = kr[s];
            svals[pos] = vr[s];
        }
    }
    __syncthreads();
    const int64_t rem = n - base;
    const int cnt = rem < MHPC_SORT_TILE ? (int)rem : MHPC_SORT_TILE;
    for (int t = tid; t < cnt; t += SORT_BLOCK) {
        const uint64_t key = skeys[t];
        const uint32_t d = (uint32_t)((key >> shift) & 255u);
        const uint64_t o = goff[d] + (uint64_t)((uint32_t)t - dstart[d]);
        if (o < (uint64_t)n) {
            kout[o] = key;
            vout[o] = svals[t];
        }
    }
}

void launch_sort(mhpc_ctx *c) {
    const int64_t n = c->n;
    if (n <= 0) return;
    const int64_t ntiles = ceil_div(n, MHPC_SORT_TILE);
    launch_fill_u32(c, c->digit_hist, 0u, MHPC_SORT_PASSES * MHPC_SORT_BUCKETS);
    launch_fill_u32(c, c->sort_ticket, 0u, MHPC_SORT_PASSES);
    launch_fill_u64(c, (unsigned long long *)c->sort_status, 0ull, ntiles * MHPC_SORT_BUCKETS);
    int hb = (int)(ceil_div(n, SORT_BLOCK * 16) < 1024 ? ceil_div(n, SORT_BLOCK * 16) : 1024);
    hipLaunchKernelGGL(k_sort_hist_all, dim3(hb), dim3(SORT_BLOCK), 0, c->stream, c->key, n,
                       c->digit_hist);
    const uint64_t *kin = c->key;
    const int32_t *vin = nullptr;
    for (int pass = 0; pass < MHPC_SORT_PASSES; pass++) {
        uint64_t *kout = (pass & 1) ? c->skey : c->keyB;
        int32_t *vout = (pass & 1) ? c->perm : c->valB;
        hipLaunchKernelGGL(k_sort_onesweep, dim3((unsigned)ntiles), dim3(SORT_BLOCK), 0, c->stream, kin,
                           vin, kout, vout, n, pass, c->digit_hist, c->sort_status, c->sort_ticket,
                           c->err);
        kin = kout;
        vin = vout;
    }
}

// k_sort.hip -- stable LSD radix sort of (64-bit key, 32-bit index) pairs.
//
// Replaces TreeNS::prepareSorting + HelperNS::sortArray = cub::DeviceRadixSort::SortPairs
// (tree.cu:1404-1413, helper.cu:150-167): same result, a stable sort by key in which
// ties keep ascending index order.  Eight passes of 8-bit digits.  Per pass:
//   count   -- digit histogram of each 2048-key tile (LDS atomics);
//   scan    -- one workgroup per digit scans the tile histograms (digit-major), adding
//              the digit's global offset from the one-read histogram of all passes;
//   scatter -- each wave ranks its 8 slots of 64 keys with a ballot multisplit (peers
//              = lanes holding the same digit), the block stages the tile in LDS in
//              digit order and writes each digit run contiguously.
// HBM bytes per pass: count reads 8 B/key, scatter reads 12 and writes 12.
#include <mutex>

#include "mhpc_internal.h"

#define SORT_BLOCK 256
#define SORT_ITEMS (MHPC_SORT_TILE / SORT_BLOCK)   // 16
#define SORT_WAVES (SORT_BLOCK / 64)

// digit histograms of passes [p_lo, p_hi) over all keys (one read of the keys)
__global__ __launch_bounds__(SORT_BLOCK) void k_sort_hist_all(const uint64_t *__restrict__ keys,
                                                              int64_t n,
                                                              uint32_t *__restrict__ hist,
                                                              uint32_t *__restrict__ check_flag, int p_lo,
                                                              int p_hi) {
    __shared__ uint32_t h[MHPC_SORT_PASSES * MHPC_SORT_BUCKETS];
    if (blockIdx.x == 0 && threadIdx.x == 0) *check_flag = 0u;
    for (int t = threadIdx.x; t < MHPC_SORT_PASSES * MHPC_SORT_BUCKETS; t += SORT_BLOCK) h[t] = 0;
    __syncthreads();
    int64_t i = blockIdx.x * (int64_t)SORT_BLOCK + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * SORT_BLOCK;
    for (; i < n; i += stride) {
        const uint64_t k = keys[i];
#pragma unroll
        for (int p = 0; p < MHPC_SORT_PASSES; p++)
            if (p >= p_lo && p < p_hi)
                atomicAdd(&h[p * MHPC_SORT_BUCKETS + (int)((k >> (MHPC_SORT_BITS * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < MHPC_SORT_PASSES * MHPC_SORT_BUCKETS; t += SORT_BLOCK)
        if (h[t]) atomicAdd(&hist[t], h[t]);
}

__global__ __launch_bounds__(SORT_BLOCK) void k_sort_count(const uint64_t *__restrict__ keys,
                                                           int64_t n, int shift,
                                                           uint32_t *__restrict__ tile_hist,
                                                           int64_t ntiles) {
    __shared__ uint32_t h[MHPC_SORT_BUCKETS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * MHPC_SORT_TILE;
#pragma unroll 4
    for (int s = 0; s < SORT_ITEMS; s++) {
        int64_t idx = base + s * SORT_BLOCK + threadIdx.x;
        if (idx < n) atomicAdd(&h[(int)((keys[idx] >> shift) & 255u)], 1u);
    }
    __syncthreads();
    tile_hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(SORT_BLOCK) void k_sort_scan(const uint32_t *__restrict__ tile_hist,
                                                          uint32_t *__restrict__ tile_off,
                                                          int64_t ntiles,
                                                          const uint32_t *__restrict__ digit_hist) {
    __shared__ uint32_t lds4[SORT_WAVES];
    const int d = blockIdx.x;
    // global offset of digit d = number of keys with a smaller digit in this pass
    uint32_t part = (threadIdx.x < (unsigned)d) ? digit_hist[threadIdx.x] : 0u;
    uint32_t dummy;
    uint32_t carry = block_excl_scan256(part, &dummy, lds4);
    block_scan_row(tile_hist + (int64_t)d * ntiles, tile_off + (int64_t)d * ntiles, ntiles, carry,
                   lds4);
}

__global__ __launch_bounds__(SORT_BLOCK) void k_sort_scatter(const uint64_t *__restrict__ kin,
                                                             const int32_t *__restrict__ vin,
                                                             uint64_t *__restrict__ kout,
                                                             int32_t *__restrict__ vout, int64_t n,
                                                             int shift,
                                                             const uint32_t *__restrict__ tile_off,
                                                             int64_t ntiles) {
    __shared__ uint64_t skeys[MHPC_SORT_TILE];
    __shared__ int32_t svals[MHPC_SORT_TILE];
    __shared__ uint32_t wcnt[SORT_WAVES][MHPC_SORT_BUCKETS];
    __shared__ uint32_t dstart[MHPC_SORT_BUCKETS];
    __shared__ uint32_t goff[MHPC_SORT_BUCKETS];
    __shared__ uint32_t lds4[SORT_WAVES];

    const int tid = threadIdx.x, w = tid / 64, lane = wave_lane();
    const int64_t base = (int64_t)blockIdx.x * MHPC_SORT_TILE;
#pragma unroll
    for (int k = 0; k < SORT_WAVES; k++) wcnt[k][tid] = 0;
    goff[tid] = tile_off[(int64_t)tid * ntiles + blockIdx.x];
    __syncthreads();

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
        const uint32_t c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
        const uint32_t tot = c0 + c1 + c2 + c3;
        __syncthreads();
        wcnt[0][tid] = 0;
        wcnt[1][tid] = c0;
        wcnt[2][tid] = c0 + c1;
        wcnt[3][tid] = c0 + c1 + c2;
        uint32_t excl;
        block_excl_scan256(tot, &excl, lds4);
        dstart[tid] = excl;
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
        const uint32_t o = goff[d] + ((uint32_t)t - dstart[d]);
        if (o < (uint64_t)n) {
            kout[o] = key;
            vout[o] = svals[t];
        }
    }
}

// the buffers of the LSD passes: input keys (values = positions), the even passes' outputs,
// the odd passes' outputs (the last pass's, so the result) and the per-pass tables
struct SortBufs {
    const uint64_t *kin;
    uint64_t *keyB;
    int32_t *valB;
    uint64_t *kout;
    int32_t *vout;
    uint32_t *tile_hist, *tile_off, *digit_hist;
};
static void sort_passes(hipStream_t stream, int64_t n, int first, int64_t ntiles, const SortBufs &b);
static SortBufs ctx_bufs(mhpc_ctx *c) {
    return SortBufs{c->key, c->keyB, c->valB, c->skey, c->perm, c->tile_hist, c->tile_off, c->digit_hist};
}

// Adjacent sorted keys equal in the bits the passes sorted (>= shift): their order in the low
// bits is not established.  Sets *flag.
__global__ __launch_bounds__(256) void k_sort_check(const uint64_t *__restrict__ skey, int64_t n, int shift,
                                                    uint32_t *__restrict__ flag) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= 1 && k < n && (skey[k] >> shift) == (skey[k - 1] >> shift)) *flag = 1u;
}

// The sort only needs the key levels at which particles are distinct.  The previous tree of
// this context had leaves no deeper than its deepest cell level + 1; with one level of margin
// the passes covering levels 1 .. depth + 2 usually suffice (4 of 8 on a 128^3 lattice).  Those
// passes run first; a check over the result finds any adjacent keys still equal in the sorted
// bits, the host reads it, and only then the full sort runs (from the untouched input keys).
// The result is the full sort's either way: distinct sorted prefixes fix the whole order.
void launch_sort(mhpc_ctx *c) {
    const int64_t n = c->n;
    if (n <= 0) return;
    const int64_t ntiles = ceil_div(n, MHPC_SORT_TILE);
    launch_fill_u32(c, c->digit_hist, 0u, MHPC_SORT_PASSES * MHPC_SORT_BUCKETS);
    int hb = (int)(ceil_div(n, SORT_BLOCK * 16) < 1024 ? ceil_div(n, SORT_BLOCK * 16) : 1024);
    const int levels = c->com_depth_hint + 2;
    const int first = levels < MHPC_LEVELS ? 3 * (MHPC_LEVELS - levels) / MHPC_SORT_BITS : 0;
    hipLaunchKernelGGL(k_sort_hist_all, dim3(hb), dim3(SORT_BLOCK), 0, c->stream, c->key, n,
                       c->digit_hist, c->sort_check, first, MHPC_SORT_PASSES);
    if (first > 0) {
        sort_passes(c->stream, n, first, ntiles, ctx_bufs(c));
        hipLaunchKernelGGL(k_sort_check, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c->stream, c->skey, n,
                           MHPC_SORT_BITS * first, c->sort_check);
        if (ctx_fetch_status(c) == hipSuccess && c->h_status->sort_check == 0u) return;
        // the full sort also needs the low passes' histograms
        hipLaunchKernelGGL(k_sort_hist_all, dim3(hb), dim3(SORT_BLOCK), 0, c->stream, c->key, n,
                           c->digit_hist, c->sort_check, 0, first);
    }
    sort_passes(c->stream, n, 0, ntiles, ctx_bufs(c));
}

// LSD passes first .. MHPC_SORT_PASSES - 1 from the input keys; the last (odd) pass writes
// kout / vout
static void sort_passes(hipStream_t stream, int64_t n, int first, int64_t ntiles, const SortBufs &b) {
    const uint64_t *kin = b.kin;
    const int32_t *vin = nullptr;
    for (int pass = first; pass < MHPC_SORT_PASSES; pass++) {
        uint64_t *kout = (pass & 1) ? b.kout : b.keyB;
        int32_t *vout = (pass & 1) ? b.vout : b.valB;
        const int shift = MHPC_SORT_BITS * pass;
        hipLaunchKernelGGL(k_sort_count, dim3((unsigned)ntiles), dim3(SORT_BLOCK), 0, stream, kin, n, shift,
                           b.tile_hist, ntiles);
        hipLaunchKernelGGL(k_sort_scan, dim3(MHPC_SORT_BUCKETS), dim3(SORT_BLOCK), 0, stream, b.tile_hist,
                           b.tile_off, ntiles, b.digit_hist + pass * MHPC_SORT_BUCKETS);
        hipLaunchKernelGGL(k_sort_scatter, dim3((unsigned)ntiles), dim3(SORT_BLOCK), 0, stream, kin, vin, kout,
                           vout, n, shift, b.tile_off, ntiles);
        kin = kout;
        vin = vout;
    }
}

// ------------------------------------------------------------------------------------
// The key sort without allocation (mhpc_sort_keys_noalloc): the same passes on caller-owned
// device arrays, with every buffer beyond the inputs and outputs carved from `scratch`, one
// static array per GPU that lives until the process ends.  Layout (byte offsets, each
// 256-aligned): even passes' keys and values, per-tile histograms and offsets, the all-pass
// digit histogram, the check word k_sort_hist_all clears.
#define NA_TILES (MHPC_NOALLOC_SORT_MAX / MHPC_SORT_TILE)
#define NA_KEYB 0
#define NA_VALB (NA_KEYB + 8 * (size_t)MHPC_NOALLOC_SORT_MAX)
#define NA_THIST (NA_VALB + 4 * (size_t)MHPC_NOALLOC_SORT_MAX)
#define NA_TOFF (NA_THIST + 4 * (size_t)MHPC_SORT_BUCKETS * NA_TILES)
#define NA_DHIST (NA_TOFF + 4 * (size_t)MHPC_SORT_BUCKETS * NA_TILES)
#define NA_CHECK (NA_DHIST + 4 * (size_t)MHPC_SORT_PASSES * MHPC_SORT_BUCKETS)
#define NA_BYTES (NA_CHECK + 256)
static_assert(MHPC_NOALLOC_SORT_MAX % MHPC_SORT_TILE == 0, "whole tiles");
static_assert(NA_VALB % 256 == 0 && NA_THIST % 256 == 0 && NA_TOFF % 256 == 0 && NA_DHIST % 256 == 0 &&
                  NA_CHECK % 256 == 0, "256-aligned parts");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

#define NA_MAX_DEVICES 64
static std::mutex na_mutex;
static unsigned char *na_ptr[NA_MAX_DEVICES];   // device address of `scratch` on each GPU

// `scratch`'s address on the current device, looked up once per device (the host-side name is
// not a device address)
static int na_scratch(unsigned char **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return MHPC_E_HIP;
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MHPC_E_HIP;
        na_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = na_ptr[dev];
    return MHPC_OK;
}

extern "C" int mhpc_sort_keys_noalloc(const uint64_t *keys, int64_t n, uint64_t *sorted_keys, int32_t *sorted,
                                      void *stream) {
    if (n < 0 || (n > 0 && (!keys || !sorted_keys || !sorted))) return MHPC_E_ARG;
    if (n > MHPC_NOALLOC_SORT_MAX) return MHPC_E_SCRATCH;
    if (n == 0) return MHPC_OK;
    unsigned char *ptr = nullptr;
    const int st = na_scratch(&ptr);
    if (st != MHPC_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t *digit_hist = reinterpret_cast<uint32_t *>(ptr + NA_DHIST);
    uint32_t *check = reinterpret_cast<uint32_t *>(ptr + NA_CHECK);
    // the digit histogram accumulates with atomics: cleared first (the previous call's counts
    // are still there)
    fill_u32_on(s, digit_hist, 0u, MHPC_SORT_PASSES * MHPC_SORT_BUCKETS);
    const int64_t ntiles = ceil_div(n, MHPC_SORT_TILE);
    const int hb = (int)(ceil_div(n, SORT_BLOCK * 16) < 1024 ? ceil_div(n, SORT_BLOCK * 16) : 1024);
    hipLaunchKernelGGL(k_sort_hist_all, dim3(hb), dim3(SORT_BLOCK), 0, s, keys, n, digit_hist, check, 0,
                       MHPC_SORT_PASSES);
    const SortBufs b{keys,
                     reinterpret_cast<uint64_t *>(ptr + NA_KEYB),
                     reinterpret_cast<int32_t *>(ptr + NA_VALB),
                     sorted_keys,
                     sorted,
                     reinterpret_cast<uint32_t *>(ptr + NA_THIST),
                     reinterpret_cast<uint32_t *>(ptr + NA_TOFF),
                     digit_hist};
    sort_passes(s, n, 0, ntiles, b);
    return hipGetLastError() == hipSuccess ? MHPC_OK : MHPC_E_HIP;
}

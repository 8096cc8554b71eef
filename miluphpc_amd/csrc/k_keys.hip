// k_keys.hip -- bounding box, space-filling-curve keys, gathers into key order, and the
// elementwise integrator kernels.  HBM-bound: one coalesced pass per kernel.
#include "mhpc_internal.h"

// ------------------------------------------------------------------------------------
// fills (used instead of hipMemsetAsync so the RHS stays capturable in a HIP graph)
__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static int grid_for(int64_t n, int block, int max_blocks) {
    int64_t g = ceil_div(n, block);
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return (int)g;
}

void fill_u32_on(hipStream_t stream, uint32_t *p, uint32_t v, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, stream, p, v, n);
}

void launch_fill_u32(mhpc_ctx *c, uint32_t *p, uint32_t v, int64_t n) { fill_u32_on(c->stream, p, v, n); }

void launch_fill_u64(mhpc_ctx *c, unsigned long long *p, unsigned long long v, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->stream, p, v, n);
}

// ------------------------------------------------------------------------------------
// a1 bounding box (tree.cu:302-437): min / max over every coordinate of every particle
// and 0 (the box is reset to 0, device_rhs.cu:25-32), then the same cube on all axes
// (CUBIC_DOMAINS, tree.cu:397-431).  Two launches, no atomics: per-block partials, then
// one block folds them.  min/max are exact, so the result does not depend on the order.
#define BBOX_BLOCKS_MAX 1024

__global__ __launch_bounds__(256) void k_bbox_partial(const double *__restrict__ x,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ z, int64_t n,
                                                      double *__restrict__ partial) {
    double lo = 0.0, hi = 0.0;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double a = x[i], b = y[i], d = z[i];
        lo = fmin(lo, fmin(a, fmin(b, d)));
        hi = fmax(hi, fmax(a, fmax(b, d)));
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    __shared__ double slo[4], shi[4];
    const int w = threadIdx.x / 64;
    if (wave_lane() == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < (int)(blockDim.x / 64); k++) {
            lo = fmin(lo, slo[k]);
            hi = fmax(hi, shi[k]);
        }
        partial[2 * blockIdx.x] = lo;
        partial[2 * blockIdx.x + 1] = hi;
    }
}

__global__ __launch_bounds__(256) void k_bbox_final(const double *__restrict__ partial, int nb,
                                                    double *__restrict__ box) {
    double lo = 0.0, hi = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        lo = fmin(lo, partial[2 * b]);
        hi = fmax(hi, partial[2 * b + 1]);
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    __shared__ double slo[4], shi[4];
    const int w = threadIdx.x / 64;
    if (wave_lane() == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < (int)(blockDim.x / 64); k++) {
            lo = fmin(lo, slo[k]);
            hi = fmax(hi, shi[k]);
        }
        box[0] = lo; box[1] = hi;
        box[2] = lo; box[3] = hi;
        box[4] = lo; box[5] = hi;
    }
}

void launch_bbox(mhpc_ctx *c) {
    int nb = grid_for(c->n, 256 * 8, BBOX_BLOCKS_MAX);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, c->stream, c->f[MHPC_F_X],
                       c->f[MHPC_F_Y], c->f[MHPC_F_Z], c->n, c->bbox_partial);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, c->stream, c->bbox_partial, nb, c->box);
}

// ------------------------------------------------------------------------------------
// a2 keys: Tree::getParticleKey (tree.cu:157-240), 21 levels of FP64 bisection, the lower
// half of an axis sets its bit (x 1, y 2, z 4); Hilbert via lebesgue2hilbert
// (tree.cu:40-53) with the reference's DirTable / HilbertTable (tree.cuh:52-118); the
// kernel fuses them in LDS into one byte per (state, octant): digit | next state << 3.
__constant__ unsigned char c_dir_table[12][8] = {
    {8, 10, 3, 3, 4, 5, 4, 5},   {2, 2, 11, 9, 4, 5, 4, 5},  {7, 6, 7, 6, 8, 10, 1, 1},
    {7, 6, 7, 6, 0, 0, 11, 9},   {0, 8, 1, 11, 6, 8, 6, 11}, {10, 0, 9, 1, 10, 7, 9, 7},
    {10, 4, 9, 4, 10, 2, 9, 3},  {5, 8, 5, 11, 2, 8, 3, 11}, {4, 9, 0, 0, 7, 9, 2, 2},
    {1, 1, 8, 5, 3, 3, 8, 6},    {11, 5, 0, 0, 11, 6, 2, 2}, {1, 1, 4, 10, 3, 3, 7, 10}};
__constant__ unsigned char c_hilbert_table[12][8] = {
    {0, 7, 3, 4, 1, 6, 2, 5}, {4, 3, 7, 0, 5, 2, 6, 1}, {6, 1, 5, 2, 7, 0, 4, 3},
    {2, 5, 1, 6, 3, 4, 0, 7}, {0, 1, 7, 6, 3, 2, 4, 5}, {6, 7, 1, 0, 5, 4, 2, 3},
    {2, 3, 5, 4, 1, 0, 6, 7}, {4, 5, 3, 2, 7, 6, 0, 1}, {0, 3, 1, 2, 7, 4, 6, 5},
    {2, 1, 3, 0, 5, 6, 4, 7}, {4, 7, 5, 6, 3, 0, 2, 1}, {6, 5, 7, 4, 1, 2, 0, 3}};

__device__ __forceinline__ uint64_t lebesgue_key(double x, double y, double z,
                                                 const double *__restrict__ box) {
    double mnx = box[0], mxx = box[1], mny = box[2], mxy = box[3], mnz = box[4], mxz = box[5];
    uint64_t key = 0;
#pragma unroll
    for (int level = 0; level < MHPC_LEVELS; level++) {
        uint64_t son = 0;
        double m = 0.5 * (mnx + mxx);
        if (x < m) { son += 1; mxx = m; } else { mnx = m; }
        m = 0.5 * (mny + mxy);
        if (y < m) { son += 2; mxy = m; } else { mny = m; }
        m = 0.5 * (mnz + mxz);
        if (z < m) { son += 4; mxz = m; } else { mnz = m; }
        key |= son << (3 * (MHPC_LEVELS - 1 - level));
    }
    return key;
}

// Also writes each particle's {x,y,z,m} as one 32-byte record in caller order, so that the
// gather into key order reads one record per particle instead of four scattered doubles.
__global__ __launch_bounds__(256) void k_keys(const double *__restrict__ x,
                                              const double *__restrict__ y,
                                              const double *__restrict__ z,
                                              const double *__restrict__ m, int64_t n,
                                              const double *__restrict__ box, int curve,
                                              uint64_t *__restrict__ key,
                                              uint64_t *__restrict__ lkey,
                                              double4 *__restrict__ xyzm) {
    __shared__ unsigned char fsm[96];
    if (threadIdx.x < 96)
        fsm[threadIdx.x] = (unsigned char)((&c_hilbert_table[0][0])[threadIdx.x] |
                                           ((&c_dir_table[0][0])[threadIdx.x] << 3));
    __syncthreads();
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i], yi = y[i], zi = z[i];
    xyzm[i] = make_double4(xi, yi, zi, m[i]);
    uint64_t leb = lebesgue_key(xi, yi, zi, box);
    if (curve == MHPC_CURVE_HILBERT) {
        uint64_t h = 0;
        int dir = 0;
#pragma unroll
        for (int lvl = MHPC_LEVELS; lvl > 0; lvl--) {
            int cell = (int)((leb >> (3 * (lvl - 1))) & 7u);
            unsigned t = fsm[dir * 8 + cell];
            h = (h << 3) + (t & 7u);
            dir = (int)(t >> 3);
        }
        key[i] = h;
        lkey[i] = leb;
    } else {
        key[i] = leb;
    }
}

void launch_keys(mhpc_ctx *c) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_keys, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                       c->f[MHPC_F_X], c->f[MHPC_F_Y], c->f[MHPC_F_Z], c->f[MHPC_F_MASS], c->n, c->box,
                       c->p.curve, c->key, c->lkey, c->xyzm);
}

// ------------------------------------------------------------------------------------
// gather into key order: the packed records {x,y,z,m} (from the caller-order copy the key
// kernel wrote) and for SPH {vx,vy,vz,0}, plus the scalar fields a phase reads.  Writes are
// coalesced; reads follow perm.
struct GatherArgs {
    const int32_t *perm;
    int64_t n;
    const double4 *xyzm;
    double4 *sxm;
    const double *vx, *vy, *vz;
    double4 *sv4;               // null when SPH is off
    int nf;
    const double *src[6];
    double *dst[6];
    const uint64_t *lkey;
    uint64_t *slkey;
};

__global__ __launch_bounds__(256) void k_gather(GatherArgs a) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const int32_t i = a.perm[k];
    if (i < 0 || (int64_t)i >= a.n) return;
    a.sxm[k] = a.xyzm[i];
    if (a.sv4) a.sv4[k] = make_double4(a.vx[i], a.vy[i], a.vz[i], 0.0);
#pragma unroll
    for (int f = 0; f < 6; f++)
        if (f < a.nf) a.dst[f][k] = a.src[f][i];
    if (a.slkey) a.slkey[k] = a.lkey[i];
}

// rhs: the RHS's own gather, after which the density kernel recomputes rho, and p and cs
// (with the EOS) -- their previous values are gathered only when the EOS reads them (a sound
// speed from the previous p, or an EOS that keeps the previous sound speed)
void launch_gather(mhpc_ctx *c, bool gravity, bool sph, bool rhs) {
    if (c->n <= 0) return;
    GatherArgs a = {};
    a.perm = c->perm;
    a.n = c->n;
    a.xyzm = c->xyzm;
    a.sxm = c->sxm;
    int nf = 0;
    auto add = [&](int field, double *dst) {
        a.src[nf] = c->f[field];
        a.dst[nf] = dst;
        nf++;
    };
    if (sph) {
        a.vx = c->f[MHPC_F_VX]; a.vy = c->f[MHPC_F_VY]; a.vz = c->f[MHPC_F_VZ];
        a.sv4 = c->sv4;
        add(MHPC_F_SML, c->sh);
        add(MHPC_F_E, c->se);
        const bool eos_fresh =
            !c->p.cs_before_p && (c->p.eos == MHPC_EOS_IDEAL_GAS || c->p.eos == MHPC_EOS_POLYTROPIC);
        if (!rhs || !eos_fresh) {
            add(MHPC_F_P, c->sp);
            add(MHPC_F_CS, c->scs);
        }
        if (!rhs) add(MHPC_F_RHO, c->srho);
    }
    (void)gravity;
    a.nf = nf;
    if (c->p.curve == MHPC_CURVE_HILBERT) {
        a.lkey = c->lkey;
        a.slkey = c->slkey;
    }
    hipLaunchKernelGGL(k_gather, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream, a);
}

// ------------------------------------------------------------------------------------
// a17 leapfrog (device_leapfrog.cu:4-58), caller order, FP64
__global__ __launch_bounds__(256) void k_leapfrog_x(int64_t n, double dt, double *x, double *y,
                                                    double *z, const double *vx, const double *vy,
                                                    const double *vz, const double *gx,
                                                    const double *gy, const double *gz,
                                                    double *gxo, double *gyo, double *gzo) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += dt * (vx[i] + 0.5 * dt * gx[i]);
    gxo[i] = gx[i];
    y[i] += dt * (vy[i] + 0.5 * dt * gy[i]);
    gyo[i] = gy[i];
    z[i] += dt * (vz[i] + 0.5 * dt * gz[i]);
    gzo[i] = gz[i];
}

__global__ __launch_bounds__(256) void k_leapfrog_v(int64_t n, double dt, double *vx, double *vy,
                                                    double *vz, const double *gx,
                                                    const double *gy, const double *gz,
                                                    const double *gxo, const double *gyo,
                                                    const double *gzo) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] += 0.5 * dt * (gx[i] + gxo[i]);
    vy[i] += 0.5 * dt * (gy[i] + gyo[i]);
    vz[i] += 0.5 * dt * (gz[i] + gzo[i]);
}

void launch_leapfrog_x(mhpc_ctx *c, double dt) {
    if (c->n <= 0) return;
    double **f = c->f;
    hipLaunchKernelGGL(k_leapfrog_x, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                       c->n, dt, f[MHPC_F_X], f[MHPC_F_Y], f[MHPC_F_Z], f[MHPC_F_VX], f[MHPC_F_VY],
                       f[MHPC_F_VZ], f[MHPC_F_G_AX], f[MHPC_F_G_AY], f[MHPC_F_G_AZ],
                       f[MHPC_F_G_AX_OLD], f[MHPC_F_G_AY_OLD], f[MHPC_F_G_AZ_OLD]);
}

void launch_leapfrog_v(mhpc_ctx *c, double dt) {
    if (c->n <= 0) return;
    double **f = c->f;
    hipLaunchKernelGGL(k_leapfrog_v, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                       c->n, dt, f[MHPC_F_VX], f[MHPC_F_VY], f[MHPC_F_VZ], f[MHPC_F_G_AX],
                       f[MHPC_F_G_AY], f[MHPC_F_G_AZ], f[MHPC_F_G_AX_OLD], f[MHPC_F_G_AY_OLD],
                       f[MHPC_F_G_AZ_OLD]);
}

// k_tree.hip -- octree from sorted keys, and the bottom-up centre-of-mass pass.
//
// The reference inserts particles with locks (TreeNS::Kernel::buildTree, tree.cu:768-1082)
// and gets a scheduling-dependent numbering of a deterministic structure: a cell for
// every key prefix (level L, first L octant digits) shared by at least two particles,
// each leaf at the first level where its prefix is unique.  With the particles in key
// order every such cell is a contiguous run [start, end].  Let l[k] be the number of
// leading octant digits that sorted keys k and k+1 share.  A level-L cell starts at k iff
// l[k-1] < L <= l[k] and ends at k iff l[k] < L <= l[k-1].  So:
//   count  -- per 1024-position tile, the number of cell starts at each level;
//   scan   -- per level, exclusive scan over tiles (one workgroup per level);
//   cells  -- every position ranks its start / end events among the level's starts
//             (ballot prefix sums) and writes cell start, end, level: cells come out in
//             level-major order, sorted by start within a level.  The same prefix sums
//             give, for every level P, the level-P cell containing the position (index
//             cell_base[P] + level-P starts at or before it - 1), hence each cell's and
//             each particle's parent;
//   link   -- every cell and particle fills its parent's child slot (octant = Lebesgue
//             digit, the reference's slot convention).
// No locks, no retries, and the same input always gives the same numbering.
// The bottom-up pass then runs level 21..0: mass, centre of mass (children in slot
// order, FP64, as calculateCentersOfMass, tree.cu:1416-1497), traceless quadrupole, the
// particles' bounding box and, for SPH, their largest smoothing length per cell.
#include "mhpc_internal.h"

#define TREE_BLOCK 256
#define TREE_ITEMS (MHPC_TREE_TILE / TREE_BLOCK)   // 4

__device__ __forceinline__ int lcp_levels(uint64_t a, uint64_t b) {
    const uint64_t x = a ^ b;
    if (x == 0) return MHPC_LEVELS;
    return (__clzll((long long)x) - 1) / 3;   // keys use bits 0..62
}

__global__ __launch_bounds__(TREE_BLOCK) void k_tree_count(const uint64_t *__restrict__ skey,
                                                           int64_t n, uint8_t *__restrict__ lcp,
                                                           uint32_t *__restrict__ tcnt,
                                                           int64_t ntiles, uint32_t *err) {
    __shared__ uint32_t cnt[MHPC_LEVELS + 1];
    if (threadIdx.x <= MHPC_LEVELS) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * MHPC_TREE_TILE;
#pragma unroll
    for (int s = 0; s < TREE_ITEMS; s++) {
        const int64_t k = base + s * TREE_BLOCK + threadIdx.x;
        if (k < n) {
            const uint64_t kk = skey[k];
            const int a = (k > 0) ? lcp_levels(skey[k - 1], kk) : 0;
            const int b = (k + 1 < n) ? lcp_levels(kk, skey[k + 1]) : 0;
            lcp[k] = (uint8_t)b;
            for (int L = a + 1; L <= b; L++) atomicAdd(&cnt[L], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x >= 1 && threadIdx.x <= MHPC_LEVELS)
        tcnt[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(TREE_BLOCK) void k_tree_scan(const uint32_t *__restrict__ tcnt,
                                                          uint32_t *__restrict__ toff,
                                                          int64_t ntiles,
                                                          int32_t *__restrict__ level_count) {
    __shared__ uint32_t lds4[4];
    const int L = blockIdx.x + 1;
    uint32_t tot = block_scan_row(tcnt + (int64_t)L * ntiles, toff + (int64_t)L * ntiles, ntiles, 0u,
                                  lds4);
    if (threadIdx.x == 0) level_count[L] = (int32_t)tot;
}

__global__ __launch_bounds__(TREE_BLOCK) void k_tree_cells(
    const uint8_t *__restrict__ lcp, int64_t n, const uint32_t *__restrict__ toff, int64_t ntiles,
    const int32_t *__restrict__ level_count, int32_t *__restrict__ cell_base_out,
    int32_t *__restrict__ cstart, int32_t *__restrict__ cend, int32_t *__restrict__ clevel,
    int32_t *__restrict__ cchild, int32_t *__restrict__ ppar, int32_t *__restrict__ cpar, int64_t cap_cells,
    uint32_t *err) {
    __shared__ int32_t cb[MHPC_LEVELS + 3];
    __shared__ uint32_t pre[TREE_ITEMS][4][MHPC_LEVELS + 1];
    const int tid = threadIdx.x, w = tid / 64, lane = wave_lane();
    if (tid == 0) {
        cb[0] = 0;
        cb[1] = 1;
        for (int L = 1; L <= MHPC_LEVELS; L++) cb[L + 1] = cb[L] + level_count[L];
        cb[MHPC_LEVELS + 2] = cb[MHPC_LEVELS + 1];
        if (blockIdx.x == 0) {
            for (int L = 0; L <= MHPC_LEVELS + 2; L++) cell_base_out[L] = cb[L];
            if ((int64_t)cb[MHPC_LEVELS + 1] > cap_cells) set_err(err, ERRF_CAP_NODES);
            // root: level 0, all particles
            cstart[0] = 0;
            cend[0] = (int32_t)(n - 1);
            clevel[0] = 0;
            for (int s = 0; s < 8; s++) cchild[s] = -1;
        }
    }
    const int64_t base = (int64_t)blockIdx.x * MHPC_TREE_TILE;
    int a[TREE_ITEMS], b[TREE_ITEMS];
    int lmax = 0;
#pragma unroll
    for (int s = 0; s < TREE_ITEMS; s++) {
        const int64_t k = base + s * TREE_BLOCK + tid;
        a[s] = (k < n && k > 0) ? (int)lcp[k - 1] : 0;
        b[s] = (k < n) ? (int)lcp[k] : 0;
        lmax = max(lmax, max(a[s], b[s]));
    }
    lmax = wave_max(lmax);
    // per (slot, wave, level) start counts
#pragma unroll
    for (int s = 0; s < TREE_ITEMS; s++) {
        for (int L = 1; L <= MHPC_LEVELS; L++) {
            const uint64_t m = (L <= lmax) ? wave_ballot(a[s] < L && L <= b[s]) : 0ull;
            if (lane == 0) pre[s][w][L] = (uint32_t)popc64(m);
        }
    }
    __syncthreads();
    if (tid >= 1 && tid <= MHPC_LEVELS) {
        uint32_t run = 0;
        for (int s = 0; s < TREE_ITEMS; s++)
            for (int ww = 0; ww < 4; ww++) {
                uint32_t v = pre[s][ww][tid];
                pre[s][ww][tid] = run;
                run += v;
            }
    }
    __syncthreads();
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int s = 0; s < TREE_ITEMS; s++) {
        const int64_t k = base + s * TREE_BLOCK + tid;
        // the particle's parent is the deepest cell containing it, of level P = max(a, b)
        const int P = max(a[s], b[s]);
        int64_t cont_prev = 0;   // the level-(L-1) cell containing k (level 0: the root)
        for (int L = 1; L <= lmax; L++) {
            const bool st = a[s] < L && L <= b[s];
            const bool en = b[s] < L && L <= a[s];
            const uint64_t m = wave_ballot(st);
            // level-L starts before k (global); with k's own start, the containing cell's rank
            const int64_t before =
                (int64_t)toff[(int64_t)L * ntiles + blockIdx.x] + pre[s][w][L] + popc64(m & lt);
            const int64_t cont = cb[L] + before + (st ? 1 : 0) - 1;   // valid for L <= P
            if (st) {
                const int64_t c = cb[L] + before;
                if (c < cap_cells) {
                    cstart[c] = (int32_t)k;
                    clevel[c] = L;
                    cpar[c] = (int32_t)cont_prev;
#pragma unroll
                    for (int q = 0; q < 8; q++) cchild[8 * c + q] = -1;
                }
            } else if (en) {
                const int64_t c = cb[L] + before - 1;
                if (c < cap_cells) cend[c] = (int32_t)k;
            }
            if (L == P && k < n) ppar[k] = (int32_t)(cont < cap_cells ? cont : cap_cells);
            cont_prev = cont;
        }
        if (P == 0 && k < n) ppar[k] = 0;
    }
}

// The octant of (x, y, z) at bisection MHPC_LEVELS + 1 of the root box: the keys' bisection
// (lower half = bit) carried one level past the key's 21 digits.
__device__ __forceinline__ int octant_past_key(double x, double y, double z, const double *__restrict__ box) {
    double mnx = box[0], mxx = box[1], mny = box[2], mxy = box[3], mnz = box[4], mxz = box[5];
    int son = 0;
    for (int level = 0; level <= MHPC_LEVELS; level++) {
        son = 0;
        double m = 0.5 * (mnx + mxx);
        if (x < m) { son += 1; mxx = m; } else { mnx = m; }
        m = 0.5 * (mny + mxy);
        if (y < m) { son += 2; mxy = m; } else { mny = m; }
        m = 0.5 * (mnz + mxz);
        if (z < m) { son += 4; mxz = m; } else { mnz = m; }
    }
    return son;
}

// Particles whose 21-digit keys are equal share a level-21 cell.  The reference's insertion
// keeps bisecting such a pair geometrically (leaves may sit one level below the keys, the
// depth limit is MAX_LEVEL + 1, tree.cu:949-963), so these particles take the child slot of
// their octant at the next bisection; two of them in the same slot would need a level-22 cell,
// which this tree does not build: that is the depth error (MHPC_E_DEPTH).
__global__ __launch_bounds__(TREE_BLOCK) void k_tree_link(
    const uint8_t *__restrict__ lcp, const uint64_t *__restrict__ slkey,
    const int32_t *__restrict__ perm, int64_t n, const int32_t *__restrict__ cell_base,
    const int32_t *__restrict__ cstart, const int32_t *__restrict__ clevel, const int32_t *__restrict__ ppar,
    const int32_t *__restrict__ cpar, int32_t *__restrict__ cchild, int32_t *__restrict__ plevel,
    int64_t cap_cells, const double4 *__restrict__ sxm, const double *__restrict__ box, uint32_t *err) {
    __shared__ int32_t cb[MHPC_LEVELS + 3];
    if (threadIdx.x < MHPC_LEVELS + 3) cb[threadIdx.x] = cell_base[threadIdx.x];
    __syncthreads();
    const int64_t ncells = cb[MHPC_LEVELS + 1];
    if (ncells > cap_cells) return;   // capacity error already flagged; cells are incomplete
    // particles, then cells 1 .. ncells-1 (grid-stride: the cell count is known here only)
    const int64_t items = n + ncells - 1;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < items;
         t += (int64_t)gridDim.x * blockDim.x) {
    if (t < n) {
        const int32_t k = (int32_t)t;
        const int a = (k > 0) ? (int)lcp[k - 1] : 0;
        const int b = (int)lcp[k];
        const int Lp = max(a, b) + 1;
        plevel[perm[k]] = Lp;
        const int32_t parent = ppar[k];
        if (parent >= cap_cells) continue;
        if (Lp > MHPC_LEVELS) {
            const double4 r = sxm[k];
            const int oct = octant_past_key(r.x, r.y, r.z, box);
            if (atomicCAS(&cchild[8 * (int64_t)parent + oct], -1, k) != -1) set_err(err, ERRF_DEPTH);
            continue;
        }
        const int digit = (int)((slkey[k] >> (3 * (MHPC_LEVELS - Lp))) & 7u);
        cchild[8 * (int64_t)parent + digit] = k;
    } else {
        const int64_t c = t - n + 1;   // cells 1 .. ncells-1
        const int L = clevel[c];
        const int32_t s = cstart[c];
        const int32_t parent = cpar[c];
        const int digit = (int)((slkey[s] >> (3 * (MHPC_LEVELS - L))) & 7u);
        cchild[8 * (int64_t)parent + digit] = (int32_t)(n + c);
    }
    }
}

void launch_tree(mhpc_ctx *c) {
    const int64_t n = c->n;
    c->nn_start_valid = false;
    if (n <= 0) return;
    const int64_t ntiles = ceil_div(n, MHPC_TREE_TILE);
    const uint64_t *slkey = (c->p.curve == MHPC_CURVE_HILBERT) ? c->slkey : c->skey;
    hipLaunchKernelGGL(k_tree_count, dim3((unsigned)ntiles), dim3(TREE_BLOCK), 0, c->stream, c->skey, n,
                       c->lcp, c->tcnt, ntiles, c->err);
    hipLaunchKernelGGL(k_tree_scan, dim3(MHPC_LEVELS), dim3(TREE_BLOCK), 0, c->stream, c->tcnt, c->toff,
                       ntiles, c->level_count);
    hipLaunchKernelGGL(k_tree_cells, dim3((unsigned)ntiles), dim3(TREE_BLOCK), 0, c->stream, c->lcp, n,
                       c->toff, ntiles, c->level_count, c->cell_base, c->cstart, c->cend, c->clevel,
                       c->cchild, c->ppar, c->cpar, c->cap_cells, c->err);
    const int64_t items = n + c->cap_cells;   // upper bound; the kernel loops over the actual count
    const int64_t link_blocks = std::min<int64_t>(ceil_div(items, TREE_BLOCK), 8192);
    hipLaunchKernelGGL(k_tree_link, dim3((unsigned)link_blocks), dim3(TREE_BLOCK), 0,
                       c->stream, c->lcp, slkey, c->perm, n, c->cell_base, c->cstart, c->clevel, c->ppar,
                       c->cpar, c->cchild, c->level, c->cap_cells, c->sxm, c->box, c->err);
}

// ------------------------------------------------------------------------------------
// bottom-up, one launch per level (cells of a level are contiguous): FP64 mass and
// centre of mass over the children in slot order (tree.cu:1416-1497), traceless
// quadrupole about the COM with the parallel-axis shift, the particles' AABB and (chmax
// set) their largest h.
// Each launch reads its level's range from cell_base on the device.
// Eight lanes per cell, lane s loading child s (all children's loads in flight at once);
// the children's terms go through LDS and the cell's first lane sums them in slot order, so
// every sum has the order and rounding of a sequential loop over the children.  An empty
// slot contributes exact zeros.
#define BU_BLOCK 256
// Levels L_hi down to L_lo; more than one level only with a single workgroup (the barrier
// between levels orders them).  With depth_out set, thread 0 also records the deepest level
// that has cells.
__global__ __launch_bounds__(BU_BLOCK) void k_bottom_up(
    int L_hi, int L_lo, int64_t n, const int32_t *__restrict__ cell_base, const int32_t *__restrict__ cchild,
    const double4 *__restrict__ sxm, double *__restrict__ cmass, double *__restrict__ cx,
    double *__restrict__ cy, double *__restrict__ cz, double *__restrict__ cquad,
    float4 *__restrict__ cquadf, double *__restrict__ cbmin, double *__restrict__ cbmax,
    int64_t cap_cells, int gravity, int quad, unsigned long long *depth_out, const double *__restrict__ sh,
    double *__restrict__ chmax) {
    // a cell's eight lanes are consecutive lanes of one wavefront, so they exchange their
    // terms through LDS with wavefront barriers only; lane s of the cell then sums one
    // quantity over the eight children in slot order (the sequential loop's rounding)
    __shared__ double term[11][BU_BLOCK];
    __shared__ double red[4][BU_BLOCK / 8];
    const int tid = threadIdx.x, sub = tid & 7, g = tid >> 3, c8 = tid & ~7;
    float *quadf = reinterpret_cast<float *>(cquadf);   // 8 floats per cell
    if (depth_out && blockIdx.x == 0 && tid == 0) {
        int d = 0;
        for (int l = 1; l <= MHPC_LEVELS; l++)
            if (cell_base[l + 1] > cell_base[l]) d = l;
        *depth_out = (unsigned long long)d;
    }
    for (int L = L_hi; L >= L_lo; L--) {
    const int64_t c0 = cell_base[L], c1a = cell_base[L + 1];
    const int64_t c1 = c1a < cap_cells ? c1a : cap_cells;
    const int64_t nthreads = (c1 - c0) * 8;
    for (int64_t b = (int64_t)blockIdx.x * BU_BLOCK; b < nthreads; b += (int64_t)gridDim.x * BU_BLOCK) {
        const int64_t t = b + tid;
        const bool valid = t < nthreads;   // the eight lanes of a cell agree
        const int64_t c = c0 + (t >> 3);
        const int32_t k = valid ? cchild[8 * c + sub] : -1;
        double px = 0.0, py = 0.0, pz = 0.0, pm = 0.0;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double hk = 0.0;   // the child's h (particle) or largest h (cell)
        if (k >= 0 && k < n) {
            const double4 r = sxm[k];
            px = r.x; py = r.y; pz = r.z; pm = r.w;
            lo[0] = hi[0] = px; lo[1] = hi[1] = py; lo[2] = hi[2] = pz;
            if (chmax) hk = sh[k];
        } else if (k >= n) {
            const int64_t cc = k - n;
            px = cx[cc]; py = cy[cc]; pz = cz[cc]; pm = cmass[cc];
            if (chmax) hk = chmax[cc];
            lo[0] = cbmin[cc]; lo[1] = cbmin[cap_cells + cc]; lo[2] = cbmin[2 * cap_cells + cc];
            hi[0] = cbmax[cc]; hi[1] = cbmax[cap_cells + cc]; hi[2] = cbmax[2 * cap_cells + cc];
            if (quad) {
#pragma unroll
                for (int e = 0; e < 6; e++) q[e] = cquad[e * cap_cells + cc];
            }
        }
        term[0][tid] = pm;
        term[1][tid] = pm * px;
        term[2][tid] = pm * py;
        term[3][tid] = pm * pz;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            term[4 + a][tid] = lo[a];
            term[7 + a][tid] = hi[a];
        }
        term[10][tid] = hk;
        wave_sync();
        if (valid) {
            if (sub < 4) {
                // lane 0: mass; lanes 1-3: mass-weighted positions
                double v = 0.0;
#pragma unroll
                for (int s = 0; s < 8; s++) v += term[sub][c8 + s];
                red[sub][g] = v;
            } else if (sub < 7) {
                const int a = sub - 4;
                double bmin = 1e300, bmax = -1e300;
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    bmin = fmin(bmin, term[4 + a][c8 + s]);
                    bmax = fmax(bmax, term[7 + a][c8 + s]);
                }
                cbmin[a * cap_cells + c] = bmin;
                cbmax[a * cap_cells + c] = bmax;
            } else if (chmax) {
                double hm = 0.0;
#pragma unroll
                for (int s = 0; s < 8; s++) hm = fmax(hm, term[10][c8 + s]);
                chmax[c] = hm;
            }
        }
        wave_sync();
        if (gravity) {
            if (valid && sub < 4) {
                const double m = red[0][g];
                double v = red[sub][g];
                if (sub == 0) {
                    cmass[c] = m;
                } else {
                    if (m > 0.0) v /= m;
                    (sub == 1 ? cx : sub == 2 ? cy : cz)[c] = v;
                    red[sub][g] = v;   // the centre of mass for the quadrupole terms
                }
            }
            wave_sync();
            if (quad) {
                double T[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                if (k != -1) {
                    const double dx = px - red[1][g], dy = py - red[2][g], dz = pz - red[3][g];
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    T[0] = q[0] + pm * (3.0 * dx * dx - d2);
                    T[1] = q[1] + pm * (3.0 * dx * dy);
                    T[2] = q[2] + pm * (3.0 * dx * dz);
                    T[3] = q[3] + pm * (3.0 * dy * dy - d2);
                    T[4] = q[4] + pm * (3.0 * dy * dz);
                    T[5] = q[5] + pm * (3.0 * dz * dz - d2);
                }
#pragma unroll
                for (int e = 0; e < 6; e++) term[e][tid] = T[e];
                wave_sync();
                if (valid) {
                    // lane e < 6: quadrupole component e; lanes 6, 7: the FP32 copy's padding
                    float f = 0.f;
                    if (sub < 6) {
                        double A = 0.0;
#pragma unroll
                        for (int s = 0; s < 8; s++) A += term[sub][c8 + s];
                        cquad[sub * cap_cells + c] = A;
                        f = (float)A;
                    }
                    quadf[8 * c + sub] = f;
                }
            }
        }
        wave_sync();   // the terms are read before the next iteration overwrites them
    }
    if (L_hi != L_lo) __syncthreads();
    }
}
void launch_bottom_up(mhpc_ctx *c, bool gravity, bool sph) {
    if (c->n <= 0) return;
    // a level holds fewer cells than particles; eight threads per cell, grid-stride beyond
    int grid = (int)ceil_div(c->n * 8, BU_BLOCK * 4);
    if (grid < 1) grid = 1;
    if (grid > 4096) grid = 4096;
    const int quad = (gravity && c->p.quadrupole) ? 1 : 0;
    auto run = [&](int hi, int lo, int g, unsigned long long *depth_out) {
        hipLaunchKernelGGL(k_bottom_up, dim3(g), dim3(BU_BLOCK), 0, c->stream, hi, lo, c->n, c->cell_base,
                           c->cchild, c->sxm, c->cmass, c->cx, c->cy, c->cz, c->cquad, c->cquadf, c->cbmin,
                           c->cbmax, c->cap_cells, gravity ? 1 : 0, quad, depth_out, sph ? c->sh : nullptr,
                           sph ? c->chmax : nullptr);
    };
    // levels below the last call's deepest (normally empty) in one single-workgroup sweep,
    // then one launch per level down to 3, then levels 2..0 (<= 73 cells) in one sweep
    int deep = c->com_depth_hint;
    if (deep < 3) deep = 3;
    if (deep > MHPC_LEVELS) deep = MHPC_LEVELS;
    if (deep < MHPC_LEVELS) run(MHPC_LEVELS, deep + 1, 1, nullptr);
    for (int L = deep; L >= 3; L--) {
        // level L holds at most min(8^L, n) cells: no more workgroups than that needs
        const int64_t cells = L < 7 ? std::min<int64_t>((int64_t)1 << (3 * L), c->n) : c->n;
        run(L, L, (int)std::min<int64_t>(grid, ceil_div(cells * 8, BU_BLOCK * 4)), nullptr);
    }
    run(2, 0, 1, c->counters + CNT_TREE_DEPTH);
}

// k_sph.hip -- fixed-radius neighbour search, plain h-iteration, density, EOS and
// internal forces, all on the key-sorted working copies.
//
// Neighbour search.  The reference walks the tree once per particle with per-thread
// stacks (SPH::Kernel::fixedRadiusNN, sph.cu:93-281).  Here one wavefront owns 64
// consecutive key-sorted particles: it walks the tree once with the group's bounding box
// grown by the group's largest h, collecting the particle ranges of small cells (each
// cell is a contiguous range of sorted positions), then streams the candidates through
// LDS in tiles of 64 with coalesced SoA loads; every lane tests every candidate of the
// tile against its own particle with the reference criterion d^2 < h_i^2 and
// d^2 < h_j^2 (FP64, exact), so the neighbour sets equal a brute-force search.
#include <algorithm>
#include <type_traits>

#include "mhpc_internal.h"

#define NN_WAVES 2
#define NN_CWIN 512   // group walk, list modes: the consumers' LDS window (in-window neighbours first)
#define NN_PB 4      // group walk, list modes: candidates read from LDS per batch
#define NN_STACK 512
#define NN_STACK_PHYS (NN_STACK + 192)
#define NN_RANGES 128
#define NN_LEAF 16           // per-particle walks: cells of at most this many particles are taken whole
#define NN_LEAF_GROUP 64     // group walks: the same for the 64-particle groups' boxes
static_assert(NN_LEAF_GROUP < NN_CWIN, "a clipped range crosses at most one end of the window");
#define NN_OUTLIER 4.0       // group-mode outlier: h above this factor x the group's smallest
#define NN_OUTLIER_MAX 32    // at most this many outliers per group are split off
#define NN_SPARSE 64         // re-lists: groups with at most this many flagged particles go per particle
#define NN_DRANGES 256       // direct form: candidate ranges per wavefront
#define NN_DSPAN 5           // direct form: start cells per axis - 1 (<= 216 cells)
#define NN_DIRECT_MAX 32768  // direct form: candidates per wavefront
#define NN_DU 8              // direct form: candidates per lane and round (loads in flight)

// per-wave LDS of the walker; TCAP = 128 gives the filtered walk room to compact the
// candidates of two 64-particle loads before a full tile of 64 is handed on
template <int TCAP, int DR>
struct NNWaveLdsT {
    int stack[NN_STACK_PHYS];
    int rstart[NN_RANGES];
    int rcnt[NN_RANGES];
    int rpre[64];
    double tx[TCAP], ty[TCAP], tz[TCAP], th2[TCAP];
    int tj[TCAP];
    int drs[DR], drp[DR];   // direct form: candidate ranges (first particle, exclusive prefix)
};
using NNWaveLds = NNWaveLdsT<64, NN_DRANGES>;
using NNWaveLdsF = NNWaveLdsT<128, 1>;

// candidate filter that keeps everything (per-particle walks test each candidate once)
struct KeepAll {
    __device__ __forceinline__ bool operator()(double, double, double, double) const { return true; }
};

struct NNArgs {
    int64_t n;
    const int32_t *perm;
    const double4 *sxm;
    const double *h;          // h used for the search (h_cur in the h-iteration)
    const int32_t *cstart, *cend, *cchild;
    const double *cbmin, *cbmax;
    int64_t cap_cells;
    // start table (levels 1..kmax) and the root box the keys were computed in
    const int32_t *stab;
    const int2 *srng;         // {first, count} of each table entry's node
    int kmax;
    const double *box;
    // list mode: entry s of particle k is nbr4[(s / 4) * stride + k].{x,y,z,w}[s % 4]
    int4 *nbr4;
    int64_t stride;
    int max_nb;
    int32_t *snoi;
    int32_t *noi_out;         // caller order
    double *sml_out;          // caller order (written when non-null)
    const int32_t *siters;
    int32_t *iters_out;       // caller order (written when non-null)
    // h-iteration mode
    double *h_next;
    uint8_t *done;
    int32_t *iters;
    const double *table;
    int tmax;
    int pass, max_counts, band_lo, band_hi;
    double sml_max;           // upper bound of h (0: none)
    // compacted lists of the particles still iterating: pass p reads list_in / count_in
    // (per-particle mode) and appends the particles that stay unconverged to list_out
    const int32_t *list_in;
    const uint32_t *count_in;
    int32_t *list_out;
    uint32_t *count_out;
    uint32_t *cnt_slot;       // split counting: neighbour count of list slot t
    const int32_t *slot_list; // split counting: the list slots it takes
    const uint32_t *slot_count;
    const uint8_t *only;      // list mode: when set, only particles with only[k] != 0
    uint8_t *dirty;           // dirty marking: particles whose list must be rebuilt
    double *h0_slot;          // pass 0 with lists: the pass-0 h of each unconverged particle
                              // (slot of list_out); read by the dirty marking
    int32_t *out_list;        // group mode: outlier particles deferred to the per-particle mode
    uint32_t *out_count;
    uint32_t *err;
    unsigned long long *counters;
    const double *chmax;      // per cell: an upper bound of its particles' h
    // h-iteration passes: the bounds the new h of the particles still iterating raise
    double *chmax_raise;
    const int32_t *ppar, *cpar;
    // tail passes: set when this pass changed the h of a particle that keeps iterating; a pass
    // whose previous pass changed none (prev_changed set and zero) repeats that pass's counts
    uint32_t *changed;
    const uint32_t *prev_changed;
};

// A tail pass after a pass that changed no h sees the same h everywhere as that pass (the
// particles that finished kept theirs), so every count it would take equals the count the
// previous pass stored in snoi: the counting kernels skip it and k_h_finish reuses snoi.
__device__ __forceinline__ bool h_pass_frozen(const NNArgs &A) {
    return A.prev_changed && *A.prev_changed == 0u;
}

// The positions a search is for -- one particle's, or the box of a group's (active) particles --
// and the largest h among them (the search radius).
struct NNCore {
    double x0, x1, y0, y1, z0, z1;
    double h;
};

// the squared distance of every neighbour test (and of the bounds that prune them):
// fma(dx, dx, fma(dy, dy, dz^2)), the oracle's form (two fused multiply-adds per test)
__device__ __forceinline__ double nn_d2(double dx, double dy, double dz) { return fma(dx, dx, fma(dy, dy, dz * dz)); }
__device__ __forceinline__ NNCore nn_point(double x, double y, double z, double h) {
    return NNCore{x, x, y, y, z, z, h};
}

// Can cell c hold a neighbour j of some particle i of the core?  A pair needs d < h_i^2 and
// d < h_j^2, and the squared distance from the core to the cell's particle box, taken per axis
// as max(0, lo_cell - hi_core, lo_core - hi_cell) and combined in the pair test's form
// (nn_d2), is at most every such pair's d (each rounding is monotone in exact values that
// grow with the per-axis distances), while h_i <= q.h and h_j <= chmax: a
// cell that fails d2 < min(q.h, chmax)^2 holds no neighbour of the core and its subtree is
// skipped.
__device__ __forceinline__ bool nn_cell_reaches(const NNArgs &A, const NNCore &q, int64_t c) {
    const int64_t cc = A.cap_cells;
    const double ex = fmax(fmax(A.cbmin[c] - q.x1, q.x0 - A.cbmax[c]), 0.0);
    const double ey = fmax(fmax(A.cbmin[cc + c] - q.y1, q.y0 - A.cbmax[cc + c]), 0.0);
    const double ez = fmax(fmax(A.cbmin[2 * cc + c] - q.z1, q.z0 - A.cbmax[2 * cc + c]), 0.0);
    const double hm = fmin(A.chmax[c], q.h);
    return nn_d2(ex, ey, ez) < hm * hm;
}

// Raise the cells' h bounds (chmax) to a particle's new h: every ancestor cell whose bound is
// smaller (an atomic maximum on the bits of the positive double, which orders like the value),
// stopping at the first ancestor that already bounds it.  Bounds never shrink within an RHS, so
// they stay upper bounds of the h every later search uses.
__device__ __forceinline__ void hmax_raise(double *chmax, const int32_t *__restrict__ ppar,
                                           const int32_t *__restrict__ cpar, int64_t cap_cells, int64_t k,
                                           double h) {
    const unsigned long long hb = (unsigned long long)__double_as_longlong(h);
    int64_t c = ppar[k];
    while (c >= 0 && c < cap_cells) {
        const unsigned long long old = atomicMax(reinterpret_cast<unsigned long long *>(chmax + c), hb);
        if (old >= hb || c == 0) break;
        c = cpar[c];
    }
}

// Level-K index along one axis of the cell containing v: the keys' bisection of [mn, mx]
// (lower half = key bit 1), counted upwards, so it is monotone in v.
__device__ __forceinline__ int nn_axis_index(double v, double mn, double mx, int K) {
    int idx = 0;
    for (int l = 0; l < K; l++) {
        const double m = 0.5 * (mn + mx);
        if (v < m) {
            mx = m;
            idx <<= 1;
        } else {
            mn = m;
            idx = (idx << 1) | 1;
        }
    }
    return idx;
}

#define NN_START_ALL (-1)     // walk start: table cells covering the box
#define NN_START_PART0 (-2)   // NN_START_PART0 - o: table cells of index = o (mod 8)

// The start level of a walk for the box [lo, hi] (see nn_start_push).
struct NNStart {
    int K, s;                          // start level K = kmax - s (< 1: whole tree)
    int ilx, ihx, ily, ihy, ilz, ihz;  // the box corners' cell indices at level kmax
};

template <int SPAN = 3>
__device__ __forceinline__ NNStart nn_start_level(const NNArgs &A, double lox, double hix, double loy,
                                                  double hiy, double loz, double hiz) {
    NNStart S = {};
    const int K0 = A.stab ? A.kmax : 0;
    S.s = K0;
    if (K0 > 0) {
        const double b0 = A.box[0], b1 = A.box[1], b2 = A.box[2], b3 = A.box[3], b4 = A.box[4], b5 = A.box[5];
        S.ilx = nn_axis_index(lox, b0, b1, K0);
        S.ihx = nn_axis_index(hix, b0, b1, K0);
        S.ily = nn_axis_index(loy, b2, b3, K0);
        S.ihy = nn_axis_index(hiy, b2, b3, K0);
        S.ilz = nn_axis_index(loz, b4, b5, K0);
        S.ihz = nn_axis_index(hiz, b4, b5, K0);
        int s = 0;
        while (s < K0 && ((S.ihx >> s) - (S.ilx >> s) > SPAN || (S.ihy >> s) - (S.ily >> s) > SPAN ||
                          (S.ihz >> s) - (S.ilz >> s) > SPAN))
            s++;
        S.s = s;
    }
    S.K = K0 - S.s;
    return S;
}

// Consumer window of a list walk: the particles [w0, w1) of the density and force kernels'
// LDS window that holds the walking group (w1 - w0 >= leaf; w0 == w1: none).  Their
// candidates are streamed first as one range, so every list holds its in-window neighbours
// first, and the tree walk skips them: a node or range inside the window is dropped (false)
// and a range of at most `leaf` (< NN_CWIN) particles is clipped to its part outside (it
// crosses at most one end of the window).  Larger nodes that cross an end are opened as usual.
__device__ __forceinline__ bool nn_win_clip(int w0, int w1, int &rs, int &rc, int leaf) {
    if (w0 == w1) return true;
    const int64_t e = (int64_t)rs + rc;
    if (rs >= w0 && e <= w1) return false;
    if (rc <= leaf) {
        if (rs < w0 && e > w0) rc = w0 - rs;
        else if (rs < w1 && e > w1) { rc = (int)(e - w1); rs = w1; }
    }
    return true;
}

// Start a walk for the box [lo, hi]: the deepest table level K <= kmax at which the box
// spans at most 4 cells per axis, and the table's node for every such cell (a cell of level
// K, a particle alone in it, or none).  Every particle inside the box lies in one of these
// cells (the keys use the same bisection), and the nodes' subtrees are disjoint, so the walk
// sees the same candidates as one from the root, in another order.  A node of at most
// `leaf` particles goes straight to the range list (its particles are candidates, as for a
// leaf the walk would take), any other onto the stack.  With nparts = 8, only cells of
// index = part (mod 8).  With a consumer window [w0, w1), its ranges come first -- the walking
// group's own particles [g0, g0 + 64) before the rest of the window -- and the nodes are
// clipped against it.  Returns the stack count; nr receives the range count.
template <typename LdsT>
__device__ __forceinline__ int nn_start_push(LdsT &L, const NNArgs &A, double lox, double hix, double loy,
                                             double hiy, double loz, double hiz, const NNCore &q, int part,
                                             int nparts, int w0, int w1, int g0, int &nr, int leaf) {
    const int lane = wave_lane();
    int node = -1;
    int2 rng = make_int2(0, 0x7fffffff);
    const NNStart S = nn_start_level(A, lox, hix, loy, hiy, loz, hiz);
    const int K = S.K, s = S.s;
    const int ilx = S.ilx, ihx = S.ihx, ily = S.ily, ihy = S.ihy, ilz = S.ilz, ihz = S.ihz;
    if (K < 1) {
        // box too large for the table (or no table): the whole tree
        if (lane == 0) node = nparts == 1 ? (int)A.n : A.cchild[part];
    } else {
        const int bx = ilx >> s, by = ily >> s, bz = ilz >> s;
        const int nx = (ihx >> s) - bx + 1, ny = (ihy >> s) - by + 1, nz = (ihz >> s) - bz + 1;
        if (lane < nx * ny * nz && lane % nparts == part) {
            const int ix = bx + lane % nx, iy = by + (lane / nx) % ny, iz = bz + lane / (nx * ny);
            uint64_t P = 0;
            for (int l = K - 1; l >= 0; l--) {
                const uint64_t d = (uint64_t)(((ix >> l) & 1) ^ 1) | ((uint64_t)(((iy >> l) & 1) ^ 1) << 1) |
                                   ((uint64_t)(((iz >> l) & 1) ^ 1) << 2);
                P = (P << 3) | d;
            }
            node = A.stab[nn_tab_off(K) + (int64_t)P];
            if (node >= 0) rng = A.srng[nn_tab_off(K) + (int64_t)P];
        }
    }
    if (node >= A.n && !nn_cell_reaches(A, q, node - A.n)) node = -1;
    if (node >= 0 && !nn_win_clip(w0, w1, rng.x, rng.y, leaf)) node = -1;
    int r0 = 0;   // the window's ranges: own group, then the window before and after it
    if (w1 > w0) {
        const int g1 = g0 + 64 < w1 ? g0 + 64 : w1;
        if (lane == 0) {
            L.rstart[0] = g0;
            L.rcnt[0] = g1 - g0;
            int r = 1;
            if (g0 > w0) {
                L.rstart[r] = w0;
                L.rcnt[r] = g0 - w0;
                r++;
            }
            if (w1 > g1) {
                L.rstart[r] = g1;
                L.rcnt[r] = w1 - g1;
            }
        }
        r0 = 1 + (g0 > w0 ? 1 : 0) + (w1 > g1 ? 1 : 0);
    }
    const uint64_t lt = lanemask_lt();
    const bool asRange = node >= 0 && rng.y <= leaf;
    const bool asNode = node >= 0 && !asRange;
    const uint64_t mR = wave_ballot(asRange), mN = wave_ballot(asNode);
    if (asRange) {
        const int pos = r0 + popc64(mR & lt);
        L.rstart[pos] = rng.x;
        L.rcnt[pos] = rng.y;
    }
    if (asNode) L.stack[popc64(mN & lt)] = node;
    nr = r0 + popc64(mR);
    return popc64(mN);
}

// Direct form of a per-particle search for the box [lo, hi]: the start cells of the box (up
// to NN_DSPAN + 1 per axis, at the deepest table level where they cover it) are taken whole as
// particle ranges -- no tree walk -- and every particle of them is handed to
// visit(valid, j, {x,y,z,m}_j, h_j), NN_DU candidates per lane and round with all their loads
// issued first; candidates come in ascending order of (round, slot, lane), and visit is called
// by all lanes together (it may use ballots).  With nparts > 1 only the start cells of index =
// part (mod nparts).  Returns false, before any visit, when the box is too large for the
// table, has more than NN_DRANGES non-empty cells or more than NN_DIRECT_MAX candidates
// (counted over the whole box, so every part of a split search decides alike).
template <typename LdsT, typename VisitFn>
__device__ __forceinline__ bool nn_direct(LdsT &L, const NNArgs &A, double lox, double hix, double loy,
                                          double hiy, double loz, double hiz, const NNCore &q, int part,
                                          int nparts, VisitFn &&visit) {
    const int lane = wave_lane();
    const uint64_t lt = lanemask_lt();
    const NNStart S = nn_start_level<NN_DSPAN>(A, lox, hix, loy, hiy, loz, hiz);
    if (S.K < 1) return false;
    const int K = S.K, s = S.s;
    const int bx = S.ilx >> s, by = S.ily >> s, bz = S.ilz >> s;
    const int nx = (S.ihx >> s) - bx + 1, ny = (S.ihy >> s) - by + 1, nz = (S.ihz >> s) - bz + 1;
    const int ncell = nx * ny * nz;
    int nr = 0, total = 0, nr_all = 0, total_all = 0;
    for (int c0 = 0; c0 < ncell; c0 += 64) {
        const int idx = c0 + lane;
        int first = 0, num = 0;
        if (idx < ncell) {
            const int ix = bx + idx % nx, iy = by + (idx / nx) % ny, iz = bz + idx / (nx * ny);
            uint64_t P = 0;
            for (int l = K - 1; l >= 0; l--) {
                const uint64_t d = (uint64_t)(((ix >> l) & 1) ^ 1) | ((uint64_t)(((iy >> l) & 1) ^ 1) << 1) |
                                   ((uint64_t)(((iz >> l) & 1) ^ 1) << 2);
                P = (P << 3) | d;
            }
            const int64_t e = nn_tab_off(K) + (int64_t)P;
            const int node = A.stab[e];
            if (node >= 0 && (node < A.n || nn_cell_reaches(A, q, node - A.n))) {
                const int2 r = A.srng[e];
                first = r.x;
                num = r.y;
            }
        }
        // the decision to decline is taken on the whole box, so all parts agree
        nr_all += popc64(wave_ballot(num > 0));
        total_all += wave_sum(num);
        if (nr_all > NN_DRANGES || total_all > NN_DIRECT_MAX) return false;
        if (idx % nparts != part) num = 0;
        const uint64_t m = wave_ballot(num > 0);
        const int incl = wave_incl_scan(num);
        const int tot = __shfl(incl, 63);
        if (num > 0) {
            const int pos = nr + popc64(m & lt);
            L.drs[pos] = first;
            L.drp[pos] = total + incl - num;
        }
        nr += popc64(m);
        total += tot;
    }
    wave_sync();
    for (int base = 0; base < total; base += 64 * NN_DU) {
        int jj[NN_DU];
#pragma unroll
        for (int u = 0; u < NN_DU; u++) {
            const int c = base + 64 * u + lane;
            jj[u] = -1;
            if (c < total) {
                int lo = 0, hi_ = nr - 1;
                while (lo < hi_) {
                    const int mid = (lo + hi_ + 1) >> 1;
                    if (L.drp[mid] <= c) lo = mid; else hi_ = mid - 1;
                }
                jj[u] = L.drs[lo] + (c - L.drp[lo]);
            }
        }
        double4 pj[NN_DU];
        double hj[NN_DU];
#pragma unroll
        for (int u = 0; u < NN_DU; u++) {
            pj[u] = make_double4(0.0, 0.0, 0.0, 0.0);
            hj[u] = 0.0;
            if (jj[u] >= 0) {
                pj[u] = A.sxm[jj[u]];
                hj[u] = A.h[jj[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < NN_DU; u++) visit(jj[u] >= 0, jj[u], pj[u], hj[u]);
    }
    wave_sync();
    return true;
}

// Walk the (sub)tree of node `start` (the root is node n) for the axis-aligned box [lo, hi]
// (the core q grown by the search radius) and hand the candidate particles to
// `tile` in LDS tiles of up to 64 (L.tx/ty/tz/th2/tj, count m).  Cells that cannot hold a
// neighbour of the core (nn_cell_reaches) are skipped with their subtrees.  Leaves of at most
// NN_LEAF particles (NN_LEAF_GROUP with FILTER: the group walks) and single particles are
// taken as ranges of sorted positions; larger cells are opened.  With FILTER, only candidates
// with keep(x, y, z, h^2) true reach the tiles (compacted in walk order, so the tile sequence
// is the unfiltered one minus the dropped candidates).  Returns false on stack overflow.
// Wave-uniform control flow.
template <bool FILTER, typename LdsT, typename KeepFn, typename TileFn>
__device__ __forceinline__ bool nn_walk_box(LdsT &L, const NNArgs &A, double lox, double hix, double loy,
                                            double hiy, double loz, double hiz, const NNCore &q, int start,
                                            int w0, int w1, int g0, KeepFn &&keep, TileFn &&tile) {
    const int lane = wave_lane();
    const int64_t n = A.n;
    // a group's box spans a few 64-particle cells per axis: taken whole, their candidates come
    // in the order of the start cells, which is also the order the density and force kernels
    // gather best (same lists as sets, fewer distinct lines per gather step)
    constexpr int leaf = FILTER ? NN_LEAF_GROUP : NN_LEAF;
    int top = 1, nr = 0;
    if (start >= 0) {
        if (lane == 0) L.stack[0] = start;
    } else {
        top = nn_start_push(L, A, lox, hix, loy, hiy, loz, hiz, q, start == NN_START_ALL ? 0 : NN_START_PART0 - start,
                            start == NN_START_ALL ? 1 : 8, w0, w1, g0, nr, leaf);
    }
    wave_sync();
    const uint64_t lt = lanemask_lt();
    int fill = 0;   // FILTER: kept candidates waiting in L.t*[0, fill)
    for (;;) {
        // walk until 64 ranges are pending or the stack is empty
        while (top > 0 && nr < 64) {
            int kpop = (NN_STACK - top) / 7;
            kpop = kpop < 1 ? 1 : kpop;
            kpop = kpop > 64 ? 64 : kpop;
            kpop = kpop > top ? top : kpop;
            const int node = lane < kpop ? L.stack[top - kpop + lane] : -1;
            top -= kpop;
            wave_sync();
            bool take = false, open = false;
            int rs = 0, rc = 0;
            int64_t c = 0;
            if (node >= 0 && node < n) {
                const double4 r = A.sxm[node];
                const double px = r.x, py = r.y, pz = r.z;
                take = px >= lox && px <= hix && py >= loy && py <= hiy && pz >= loz && pz <= hiz;
                rs = node;
                rc = 1;
                take = take && nn_win_clip(w0, w1, rs, rc, leaf);
            } else if (node >= n) {
                c = node - n;
                const int64_t cc = A.cap_cells;
                const bool overlap = A.cbmin[c] <= hix && A.cbmax[c] >= lox &&
                                     A.cbmin[cc + c] <= hiy && A.cbmax[cc + c] >= loy &&
                                     A.cbmin[2 * cc + c] <= hiz && A.cbmax[2 * cc + c] >= loz &&
                                     nn_cell_reaches(A, q, c);
                if (overlap) {
                    rs = A.cstart[c];
                    rc = A.cend[c] - rs + 1;
                    if (nn_win_clip(w0, w1, rs, rc, leaf)) {
                        if (rc <= leaf) take = true; else open = true;
                    }
                }
            }
            const uint64_t mT = wave_ballot(take);
            if (take) {
                const int pos = nr + popc64(mT & lt);
                L.rstart[pos] = rs;
                L.rcnt[pos] = rc;
            }
            nr += popc64(mT);
            int ch[8];
            int nc = 0;
            if (open) {
                const int4 *cp = reinterpret_cast<const int4 *>(A.cchild + 8 * c);
                const int4 a = cp[0], b = cp[1];
                ch[0] = a.x; ch[1] = a.y; ch[2] = a.z; ch[3] = a.w;
                ch[4] = b.x; ch[5] = b.y; ch[6] = b.z; ch[7] = b.w;
#pragma unroll
                for (int s = 0; s < 8; s++) nc += ch[s] != -1;
            }
            const int incl = wave_incl_scan(nc);
            const int tot = __shfl(incl, 63);
            if (top + tot > NN_STACK_PHYS) return false;
            if (open) {
                int pos = top + incl - nc;
#pragma unroll
                for (int s = 7; s >= 0; s--)
                    if (ch[s] != -1) L.stack[pos++] = ch[s];
            }
            top += tot;
            wave_sync();
        }
        if (nr == 0 && top == 0) {
            if (FILTER && fill > 0) tile(fill);
            return true;
        }
        // stream the candidates of up to 64 ranges through LDS
        const int np = nr < 64 ? nr : 64;
        const int v = lane < np ? L.rcnt[lane] : 0;
        const int incl = wave_incl_scan(v);
        const int total = __shfl(incl, 63);
        const int pre = incl - v;
        L.rpre[lane] = pre;
        wave_sync();
        for (int base = 0; base < total; base += 64) {
            const int t = base + lane;
            // the ranges holding this chunk's first and last candidates bound every lane's
            // search (prefixes are non-decreasing): usually one or two ranges apart
            const int rf = popc64(wave_ballot(lane < np && pre <= base)) - 1;
            const int rl = popc64(wave_ballot(lane < np && pre <= base + 63)) - 1;
            if constexpr (!FILTER) {
                if (t < total) {
                    int rlo = rf, rhi = rl;
                    while (rlo < rhi) {
                        const int mid = (rlo + rhi + 1) >> 1;
                        if (L.rpre[mid] <= t) rlo = mid; else rhi = mid - 1;
                    }
                    const int j = L.rstart[rlo] + (t - L.rpre[rlo]);
                    const double hj = A.h[j];
                    const double4 r = A.sxm[j];
                    L.tx[lane] = r.x;
                    L.ty[lane] = r.y;
                    L.tz[lane] = r.z;
                    L.th2[lane] = hj * hj;
                    L.tj[lane] = j;
                }
                wave_sync();
                tile((total - base) < 64 ? (total - base) : 64);
                wave_sync();
            } else {
                bool kp = false;
                double rx = 0.0, ry = 0.0, rz = 0.0, h2 = 0.0;
                int j = 0;
                if (t < total) {
                    int rlo = rf, rhi = rl;
                    while (rlo < rhi) {
                        const int mid = (rlo + rhi + 1) >> 1;
                        if (L.rpre[mid] <= t) rlo = mid; else rhi = mid - 1;
                    }
                    j = L.rstart[rlo] + (t - L.rpre[rlo]);
                    const double hj = A.h[j];
                    const double4 r = A.sxm[j];
                    rx = r.x;
                    ry = r.y;
                    rz = r.z;
                    h2 = hj * hj;
                    kp = keep(rx, ry, rz, h2);
                }
                const uint64_t mk = wave_ballot(kp);
                if (kp) {
                    const int pos = fill + popc64(mk & lt);   // < fill + 64 <= 127
                    L.tx[pos] = rx;
                    L.ty[pos] = ry;
                    L.tz[pos] = rz;
                    L.th2[pos] = h2;
                    L.tj[pos] = j;
                }
                fill += popc64(mk);
                wave_sync();
                if (fill >= 64) {
                    tile(64);
                    wave_sync();
                    // move the overflow [64, fill) to the front (disjoint ranges)
                    const int rest = fill - 64;
                    if (lane < rest) {
                        const double a0 = L.tx[64 + lane], a1 = L.ty[64 + lane], a2 = L.tz[64 + lane];
                        const double a3 = L.th2[64 + lane];
                        const int a4 = L.tj[64 + lane];
                        L.tx[lane] = a0;
                        L.ty[lane] = a1;
                        L.tz[lane] = a2;
                        L.th2[lane] = a3;
                        L.tj[lane] = a4;
                    }
                    fill = rest;
                    wave_sync();
                }
            }
        }
        const int rest = nr - np;
        if (lane < rest) {
            const int a = L.rstart[64 + lane], b = L.rcnt[64 + lane];
            L.rstart[lane] = a;
            L.rcnt[lane] = b;
        }
        nr = rest;
        wave_sync();
    }
}

// the plain h-iteration's per-particle update after a count N (sph.cu:909-935 in the
// Jacobi form): finish when N is in the band (or at the last pass) and below max_nb,
// else scale h by 1/2 (1 + (n_target / N)^(1/3)) from the table.  Returns fin.
__device__ __forceinline__ bool h_update(const NNArgs &A, int N, double hi, double &hn) {
    const bool inband = N >= A.band_lo && N <= A.band_hi;
    const bool last = A.pass == A.max_counts - 1;
    bool fin = false;
    hn = hi;
    if ((inband || last) && N < A.max_nb) {
        fin = true;
    } else {
        int Nc = N < 1 ? 1 : N;
        Nc = Nc > A.tmax ? A.tmax : Nc;
        hn = hi * A.table[Nc];
        if (A.sml_max > 0.0 && hn > A.sml_max) hn = A.sml_max;
        if (last) fin = true;
    }
    if (fin && (N >= A.max_nb || N == 0)) set_err(A.err, ERRF_CAP_NEIGHBOURS);
    return fin;
}

// Group mode: one wavefront owns 64 consecutive sorted particles and walks once with the
// bounding box of its active particles grown by their largest h.
// MODE 0: build lists (only the particles flagged in `only` when it is set);
// MODE 1: one Jacobi pass of the h-iteration (count only; with list_out set, the
// unconverged particles are appended to it for the next pass);
// MODE 2: MODE 1 that also writes every particle's list with the h of this pass
// (speculative: kept for the particles whose list cannot change afterwards)
template <int MODE>
__global__ __launch_bounds__(64 * NN_WAVES) void k_nn_walk(NNArgs A) {
    __shared__ NNWaveLdsF lds_all[NN_WAVES];
    const int w = threadIdx.x / 64, lane = wave_lane();
    NNWaveLdsF &L = lds_all[w];
    const int64_t group = xcd_block(blockIdx.x, gridDim.x) * NN_WAVES + w;
    const int64_t k = group * 64 + lane;
    const int64_t n = A.n;
    if (group * 64 >= n) return;
    const bool valid = k < n;
    bool active = valid;
    if (MODE >= 1 && valid) active = A.done[k] == 0;
    if (MODE == 0 && valid && A.only) active = A.only[k] != 0;
    uint64_t mActive = wave_ballot(active);
    if (mActive == 0) {
        if (MODE >= 1 && valid) A.h_next[k] = A.h[k];
        return;
    }
    if (MODE == 0 && A.only && A.out_list && popc64(mActive) <= NN_SPARSE) {
        // a re-list of a few flagged particles: 64 lanes testing candidates for one or two of
        // them is wasted, so they go to the per-particle mode (lanes = candidates)
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(A.out_count, (uint32_t)popc64(mActive));
        b = __shfl(b, 0);
        if (active) A.out_list[b + popc64(mActive & lanemask_lt())] = (int32_t)k;
        return;
    }
    const double4 own = valid ? A.sxm[k] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double xi = own.x, yi = own.y, zi = own.z;
    const double hi = valid ? A.h[k] : 0.0;
    const double hi2 = hi * hi;
    // outliers: a few active particles with h above NN_OUTLIER x the group's smallest would
    // grow the search box of all 64; they go to the per-particle mode instead
    bool outlier = false;
    if (A.out_list) {
        const double hmin = wave_min(active ? hi : 1e300);
        const bool big = active && hi > NN_OUTLIER * hmin;
        const uint64_t mO = wave_ballot(big);
        if (mO != 0 && popc64(mO) <= NN_OUTLIER_MAX) {
            outlier = big;
            uint32_t b = 0;
            if (lane == 0) b = atomicAdd(A.out_count, (uint32_t)popc64(mO));
            b = __shfl(b, 0);
            if (outlier) A.out_list[b + popc64(mO & lanemask_lt())] = (int32_t)k;
            active = active && !outlier;
            mActive = wave_ballot(active);
        }
    }
    // search box: bounding box of the active lanes grown by their largest h (plus a
    // relative margin far above FP64 rounding, so pruning never drops a true pair)
    const double hmax = wave_max(active ? hi : 0.0);
    const double grow = hmax * (1.0 + 1e-9);
    const double bx0 = wave_min(active ? xi : 1e300), bx1 = wave_max(active ? xi : -1e300);
    const double by0 = wave_min(active ? yi : 1e300), by1 = wave_max(active ? yi : -1e300);
    const double bz0 = wave_min(active ? zi : 1e300), bz1 = wave_max(active ? zi : -1e300);
    const double lox = bx0 - grow, hix = bx1 + grow;
    const double loy = by0 - grow, hiy = by1 + grow;
    const double loz = bz0 - grow, hiz = bz1 + grow;
    // candidate filter: squared distance from the box of the active positions, below both
    // hmax^2 and h_j^2.  Per axis fl(b0 - x) <= fl(x_i - x) for every x_i >= b0 (rounding is
    // monotone), and nn_d2 is monotone in each |axis distance|, so this distance never exceeds
    // the pair test's d for any active i, and every
    // pair with d < h_i^2 <= hmax^2 and d < h_j^2 passes: the tiles lose only non-pairs.
    const double hm2 = hmax * hmax;
    auto keep = [&](double x, double y, double z, double h2) {
        const double ex = fmax(fmax(bx0 - x, x - bx1), 0.0);
        const double ey = fmax(fmax(by0 - y, y - by1), 0.0);
        const double ez = fmax(fmax(bz0 - z, z - bz1), 0.0);
        const double d = nn_d2(ex, ey, ez);
        return d < hm2 && d < h2;
    };

    int count = 0;
    // the last three hits, oldest first (a shift register): on every fourth hit they and the
    // new entry go out as one int4; the stale values left after a store are shifted out by
    // the next three hits before they are stored again
    int p0 = 0, p1 = 0, p2 = 0;
    int4 *wp = A.nbr4 + k;        // where the next four go
    // the count & 3 pending entries, in order, for the final partial int4
    auto tail4 = [&]() {
        const int r = count & 3;
        return r == 1 ? make_int4(p2, 0, 0, 0) : r == 2 ? make_int4(p1, p2, 0, 0) : make_int4(p0, p1, p2, 0);
    };
    // list modes: the group's consumer window (NN_CWIN-aligned) is streamed first
    int w0 = 0, w1 = 0;
    if (MODE != 1) {
        w0 = (int)(group * 64 / NN_CWIN * NN_CWIN);
        w1 = (int)(w0 + NN_CWIN < n ? w0 + NN_CWIN : n);
        if (w1 - w0 < NN_LEAF_GROUP) w0 = w1 = 0;
    }
    // with the window on, the group's own particles are the first candidates streamed, so all of
    // them that pass the filter are in the first tile: only its pair tests must skip j == k
    bool first_tile = true;
    const NNCore core = {bx0, bx1, by0, by1, bz0, bz1, hmax};
    const bool ok = nn_walk_box<true>(L, A, lox, hix, loy, hiy, loz, hiz, core, NN_START_ALL, w0, w1,
                                      (int)(group * 64), keep, [&](int m) {
        if (!active) return;
        if (MODE == 1) {
            for (int e = 0; e < m; e++) {
                const double dx = xi - L.tx[e];
                const double dy = yi - L.ty[e];
                const double dz = zi - L.tz[e];
                const double d = nn_d2(dx, dy, dz);
                count += (d < hi2 && d < L.th2[e] && L.tj[e] != (int)k) ? 1 : 0;
            }
            return;
        }
        // list modes: candidates come as LDS broadcasts, NN_PB per batch with all their reads
        // issued first (the appends below are branches, across which the compiler does not
        // move loads).  The append is predicated (three selects shift the pending entries);
        // only the 16-byte store of every fourth entry is a branch.  The capacity check of
        // the stores is needed only when some lane could reach max_nb within this tile.
        auto pairs = [&](auto check_cap, auto check_self) {
            for (int e = 0; e < m; e += NN_PB) {
                int js[NN_PB];
                double xs[NN_PB], ys[NN_PB], zs[NN_PB], h2s[NN_PB];
#pragma unroll
                for (int u = 0; u < NN_PB; u++) {
                    const int eu = e + u < m ? e + u : m - 1;   // m >= 1: a tile is never empty
                    xs[u] = L.tx[eu];
                    ys[u] = L.ty[eu];
                    zs[u] = L.tz[eu];
                    h2s[u] = L.th2[eu];
                    js[u] = L.tj[eu];
                }
#pragma unroll
                for (int u = 0; u < NN_PB; u++) {
                    if (e + u < m) {
                        const int j = js[u];
                        const double dx = xi - xs[u];
                        const double dy = yi - ys[u];
                        const double dz = zi - zs[u];
                        const double d = nn_d2(dx, dy, dz);
                        const bool hit =
                            (d < hi2) & (d < h2s[u]) & (!decltype(check_self)::value || j != (int)k);
                        if (hit && (count & 3) == 3) {
                            if (!decltype(check_cap)::value || count < A.max_nb) *wp = make_int4(p0, p1, p2, j);
                            wp += A.stride;
                        }
                        p0 = hit ? p1 : p0;
                        p1 = hit ? p2 : p1;
                        p2 = hit ? j : p2;
                        count += hit ? 1 : 0;
                    }
                }
            }
        };
        const bool self = first_tile || w0 == w1;
        first_tile = false;
        if (wave_ballot(count + m > A.max_nb) == 0) {
            if (self) pairs(std::false_type{}, std::true_type{});
            else pairs(std::false_type{}, std::false_type{});
        } else {
            if (self) pairs(std::true_type{}, std::true_type{});
            else pairs(std::true_type{}, std::false_type{});
        }
    });
    if (!ok) {
        if (lane == 0) set_err(A.err, ERRF_STACK);
        return;
    }
    if (MODE == 2 && active && (count & 3) != 0 && count < A.max_nb)
        A.nbr4[(int64_t)(count >> 2) * A.stride + k] = tail4();
    if (MODE == 0 && A.only) {
        // re-list of flagged particles: counts and caller-order outputs are published for
        // all particles by the density kernel afterwards
        if (active) {
            if ((count & 3) != 0 && count < A.max_nb)
                A.nbr4[(int64_t)(count >> 2) * A.stride + k] = tail4();
            A.snoi[k] = count;
            if (count > A.max_nb) set_err(A.err, ERRF_CAP_NEIGHBOURS);
        }
    } else if (MODE == 0) {
        const bool own_list = valid && !outlier;
        if (own_list && (count & 3) != 0 && count < A.max_nb)
            A.nbr4[(int64_t)(count >> 2) * A.stride + k] = tail4();
        if (own_list) {
            A.snoi[k] = count;
            const int32_t i = A.perm[k];
            A.noi_out[i] = count;
            if (A.sml_out) A.sml_out[i] = hi;
            if (A.iters_out) A.iters_out[i] = A.siters[k];
            if (count > A.max_nb) set_err(A.err, ERRF_CAP_NEIGHBOURS);
        }
        unsigned long long s = (unsigned long long)wave_sum(own_list ? count : 0);
        int mx = wave_max(own_list ? count : 0);
        if (lane == 0) {
            atomicAdd(&cnt_slot(A.counters)[CNT_NEIGHBOURS], s);
            atomicMax(&cnt_slot(A.counters)[CNT_MAX_NOI], (unsigned long long)mx);
        }
    } else {
        bool stay = false;
        if (valid && !outlier) {
            if (active) {
                A.iters[k] = A.pass + 1;
                A.snoi[k] = count;
                double hn;
                const bool fin = h_update(A, count, hi, hn);
                A.h_next[k] = hn;
                if (fin) A.done[k] = 1;
                stay = !fin;
            } else {
                A.h_next[k] = hi;
            }
        }
        if (A.list_out) {
            // append the unconverged particles (wave-aggregated; list order is free, the
            // next pass's results do not depend on it)
            const uint64_t mS = wave_ballot(stay);
            uint32_t b = 0;
            if (lane == 0 && mS) b = atomicAdd(A.count_out, (uint32_t)popc64(mS));
            b = __shfl(b, 0);
            if (stay) {
                const uint32_t pos = b + popc64(mS & lanemask_lt());
                A.list_out[pos] = (int32_t)k;
                if (A.h0_slot) A.h0_slot[pos] = hi;   // the h this pass's list was built with
            }
        }
        if (lane == 0) atomicAdd(&cnt_slot(A.counters)[CNT_H_COUNTS], (unsigned long long)popc64(mActive));
    }
}

// Per-particle mode: one wavefront per listed particle, searching the particle's own box
// x +- h with the 64 lanes testing 64 candidates at a time.  Used for the outliers of the
// group mode (particles whose h is far above the rest of their group) and for list fix-ups.
// MODE 1: count and h update (identical to the group mode's; the new h goes to h_next[k]);
// MODE 2: MODE 1 plus the list; MODE 0: the list only.  Lists are written in candidate
// order, one 4-byte entry per lane.
#define NN1_WAVES 4
#define NN1_MAX_BLOCKS 1024   // about one resident round: 4096 wavefronts loop over the list
template <int MODE>
__global__ __launch_bounds__(64 * NN1_WAVES) void k_nn_one(NNArgs A) {
    __shared__ NNWaveLds lds_all[NN1_WAVES];
    const int w = threadIdx.x / 64, lane = wave_lane();
    NNWaveLds &L = lds_all[w];
    const uint32_t cnt = *A.count_in;
    const int64_t nwaves = (int64_t)gridDim.x * NN1_WAVES;
    const uint64_t lt = lanemask_lt();
    int32_t *nbr32 = reinterpret_cast<int32_t *>(A.nbr4);
    for (int64_t t = (int64_t)blockIdx.x * NN1_WAVES + w; t < cnt; t += nwaves) {
        const int64_t k = A.list_in[t];
        const double4 own = A.sxm[k];
        const double hi = A.h[k];
        const double hi2 = hi * hi;
        const double grow = hi * (1.0 + 1e-9);
        int count = 0;
        // one candidate per lane: pair test, then the hits appended in lane order
        auto take = [&](bool v, int j, double x, double y, double z, double h2) {
            bool hit = false;
            if (v) {
                const double dx = own.x - x;
                const double dy = own.y - y;
                const double dz = own.z - z;
                const double d = nn_d2(dx, dy, dz);
                hit = d < hi2 && d < h2 && j != (int)k;
            }
            const uint64_t mh = wave_ballot(hit);
            if (MODE != 1 && hit) {
                const int s = count + popc64(mh & lt);
                if (s < A.max_nb) nbr32[((int64_t)(s >> 2) * A.stride + k) * 4 + (s & 3)] = j;
            }
            count += popc64(mh);
        };
        const double lox = own.x - grow, hix = own.x + grow, loy = own.y - grow, hiy = own.y + grow;
        const double loz = own.z - grow, hiz = own.z + grow;
        const bool direct = nn_direct(L, A, lox, hix, loy, hiy, loz, hiz, nn_point(own.x, own.y, own.z, hi), 0, 1,
                                      [&](bool v, int j, const double4 &p, double h) { take(v, j, p.x, p.y, p.z, h * h); });
        if (!direct) {
            const bool ok = nn_walk_box<false>(L, A, lox, hix, loy, hiy, loz, hiz, nn_point(own.x, own.y, own.z, hi), NN_START_ALL, 0, 0, -1, KeepAll(), [&](int m) {
                const bool v = lane < m;
                take(v, v ? L.tj[lane] : -1, v ? L.tx[lane] : 0.0, v ? L.ty[lane] : 0.0, v ? L.tz[lane] : 0.0,
                     v ? L.th2[lane] : 0.0);
            });
            if (!ok) {
                if (lane == 0) set_err(A.err, ERRF_STACK);
                return;
            }
        }
        if (MODE == 0) {
            if (lane == 0) {
                A.snoi[k] = count;
                if (A.noi_out) A.noi_out[A.perm[k]] = count;
                if (count > A.max_nb) set_err(A.err, ERRF_CAP_NEIGHBOURS);
                if (A.counters) {
                    atomicAdd(&cnt_slot(A.counters)[CNT_NEIGHBOURS], (unsigned long long)count);
                    atomicMax(&cnt_slot(A.counters)[CNT_MAX_NOI], (unsigned long long)count);
                }
            }
        } else {
            double hn;
            const bool fin = h_update(A, count, hi, hn);
            if (lane == 0) {
                A.iters[k] = A.pass + 1;
                A.snoi[k] = count;
                A.h_next[k] = hn;
                if (fin) A.done[k] = 1;
                else {
                    const uint32_t pos = atomicAdd(A.count_out, 1u);
                    A.list_out[pos] = (int32_t)k;
                    if (A.h0_slot) A.h0_slot[pos] = hi;
                }
                atomicAdd(&cnt_slot(A.counters)[CNT_H_COUNTS], 1ull);
            }
        }
        wave_sync();
    }
}

// Dirty marking after a speculative pass 0: for every particle j whose h changed after
// pass 0 (list_in, with its pass-0 h in h0_slot), flag j and every particle i whose pair
// membership with j changed, (d < min(h_i, h_j0)) != (d < min(h_i, h_j1)) with h_j1 the final
// h: the only pairs whose membership d < h_i and d < h_j can differ from pass 0 (an unflagged
// particle keeps its pass-0 h).  Such an i lies within max(h_j0, h_j1) of x_j.  One wavefront
// per changed particle; lanes are candidates; d is the pair test's (symmetric) value.
__global__ __launch_bounds__(64 * NN1_WAVES) void k_nn_mark_dirty(NNArgs A) {
    __shared__ NNWaveLds lds_all[NN1_WAVES];
    const int w = threadIdx.x / 64, lane = wave_lane();
    NNWaveLds &L = lds_all[w];
    const uint32_t cnt = *A.count_in;
    const int64_t nwaves = (int64_t)gridDim.x * NN1_WAVES;
    for (int64_t t = (int64_t)blockIdx.x * NN1_WAVES + w; t < cnt; t += nwaves) {
        const int64_t jc = A.list_in[t];
        const double4 own = A.sxm[jc];
        const double h0 = A.h0_slot[t], h1 = A.h[jc];
        const double h02 = h0 * h0, h12 = h1 * h1;
        const double hq = h0 > h1 ? h0 : h1;
        const double grow = hq * (1.0 + 1e-9);
        if (lane == 0) A.dirty[jc] = 1;
        auto mark = [&](bool v, int i, double x, double y, double z, double hi2) {
            if (v) {
                const double dx = own.x - x;
                const double dy = own.y - y;
                const double dz = own.z - z;
                const double d = nn_d2(dx, dy, dz);
                if (((d < hi2) & (d < h02)) != ((d < hi2) & (d < h12))) A.dirty[i] = 1;
            }
        };
        const double lox = own.x - grow, hix = own.x + grow, loy = own.y - grow, hiy = own.y + grow;
        const double loz = own.z - grow, hiz = own.z + grow;
        const bool direct = nn_direct(L, A, lox, hix, loy, hiy, loz, hiz, nn_point(own.x, own.y, own.z, hq), 0, 1,
                                      [&](bool v, int i, const double4 &p, double h) { mark(v, i, p.x, p.y, p.z, h * h); });
        if (!direct) {
            const bool ok = nn_walk_box<false>(L, A, lox, hix, loy, hiy, loz, hiz, nn_point(own.x, own.y, own.z, hq), NN_START_ALL, 0, 0, -1, KeepAll(), [&](int m) {
                if (lane < m) mark(true, L.tj[lane], L.tx[lane], L.ty[lane], L.tz[lane], L.th2[lane]);
            });
            if (!ok) {
                if (lane == 0) set_err(A.err, ERRF_STACK);
                return;
            }
        }
        wave_sync();
    }
}

// Counting pass of the h-iteration for the listed particles, direct form (nn_direct): one
// wavefront per particle tests every particle of the start cells of its box x +- h, with no
// tree walk.  (Sharing large boxes among several wavefronts gained 0.07 ms on an evolved state
// and cost more than that on the others: the extra wavefronts of the small boxes.)  A box the
// direct form declines (too large for the table, too many cells or candidates) is appended to
// hard_list for k_nn_count_split.  The count goes to cnt_slot[t] (integer, exact).
__global__ __launch_bounds__(64 * NN1_WAVES) void k_nn_count_direct(NNArgs A) {
    __shared__ NNWaveLds lds_all[NN1_WAVES];
    const int w = threadIdx.x / 64, lane = wave_lane();
    NNWaveLds &L = lds_all[w];
    if (h_pass_frozen(A)) return;
    const uint32_t cnt = *A.count_in;
    const int64_t nwaves = (int64_t)gridDim.x * NN1_WAVES;
    for (int64_t t = (int64_t)blockIdx.x * NN1_WAVES + w; t < cnt; t += nwaves) {
        const int64_t k = A.list_in[t];
        const double4 own = A.sxm[k];
        const double hi = A.h[k];
        const double hi2 = hi * hi;
        const double grow = hi * (1.0 + 1e-9);
        int count = 0;
        const bool direct = nn_direct(L, A, own.x - grow, own.x + grow, own.y - grow, own.y + grow, own.z - grow,
                                      own.z + grow, nn_point(own.x, own.y, own.z, hi), 0, 1,
                                      [&](bool v, int j, const double4 &p, double h) {
            if (v) {
                const double dx = own.x - p.x;
                const double dy = own.y - p.y;
                const double dz = own.z - p.z;
                const double d = nn_d2(dx, dy, dz);
                count += (d < hi2 && d < h * h && j != (int)k) ? 1 : 0;
            }
        });
        if (!direct) {
            if (lane == 0) A.list_out[atomicAdd(A.count_out, 1u)] = (int32_t)t;   // hard_list (slots)
            continue;
        }
        count = wave_sum(count);
        if (lane == 0) A.cnt_slot[t] = (uint32_t)count;
        wave_sync();
    }
}

// Counting pass of the h-iteration for the particles the direct form left (slots in
// slot_list), split eight ways: wavefront (t, o) counts particle list_in[t]'s neighbours
// under the start cells of index = o (mod 8) (the root's child o when the box is too large
// for the start table) and adds them to cnt_slot[t] (integer sums: exact and order-free).
// A particle with a large h (a free-boundary particle whose h keeps growing) is so shared
// by eight wavefronts; k_h_finish then applies the update.
__global__ __launch_bounds__(64 * NN1_WAVES) void k_nn_count_split(NNArgs A) {
    __shared__ NNWaveLds lds_all[NN1_WAVES];
    const int w = threadIdx.x / 64, lane = wave_lane();
    NNWaveLds &L = lds_all[w];
    if (h_pass_frozen(A)) return;
    const int64_t work = 8 * (int64_t)*A.slot_count;
    const int64_t nwaves = (int64_t)gridDim.x * NN1_WAVES;
    for (int64_t tt = (int64_t)blockIdx.x * NN1_WAVES + w; tt < work; tt += nwaves) {
        const int64_t t = A.slot_list[tt >> 3];
        const int o = (int)(tt & 7);
        const int64_t k = A.list_in[t];
        const double4 own = A.sxm[k];
        const double hi = A.h[k];
        const double hi2 = hi * hi;
        const double grow = hi * (1.0 + 1e-9);
        const int start = NN_START_PART0 - o;   // eighth o of the start cells
        int count = 0;
        const bool ok = nn_walk_box<false>(L, A, own.x - grow, own.x + grow, own.y - grow, own.y + grow, own.z - grow,
                                    own.z + grow, nn_point(own.x, own.y, own.z, hi), start, 0, 0, -1, KeepAll(), [&](int m) {
            bool hit = false;
            if (lane < m) {
                const double dx = own.x - L.tx[lane];
                const double dy = own.y - L.ty[lane];
                const double dz = own.z - L.tz[lane];
                const double d = nn_d2(dx, dy, dz);
                hit = d < hi2 && d < L.th2[lane] && L.tj[lane] != (int)k;
            }
            count += popc64(wave_ballot(hit));
        });
        if (!ok) {
            if (lane == 0) set_err(A.err, ERRF_STACK);
            return;
        }
        if (lane == 0 && count > 0) atomicAdd(&A.cnt_slot[t], (uint32_t)count);
        wave_sync();
    }
}

// the h update of the listed particles from their counts (after the counting pass, so every
// count saw the h of this pass; each thread writes only its own particle's h): clears
// cnt_slot for the next pass and appends the unconverged particles to list_out
__global__ __launch_bounds__(256) void k_h_finish(NNArgs A) {
    const uint32_t cnt = *A.count_in;
    const bool frozen = h_pass_frozen(A);
    const int lane = wave_lane();
    const uint64_t lt = lanemask_lt();
    for (int64_t t0 = blockIdx.x * (int64_t)blockDim.x; t0 < cnt; t0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = t0 + threadIdx.x;
        const bool valid = t < cnt;
        bool stay = false, changed = false;
        int64_t k = 0;
        if (valid) {
            k = A.list_in[t];
            const int N = frozen ? A.snoi[k] : (int)A.cnt_slot[t];
            A.cnt_slot[t] = 0u;
            const double hi = A.h[k];
            double hn;
            const bool fin = h_update(A, N, hi, hn);
            A.iters[k] = A.pass + 1;
            A.snoi[k] = N;
            A.h_next[k] = hn;
            if (fin) A.done[k] = 1;
            stay = !fin;
            // the next pass counts with hn: the cell bounds must cover it (they cover hi)
            changed = stay && hn != hi;
            if (changed && A.chmax_raise) hmax_raise(A.chmax_raise, A.ppar, A.cpar, A.cap_cells, k, hn);
        }
        // wave-aggregated list append and counter (list order is free)
        const uint64_t mS = wave_ballot(stay);
        uint32_t b = 0;
        if (lane == 0 && mS != 0) b = atomicAdd(A.count_out, (uint32_t)popc64(mS));
        b = __shfl(b, 0);
        if (stay) A.list_out[b + popc64(mS & lt)] = (int32_t)k;
        const uint64_t mV = wave_ballot(valid);
        if (lane == 0 && mV != 0) atomicAdd(&cnt_slot(A.counters)[CNT_H_COUNTS], (unsigned long long)popc64(mV));
        if (A.changed && wave_ballot(changed) != 0 && lane == 0) *A.changed = 1u;
    }
}

// Start table of the current tree: a particle alone in its level-l prefix for every
// l >= its leaf level, and every cell of level <= kmax, written at their Lebesgue prefix.
__global__ __launch_bounds__(256) void k_nn_start_table(int64_t n, int kmax, const uint8_t *__restrict__ lcp,
                                                        const uint64_t *__restrict__ slkey,
                                                        const int32_t *__restrict__ cell_base,
                                                        const int32_t *__restrict__ cstart,
                                                        const int32_t *__restrict__ cend,
                                                        const int32_t *__restrict__ clevel, int64_t cap_cells,
                                                        int32_t *__restrict__ tab, int2 *__restrict__ rng) {
    // particles, then the cells of levels 1..kmax (cells are numbered level by level, so these
    // are cells 1 .. cell_base[kmax + 1] - 1); grid-stride over the actual counts
    const int64_t ncells = cell_base[MHPC_LEVELS + 1];
    if (ncells > cap_cells) return;   // capacity error already flagged; cells are incomplete
    const int64_t ctop = cell_base[kmax + 1] < ncells ? cell_base[kmax + 1] : ncells;
    const int64_t items = n + (ctop > 1 ? ctop - 1 : 0);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < items;
         t += (int64_t)gridDim.x * blockDim.x) {
        if (t < n) {
            const int a = t > 0 ? (int)lcp[t - 1] : 0;
            const int b = (int)lcp[t];
            const int lp = (a > b ? a : b) + 1;
            const uint64_t key = slkey[t];
            for (int l = lp; l <= kmax; l++) {
                const int64_t e = nn_tab_off(l) + (int64_t)(key >> (3 * (MHPC_LEVELS - l)));
                tab[e] = (int32_t)t;
                rng[e] = make_int2((int)t, 1);
            }
        } else {
            const int64_t c = t - n + 1;   // cells 1 .. ctop-1 (cell 0 is the root)
            const int L = clevel[c];
            if (L < 1 || L > kmax) continue;
            const int32_t s0 = cstart[c];
            const int64_t e = nn_tab_off(L) + (int64_t)(slkey[s0] >> (3 * (MHPC_LEVELS - L)));
            tab[e] = (int32_t)(n + c);
            rng[e] = make_int2((int)s0, (int)(cend[c] - s0 + 1));
        }
    }
}

static void ensure_nn_start(mhpc_ctx *c) {
    if (c->nn_start_valid || c->n <= 0) return;
    launch_fill_u32(c, reinterpret_cast<uint32_t *>(c->nn_start), 0xffffffffu, nn_tab_off(c->nn_kmax + 1));
    const uint64_t *slkey = (c->p.curve == MHPC_CURVE_HILBERT) ? c->slkey : c->skey;
    const int64_t items = c->n + c->cap_cells;
    hipLaunchKernelGGL(k_nn_start_table, dim3((unsigned)std::min<int64_t>(ceil_div(items, 256), 8192)), dim3(256),
                       0, c->stream, c->n,
                       c->nn_kmax, c->lcp, slkey, c->cell_base, c->cstart, c->cend, c->clevel, c->cap_cells,
                       c->nn_start, c->nn_start_rng);
    c->nn_start_valid = true;
}

static NNArgs nn_common(mhpc_ctx *c) {
    ensure_nn_start(c);
    NNArgs a = {};
    a.stab = c->nn_start;
    a.srng = c->nn_start_rng;
    a.kmax = c->nn_kmax;
    a.box = c->box;
    a.n = c->n;
    a.perm = c->perm;
    a.sxm = c->sxm;
    a.cstart = c->cstart; a.cend = c->cend; a.cchild = c->cchild;
    a.cbmin = c->cbmin; a.cbmax = c->cbmax;
    a.cap_cells = c->cap_cells;
    a.max_nb = c->p.max_nb;
    a.err = c->err;
    a.counters = c->counters;
    a.chmax = c->chmax;
    return a;
}

static unsigned nn_grid(mhpc_ctx *c) { return (unsigned)ceil_div(ceil_div(c->n, 64), NN_WAVES); }

// done flags, iteration counts and dirty flags to zero; the per-pass list counters, the
// per-pass changed flags and the outlier counters (pass 0, list fix-up) to zero
static_assert(offsetof(DevStatus, hchanged) == offsetof(DevStatus, hlist_count) + sizeof(DevStatus::hlist_count),
              "k_h_init zeroes the list counters and the changed flags as one range");
__global__ void k_h_init(int64_t n, uint8_t *done, int32_t *iters, uint8_t *dirty, uint32_t *list_counts,
                         int nlc, uint32_t *ocount, uint32_t *cnt_slot) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) {
        done[k] = 0;
        iters[k] = 0;
        dirty[k] = 0;
        cnt_slot[k] = 0u;
    }
    if (k < nlc) list_counts[k] = 0;
    if (k < 2) ocount[k] = 0;
}

// the raise for the listed particles (after pass 0: the particles still iterating are the ones
// whose h changed)
__global__ __launch_bounds__(256) void k_hmax_raise(const int32_t *__restrict__ list, const uint32_t *count,
                                                    const double *__restrict__ h, const int32_t *__restrict__ ppar,
                                                    const int32_t *__restrict__ cpar, double *chmax,
                                                    int64_t cap_cells) {
    const uint32_t cnt = *count;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cnt; t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = list[t];
        hmax_raise(chmax, ppar, cpar, cap_cells, k, h[k]);
    }
}

static void launch_hmax_raise(mhpc_ctx *c, const int32_t *list, const uint32_t *count) {
    hipLaunchKernelGGL(k_hmax_raise, dim3(64), dim3(256), 0, c->stream, list, count, c->sh, c->ppar, c->cpar,
                       c->chmax, c->cap_cells);
}

// a9: the plain h-iteration as Jacobi passes; sh holds the final h afterwards.  Pass 0
// runs in group mode over all particles and lists the unconverged ones; later passes run
// in per-particle mode over that list only.
void launch_h_iterate(mhpc_ctx *c, int passes, bool speculative, hipEvent_t *walk_events) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_h_init, dim3((unsigned)ceil_div(c->n > 64 ? c->n : 64, 256)), dim3(256), 0, c->stream,
                       c->n, c->sdone, c->siters, c->sdirty, c->hlist_count, 3 * (MHPC_MAX_PASSES + 1), c->ocount,
                       c->hcnt);
    NNArgs a = nn_common(c);
    a.done = c->sdone;
    a.iters = c->siters;
    a.snoi = c->snoi;
    a.table = c->htable;
    a.tmax = MHPC_HTABLE_MAX;
    a.max_counts = passes;
    a.band_lo = c->p.band_lo;
    a.band_hi = c->p.band_hi;
    a.sml_max = c->p.sml_max;
    // pass 0, group mode; its list of unconverged particles (hlist[2]) is also the list
    // of particles whose h changed after pass 0
    a.pass = 0;
    a.h = c->sh;
    a.h_next = c->sh2;
    a.list_out = c->hlist[2];
    a.count_out = c->hlist_count + 0;
    a.h0_slot = c->h0slot;
    const unsigned g1 = (unsigned)std::min<int64_t>(ceil_div(c->n, NN1_WAVES), NN1_MAX_BLOCKS);
    a.out_list = c->olist;
    a.out_count = c->ocount;   // zeroed by k_h_init
    if (walk_events) (void)hipEventRecord(walk_events[0], c->stream);
    if (speculative) {
        a.nbr4 = reinterpret_cast<int4 *>(c->nbr);
        a.stride = c->cap_pad;
        hipLaunchKernelGGL(k_nn_walk<2>, dim3(nn_grid(c)), dim3(64 * NN_WAVES), 0, c->stream, a);
    } else {
        hipLaunchKernelGGL(k_nn_walk<1>, dim3(nn_grid(c)), dim3(64 * NN_WAVES), 0, c->stream, a);
    }
    if (walk_events) (void)hipEventRecord(walk_events[1], c->stream);
    // the group mode's outliers, same pass, per-particle mode
    {
        NNArgs o = a;
        o.out_list = nullptr;
        o.list_in = c->olist;
        o.count_in = c->ocount;
        if (speculative) hipLaunchKernelGGL(k_nn_one<2>, dim3(g1), dim3(64 * NN1_WAVES), 0, c->stream, o);
        else hipLaunchKernelGGL(k_nn_one<1>, dim3(g1), dim3(64 * NN1_WAVES), 0, c->stream, o);
    }
    a.nbr4 = nullptr;
    a.out_list = nullptr;
    a.h0_slot = nullptr;   // the pass-0 slots stay for the dirty marking
    std::swap(c->sh, c->sh2);
    // the particles still iterating are the ones whose h changed
    launch_hmax_raise(c, c->hlist[2], c->hlist_count + 0);
    // passes 1.., per-particle mode over the list of the previous pass
    for (int pass = 1; pass < passes; pass++) {
        a.pass = pass;
        a.h = c->sh;
        a.list_in = pass == 1 ? c->hlist[2] : c->hlist[(pass - 1) & 1];
        a.count_in = c->hlist_count + (pass - 1);
        a.list_out = c->hlist[pass & 1];
        a.count_out = c->hlist_count + pass;
        a.h_next = c->sh;   // k_h_finish updates h in place once every count is done
        a.cnt_slot = c->hcnt;
        a.changed = c->dstatus->hchanged + pass;   // zeroed by k_h_init
        a.prev_changed = pass >= 2 ? c->dstatus->hchanged + (pass - 1) : nullptr;
        // direct counts; the particles they cannot take go to the split walk
        uint32_t *hard_count = c->hlist_count + MHPC_MAX_PASSES + 1 + pass;   // zeroed by k_h_init
        NNArgs d = a;
        d.list_out = c->hhard;
        d.count_out = hard_count;
        hipLaunchKernelGGL(k_nn_count_direct, dim3(g1), dim3(64 * NN1_WAVES), 0, c->stream, d);
        a.slot_list = c->hhard;
        a.slot_count = hard_count;
        // the split work is 8 wavefronts per particle the direct form declined: usually none
        hipLaunchKernelGGL(k_nn_count_split, dim3(std::min(g1, 256u)), dim3(64 * NN1_WAVES), 0, c->stream, a);
        NNArgs f = a;
        f.chmax_raise = c->chmax;
        f.ppar = c->ppar;
        f.cpar = c->cpar;
        hipLaunchKernelGGL(k_h_finish, dim3(64), dim3(256), 0, c->stream, f);
    }
}

// After a speculative h-iteration (launch_h_iterate with lists from pass 0; its k_h_init
// zeroed the dirty flags and the second outlier counter): re-list exactly the particles whose
// pass-0 list may differ from the list with the final h (the changed particles and every
// particle within its own h of one).  The lists then equal those of a full list walk with the
// final h as sets (order may differ); the density kernel publishes the counts.
void launch_neighbours_fixup(mhpc_ctx *c) {
    if (c->n <= 0) return;
    NNArgs a = nn_common(c);
    a.h = c->sh;
    a.list_in = c->hlist[2];
    a.count_in = c->hlist_count + 0;
    a.dirty = c->sdirty;
    a.h0_slot = c->h0slot;
    const unsigned g1 = (unsigned)std::min<int64_t>(ceil_div(c->n, NN1_WAVES), NN1_MAX_BLOCKS);
    hipLaunchKernelGGL(k_nn_mark_dirty, dim3(g1), dim3(64 * NN1_WAVES), 0, c->stream, a);
    NNArgs b = nn_common(c);
    b.h = c->sh;
    b.nbr4 = reinterpret_cast<int4 *>(c->nbr);
    b.stride = c->cap_pad;
    b.snoi = c->snoi;
    b.only = c->sdirty;
    b.out_list = c->olist;
    b.out_count = c->ocount + 1;
    hipLaunchKernelGGL(k_nn_walk<0>, dim3(nn_grid(c)), dim3(64 * NN_WAVES), 0, c->stream, b);
    b.out_list = nullptr;
    b.list_in = c->olist;
    b.count_in = c->ocount + 1;
    b.counters = nullptr;   // the density kernel counts every particle
    b.noi_out = nullptr;
    hipLaunchKernelGGL(k_nn_one<0>, dim3(g1), dim3(64 * NN1_WAVES), 0, c->stream, b);
    // h, counts and the neighbour sums go to caller order with the density kernel
    // (launch_density_eos(c, true)), which runs next and visits every particle
}

void launch_neighbours(mhpc_ctx *c) {
    if (c->n <= 0) return;
    NNArgs a = nn_common(c);
    a.h = c->sh;
    a.nbr4 = reinterpret_cast<int4 *>(c->nbr);
    a.stride = c->cap_pad;
    a.snoi = c->snoi;
    a.noi_out = c->noi;
    a.sml_out = nullptr;
    a.siters = c->siters;
    a.iters_out = nullptr;
    launch_fill_u32(c, c->ocount, 0u, 1);
    a.out_list = c->olist;
    a.out_count = c->ocount;
    hipLaunchKernelGGL(k_nn_walk<0>, dim3(nn_grid(c)), dim3(64 * NN_WAVES), 0, c->stream, a);
    a.out_list = nullptr;
    a.list_in = c->olist;
    a.count_in = c->ocount;
    const unsigned g1 = (unsigned)std::min<int64_t>(ceil_div(c->n, NN1_WAVES), NN1_MAX_BLOCKS);
    hipLaunchKernelGGL(k_nn_one<0>, dim3(g1), dim3(64 * NN1_WAVES), 0, c->stream, a);
}

// ------------------------------------------------------------------------------------
// a10 cubic spline (kernel.cu:50-83), support radius h.  FP64 form: the reference's
// expressions in the same order; FP32 form: the same formula in float.
#define MHPC_8_OVER_PI_F 2.546479089470325f

__device__ __forceinline__ void cubic32(float r, float h, float &W, float &dWdr) {
    const float hinv = 1.0f / h;
    const float q = r * hinv;
    const float f = MHPC_8_OVER_PI_F * hinv * hinv * hinv;
    const float omq = 1.0f - q;
    if (q > 1.0f) {
        W = 0.f;
        dWdr = 0.f;
    } else if (q > 0.5f) {
        W = 2.f * f * omq * omq * omq;
        dWdr = -6.f * f * hinv * omq * omq;
    } else {
        W = f * __builtin_fmaf(6.f * q * q, q - 1.0f, 1.0f);
        dWdr = 6.f * f * hinv * q * __builtin_fmaf(3.f, q, -2.f);
    }
}

// ------------------------------------------------------------------------------------
// a11 density (density.cu:8-72): rho_i = m_i W(0, h_i) + sum_j m_j W(r_ij, h_ij).
// Neighbour records are read as one 32-byte {x,y,z,m} plus h; the loop is unrolled so a
// wave keeps several independent gathers in flight.
#define K8PI 2.5464790894703255   // 8 / pi

// 1/x and 1/sqrt(x) from the hardware estimates (v_rcp_f64, v_rsq_f64) refined by two
// Newton-Raphson steps each, for positive normal x: a few FP64 ulp, several instructions
// fewer than a correctly rounded division / square root
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

__device__ __forceinline__ double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    double e = fma(-hx, y * y, 0.5);
    y = fma(y, e, y);
    e = fma(-hx, y * y, 0.5);
    return fma(y, e, y);
}

// a10, the five kernels of SPH::SmoothingKernel (kernel.cu:14-183, 3D branches; Dehnen & Aly
// 2012 for the Wendland functions), selected by mhpc_params.smoothing_kernel as the
// reference's Smoothing::Kernel (parameter.h:177-180).  Normalised form: for q = r/h <= 1,
// W = sigma/h^3 w(q) and dW/dr = sigma/h^4 w'(q); both are 0 beyond.  The cubic spline keeps
// its own inlined form in the density and force loops; the others go through kernel_shape.
// The Wendland C4 gradient is the reference's expression (kernel.cu:140), which is not the
// derivative of its W; it is followed as written so that results are the reference's.
__device__ __forceinline__ double kernel_sigma(int kind) {
    switch (kind) {
        case MHPC_KERNEL_SPIKY: return 15.0 / M_PI;
        case MHPC_KERNEL_WENDLAND_C2: return 21.0 / (2.0 * M_PI);
        case MHPC_KERNEL_WENDLAND_C4: return 495.0 / (32.0 * M_PI);
        case MHPC_KERNEL_WENDLAND_C6: return 1365.0 / (64.0 * M_PI);
        default: return 8.0 / M_PI;
    }
}

template <typename T>
__device__ __forceinline__ void kernel_shape(int kind, T q, T &w, T &dw) {
    const T t = T(1) - q;
    const T t2 = t * t, t3 = t2 * t;
    switch (kind) {
        case MHPC_KERNEL_SPIKY:
            w = t3;
            dw = T(-3) * t2;
            break;
        case MHPC_KERNEL_WENDLAND_C2:
            w = t2 * t2 * (T(1) + T(4) * q);
            dw = T(-20) * q * t3;
            break;
        case MHPC_KERNEL_WENDLAND_C4:
            w = t3 * t3 * (T(1) + q * (T(6) + T(35.0 / 3.0) * q));
            dw = T(-6) * t3 * t2 * (T(1) + q * q * (T(-35) + T(105) * q));
            break;
        case MHPC_KERNEL_WENDLAND_C6:
            w = t2 * t2 * t2 * t2 * (T(1) + q * (T(8) + q * (T(25) + T(32) * q)));
            dw = T(-22) * q * t3 * t2 * t2 * (T(1) + q * (T(7) + T(16) * q));
            break;
        default:   // cubic spline
            if (q <= T(0.5)) {
                w = T(1) + T(6) * q * q * (q - T(1));
                dw = T(6) * q * (T(3) * q - T(2));
            } else {
                w = T(2) * t3;
                dw = T(-6) * t2;
            }
            break;
    }
}

// ------------------------------------------------------------------------------------
// a12 / a13 sound speed (soundspeed.cu:38-97) and pressure (pressure.cu:6-87), fused;
// also writes the records the force loop gathers: cs into {vx,vy,vz,cs} and
// {rho, p/rho^2, m/rho, h}.
// cs_const > 0: the constant sound speed of MHPC_EOS_ISOTHERMAL_CONST_CS (otherwise that EOS
// keeps each particle's uploaded cs)
__device__ __forceinline__ double eos_cs(int eos, double gamma, double K, double cs_const, double rho, double p,
                                         double cs) {
    switch (eos) {
        case MHPC_EOS_POLYTROPIC: return sqrt(K * pow(rho, gamma - 1.0));
        case MHPC_EOS_IDEAL_GAS: return sqrt(gamma * p / rho);
        case MHPC_EOS_LOCALLY_ISOTHERMAL: return 0.0;
        case MHPC_EOS_ISOTHERMAL_CONST_CS: return cs_const > 0.0 ? cs_const : cs;
        default: return cs;
    }
}

__device__ __forceinline__ double eos_p(int eos, double gamma, double K, double rho, double e,
                                        double cs) {
    switch (eos) {
        case MHPC_EOS_POLYTROPIC: return K * pow(rho, gamma);
        case MHPC_EOS_ISOTHERMAL: return 41255.407 * rho;
        case MHPC_EOS_IDEAL_GAS: return (gamma - 1) * rho * e;
        case MHPC_EOS_LOCALLY_ISOTHERMAL:
        case MHPC_EOS_ISOTHERMAL_CONST_CS: return cs * cs * rho;
        default: return 0.0;
    }
}

// EOS of one particle and the records the force loop reads (shared by k_eos and the
// density kernel's fused form)
struct EosArgs {
    int eos, cs_before_p;
    double gamma, K, cs_const;
    const double *se;
    double *sp, *scs;
    double4 *sv4, *sr4;
    double *p_out, *cs_out;
};

__device__ __forceinline__ void eos_particle(const EosArgs &E, int64_t k, int32_t i, double rho, double m,
                                             double h) {
    const double e = E.se[k];
    double p = E.sp[k], cs = E.scs[k];
    if (E.eos == MHPC_EOS_ISOTHERMAL_CONST_CS) cs = eos_cs(E.eos, E.gamma, E.K, E.cs_const, rho, p, cs);
    if (E.cs_before_p) {
        cs = eos_cs(E.eos, E.gamma, E.K, E.cs_const, rho, p, cs);
        p = eos_p(E.eos, E.gamma, E.K, rho, e, cs);
    } else {
        p = eos_p(E.eos, E.gamma, E.K, rho, e, cs);
        cs = eos_cs(E.eos, E.gamma, E.K, E.cs_const, rho, p, cs);
    }
    E.sp[k] = p;
    E.scs[k] = cs;
    E.sv4[k].w = cs;
    E.sr4[k] = make_double4(rho, p / (rho * rho), m / rho, h);
    E.p_out[i] = p;
    E.cs_out[i] = cs;
}

static EosArgs eos_args(mhpc_ctx *c) {
    EosArgs E;
    E.eos = c->p.eos;
    E.cs_before_p = c->p.cs_before_p;
    E.gamma = c->p.gamma;
    E.K = c->p.K;
    E.cs_const = c->p.cs_const;
    E.se = c->se;
    E.sp = c->sp;
    E.scs = c->scs;
    E.sv4 = c->sv4;
    E.sr4 = c->sr4;
    E.p_out = c->f[MHPC_F_P];
    E.cs_out = c->f[MHPC_F_CS];
    return E;
}

// one density pair term: FP64 adds m_j W into acc; FP32 adds it (in FP32, from FP64
// differences) into the Kahan accumulator acc32
template <bool FP64, bool CUBIC>
__device__ __forceinline__ void density_pair(const double4 &pi, double hi, const double4 &pj, double hj,
                                             int sml_avg, int kind, double &acc, Kahan &acc32) {
    if constexpr (!CUBIC) {
        if constexpr (FP64) {
            const double hb = sml_avg ? 0.5 * (hi + hj) : hi;
            const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            const double ih = rcp_nr(hb);
            const double q = r2 * rsq_nr(r2) * ih;
            if (q <= 1.0) {
                double w, dw;
                kernel_shape<double>(kind, q, w, dw);
                acc = fma(pj.w, kernel_sigma(kind) * (ih * ih * ih) * w, acc);
            }
        } else {
            const float hb = (float)(sml_avg ? 0.5 * (hi + hj) : hi);
            const float dx = (float)(pi.x - pj.x), dy = (float)(pi.y - pj.y), dz = (float)(pi.z - pj.z);
            const float ih = 1.0f / hb;
            const float q = sqrtf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz))) * ih;
            if (q <= 1.0f) {
                float w, dw;
                kernel_shape<float>(kind, q, w, dw);
                acc32.add((float)pj.w * ((float)kernel_sigma(kind) * ih * ih * ih * w));
            }
        }
        return;
    }
    if constexpr (FP64) {
        const double hb = sml_avg ? 0.5 * (hi + hj) : hi;
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        const double ih = rcp_nr(hb);
        const double q = r2 * rsq_nr(r2) * ih;
        const double f = K8PI * (ih * ih * ih);
        double W = 0.0;
        if (q <= 0.5) W = f * fma(6.0 * q * q, q - 1.0, 1.0);
        else if (q <= 1.0) { const double t = 1.0 - q; W = 2.0 * f * (t * t * t); }
        acc = fma(pj.w, W, acc);
    } else {
        const float hb = (float)(sml_avg ? 0.5 * (hi + hj) : hi);
        const float dx = (float)(pi.x - pj.x), dy = (float)(pi.y - pj.y), dz = (float)(pi.z - pj.z);
        const float rr = sqrtf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)));
        float W, dW;
        cubic32(rr, hb, W, dW);
        acc32.add((float)pj.w * W);
    }
}

// caller-order h, neighbour and iteration counts and the neighbour sum / maximum, published by
// the density kernel when noi is set (the RHS with lists from pass 0, after the list fix-up)
struct DensPub {
    const int32_t *siters;
    double *sml;
    int32_t *noi, *iters;
    unsigned long long *counters;
};

// Neighbour window.  Per-lane gathers of neighbour records are bound by the L1 address path
// (each 16-byte load of a wave touches ~50 distinct cache lines).  A block covers a
// contiguous range of sorted particles -- a compact region of the curve -- and stages that
// range's records in LDS: a neighbour j inside the window is read from LDS at j - base (no
// lookup), the others are gathered (the masked loads of the out-of-window lanes are all the
// address path sees).  On a lattice 67 % of the pairs fall in a 512-particle window and 58 %
// in a 256-particle one.  Same arithmetic and order as a pure gather.
#define DENS_WIN 512

// EOS: the sound speed / pressure of the particle follow in the same kernel (the RHS path)
template <bool FP64, bool EOS, bool CUBIC>
__global__ __launch_bounds__(DENS_WIN) void k_density(int64_t n, const double4 *__restrict__ sxm,
                                                      const double *__restrict__ sh,
                                                      const int32_t *__restrict__ snoi,
                                                      const int32_t *__restrict__ nbr, int64_t stride,
                                                      int max_nb, int sml_avg, int kind,
                                                      const int32_t *__restrict__ perm,
                                                      double *__restrict__ srho, double *__restrict__ rho,
                                                      uint32_t *err, EosArgs E, DensPub P) {
    __shared__ double4 wp[DENS_WIN];
    __shared__ double wh[DENS_WIN];
    const int64_t base = xcd_block(blockIdx.x, gridDim.x) * (int64_t)DENS_WIN;
    const int tid = threadIdx.x;
    const int64_t k = base + tid;
    if (k < n) {
        wp[tid] = sxm[k];
        wh[tid] = sh[k];
    }
    __syncthreads();
    int pub_cnt = 0;   // the neighbour count this thread publishes (0 beyond n)
    if (k < n) {
        const double4 pi = wp[tid];
        const double hi = wh[tid];
        int cnt = snoi[k];
        cnt = cnt > max_nb ? max_nb : cnt;
        const int4 *nrow = reinterpret_cast<const int4 *>(nbr) + k;   // row r at nrow[r * stride]
        double acc = 0.0;
        Kahan acc32;
        // two list rows (eight neighbours) per iteration: the eight records are requested
        // together, then the pairs are evaluated in list order; the next two rows are
        // requested a step ahead.  The window copy is read unconditionally (clamped index),
        // the gather only where j is outside the window.
        int4 qn0 = make_int4(0, 0, 0, 0), qn1 = qn0;
        if (cnt > 0) qn0 = nrow[0];
        if (cnt > 4) qn1 = nrow[stride];
        for (int r = 0; 4 * r < cnt; r += 2) {
            const int4 qa = qn0, qb = qn1;
            const int m = cnt - 4 * r;   // entries left (>= 1)
            if (m > 8) qn0 = nrow[(int64_t)(r + 2) * stride];
            if (m > 12) qn1 = nrow[(int64_t)(r + 3) * stride];
            const int32_t jj[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
            double4 pj[8];
            double hj[8];
    #pragma unroll
            for (int e = 0; e < 8; e++) {
                const int32_t j = e < m ? jj[e] : (int32_t)k;
                const uint64_t o = (uint64_t)((int64_t)j - base);
                const bool in = o < (uint64_t)DENS_WIN;
                const int oc = in ? (int)o : 0;
                pj[e] = wp[oc];
                hj[e] = wh[oc];
                if (!in) {
                    pj[e] = sxm[j];
                    hj[e] = sh[j];
                }
            }
    #pragma unroll
            for (int e = 0; e < 8; e++)
                if (e < m) density_pair<FP64, CUBIC>(pi, hi, pj[e], hj[e], sml_avg, kind, acc, acc32);
        }
        const double self = CUBIC ? K8PI / (hi * hi * hi) : kernel_sigma(kind) / (hi * hi * hi);   // W(0)
        const double r = pi.w * self + (FP64 ? acc : acc32.value());
        srho[k] = r;
        const int32_t i = perm[k];
        rho[i] = r;
        if (!(r > 0.0)) set_err(err, ERRF_RHO);
        if (EOS) eos_particle(E, k, i, r, pi.w, hi);
        if (P.noi) {
            pub_cnt = snoi[k];
            P.noi[i] = pub_cnt;
            P.sml[i] = hi;
            P.iters[i] = P.siters[k];
        }
    }
    if (P.noi) {
        // every thread of the block takes part: one sum / max and one pair of atomics per block
        // (a pair per wavefront would queue tens of thousands of atomics on one address)
        __shared__ unsigned long long psum[DENS_WIN / 64];
        __shared__ int pmax[DENS_WIN / 64];
        const unsigned long long sum = wave_sum((unsigned long long)pub_cnt);
        const int mx = wave_max(pub_cnt);
        if (wave_lane() == 0) {
            psum[tid / 64] = sum;
            pmax[tid / 64] = mx;
        }
        __syncthreads();
        if (tid == 0) {
            unsigned long long t = 0;
            int m = 0;
            for (int w = 0; w < DENS_WIN / 64; w++) {
                t += psum[w];
                m = pmax[w] > m ? pmax[w] : m;
            }
            atomicAdd(&cnt_slot(P.counters)[CNT_NEIGHBOURS], t);
            atomicMax(&cnt_slot(P.counters)[CNT_MAX_NOI], (unsigned long long)m);
        }
    }
}

template <bool FP64, bool EOS>
static void launch_density_p(mhpc_ctx *c, const EosArgs &E, const DensPub &P) {
    const unsigned g = (unsigned)ceil_div(c->n, DENS_WIN);
    const int kind = c->p.smoothing_kernel;
    if (kind == MHPC_KERNEL_CUBIC_SPLINE)
        hipLaunchKernelGGL((k_density<FP64, EOS, true>), dim3(g), dim3(DENS_WIN), 0, c->stream, c->n, c->sxm, c->sh,
                           c->snoi, c->nbr, c->cap_pad, c->p.max_nb, c->p.sml_avg, kind, c->perm, c->srho,
                           c->f[MHPC_F_RHO], c->err, E, P);
    else
        hipLaunchKernelGGL((k_density<FP64, EOS, false>), dim3(g), dim3(DENS_WIN), 0, c->stream, c->n, c->sxm,
                           c->sh, c->snoi, c->nbr, c->cap_pad, c->p.max_nb, c->p.sml_avg, kind, c->perm, c->srho,
                           c->f[MHPC_F_RHO], c->err, E, P);
}

template <bool EOS>
static void launch_density_t(mhpc_ctx *c, bool publish) {
    if (c->n <= 0) return;
    const EosArgs E = eos_args(c);
    DensPub P = {};
    if (publish) P = DensPub{c->siters, c->f[MHPC_F_SML], c->noi, c->h_iters, c->counters};
    if (c->p.precision == MHPC_PREC_FP64) launch_density_p<true, EOS>(c, E, P);
    else launch_density_p<false, EOS>(c, E, P);
}

void launch_density(mhpc_ctx *c) { launch_density_t<false>(c, false); }

void launch_density_eos(mhpc_ctx *c, bool publish) { launch_density_t<true>(c, publish); }

__global__ __launch_bounds__(256) void k_eos(int64_t n, EosArgs E, const double *__restrict__ srho,
                                             const double4 *__restrict__ sxm, const double *__restrict__ sh,
                                             const int32_t *__restrict__ perm) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    eos_particle(E, k, perm[k], srho[k], sxm[k].w, sh[k]);
}

void launch_eos(mhpc_ctx *c) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_eos, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->n, eos_args(c),
                       c->srho, c->sxm, c->sh, c->perm);
}

// ------------------------------------------------------------------------------------
// a14 internal forces (internal_forces.cu:4-668): SPH_EQU_VERSION 1 pressure term,
// Monaghan viscosity, drho/dt, de/dt (INTEGRATE_ENERGY), dh/dt (INTEGRATE_SML).
// Per pair: grad W = dW/dr (r_ij / r); pressure factor pf = -m_j (p_i/rho_i^2 +
// p_j/rho_j^2); a_i += (pf - m_j Pi_ij) grad W; drho_i += rho_i (m_j/rho_j) v_ij.gradW;
// dh_i -= h_i/3 (m_j/rho_j) v_ij.gradW; de_i += 1/2 (m_j Pi_ij - pf) v_ij.gradW (the
// reference's 1/2 m_j Pi v.gradW + 1/2 a^p_ij . (-v_ij), with a^p_ij = pf gradW).
struct ForceArgs {
    int64_t n;
    const double4 *sxm, *sv4, *sr4;
    const double *sh;
    const int32_t *snoi, *nbr;
    int64_t stride;
    int max_nb, sml_avg, kind;
    double alpha, beta;
    const int32_t *perm;
    double *ax, *ay, *az, *drhodt, *dedt, *dsmldt, *muijmax;
    double *curlv, *gradrho;   // learned-h features (FEAT)
};

// Per-lane force accumulators.  FEAT also accumulates curl v = sum (m_j/rho_j) grad W x
// (v_j - v_i) and grad rho = sum m_j grad W, the learned-h features (plain sums in the pair
// precision).
template <bool FP64>
struct ForceAcc;
template <>
struct ForceAcc<true> {
    double a0 = 0, a1 = 0, a2 = 0, dvv = 0, de = 0;
};
template <>
struct ForceAcc<false> {
    Kahan a0, a1, a2, drho, de, dsml;
};

struct ForceI {   // particle i's data and the constants of the pair loop
    double4 p, v, r;
    double h;
    int sml_avg, kind;
    double alpha, beta;
};

// one pair (i, j) of the force loop; j's records {x,y,z,m}, {vx,vy,vz,cs},
// {rho, p/rho^2, m/rho, h}
template <bool FP64, bool FEAT, bool CUBIC>
__device__ __forceinline__ void force_pair(const ForceI &I, const double4 &pj, const double4 &vj, const double4 &rj,
                                           ForceAcc<FP64> &acc, double &mumax, double *feat) {
    const double4 pi = I.p, vi = I.v, ri = I.r;
    if constexpr (FP64) {
        const double hi = I.h;
        const double hj = rj.w;
        const double hb = 0.5 * (hi + hj);
        const double sml = I.sml_avg ? hb : hi;
        const double d0 = pi.x - pj.x, d1 = pi.y - pj.y, d2 = pi.z - pj.z;
        const double r2 = fma(d0, d0, fma(d1, d1, d2 * d2));
        const double rinv = rsq_nr(r2);
        const double ih = rcp_nr(sml);
        const double q = r2 * rinv * ih;
        const double ih2 = ih * ih;
        // grad W = (dW/dr / r) r_ij
        double gs = 0.0;
        if constexpr (CUBIC) {
            // for q <= 1/2, dW/dr / r = 6 f/h (3q - 2) q / r with q / r = 1/h
            const double fih = K8PI * (ih2 * ih2);
            if (q <= 0.5) gs = fih * ih * fma(18.0, q, -12.0);
            else if (q <= 1.0) { const double t = 1.0 - q; gs = -6.0 * fih * (t * t) * rinv; }
        } else if (q <= 1.0) {
            double w, dw;
            kernel_shape<double>(I.kind, q, w, dw);
            gs = kernel_sigma(I.kind) * (ih2 * ih2) * dw * rinv;
        }
        const double g0 = gs * d0, g1 = gs * d1, g2 = gs * d2;
        const double v0 = vi.x - vj.x, v1 = vi.y - vj.y, v2 = vi.z - vj.z;
        const double vv = fma(v0, g0, fma(v1, g1, v2 * g2));
        const double vr = fma(v0, d0, fma(v1, d1, v2 * d2));
        double mpij = 0.0;
        if (vr < 0) {
            const double csbar = 0.5 * (vi.w + vj.w);
            const double mu = hb * vr * rcp_nr(fma(hb * hb, 1e-2, r2));
            mumax = fmax(mumax, mu);
            const double rhobar = 0.5 * (ri.x + rj.x);
            mpij = pj.w * fma(I.beta, mu, -I.alpha * csbar) * mu * rcp_nr(rhobar);
        }
        const double pf = -pj.w * (ri.y + rj.y);
        const double fa = pf - mpij;
        acc.a0 = fma(fa, g0, acc.a0);
        acc.a1 = fma(fa, g1, acc.a1);
        acc.a2 = fma(fa, g2, acc.a2);
        acc.dvv = fma(rj.z, vv, acc.dvv);   // the drho/dt and dh/dt sums are the same sum
        acc.de = fma(mpij - pf, vv, acc.de);
        if (FEAT) {
            feat[0] += rj.z * (v1 * g2 - v2 * g1);
            feat[1] += rj.z * (v2 * g0 - v0 * g2);
            feat[2] += rj.z * (v0 * g1 - v1 * g0);
            feat[3] += pj.w * g0;
            feat[4] += pj.w * g1;
            feat[5] += pj.w * g2;
        }
    } else {
        const float hif = (float)I.h, rhoif = (float)ri.x, csif = (float)vi.w, prif = (float)ri.y;
        const float alpha = (float)I.alpha, beta = (float)I.beta;
        const float hj = (float)rj.w;
        const float hb = 0.5f * (hif + hj);
        const float sml = I.sml_avg ? hb : hif;
        const float d0 = (float)(pi.x - pj.x), d1 = (float)(pi.y - pj.y), d2 = (float)(pi.z - pj.z);
        const float v0 = (float)(vi.x - vj.x), v1 = (float)(vi.y - vj.y), v2 = (float)(vi.z - vj.z);
        const float r2 = __builtin_fmaf(d0, d0, __builtin_fmaf(d1, d1, d2 * d2));
        const float r = sqrtf(r2);
        float W, dWdr;
        if constexpr (CUBIC) {
            cubic32(r, sml, W, dWdr);
        } else {
            const float ih = 1.0f / sml;
            const float q = r * ih;
            float w = 0.f, dw = 0.f;
            if (q <= 1.0f) kernel_shape<float>(I.kind, q, w, dw);
            dWdr = (float)kernel_sigma(I.kind) * (ih * ih) * (ih * ih) * dw;
        }
        const float gs = dWdr / r;
        const float g0 = gs * d0, g1 = gs * d1, g2 = gs * d2;
        const float vv = __builtin_fmaf(v0, g0, __builtin_fmaf(v1, g1, v2 * g2));
        const float vr = __builtin_fmaf(v0, d0, __builtin_fmaf(v1, d1, v2 * d2));
        const float mj = (float)pj.w;
        float mpij = 0.f;
        if (vr < 0.f) {
            const float csbar = 0.5f * (csif + (float)vj.w);
            const float mu = hb * vr / __builtin_fmaf(hb * hb, 1e-2f, r2);
            mumax = fmax(mumax, (double)mu);
            const float rhobar = 0.5f * (rhoif + (float)rj.x);
            mpij = mj * (beta * mu - alpha * csbar) * mu / rhobar;
        }
        const float pf = -mj * (prif + (float)rj.y);
        const float fa = pf - mpij;
        acc.a0.add(fa * g0);
        acc.a1.add(fa * g1);
        acc.a2.add(fa * g2);
        const float mr_vv = (float)rj.z * vv;
        acc.drho.add(mr_vv);
        acc.dsml.add(mr_vv);
        acc.de.add((mpij - pf) * vv);
        if (FEAT) {
            const float vjr = (float)rj.z;
            feat[0] += vjr * (v1 * g2 - v2 * g1);
            feat[1] += vjr * (v2 * g0 - v0 * g2);
            feat[2] += vjr * (v0 * g1 - v1 * g0);
            feat[3] += mj * g0;
            feat[4] += mj * g1;
            feat[5] += mj * g2;
        }
    }
}

// One block per window of FORCE_WIN consecutive sorted particles, their records staged in
// LDS (see the density kernel): in-window neighbours are read there, the others gathered.
#define FORCE_WIN 512
static_assert(FORCE_WIN == NN_CWIN && DENS_WIN == NN_CWIN, "lists hold the neighbours of the consumers' window first");

template <bool FP64, bool FEAT, bool CUBIC>
__global__ __launch_bounds__(FORCE_WIN) void k_forces(ForceArgs A) {
    __shared__ double4 wp[FORCE_WIN], wv[FORCE_WIN], wr[FORCE_WIN];
    const int64_t base = xcd_block(blockIdx.x, gridDim.x) * (int64_t)FORCE_WIN;
    const int tid = threadIdx.x;
    const int64_t k = base + tid;
    if (k < A.n) {
        wp[tid] = A.sxm[k];
        wv[tid] = A.sv4[k];
        wr[tid] = A.sr4[k];
    }
    __syncthreads();
    if (k >= A.n) return;
    ForceI I;
    I.p = wp[tid];
    I.v = wv[tid];
    I.r = wr[tid];
    I.h = A.sh[k];
    I.sml_avg = A.sml_avg;
    I.kind = A.kind;
    I.alpha = A.alpha;
    I.beta = A.beta;
    int cnt = A.snoi[k];
    cnt = cnt > A.max_nb ? A.max_nb : cnt;
    ForceAcc<FP64> acc;
    double mumax = 0.0;
    double feat[6] = {0, 0, 0, 0, 0, 0};
    const int4 *nrow = reinterpret_cast<const int4 *>(A.nbr) + k;   // row r at nrow[r * stride]
    const int64_t stride = A.stride;
    // two neighbours per step, both requested before either pair is evaluated (see the
    // density kernel); list rows a row ahead
    int4 qn = make_int4(0, 0, 0, 0);
    if (cnt > 0) qn = nrow[0];
    for (int r = 0; 4 * r < cnt; r++) {
        const int4 q4 = qn;
        const int m = cnt - 4 * r;
        if (m > 4) qn = nrow[(int64_t)(r + 1) * stride];
        const int32_t jj[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
        for (int h = 0; h < 4; h += 2) {
            double4 pj[2], vj[2], rj[2];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int32_t j = h + e < m ? jj[h + e] : (int32_t)k;
                const uint64_t o = (uint64_t)((int64_t)j - base);
                const bool in = o < (uint64_t)FORCE_WIN;
                const int oc = in ? (int)o : 0;
                pj[e] = wp[oc];
                vj[e] = wv[oc];
                rj[e] = wr[oc];
                if (!in) {
                    pj[e] = A.sxm[j];
                    vj[e] = A.sv4[j];
                    rj[e] = A.sr4[j];
                }
            }
#pragma unroll
            for (int e = 0; e < 2; e++)
                if (h + e < m) force_pair<FP64, FEAT, CUBIC>(I, pj[e], vj[e], rj[e], acc, mumax, feat);
        }
    }
    double out[6];
    if constexpr (FP64) {
        out[0] = acc.a0; out[1] = acc.a1; out[2] = acc.a2;
        out[3] = I.r.x * acc.dvv;
        out[4] = 0.5 * acc.de;
        out[5] = -(1.0 / 3.0) * I.h * acc.dvv;
    } else {
        out[0] = acc.a0.value(); out[1] = acc.a1.value(); out[2] = acc.a2.value();
        out[3] = I.r.x * acc.drho.value();
        out[4] = 0.5 * acc.de.value();
        out[5] = -(1.0 / 3.0) * I.h * acc.dsml.value();
    }
    const int32_t i = A.perm[k];
    A.ax[i] = out[0];
    A.ay[i] = out[1];
    A.az[i] = out[2];
    A.drhodt[i] = out[3];
    A.dedt[i] = out[4];
    A.dsmldt[i] = out[5];
    A.muijmax[i] = mumax;
    if (FEAT) {
        A.curlv[i] = sqrt(feat[0] * feat[0] + feat[1] * feat[1] + feat[2] * feat[2]);
        A.gradrho[i] = sqrt(feat[3] * feat[3] + feat[4] * feat[4] + feat[5] * feat[5]);
    }
}

template <bool FP64, bool FEAT>
static void launch_forces_p(mhpc_ctx *c, const ForceArgs &a) {
    const unsigned g = (unsigned)ceil_div(c->n, FORCE_WIN);
    if (a.kind == MHPC_KERNEL_CUBIC_SPLINE)
        hipLaunchKernelGGL((k_forces<FP64, FEAT, true>), dim3(g), dim3(FORCE_WIN), 0, c->stream, a);
    else
        hipLaunchKernelGGL((k_forces<FP64, FEAT, false>), dim3(g), dim3(FORCE_WIN), 0, c->stream, a);
}

void launch_forces(mhpc_ctx *c, bool features) {
    if (c->n <= 0) return;
    ForceArgs a;
    a.n = c->n;
    a.sxm = c->sxm; a.sv4 = c->sv4; a.sr4 = c->sr4; a.sh = c->sh;
    a.snoi = c->snoi; a.nbr = c->nbr; a.stride = c->cap_pad;
    a.max_nb = c->p.max_nb; a.sml_avg = c->p.sml_avg; a.kind = c->p.smoothing_kernel;
    a.alpha = c->p.alpha; a.beta = c->p.beta;
    a.perm = c->perm;
    a.ax = c->f[MHPC_F_AX]; a.ay = c->f[MHPC_F_AY]; a.az = c->f[MHPC_F_AZ];
    a.drhodt = c->f[MHPC_F_DRHODT]; a.dedt = c->f[MHPC_F_DEDT]; a.dsmldt = c->f[MHPC_F_DSMLDT];
    a.muijmax = c->f[MHPC_F_MUIJMAX];
    a.curlv = c->f[MHPC_F_CURLV];
    a.gradrho = c->f[MHPC_F_GRADRHO];
    if (c->p.precision == MHPC_PREC_FP64) {
        if (features) launch_forces_p<true, true>(c, a);
        else launch_forces_p<true, false>(c, a);
    } else {
        if (features) launch_forces_p<false, true>(c, a);
        else launch_forces_p<false, false>(c, a);
    }
}

// ------------------------------------------------------------------------------------
// neighbour lists back to the reference layout nnl[i * max_nb + k], original indices
__global__ __launch_bounds__(256) void k_nnl_to_caller(int64_t n, const int32_t *__restrict__ perm,
                                                       const int32_t *__restrict__ snoi,
                                                       const int32_t *__restrict__ nbr, int64_t stride,
                                                       int cap_nb, int32_t *__restrict__ nnl,
                                                       int32_t *__restrict__ noi, int max_nb) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t i = perm[k];
    int cnt = snoi[k];
    noi[i] = cnt;
    cnt = cnt > cap_nb ? cap_nb : cnt;
    cnt = cnt > max_nb ? max_nb : cnt;
    for (int s = 0; s < cnt; s++) {
        const int32_t j = nbr[((int64_t)(s >> 2) * stride + k) * 4 + (s & 3)];
        nnl[(int64_t)i * max_nb + s] = perm[j];
    }
}

void launch_neighbours_to_caller(mhpc_ctx *c, int32_t *dst_nnl, int32_t *dst_noi, int max_nb) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_nnl_to_caller, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream,
                       c->n, c->perm, c->snoi, c->nbr, c->cap_pad, c->p.max_nb, dst_nnl, dst_noi, max_nb);
}

// after the h-iteration: final h, iteration counts and last counts to caller order
__global__ __launch_bounds__(256) void k_scatter_h(int64_t n, const int32_t *__restrict__ perm,
                                                   const double *__restrict__ sh,
                                                   const int32_t *__restrict__ siters,
                                                   const int32_t *__restrict__ snoi,
                                                   double *__restrict__ sml, int32_t *__restrict__ iters,
                                                   int32_t *__restrict__ noi) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t i = perm[k];
    sml[i] = sh[k];
    iters[i] = siters[k];
    noi[i] = snoi[k];
}

void launch_scatter_h(mhpc_ctx *c) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_scatter_h, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       c->perm, c->sh, c->siters, c->snoi, c->f[MHPC_F_SML], c->h_iters, c->noi);
}

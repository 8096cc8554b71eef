/*
 * oracle.c -- FP64 CPU restatement of milupHPC's single-subdomain RHS.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP path.  Compiled with
 * -ffp-contract=off so every product and sum rounds exactly as written, which is what
 * the bit-exact checks (keys, tree, COM, MAC decisions, neighbour sets) rely on.
 * Reference file:line citations are relative to the milupHPC source tree.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------
 * a1: bounding box.  tree.cu:302-437: per-axis min/max, combined with the box reset to
 * 0 (device_rhs.cu:25-32), then CUBIC_DOMAINS (tree.cu:397-431) makes all three axes the
 * global min / max.  Result: min = min(0, all coords), max = max(0, all coords).
 * box = {minX, maxX, minY, maxY, minZ, maxZ}.
 * ---------------------------------------------------------------------------------- */
void orc_bbox(int64_t n, const double *x, const double *y, const double *z, double box[6]) {
    double lo = 0.0, hi = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (x[i] < lo) lo = x[i];
        if (x[i] > hi) hi = x[i];
        if (y[i] < lo) lo = y[i];
        if (y[i] > hi) hi = y[i];
        if (z[i] < lo) lo = z[i];
        if (z[i] > hi) hi = z[i];
    }
    box[0] = box[2] = box[4] = lo;
    box[1] = box[3] = box[5] = hi;
}

/* ------------------------------------------------------------------------------------
 * a2: keys.  Tree::getParticleKey, tree.cu:157-240.  21 levels of FP64 bisection; the
 * LOWER half of each axis sets the bit (x: 1, y: 2, z: 4).  The reference loop runs a
 * 22nd pass whose shift count is negative (UB on the host, a no-op on its GPU); it
 * contributes nothing, so this restatement stops after 21 levels.
 * ---------------------------------------------------------------------------------- */
uint64_t orc_lebesgue_key(double x, double y, double z, const double box[6]) {
    double min_x = box[0], max_x = box[1];
    double min_y = box[2], max_y = box[3];
    double min_z = box[4], max_z = box[5];
    uint64_t key = 0;
    for (int level = 0; level < 21; level++) {
        uint64_t son = 0;
        double mx = 0.5 * (min_x + max_x);
        if (x < mx) { son += 1; max_x = mx; } else { min_x = mx; }
        double my = 0.5 * (min_y + max_y);
        if (y < my) { son += 2; max_y = my; } else { min_y = my; }
        double mz = 0.5 * (min_z + max_z);
        if (z < mz) { son += 4; max_z = mz; } else { min_z = mz; }
        key |= son << (3 * (20 - level));
    }
    return key;
}

/* tree.cuh:52-64 (DirTable) and tree.cuh:107-118 (HilbertTable), 3D. */
static const unsigned char kDirTable[12][8] = {
    {8, 10, 3, 3, 4, 5, 4, 5},   {2, 2, 11, 9, 4, 5, 4, 5},  {7, 6, 7, 6, 8, 10, 1, 1},
    {7, 6, 7, 6, 0, 0, 11, 9},   {0, 8, 1, 11, 6, 8, 6, 11}, {10, 0, 9, 1, 10, 7, 9, 7},
    {10, 4, 9, 4, 10, 2, 9, 3},  {5, 8, 5, 11, 2, 8, 3, 11}, {4, 9, 0, 0, 7, 9, 2, 2},
    {1, 1, 8, 5, 3, 3, 8, 6},    {11, 5, 0, 0, 11, 6, 2, 2}, {1, 1, 4, 10, 3, 3, 7, 10}};
static const unsigned char kHilbertTable[12][8] = {
    {0, 7, 3, 4, 1, 6, 2, 5}, {4, 3, 7, 0, 5, 2, 6, 1}, {6, 1, 5, 2, 7, 0, 4, 3},
    {2, 5, 1, 6, 3, 4, 0, 7}, {0, 1, 7, 6, 3, 2, 4, 5}, {6, 7, 1, 0, 5, 4, 2, 3},
    {2, 3, 5, 4, 1, 0, 6, 7}, {4, 5, 3, 2, 7, 6, 0, 1}, {0, 3, 1, 2, 7, 4, 6, 5},
    {2, 1, 3, 0, 5, 6, 4, 7}, {4, 7, 5, 6, 3, 0, 2, 1}, {6, 5, 7, 4, 1, 2, 0, 3}};

/* KeyNS::lebesgue2hilbert(key, maxLevel=21, level=21), tree.cu:40-53 */
uint64_t orc_lebesgue2hilbert(uint64_t lebesgue) {
    uint64_t h = 0;
    int dir = 0;
    for (int lvl = 21; lvl > 0; lvl--) {
        int cell = (int)((lebesgue >> (3 * (lvl - 1))) & 7u);
        h = (h << 3) + kHilbertTable[dir][cell];
        dir = kDirTable[dir][cell];
    }
    return h;
}

/* SubDomainKeyTreeNS::Kernel::getParticleKeys, subdomain.cu:849-895 */
void orc_keys(int64_t n, const double *x, const double *y, const double *z,
              const double box[6], int curve, uint64_t *keys) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint64_t k = orc_lebesgue_key(x[i], y[i], z[i], box);
        keys[i] = (curve == ORC_CURVE_HILBERT) ? orc_lebesgue2hilbert(k) : k;
    }
}

/* ------------------------------------------------------------------------------------
 * a3: stable sort of (key, index), i.e. cub::DeviceRadixSort::SortPairs with values
 * start[i] = i (tree.cu:1404-1413, helper.cu:150-167).  LSD radix, 16-bit digits: each
 * pass is a stable counting sort, so ties keep ascending index order.
 * ---------------------------------------------------------------------------------- */
void orc_sort(int64_t n, const uint64_t *keys, int32_t *perm) {
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * 65537);
    for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
    int32_t *src = perm, *dst = tmp;
    for (int pass = 0; pass < 4; pass++) {
        int sh = 16 * pass;
        memset(cnt, 0, sizeof(int64_t) * 65537);
        for (int64_t i = 0; i < n; i++) cnt[((keys[src[i]] >> sh) & 0xffffu) + 1]++;
        for (int d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < n; i++) dst[cnt[(keys[src[i]] >> sh) & 0xffffu]++] = src[i];
        int32_t *t = src; src = dst; dst = t;
    }
    /* 4 passes: result is back in perm */
    free(tmp);
    free(cnt);
}

/* ------------------------------------------------------------------------------------
 * a5: octree by insertion, TreeNS::Kernel::buildTree (tree.cu:768-1082), restated
 * serially.  Same bisection as the keys; a collision in a slot creates a cell whose
 * position is the geometric centre of its box (tree.cu:995-1001, COMPUTE_DIRECTLY 0),
 * and subdivides until the two particles separate.  level[cell] = depth (root 0),
 * level[leaf] = depth of its parent + 1.  Depth > MAX_LEVEL+1 = 22 is the reference's
 * assertion (tree.cu:949-963) and returns ORC_E_DEPTH.
 * ---------------------------------------------------------------------------------- */
static int octant_shrink(double x, double y, double z, double *mn, double *mx) {
    int s = 0;
    double c = 0.5 * (mn[0] + mx[0]);
    if (x < c) { s += 1; mx[0] = c; } else { mn[0] = c; }
    c = 0.5 * (mn[1] + mx[1]);
    if (y < c) { s += 2; mx[1] = c; } else { mn[1] = c; }
    c = 0.5 * (mn[2] + mx[2]);
    if (z < c) { s += 4; mx[2] = c; } else { mn[2] = c; }
    return s;
}

static int octant_of(double x, double y, double z, const double *mn, const double *mx) {
    int s = 0;
    if (x < 0.5 * (mn[0] + mx[0])) s += 1;
    if (y < 0.5 * (mn[1] + mx[1])) s += 2;
    if (z < 0.5 * (mn[2] + mx[2])) s += 4;
    return s;
}

int orc_tree_build(int64_t n, const double *x, const double *y, const double *z,
                   const double box[6], int64_t cap_cells,
                   int32_t *child, int32_t *level, double *nx, double *ny, double *nz,
                   int64_t *ncells) {
    int64_t total = n + cap_cells;
    for (int64_t i = 0; i < 8 * total; i++) child[i] = -1;
    for (int64_t i = 0; i < total; i++) level[i] = -1;
    for (int64_t i = 0; i < n; i++) { nx[i] = x[i]; ny[i] = y[i]; nz[i] = z[i]; }
    const int64_t root = n;
    int64_t next = n + 1;
    level[root] = 0;
    nx[root] = 0.5 * (box[0] + box[1]);
    ny[root] = 0.5 * (box[2] + box[3]);
    nz[root] = 0.5 * (box[4] + box[5]);
    *ncells = 1;
    for (int64_t p = 0; p < n; p++) {
        double mn[3] = {box[0], box[2], box[4]}, mx[3] = {box[1], box[3], box[5]};
        int64_t node = root;
        int lev = 0;
        int slot = octant_shrink(x[p], y[p], z[p], mn, mx);
        for (;;) {
            int32_t c = child[8 * node + slot];
            if (c == -1) {
                child[8 * node + slot] = (int32_t)p;
                level[p] = lev + 1;
                break;
            }
            if (c >= n) { /* descend into a cell */
                node = c;
                lev++;
                slot = octant_shrink(x[p], y[p], z[p], mn, mx);
                continue;
            }
            /* slot holds particle c: subdivide until c and p separate */
            for (;;) {
                if (next >= total) return ORC_E_CAP_NODES;
                int64_t cell = next++;
                lev++;
                if (lev > 22) return ORC_E_DEPTH;
                child[8 * node + slot] = (int32_t)cell;
                level[cell] = lev;
                nx[cell] = 0.5 * (mn[0] + mx[0]);
                ny[cell] = 0.5 * (mn[1] + mx[1]);
                nz[cell] = 0.5 * (mn[2] + mx[2]);
                int so = octant_of(x[c], y[c], z[c], mn, mx);
                int sp = octant_shrink(x[p], y[p], z[p], mn, mx);
                node = cell;
                if (so != sp) {
                    child[8 * cell + so] = c;
                    level[c] = lev + 1;
                    child[8 * cell + sp] = (int32_t)p;
                    level[p] = lev + 1;
                    break;
                }
                slot = sp;
            }
            break;
        }
    }
    *ncells = next - n;
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------
 * a6: centres of mass.  TreeNS::Kernel::calculateCentersOfMass (tree.cu:1416-1497),
 * called for level = 21..1 (miluphpc.cpp:1205-1210): mass = sum of children's masses,
 * x = sum(m_c x_c) / mass (only if mass > 0), children in slot order 0..7.  The root
 * (level 0) is never computed by the reference; this restatement does compute it (the
 * HIP walk starts at the root and always opens it).
 * nm/nx/ny/nz are node arrays of length n + ncells (particles first).
 * ---------------------------------------------------------------------------------- */
void orc_com(int64_t n, int64_t ncells, const int32_t *child, const int32_t *level,
             double *nm, double *nx, double *ny, double *nz) {
    for (int lev = 22; lev >= 0; lev--) {
#pragma omp parallel for schedule(static)
        for (int64_t c = n; c < n + ncells; c++) {
            if (level[c] != lev) continue;
            double m = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
            for (int k = 0; k < 8; k++) {
                int32_t ch = child[8 * c + k];
                if (ch == -1) continue;
                sx += nm[ch] * nx[ch];
                sy += nm[ch] * ny[ch];
                sz += nm[ch] * nz[ch];
                m += nm[ch];
            }
            if (m > 0.0) { sx /= m; sy /= m; sz /= m; }
            nm[c] = m;
            nx[c] = sx;
            ny[c] = sy;
            nz[c] = sz;
        }
    }
}

/* Traceless quadrupole about the COM, Q_ij = sum m (3 d_i d_j - d^2 delta_ij), combined
 * bottom-up with the parallel-axis shift Q_ij += Q_child_ij + m_c (3 d_i d_j - d^2
 * delta_ij), d = x_child - x_cell.  Order of components: xx, xy, xz, yy, yz, zz. */
void orc_quadrupole(int64_t n, int64_t ncells, const int32_t *child, const int32_t *level,
                    const double *nm, const double *nx, const double *ny, const double *nz,
                    double *q) {
    for (int64_t i = 0; i < 6 * n; i++) q[i] = 0.0;
    for (int lev = 22; lev >= 0; lev--) {
#pragma omp parallel for schedule(static)
        for (int64_t c = n; c < n + ncells; c++) {
            if (level[c] != lev) continue;
            double a[6] = {0, 0, 0, 0, 0, 0};
            for (int k = 0; k < 8; k++) {
                int32_t ch = child[8 * c + k];
                if (ch == -1) continue;
                double dx = nx[ch] - nx[c], dy = ny[ch] - ny[c], dz = nz[ch] - nz[c];
                double d2 = dx * dx + dy * dy + dz * dz;
                double m = nm[ch];
                const double *qc = q + 6 * (int64_t)ch;
                a[0] += qc[0] + m * (3.0 * dx * dx - d2);
                a[1] += qc[1] + m * (3.0 * dx * dy);
                a[2] += qc[2] + m * (3.0 * dx * dz);
                a[3] += qc[3] + m * (3.0 * dy * dy - d2);
                a[4] += qc[4] + m * (3.0 * dy * dz);
                a[5] += qc[5] + m * (3.0 * dz * dz - d2);
            }
            for (int k = 0; k < 6; k++) q[6 * c + k] = a[k];
        }
    }
}

/* ------------------------------------------------------------------------------------
 * a7': Gravity::Kernel::computeForces_v1 (gravity.cu:252-396): per-particle DFS from the
 * root; cellSize[d] = 4 radius^2 / 4^d; accept a child if it is a leaf or
 * (d^2 + smoothing) theta^2 > cellSize[depth]; skip the particle itself.
 * u[i] -= 0.5 m_c m_i / d (gravity.cu:360-366).  smoothing is eps^2.
 * ---------------------------------------------------------------------------------- */
int orc_bh_v1(int64_t n, int64_t ncells, const int32_t *child, const double *nm,
              const double *nx, const double *ny, const double *nz, double radius,
              double theta, double smoothing, double *ax, double *ay, double *az, double *u) {
    (void)ncells;
    double cellSize[128];
    cellSize[0] = 4.0 * radius * radius;
    for (int i = 1; i < 128; i++) cellSize[i] = cellSize[i - 1] * 0.25;
    const double thetasq = theta * theta;
    int err = ORC_OK;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        int curNode[128], curChild[128];
        double px = nx[i], py = ny[i], pz = nz[i];
        double axi = 0, ayi = 0, azi = 0, ui = 0;
        int depth = 1;
        curNode[depth] = (int)n; /* root */
        curChild[depth] = 0;
        do {
            int childNumber = curChild[depth];
            int node = curNode[depth];
            while (childNumber < 8) {
                int c;
                do {
                    c = child[8 * (int64_t)node + childNumber];
                    childNumber++;
                } while (c == -1 && childNumber < 8);
                if (c != -1 && c != i) {
                    double dx = nx[c] - px;
                    double dist = dx * dx + smoothing;
                    double dy = ny[c] - py;
                    dist += dy * dy;
                    double dz = nz[c] - pz;
                    dist += dz * dz;
                    if (c < n || dist * thetasq > cellSize[depth]) {
                        dist = sqrt(dist);
                        double f = nm[c] / (dist * dist * dist);
                        axi += f * dx;
                        ayi += f * dy;
                        azi += f * dz;
                        if (u) ui -= 0.5 * (nm[c] * nm[i]) / dist;
                    } else {
                        curChild[depth] = childNumber;
                        curNode[depth] = node;
                        depth++;
                        if (depth >= 127) {
#pragma omp atomic write
                            err = ORC_E_STACK;
                            depth = 0;
                            break;
                        }
                        childNumber = 0;
                        node = c;
                    }
                }
            }
            depth--;
        } while (depth > 0);
        ax[i] = axi;
        ay[i] = ayi;
        az[i] = azi;
        if (u) u[i] = ui;
    }
    return err;
}

/* Group MAC walk.  s^2(L) = (4 R^2) 4^-L with R = half the cube side (the v1 cellSize
 * table), d = distance from the cell's COM to the group's axis-aligned bounding box,
 * computed per axis as (lo - c) if c < lo, (c - hi) if c > hi, else 0.  mac 0 accepts
 * iff (d^2 + eps^2) theta^2 > s^2; mac 1 accepts iff s^2/theta <= d^2 + eps^2, where
 * s^2/theta = (R^2/theta) 4^-(L-1) as the v2 kernel's depth stack holds it.  Leaves are
 * always accepted.  Accepting for the group implies that every member's own v1/v2 test
 * would accept (d_box <= d_i), so the group walk is never less accurate per particle.
 * Monopole f = m r^-3 with r^2 = d^2 + eps^2; quadrupole adds -Q.d r^-5 +
 * 5/2 (d.Q.d) d r^-7.  phi_i = -sum m/r over all interactions except the particle itself. */
/* the walk over the groups gsel[0..ngsel) (all groups when gsel is NULL) */
static int bh_group_walk(int64_t n, const int32_t *perm, int group, const int32_t *child,
                         const int32_t *level, const double *nm, const double *nx, const double *ny,
                         const double *nz, const double *q, const double box[6], double theta,
                         double smoothing, int mac, double *ax, double *ay, double *az, double *phi,
                         int64_t *n_cell_int, int64_t *n_part_int, const int64_t *gsel, int64_t ngsel) {
    if (group <= 0 || mac < 0 || mac > 3) return ORC_E_ARG;
    const int form = mac & 1, vote = (mac & 2) != 0;
    const double R = 0.5 * (box[1] - box[0]);
    const double s2_0 = 4.0 * R * R;
    const double dp_1 = R * R / theta;
    const double theta2 = theta * theta;
    int64_t ngroups = gsel ? ngsel : (n + group - 1) / group;
    int64_t cell_int = 0, part_int = 0;
    int err = ORC_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cell_int, part_int)
    for (int64_t gi = 0; gi < ngroups; gi++) {
        const int64_t g = gsel ? gsel[gi] : gi;
        int64_t k0 = g * group, k1 = k0 + group;
        if (k1 > n) k1 = n;
        int cnt = (int)(k1 - k0);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        double *acc = (double *)calloc((size_t)cnt * 4, sizeof(double));
        int *stack = (int *)malloc(sizeof(int) * 8 * 64 * 32);
        for (int k = 0; k < cnt; k++) {
            int32_t i = perm[k0 + k];
            if (nx[i] < lo[0]) lo[0] = nx[i];
            if (nx[i] > hi[0]) hi[0] = nx[i];
            if (ny[i] < lo[1]) lo[1] = ny[i];
            if (ny[i] > hi[1]) hi[1] = ny[i];
            if (nz[i] < lo[2]) lo[2] = nz[i];
            if (nz[i] > hi[2]) hi[2] = nz[i];
        }
        int top = 0;
        stack[top++] = (int)n; /* root */
        while (top > 0) {
            int c = stack[--top];
            int accept;
            if (c < n) {
                accept = 1;
                part_int += cnt;
            } else {
                int L = level[c];
                const double s2 = s2_0 * ldexp(1.0, -2 * L);
                const double dp = dp_1 * ldexp(1.0, -2 * (L - 1));
                if (!vote) {
                    double cc[3] = {nx[c], ny[c], nz[c]};
                    double dd[3];
                    for (int a = 0; a < 3; a++) {
                        if (cc[a] < lo[a]) dd[a] = lo[a] - cc[a];
                        else if (cc[a] > hi[a]) dd[a] = cc[a] - hi[a];
                        else dd[a] = 0.0;
                    }
                    double d2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
                    accept = form == 0 ? (d2 + smoothing) * theta2 > s2 : dp <= d2 + smoothing;
                } else {
                    /* every member tests the cell on its own distance; all must accept */
                    accept = 1;
                    for (int k = 0; k < cnt && accept; k++) {
                        int32_t i = perm[k0 + k];
                        double dx = nx[c] - nx[i], dy = ny[c] - ny[i], dz = nz[c] - nz[i];
                        double d2 = dx * dx + dy * dy + dz * dz;
                        accept = form == 0 ? (d2 + smoothing) * theta2 > s2 : dp <= d2 + smoothing;
                    }
                }
                if (accept) cell_int += cnt;
            }
            if (accept) {
                double m = nm[c];
                const double *qc = (q && c >= n) ? q + 6 * (int64_t)c : NULL;
                for (int k = 0; k < cnt; k++) {
                    int32_t i = perm[k0 + k];
                    double dx = nx[c] - nx[i], dy = ny[c] - ny[i], dz = nz[c] - nz[i];
                    double r2 = dx * dx + dy * dy + dz * dz + smoothing;
                    if (!(r2 > 0.0)) continue;
                    double rinv = 1.0 / sqrt(r2);
                    double rinv2 = rinv * rinv;
                    double mr3 = m * rinv * rinv2;
                    double fx = mr3 * dx, fy = mr3 * dy, fz = mr3 * dz;
                    double ph = (c == i) ? 0.0 : -m * rinv;
                    if (qc) {
                        double qdx = qc[0] * dx + qc[1] * dy + qc[2] * dz;
                        double qdy = qc[1] * dx + qc[3] * dy + qc[4] * dz;
                        double qdz = qc[2] * dx + qc[4] * dy + qc[5] * dz;
                        double dqd = dx * qdx + dy * qdy + dz * qdz;
                        double rinv5 = rinv2 * rinv2 * rinv;
                        double rinv7 = rinv5 * rinv2;
                        double t = 2.5 * dqd * rinv7;
                        fx += -qdx * rinv5 + t * dx;
                        fy += -qdy * rinv5 + t * dy;
                        fz += -qdz * rinv5 + t * dz;
                        ph += -0.5 * dqd * rinv5;
                    }
                    acc[4 * k + 0] += fx;
                    acc[4 * k + 1] += fy;
                    acc[4 * k + 2] += fz;
                    acc[4 * k + 3] += ph;
                }
            } else {
                /* push children in reverse slot order so slot 0 is visited first */
                for (int s = 7; s >= 0; s--) {
                    int32_t ch = child[8 * (int64_t)c + s];
                    if (ch == -1) continue;
                    if (top >= 8 * 64 * 32) {
#pragma omp atomic write
                        err = ORC_E_STACK;
                        top = 0;
                        break;
                    }
                    stack[top++] = ch;
                }
            }
        }
        for (int k = 0; k < cnt; k++) {
            int32_t i = perm[k0 + k];
            ax[i] = acc[4 * k + 0];
            ay[i] = acc[4 * k + 1];
            az[i] = acc[4 * k + 2];
            if (phi) phi[i] = acc[4 * k + 3];
        }
        free(acc);
        free(stack);
    }
    if (n_cell_int) *n_cell_int = cell_int;
    if (n_part_int) *n_part_int = part_int;
    return err;
}

int orc_bh_group(int64_t n, const int32_t *perm, int group, int64_t ncells,
                 const int32_t *child, const int32_t *level, const double *nm,
                 const double *nx, const double *ny, const double *nz, const double *q,
                 const double box[6], double theta, double smoothing, int mac,
                 double *ax, double *ay, double *az, double *phi,
                 int64_t *n_cell_int, int64_t *n_part_int) {
    (void)ncells;
    return bh_group_walk(n, perm, group, child, level, nm, nx, ny, nz, q, box, theta, smoothing, mac,
                         ax, ay, az, phi, n_cell_int, n_part_int, NULL, 0);
}

int orc_bh_group_subset(int64_t n, const int32_t *perm, int group, const int32_t *child,
                        const int32_t *level, const double *nm, const double *nx, const double *ny,
                        const double *nz, const double *q, const double box[6], double theta,
                        double smoothing, int mac, int64_t ngsel, const int64_t *gsel, double *ax,
                        double *ay, double *az, double *phi, int64_t *n_cell_int, int64_t *n_part_int) {
    const int64_t ngroups = (n + group - 1) / (group > 0 ? group : 1);
    for (int64_t t = 0; t < ngsel; t++)
        if (gsel[t] < 0 || gsel[t] >= ngroups) return ORC_E_ARG;
    return bh_group_walk(n, perm, group, child, level, nm, nx, ny, nz, q, box, theta, smoothing, mac,
                         ax, ay, az, phi, n_cell_int, n_part_int, gsel, ngsel);
}

/* FP64 direct summation, the BH accuracy reference (softened with eps^2 = smoothing). */
void orc_direct(int64_t n, const double *x, const double *y, const double *z,
                const double *m, double smoothing, int64_t ntarget, const int32_t *targets,
                double *ax, double *ay, double *az, double *phi) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t t = 0; t < ntarget; t++) {
        int64_t i = targets ? targets[t] : t;
        double sx = 0, sy = 0, sz = 0, sp = 0;
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r2 = dx * dx + dy * dy + dz * dz + smoothing;
            double rinv = 1.0 / sqrt(r2);
            double f = m[j] * rinv * rinv * rinv;
            sx += f * dx;
            sy += f * dy;
            sz += f * dz;
            sp -= m[j] * rinv;
        }
        ax[t] = sx;
        ay[t] = sy;
        az[t] = sz;
        if (phi) phi[t] = sp;
    }
}

/* ------------------------------------------------------------------------------------
 * a8: neighbour criterion of SPH::Kernel::fixedRadiusNN (sph.cu:93-281): j != i is a
 * neighbour of i iff d^2 < h_i^2 and d^2 < h_j^2, d^2 = fma(dx, dx, fma(dy, dy, dz*dz))
 * in FP64 with dx = x_i - x_j (nn_d2: the reference's CUDA build contracts products into fused
 * multiply-adds by default, so its d^2 is some fused form; this one is fixed here and on the
 * GPU).  Lists here are sorted by j.
 * ---------------------------------------------------------------------------------- */
static inline double nn_d2(double dx, double dy, double dz) { return fma(dx, dx, fma(dy, dy, dz * dz)); }

static inline int nn_pair(const double *x, const double *y, const double *z,
                          const double *h, int64_t i, int64_t j) {
    double dx = x[i] - x[j];
    double dy = y[i] - y[j];
    double dz = z[i] - z[j];
    double d = nn_d2(dx, dy, dz);
    return d < h[i] * h[i] && d < h[j] * h[j];
}

int orc_nn_brute(int64_t n, const double *x, const double *y, const double *z,
                 const double *h, int max_nb, int32_t *nnl, int32_t *noi) {
    int err = ORC_OK;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) {
        int k = 0;
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            if (nn_pair(x, y, z, h, i, j)) {
                if (nnl && k < max_nb) nnl[i * max_nb + k] = (int32_t)j;
                k++;
            }
        }
        if (k > max_nb) {
#pragma omp atomic write
            err = ORC_E_CAP_NEIGHBOURS;
        }
        noi[i] = k;
    }
    return err;
}

/* Uniform-grid search with cell side >= max h; the same criterion as nn_pair, so the
 * lists equal orc_nn_brute's exactly (tests check this). */
typedef struct {
    double lo[3];
    double cell;
    int64_t dim[3];
    int64_t *start; /* ncell + 1 */
    int32_t *idx;   /* particles bucketed by cell, ascending index inside a cell */
} grid_t;

static int64_t grid_coord(double v, double lo, double cell, int64_t dim) {
    int64_t c = (int64_t)floor((v - lo) / cell);
    if (c < 0) c = 0;
    if (c >= dim) c = dim - 1;
    return c;
}

static void grid_build(grid_t *g, int64_t n, const double *x, const double *y,
                       const double *z, double cell) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; i++) {
        if (x[i] < lo[0]) lo[0] = x[i];
        if (x[i] > hi[0]) hi[0] = x[i];
        if (y[i] < lo[1]) lo[1] = y[i];
        if (y[i] > hi[1]) hi[1] = y[i];
        if (z[i] < lo[2]) lo[2] = z[i];
        if (z[i] > hi[2]) hi[2] = z[i];
    }
    int64_t total = 1;
    for (int a = 0; a < 3; a++) {
        g->lo[a] = lo[a];
        int64_t d = (int64_t)((hi[a] - lo[a]) / cell) + 1;
        if (d < 1) d = 1;
        /* keep the grid at a sane size; cells only get larger (still exact) */
        while (d > 256) { cell *= 2.0; d = (int64_t)((hi[a] - lo[a]) / cell) + 1; }
        g->dim[a] = d;
    }
    for (int a = 0; a < 3; a++) g->dim[a] = (int64_t)((hi[a] - lo[a]) / cell) + 1;
    g->cell = cell;
    total = g->dim[0] * g->dim[1] * g->dim[2];
    g->start = (int64_t *)calloc((size_t)total + 1, sizeof(int64_t));
    g->idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t *cid = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) {
        int64_t cx = grid_coord(x[i], g->lo[0], cell, g->dim[0]);
        int64_t cy = grid_coord(y[i], g->lo[1], cell, g->dim[1]);
        int64_t cz = grid_coord(z[i], g->lo[2], cell, g->dim[2]);
        cid[i] = (cz * g->dim[1] + cy) * g->dim[0] + cx;
        g->start[cid[i] + 1]++;
    }
    for (int64_t c = 0; c < total; c++) g->start[c + 1] += g->start[c];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)total);
    memcpy(fill, g->start, sizeof(int64_t) * (size_t)total);
    for (int64_t i = 0; i < n; i++) g->idx[fill[cid[i]]++] = (int32_t)i;
    free(fill);
    free(cid);
}

static void grid_free(grid_t *g) {
    free(g->start);
    free(g->idx);
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* count (and optionally list, sorted by j) neighbours of i via the grid; active may be
 * NULL; returns the count */
static int grid_query(const grid_t *g, const double *x, const double *y, const double *z,
                      const double *h, int64_t i, int max_nb, int32_t *row) {
    int64_t c0[3], c1[3];
    double p[3] = {x[i], y[i], z[i]};
    for (int a = 0; a < 3; a++) {
        c0[a] = grid_coord(p[a] - h[i], g->lo[a], g->cell, g->dim[a]);
        c1[a] = grid_coord(p[a] + h[i], g->lo[a], g->cell, g->dim[a]);
    }
    int k = 0;
    for (int64_t cz = c0[2]; cz <= c1[2]; cz++)
        for (int64_t cy = c0[1]; cy <= c1[1]; cy++)
            for (int64_t cx = c0[0]; cx <= c1[0]; cx++) {
                int64_t c = (cz * g->dim[1] + cy) * g->dim[0] + cx;
                for (int64_t s = g->start[c]; s < g->start[c + 1]; s++) {
                    int32_t j = g->idx[s];
                    if (j == i) continue;
                    if (nn_pair(x, y, z, h, i, j)) {
                        if (row && k < max_nb) row[k] = j;
                        k++;
                    }
                }
            }
    if (row) qsort(row, (size_t)(k < max_nb ? k : max_nb), sizeof(int32_t), cmp_i32);
    return k;
}

static double max_of(int64_t n, const double *v) {
    double m = 0.0;
    for (int64_t i = 0; i < n; i++)
        if (v[i] > m) m = v[i];
    return m;
}

int orc_nn_grid(int64_t n, const double *x, const double *y, const double *z,
                const double *h, int max_nb, int32_t *nnl, int32_t *noi) {
    if (n <= 0) return ORC_OK;
    grid_t g;
    double hmax = max_of(n, h);
    grid_build(&g, n, x, y, z, hmax > 0 ? hmax : 1.0);
    int err = ORC_OK;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        int k = grid_query(&g, x, y, z, h, i, max_nb, nnl ? nnl + i * max_nb : NULL);
        if (k > max_nb) {
#pragma omp atomic write
            err = ORC_E_CAP_NEIGHBOURS;
        }
        noi[i] = k;
    }
    grid_free(&g);
    return err;
}

/* The reference tree walk itself (sph.cu:93-281): r starts at diam/2 for the root's
 * children and halves per level; descend iff |dx|,|dy|,|dz| < r + h_i measured to the
 * node's stored position (cell centre, or COM once computed).  Lists in DFS order. */
int orc_nn_refwalk(int64_t n, const double *x, const double *y, const double *z,
                   const double *h, const int32_t *child, const double *nx,
                   const double *ny, const double *nz, double diam, int max_nb,
                   int32_t *nnl, int32_t *noi) {
    int err = ORC_OK;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        int curNode[130], curChild[130];
        int depth = 0;
        curNode[0] = (int)n;
        curChild[0] = 0;
        int k = 0;
        double r = diam * 0.5;
        double sml = h[i];
        double idist = r + sml;
        do {
            int childNumber = curChild[depth];
            int node = curNode[depth];
            while (childNumber < 8) {
                int c = child[8 * (int64_t)node + childNumber];
                childNumber++;
                if (c != -1 && c != i) {
                    double dx = x[i] - nx[c];
                    double dy = y[i] - ny[c];
                    double dz = z[i] - nz[c];
                    if (c < n) {
                        double d = nn_d2(dx, dy, dz);
                        if (d < sml * sml && d < h[c] * h[c]) {
                            if (k < max_nb) nnl[i * max_nb + k] = c;
                            k++;
                        }
                    } else if (fabs(dx) < idist && fabs(dy) < idist && fabs(dz) < idist) {
                        curChild[depth] = childNumber;
                        curNode[depth] = node;
                        depth++;
                        r *= 0.5;
                        idist = r + sml;
                        if (depth > 128) {
#pragma omp atomic write
                            err = ORC_E_STACK;
                            depth = -1;
                            break;
                        }
                        childNumber = 0;
                        node = c;
                    }
                }
            }
            if (depth < 0) break;
            depth--;
            r *= 2.0;
            idist = r + sml;
        } while (depth >= 0);
        if (k > max_nb) {
#pragma omp atomic write
            err = ORC_E_CAP_NEIGHBOURS;
        }
        noi[i] = k;
    }
    return err;
}

/* ------------------------------------------------------------------------------------
 * a9: plain h-iteration, Jacobi form of SPH::Kernel::fixedRadiusNN_variableSML
 * (sph.cu:739-961).  Every pass counts, for each unconverged particle, N = #{j != i :
 * d^2 < h_i^2 and d^2 < h_j^2} with every h_j taken from the start of the pass (the
 * reference reads neighbours' h while other threads write them, sph.cu:846/916, which
 * makes it order-dependent).  A particle stops when (N is in [band_lo, band_hi] or this
 * was its max_counts-th count) and N < max_nb; otherwise h *= table[min(max(N,1), tmax)]
 * where table[N] = 0.5 (1 + (N_t/N)^(1/3)) (sph.cu:918-930).  iters = counts done.
 * Returns ORC_E_CAP_NEIGHBOURS if a particle ends with N >= max_nb or N == 0
 * (the reference's cudaTerminate, sph.cu:944-953).
 * ---------------------------------------------------------------------------------- */
void orc_h_factor_table(int n_target, int tmax, double *table) {
    table[0] = 0.0;
    for (int k = 1; k <= tmax; k++)
        table[k] = 0.5 * (1.0 + pow((double)n_target / (double)k, 1.0 / 3.0));
}

int orc_h_iterate(int64_t n, const double *x, const double *y, const double *z, double *h,
                  int n_target, int band_lo, int band_hi, int max_counts, int max_nb,
                  const double *table, int tmax, double sml_max, int32_t *iters, int32_t *noi) {
    (void)n_target;
    if (n <= 0) return ORC_OK;
    double *hn = (double *)malloc(sizeof(double) * (size_t)n);
    unsigned char *done = (unsigned char *)calloc((size_t)n, 1);
    for (int64_t i = 0; i < n; i++) iters[i] = 0;
    int err = ORC_OK;
    for (int pass = 0; pass < max_counts; pass++) {
        memcpy(hn, h, sizeof(double) * (size_t)n);
        grid_t g;
        double hmax = max_of(n, h);
        grid_build(&g, n, x, y, z, hmax > 0 ? hmax : 1.0);
        int64_t active = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : active)
        for (int64_t i = 0; i < n; i++) {
            if (done[i]) continue;
            int N = grid_query(&g, x, y, z, h, i, 0, NULL);
            noi[i] = N;
            iters[i] = pass + 1;
            int inband = (N >= band_lo && N <= band_hi);
            if ((inband || pass == max_counts - 1) && N < max_nb) {
                done[i] = 1;
            } else {
                int Nc = N < 1 ? 1 : (N > tmax ? tmax : N);
                hn[i] = h[i] * table[Nc];
                if (sml_max > 0.0 && hn[i] > sml_max) hn[i] = sml_max;
                if (pass == max_counts - 1) done[i] = 1;
                active++;
            }
        }
        grid_free(&g);
        memcpy(h, hn, sizeof(double) * (size_t)n);
        if (active == 0) break;
    }
    for (int64_t i = 0; i < n; i++)
        if (noi[i] >= max_nb || noi[i] == 0) err = ORC_E_CAP_NEIGHBOURS;
    free(hn);
    free(done);
    return err;
}

/* ------------------------------------------------------------------------------------
 * a10: SPH::SmoothingKernel::cubicSpline (kernel.cu:50-83), 3D, support radius h.
 * ---------------------------------------------------------------------------------- */
void orc_cubic_spline(double r2, double dx, double dy, double dz, double h, double *W,
                      double dWdx[3], double *dWdr) {
    double r = sqrt(r2);
    double q = r / h;
    double f = 8. / M_PI * 1. / (h * h * h);
    *W = 0.0;
    *dWdr = 0.0;
    if (q > 1) {
        *W = 0;
        *dWdr = 0.0;
    } else if (q > 0.5) {
        *W = 2. * f * (1. - q) * (1. - q) * (1 - q);
        *dWdr = -6. * f * 1. / h * (1. - q) * (1. - q);
    } else {
        *W = f * (6. * q * q * q - 6. * q * q + 1.);
        *dWdr = 6. * f / h * (3 * q * q - 2 * q);
    }
    dWdx[0] = *dWdr / r * dx;
    dWdx[1] = *dWdr / r * dy;
    dWdx[2] = *dWdr / r * dz;
}

/* a10, the other four kernels of SPH::SmoothingKernel (kernel.cu:14-48 spiky, :86-113
 * wendlandc2, :116-145 wendlandc4, :149-183 wendlandc6; Dehnen & Aly 2012), their 3D
 * branches, selected as the reference's Smoothing::Kernel (parameter.h:177-180,
 * kernel_handler.cu:13-36).  Each is W = 0, dW/dr = 0 beyond r = h and grad W = dW/dr / r dx.
 * The Wendland C4 gradient restates the reference's expression (kernel.cu:140) as written:
 * it is not the derivative of that kernel's W (which is -(56/3) sigma/h^4 q (1-q)^5 (1+5q)),
 * so a finite-difference check cannot pin it -- parity unpinned for that gradient.
 * Returns ORC_E_ARG for an unknown kind. */
int orc_sph_kernel(int kind, double r2, double dx, double dy, double dz, double h, double *W,
                   double dWdx[3], double *dWdr) {
    if (kind == ORC_KERNEL_CUBIC_SPLINE) {
        orc_cubic_spline(r2, dx, dy, dz, h, W, dWdx, dWdr);
        return ORC_OK;
    }
    const double r = sqrt(r2);
    const double q = r / h;
    const double h3 = h * h * h, h4 = h3 * h;
    const double t = 1.0 - q;
    *W = 0.0;
    *dWdr = 0.0;
    if (q <= 1.0) {
        switch (kind) {
            case ORC_KERNEL_SPIKY:
                *W = 15. / (M_PI * h3) * t * t * t;
                *dWdr = -45. / (M_PI * h4) * t * t;
                break;
            case ORC_KERNEL_WENDLAND_C2:
                *W = 21. / (2 * M_PI * h3) * t * t * t * t * (1 + 4 * q);
                *dWdr = -210. / (M_PI * h4) * q * t * t * t;
                break;
            case ORC_KERNEL_WENDLAND_C4:
                *W = 495. / (32. * M_PI * h3) * t * t * t * t * t * t * (1. + 6. * q + 35. / 3. * q * q);
                *dWdr = -1485. / (16. * M_PI * h4) * t * t * t * t * t * (1. - 35. * q * q + 105. * q * q * q);
                break;
            case ORC_KERNEL_WENDLAND_C6:
                *W = 1365. / (64. * M_PI * h3) * t * t * t * t * t * t * t * t *
                     (1. + 8. * q + 25. * q * q + 32 * q * q * q);
                *dWdr = -15015. / (32. * M_PI * h4) * q * t * t * t * t * t * t * t * (1. + 7 * q + 16 * q * q);
                break;
            default: return ORC_E_ARG;
        }
    } else if (kind < ORC_KERNEL_SPIKY || kind > ORC_KERNEL_WENDLAND_C6) {
        return ORC_E_ARG;
    }
    dWdx[0] = *dWdr / r * dx;
    dWdx[1] = *dWdr / r * dy;
    dWdx[2] = *dWdr / r * dz;
    return ORC_OK;
}

/* a11: SPH::Kernel::calculateDensity (density.cu:8-72): self term m_i W(0, h_i) plus
 * m_j W(r_ij, h) over the list, h = (h_i + h_j)/2 when sml_avg (VARIABLE_SML ||
 * INTEGRATE_SML) else h_i.  rho <= 0 is the reference's termination (density.cu:68). */
int orc_density(int64_t n, const double *x, const double *y, const double *z,
                const double *m, const double *h, const int32_t *nnl, const int32_t *noi,
                int max_nb, int sml_avg, int kernel, double *rho) {
    if (kernel < ORC_KERNEL_SPIKY || kernel > ORC_KERNEL_WENDLAND_C6) return ORC_E_ARG;
    int err = ORC_OK;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double W, dWdx[3], dWdr;
        orc_sph_kernel(kernel, 0.0, 0.0, 0.0, 0.0, h[i], &W, dWdx, &dWdr);
        double r = m[i] * W;
        for (int k = 0; k < noi[i]; k++) {
            int32_t j = nnl[i * max_nb + k];
            double sml = sml_avg ? 0.5 * (h[i] + h[j]) : h[i];
            double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            double r2 = dx * dx + dy * dy + dz * dz;
            orc_sph_kernel(kernel, r2, dx, dy, dz, sml, &W, dWdx, &dWdr);
            r += m[j] * W;
        }
        rho[i] = r;
        if (!(r > 0.0)) {
#pragma omp atomic write
            err = ORC_E_RHO_NONPOSITIVE;
        }
    }
    return err;
}

/* a12: SPH::Kernel::calculateSoundSpeed (soundspeed.cu:38-97), run before the pressure
 * so the ideal gas uses the previous p.  Isothermal (3) and constant-cs (100) leave cs. */
void orc_soundspeed(int64_t n, int eos, double gamma, double K, const double *rho,
                    const double *p, double *cs) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        switch (eos) {
            case ORC_EOS_POLYTROPIC: cs[i] = sqrt(K * pow(rho[i], gamma - 1.0)); break;
            case ORC_EOS_IDEAL_GAS: cs[i] = sqrt(gamma * p[i] / rho[i]); break;
            case ORC_EOS_LOCALLY_ISOTHERMAL: cs[i] = 0; break;
            default: break;
        }
    }
}

/* a13: SPH::Kernel::calculatePressure and the EOS functions (pressure.cu:6-87). */
void orc_pressure(int64_t n, int eos, double gamma, double K, const double *rho,
                  const double *e, const double *cs, double *p) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        switch (eos) {
            case ORC_EOS_POLYTROPIC: p[i] = K * pow(rho[i], gamma); break;
            case ORC_EOS_ISOTHERMAL: p[i] = 41255.407 * rho[i]; break;
            case ORC_EOS_IDEAL_GAS: p[i] = (gamma - 1) * rho[i] * e[i]; break;
            case ORC_EOS_LOCALLY_ISOTHERMAL:
            case ORC_EOS_ISOTHERMAL_CONST_CS: p[i] = cs[i] * cs[i] * rho[i]; break;
            default: p[i] = 0.0; break;
        }
    }
}

/* a14: SPH::Kernel::internalForces (internal_forces.cu:4-668), hydro with
 * SPH_EQU_VERSION 1, Monaghan viscosity (eps 1e-2), INTEGRATE_ENERGY and INTEGRATE_SML
 * outputs.  Particles with noi < 1 get zero derivatives. */
void orc_forces(int64_t n, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *m,
                const double *h, const double *rho, const double *p, const double *cs,
                const int32_t *nnl, const int32_t *noi, int max_nb, int sml_avg, int kernel,
                double alpha, double beta, double *ax, double *ay, double *az,
                double *drhodt, double *dedt, double *dsmldt, double *muijmax) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double acc[3] = {0, 0, 0};
        double drho = 0.0, de = 0.0, dsml = 0.0, mumax = 0.0;
        for (int k = 0; k < noi[i]; k++) {
            int32_t j = nnl[i * max_nb + k];
            double sml = sml_avg ? 0.5 * (h[i] + h[j]) : h[i];
            double dr[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
            double r2 = 0.0;
            for (int e = 0; e < 3; e++) r2 += dr[e] * dr[e];
            double W, dWdx[3], dWdr;
            orc_sph_kernel(kernel, r2, dr[0], dr[1], dr[2], sml, &W, dWdx, &dWdr);
            double dv[3] = {vx[i] - vx[j], vy[i] - vy[j], vz[i] - vz[j]};
            double vvnablaW = dv[0] * dWdx[0];
            vvnablaW += dv[1] * dWdx[1];
            vvnablaW += dv[2] * dWdx[2];
            double rr = 0.0, vr = 0.0;
            for (int e = 0; e < 3; e++) {
                rr += dr[e] * dr[e];
                vr += dv[e] * dr[e];
            }
            double pij = 0.0;
            if (vr < 0) {
                double csbar = 0.5 * (cs[i] + cs[j]);
                double smooth = 0.5 * (h[i] + h[j]);
                const double eps_artvisc = 1e-2;
                double mu = smooth * vr / (rr + smooth * smooth * eps_artvisc);
                if (mu > mumax) mumax = mu;
                double rhobar = 0.5 * (rho[i] + rho[j]);
                pij = (beta * mu - alpha * csbar) * mu / rhobar;
            }
            double accelsj[3];
            for (int d = 0; d < 3; d++) {
                accelsj[d] = -m[j] * (p[i] / (rho[i] * rho[i]) + p[j] / (rho[j] * rho[j])) * dWdx[d];
                acc[d] += accelsj[d];
            }
            for (int d = 0; d < 3; d++) acc[d] += m[j] * (-pij) * dWdx[d];
            drho += rho[i] / rho[j] * m[j] * vvnablaW;
            dsml -= 1. / 3 * h[i] * m[j] / rho[j] * vvnablaW;
            de += 0.5 * m[j] * pij * vvnablaW;
            de += 0.5 * accelsj[0] * -dv[0];
            de += 0.5 * accelsj[1] * -dv[1];
            de += 0.5 * accelsj[2] * -dv[2];
        }
        ax[i] = acc[0];
        ay[i] = acc[1];
        az[i] = acc[2];
        drhodt[i] = drho;
        dedt[i] = de;
        dsmldt[i] = dsml;
        if (muijmax) muijmax[i] = mumax;
    }
}

/* ------------------------------------------------------------------------------------
 * a17: LeapfrogNS::Kernel::updateX / updateV (device_leapfrog.cu:4-58).
 * ---------------------------------------------------------------------------------- */
void orc_leapfrog_update_x(int64_t n, double dt, double *x, double *y, double *z,
                           const double *vx, const double *vy, const double *vz,
                           const double *gx, const double *gy, const double *gz,
                           double *gx_old, double *gy_old, double *gz_old) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        x[i] += dt * (vx[i] + 0.5 * dt * gx[i]);
        gx_old[i] = gx[i];
        y[i] += dt * (vy[i] + 0.5 * dt * gy[i]);
        gy_old[i] = gy[i];
        z[i] += dt * (vz[i] + 0.5 * dt * gz[i]);
        gz_old[i] = gz[i];
    }
}

void orc_leapfrog_update_v(int64_t n, double dt, double *vx, double *vy, double *vz,
                           const double *gx, const double *gy, const double *gz,
                           const double *gx_old, const double *gy_old, const double *gz_old) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        vx[i] += 0.5 * dt * (gx[i] + gx_old[i]);
        vy[i] += 0.5 * dt * (gy[i] + gy_old[i]);
        vz[i] += 0.5 * dt * (gz[i] + gz_old[i]);
    }
}

/* ------------------------------------------------------------------------------------
 * Predictor-corrector Euler, PredictorCorrectorEulerNS::Kernel::predictor / corrector
 * (device_predictor_corrector_euler.cu:130-291), hydro fields, INTEGRATE_ENERGY,
 * h from the predicted state (INTEGRATE_SML 0).  s[] = x,y,z,vx,vy,vz,ax,ay,az,gx,gy,gz,
 * e,dedt; b[] = saved x,y,z,vx,vy,vz,ax,ay,az,e,dedt.
 * ---------------------------------------------------------------------------------- */
void orc_pc_predictor(int64_t n, double dt, double **s, double **b) {
    for (int64_t i = 0; i < n; i++) {
        for (int k = 0; k < 9; k++) b[k][i] = s[k][i];
        b[9][i] = s[12][i];
        b[10][i] = s[13][i];
        for (int d = 0; d < 3; d++) {
            s[d][i] = b[d][i] + dt * b[3 + d][i];
            s[3 + d][i] = b[3 + d][i] + dt * (b[6 + d][i] + s[9 + d][i]);
        }
        double e = b[9][i] + dt * b[10][i];
        s[12][i] = e < 1e-6 ? 1e-6 : e;
    }
}

void orc_pc_corrector(int64_t n, double dt, double **s, double **b) {
    for (int64_t i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) {
            const double vp = s[3 + d][i], ap = s[6 + d][i], g = s[9 + d][i];
            s[d][i] = b[d][i] + dt / 2 * (vp + b[3 + d][i]);
            s[3 + d][i] = b[3 + d][i] + dt / 2 * (ap + b[6 + d][i] + 2 * g);
            s[6 + d][i] = 0.5 * (ap + b[6 + d][i]) + g;
        }
        double e = b[9][i] + dt / 2 * (s[13][i] + b[10][i]);
        s[12][i] = e < 1e-6 ? 1e-6 : e;
        s[13][i] = 0.5 * (s[13][i] + b[10][i]);
    }
}

/* setTimeStep's per-particle terms (device_predictor_corrector_euler.cu:315-455) */
void orc_time_step_terms(int64_t n, const double *ax, const double *ay, const double *az,
                         const double *gx, const double *gy, const double *gz, const double *h,
                         const double *cs, const double *mu, const double *vx, const double *vy,
                         const double *vz, double alpha, double beta, double terms[5]) {
    double forces = 1.7976931348623157e308, courant = forces, artvisc = forces, vmax = 0, hmax = 0;
    for (int64_t i = 0; i < n; i++) {
        double a0 = ax[i] + gx[i], a1 = ay[i] + gy[i], a2 = az[i] + gz[i];
        double t = a0 * a0;
        t += a1 * a1;
        t += a2 * a2;
        double v = sqrt(h[i] / sqrt(t));
        if (v < forces) forces = v;
        v = h[i] / cs[i];
        if (v < courant) courant = v;
        v = h[i] / (cs[i] + 1.2 * (alpha * cs[i] + beta * mu[i]));
        if (v < artvisc) artvisc = v;
        v = sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
        if (v > vmax) vmax = v;
        if (h[i] > hmax) hmax = h[i];
    }
    terms[0] = forces;
    terms[1] = courant;
    terms[2] = artvisc;
    terms[3] = vmax;
    terms[4] = hmax;
}

/* ------------------------------------------------------------------------------------
 * Learned smoothing length MLP (build extension, no reference counterpart):
 * A1 = tanh(W1 X + b1), A2 = tanh(W2 A1 + b2), Y = W3 . A2 + b3; loss = mean (Y - T)^2 / 2.
 * W1 is [nh][nin], W2 [nh][nh], W3 [nh]; X is [nb][nin].  The backward returns the loss
 * and writes d loss / d parameter.
 * ---------------------------------------------------------------------------------- */
void orc_sph_features(int64_t n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz, const double *m,
                      const double *h, const double *rho, const int32_t *nnl, const int32_t *noi,
                      int max_nb, int sml_avg, int kernel, double *curl, double *gradrho) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double c[3] = {0, 0, 0}, g[3] = {0, 0, 0};
        for (int k = 0; k < noi[i]; k++) {
            const int32_t j = nnl[i * max_nb + k];
            const double sml = sml_avg ? 0.5 * (h[i] + h[j]) : h[i];
            const double dr[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
            const double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            double W, gW[3], dWdr;
            orc_sph_kernel(kernel, r2, dr[0], dr[1], dr[2], sml, &W, gW, &dWdr);
            const double V = m[j] / rho[j];
            const double dv[3] = {vx[j] - vx[i], vy[j] - vy[i], vz[j] - vz[i]};
            c[0] += V * (gW[1] * dv[2] - gW[2] * dv[1]);
            c[1] += V * (gW[2] * dv[0] - gW[0] * dv[2]);
            c[2] += V * (gW[0] * dv[1] - gW[1] * dv[0]);
            for (int e = 0; e < 3; e++) g[e] += m[j] * gW[e];
        }
        for (int e = 0; e < 3; e++) {
            curl[3 * i + e] = c[e];
            gradrho[3 * i + e] = g[e];
        }
    }
}

void orc_mlp_forward(int64_t nb, int nin, int nh, const double *X, const double *W1,
                     const double *b1, const double *W2, const double *b2, const double *W3,
                     double b3, double *Y, double *A1, double *A2) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < nb; s++) {
        double *a1 = A1 + s * nh, *a2 = A2 + s * nh;
        for (int o = 0; o < nh; o++) {
            double t = b1[o];
            for (int k = 0; k < nin; k++) t += W1[o * nin + k] * X[s * nin + k];
            a1[o] = tanh(t);
        }
        for (int o = 0; o < nh; o++) {
            double t = b2[o];
            for (int k = 0; k < nh; k++) t += W2[o * nh + k] * a1[k];
            a2[o] = tanh(t);
        }
        double y = b3;
        for (int k = 0; k < nh; k++) y += W3[k] * a2[k];
        Y[s] = y;
    }
}

double orc_mlp_backward(int64_t nb, int nin, int nh, const double *X, const double *T,
                        const double *W1, const double *b1, const double *W2,
                        const double *b2, const double *W3, double b3,
                        double *gW1, double *gb1, double *gW2, double *gb2, double *gW3,
                        double *gb3) {
    double *A1 = (double *)malloc(sizeof(double) * (size_t)(nb * nh));
    double *A2 = (double *)malloc(sizeof(double) * (size_t)(nb * nh));
    double *Y = (double *)malloc(sizeof(double) * (size_t)nb);
    orc_mlp_forward(nb, nin, nh, X, W1, b1, W2, b2, W3, b3, Y, A1, A2);
    memset(gW1, 0, sizeof(double) * (size_t)(nh * nin));
    memset(gb1, 0, sizeof(double) * (size_t)nh);
    memset(gW2, 0, sizeof(double) * (size_t)(nh * nh));
    memset(gb2, 0, sizeof(double) * (size_t)nh);
    memset(gW3, 0, sizeof(double) * (size_t)nh);
    *gb3 = 0.0;
    double loss = 0.0;
    double *d2 = (double *)malloc(sizeof(double) * (size_t)nh);
    double *d1 = (double *)malloc(sizeof(double) * (size_t)nh);
    for (int64_t s = 0; s < nb; s++) {
        const double *a1 = A1 + s * nh, *a2 = A2 + s * nh;
        double e = Y[s] - T[s];
        loss += 0.5 * e * e;
        double dy = e / (double)nb;
        *gb3 += dy;
        for (int k = 0; k < nh; k++) {
            gW3[k] += dy * a2[k];
            d2[k] = dy * W3[k] * (1.0 - a2[k] * a2[k]);
        }
        for (int o = 0; o < nh; o++) {
            gb2[o] += d2[o];
            for (int k = 0; k < nh; k++) gW2[o * nh + k] += d2[o] * a1[k];
        }
        for (int k = 0; k < nh; k++) {
            double t = 0.0;
            for (int o = 0; o < nh; o++) t += W2[o * nh + k] * d2[o];
            d1[k] = t * (1.0 - a1[k] * a1[k]);
        }
        for (int o = 0; o < nh; o++) {
            gb1[o] += d1[o];
            for (int k = 0; k < nin; k++) gW1[o * nin + k] += d1[o] * X[s * nin + k];
        }
    }
    free(d1);
    free(d2);
    free(A1);
    free(A2);
    free(Y);
    return loss / (double)nb;
}

/*
 * baseline.c -- the host-core baseline: the oracle's composite passes (oracle.c) timed at C
 * level, so no Python marshalling falls inside the timed region.  Test and benchmark
 * infrastructure like the rest of oracle/: nothing in the product links or calls it.
 */
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

#define BENCH_TMAX 4096

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int orc_bench_sph_rhs(int64_t n, const double *x, const double *y, const double *z,
                      const double *vx, const double *vy, const double *vz, const double *m,
                      const double *h0, const double *e, const double *p0, const double *cs0,
                      int n_target, int band_lo, int band_hi, int max_counts, int max_nb, double sml_max,
                      int eos, double gamma, double K, double alpha, double beta, int reps, double *seconds) {
    if (n < 1 || reps < 1 || max_nb < 1) return ORC_E_ARG;
    const size_t N = (size_t)n;
    double *table = malloc(sizeof(double) * (BENCH_TMAX + 1));
    double *h = malloc(sizeof(double) * N), *rho = malloc(sizeof(double) * N);
    double *p = malloc(sizeof(double) * N), *cs = malloc(sizeof(double) * N);
    double *out = malloc(sizeof(double) * N * 7);
    int32_t *iters = malloc(sizeof(int32_t) * N), *noi = malloc(sizeof(int32_t) * N);
    int32_t *nnl = malloc(sizeof(int32_t) * N * (size_t)max_nb);
    int st = ORC_OK;
    if (!table || !h || !rho || !p || !cs || !out || !iters || !noi || !nnl) {
        st = ORC_E_ARG;
        goto done;
    }
    orc_h_factor_table(n_target, BENCH_TMAX, table);
    for (int r = 0; r < reps && st == ORC_OK; r++) {
        const double t0 = now_s();
        memcpy(h, h0, sizeof(double) * N);
        memcpy(p, p0, sizeof(double) * N);
        memcpy(cs, cs0, sizeof(double) * N);
        st = orc_h_iterate(n, x, y, z, h, n_target, band_lo, band_hi, max_counts, max_nb, table, BENCH_TMAX,
                           sml_max, iters, noi);
        if (st == ORC_OK) st = orc_nn_grid(n, x, y, z, h, max_nb, nnl, noi);
        if (st == ORC_OK) st = orc_density(n, x, y, z, m, h, nnl, noi, max_nb, 1, ORC_KERNEL_CUBIC_SPLINE, rho);
        if (st == ORC_OK) {
            orc_soundspeed(n, eos, gamma, K, rho, p, cs);
            orc_pressure(n, eos, gamma, K, rho, e, cs, p);
            orc_forces(n, x, y, z, vx, vy, vz, m, h, rho, p, cs, nnl, noi, max_nb, 1, ORC_KERNEL_CUBIC_SPLINE, alpha,
                       beta, out, out + N, out + 2 * N, out + 3 * N, out + 4 * N, out + 5 * N, out + 6 * N);
        }
        seconds[r] = now_s() - t0;
    }
done:
    free(table);
    free(h);
    free(rho);
    free(p);
    free(cs);
    free(out);
    free(iters);
    free(noi);
    free(nnl);
    return st;
}

int orc_bench_gravity(int64_t n, const double *x, const double *y, const double *z, const double *m,
                      double theta, double smoothing, int curve, int mac, int quad, int group,
                      int64_t sample_every, double seconds[2], int64_t *walked, int64_t counts[2]) {
    if (n < 1 || group < 1 || sample_every < 1) return ORC_E_ARG;
    const size_t N = (size_t)n;
    const int64_t cap = 2 * n + 64, tot = n + cap;
    uint64_t *keys = malloc(sizeof(uint64_t) * N);
    int32_t *perm = malloc(sizeof(int32_t) * N);
    int32_t *child = malloc(sizeof(int32_t) * 8 * (size_t)tot), *level = malloc(sizeof(int32_t) * (size_t)tot);
    double *nx = malloc(sizeof(double) * (size_t)tot), *ny = malloc(sizeof(double) * (size_t)tot);
    double *nz = malloc(sizeof(double) * (size_t)tot), *nm = calloc((size_t)tot, sizeof(double));
    double *q = quad ? calloc(6 * (size_t)tot, sizeof(double)) : NULL;
    double *acc = malloc(sizeof(double) * N * 3);
    const int64_t ngroups = (n + group - 1) / group;
    const int64_t nsel = (ngroups + sample_every - 1) / sample_every;
    int64_t *gsel = malloc(sizeof(int64_t) * (size_t)nsel);
    int st = ORC_OK;
    if (!keys || !perm || !child || !level || !nx || !ny || !nz || !nm || (quad && !q) || !acc || !gsel) {
        st = ORC_E_ARG;
        goto done;
    }
    for (int64_t t = 0; t < nsel; t++) gsel[t] = t * sample_every;
    double t0 = now_s();
    double box[6];
    int64_t ncells = 0;
    orc_bbox(n, x, y, z, box);
    orc_keys(n, x, y, z, box, curve, keys);
    orc_sort(n, keys, perm);
    st = orc_tree_build(n, x, y, z, box, cap, child, level, nx, ny, nz, &ncells);
    if (st == ORC_OK) {
        memcpy(nm, m, sizeof(double) * N);
        orc_com(n, ncells, child, level, nm, nx, ny, nz);
        if (quad) orc_quadrupole(n, ncells, child, level, nm, nx, ny, nz, q);
    }
    seconds[0] = now_s() - t0;
    if (st == ORC_OK) {
        t0 = now_s();
        st = orc_bh_group_subset(n, perm, group, child, level, nm, nx, ny, nz, q, box, theta, smoothing, mac,
                                 nsel, gsel, acc, acc + N, acc + 2 * N, NULL, &counts[0], &counts[1]);
        seconds[1] = now_s() - t0;
        *walked = nsel;
    }
done:
    free(keys);
    free(perm);
    free(child);
    free(level);
    free(nx);
    free(ny);
    free(nz);
    free(nm);
    free(q);
    free(acc);
    free(gsel);
    return st;
}

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
#include "mhpc_internal.h"

#define NN_WAVES 2
#define NN_STACK 512
#define NN_STACK_PHYS (NN_STACK + 192)
#define NN_RANGES 128
#define NN_LEAF 16

struct NNWaveLds {
    int stack[NN_STACK_PHYS];
    int rstart[NN_RANGES];
    int rcnt[NN_RANGES];
    int rpre[64];
    double tx[64], ty[64], tz[64], th2[64];
    int tj[64];
};

struct NNArgs {
    int64_t n;
    const int32_t *perm;
    const double *sx, *sy, *sz;
    const double *h;          // h used for the search (h_cur in the h-iteration)
    const int32_t *cstart, *cend, *cchild;
    const double *cbmin, *cbmax;
    int64_t cap_cells;
    // list mode
    int32_t *nbr;
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
    uint32_t *err;
    unsigned long long *counters;
};

// MODE 0: build lists; MODE 1: one Jacobi pass of the h-iteration (count only)
template <int MODE>
__global__ __launch_bounds__(64 * NN_WAVES) void k_nn_walk(NNArgs A) {
    __shared__ NNWaveLds lds_all[NN_WAVES];
    const int w = threadIdx.x / 64, lane = wave_lane();
    NNWaveLds &L = lds_all[w];
    const int64_t group = (int64_t)blockIdx.x * NN_WAVES + w;
    const int64_t k = group * 64 + lane;
    const int64_t n = A.n;
    if (group * 64 >= n) return;
    const bool valid = k < n;
    bool active = valid;
    if (MODE == 1 && valid) active = A.done[k] == 0;
    const uint64_t mActive = wave_ballot(active);
    if (mActive == 0) {
        if (MODE == 1 && valid) A.h_next[k] = A.h[k];
        return;
    }
    const double xi = valid ? A.sx[k] : 0.0, yi = valid ? A.sy[k] : 0.0, zi = valid ? A.sz[k] : 0.0;
    const double hi = valid ? A.h[k] : 0.0;
    const double hi2 = hi * hi;
    // search box: bounding box of the active lanes grown by their largest h (plus a
    // relative margin far above FP64 rounding, so pruning never drops a true pair)
    const double hmax = wave_max(active ? hi : 0.0);
    const double grow = hmax * (1.0 + 1e-9);
    const double lox = wave_min(active ? xi : 1e300) - grow, hix = wave_max(active ? xi : -1e300) + grow;
    const double loy = wave_min(active ? yi : 1e300) - grow, hiy = wave_max(active ? yi : -1e300) + grow;
    const double loz = wave_min(active ? zi : 1e300) - grow, hiz = wave_max(active ? zi : -1e300) + grow;

    int count = 0;
    int top = 1, nr = 0;
    if (lane == 0) L.stack[0] = (int)n;
    wave_sync();
    const uint64_t lt = lanemask_lt();
    bool overflow = false;
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
                const double px = A.sx[node], py = A.sy[node], pz = A.sz[node];
                take = px >= lox && px <= hix && py >= loy && py <= hiy && pz >= loz && pz <= hiz;
                rs = node;
                rc = 1;
            } else if (node >= n) {
                c = node - n;
                const int64_t cc = A.cap_cells;
                const bool overlap = A.cbmin[c] <= hix && A.cbmax[c] >= lox &&
                                     A.cbmin[cc + c] <= hiy && A.cbmax[cc + c] >= loy &&
                                     A.cbmin[2 * cc + c] <= hiz && A.cbmax[2 * cc + c] >= loz;
                if (overlap) {
                    rs = A.cstart[c];
                    rc = A.cend[c] - rs + 1;
                    if (rc <= NN_LEAF) take = true; else open = true;
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
            if (top + tot > NN_STACK_PHYS) {
                overflow = true;
                break;
            }
            if (open) {
                int pos = top + incl - nc;
#pragma unroll
                for (int s = 7; s >= 0; s--)
                    if (ch[s] != -1) L.stack[pos++] = ch[s];
            }
            top += tot;
            wave_sync();
        }
        if (overflow || (nr == 0 && top == 0)) break;
        // stream the candidates of up to 64 ranges through LDS
        const int np = nr < 64 ? nr : 64;
        const int v = lane < np ? L.rcnt[lane] : 0;
        const int incl = wave_incl_scan(v);
        const int total = __shfl(incl, 63);
        L.rpre[lane] = incl - v;
        wave_sync();
        for (int base = 0; base < total; base += 64) {
            const int t = base + lane;
            if (t < total) {
                int rlo = 0, rhi = np - 1;
                while (rlo < rhi) {
                    const int mid = (rlo + rhi + 1) >> 1;
                    if (L.rpre[mid] <= t) rlo = mid; else rhi = mid - 1;
                }
                const int j = L.rstart[rlo] + (t - L.rpre[rlo]);
                const double hj = A.h[j];
                L.tx[lane] = A.sx[j];
                L.ty[lane] = A.sy[j];
                L.tz[lane] = A.sz[j];
                L.th2[lane] = hj * hj;
                L.tj[lane] = j;
            }
            wave_sync();
            const int m = (total - base) < 64 ? (total - base) : 64;
            if (active) {
                for (int e = 0; e < m; e++) {
                    const double dx = xi - L.tx[e];
                    const double dy = yi - L.ty[e];
                    const double dz = zi - L.tz[e];
                    const double d = dx * dx + dy * dy + dz * dz;
                    const int j = L.tj[e];
                    if (d < hi2 && d < L.th2[e] && j != (int)k) {
                        if (MODE == 0 && count < A.max_nb) A.nbr[(int64_t)count * A.stride + k] = j;
                        count++;
                    }
                }
            }
            wave_sync();
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
    if (overflow) {
        if (lane == 0) set_err(A.err, ERRF_STACK);
        return;
    }
    if (MODE == 0) {
        if (valid) {
            A.snoi[k] = count;
            const int32_t i = A.perm[k];
            A.noi_out[i] = count;
            if (A.sml_out) A.sml_out[i] = hi;
            if (A.iters_out) A.iters_out[i] = A.siters[k];
            if (count > A.max_nb) set_err(A.err, ERRF_CAP_NEIGHBOURS);
        }
        unsigned long long s = (unsigned long long)wave_sum(valid ? count : 0);
        int mx = wave_max(valid ? count : 0);
        if (lane == 0) {
            atomicAdd(&A.counters[CNT_NEIGHBOURS], s);
            atomicMax(&A.counters[CNT_MAX_NOI], (unsigned long long)mx);
        }
    } else {
        if (valid) {
            if (active) {
                const int N = count;
                A.iters[k] = A.pass + 1;
                A.snoi[k] = N;
                const bool inband = N >= A.band_lo && N <= A.band_hi;
                const bool last = A.pass == A.max_counts - 1;
                bool fin = false;
                double hn = hi;
                if ((inband || last) && N < A.max_nb) {
                    fin = true;
                } else {
                    int Nc = N < 1 ? 1 : N;
                    Nc = Nc > A.tmax ? A.tmax : Nc;
                    hn = hi * A.table[Nc];
                    if (last) fin = true;
                }
                A.h_next[k] = hn;
                if (fin) {
                    A.done[k] = 1;
                    if (N >= A.max_nb || N == 0) set_err(A.err, ERRF_CAP_NEIGHBOURS);
                }
            } else {
                A.h_next[k] = hi;
            }
        }
        if (lane == 0) atomicAdd(&A.counters[CNT_H_COUNTS], (unsigned long long)popc64(mActive));
    }
}

static NNArgs nn_common(mhpc_ctx *c) {
    NNArgs a = {};
    a.n = c->n;
    a.perm = c->perm;
    a.sx = c->sx; a.sy = c->sy; a.sz = c->sz;
    a.cstart = c->cstart; a.cend = c->cend; a.cchild = c->cchild;
    a.cbmin = c->cbmin; a.cbmax = c->cbmax;
    a.cap_cells = c->cap_cells;
    a.max_nb = c->p.max_nb;
    a.err = c->err;
    a.counters = c->counters;
    return a;
}

static unsigned nn_grid(mhpc_ctx *c) { return (unsigned)ceil_div(ceil_div(c->n, 64), NN_WAVES); }

__global__ void k_h_init(int64_t n, uint8_t *done, int32_t *iters) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n) {
        done[k] = 0;
        iters[k] = 0;
    }
}

// a9: the plain h-iteration as Jacobi passes; sh holds the final h afterwards
void launch_h_iterate(mhpc_ctx *c, int passes) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_h_init, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       c->sdone, c->siters);
    NNArgs a = nn_common(c);
    a.done = c->sdone;
    a.iters = c->siters;
    a.snoi = c->snoi;
    a.table = c->htable;
    a.tmax = MHPC_HTABLE_MAX;
    a.max_counts = passes;
    a.band_lo = c->p.band_lo;
    a.band_hi = c->p.band_hi;
    for (int pass = 0; pass < passes; pass++) {
        a.pass = pass;
        a.h = c->sh;
        a.h_next = c->sh2;
        hipLaunchKernelGGL(k_nn_walk<1>, dim3(nn_grid(c)), dim3(64 * NN_WAVES), 0, c->stream, a);
        double *t = c->sh;
        c->sh = c->sh2;
        c->sh2 = t;
    }
}

void launch_neighbours(mhpc_ctx *c) {
    if (c->n <= 0) return;
    NNArgs a = nn_common(c);
    a.h = c->sh;
    a.nbr = c->nbr;
    a.stride = c->cap_pad;
    a.snoi = c->snoi;
    a.noi_out = c->noi;
    a.sml_out = nullptr;
    a.siters = c->siters;
    a.iters_out = nullptr;
    hipLaunchKernelGGL(k_nn_walk<0>, dim3(nn_grid(c)), dim3(64 * NN_WAVES), 0, c->stream, a);
}



// ------------------------------------------------------------------------------------
// a10 cubic spline (kernel.cu:50-83), support radius h.  FP64 form: the reference's
// expressions in the same order; FP32 form: the same formula in float.
__device__ __forceinline__ double w_cubic64(double r, double h) {
    const double q = r / h;
    const double f = 8. / M_PI * 1. / (h * h * h);
    if (q > 1) return 0.0;
    if (q > 0.5) return 2. * f * (1. - q) * (1. - q) * (1 - q);
    return f * (6. * q * q * q - 6. * q * q + 1.);
}

__device__ __forceinline__ double dwdr_cubic64(double r, double h) {
    const double q = r / h;
    const double f = 8. / M_PI * 1. / (h * h * h);
    if (q > 1) return 0.0;
    if (q > 0.5) return -6. * f * 1. / h * (1. - q) * (1. - q);
    return 6. * f / h * (3 * q * q - 2 * q);
}

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
// a11 density (density.cu:8-72): rho_i = m_i W(0, h_i) + sum_j m_j W(r_ij, h_ij)
template <bool FP64>
__global__ __launch_bounds__(256) void k_density(int64_t n, const double *__restrict__ sx,
                                                 const double *__restrict__ sy,
                                                 const double *__restrict__ sz,
                                                 const double *__restrict__ sm,
                                                 const double *__restrict__ sh,
                                                 const int32_t *__restrict__ snoi,
                                                 const int32_t *__restrict__ nbr, int64_t stride,
                                                 int max_nb, int sml_avg, const int32_t *__restrict__ perm,
                                                 double *__restrict__ srho, double *__restrict__ rho,
                                                 uint32_t *err) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double xi = sx[k], yi = sy[k], zi = sz[k], hi = sh[k];
    int cnt = snoi[k];
    cnt = cnt > max_nb ? max_nb : cnt;
    double r;
    if constexpr (FP64) {
        r = sm[k] * w_cubic64(0.0, hi);
        for (int s = 0; s < cnt; s++) {
            const int32_t j = nbr[(int64_t)s * stride + k];
            const double sml = sml_avg ? 0.5 * (hi + sh[j]) : hi;
            const double dx = xi - sx[j], dy = yi - sy[j], dz = zi - sz[j];
            const double r2 = dx * dx + dy * dy + dz * dz;
            r += sm[j] * w_cubic64(sqrt(r2), sml);
        }
    } else {
        Kahan acc;
        for (int s = 0; s < cnt; s++) {
            const int32_t j = nbr[(int64_t)s * stride + k];
            const float sml = (float)(sml_avg ? 0.5 * (hi + sh[j]) : hi);
            const float dx = (float)(xi - sx[j]), dy = (float)(yi - sy[j]), dz = (float)(zi - sz[j]);
            const float rr = sqrtf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)));
            float W, dW;
            cubic32(rr, sml, W, dW);
            acc.add((float)sm[j] * W);
        }
        r = sm[k] * w_cubic64(0.0, hi) + acc.value();
    }
    srho[k] = r;
    rho[perm[k]] = r;
    if (!(r > 0.0)) set_err(err, ERRF_RHO);
}

void launch_density(mhpc_ctx *c) {
    if (c->n <= 0) return;
    const unsigned g = (unsigned)ceil_div(c->n, 256);
    if (c->p.precision == MHPC_PREC_FP64)
        hipLaunchKernelGGL(k_density<true>, dim3(g), dim3(256), 0, c->stream, c->n, c->sx, c->sy, c->sz,
                           c->sm, c->sh, c->snoi, c->nbr, c->cap_pad, c->p.max_nb, c->p.sml_avg,
                           c->perm, c->srho, c->f[MHPC_F_RHO], c->err);
    else
        hipLaunchKernelGGL(k_density<false>, dim3(g), dim3(256), 0, c->stream, c->n, c->sx, c->sy, c->sz,
                           c->sm, c->sh, c->snoi, c->nbr, c->cap_pad, c->p.max_nb, c->p.sml_avg,
                           c->perm, c->srho, c->f[MHPC_F_RHO], c->err);
}

// ------------------------------------------------------------------------------------
// a12 / a13 sound speed (soundspeed.cu:38-97) and pressure (pressure.cu:6-87), fused.
__device__ __forceinline__ double eos_cs(int eos, double gamma, double K, double rho, double p,
                                         double cs) {
    switch (eos) {
        case MHPC_EOS_POLYTROPIC: return sqrt(K * pow(rho, gamma - 1.0));
        case MHPC_EOS_IDEAL_GAS: return sqrt(gamma * p / rho);
        case MHPC_EOS_LOCALLY_ISOTHERMAL: return 0.0;
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

__global__ __launch_bounds__(256) void k_eos(int64_t n, int eos, double gamma, double K,
                                             int cs_before_p, const double *__restrict__ srho,
                                             const double *__restrict__ se, double *__restrict__ sp,
                                             double *__restrict__ scs, const int32_t *__restrict__ perm,
                                             double *__restrict__ p_out, double *__restrict__ cs_out) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double rho = srho[k], e = se[k];
    double p = sp[k], cs = scs[k];
    if (cs_before_p) {
        cs = eos_cs(eos, gamma, K, rho, p, cs);
        p = eos_p(eos, gamma, K, rho, e, cs);
    } else {
        p = eos_p(eos, gamma, K, rho, e, cs);
        cs = eos_cs(eos, gamma, K, rho, p, cs);
    }
    sp[k] = p;
    scs[k] = cs;
    const int32_t i = perm[k];
    p_out[i] = p;
    cs_out[i] = cs;
}

void launch_eos(mhpc_ctx *c) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_eos, dim3((unsigned)ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       c->p.eos, c->p.gamma, c->p.K, c->p.cs_before_p, c->srho, c->se, c->sp, c->scs,
                       c->perm, c->f[MHPC_F_P], c->f[MHPC_F_CS]);
}

// ------------------------------------------------------------------------------------
// a14 internal forces (internal_forces.cu:4-668): SPH_EQU_VERSION 1 pressure term,
// Monaghan viscosity, drho/dt, de/dt (INTEGRATE_ENERGY), dh/dt (INTEGRATE_SML).
struct ForceArgs {
    int64_t n;
    const double *sx, *sy, *sz, *svx, *svy, *svz, *sm, *sh, *srho, *sp, *scs;
    const int32_t *snoi, *nbr;
    int64_t stride;
    int max_nb, sml_avg;
    double alpha, beta;
    const int32_t *perm;
    double *ax, *ay, *az, *drhodt, *dedt, *dsmldt;
};

template <bool FP64>
__global__ __launch_bounds__(256) void k_forces(ForceArgs A) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= A.n) return;
    const double xi = A.sx[k], yi = A.sy[k], zi = A.sz[k];
    const double vxi = A.svx[k], vyi = A.svy[k], vzi = A.svz[k];
    const double hi = A.sh[k], rhoi = A.srho[k], pi = A.sp[k], csi = A.scs[k];
    int cnt = A.snoi[k];
    cnt = cnt > A.max_nb ? A.max_nb : cnt;
    double out[6];
    if constexpr (FP64) {
        double acc0 = 0, acc1 = 0, acc2 = 0, drho = 0, de = 0, dsml = 0;
        const double pri = pi / (rhoi * rhoi);
        for (int s = 0; s < cnt; s++) {
            const int32_t j = A.nbr[(int64_t)s * A.stride + k];
            const double hj = A.sh[j], rhoj = A.srho[j], mj = A.sm[j];
            const double sml = A.sml_avg ? 0.5 * (hi + hj) : hi;
            const double dr0 = xi - A.sx[j], dr1 = yi - A.sy[j], dr2 = zi - A.sz[j];
            double r2 = 0.0;
            r2 += dr0 * dr0;
            r2 += dr1 * dr1;
            r2 += dr2 * dr2;
            const double r = sqrt(r2);
            const double dWdr = dwdr_cubic64(r, sml);
            const double g0 = dWdr / r * dr0, g1 = dWdr / r * dr1, g2 = dWdr / r * dr2;
            const double dv0 = vxi - A.svx[j], dv1 = vyi - A.svy[j], dv2 = vzi - A.svz[j];
            double vvnablaW = dv0 * g0;
            vvnablaW += dv1 * g1;
            vvnablaW += dv2 * g2;
            double rr = 0.0, vr = 0.0;
            rr += dr0 * dr0; vr += dv0 * dr0;
            rr += dr1 * dr1; vr += dv1 * dr1;
            rr += dr2 * dr2; vr += dv2 * dr2;
            double pij = 0.0;
            if (vr < 0) {
                const double csbar = 0.5 * (csi + A.scs[j]);
                const double smooth = 0.5 * (hi + hj);
                const double mu = smooth * vr / (rr + smooth * smooth * 1e-2);
                const double rhobar = 0.5 * (rhoi + rhoj);
                pij = (A.beta * mu - A.alpha * csbar) * mu / rhobar;
            }
            const double pf = -mj * (pri + A.sp[j] / (rhoj * rhoj));
            const double as0 = pf * g0, as1 = pf * g1, as2 = pf * g2;
            acc0 += as0; acc1 += as1; acc2 += as2;
            acc0 += mj * (-pij) * g0;
            acc1 += mj * (-pij) * g1;
            acc2 += mj * (-pij) * g2;
            drho += rhoi / rhoj * mj * vvnablaW;
            dsml -= 1. / 3 * hi * mj / rhoj * vvnablaW;
            de += 0.5 * mj * pij * vvnablaW;
            de += 0.5 * as0 * -dv0;
            de += 0.5 * as1 * -dv1;
            de += 0.5 * as2 * -dv2;
        }
        out[0] = acc0; out[1] = acc1; out[2] = acc2; out[3] = drho; out[4] = de; out[5] = dsml;
    } else {
        Kahan a0, a1, a2, drho, de, dsml;
        const float hif = (float)hi, rhoif = (float)rhoi, csif = (float)csi;
        const float pri = (float)(pi / (rhoi * rhoi));
        const float alpha = (float)A.alpha, beta = (float)A.beta;
        for (int s = 0; s < cnt; s++) {
            const int32_t j = A.nbr[(int64_t)s * A.stride + k];
            const float hj = (float)A.sh[j], rhoj = (float)A.srho[j], mj = (float)A.sm[j];
            const float sml = A.sml_avg ? (float)(0.5 * (hi + A.sh[j])) : hif;
            const float dr0 = (float)(xi - A.sx[j]), dr1 = (float)(yi - A.sy[j]), dr2 = (float)(zi - A.sz[j]);
            const float dv0 = (float)(vxi - A.svx[j]), dv1 = (float)(vyi - A.svy[j]),
                        dv2 = (float)(vzi - A.svz[j]);
            const float rr = __builtin_fmaf(dr0, dr0, __builtin_fmaf(dr1, dr1, dr2 * dr2));
            const float r = sqrtf(rr);
            float W, dWdr;
            cubic32(r, sml, W, dWdr);
            const float gs = dWdr / r;
            const float g0 = gs * dr0, g1 = gs * dr1, g2 = gs * dr2;
            const float vvnablaW = __builtin_fmaf(dv0, g0, __builtin_fmaf(dv1, g1, dv2 * g2));
            const float vr = __builtin_fmaf(dv0, dr0, __builtin_fmaf(dv1, dr1, dv2 * dr2));
            float pij = 0.f;
            if (vr < 0.f) {
                const float csbar = 0.5f * (csif + (float)A.scs[j]);
                const float smooth = 0.5f * (hif + hj);
                const float mu = smooth * vr / __builtin_fmaf(smooth * smooth, 1e-2f, rr);
                const float rhobar = 0.5f * (rhoif + rhoj);
                pij = (beta * mu - alpha * csbar) * mu / rhobar;
            }
            const float pf = -mj * (pri + (float)(A.sp[j] / (A.srho[j] * A.srho[j])));
            const float as0 = pf * g0, as1 = pf * g1, as2 = pf * g2;
            const float mp = -mj * pij;
            a0.add(as0 + mp * g0);
            a1.add(as1 + mp * g1);
            a2.add(as2 + mp * g2);
            drho.add(rhoif / rhoj * mj * vvnablaW);
            dsml.add(-(1.f / 3.f) * hif * mj / rhoj * vvnablaW);
            de.add(0.5f * mj * pij * vvnablaW -
                   0.5f * __builtin_fmaf(as0, dv0, __builtin_fmaf(as1, dv1, as2 * dv2)));
        }
        out[0] = a0.value(); out[1] = a1.value(); out[2] = a2.value();
        out[3] = drho.value(); out[4] = de.value(); out[5] = dsml.value();
    }
    const int32_t i = A.perm[k];
    A.ax[i] = out[0];
    A.ay[i] = out[1];
    A.az[i] = out[2];
    A.drhodt[i] = out[3];
    A.dedt[i] = out[4];
    A.dsmldt[i] = out[5];
}

void launch_forces(mhpc_ctx *c) {
    if (c->n <= 0) return;
    ForceArgs a;
    a.n = c->n;
    a.sx = c->sx; a.sy = c->sy; a.sz = c->sz;
    a.svx = c->svx; a.svy = c->svy; a.svz = c->svz;
    a.sm = c->sm; a.sh = c->sh; a.srho = c->srho; a.sp = c->sp; a.scs = c->scs;
    a.snoi = c->snoi; a.nbr = c->nbr; a.stride = c->cap_pad;
    a.max_nb = c->p.max_nb; a.sml_avg = c->p.sml_avg;
    a.alpha = c->p.alpha; a.beta = c->p.beta;
    a.perm = c->perm;
    a.ax = c->f[MHPC_F_AX]; a.ay = c->f[MHPC_F_AY]; a.az = c->f[MHPC_F_AZ];
    a.drhodt = c->f[MHPC_F_DRHODT]; a.dedt = c->f[MHPC_F_DEDT]; a.dsmldt = c->f[MHPC_F_DSMLDT];
    const unsigned g = (unsigned)ceil_div(c->n, 256);
    if (c->p.precision == MHPC_PREC_FP64)
        hipLaunchKernelGGL(k_forces<true>, dim3(g), dim3(256), 0, c->stream, a);
    else
        hipLaunchKernelGGL(k_forces<false>, dim3(g), dim3(256), 0, c->stream, a);
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
    for (int s = 0; s < cnt; s++) nnl[(int64_t)i * max_nb + s] = perm[nbr[(int64_t)s * stride + k]];
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

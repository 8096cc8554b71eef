// k_gravity.hip -- Barnes-Hut force, wavefront-cooperative.
//
// The reference walks the tree once per particle (computeForces_v1, gravity.cu:252-396)
// or once per warp of 32 sorted particles with a warp vote on the acceptance test
// (computeForces_v2, gravity.cu:726-903).  Here one wavefront owns a GROUP of 64
// consecutive key-sorted particles and walks the tree once for all of them:
//   * the 64 lanes test up to 64 stacked nodes at a time (FP64, exact, the acceptance
//     test the oracle restates), so the traversal itself is parallel across lanes.  Vote
//     form (the reference v2's warp vote, gravity.cu:836, over the group of 64): a cell is
//     accepted only if every member accepts it with its own distance.  A lane's cell that
//     passes the test against the group's bounding box passes it for every member (the box
//     distance is a lower bound of each member's, also after rounding), so only the other
//     cells are put to the vote: for each, all 64 lanes test it against their own particle
//     and one ballot decides.  Box form: the bounding-box test alone (one test per group);
//     it opens more cells;
//   * leaves and accepted cells go to two LDS interaction lists shared by the wave;
//   * each time a list holds 64 entries all 64 lanes evaluate it, every lane reading
//     the same LDS entry (broadcast) for its own particle.  FP32 mode: entries are
//     staged as FP32 positions relative to the group's centre (each rounded once from the
//     FP64 difference), each batch of up to 64 entries is summed in FP32 and the batch
//     sums are added with Kahan compensation.  FP64 mode: plain FP64 throughout.
// The stack lives in LDS; the number of nodes popped per round shrinks as the stack
// fills so it cannot overflow (a hard bound still sets ERRF_STACK).
#include "mhpc_internal.h"

#define BH_WAVES 2                       // waves (groups) per block
static_assert(MHPC_BH_GROUP == 64, "one Barnes-Hut group per wavefront");
#define BH_STACK 768                     // soft stack capacity per wave
#define BH_STACK_PHYS (BH_STACK + 192)   // hard capacity
#define BH_LIST 128

// FP64 mode: list entries keep FP64 positions
struct BHWaveLds64 {
    int stack[BH_STACK_PHYS];
    // particle list
    double px[BH_LIST], py[BH_LIST], pz[BH_LIST];
    double pm[BH_LIST];
    int pid[BH_LIST];
    // cell list
    double cx[BH_LIST], cy[BH_LIST], cz[BH_LIST];
    double cm[BH_LIST];
    float4 q0[BH_LIST];
    float2 q1[BH_LIST];
};

// FP32 mode: list entries are FP32 {x, y, z, m} relative to the group's centre (rounded
// once per entry from the FP64 difference, instead of once per lane and interaction)
struct BHWaveLds32 {
    int stack[BH_STACK_PHYS];
    float4 p[BH_LIST];
    int pid[BH_LIST];
    float4 c[BH_LIST];
    float4 q0[BH_LIST];
    float2 q1[BH_LIST];
};

template <bool FP64>
struct BHLds;
template <>
struct BHLds<true> { typedef BHWaveLds64 type; };
template <>
struct BHLds<false> { typedef BHWaveLds32 type; };

template <bool FP64>
struct BHAcc;

template <>
struct BHAcc<false> {
    Kahan x, y, z, phi;
};

template <>
struct BHAcc<true> {
    Plain64 x, y, z, phi;
};

// 1/sqrt(x) for x >= 0, 0 for x = 0, from the hardware reciprocal square root: for normal x
// the same value rsqrtf gives, without the compiler's scaling for denormal inputs (four
// instructions per interaction).  x = r^2 + eps^2 is normal unless eps = 0 and two particles
// are closer than ~1e-19; such x are raised to the smallest normal float.
__device__ __forceinline__ float rsq_pos(float x) {
    return x > 0.f ? __builtin_amdgcn_rsqf(fmaxf(x, 1.17549435e-38f)) : 0.f;
}

// FP32 batches: plain FP32 sums over the (at most 64) entries of one batch, then one
// compensated (Kahan) addition of each batch sum into the running total
template <bool POT>
__device__ __forceinline__ void eval_particles32(const BHWaveLds32 &L, int cnt, float px, float py, float pz,
                                                 int self, float eps2, BHAcc<false> &acc) {
    float bx = 0.f, by = 0.f, bz = 0.f, bp = 0.f;
    for (int e = 0; e < cnt; e++) {
        const float4 q = L.p[e];
        const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float rinv = rsq_pos(r2);
        const float mr = q.w * rinv;
        const float mr3 = mr * rinv * rinv;
        bx = __builtin_fmaf(mr3, dx, bx);
        by = __builtin_fmaf(mr3, dy, by);
        bz = __builtin_fmaf(mr3, dz, bz);
        if (POT && L.pid[e] != self) bp -= mr;
    }
    acc.x.add(bx);
    acc.y.add(by);
    acc.z.add(bz);
    if (POT) acc.phi.add(bp);
}

template <bool POT, bool QUAD>
__device__ __forceinline__ void eval_cells32(const BHWaveLds32 &L, int cnt, float px, float py, float pz,
                                             float eps2, BHAcc<false> &acc) {
    float bx = 0.f, by = 0.f, bz = 0.f, bp = 0.f;
    for (int e = 0; e < cnt; e++) {
        const float4 q = L.c[e];
        const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        const float rinv = rsq_pos(r2);
        const float rinv2 = rinv * rinv;
        const float mr = q.w * rinv;
        const float mr3 = mr * rinv2;
        bx = __builtin_fmaf(mr3, dx, bx);
        by = __builtin_fmaf(mr3, dy, by);
        bz = __builtin_fmaf(mr3, dz, bz);
        if (POT) bp -= mr;
        if (QUAD) {
            const float4 a = L.q0[e];
            const float2 b = L.q1[e];
            const float qdx = __builtin_fmaf(a.x, dx, __builtin_fmaf(a.y, dy, a.z * dz));
            const float qdy = __builtin_fmaf(a.y, dx, __builtin_fmaf(a.w, dy, b.x * dz));
            const float qdz = __builtin_fmaf(a.z, dx, __builtin_fmaf(b.x, dy, b.y * dz));
            const float dqd = __builtin_fmaf(dx, qdx, __builtin_fmaf(dy, qdy, dz * qdz));
            const float rinv5 = rinv2 * rinv2 * rinv;
            const float t = 2.5f * dqd * rinv5 * rinv2;
            bx = __builtin_fmaf(t, dx, __builtin_fmaf(-qdx, rinv5, bx));
            by = __builtin_fmaf(t, dy, __builtin_fmaf(-qdy, rinv5, by));
            bz = __builtin_fmaf(t, dz, __builtin_fmaf(-qdz, rinv5, bz));
            if (POT) bp = __builtin_fmaf(-0.5f * dqd, rinv5, bp);
        }
    }
    acc.x.add(bx);
    acc.y.add(by);
    acc.z.add(bz);
    if (POT) acc.phi.add(bp);
}

// particle-particle (monopole) batch, FP64 mode
template <bool FP64, bool POT>
__device__ __forceinline__ void eval_particles(const BHWaveLds64 &L, int cnt, double px, double py,
                                               double pz, int self, double eps2, BHAcc<FP64> &acc) {
    for (int e = 0; e < cnt; e++) {
        if constexpr (FP64) {
            const double dx = L.px[e] - px, dy = L.py[e] - py, dz = L.pz[e] - pz;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double rinv = r2 > 0.0 ? 1.0 / sqrt(r2) : 0.0;
            const double mr3 = L.pm[e] * rinv * rinv * rinv;
            acc.x.add(mr3 * dx);
            acc.y.add(mr3 * dy);
            acc.z.add(mr3 * dz);
            if (POT && L.pid[e] != self) acc.phi.add(-L.pm[e] * rinv);
        } else {
            const float dx = (float)(L.px[e] - px), dy = (float)(L.py[e] - py),
                        dz = (float)(L.pz[e] - pz);
            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, (float)eps2)));
            const float rinv = rsq_pos(r2);
            const float mr = (float)L.pm[e] * rinv;
            const float mr3 = mr * rinv * rinv;
            acc.x.add(mr3 * dx);
            acc.y.add(mr3 * dy);
            acc.z.add(mr3 * dz);
            if (POT && L.pid[e] != self) acc.phi.add(-mr);
        }
    }
}

// particle-cell batch: monopole (+ traceless quadrupole)
template <bool FP64, bool POT, bool QUAD>
__device__ __forceinline__ void eval_cells(const BHWaveLds64 &L, int cnt, double px, double py,
                                           double pz, double eps2, BHAcc<FP64> &acc) {
    for (int e = 0; e < cnt; e++) {
        if constexpr (FP64) {
            const double dx = L.cx[e] - px, dy = L.cy[e] - py, dz = L.cz[e] - pz;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double rinv = r2 > 0.0 ? 1.0 / sqrt(r2) : 0.0;
            const double rinv2 = rinv * rinv;
            const double m = L.cm[e];
            const double mr3 = m * rinv * rinv2;
            double fx = mr3 * dx, fy = mr3 * dy, fz = mr3 * dz;
            double ph = -m * rinv;
            if (QUAD) {
                const float4 a = L.q0[e];
                const float2 b = L.q1[e];
                const double qxx = a.x, qxy = a.y, qxz = a.z, qyy = a.w, qyz = b.x, qzz = b.y;
                const double qdx = qxx * dx + qxy * dy + qxz * dz;
                const double qdy = qxy * dx + qyy * dy + qyz * dz;
                const double qdz = qxz * dx + qyz * dy + qzz * dz;
                const double dqd = dx * qdx + dy * qdy + dz * qdz;
                const double rinv5 = rinv2 * rinv2 * rinv;
                const double t = 2.5 * dqd * rinv5 * rinv2;
                fx += -qdx * rinv5 + t * dx;
                fy += -qdy * rinv5 + t * dy;
                fz += -qdz * rinv5 + t * dz;
                ph += -0.5 * dqd * rinv5;
            }
            acc.x.add(fx);
            acc.y.add(fy);
            acc.z.add(fz);
            if (POT) acc.phi.add(ph);
        } else {
            const float dx = (float)(L.cx[e] - px), dy = (float)(L.cy[e] - py),
                        dz = (float)(L.cz[e] - pz);
            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, (float)eps2)));
            const float rinv = rsq_pos(r2);
            const float rinv2 = rinv * rinv;
            const float m = (float)L.cm[e];
            const float mr = m * rinv;
            const float mr3 = mr * rinv2;
            float fx = mr3 * dx, fy = mr3 * dy, fz = mr3 * dz;
            float ph = -mr;
            if (QUAD) {
                const float4 a = L.q0[e];
                const float2 b = L.q1[e];
                const float qdx = __builtin_fmaf(a.x, dx, __builtin_fmaf(a.y, dy, a.z * dz));
                const float qdy = __builtin_fmaf(a.y, dx, __builtin_fmaf(a.w, dy, b.x * dz));
                const float qdz = __builtin_fmaf(a.z, dx, __builtin_fmaf(b.x, dy, b.y * dz));
                const float dqd = __builtin_fmaf(dx, qdx, __builtin_fmaf(dy, qdy, dz * qdz));
                const float rinv5 = rinv2 * rinv2 * rinv;
                const float t = 2.5f * dqd * rinv5 * rinv2;
                fx = __builtin_fmaf(t, dx, __builtin_fmaf(-qdx, rinv5, fx));
                fy = __builtin_fmaf(t, dy, __builtin_fmaf(-qdy, rinv5, fy));
                fz = __builtin_fmaf(t, dz, __builtin_fmaf(-qdz, rinv5, fz));
                ph = __builtin_fmaf(-0.5f * dqd, rinv5, ph);
            }
            acc.x.add(fx);
            acc.y.add(fy);
            acc.z.add(fz);
            if (POT) acc.phi.add(ph);
        }
    }
}

// a uniform lane's double (l must be the same in every lane)
__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

struct BHArgs {
    int64_t n;
    const int32_t *perm;
    const double4 *sxm;
    const int32_t *clevel, *cchild;
    const double *cmass, *cx, *cy, *cz;
    const float4 *cquadf;
    const double *box;
    double theta, eps2, G;
    int mac;
    double *gax, *gay, *gaz, *u;
    uint32_t *err;
    unsigned long long *counters;
};

template <bool FP64, bool POT, bool QUAD, bool VOTE>
__global__ __launch_bounds__(64 * BH_WAVES) void k_bh_walk(BHArgs A) {
    typedef typename BHLds<FP64>::type Lds;
    __shared__ Lds lds_all[BH_WAVES];
    const int w = threadIdx.x / 64, lane = wave_lane();
    Lds &L = lds_all[w];
    const int64_t group = xcd_block(blockIdx.x, gridDim.x) * BH_WAVES + w;
    const int64_t k = group * 64 + lane;
    const int64_t n = A.n;
    if (group * 64 >= n) return;   // whole wave exits together
    const bool valid = k < n;
    const double4 own = valid ? A.sxm[k] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double px = own.x, py = own.y, pz = own.z;
    // group bounding box (exact)
    const double glo_x = wave_min(valid ? px : 1e300), ghi_x = wave_max(valid ? px : -1e300);
    const double glo_y = wave_min(valid ? py : 1e300), ghi_y = wave_max(valid ? py : -1e300);
    const double glo_z = wave_min(valid ? pz : 1e300), ghi_z = wave_max(valid ? pz : -1e300);
    const int nvalid = popc64(wave_ballot(valid));
    // acceptance constants, same expressions as the oracle
    const double R = 0.5 * (A.box[1] - A.box[0]);
    const double s2_0 = 4.0 * R * R;
    const double dp_1 = R * R / A.theta;
    const double theta2 = A.theta * A.theta;
    const double eps2 = A.eps2;
    // FP32 mode: positions relative to the group's centre
    const double refx = 0.5 * (glo_x + ghi_x), refy = 0.5 * (glo_y + ghi_y), refz = 0.5 * (glo_z + ghi_z);
    const float pxr = (float)(px - refx), pyr = (float)(py - refy), pzr = (float)(pz - refz);
    const float eps2f = (float)eps2;

    BHAcc<FP64> acc;
    int top = 1, pl_n = 0, cl_n = 0;
    long long n_cells = 0, n_parts = 0;
    if (lane == 0) L.stack[0] = (int)n;   // root
    wave_sync();
    const uint64_t lt = lanemask_lt();
    bool overflow = false;
    while (top > 0) {
        int kpop = (BH_STACK - top) / 7;
        kpop = kpop < 1 ? 1 : kpop;
        kpop = kpop > 64 ? 64 : kpop;
        kpop = kpop > top ? top : kpop;
        const int node = lane < kpop ? L.stack[top - kpop + lane] : -1;
        top -= kpop;
        wave_sync();
        const bool isPart = node >= 0 && node < n;
        const bool isCell = node >= n;
        bool accept = false, open = false;
        double ccx = 0, ccy = 0, ccz = 0;
        int64_t c = 0;
        int lev = 0;
        // the acceptance test of distance^2 d2 for a cell of level L (same expressions as the
        // oracle)
        auto mac_ok = [&](double d2, int L) {
            if (A.mac == MHPC_MAC_CLASSIC) return (d2 + eps2) * theta2 > s2_0 * ldexp(1.0, -2 * L);
            return dp_1 * ldexp(1.0, -2 * (L - 1)) <= d2 + eps2;
        };
        // everything the round may need of its node is requested together, before the test
        // decides which of it is used: one memory round trip per round instead of one for the
        // test and one more for the children or the list entries
        int4 cha = make_int4(-1, -1, -1, -1), chb = cha;
        double cmass = 0.0;
        float4 qa = make_float4(0.f, 0.f, 0.f, 0.f), qb = qa;
        double4 prec = make_double4(0.0, 0.0, 0.0, 0.0);
        if (isPart) prec = A.sxm[node];
        if (isCell) {
            c = node - n;
            const int4 *cp = reinterpret_cast<const int4 *>(A.cchild + 8 * c);
            cha = cp[0];
            chb = cp[1];
            cmass = A.cmass[c];
            if (QUAD) {
                qa = A.cquadf[2 * c];
                qb = A.cquadf[2 * c + 1];
            }
            ccx = A.cx[c];
            ccy = A.cy[c];
            ccz = A.cz[c];
            lev = A.clevel[c];
            double dx = 0.0, dy = 0.0, dz = 0.0;
            if (ccx < glo_x) dx = glo_x - ccx; else if (ccx > ghi_x) dx = ccx - ghi_x;
            if (ccy < glo_y) dy = glo_y - ccy; else if (ccy > ghi_y) dy = ccy - ghi_y;
            if (ccz < glo_z) dz = glo_z - ccz; else if (ccz > ghi_z) dz = ccz - ghi_z;
            accept = mac_ok(dx * dx + dy * dy + dz * dz, lev);
        }
        if constexpr (VOTE) {
            // cells the box test rejects: every member tests each with its own distance
            uint64_t mAmb = wave_ballot(isCell && !accept);
            uint64_t mVote = 0;
            while (mAmb) {
                const int b = __builtin_ctzll(mAmb);
                mAmb &= mAmb - 1;
                const double bx = readlane_d(ccx, b), by = readlane_d(ccy, b), bz = readlane_d(ccz, b);
                const int bl = __builtin_amdgcn_readlane(lev, b);
                const double dx = bx - px, dy = by - py, dz = bz - pz;
                const bool ok = !valid || mac_ok(dx * dx + dy * dy + dz * dz, bl);
                if (wave_ballot(ok) == ~0ull) mVote |= 1ull << b;
            }
            accept = accept || ((mVote >> lane) & 1ull);
        }
        open = isCell && !accept;
        const uint64_t mP = wave_ballot(isPart), mA = wave_ballot(accept);
        if (isPart) {
            const int pos = pl_n + popc64(mP & lt);
            const double4 r = prec;
            if constexpr (FP64) {
                L.px[pos] = r.x;
                L.py[pos] = r.y;
                L.pz[pos] = r.z;
                L.pm[pos] = r.w;
            } else {
                L.p[pos] = make_float4((float)(r.x - refx), (float)(r.y - refy), (float)(r.z - refz), (float)r.w);
            }
            L.pid[pos] = node;
        }
        if (accept) {
            const int pos = cl_n + popc64(mA & lt);
            if constexpr (FP64) {
                L.cx[pos] = ccx;
                L.cy[pos] = ccy;
                L.cz[pos] = ccz;
                L.cm[pos] = cmass;
            } else {
                L.c[pos] = make_float4((float)(ccx - refx), (float)(ccy - refy), (float)(ccz - refz),
                                       (float)cmass);
            }
            if (QUAD) {
                L.q0[pos] = qa;
                L.q1[pos] = make_float2(qb.x, qb.y);
            }
        }
        pl_n += popc64(mP);
        cl_n += popc64(mA);
        n_parts += popc64(mP);
        n_cells += popc64(mA);
        // push the children of opened cells
        int ch[8];
        int nc = 0;
        if (open) {
            ch[0] = cha.x; ch[1] = cha.y; ch[2] = cha.z; ch[3] = cha.w;
            ch[4] = chb.x; ch[5] = chb.y; ch[6] = chb.z; ch[7] = chb.w;
#pragma unroll
            for (int s = 0; s < 8; s++) nc += ch[s] != -1;
        }
        const int incl = wave_incl_scan(nc);
        const int tot = __shfl(incl, 63);
        if (top + tot > BH_STACK_PHYS) {
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
        if (pl_n >= 64) {
            if constexpr (FP64) eval_particles<FP64, POT>(L, 64, px, py, pz, (int)k, eps2, acc);
            else eval_particles32<POT>(L, 64, pxr, pyr, pzr, (int)k, eps2f, acc);
            wave_sync();
            const int rest = pl_n - 64;
            if (lane < rest) {
                if constexpr (FP64) {
                    const double a = L.px[64 + lane], b = L.py[64 + lane], d = L.pz[64 + lane];
                    const double m = L.pm[64 + lane];
                    L.px[lane] = a; L.py[lane] = b; L.pz[lane] = d; L.pm[lane] = m;
                } else {
                    const float4 a = L.p[64 + lane];
                    L.p[lane] = a;
                }
                const int id = L.pid[64 + lane];
                L.pid[lane] = id;
            }
            pl_n = rest;
            wave_sync();
        }
        if (cl_n >= 64) {
            if constexpr (FP64) eval_cells<FP64, POT, QUAD>(L, 64, px, py, pz, eps2, acc);
            else eval_cells32<POT, QUAD>(L, 64, pxr, pyr, pzr, eps2f, acc);
            wave_sync();
            const int rest = cl_n - 64;
            if (lane < rest) {
                if constexpr (FP64) {
                    const double a = L.cx[64 + lane], b = L.cy[64 + lane], d = L.cz[64 + lane];
                    const double m = L.cm[64 + lane];
                    L.cx[lane] = a; L.cy[lane] = b; L.cz[lane] = d; L.cm[lane] = m;
                } else {
                    const float4 a = L.c[64 + lane];
                    L.c[lane] = a;
                }
                if (QUAD) {
                    const float4 q0 = L.q0[64 + lane];
                    const float2 q1 = L.q1[64 + lane];
                    L.q0[lane] = q0;
                    L.q1[lane] = q1;
                }
            }
            cl_n = rest;
            wave_sync();
        }
    }
    if (overflow) {
        if (lane == 0) set_err(A.err, ERRF_STACK);
        return;
    }
    if constexpr (FP64) {
        if (pl_n > 0) eval_particles<FP64, POT>(L, pl_n, px, py, pz, (int)k, eps2, acc);
        if (cl_n > 0) eval_cells<FP64, POT, QUAD>(L, cl_n, px, py, pz, eps2, acc);
    } else {
        if (pl_n > 0) eval_particles32<POT>(L, pl_n, pxr, pyr, pzr, (int)k, eps2f, acc);
        if (cl_n > 0) eval_cells32<POT, QUAD>(L, cl_n, pxr, pyr, pzr, eps2f, acc);
    }
    if (valid) {
        const int32_t i = A.perm[k];
        A.gax[i] = A.G * acc.x.value();
        A.gay[i] = A.G * acc.y.value();
        A.gaz[i] = A.G * acc.z.value();
        if (POT) A.u[i] = 0.5 * own.w * A.G * acc.phi.value();
    }
    if (lane == 0) {
        atomicAdd(&cnt_slot(A.counters)[CNT_CELL_INT], (unsigned long long)(n_cells * nvalid));
        atomicAdd(&cnt_slot(A.counters)[CNT_PART_INT], (unsigned long long)(n_parts * nvalid));
    }
}

template <bool FP64, bool POT, bool QUAD>
static void launch_bh(mhpc_ctx *c, const BHArgs &a) {
    const int64_t groups = ceil_div(c->n, 64);
    const dim3 grid((unsigned)ceil_div(groups, BH_WAVES)), block(64 * BH_WAVES);
    if (c->p.bh_accept == MHPC_BH_VOTE)
        hipLaunchKernelGGL((k_bh_walk<FP64, POT, QUAD, true>), grid, block, 0, c->stream, a);
    else
        hipLaunchKernelGGL((k_bh_walk<FP64, POT, QUAD, false>), grid, block, 0, c->stream, a);
}

// ------------------------------------------------------------------------------------
// a7' computeForces_v1 (gravity.cu:252-396) as the reference runs it: one thread per particle
// walks the tree depth first from the root, resuming each level at its next child slot
// (curNode / curChild per depth), skipping itself; a child is taken if it is a particle or
// (d^2 + eps^2) theta^2 > cellSize[depth], cellSize[d] = 4 R^2 / 4^d, else descended into.
// Monopole, FP64, in the reference's order and expressions (f = m / d^3 with d the square
// root of the softened distance; u -= 1/2 m_c m_i / d).  The per-depth resume state lives in
// LDS; the tree has at most MHPC_LEVELS + 1 levels below the root.
#define BH1_THREADS 64
#define BH1_DEPTH (MHPC_LEVELS + 3)

template <bool POT>
__global__ __launch_bounds__(BH1_THREADS) void k_bh_v1(BHArgs A) {
    __shared__ int cur_node[BH1_DEPTH][BH1_THREADS];
    __shared__ unsigned char cur_child[BH1_DEPTH][BH1_THREADS];
    const int t = threadIdx.x;
    const int64_t k = blockIdx.x * (int64_t)BH1_THREADS + t;
    const int64_t n = A.n;
    // lanes past the last particle walk nothing but take part in the wave's counter sums
    const bool valid = k < n;
    const double4 own = valid ? A.sxm[k] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double px = own.x, py = own.y, pz = own.z;
    const double R = 0.5 * (A.box[1] - A.box[0]);
    const double cs0 = 4.0 * R * R;
    const double thetasq = A.theta * A.theta;
    double axi = 0.0, ayi = 0.0, azi = 0.0, ui = 0.0;
    long long n_cells = 0, n_parts = 0;
    int depth = valid ? 1 : 0;
    cur_node[1][t] = (int)n;   // root
    cur_child[1][t] = 0;
    bool overflow = false;
    while (depth > 0) {
        int child_number = cur_child[depth][t];
        int node = cur_node[depth][t];
        // cellSize[depth] of the reference's table: 4 R^2 / 4^depth, exact powers of 4
        double cell_size = cs0 * ldexp(1.0, -2 * depth);
        while (child_number < 8) {
            int c;
            do {
                c = node >= n ? A.cchild[8 * (int64_t)(node - n) + child_number] : -1;
                child_number++;
            } while (c == -1 && child_number < 8);
            if (c != -1 && c != (int)k) {
                double cx, cy, cz, cm;
                if (c < n) {
                    const double4 r = A.sxm[c];
                    cx = r.x; cy = r.y; cz = r.z; cm = r.w;
                } else {
                    const int64_t q = c - n;
                    cx = A.cx[q]; cy = A.cy[q]; cz = A.cz[q]; cm = A.cmass[q];
                }
                const double dx = cx - px;
                double dist = dx * dx + A.eps2;
                const double dy = cy - py;
                dist += dy * dy;
                const double dz = cz - pz;
                dist += dz * dz;
                if (c < n || dist * thetasq > cell_size) {
                    dist = sqrt(dist);
                    const double f = cm / (dist * dist * dist);
                    axi += f * dx;
                    ayi += f * dy;
                    azi += f * dz;
                    if (POT) ui -= 0.5 * (cm * own.w) / dist;
                    if (c < n) n_parts++; else n_cells++;
                } else {
                    cur_child[depth][t] = (unsigned char)child_number;
                    cur_node[depth][t] = node;
                    depth++;
                    if (depth >= BH1_DEPTH) {
                        overflow = true;
                        depth = 0;
                        break;
                    }
                    child_number = 0;
                    node = c;
                    cell_size = cs0 * ldexp(1.0, -2 * depth);
                }
            }
        }
        depth--;
    }
    if (overflow) set_err(A.err, ERRF_STACK);
    if (valid) {
        const int32_t i = A.perm[k];
        A.gax[i] = A.G * axi;
        A.gay[i] = A.G * ayi;
        A.gaz[i] = A.G * azi;
        if (POT) A.u[i] = A.G * ui;
    }
    const unsigned long long sc = (unsigned long long)wave_sum(n_cells);
    const unsigned long long sp = (unsigned long long)wave_sum(n_parts);
    if (wave_lane() == 0) {
        atomicAdd(&cnt_slot(A.counters)[CNT_CELL_INT], sc);
        atomicAdd(&cnt_slot(A.counters)[CNT_PART_INT], sp);
    }
}

void launch_gravity(mhpc_ctx *c, bool potential) {
    if (c->n <= 0) return;
    BHArgs a;
    a.n = c->n;
    a.perm = c->perm;
    a.sxm = c->sxm;
    a.clevel = c->clevel;
    a.cchild = c->cchild;
    a.cmass = c->cmass; a.cx = c->cx; a.cy = c->cy; a.cz = c->cz;
    a.cquadf = c->cquadf;
    a.box = c->box;
    a.theta = c->p.theta;
    a.eps2 = c->p.smoothing;
    a.G = c->p.G;
    a.mac = c->p.mac;
    a.gax = c->f[MHPC_F_G_AX]; a.gay = c->f[MHPC_F_G_AY]; a.gaz = c->f[MHPC_F_G_AZ];
    a.u = c->f[MHPC_F_U];
    a.err = c->err;
    a.counters = c->counters;
    if (c->p.bh_accept == MHPC_BH_PER_PARTICLE) {
        const dim3 grid((unsigned)ceil_div(c->n, BH1_THREADS)), block(BH1_THREADS);
        if (potential) hipLaunchKernelGGL(k_bh_v1<true>, grid, block, 0, c->stream, a);
        else hipLaunchKernelGGL(k_bh_v1<false>, grid, block, 0, c->stream, a);
        return;
    }
    const bool fp64 = c->p.precision == MHPC_PREC_FP64;
    const bool quad = c->p.quadrupole != 0;
    if (fp64) {
        if (potential) {
            if (quad) launch_bh<true, true, true>(c, a); else launch_bh<true, true, false>(c, a);
        } else {
            if (quad) launch_bh<true, false, true>(c, a); else launch_bh<true, false, false>(c, a);
        }
    } else {
        if (potential) {
            if (quad) launch_bh<false, true, true>(c, a); else launch_bh<false, true, false>(c, a);
        } else {
            if (quad) launch_bh<false, false, true>(c, a); else launch_bh<false, false, false>(c, a);
        }
    }
}

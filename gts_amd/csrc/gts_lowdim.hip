// Persistent kNN over low-dimensional vector rows (at most 8 dimensions):
// one launch answers a whole batch of queries with uneven k and uneven
// candidate counts.  Each wave takes the next query from a work counter and
// walks the pivot tree depth-first: at a split node it computes the query's
// FP32 distance to the node's pivot, tests the T children's rings in
// parallel (one lane per child; the ring test and FP32 error margin of the
// level-by-level search, nodeProcessKnn's triangle-inequality rule,
// search_v2.cuh:139-148), verifies the objects of the kept leaf children
// (one lane per object, FP64 distances) and pushes the kept inner children on
// a stack in LDS, nearest lower bound on top.  The query's k best (FP64
// distance, ID) pairs live one per lane in ascending order; new candidates
// are bitonic-sorted across the lanes and merged in, and the k-th distance
// (or the learned radius while it is smaller) bounds every later ring test
// and every popped node's stored lower bound.  If fewer than k objects lie
// within the learned radius the query walks again without it.  Results are
// the k smallest FP64 (distance, ID) keys, as the level-by-level search
// returns them.  In two dimensions the walk visits a few dozen nodes per
// query, so the launch is bound by the latency of its dependent loads; many
// waves per CU (blocks of 4 waves, a few blocks per CU chosen by the
// autotuner) hide it.
#include <hip/hip_runtime.h>

#include <climits>

#include "gts_dist.h"
#include "gts_index.h"
#include "gts_lowdim.h"

namespace gts {

namespace {

constexpr int LD_WAVES = 4;     // waves per block
constexpr int LD_SD = 512;      // stack entries per wave (node, stored lower bound)

__device__ __forceinline__ bool key_less(double a, int ia, double b, int ib)
{
    return a < b || (a == b && ia < ib);
}

// One compare-exchange step of a bitonic network across the wave: lanes l and
// l ^ j swap keys when out of order; `up`: l's pair sorts ascending.
__device__ __forceinline__ void cmpx(double &d, int &id, int lane, int j, bool up)
{
    const double od = __shfl_xor(d, j, 64);
    const int oi = __shfl_xor(id, j, 64);
    const bool lower = (lane & j) == 0;
    const bool take = lower == up ? key_less(od, oi, d, id) : key_less(d, id, od, oi);
    if (take) {
        d = od;
        id = oi;
    }
}

__device__ __forceinline__ void sort64(double &d, int &id, int lane)
{
#pragma unroll
    for (int kk = 2; kk <= 64; kk <<= 1)
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1) cmpx(d, id, lane, j, (lane & kk) == 0);
}

// (d, id): the wave's ascending list; (nd, ni): an ascending list of new keys.
// The 64 smallest of both, ascending: the lane-wise minimum of the list and
// the reversed new list is bitonic, and one merge sorts it.
__device__ __forceinline__ void merge64(double &d, int &id, double nd, int ni, int lane)
{
    const double rd = __shfl(nd, 63 - lane, 64);
    const int ri = __shfl(ni, 63 - lane, 64);
    if (key_less(rd, ri, d, id)) {
        d = rd;
        id = ri;
    }
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) cmpx(d, id, lane, j, true);
}

template <int M>
__global__ void __launch_bounds__(LD_WAVES * 64) k_knn_lowdim(DevTree t, int height, const int32_t *__restrict__ qpos,
                                                             const int32_t *__restrict__ kq,
                                                             const float *__restrict__ bnd,
                                                             const int32_t *__restrict__ orig, int64_t nq, int kstride,
                                                             ErrModel em, int32_t *ids, float *dists, float *kth,
                                                             uint32_t *work, unsigned long long *verified)
{
    __shared__ int32_t s_node[LD_WAVES][LD_SD];
    __shared__ float s_lb[LD_WAVES][LD_SD];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int T = t.T, dim = t.dim, dp = t.dp;
    unsigned long long nver = 0;
    for (;;) {
        uint32_t qn = 0;
        if (lane == 0) qn = atomicAdd(work, 1u);
        const int64_t q = (int64_t)__shfl((int)qn, 0, 64);
        if (q >= nq) break;   // every wave of the grid ends here once the counter passes nq
        const int32_t qp = qpos[q];
        const int k = kq[q];
        float qv[LOWDIM_MAX_DIM];
#pragma unroll
        for (int j = 0; j < LOWDIM_MAX_DIM; j++) qv[j] = j < dim ? t.X[(int64_t)qp * dp + j] : 0.0f;
        double b0 = bnd ? (double)bnd[q] : INFINITY;
        if (!(b0 < INFINITY)) b0 = INFINITY;   // NaN or +inf: no radius
        double d = INFINITY;
        int id = INT_MAX;
        for (int round = 0; round < 2; round++) {
            d = INFINITY;
            id = INT_MAX;
            // the current bound: the k-th best distance, or the radius while smaller
            auto bound = [&]() { return fmin(b0, __shfl(d, k - 1, 64)); };
            auto leaf = [&](int32_t lid, int32_t sz) {
                for (int o = 0; o < sz; o += 64) {
                    const double B = bound();
                    double nd = INFINITY;
                    int ni = INT_MAX;
                    if (o + lane < sz) {
                        const int64_t pos = (int64_t)lid + o + lane;
                        const float *x = t.X + pos * dp;
                        // the FP64 distance as the top-k selections form it
                        // (group_dist64_v4: one sum per 4-float group, then
                        // the groups' sums combined), so both return the
                        // same keys
                        double lo = 0.0, hi = 0.0;
                        for (int j = 0; j < dim && j < 4; j++) lo = Acc<M>::step64(lo, qv[j], x[j]);
                        for (int j = 4; j < dim; j++) hi = Acc<M>::step64(hi, qv[j], x[j]);
                        const double v = Acc<M>::fin64(dp > 4 ? Acc<M>::comb64(lo, hi) : lo);
                        nver++;
                        if (v <= B) {
                            nd = v;
                            ni = t.perm[pos];
                        }
                    }
                    if (!__builtin_amdgcn_ballot_w64(ni != INT_MAX)) continue;
                    sort64(nd, ni, lane);
                    merge64(d, id, nd, ni, lane);
                }
            };
            int top = 0;
            if (height <= 1) {
                leaf(t.nlid[0], t.nsize[0]);
            } else {
                if (lane == 0) {
                    s_node[w][0] = 0;
                    s_lb[w][0] = -INFINITY;
                }
                top = 1;
            }
            while (top > 0) {
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                top--;
                const int32_t P = s_node[w][top];
                const float lbp = s_lb[w][top];
                const float Bf = __double2float_ru(bound());
                // the stored bound against the current radius (the ring test's
                // margin, rewritten for a smaller bound: see below)
                if (lbp > Bf * (1.0f + 2.0f * em.rel)) continue;
                const float *pv = t.X + (int64_t)t.npiv[P] * dp;
                float a32 = 0.0f;
                for (int j = 0; j < dim; j++) a32 = Acc<M>::step(a32, qv[j], pv[j]);
                const float dq = Acc<M>::fin(a32);
                const uint32_t ne = t.cne[P], lf = t.cleaf[P];
                bool kept = false;
                float lbc = 0.0f;
                if (lane < T) {
                    // child `lane` holds the objects whose pivot keys lie in
                    // [ring[lane], ring[lane + 1]] (the last child: no upper end)
                    const float *ring = t.ring + (int64_t)P * (T + 1);
                    const float lo = ring[lane];
                    float lb = lo - dq, hv = 0.0f;
                    if (lane < T - 1) {
                        hv = ring[lane + 1];
                        lb = fmaxf(lb, dq - hv);
                    }
                    const float bm = isinf(Bf) ? 0.0f : Bf;
                    const float margin = 2.0f * em.rel * (dq + lo + hv + bm) + 2.0f * em.abs;
                    kept = ((ne >> lane) & 1u) && !(lb > Bf + margin);
                    // lb > B' + 2 rel (dq + lo + hv + B') + 2 abs, for a later B',
                    // is lbc > B' (1 + 2 rel)
                    lbc = lb - 2.0f * em.rel * (dq + lo + hv) - 2.0f * em.abs;
                }
                const uint32_t km = (uint32_t)__builtin_amdgcn_ballot_w64(kept);
                for (uint32_t lk = km & lf; lk; lk &= lk - 1) {
                    const int32_t L = P * T + 1 + __builtin_ctz(lk);
                    leaf(t.nlid[L], t.nsize[L]);
                }
                const uint32_t ik = km & ~lf;
                if (ik) {
                    // the kept inner children, the largest stored bound deepest
                    // (ties: the higher child first), so the nearest pops first
                    int rank = 0;
                    for (int o = 0; o < T; o++) {
                        const float v = __shfl(lbc, o, 64);
                        rank += ((ik >> o) & 1u) && (v > lbc || (v == lbc && o > lane));
                    }
                    if (lane < T && ((ik >> lane) & 1u)) {
                        s_node[w][top + rank] = P * T + 1 + lane;
                        s_lb[w][top + rank] = lbc;
                    }
                    top += __popc(ik);
                }
            }
            const int found = __popcll(__builtin_amdgcn_ballot_w64(id != INT_MAX));
            if (found >= k || !(b0 < INFINITY)) break;
            b0 = INFINITY;   // fewer than k objects within the radius: walk again without it
        }
        const int32_t row = orig[q];
        if (row >= 0) {
            const bool ok = lane < k && id != INT_MAX;
            for (int i = lane; i < kstride; i += 64) {
                ids[(int64_t)row * kstride + i] = (i == lane && ok) ? id : -1;
                dists[(int64_t)row * kstride + i] = (i == lane && ok) ? (float)d : INFINITY;
            }
            const double dk = __shfl(d, k - 1, 64);
            const int ik = __shfl(id, k - 1, 64);
            if (lane == 0) kth[row] = ik != INT_MAX ? (float)dk : INFINITY;
        }
        __builtin_amdgcn_wave_barrier();   // the wave's stack is reused by its next query
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nver += __shfl_xor(nver, o, 64);
    if (lane == 0 && nver) atomicAdd(verified, nver);
}

}  // namespace

bool lowdim_supported(const gts_index *I, int kstride)
{
    if (I->metric != GTS_METRIC_L2 && I->metric != GTS_METRIC_L1 && I->metric != GTS_METRIC_LINF) return false;
    if (I->dp > LOWDIM_MAX_DIM || kstride > LOWDIM_MAX_K || I->T > 32) return false;
    // the depth-first stack: at most T - 1 waiting siblings per level, and the node being expanded
    return (int64_t)I->height * (I->T - 1) + 1 <= LD_SD;
}

void lowdim_launch(gts_index *I, const int32_t *qpos, const int32_t *kq, const float *bnd, const int32_t *orig,
                   int64_t nq, int kstride, int32_t *ids, float *dists, float *kth, uint32_t *work,
                   unsigned long long *verified, int bpc)
{
    int ncu = 0;
    GTS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, I->device));
    const int64_t want = (int64_t)std::max(ncu, 1) * std::max(bpc, 1);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (nq + LD_WAVES - 1) / LD_WAVES));
    const DevTree t = I->dev();
    switch (I->metric) {
    case GTS_METRIC_L2:
        k_knn_lowdim<GTS_METRIC_L2><<<grid, LD_WAVES * 64, 0, I->st>>>(t, I->height, qpos, kq, bnd, orig, nq, kstride,
                                                                      I->em, ids, dists, kth, work, verified);
        break;
    case GTS_METRIC_L1:
        k_knn_lowdim<GTS_METRIC_L1><<<grid, LD_WAVES * 64, 0, I->st>>>(t, I->height, qpos, kq, bnd, orig, nq, kstride,
                                                                      I->em, ids, dists, kth, work, verified);
        break;
    case GTS_METRIC_LINF:
        k_knn_lowdim<GTS_METRIC_LINF><<<grid, LD_WAVES * 64, 0, I->st>>>(t, I->height, qpos, kq, bnd, orig, nq,
                                                                        kstride, I->em, ids, dists, kth, work,
                                                                        verified);
        break;
    default: fail(GTS_E_ARG, "the persistent kNN launch covers L2, L1 and Linf rows");
    }
    GTS_HIP(hipGetLastError());
}

}  // namespace gts

// Range and kNN drivers: per-batch setup, reruns, result assembly.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>

#include "gts_cells.h"
#include "gts_api_internal.h"
#include "gts_dist.h"
#include "gts_lowdim.h"
#include "gts_search.h"
#include "gts_verify2.h"

namespace gts {

namespace {

unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

// Query IDs outside [0, n) and k outside [1, kcap] (device inputs of
// gts_knn_query_dev) set *argerr, which the pass's publish reports, and are
// clamped so that no kernel of the call reads or writes out of bounds.
__global__ void k_prep(const int32_t *inv, const int32_t *qids, const int32_t *ks, int kfix,
                       int64_t nq, float b0, int32_t *qpos, int32_t *kq, int32_t *orig, float *bnd,
                       int64_t n, int kcap, int32_t *argerr)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
         i += (int64_t)gridDim.x * blockDim.x) {
        int32_t q = qids[i], k = ks ? ks[i] : kfix;
        if (q < 0 || q >= n || (kq && (k < 1 || k > kcap))) {
            atomicOr(argerr, 1);
            q = min(max(q, 0), (int32_t)(n - 1));
            k = min(max(k, 1), kcap);
        }
        qpos[i] = inv[q];
        if (kq) kq[i] = k;
        if (orig) orig[i] = (int32_t)i;
        bnd[i] = b0;
    }
}

// The device flag of the kNN argument check (k_prep, k_qplace_fields),
// allocated and cleared on first use, outside graph capture.
static int32_t *arg_flag(gts_index *I)
{
    if (!I->argerr.p) {
        int32_t *p = I->argerr.ensure(1);
        GTS_HIP(hipMemsetAsync(p, 0, I->argerr.cap * sizeof(int32_t), I->st));
    }
    return I->argerr.p;
}

// A pass's six scalars (src = scal + 9: verified pairs, range results, live
// chunks, failures, candidates, speculative rerun) and the reruns' live
// chunks (src[-1] = scal[8]) copied into the coherent pinned host block
// (hpin[0..5], hpin[7]), with the argument flag (hpin[6]), then the
// incremented publish counter (scal[15]; system-scope
// release), which the host polls: no copy engine and no driver wait between
// a pass and the host's decision.  The counter lives on the device, so a
// graph can end with this kernel.
__global__ void k_publish(int64_t *src, int64_t *dst, int32_t *argerr)
{
    if (threadIdx.x == 0) {
        for (int i = 0; i < 6; i++) dst[i] = src[i];
        dst[7] = src[-1];
        // the argument check's flag (then cleared for the next call)
        dst[6] = *argerr;
        *argerr = 0;
        const int64_t seq = src[6] + 1;
        src[6] = seq;
        __threadfence_system();
        __hip_atomic_store(dst + 15, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Waits until the stream's next k_publish is visible (the counter in hpin[15]
// differs from I->pub_seen), and after a second falls back to the driver wait
// (which reports a failed launch).
static void await_publish(gts_index *I)
{
    volatile int64_t *flag = I->hpin + 15;
    const int64_t seen = I->pub_seq;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t it = 0; *flag == seen; it++) {
        if ((it & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
            GTS_HIP(hipStreamSynchronize(I->st));
            if (*flag == seen) fail(GTS_E_IO, "pass scalars not published");
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    I->pub_seq = *flag;
}

// The pass's end: the call's end event (complete once the host sees the
// publish), then the publish.
static void end_pass(gts_index *I, bool external)
{
    hipStream_t st = I->st;
    if (I->timing)
        GTS_HIP(external ? hipEventRecordWithFlags(I->call_ev[1], st, hipEventRecordExternal)
                         : hipEventRecord(I->call_ev[1], st));
    k_publish<<<1, 64, 0, st>>>(I->scal.ensure(16) + 9, I->hpin, arg_flag(I));
    GTS_HIP(hipGetLastError());
}

// Rerun batch: the failed queries, in fail[] order; status 2 (fewer than k
// results within the radius) doubles the radius, from the 7th attempt on (or
// for a non-finite radius) the radius is unbounded (exact fallback).
__global__ void k_retry_gather(const int32_t *fail, int64_t nf, const int32_t *status, int attempt,
                               const int32_t *qpos, const int32_t *kq, const int32_t *orig, const float *bnd,
                               int32_t *t, float *tf)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t j = fail[i];
        t[i] = qpos[j];
        t[nf + i] = kq[j];
        t[2 * nf + i] = orig[j];
        float b = bnd[j];
        if (status[j] == 2) b = (attempt >= 6 || !isfinite(b)) ? INFINITY : (b > 0.0f ? 2.0f * b : 1e-30f);
        tf[i] = b;
    }
}

// Padded rerun batch of the cell scan in one workgroup (npad <= blockDim):
// the failed queries of the previous pass (fail[], count *nfail; fields
// sqpos..sbnd, status) as k_retry_gather takes them, then inert queries up to
// npad (a valid position, k = 1, output row -1, radius NaN); all ordered by
// the scan row of the query (rank of the (scan row, slot) key) into qpos..bnd
// (which may alias the previous pass's fields: every read precedes the first
// write); candidate counts and nbits words of the live-chunk bitmap cleared.
// Ordinary rerun (spec null): the pass scalars cleared.  Speculative rerun
// (spec: the first pass graph's): only with 1 <= *nfail <= npad; then *spec =
// *nfail and only the failure count is cleared (the pass totals accumulate);
// otherwise every slot is inert and *spec = 0, *nfail and fail[] untouched.
__global__ void __launch_bounds__(256) k_rerun_prepare(const int32_t *fail, unsigned long long *nfail, int npad,
                                                       const int32_t *status, int attempt, const int32_t *pos_scan,
                                                       const int32_t *sqpos, const int32_t *skq,
                                                       const int32_t *sorig, const float *sbnd, int32_t *qpos,
                                                       int32_t *kq, int32_t *orig, float *bnd, int32_t *ccount,
                                                       int64_t *scal8, uint32_t *bits, int64_t nbits, int64_t *spec)
{
    __shared__ int32_t s_qp[256], s_k[256], s_o[256];
    __shared__ float s_b[256];
    __shared__ uint64_t s_key[256];
    const int i = threadIdx.x;
    const unsigned long long nf0 = *nfail;
    const bool run = !spec || (nf0 >= 1 && nf0 <= (unsigned long long)npad);
    const int nf = run ? (int)nf0 : 0;
    if (i < npad) {
        if (i < nf) {
            const int32_t j = fail[i];
            s_qp[i] = sqpos[j];
            s_k[i] = skq[j];
            s_o[i] = sorig[j];
            float b = sbnd[j];
            if (status[j] == 2) b = (attempt >= 6 || !isfinite(b)) ? INFINITY : (b > 0.0f ? 2.0f * b : 1e-30f);
            s_b[i] = b;
        } else {
            s_qp[i] = nf > 0 ? sqpos[fail[0]] : 0;
            s_k[i] = 1;
            s_o[i] = -1;
            s_b[i] = __int_as_float(0x7fc00000);   // NaN
        }
        s_key[i] = ((uint64_t)(uint32_t)pos_scan[s_qp[i]] << 32) | (uint32_t)i;
    }
    __syncthreads();   // every read of the previous batch is done
    if (i < npad) {
        int rank = 0;
        for (int u = 0; u < npad; u++) rank += s_key[u] < s_key[i];
        qpos[rank] = s_qp[i];
        kq[rank] = s_k[i];
        orig[rank] = s_o[i];
        bnd[rank] = s_b[i];
        ccount[i] = 0;
    }
    if (!spec) {
        // reruns' live chunks, verified pairs, range results, live chunks,
        // failures, candidates, speculative rerun
        if (i < 7) scal8[i] = 0;
    } else if (i == 0) {
        *spec = nf;
        if (run) *nfail = 0;
    }
    for (int64_t w = i; w < nbits; w += blockDim.x) bits[w] = 0u;
}

__global__ void k_retry_scatter(int64_t nf, const int32_t *t, const float *tf, int32_t *qpos, int32_t *kq,
                                int32_t *orig, float *bnd)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += (int64_t)gridDim.x * blockDim.x) {
        qpos[i] = t[i];
        kq[i] = t[nf + i];
        orig[i] = t[2 * nf + i];
        bnd[i] = tf[i];
    }
}

__global__ void k_gather_f32(const int32_t *idx, const float *src, float *dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

template <class T>
T *scratch(DBuf<T> &b, int64_t n) { return b.ensure(n > 0 ? n : 1); }

// FP64 distances of the reported range results (query << 32 | ID), one wave
// per result: the angular metric's FP32 angle is far from 1e-5 relative near
// 0 degrees, so its reported distances come from FP64 (the identical object
// is at distance 0, search_v2.cuh:338-340).
template <int M>
__global__ void k_result_dist64(DevTree t, const int32_t *inv, const int32_t *qpos, const uint64_t *res_k,
                                float *res_d, int64_t nres)
{
    const int lane = threadIdx.x & 63;
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t i = w0; i < nres; i += nw) {
        const int64_t qp = qpos[(int64_t)(res_k[i] >> 32)];
        const int64_t xp = inv[(uint32_t)res_k[i]];
        const double d = xp == qp ? 0.0 : group_dist64<M, 64>(t.X + qp * t.dp, t.X + xp * t.dp, t.dim, lane);
        if (lane == 0) res_d[i] = (float)d;
    }
}

// strided sample of ns of the nq batch queries
__global__ void k_sample(const int32_t *qpos, const float *bnd, const int32_t *kq, int64_t nq, int64_t ns,
                         int32_t *sq, float *sb, int32_t *sk)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i * nq / ns;
        sq[i] = qpos[j];
        sb[i] = bnd[j];
        sk[i] = kq[j];
    }
}

}  // namespace

void release_scratch(gts_index *I)
{
    for (int i = 0; i < 2; i++) { I->fr_q[i].release(); I->fr_n[i].release(); }
    I->fr_dq.release(); I->cnt_int.release(); I->qcnt.release(); I->leafmask.release();
    I->off_int.release(); I->off_leaf.release(); I->qoff.release(); I->rec_q.release();
    I->rec_n.release(); I->rec_m.release(); I->grp_q.release(); I->grp_n.release();
    I->grp_m.release(); I->node_cnt.release(); I->node_off.release(); I->tile_off.release();
    I->rs_k0.release(); I->rs_k1.release(); I->rs_v0.release(); I->rs_v1.release();
    I->tmp64.release(); I->qpos.release(); I->bnd.release(); I->kq.release(); I->cand.release();
    I->ccount.release(); I->band_q.release(); I->band_p.release(); I->counts.release();
    I->res_k.release(); I->res_k2.release(); I->res_d.release(); I->res_v.release();
    I->res_v2.release(); I->status.release(); I->scal.release(); I->dqid.release();
    I->dks.release(); I->dids.release(); I->ddist.release(); I->dkth.release();
    I->dfeat.release(); I->dpred.release(); I->orig.release(); I->death.release();
    I->dense.release();
    I->par_of.release();
    I->plan_q.release(); I->plan_k.release(); I->plan_b.release(); I->qimg.release(); I->ftab.release();
    I->fail.release(); I->cell_bits.release(); I->cell_list.release(); I->cell_cnt.release();
    I->cell_pcnt.release();
    I->argerr.release();
    I->ex_k.release();
    I->ex_v.release();
    I->rp_piv.release(); I->rp_off.release(); I->ld_work.release(); I->tile_slot.release();
    I->rp_n = -1;
    I->qperm.release(); I->qrank.release(); I->qtmp.release(); I->qtmp_f.release();
    I->qrank_zero = nullptr;
    I->rq_i.release(); I->rq_b.release(); I->rq_cand.release();
    if (I->model.w) (void)hipFree(I->model.w);
    if (I->model.wb) (void)hipFree(I->model.wb);
    if (I->model.piv) (void)hipFree(I->model.piv);
    if (I->model.mu) (void)hipFree(I->model.mu);
    for (void *p : {(void *)I->model.phi, (void *)I->model.pn2})
        if (p) (void)hipFree(p);
    I->model = LearnedModel{};
}

static void add_stats(gts_stats &a, const gts_stats &b)
{
    a.frontier_pairs += b.frontier_pairs;
    a.leaf_records += b.leaf_records;
    a.verified_pairs += b.verified_pairs;
    a.candidates += b.candidates;
    a.band_pairs += b.band_pairs;
    a.retries += b.retries;
    a.ms_verify += b.ms_verify;
    a.verify_launches += b.verify_launches;
    a.exhaustive_passes += b.exhaustive_passes;
    a.graph_replays += b.graph_replays;
    a.live_chunks += b.live_chunks;
    a.ms_verify_batch += b.ms_verify_batch;
    a.verify_launches_batch += b.verify_launches_batch;
    a.live_chunks_batch += b.live_chunks_batch;
}

static void merge_call_stats(gts_stats &a, const gts_stats &b)
{
    add_stats(a, b);
    a.ms_total += b.ms_total;
    a.ms_cells_build += b.ms_cells_build;
    a.dense_rows = std::max(a.dense_rows, b.dense_rows);
    a.dense_persistent = std::max(a.dense_persistent, b.dense_persistent);
    a.cell_scan = std::max(a.cell_scan, b.cell_scan);
}

// Memory budget of a query call: the index's setting, or half of the free
// device memory, the share searchIndexRnnV2 / searchIndexKnnV2 allocate
// (search_v2.cuh:791-805, 984-1001).
static int64_t call_budget(gts_index *I)
{
    if (I->mem_budget > 0) return I->mem_budget;
    size_t fr = 0, tot = 0;
    GTS_HIP(hipMemGetInfo(&fr, &tot));
    return std::max<int64_t>((int64_t)(fr / 2), 1);
}

// The batch splitter of every query call: run(q0, n) searches queries
// [q0, q0 + n) of the call's batch.  The batch first goes in sub-batches whose
// per-query buffers (per_query bytes each, at least min_slots of them) fit the
// budget; a sub-batch whose search would exceed it (BudgetExceeded, raised
// before any of its results are written) is halved and its halves run first
// half first, depth-first from a stack of query ranges, as the reference
// splits its query range by level (search_v2.cuh:818-857, 921-966,
// 1177-1221).  Results do not depend on the split.  One sub-batch of the
// whole batch, the case without a split, runs with no extra work; otherwise
// the sub-batches' statistics are summed into the call's (budget_batches
// counts them).
template <class F>
static void budget_split(gts_index *I, int64_t nq, int64_t per_query, int64_t min_slots, F run)
{
    struct Guard {
        gts_index *I;
        ~Guard() { I->budget_limit = I->budget_fixed = 0; }
    } guard{I};
    if (nq <= 0) {
        I->stats = gts_stats{};
        return;
    }
    const int64_t B = call_budget(I);
    auto fixed = [&](int64_t n) { return std::max(n, min_slots) * per_query; };
    int64_t first = nq;
    while (first > 1 && fixed(first) > B) first = (first + 1) / 2;
    I->budget_limit = B;
    if (first == nq && fixed(nq) <= B) {
        try {
            I->budget_fixed = fixed(nq);
            run(0, nq);
            I->stats.budget_batches = 1;
            return;
        } catch (const BudgetExceeded &) {
            first = (nq + 1) / 2;
        }
    }
    gts_stats total{};
    int64_t batches = 0;
    std::vector<std::pair<int64_t, int64_t>> stack;
    for (int64_t a = ((nq - 1) / first) * first; a >= 0; a -= first) stack.push_back({a, std::min(first, nq - a)});
    while (!stack.empty()) {
        const auto [a, n] = stack.back();
        stack.pop_back();
        try {
            if (fixed(n) > B) throw BudgetExceeded{};
            I->budget_fixed = fixed(n);
            run(a, n);
        } catch (const BudgetExceeded &) {
            if (n == 1) fail(GTS_E_OOM, "the memory budget (" + std::to_string(B) + " bytes) holds no one query's search");
            stack.push_back({a + n / 2, n - n / 2});
            stack.push_back({a, n / 2});
            continue;
        }
        finalize_timing(I);
        merge_call_stats(total, I->stats);
        batches++;
    }
    total.budget_batches = batches;
    I->stats = total;
}

// Range search over device query positions; returns counts (and, if
// requested, the unsorted (query, ID) results in res_k/res_d with *nres).
static void range_core(gts_index *I, const int32_t *qpos, float *bnd, int64_t nq, bool want_ids,
                       int64_t *nres, gts_stats &stats)
{
    hipStream_t st = I->st;
    Search S;
    S.idx = I;
    S.nq = nq;
    S.qpos = qpos;
    S.bnd = bnd;
    S.kq = nullptr;
    S.expand(false);
    int64_t *scal = I->scal.ensure(16);
    // first guesses of the band and result buffers (8 B a band pair; 28 B a
    // result: key, distance, sort copies), within the memory budget's room;
    // growth beyond it splits the batch
    const int64_t room = budget_room(I);
    int64_t band_cap = std::max<int64_t>(I->band_q.cap,
                                         std::min(std::max<int64_t>(1 << 16, nq * 16), std::max<int64_t>(1024, room / 32)));
    int64_t res_cap = want_ids ? std::max<int64_t>(I->res_k.cap, std::min(std::max<int64_t>(1 << 20, nq * 64),
                                                                          std::max<int64_t>(1024, room / 56)))
                               : 0;
    for (int pass = 0; pass < 3; pass++) {
        VerifyArgs va;
        va.counts = scratch(I->counts, nq);
        va.band_q = scratch(I->band_q, band_cap);
        va.band_p = scratch(I->band_p, band_cap);
        va.band_cap = band_cap;
        va.band_n = scal + 8;
        va.verified = scal + 9;
        va.res_n = scal + 10;
        if (want_ids) {
            va.res_k = scratch(I->res_k, res_cap);
            va.res_d = scratch(I->res_d, res_cap);
            va.res_cap = res_cap;
        }
        GTS_HIP(hipMemsetAsync(va.counts, 0, nq * sizeof(int64_t), st));
        GTS_HIP(hipMemsetAsync(scal + 8, 0, 3 * sizeof(int64_t), st));
        const int mode = want_ids ? 1 : 0;
        S.verify(mode, va);
        int64_t h[3];
        GTS_HIP(hipMemcpyAsync(h, scal + 8, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
        if (h[0] > band_cap) {
            budget_check(I, (h[0] * 8 + res_cap * 28) * 5 / 4);
            band_cap = h[0];
            continue;
        }
        S.band(mode, va, h[0]);
        GTS_HIP(hipMemcpyAsync(h + 2, scal + 10, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
        if (want_ids && h[2] > res_cap) {
            budget_check(I, (band_cap * 8 + h[2] * 28) * 5 / 4);
            res_cap = h[2];
            continue;
        }
        S.stats.verified_pairs += h[1];
        S.stats.band_pairs += h[0];
        *nres = h[2];
        if (want_ids && h[2] > 0 && I->metric == GTS_METRIC_ANGULAR) {
            const unsigned g = (unsigned)std::min<int64_t>((h[2] + 3) / 4, 65536);
            k_result_dist64<GTS_METRIC_ANGULAR><<<g, 256, 0, st>>>(I->dev(), I->inv, qpos, va.res_k, va.res_d, h[2]);
            GTS_HIP(hipGetLastError());
        }
        add_stats(stats, S.stats);
        return;
    }
    fail(GTS_E_CAPACITY, "range result buffers did not converge");
}

static void range_run_batch(gts_index *I, const int32_t *qids_host, int64_t nq, float r, bool want_ids,
                            RangeOut *out)
{
    hipStream_t st = I->st;
    // the call's events, released on every path (a split sub-batch throws)
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    GTS_HIP(hipEventCreate(&ev.e[0]));
    GTS_HIP(hipEventCreate(&ev.e[1]));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    I->stats = gts_stats{};
    I->pend_timing = 0;   // an earlier kNN call's event timings are not this call's
    int32_t *dq = scratch(I->dqid, nq);
    GTS_HIP(hipMemcpyAsync(dq, qids_host, nq * 4, hipMemcpyHostToDevice, st));
    GTS_HIP(hipEventRecord(e0, st));
    int32_t *qpos = scratch(I->qpos, nq);
    float *bnd = scratch(I->bnd, nq);
    k_prep<<<grid_for(nq, 256), 256, 0, st>>>(I->inv, dq, nullptr, 0, nq, r, qpos, nullptr, nullptr, bnd, I->n,
                                              INT_MAX, arg_flag(I));
    GTS_HIP(hipGetLastError());
    int64_t nres = 0;
    range_core(I, qpos, bnd, nq, want_ids, &nres, I->stats);
    out->counts.resize(nq);
    GTS_HIP(hipMemcpyAsync(out->counts.data(), I->counts.p, nq * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    if (want_ids && nres > 0) {
        // sort results by (query, ID)
        uint64_t *k0 = I->res_k.p, *k1 = scratch(I->res_k2, nres);
        int32_t *v0 = scratch(I->res_v, nres), *v1 = scratch(I->res_v2, nres);
        fill_iota_i32(v0, nres, st);
        int qb = 1;
        while (qb < 62 && ((int64_t)1 << qb) < nq) qb++;
        uint64_t *ks;
        int32_t *vs;
        radix_sort_u64(k0, v0, k1, v1, nres, 32 + qb, nullptr, 0, st, &ks, &vs);
        std::vector<uint64_t> hk(nres);
        std::vector<int32_t> hv(nres);
        std::vector<float> hd(nres);
        GTS_HIP(hipMemcpyAsync(hk.data(), ks, nres * 8, hipMemcpyDeviceToHost, st));
        GTS_HIP(hipMemcpyAsync(hv.data(), vs, nres * 4, hipMemcpyDeviceToHost, st));
        GTS_HIP(hipMemcpyAsync(hd.data(), I->res_d.p, nres * 4, hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
        out->ids.resize(nres);
        out->dists.resize(nres);
        for (int64_t i = 0; i < nres; i++) {
            out->ids[i] = (int32_t)(uint32_t)hk[i];
            out->dists[i] = hd[hv[i]];
        }
    }
    GTS_HIP(hipEventRecord(e1, st));
    GTS_HIP(hipStreamSynchronize(st));
    float ms = 0;
    GTS_HIP(hipEventElapsedTime(&ms, e0, e1));
    I->stats.ms_total = ms;
}

// per-query buffers of a range batch: counts, positions, radii, the children
// count and offsets, the device query IDs
constexpr int64_t RANGE_QUERY_BYTES = 40;

void range_run(gts_index *I, const int32_t *qids_host, int64_t nq, float r, bool want_ids, RangeOut *out)
{
    out->counts.clear();
    out->ids.clear();
    out->dists.clear();
    budget_split(I, nq, RANGE_QUERY_BYTES, 1, [&](int64_t q0, int64_t n) {
        RangeOut part;
        range_run_batch(I, qids_host + q0, n, r, want_ids, &part);
        out->counts.insert(out->counts.end(), part.counts.begin(), part.counts.end());
        out->ids.insert(out->ids.end(), part.ids.begin(), part.ids.end());
        out->dists.insert(out->dists.end(), part.dists.begin(), part.dists.end());
    });
}

static void range_count_batch(gts_index *I, const int32_t *d_qpos, const int32_t *d_qop, int64_t nq, float r,
                              const int32_t *d_death, int64_t *h_counts)
{
    hipStream_t st = I->st;
    I->stats = gts_stats{};
    I->pend_timing = 0;
    float *bnd = scratch(I->bnd, nq);
    fill_f32(bnd, r, nq, st);
    I->death_ptr = d_death;
    I->qop_ptr = d_qop;
    int64_t nres = 0;
    try {
        range_core(I, d_qpos, bnd, nq, false, &nres, I->stats);
    } catch (...) {
        I->death_ptr = nullptr;
        I->qop_ptr = nullptr;
        throw;
    }
    I->death_ptr = nullptr;
    I->qop_ptr = nullptr;
    GTS_HIP(hipMemcpyAsync(h_counts, I->counts.p, nq * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
}

void range_count_dev(gts_index *I, const int32_t *d_qpos, const int32_t *d_qop, int64_t nq, float r,
                     const int32_t *d_death, int64_t *h_counts)
{
    const gts_stats before = I->stats;   // the update call's totals so far
    budget_split(I, nq, RANGE_QUERY_BYTES, 1, [&](int64_t q0, int64_t n) {
        range_count_batch(I, d_qpos + q0, d_qop ? d_qop + q0 : nullptr, n, r, d_death, h_counts + q0);
    });
    gts_stats total = before;
    merge_call_stats(total, I->stats);
    total.budget_batches = before.budget_batches + I->stats.budget_batches;
    I->stats = total;
}

// Plan of a learned-radius kNN batch in the BF16 MFMA form.  A strided sample
// of 256 queries is expanded through the tree; when its surviving leaf records
// sit on one level and cover at least half of the (query, parent) pairs, the
// batch would take the dense kernel over every row anyway and the pruning
// removes no verification work, so the batch skips it (exhaustive plan).  The
// decision is kept for the next batches of the same shape (batch size rounded
// to a power of two, k) and resampled every PLAN_REUSE batches (a rebuild or
// retraining clears it); the plan never changes results, only the work done.
// A resampled batch runs eagerly (no graph replay, ~1 ms more at 10k queries),
// so the decision is reused for 256 batches, not 16.
constexpr int PLAN_REUSE = 256;
static bool plan_exhaustive(gts_index *I, const int32_t *qpos, const float *bnd, const int32_t *kq,
                            int64_t nq, uint64_t shape)
{
    constexpr int64_t NS = 256;   // plan_known_exhaustive repeats the size test
    if (I->knn_plan == GTS_PLAN_EXHAUSTIVE || I->knn_plan == GTS_PLAN_CELLS) {
        if (!v2_supported(I, 2)) fail(GTS_E_ARG, "the exhaustive plan needs L2, BF16X3 and dimension <= 512");
        return true;
    }
    if (I->knn_plan == GTS_PLAN_PRUNED) return false;
    if (!v2_supported(I, 2) || I->height <= 1 || nq < 4 * NS) return false;
    auto memo = I->plan_memo.find(shape);
    if (memo != I->plan_memo.end() && ++memo->second.second < PLAN_REUSE) return memo->second.first;
    hipStream_t st = I->st;
    int32_t *sq = scratch(I->plan_q, NS), *sk = scratch(I->plan_k, NS);
    float *sb = scratch(I->plan_b, NS);
    k_sample<<<grid_for(NS, 256), 256, 0, st>>>(qpos, bnd, kq, nq, NS, sq, sb, sk);
    GTS_HIP(hipGetLastError());
    Search S;
    S.idx = I;
    S.nq = NS;
    S.qpos = sq;
    S.bnd = sb;
    S.kq = sk;
    S.expand(true);
    I->stats.frontier_pairs += S.stats.frontier_pairs;
    const int64_t nP = (int64_t)S.rec_pmax - S.rec_pmin + 1;
    const bool ex = S.rec_levels == 1 && S.rec_pmin >= 0 && S.nrec * 2 >= NS * nP;
    I->plan_memo[shape] = {ex, 0};
    return ex;
}

// True when plan_exhaustive would choose the exhaustive plan for this batch
// without sampling: a forced plan, or a reusable automatic decision, whose
// use is counted when `consume` is set.
static bool plan_known_exhaustive(gts_index *I, int64_t nq, uint64_t shape, bool consume)
{
    if (I->knn_plan == GTS_PLAN_EXHAUSTIVE || I->knn_plan == GTS_PLAN_CELLS) return v2_supported(I, 2);
    if (I->knn_plan == GTS_PLAN_PRUNED || !v2_supported(I, 2) || I->height <= 1 || nq < 4 * 256) return false;
    auto memo = I->plan_memo.find(shape);
    if (memo == I->plan_memo.end() || !memo->second.first || memo->second.second + 1 >= PLAN_REUSE) return false;
    if (consume) memo->second.second++;
    return true;
}

// The exhaustive plan verifies through the cell scan unless EXHAUSTIVE is
// forced; the partition is built on first use.
static bool use_cells(gts_index *I)
{
    if (I->knn_plan != GTS_PLAN_EXHAUSTIVE && cells_supported(I)) {
        cells_build(I);
        return true;
    }
    if (I->knn_plan == GTS_PLAN_CELLS)
        fail(GTS_E_ARG, "the cell scan needs L2, BF16X3, no pending deletions and dimension <= 512");
    return false;
}

static void mix_key(void *ctx, uint64_t v)
{
    uint64_t &k = *(uint64_t *)ctx;
    k = (k ^ v) * 1099511628211ull;
}

constexpr int64_t QPAD = 64;                // rerun batches are padded to whole query blocks
constexpr int64_t MAX_PADDED_RERUN = 256;   // larger reruns gather on the host's count

// Rerun pass of the cell scan for at most npad queries, issued without host
// synchronisation: the failed queries of the previous pass (fail[], count on
// the device) gathered with doubled radii and padded to npad with inert
// queries (output row -1, radius NaN: no live chunk, no candidate, no
// output), ordered, verified and ranked; failures listed again.  Speculative
// (the tail of the first pass graph): into the rq_* buffers, leaving the
// first pass's fields intact, and only when 1..npad queries failed (scal[14]
// = how many were rerun, else 0 and the pass does no work).
static void rerun_pass(gts_index *I, int64_t npad, int attempt, int kstride, int32_t *d_ids, float *d_dists,
                       float *d_kth, bool spec = false)
{
    hipStream_t st = I->st;
    const int cap = topk_cap_limit();
    int32_t *qpos = scratch(I->qpos, npad), *kq = scratch(I->kq, npad), *orig = scratch(I->orig, npad);
    float *bnd = scratch(I->bnd, npad);
    int32_t *failq = scratch(I->fail, npad), *status = scratch(I->status, npad);
    int64_t *scal = I->scal.ensure(16);
    VerifyArgs va;
    va.cand = scratch(I->cand, npad * (int64_t)cap);
    va.ccount = scratch(I->ccount, npad);
    va.cap = cap;
    va.verified = scal + 9;
    va.live = scal + 8;   // the reruns' own count
    va.ncand = scal + 13;
    int32_t *dq = qpos, *dk = kq, *dorig = orig, *dstatus = status;
    float *db = bnd;
    if (spec) {
        int32_t *r = I->rq_i.ensure(5 * QPAD);
        dq = r;
        dk = r + QPAD;
        dorig = r + 2 * QPAD;
        dstatus = r + 3 * QPAD;
        va.ccount = r + 4 * QPAD;
        db = I->rq_b.ensure(QPAD);
        va.cand = I->rq_cand.ensure(QPAD * (int64_t)cap);
    }
    int64_t nbits = 0;
    uint32_t *bits = cells_bitmap(I, npad, &nbits);
    k_rerun_prepare<<<1, 256, 0, st>>>(failq, (unsigned long long *)(scal + 12), (int)npad, status, attempt,
                                       I->cells.pos_scan, qpos, kq, orig, bnd, dq, dk, dorig, db, va.ccount, scal + 8,
                                       bits, nbits, spec ? scal + 14 : nullptr);
    GTS_HIP(hipGetLastError());
    Search S;
    S.idx = I;
    S.nq = npad;
    S.qpos = dq;
    S.bnd = db;
    S.kq = dk;
    cells_launch(I, S, va, I->timing ? I->timing_ev[0] : nullptr, I->timing ? I->timing_ev[1] : nullptr, true,
                 CELLS_CLEARED, true);
    if (!spec && I->timing) I->timing_pending = true;   // speculative: timed only when it ran (scal[14] > 0)
    S.topk(va, 1, dorig, kstride, d_ids, d_dists, d_kth, dstatus, failq, (unsigned long long *)(scal + 12), false);
}

// The speculative rerun's fields (rq_*) copied over the first pass's first
// QPAD slots, where the next rerun finds the queries it lists as failed.
__global__ void k_spec_adopt(const int32_t *r, const float *rb, int32_t *qpos, int32_t *kq, int32_t *orig,
                             int32_t *status, float *bnd)
{
    const int i = threadIdx.x;
    if (i < 64) {
        qpos[i] = r[i];
        kq[i] = r[64 + i];
        orig[i] = r[128 + i];
        status[i] = r[192 + i];
        bnd[i] = rb[i];
    }
}

// First pass of a learned kNN batch under the exhaustive plan, stream
// operations only (no host synchronisation, no allocation once the buffers
// exist), so that it can be captured into a graph: query preparation, radius
// network, counter resets (fill kernels), dense verification between the
// events ev[0], ev[1] (none when null), top-k.
static void first_pass_exhaustive(gts_index *I, const int32_t *d_qids, int64_t nq, const int32_t *d_ks, int kfix,
                                  int kstride, int32_t *d_ids, float *d_dists, float *d_kth, hipEvent_t *ev)
{
    hipStream_t st = I->st;
    const int cap = topk_cap_limit();
    int32_t *qpos = scratch(I->qpos, nq), *kq = scratch(I->kq, nq), *orig = scratch(I->orig, nq);
    float *bnd = scratch(I->bnd, nq);
    int64_t *scal = I->scal.ensure(16);
    VerifyArgs va;
    va.cand = scratch(I->cand, nq * (int64_t)cap);
    va.ccount = scratch(I->ccount, nq);
    va.cap = cap;
    va.verified = scal + 9;
    va.live = scal + 11;
    va.ncand = scal + 13;
    int ready = 0;
    if (I->cells_plan) {
        // the batch in the cell scan's query order, counters and bitmap
        // cleared; the radius network then stages the query image that the
        // scan reads
        cells_prepare_batch(I, d_qids, d_ks, kfix, std::min(kstride, topk_cap_limit()), arg_flag(I), nq, qpos, kq,
                            orig, bnd, va.ccount, (int32_t *)(scal + 8));
        ready = CELLS_CLEARED | (learned_predict(I, qpos, kq, nq, bnd) ? CELLS_STAGED : 0);
    } else {
        k_prep<<<grid_for(nq, 256), 256, 0, st>>>(I->inv, d_qids, d_ks, kfix, nq, INFINITY, qpos, kq, orig, bnd,
                                                  I->n, std::min(kstride, topk_cap_limit()), arg_flag(I));
        GTS_HIP(hipGetLastError());
        learned_predict(I, qpos, kq, nq, bnd);
        fill_i32(va.ccount, 0, nq, st);
        fill_i32((int32_t *)(scal + 8), 0, 14, st);
    }
    Search S;
    S.idx = I;
    S.nq = nq;
    S.qpos = qpos;
    S.bnd = bnd;
    S.kq = kq;
    // external records: under capture they become event nodes of the graph
    if (I->cells_plan) {
        cells_launch(I, S, va, ev[0], ev[1], true, ready);
    } else {
        if (ev[0]) GTS_HIP(hipEventRecordWithFlags(ev[0], st, hipEventRecordExternal));
        dense_all_launch(I, S, va);
        if (ev[1]) GTS_HIP(hipEventRecordWithFlags(ev[1], st, hipEventRecordExternal));
    }
    int32_t *status = scratch(I->status, nq);
    // the cell scan's verification counts the candidates itself
    S.topk(va, 1, orig, kstride, d_ids, d_dists, d_kth, status, scratch(I->fail, nq),
           (unsigned long long *)(scal + 12), !I->cells_plan);
    // the first rerun, speculatively, before the host sees the failure count
    if (I->cells_plan) rerun_pass(I, QPAD, 0, kstride, d_ids, d_dists, d_kth, true);
    end_pass(I, true);
}

void graph_reset(gts_index *I)
{
    for (GraphEntry &g : I->graphs) (void)hipGraphExecDestroy(g.exec);
    I->graphs.clear();
    I->graph_pending.clear();
}

// Graph cache of the index: `issue` (stream operations only) runs as a replay
// of its captured graph when `key` was captured before; on its second sighting
// it is captured, instantiated and launched.  Returns false, having issued
// nothing, on a first sighting or when capture fails (graphs then stay off for
// this index and the caller issues ordinary launches).
template <class F>
static bool graph_run(gts_index *I, uint64_t key, F issue)
{
    hipStream_t st = I->st;
    constexpr size_t MAX_GRAPHS = 4, MAX_PENDING = 8;
    for (size_t i = 0; i < I->graphs.size(); i++) {
        if (I->graphs[i].key != key) continue;
        std::rotate(I->graphs.begin(), I->graphs.begin() + i, I->graphs.begin() + i + 1);
        GTS_HIP(hipGraphLaunch(I->graphs[0].exec, st));
        return true;
    }
    auto pend = std::find(I->graph_pending.begin(), I->graph_pending.end(), key);
    if (pend == I->graph_pending.end()) {   // first sighting: ordinary launches, capture next time
        I->graph_pending.push_back(key);
        if (I->graph_pending.size() > MAX_PENDING) I->graph_pending.erase(I->graph_pending.begin());
        return false;
    }
    I->graph_pending.erase(pend);
    if (I->graphs.size() == MAX_GRAPHS) {   // drop the least recently used
        (void)hipGraphExecDestroy(I->graphs.back().exec);
        I->graphs.pop_back();
    }
    hipGraph_t g = nullptr;
    GTS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        issue();
    } catch (...) {
        // nothing captured ran; this index keeps to ordinary launches, which
        // report any error of their own
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        I->use_graph = false;
        return false;
    }
    if (hipStreamEndCapture(st, &g) != hipSuccess || !g) {
        (void)hipGetLastError();
        I->use_graph = false;
        return false;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        (void)hipGetLastError();
        I->use_graph = false;
        return false;
    }
    // a captured dense launch recorded its schedule
    I->graphs.insert(I->graphs.begin(), GraphEntry{key, ex, I->last_sched});
    GTS_HIP(hipGraphLaunch(ex, st));
    return true;
}

// Runs the first pass as a graph replay when the batch repeats: returns false
// (nothing issued) when the pass must run as ordinary launches this time.
static bool first_pass_graph(gts_index *I, const int32_t *d_qids, int64_t nq, const int32_t *d_ks, int kfix,
                             int kstride, int32_t *d_ids, float *d_dists, float *d_kth, uint64_t shape)
{
    if (!I->use_graph || I->verify_form == GTS_VERIFY_FP32 || !v2_supported(I, 2)) return false;
    if (!plan_known_exhaustive(I, nq, shape, false)) return false;
    const bool cells = use_cells(I);
    if (!cells && !dense_schedule_known(I, nq)) return false;
    I->cells_plan = cells;
    // every buffer at its size for this batch, so the capture allocates nothing
    const int cap = topk_cap_limit();
    const int32_t *ptrs[] = {scratch(I->qpos, nq), scratch(I->kq, nq), scratch(I->orig, nq), scratch(I->status, nq),
                             scratch(I->ccount, nq), scratch(I->fail, nq)};
    const float *bnd = scratch(I->bnd, nq);
    const uint64_t *cand = scratch(I->cand, nq * (int64_t)cap);
    const int64_t *scal = I->scal.ensure(16);
    const float *feat = I->dfeat.ensure(nq * 16), *pred = I->dpred.ensure(nq);
    if (I->radius_src == GTS_RADIUS_PIVOTS) pivot_radius_buffers(I, nq);
    const float *ftab = I->ftab.ensure(std::min<int64_t>(nq, 16384) * I->model.npiv);
    const int64_t nqb = (nq + 63) / 64;
    const uint16_t *qimg = I->qimg.ensure(nqb * 2 * 64 * (I->dk + 8));
    uint64_t key = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { mix_key(&key, v); };
    mix(cells ? 1 : 2);
    if (cells) {
        cells_reserve(I, nq, mix_key, &key);
        cells_ranks_ready(I, nq);
        // the speculative rerun's buffers
        for (const void *p : {(const void *)I->rq_i.ensure(5 * QPAD), (const void *)I->rq_b.ensure(QPAD),
                              (const void *)I->rq_cand.ensure(QPAD * (int64_t)cap)})
            mix((uint64_t)(uintptr_t)p);
    }
    for (const int32_t *p : ptrs) mix((uint64_t)(uintptr_t)p);
    for (const void *p : {(const void *)bnd, (const void *)cand, (const void *)scal, (const void *)feat,
                          (const void *)pred, (const void *)ftab, (const void *)qimg, (const void *)d_qids,
                          (const void *)d_ks, (const void *)d_ids, (const void *)d_dists, (const void *)d_kth,
                          (const void *)I->X, (const void *)I->Xhi, (const void *)I->model.wb,
                          (const void *)I->model.piv, (const void *)I->death_ptr, (const void *)I->qop_ptr,
                          (const void *)I->rp_piv.p, (const void *)I->rp_off.p})
        mix((uint64_t)(uintptr_t)p);
    mix((uint64_t)I->radius_src);
    mix((uint64_t)nq);
    mix((uint64_t)(uint32_t)kfix);
    mix((uint64_t)(uint32_t)kstride);
    mix(shape);
    mix((uint64_t)I->forced_sched.rows * 2 + I->forced_sched.persist);
    mix(I->timing ? 1 : 0);
    for (hipEvent_t &e : I->graph_ev)
        if (!e) GTS_HIP(hipEventCreate(&e));
    hipEvent_t untimed[2] = {nullptr, nullptr};
    const bool ran = graph_run(I, key, [&]() {
        first_pass_exhaustive(I, d_qids, nq, d_ks, kfix, kstride, d_ids, d_dists, d_kth,
                              I->timing ? I->graph_ev : untimed);
    });
    if (!ran) return false;
    plan_known_exhaustive(I, nq, shape, true);
    I->last_sched = I->graphs[0].sched;
    I->last_sched_nq = nq;
    return true;
}

// The rerun pass as a graph replay when possible, else as ordinary launches.
static void rerun_pass_graph(gts_index *I, int64_t npad, int attempt, int kstride, int32_t *d_ids, float *d_dists,
                             float *d_kth)
{
    if (I->use_graph) {
        const int cap = topk_cap_limit();
        uint64_t key = 7809847782465536322ull;
        auto mix = [&](uint64_t v) { mix_key(&key, v); };
        cells_reserve(I, npad, mix_key, &key);
        for (const void *p : {(const void *)scratch(I->qpos, npad), (const void *)scratch(I->kq, npad),
                              (const void *)scratch(I->orig, npad), (const void *)scratch(I->bnd, npad),
                              (const void *)scratch(I->fail, npad), (const void *)scratch(I->status, npad),
                              (const void *)scratch(I->ccount, npad), (const void *)scratch(I->cand, npad * (int64_t)cap),
                              (const void *)I->scal.ensure(16), (const void *)d_ids, (const void *)d_dists,
                              (const void *)d_kth, (const void *)I->X, (const void *)I->Xhi})
            mix((uint64_t)(uintptr_t)p);
        mix((uint64_t)npad);
        mix((uint64_t)attempt);
        mix((uint64_t)(uint32_t)kstride);
        mix(I->timing ? 1 : 0);
        if (graph_run(I, key, [&]() { rerun_pass(I, npad, attempt, kstride, d_ids, d_dists, d_kth); })) {
            I->timing_pending = I->timing;
            return;
        }
    }
    rerun_pass(I, npad, attempt, kstride, d_ids, d_dists, d_kth);
}

// Event timings of a kNN call are read after the call, when its stats are
// asked for (the events are complete then): no event queries on the
// critical path of a replayed step.  A later pass that re-records an event
// pair reads the pending one first.
void finalize_timing(gts_index *I, int bits)
{
    const int b = I->pend_timing & bits;
    float ms = 0;
    auto add_verify = [&](hipEvent_t e0, hipEvent_t e1) {
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
            I->stats.ms_verify += ms;
            I->stats.verify_launches++;
        } else {
            (void)hipGetLastError();   // untimed: reported as 0 ms
        }
    };
    if (b & 1) {   // a captured first pass: the batch's own verification launches
        add_verify(I->graph_ev[0], I->graph_ev[1]);
        if (ms > 0.0f) {
            I->stats.ms_verify_batch += ms;
            I->stats.verify_launches_batch++;
        }
    }
    if (b & 2) add_verify(I->timing_ev[0], I->timing_ev[1]);
    if (b & 4) {
        GTS_HIP(hipEventSynchronize(I->call_ev[1]));   // recorded before the last publish: already reached
        if (hipEventElapsedTime(&ms, I->call_ev[0], I->call_ev[1]) == hipSuccess) I->stats.ms_total = ms;
        else (void)hipGetLastError();
    }
    I->pend_timing &= ~bits;
}

// ---- exact fallback ------------------------------------------------------
// Queries still failing after EXACT_AFTER passes (a band of exact ties wider
// than the candidate capacity keeps overflowing, whatever the bound) are
// answered exactly: the query's distance to every live object, in ID order,
// as a 64-bit key (FP64 distance bits, which order as integers for
// non-negative values; edit distances exact), a stable radix sort (ties
// stay in ID order), and the first k.  It is a rare path: O(n) work and an
// n-key sort per query.
constexpr int EXACT_AFTER = 12;

template <int M>
__global__ void __launch_bounds__(256) k_exact_keys(DevTree t, const int32_t *inv, int64_t qp, uint64_t *keys,
                                                    int32_t *vals)
{
    constexpr int G = 8;   // lanes per object
    const int gl = threadIdx.x % G;
    const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G, ng = (int64_t)gridDim.x * blockDim.x / G;
    for (int64_t id = g0; id < t.n; id += ng) {   // every lane of a group takes the same objects
        const int64_t pos = inv[id];
        double d;
        if constexpr (M == GTS_METRIC_EDIT) {
            d = (double)gts_edit_distance(t.S + qp * t.sstride, t.slen[qp], t.S + pos * t.sstride, t.slen[pos]);
        } else {
            d = group_dist64<M, G>(t.X + qp * t.dp, t.X + pos * t.dp, t.dim, gl);
        }
        if (pos == qp) d = 0.0;
        const bool gone = t.death && t.death[pos] != INT32_MAX;   // deleted objects are not results
        if (gl == 0) {
            keys[id] = gone ? ~0ull : (uint64_t)__double_as_longlong(d);
            vals[id] = (int32_t)id;
        }
    }
}

// the first k of the sorted keys into output row orow (then -1 / +inf)
__global__ void k_exact_emit(const uint64_t *keys, const int32_t *vals, int64_t n, int k, int64_t orow, int kstride,
                             int32_t *ids, float *dists, float *kth)
{
    for (int i = threadIdx.x; i < kstride; i += blockDim.x) {
        const bool ok = i < k && i < n && keys[i] != ~0ull;
        ids[orow * kstride + i] = ok ? vals[i] : -1;
        dists[orow * kstride + i] = ok ? (float)__longlong_as_double((long long)keys[i]) : INFINITY;
    }
    if (threadIdx.x == 0 && kth) {
        const bool ok = k <= n && keys[k - 1] != ~0ull;
        kth[orow] = ok ? (float)__longlong_as_double((long long)keys[k - 1]) : INFINITY;
    }
}

static void knn_exact_fallback(gts_index *I, const int32_t *failq, int64_t nfail, const int32_t *qpos,
                               const int32_t *kq, const int32_t *orig, int64_t nl, int kstride, int32_t *d_ids,
                               float *d_dists, float *d_kth)
{
    hipStream_t st = I->st;
    std::vector<int32_t> hf(nfail), hq(nl), hk(nl), ho(nl);
    GTS_HIP(hipMemcpyAsync(hf.data(), failq, nfail * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipMemcpyAsync(hq.data(), qpos, nl * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipMemcpyAsync(hk.data(), kq, nl * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipMemcpyAsync(ho.data(), orig, nl * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    const int64_t n = I->n;
    uint64_t *k0 = I->ex_k.ensure(2 * n), *k1 = k0 + n;
    int32_t *v0 = I->ex_v.ensure(2 * n), *v1 = v0 + n;
    const DevTree t = I->dev();
    for (int64_t f = 0; f < nfail; f++) {
        const int32_t b = hf[f];
        if (b < 0 || b >= nl || ho[b] < 0) continue;   // an inert padding slot
        const unsigned grid = grid_for(n * 8, 256);
        switch (I->metric) {
        case GTS_METRIC_L1: k_exact_keys<GTS_METRIC_L1><<<grid, 256, 0, st>>>(t, I->inv, hq[b], k0, v0); break;
        case GTS_METRIC_LINF: k_exact_keys<GTS_METRIC_LINF><<<grid, 256, 0, st>>>(t, I->inv, hq[b], k0, v0); break;
        case GTS_METRIC_ANGULAR:
            k_exact_keys<GTS_METRIC_ANGULAR><<<grid, 256, 0, st>>>(t, I->inv, hq[b], k0, v0);
            break;
        case GTS_METRIC_EDIT: k_exact_keys<GTS_METRIC_EDIT><<<grid, 256, 0, st>>>(t, I->inv, hq[b], k0, v0); break;
        default: k_exact_keys<GTS_METRIC_L2><<<grid, 256, 0, st>>>(t, I->inv, hq[b], k0, v0); break;
        }
        GTS_HIP(hipGetLastError());
        uint64_t *ks = nullptr;
        int32_t *vs = nullptr;
        radix_sort_u64(k0, v0, k1, v1, n, 64, nullptr, 0, st, &ks, &vs);
        k_exact_emit<<<1, 256, 0, st>>>(ks, vs, n, hk[b], ho[b], kstride, d_ids, d_dists, d_kth);
        GTS_HIP(hipGetLastError());
    }
    GTS_HIP(hipStreamSynchronize(st));
}

// ---- the persistent low-dimensional pass ----------------------------------
// Learned kNN batches over rows of at most 8 dimensions (the automatic plan or
// PERSISTENT) run as one persistent launch (gts_lowdim.hip) instead of the
// level-by-level passes.

static bool use_lowdim(const gts_index *I, int learned, int kstride)
{
    if (!learned || (I->knn_plan != GTS_PLAN_AUTO && I->knn_plan != GTS_PLAN_PERSISTENT)) return false;
    if (lowdim_supported(I, kstride)) return true;
    if (I->knn_plan == GTS_PLAN_PERSISTENT)
        fail(GTS_E_ARG, "the persistent plan needs L2, L1 or Linf rows of at most 8 dimensions, kstride <= 64 and "
                        "a tree of at most 512 / (order - 1) levels");
    return false;
}

// Stream operations only once the buffers exist, so that the pass can be
// captured: query preparation, the radius (network or level pivots), counter
// resets by fill kernels, the launch between the events ev[0], ev[1] (none
// when null), the publish of the pass scalars.
static void lowdim_issue(gts_index *I, const int32_t *d_qids, int64_t nq, const int32_t *d_ks, int kfix, int kstride,
                         int32_t *d_ids, float *d_dists, float *d_kth, int bpc, const hipEvent_t *ev, bool external)
{
    hipStream_t st = I->st;
    int32_t *qpos = scratch(I->qpos, nq), *kq = scratch(I->kq, nq), *orig = scratch(I->orig, nq);
    float *bnd = scratch(I->bnd, nq);
    int64_t *scal = I->scal.ensure(16);
    uint32_t *work = I->ld_work.ensure(1);
    k_prep<<<grid_for(nq, 256), 256, 0, st>>>(I->inv, d_qids, d_ks, kfix, nq, INFINITY, qpos, kq, orig, bnd, I->n,
                                              std::min(kstride, topk_cap_limit()), arg_flag(I));
    GTS_HIP(hipGetLastError());
    learned_predict(I, qpos, kq, nq, bnd);
    fill_i32((int32_t *)work, 0, 1, st);
    fill_i32((int32_t *)(scal + 8), 0, 14, st);
    auto rec = [&](hipEvent_t e) {
        if (e) GTS_HIP(external ? hipEventRecordWithFlags(e, st, hipEventRecordExternal) : hipEventRecord(e, st));
    };
    rec(ev[0]);
    lowdim_launch(I, qpos, kq, bnd, orig, nq, kstride, d_ids, d_dists, d_kth, work, (unsigned long long *)(scal + 9),
                  bpc);
    rec(ev[1]);
    end_pass(I, external);
}

// The batch: the launch schedule (blocks per CU) of the batch shape from the
// autotuner, which times each candidate on the first batch of a shape (every
// run writes the same results); a repeated batch replays its captured pass.
static void lowdim_run(gts_index *I, const int32_t *d_qids, int64_t nq, const int32_t *d_ks, int kfix, int kstride,
                       int32_t *d_ids, float *d_dists, float *d_kth)
{
    int qlog = 0;
    while (((int64_t)1 << qlog) < nq) qlog++;
    const uint64_t skey = ((uint64_t)qlog << 32) | (uint32_t)kstride;
    hipEvent_t untimed[2] = {nullptr, nullptr};
    const int64_t *h = I->hpin;
    auto publish_checked = [&]() {
        await_publish(I);
        if (h[6])
            fail(GTS_E_ARG, "query IDs outside [0, n) or k outside [1, min(kstride, " +
                                std::to_string(topk_cap_limit()) + ")]");
    };
    auto eager = [&](int bpc, bool timed) {
        lowdim_issue(I, d_qids, nq, d_ks, kfix, kstride, d_ids, d_dists, d_kth, bpc,
                     timed ? I->timing_ev : untimed, false);
        publish_checked();
        float ms = -1.0f;
        if (timed && hipEventElapsedTime(&ms, I->timing_ev[0], I->timing_ev[1]) != hipSuccess) {
            (void)hipGetLastError();
            ms = -1.0f;
        }
        return ms;
    };
    auto it = I->lowdim_sched.find(skey);
    if (it == I->lowdim_sched.end()) {
        int pick = 4;
        if (I->autotune) {
            float best = INFINITY;
            for (int bpc : {2, 4, 8}) {
                const float ms = eager(bpc, true);
                if (ms >= 0.0f && ms < best) {
                    best = ms;
                    pick = bpc;
                }
            }
        } else {
            eager(pick, false);
        }
        I->lowdim_sched[skey] = pick;
        I->stats.verified_pairs += h[0];
        I->stats.verify_launches++;
        return;
    }
    const int bpc = it->second;
    bool ran = false;
    if (I->use_graph) {
        // every buffer at its size for this batch, so the capture allocates nothing
        const int32_t *ptrs[] = {scratch(I->qpos, nq), scratch(I->kq, nq), scratch(I->orig, nq)};
        const float *bnd = scratch(I->bnd, nq);
        const int64_t *scal = I->scal.ensure(16);
        const uint32_t *work = I->ld_work.ensure(1);
        const float *feat = I->dfeat.ensure(nq * 16), *pred = I->dpred.ensure(nq);
        if (I->radius_src == GTS_RADIUS_PIVOTS) pivot_radius_buffers(I, nq);
        const float *ftab = I->ftab.ensure(std::min<int64_t>(nq, 16384) * std::max(I->model.npiv, 1));
        uint64_t key = 5381ull * 0x9E3779B97F4A7C15ull;
        auto mix = [&](uint64_t v) { mix_key(&key, v); };
        for (const int32_t *p : ptrs) mix((uint64_t)(uintptr_t)p);
        for (const void *p : {(const void *)bnd, (const void *)scal, (const void *)work, (const void *)feat,
                              (const void *)pred, (const void *)ftab, (const void *)d_qids, (const void *)d_ks,
                              (const void *)d_ids, (const void *)d_dists, (const void *)d_kth, (const void *)I->X,
                              (const void *)I->model.wb, (const void *)I->model.piv, (const void *)I->rp_piv.p,
                              (const void *)I->rp_off.p})
            mix((uint64_t)(uintptr_t)p);
        for (uint64_t v : {(uint64_t)nq, (uint64_t)(uint32_t)kfix, (uint64_t)(uint32_t)kstride, (uint64_t)bpc,
                           (uint64_t)I->radius_src, (uint64_t)(I->timing ? 1 : 0), (uint64_t)0x10D1})
            mix(v);
        for (hipEvent_t &e : I->graph_ev)
            if (!e) GTS_HIP(hipEventCreate(&e));
        ran = graph_run(I, key, [&]() {
            lowdim_issue(I, d_qids, nq, d_ks, kfix, kstride, d_ids, d_dists, d_kth, bpc,
                         I->timing ? I->graph_ev : untimed, true);
        });
    }
    if (ran) {
        publish_checked();
        I->stats.graph_replays++;
        if (I->timing) I->pend_timing |= 1;
    } else {
        const float ms = eager(bpc, I->timing);
        if (ms >= 0.0f) {
            I->stats.ms_verify += ms;
            I->stats.verify_launches++;
            I->stats.ms_verify_batch += ms;
            I->stats.verify_launches_batch++;
        }
    }
    I->stats.verified_pairs += h[0];
}

static void knn_run_batch(gts_index *I, const int32_t *d_qids, int64_t nq, const int32_t *d_ks, int kfix,
                          int kstride, int32_t *d_ids, float *d_dists, float *d_kth, int learned, int verify)
{
    hipStream_t st = I->st;
    for (hipEvent_t *e : {&I->call_ev[0], &I->call_ev[1], &I->timing_ev[0], &I->timing_ev[1]})
        if (!*e) GTS_HIP(hipEventCreate(e));
    if (!I->hpin) {
        GTS_HIP(hipHostMalloc((void **)&I->hpin, 16 * sizeof(int64_t), hipHostMallocCoherent));
        for (int i = 0; i < 16; i++) I->hpin[i] = 0;
    }
    {
        int64_t *scal = I->scal.ensure(16);
        // a call that ended in an error may have left work in flight: drain
        // it, so that its publish cannot pass for this call's
        int32_t *flag = arg_flag(I);
        if (!I->pub_clean) {
            GTS_HIP(hipStreamSynchronize(st));
            I->qrank_zero = nullptr;   // the failed call may have left ranks set
            GTS_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), st));   // and the argument flag
        }
        if (I->pub_scal != scal) {
            // the device publish counter (scal[15]) of a fresh block
            GTS_HIP(hipMemsetAsync(scal + 15, 0, sizeof(int64_t), st));
            GTS_HIP(hipStreamSynchronize(st));
            I->pub_scal = scal;
        }
        I->pub_clean = false;
        I->pub_seq = ((volatile int64_t *)I->hpin)[15];   // every earlier publish has been seen
    }
    if (I->timing) GTS_HIP(hipEventRecord(I->call_ev[0], st));
    I->stats = gts_stats{};
    I->pend_timing = 0;
    I->verify_form = verify;
    I->cells_plan = false;
    I->timing_pending = false;
    I->last_sched = DenseSched{};
    I->last_sched_nq = 0;
    if (!learned && I->knn_plan == GTS_PLAN_EXHAUSTIVE) fail(GTS_E_ARG, "the exhaustive plan needs the learned radius");
    if (use_lowdim(I, learned, kstride)) {
        lowdim_run(I, d_qids, nq, d_ks, kfix, kstride, d_ids, d_dists, d_kth);
        if (I->timing) I->pend_timing |= 4;
        I->pub_clean = true;
        return;
    }
    // per-query buffers sized for the batch and for a padded rerun batch up
    // front: later passes of the call must not reallocate (and so lose) them
    {
        const int64_t ncap = std::max<int64_t>(nq, MAX_PADDED_RERUN);
        scratch(I->qpos, ncap);
        scratch(I->kq, ncap);
        scratch(I->orig, ncap);
        scratch(I->bnd, ncap);
        scratch(I->fail, ncap);
        scratch(I->status, ncap);
        scratch(I->ccount, ncap);
        scratch(I->cand, ncap * (int64_t)topk_cap_limit());
    }
    int qlog = 0;
    while (((int64_t)1 << qlog) < nq) qlog++;
    const uint64_t shape = ((uint64_t)qlog << 32) | (uint32_t)(d_ks ? -1 : kfix);
    const bool graphed = learned && first_pass_graph(I, d_qids, nq, d_ks, kfix, kstride, d_ids, d_dists, d_kth, shape);
    int32_t *qpos = scratch(I->qpos, nq);
    float *bnd = scratch(I->bnd, nq);
    int32_t *kq = scratch(I->kq, nq);
    int32_t *orig = scratch(I->orig, nq);
    int32_t *failq = scratch(I->fail, nq);
    bool exhaustive = graphed;
    if (!graphed) {
        k_prep<<<grid_for(nq, 256), 256, 0, st>>>(I->inv, d_qids, d_ks, kfix, nq, INFINITY, qpos, kq, orig, bnd,
                                                  I->n, std::min(kstride, topk_cap_limit()), arg_flag(I));
        GTS_HIP(hipGetLastError());
        if (learned) learned_predict(I, qpos, kq, nq, bnd);
        // reruns keep the plan of the first pass
        exhaustive = learned && plan_exhaustive(I, qpos, bnd, kq, nq, shape);
        I->cells_plan = exhaustive && use_cells(I);
    }
    const int cap = topk_cap_limit();
    int64_t nl = nq;
    int64_t *scal = I->scal.ensure(16);
    int64_t padded = 0;   // > 0: this pass is a padded rerun of that many slots, gathered on the device
    for (int attempt = 0;; attempt++) {
        Search S;
        S.idx = I;
        S.nq = nl;
        S.qpos = qpos;
        S.bnd = bnd;
        S.kq = kq;
        VerifyArgs va;
        va.cand = scratch(I->cand, nl * (int64_t)cap);
        va.ccount = scratch(I->ccount, nl);
        va.cap = cap;
        va.verified = scal + 9;
        va.live = scal + 11;
        va.ncand = scal + 13;
        int32_t *status = scratch(I->status, nl);
        if (attempt > 0) finalize_timing(I, 2);   // this pass may re-record timing_ev
        if (padded) {
            rerun_pass_graph(I, nl, attempt - 1, kstride, d_ids, d_dists, d_kth);
            S.stats.exhaustive_passes++;
        } else if (!(attempt == 0 && graphed)) {
            if (!exhaustive) S.expand(true);
            if (I->cells_plan) cells_sort_queries(I, qpos, kq, bnd, orig, nl);
            fill_i32(va.ccount, 0, nl, st);
            fill_i32((int32_t *)(scal + 8), 0, 14, st);
            if (exhaustive) S.verify_all(va);
            else S.verify(2, va);
            S.topk(va, learned ? 1 : 0, orig, kstride, d_ids, d_dists, d_kth, status, failq,
                   (unsigned long long *)(scal + 12), !I->cells_plan);
        }
        // one readback per pass: verified pairs, live chunks, failures,
        // candidates, queries of the speculative rerun
        if (!(attempt == 0 && graphed)) end_pass(I, false);   // a graphed first pass ends with its own
        await_publish(I);
        const int64_t *h = I->hpin;
        float ms = 0;
        if (attempt == 0 && graphed) {
            // the replayed graph ran the first pass; its events bracket the
            // verification kernel (read with the stats)
            if (I->timing) I->pend_timing |= 1;
            S.stats.exhaustive_passes++;
            S.stats.graph_replays++;
        } else if (I->timing_pending) {
            if (hipEventElapsedTime(&ms, I->timing_ev[0], I->timing_ev[1]) == hipSuccess) {
                S.stats.ms_verify += ms;
                S.stats.verify_launches++;
                if (attempt == 0) {
                    S.stats.ms_verify_batch += ms;
                    S.stats.verify_launches_batch++;
                }
            }
            I->timing_pending = false;
        }
        S.stats.verified_pairs += h[0];
        S.stats.live_chunks += h[2] + h[7];
        if (attempt == 0) S.stats.live_chunks_batch += h[2];
        S.stats.candidates += h[4];
        if (attempt == 0 && h[6])
            fail(GTS_E_ARG, "query IDs outside [0, n) or k outside [1, min(kstride, " +
                                std::to_string(topk_cap_limit()) + ")]");
        const int64_t nfail = h[3];
        const int64_t spec = (attempt == 0 && graphed) ? h[5] : 0;
        if (spec > 0) {
            // the graph's speculative rerun took the first pass's failures;
            // nfail counts its own, which index its rq_* slots
            S.stats.exhaustive_passes++;
            if (I->timing) I->pend_timing |= 2;
            I->stats.retries += spec;
            attempt++;
        }
        add_stats(I->stats, S.stats);
        if (nfail == 0) break;
        if (attempt >= EXACT_AFTER && !padded && spec == 0) {
            knn_exact_fallback(I, failq, nfail, qpos, kq, orig, nl, kstride, d_ids, d_dists, d_kth);
            I->stats.retries += nfail;
            break;
        }
        if (attempt >= 40) fail(GTS_E_CAPACITY, "kNN reruns did not converge");
        I->stats.retries += nfail;
        if (spec > 0) {
            k_spec_adopt<<<1, 64, 0, st>>>(I->rq_i.ensure(5 * QPAD), I->rq_b.ensure(QPAD), qpos, kq, orig,
                                           scratch(I->status, QPAD), bnd);
            GTS_HIP(hipGetLastError());
        }
        if (attempt >= 8 && I->verify_form == GTS_VERIFY_BF16X3) {
            // The BF16X3 bound grows with the largest row norm: on sets whose
            // norms span many decades the bands of small queries can keep
            // overflowing.  The remaining queries finish in the FP32 form
            // (relative bound) through the tree.
            I->verify_form = GTS_VERIFY_FP32;
            I->cells_plan = false;
            exhaustive = false;
        }
        if (I->cells_plan && nfail <= MAX_PADDED_RERUN) {
            padded = (nfail + QPAD - 1) / QPAD * QPAD;
            nl = padded;
            continue;
        }
        padded = 0;
        int32_t *t = I->qtmp.ensure(3 * nfail);
        float *tf = I->qtmp_f.ensure(nfail);
        k_retry_gather<<<grid_for(nfail, 256), 256, 0, st>>>(failq, nfail, status, attempt, qpos, kq, orig, bnd, t,
                                                             tf);
        k_retry_scatter<<<grid_for(nfail, 256), 256, 0, st>>>(nfail, t, tf, qpos, kq, orig, bnd);
        GTS_HIP(hipGetLastError());
        nl = nfail;
    }
    if (I->timing) I->pend_timing |= 4;
    I->stats.dense_rows = I->last_sched.rows;
    I->stats.dense_persistent = I->last_sched.persist;
    I->stats.cell_scan = I->cells_plan ? I->cells.ncell : 0;
    I->cells_plan = false;
    I->pub_clean = true;
}

// per-query buffers of a kNN batch: the candidate list (TOPK_CAP keys) and its
// count, position, k, output row, radius, status, failure slot; a batch holds
// at least MAX_PADDED_RERUN of them
void knn_run(gts_index *I, const int32_t *d_qids, int64_t nq, const int32_t *d_ks, int kfix, int kstride,
             int32_t *d_ids, float *d_dists, float *d_kth, int learned, int verify)
{
    // (the persistent low-dimensional pass keeps no candidate lists: position,
    // k, output row, radius)
    const bool low = use_lowdim(I, learned, kstride);
    const int64_t per_query = low ? 16 : (int64_t)topk_cap_limit() * 8 + 32;
    budget_split(I, nq, per_query, low ? 1 : MAX_PADDED_RERUN, [&](int64_t q0, int64_t n) {
        knn_run_batch(I, d_qids + q0, n, d_ks ? d_ks + q0 : nullptr, kfix, kstride, d_ids + q0 * kstride,
                      d_dists + q0 * kstride, d_kth + q0, learned, verify);
    });
}

}  // namespace gts

// Update path: tombstones, a lazy cache list of inserts and a rebuild when it
// fills (updateIndexRnn, update.cuh:449-638).
//
// Semantics kept from the reference:
//   op 0 insert : the row `id` of the current data joins the cache list
//                 (update.cuh:481); a full list triggers a rebuild over the live
//                 tree rows in physical order followed by the cached rows
//                 (getNewData + indexConstru, update.cuh:484-540);
//   op 1 delete : `id` is logical = rank among live tree rows, then cache
//                 entries (findIdx, update.cuh:118-143); a tree row is
//                 tombstoned, a cache entry is removed (update.cuh:546-582);
//   op 2 query  : `id` is a physical row (update.cuh:588); the count is the live
//                 tree rows within r (leafProcessRnnUpdate skips tombstones,
//                 update.cuh:244) plus the cache entries within r
//                 (searchNaiveRnn, search_naive.cuh:208-251).
// Extension: op 3 inserts row `id` of the caller's `fresh` matrix.
// Differences: consecutive query ops see the same state, so they run as one
// batched range search; the cache count is recomputed for every query (the
// reference reuses a stale count when the cache is empty, search_naive.cuh
// is skipped and rnum keeps its last value, update.cuh:592-597); invalid IDs
// are errors instead of out-of-bounds reads.
#include <algorithm>
#include <cstring>

#include "gts_api_internal.h"
#include "gts_dist.h"
#include "gts_search.h"

namespace gts {

namespace {

unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

// Fenwick tree over live flags: rank/select for logical IDs.
struct Fenwick {
    std::vector<int32_t> t;
    int64_t n = 0;
    void init(int64_t nn)
    {
        n = nn;
        t.assign(nn + 1, 0);
        for (int64_t i = 1; i <= nn; i++) {
            t[i] += 1;
            int64_t j = i + (i & -i);
            if (j <= nn) t[j] += t[i];
        }
    }
    void add(int64_t i, int v)
    {
        for (i++; i <= n; i += i & -i) t[i] += v;
    }
    // smallest index whose prefix count exceeds k (k-th live row, 0-based)
    int64_t select(int64_t k) const
    {
        int64_t pos = 0, step = 1;
        while (step * 2 <= n) step *= 2;
        for (; step > 0; step >>= 1) {
            if (pos + step <= n && t[pos + step] <= k) {
                pos += step;
                k -= t[pos];
            }
        }
        return pos;  // 0-based index
    }
};

__global__ void k_qpos(const int32_t *inv, const int32_t *ids, int64_t n, int32_t *qpos)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        qpos[i] = inv[ids[i]];
}

__global__ void k_set_dead(uint8_t *dead, const int32_t *pos, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dead[pos[i]] = 1;
}

// counts[q] += #{cache rows j : d(q, C_j) <= r}, exact FP64 (one wave per pair)
template <int M>
__global__ void __launch_bounds__(256) k_cache_count(DevTree t, const int32_t *qpos, int64_t nq,
                                                     const float *C, int nc, double r,
                                                     unsigned long long *counts)
{
    const int lane = threadIdx.x & 63;
    int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t p = w; p < nq * nc; p += nw) {
        int64_t q = p / nc;
        int j = (int)(p % nc);
        double d = group_dist64<M, 64>(t.X + (int64_t)qpos[q] * t.dp, C + (int64_t)j * t.dp, t.dim, lane);
        if (lane == 0 && d <= r) atomicAdd(&counts[q], 1ull);
    }
}

__global__ void k_cache_count_edit(DevTree t, const int32_t *qpos, int64_t nq, const uint8_t *Cs,
                                   const int32_t *Cl, int nc, double r, unsigned long long *counts)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nq * nc;
         p += (int64_t)gridDim.x * blockDim.x) {
        int64_t q = p / nc;
        int j = (int)(p % nc);
        int32_t qp = qpos[q];
        int d = gts_edit_distance(t.S + (int64_t)qp * t.sstride, t.slen[qp], Cs + (int64_t)j * t.sstride, Cl[j]);
        if ((double)d <= r) atomicAdd(&counts[q], 1ull);
    }
}

}  // namespace

void update_run(gts_index *I, const gts_update_op *ops, int64_t nops, float r, int cache_cap,
                const float *fresh, int64_t nfresh, int64_t *counts_out, int64_t cap,
                int64_t *nqueries)
{
    hipStream_t st = I->st;
    const bool is_str = I->metric == GTS_METRIC_EDIT;
    const int dim = I->data.dim;
    I->stats = gts_stats{};
    int64_t tree_size = I->data.n;
    std::vector<uint8_t> dead(tree_size, 0);
    Fenwick fw;
    fw.init(tree_size);
    int64_t live = tree_size;
    // cache (host copies for rebuilds; device copies for the brute-force count)
    std::vector<float> cvec;
    std::vector<uint8_t> cstr;
    std::vector<int32_t> clen;
    int in_size = 0;
    std::vector<int32_t> hinv;  // object ID -> tree position, host copy
    auto load_inv = [&]() {
        hinv.resize(I->n);
        GTS_HIP(hipMemcpyAsync(hinv.data(), I->inv, I->n * 4, hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
    };
    load_inv();
    auto new_dead = [&]() {
        uint8_t *d = I->dead.ensure(I->n);
        GTS_HIP(hipMemsetAsync(d, 0, I->n, st));
    };
    new_dead();
    float *dC = nullptr;
    uint8_t *dCs = nullptr;
    int32_t *dCl = nullptr;
    auto alloc_cache = [&]() {
        if (dC) (void)hipFree(dC);
        if (dCs) (void)hipFree(dCs);
        if (dCl) (void)hipFree(dCl);
        dC = nullptr; dCs = nullptr; dCl = nullptr;
        if (is_str) {
            GTS_HIP(hipMalloc((void **)&dCs, (size_t)cache_cap * I->sstride));
            GTS_HIP(hipMalloc((void **)&dCl, (size_t)cache_cap * 4));
        } else {
            GTS_HIP(hipMalloc((void **)&dC, (size_t)cache_cap * I->dp * 4));
            GTS_HIP(hipMemsetAsync(dC, 0, (size_t)cache_cap * I->dp * 4, st));
        }
    };
    auto upload_cache = [&](int from) {
        if (in_size <= from) return;
        if (is_str) {
            std::vector<uint8_t> pad((size_t)(in_size - from) * I->sstride, 0);
            for (int j = from; j < in_size; j++)
                memcpy(&pad[(size_t)(j - from) * I->sstride], &cstr[(size_t)j * I->data.stride], clen[j]);
            GTS_HIP(hipMemcpyAsync(dCs + (size_t)from * I->sstride, pad.data(), pad.size(), hipMemcpyHostToDevice, st));
            GTS_HIP(hipMemcpyAsync(dCl + from, clen.data() + from, (in_size - from) * 4, hipMemcpyHostToDevice, st));
        } else {
            std::vector<float> pad((size_t)(in_size - from) * I->dp, 0.f);
            for (int j = from; j < in_size; j++)
                memcpy(&pad[(size_t)(j - from) * I->dp], &cvec[(size_t)j * dim], dim * 4);
            GTS_HIP(hipMemcpyAsync(dC + (size_t)from * I->dp, pad.data(), pad.size() * 4, hipMemcpyHostToDevice, st));
        }
        GTS_HIP(hipStreamSynchronize(st));
    };
    alloc_cache();

    std::vector<int32_t> qbatch;      // physical row IDs of pending query ops
    std::vector<int32_t> dbatch;      // positions to tombstone before the next search
    int64_t nout = 0;
    auto apply_dead = [&]() {
        if (dbatch.empty()) return;
        int32_t *dp = I->dqid.ensure(dbatch.size());
        GTS_HIP(hipMemcpyAsync(dp, dbatch.data(), dbatch.size() * 4, hipMemcpyHostToDevice, st));
        k_set_dead<<<grid_for(dbatch.size(), 256), 256, 0, st>>>(I->dead.p, dp, (int64_t)dbatch.size());
        GTS_HIP(hipGetLastError());
        GTS_HIP(hipStreamSynchronize(st));
        dbatch.clear();
    };
    auto flush = [&]() {
        if (qbatch.empty()) return;
        apply_dead();
        const int64_t nb = (int64_t)qbatch.size();
        int32_t *dq = I->dks.ensure(nb);
        int32_t *qp = I->kq.ensure(nb);
        GTS_HIP(hipMemcpyAsync(dq, qbatch.data(), nb * 4, hipMemcpyHostToDevice, st));
        k_qpos<<<grid_for(nb, 256), 256, 0, st>>>(I->inv, dq, nb, qp);
        GTS_HIP(hipGetLastError());
        std::vector<int64_t> hc(nb);
        range_count_dev(I, qp, nb, r, I->dead.p, hc.data());
        if (in_size > 0) {
            // reuse the range count buffer: zero, add cache hits, read
            unsigned long long *cc = (unsigned long long *)I->counts.ensure(nb);
            GTS_HIP(hipMemsetAsync(cc, 0, nb * 8, st));
            DevTree t = I->dev();
            if (is_str) {
                k_cache_count_edit<<<grid_for(nb * in_size, 256), 256, 0, st>>>(t, qp, nb, dCs, dCl, in_size, (double)r, cc);
            } else {
                GTS_METRIC_SWITCH(I->metric, {
                    k_cache_count<M_><<<grid_for(nb * in_size * 64, 256), 256, 0, st>>>(t, qp, nb, dC, in_size, (double)r, cc);
                });
            }
            GTS_HIP(hipGetLastError());
            std::vector<int64_t> add(nb);
            GTS_HIP(hipMemcpyAsync(add.data(), cc, nb * 8, hipMemcpyDeviceToHost, st));
            GTS_HIP(hipStreamSynchronize(st));
            for (int64_t i = 0; i < nb; i++) hc[i] += add[i];
        }
        for (int64_t i = 0; i < nb; i++) {
            if (counts_out && nout < cap) counts_out[nout] = hc[i];
            nout++;
        }
        qbatch.clear();
    };
    auto rebuild = [&]() {
        gts_dataset nd;
        nd.metric = I->data.metric;
        nd.dim = I->data.dim;
        nd.stride = I->data.stride;
        int64_t m = live + in_size;
        nd.n = m;
        if (is_str) {
            nd.str.assign((size_t)m * nd.stride, 0);
            nd.len.resize(m);
        } else {
            nd.vec.resize((size_t)m * dim);
        }
        int64_t o = 0;
        for (int64_t i = 0; i < tree_size; i++) {
            if (dead[i]) continue;
            if (is_str) {
                memcpy(&nd.str[(size_t)o * nd.stride], &I->data.str[(size_t)i * nd.stride], I->data.len[i]);
                nd.len[o] = I->data.len[i];
            } else {
                memcpy(&nd.vec[(size_t)o * dim], &I->data.vec[(size_t)i * dim], dim * 4);
            }
            o++;
        }
        for (int j = 0; j < in_size; j++, o++) {
            if (is_str) {
                memcpy(&nd.str[(size_t)o * nd.stride], &cstr[(size_t)j * nd.stride], clen[j]);
                nd.len[o] = clen[j];
            } else {
                memcpy(&nd.vec[(size_t)o * dim], &cvec[(size_t)j * dim], dim * 4);
            }
        }
        I->data = std::move(nd);
        dbatch.clear();  // pending tombstones refer to the old tree; its dead rows are dropped
        build_tree(I);
        tree_size = I->data.n;
        dead.assign(tree_size, 0);
        fw.init(tree_size);
        live = tree_size;
        in_size = 0;
        cvec.clear(); cstr.clear(); clen.clear();
        new_dead();
        alloc_cache();
        load_inv();
        I->stats.retries++;  // counts rebuilds
    };

    for (int64_t op = 0; op < nops; op++) {
        const int32_t fl = ops[op].flag, id = ops[op].id;
        if (fl == 2) {
            if (id < 0 || id >= tree_size)
                fail(GTS_E_ARG, "query op " + std::to_string(op) + ": row " + std::to_string(id) + " out of range");
            qbatch.push_back(id);
            continue;
        }
        flush();
        if (fl == 0 || fl == 3) {
            const int before = in_size;
            if (fl == 0) {
                if (id < 0 || id >= tree_size)
                    fail(GTS_E_ARG, "insert op " + std::to_string(op) + ": row " + std::to_string(id) + " out of range");
                if (is_str) {
                    cstr.resize((size_t)(in_size + 1) * I->data.stride, 0);
                    memcpy(&cstr[(size_t)in_size * I->data.stride], &I->data.str[(size_t)id * I->data.stride], I->data.len[id]);
                    clen.push_back(I->data.len[id]);
                } else {
                    cvec.insert(cvec.end(), I->data.vec.begin() + (size_t)id * dim, I->data.vec.begin() + (size_t)(id + 1) * dim);
                }
            } else {
                if (is_str || !fresh || id < 0 || id >= nfresh)
                    fail(GTS_E_ARG, "fresh insert op " + std::to_string(op) + " invalid");
                cvec.insert(cvec.end(), fresh + (size_t)id * dim, fresh + (size_t)(id + 1) * dim);
            }
            in_size++;
            if (in_size == cache_cap) rebuild();
            else upload_cache(before);
        } else if (fl == 1) {
            if (id < 0 || id >= live + in_size)
                fail(GTS_E_ARG, "delete op " + std::to_string(op) + ": logical ID " + std::to_string(id) + " out of range");
            if (id < live) {
                int64_t phys = fw.select(id);
                dead[phys] = 1;
                fw.add(phys, -1);
                live--;
                dbatch.push_back(hinv[phys]);  // tombstones are kept by tree position
            } else {
                int j = (int)(id - live);
                if (is_str) {
                    cstr.erase(cstr.begin() + (size_t)j * I->data.stride, cstr.begin() + (size_t)(j + 1) * I->data.stride);
                    clen.erase(clen.begin() + j);
                } else {
                    cvec.erase(cvec.begin() + (size_t)j * dim, cvec.begin() + (size_t)(j + 1) * dim);
                }
                in_size--;
                upload_cache(j);
            }
        } else {
            fail(GTS_E_ARG, "op " + std::to_string(op) + ": unknown flag " + std::to_string(fl));
        }
    }
    flush();
    if (dC) (void)hipFree(dC);
    if (dCs) (void)hipFree(dCs);
    if (dCl) (void)hipFree(dCl);
    if (nqueries) *nqueries = nout;
    if (counts_out && nout > cap) fail(GTS_E_CAPACITY, "counts buffer too small");
}

}  // namespace gts

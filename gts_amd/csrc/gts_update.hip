// Update path: deletion stamps, a lazy cache list of inserts and a rebuild
// when it fills (updateIndexRnn, update.cuh:449-638).
//
// Semantics kept from the reference, op by op:
//   op 0 insert : the row `id` of the current data joins the cache list
//                 (update.cuh:481); a full list triggers a rebuild over the live
//                 tree rows in physical order followed by the cached rows
//                 (getNewData + indexConstru, update.cuh:484-540);
//   op 1 delete : `id` is logical = rank among live tree rows, then cache
//                 entries (findIdx, update.cuh:118-143); a tree row is
//                 deleted, a cache entry is removed (update.cuh:546-582);
//   op 2 query  : `id` is a physical row (update.cuh:588); the count is the live
//                 tree rows within r (leafProcessRnnUpdate skips deleted rows,
//                 update.cuh:244) plus the cache entries within r
//                 (searchNaiveRnn, search_naive.cuh:208-251).
// Extension: op 3 inserts row `id` of the caller's `fresh` matrix.
//
// Batching with snapshot visibility: query ops are collected, up to 4,096, across
// inserts and deletes; each keeps its op index.  Deleting a tree row stamps
// its position with the op index, a cache entry carries the op indexes of its
// insert and removal, and the batch search hides from query q every row
// deleted before q and every cache entry not live at q.  So each query sees
// exactly the state at its position in the stream, as in the sequential
// reference.  A rebuild first completes the pending batch.
// Differences: the cache count is recomputed for every query (the reference
// reuses a stale count when the cache is empty, search_naive.cuh is skipped
// and rnum keeps its last value, update.cuh:592-597); invalid IDs are errors
// instead of out-of-bounds reads.
#include <algorithm>
#include <chrono>
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

__global__ void k_set_death(int32_t *death, const int32_t *pos, const int32_t *op, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        death[pos[i]] = op[i];
}

// counts[q] += #{cache entries j live at q's op : d(q, C_j) <= r}, exact FP64
// (one wave per pair); entry j is live for op indexes in (born_j, died_j)
template <int M>
__global__ void __launch_bounds__(256) k_cache_count(DevTree t, const int32_t *qpos, const int32_t *qop, int64_t nq,
                                                     const float *C, const int32_t *born, const int32_t *died,
                                                     int nc, double r, unsigned long long *counts)
{
    const int lane = threadIdx.x & 63;
    int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t p = w; p < nq * nc; p += nw) {
        int64_t q = p / nc;
        int j = (int)(p % nc);
        if (!(born[j] < qop[q] && qop[q] < died[j])) continue;   // wave-uniform
        double d = group_dist64<M, 64>(t.X + (int64_t)qpos[q] * t.dp, C + (int64_t)j * t.dp, t.dim, lane);
        if (lane == 0 && d <= r) atomicAdd(&counts[q], 1ull);
    }
}

__global__ void k_cache_count_edit(DevTree t, const int32_t *qpos, const int32_t *qop, int64_t nq,
                                   const uint8_t *Cs, const int32_t *Cl, const int32_t *born,
                                   const int32_t *died, int nc, double r, unsigned long long *counts)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nq * nc;
         p += (int64_t)gridDim.x * blockDim.x) {
        int64_t q = p / nc;
        int j = (int)(p % nc);
        if (!(born[j] < qop[q] && qop[q] < died[j])) continue;
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
    constexpr int64_t QBATCH = 4096;   // query ops per batch search
    hipStream_t st = I->st;
    const bool is_str = I->metric == GTS_METRIC_EDIT;
    const int dim = I->data.dim;
    const size_t srow = is_str ? (size_t)I->data.stride : (size_t)dim;   // host row size (bytes or floats)
    if (nops > (int64_t)INT32_MAX - 1) fail(GTS_E_ARG, "too many ops");
    I->stats = gts_stats{};
    // wall time split as the reference reports it: the query ops' searches
    // (flush) and everything else (inserts, deletes, rebuilds)
    const auto t_start = std::chrono::steady_clock::now();
    double s_search = 0.0;
    int64_t tree_size = I->data.n;
    std::vector<uint8_t> dead(tree_size, 0);
    Fenwick fw;
    fw.init(tree_size);
    int64_t live = tree_size;
    std::vector<int32_t> hinv;  // object ID (= physical row) -> tree position, host copy
    auto load_inv = [&]() {
        hinv.resize(I->n);
        GTS_HIP(hipMemcpyAsync(hinv.data(), I->inv, I->n * 4, hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
    };
    auto new_death = [&]() {
        fill_i32(I->death.ensure(I->n), INT32_MAX, I->n, st);
    };
    load_inv();
    new_death();
    // cache entries since the last rebuild, in insertion order (kept until the
    // next flush after their removal, for the queries that still see them)
    std::vector<float> cvec;
    std::vector<uint8_t> cstr;
    std::vector<int32_t> clen, born, died;
    std::vector<int32_t> live_cache;   // entry indexes of the live cache, in order
    DBuf<float> dC;
    DBuf<uint8_t> dCs;
    DBuf<int32_t> dCl, dborn, ddied;
    int64_t uploaded = 0;              // entries whose rows are on the device
    auto nent = [&]() { return (int64_t)born.size(); };

    std::vector<int32_t> qrow, qopv;   // pending query ops: physical row, op index
    std::vector<int64_t> qslot;
    std::vector<int32_t> dpos, dop;    // pending deletion stamps: position, op index
    std::vector<int64_t> out;          // counts by query op
    auto flush = [&]() {
        if (qrow.empty()) return;
        const auto t0 = std::chrono::steady_clock::now();
        if (!dpos.empty()) {
            int32_t *p = I->dids.ensure(2 * (int64_t)dpos.size());
            GTS_HIP(hipMemcpyAsync(p, dpos.data(), dpos.size() * 4, hipMemcpyHostToDevice, st));
            GTS_HIP(hipMemcpyAsync(p + dpos.size(), dop.data(), dop.size() * 4, hipMemcpyHostToDevice, st));
            k_set_death<<<grid_for((int64_t)dpos.size(), 256), 256, 0, st>>>(I->death.p, p, p + dpos.size(),
                                                                            (int64_t)dpos.size());
            GTS_HIP(hipGetLastError());
            dpos.clear();
            dop.clear();
        }
        const int64_t nb = (int64_t)qrow.size();
        int32_t *dq = I->dks.ensure(2 * nb);
        int32_t *qopd = dq + nb;
        int32_t *qp = I->kq.ensure(nb);
        GTS_HIP(hipMemcpyAsync(dq, qrow.data(), nb * 4, hipMemcpyHostToDevice, st));
        GTS_HIP(hipMemcpyAsync(qopd, qopv.data(), nb * 4, hipMemcpyHostToDevice, st));
        k_qpos<<<grid_for(nb, 256), 256, 0, st>>>(I->inv, dq, nb, qp);
        GTS_HIP(hipGetLastError());
        std::vector<int64_t> hc(nb);
        range_count_dev(I, qp, qopd, nb, r, I->death.p, hc.data());
        const int64_t ne = nent();
        if (ne > 0) {
            // cache rows not yet on the device, then every entry's stamps
            if (uploaded < ne) {
                const int64_t k = ne - uploaded;
                if (is_str) {
                    uint8_t *d = dCs.ensure(ne * I->sstride);
                    int32_t *l = dCl.ensure(ne);
                    std::vector<uint8_t> pad((size_t)k * I->sstride, 0);
                    for (int64_t j = 0; j < k; j++)
                        memcpy(&pad[(size_t)j * I->sstride], &cstr[(size_t)(uploaded + j) * srow], clen[uploaded + j]);
                    GTS_HIP(hipMemcpyAsync(d + uploaded * I->sstride, pad.data(), pad.size(), hipMemcpyHostToDevice, st));
                    GTS_HIP(hipMemcpyAsync(l + uploaded, clen.data() + uploaded, k * 4, hipMemcpyHostToDevice, st));
                } else {
                    float *d = dC.ensure(ne * I->dp);
                    std::vector<float> pad((size_t)k * I->dp, 0.f);
                    for (int64_t j = 0; j < k; j++)
                        memcpy(&pad[(size_t)j * I->dp], &cvec[(size_t)(uploaded + j) * srow], dim * 4);
                    GTS_HIP(hipMemcpyAsync(d + uploaded * I->dp, pad.data(), pad.size() * 4, hipMemcpyHostToDevice, st));
                }
                uploaded = ne;
            }
            int32_t *b = dborn.ensure(ne), *dd = ddied.ensure(ne);
            GTS_HIP(hipMemcpyAsync(b, born.data(), ne * 4, hipMemcpyHostToDevice, st));
            GTS_HIP(hipMemcpyAsync(dd, died.data(), ne * 4, hipMemcpyHostToDevice, st));
            unsigned long long *cc = (unsigned long long *)I->counts.ensure(nb);
            GTS_HIP(hipMemsetAsync(cc, 0, nb * 8, st));
            DevTree t = I->dev();
            if (is_str) {
                k_cache_count_edit<<<grid_for(nb * ne, 256), 256, 0, st>>>(t, qp, qopd, nb, dCs.p, dCl.p, b, dd,
                                                                           (int)ne, (double)r, cc);
            } else {
                GTS_METRIC_SWITCH(I->metric, {
                    k_cache_count<M_><<<grid_for(nb * ne * 64, 256), 256, 0, st>>>(t, qp, qopd, nb, dC.p, b, dd,
                                                                                  (int)ne, (double)r, cc);
                });
            }
            GTS_HIP(hipGetLastError());
            std::vector<int64_t> add(nb);
            GTS_HIP(hipMemcpyAsync(add.data(), cc, nb * 8, hipMemcpyDeviceToHost, st));
            GTS_HIP(hipStreamSynchronize(st));
            for (int64_t i = 0; i < nb; i++) hc[i] += add[i];
        }
        for (int64_t i = 0; i < nb; i++) out[qslot[i]] = hc[i];
        qrow.clear();
        qopv.clear();
        qslot.clear();
        // drop removed cache entries: no pending query sees them any more
        if (ne > 0 && (int64_t)live_cache.size() < ne) {
            std::vector<float> nv;
            std::vector<uint8_t> ns;
            std::vector<int32_t> nl, nbn;
            for (int32_t e : live_cache) {
                if (is_str) {
                    ns.insert(ns.end(), cstr.begin() + (size_t)e * srow, cstr.begin() + (size_t)(e + 1) * srow);
                    nl.push_back(clen[e]);
                } else {
                    nv.insert(nv.end(), cvec.begin() + (size_t)e * srow, cvec.begin() + (size_t)(e + 1) * srow);
                }
                nbn.push_back(born[e]);
            }
            cvec.swap(nv);
            cstr.swap(ns);
            clen.swap(nl);
            born.swap(nbn);
            died.assign(born.size(), INT32_MAX);
            for (size_t j = 0; j < live_cache.size(); j++) live_cache[j] = (int32_t)j;
            uploaded = 0;
        }
        s_search += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    auto rebuild = [&]() {
        flush();
        dpos.clear();   // stamps of the old tree's positions (no query left to see them)
        dop.clear();
        gts_dataset nd;
        nd.metric = I->data.metric;
        nd.dim = I->data.dim;
        nd.stride = I->data.stride;
        const int64_t m = live + (int64_t)live_cache.size();
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
        for (int32_t e : live_cache) {
            if (is_str) {
                memcpy(&nd.str[(size_t)o * nd.stride], &cstr[(size_t)e * srow], clen[e]);
                nd.len[o] = clen[e];
            } else {
                memcpy(&nd.vec[(size_t)o * dim], &cvec[(size_t)e * srow], dim * 4);
            }
            o++;
        }
        I->data = std::move(nd);
        build_tree(I);
        tree_size = I->data.n;
        dead.assign(tree_size, 0);
        fw.init(tree_size);
        live = tree_size;
        cvec.clear(); cstr.clear(); clen.clear(); born.clear(); died.clear(); live_cache.clear();
        uploaded = 0;
        new_death();
        load_inv();
        I->stats.retries++;  // counts rebuilds
    };

    for (int64_t op = 0; op < nops; op++) {
        const int32_t fl = ops[op].flag, id = ops[op].id;
        if (fl == 2) {
            I->stats.update_queries++;
            if (id < 0 || id >= tree_size)
                fail(GTS_E_ARG, "query op " + std::to_string(op) + ": row " + std::to_string(id) + " out of range");
            qrow.push_back(id);
            qopv.push_back((int32_t)op);
            qslot.push_back((int64_t)out.size());
            out.push_back(0);
            if ((int64_t)qrow.size() == QBATCH) flush();
        } else if (fl == 0 || fl == 3) {
            I->stats.update_changes++;
            if (fl == 0) {
                if (id < 0 || id >= tree_size)
                    fail(GTS_E_ARG, "insert op " + std::to_string(op) + ": row " + std::to_string(id) + " out of range");
                if (is_str) {
                    cstr.resize((size_t)(nent() + 1) * srow, 0);
                    memcpy(&cstr[(size_t)nent() * srow], &I->data.str[(size_t)id * srow], I->data.len[id]);
                    clen.push_back(I->data.len[id]);
                } else {
                    cvec.insert(cvec.end(), I->data.vec.begin() + (size_t)id * dim, I->data.vec.begin() + (size_t)(id + 1) * dim);
                }
            } else {
                if (is_str || !fresh || id < 0 || id >= nfresh)
                    fail(GTS_E_ARG, "fresh insert op " + std::to_string(op) + " invalid");
                cvec.insert(cvec.end(), fresh + (size_t)id * dim, fresh + (size_t)(id + 1) * dim);
            }
            live_cache.push_back((int32_t)nent());
            born.push_back((int32_t)op);
            died.push_back(INT32_MAX);
        } else if (fl == 1) {
            I->stats.update_changes++;
            if (id < 0 || id >= live + (int64_t)live_cache.size())
                fail(GTS_E_ARG, "delete op " + std::to_string(op) + ": logical ID " + std::to_string(id) + " out of range");
            if (id < live) {
                const int64_t phys = fw.select(id);
                dead[phys] = 1;
                fw.add(phys, -1);
                live--;
                dpos.push_back(hinv[phys]);   // stamps are kept by tree position
                dop.push_back((int32_t)op);
            } else {
                const int64_t j = id - live;
                died[live_cache[j]] = (int32_t)op;
                live_cache.erase(live_cache.begin() + j);
            }
        } else {
            fail(GTS_E_ARG, "op " + std::to_string(op) + ": unknown flag " + std::to_string(fl));
        }
        // a full cache list triggers a rebuild (update.cuh:484-540), and so
        // does every period-th op before the last when a period is set
        const int64_t per = I->rebuild_period;
        if ((int64_t)live_cache.size() == cache_cap || (per > 0 && op + 1 < nops && (op + 1) % per == 0))
            rebuild();
    }
    flush();
    const double s_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    I->stats.ms_update_search = 1e3 * s_search;
    I->stats.ms_update_apply = 1e3 * std::max(0.0, s_total - s_search);
    I->stats.ms_total = 1e3 * s_total;
    const int64_t nout = (int64_t)out.size();
    if (counts_out)
        for (int64_t i = 0; i < nout && i < cap; i++) counts_out[i] = out[i];
    if (nqueries) *nqueries = nout;
    if (counts_out && nout > cap) fail(GTS_E_CAPACITY, "counts buffer too small");
}

}  // namespace gts

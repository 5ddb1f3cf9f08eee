// C ABI of the GTS index (include/gts.h): file formats, index lifecycle and
// the query drivers.  Every entry point converts internal exceptions into a
// status code and a thread-local message (the reference exits instead:
// tree.cuh:19-29, file.cuh:31-35, update.cuh:45-50).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <new>
#include <sstream>
#include <thread>

#include "gts_cells.h"
#include "gts_api_internal.h"
#include "gts_search.h"
#include "gts_edit.h"

using gts::Error;
using gts::fail;

static thread_local std::string g_err;

#define API_BEGIN try {
#define API_END                                   \
    }                                             \
    catch (const Error &e)                        \
    {                                             \
        g_err = e.msg;                            \
        return e.code;                            \
    }                                             \
    catch (const std::bad_alloc &)                \
    {                                             \
        g_err = "host allocation failed";         \
        return GTS_E_OOM;                         \
    }                                             \
    catch (const std::exception &e)               \
    {                                             \
        g_err = e.what();                         \
        return GTS_E_ARG;                         \
    }                                             \
    return GTS_OK;

extern "C" const char *gts_last_error(void) { return g_err.c_str(); }
extern "C" int gts_version(void) { return 100; }

// ------------------------------------------------------------------ files
static std::string read_file(const char *path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) fail(GTS_E_IO, std::string("cannot open ") + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// next line [p, e) of buf starting at *pos; returns false at the end
static bool next_line(const std::string &buf, size_t *pos, const char **p, size_t *len)
{
    if (*pos >= buf.size()) return false;
    size_t s = *pos, e = buf.find('\n', s);
    if (e == std::string::npos) e = buf.size();
    *p = buf.data() + s;
    *len = e - s;
    if (*len > 0 && (*p)[*len - 1] == '\r') (*len)--;
    *pos = e + 1;
    return true;
}

extern "C" int gts_dataset_load(const char *path, gts_dataset **out)
{
    API_BEGIN
    if (!path || !out) fail(GTS_E_ARG, "null argument");
    std::string buf = read_file(path);
    size_t pos = 0;
    const char *lp;
    size_t ll;
    if (!next_line(buf, &pos, &lp, &ll)) fail(GTS_E_FORMAT, "empty data file");
    std::string head(lp, ll);
    long long dim = -1, n = -1, metric = -1;
    if (sscanf(head.c_str(), "%lld %lld %lld", &dim, &n, &metric) != 3 || n < 1 || dim < 0)
        fail(GTS_E_FORMAT, "bad header (want: dim N metric): " + head);
    if (metric != 0 && metric != 1 && metric != 2 && metric != 5 && metric != 6)
        fail(GTS_E_FORMAT, "unknown metric code " + std::to_string(metric));
    auto *ds = new gts_dataset();
    ds->n = n;
    ds->metric = (int)metric;
    try {
        if (metric == 6) {
            std::vector<std::string> rows;
            rows.reserve(n);
            size_t mx = 1;
            for (long long i = 0; i < n; i++) {
                if (!next_line(buf, &pos, &lp, &ll)) fail(GTS_E_FORMAT, "fewer strings than the header says");
                size_t L = strnlen(lp, ll);  // the reference keeps strlen() of the line
                if (L > GTS_EDIT_MAXLEN) fail(GTS_E_FORMAT, "string longer than 255 bytes");
                rows.emplace_back(lp, L);
                mx = std::max(mx, L);
            }
            ds->stride = (int)mx;
            ds->dim = (int)mx;
            ds->str.assign((size_t)n * mx, 0);
            ds->len.resize(n);
            for (long long i = 0; i < n; i++) {
                memcpy(&ds->str[(size_t)i * mx], rows[i].data(), rows[i].size());
                ds->len[i] = (int32_t)rows[i].size();
            }
        } else {
            if (dim < 1) fail(GTS_E_FORMAT, "dimension must be positive");
            ds->dim = (int)dim;
            ds->vec.resize((size_t)n * dim);
            for (long long i = 0; i < n; i++) {
                if (!next_line(buf, &pos, &lp, &ll)) fail(GTS_E_FORMAT, "fewer rows than the header says");
                std::string line(lp, ll);
                const char *c = line.c_str();
                char *end;
                for (long long j = 0; j < dim; j++) {
                    float v = strtof(c, &end);
                    if (end == c) fail(GTS_E_FORMAT, "row " + std::to_string(i) + " has fewer than dim values");
                    ds->vec[(size_t)i * dim + j] = v;
                    c = end;
                }
            }
        }
    } catch (...) {
        delete ds;
        throw;
    }
    *out = ds;
    API_END
}

extern "C" int gts_dataset_from_vectors(const float *rows, int64_t n, int dim, int metric,
                                        gts_dataset **out)
{
    API_BEGIN
    if (!rows || !out || n < 1 || dim < 1) fail(GTS_E_ARG, "bad dataset arguments");
    if (metric != 0 && metric != 1 && metric != 2 && metric != 5)
        fail(GTS_E_ARG, "vector metric must be 0, 1, 2 or 5");
    auto *ds = new gts_dataset();
    ds->n = n;
    ds->dim = dim;
    ds->metric = metric;
    ds->vec.assign(rows, rows + (size_t)n * dim);
    *out = ds;
    API_END
}

extern "C" int gts_dataset_from_strings(const uint8_t *chars, const int32_t *lens, int64_t n,
                                        int stride, gts_dataset **out)
{
    API_BEGIN
    if (!chars || !lens || !out || n < 1 || stride < 1) fail(GTS_E_ARG, "bad dataset arguments");
    auto *ds = new gts_dataset();
    ds->n = n;
    ds->metric = GTS_METRIC_EDIT;
    int mx = 1;
    for (int64_t i = 0; i < n; i++) {
        if (lens[i] < 0 || lens[i] > stride || lens[i] > GTS_EDIT_MAXLEN) {
            delete ds;
            fail(GTS_E_ARG, "string length out of range");
        }
        mx = std::max(mx, (int)lens[i]);
    }
    ds->stride = mx;
    ds->dim = mx;
    ds->str.assign((size_t)n * mx, 0);
    ds->len.assign(lens, lens + n);
    for (int64_t i = 0; i < n; i++) memcpy(&ds->str[(size_t)i * mx], chars + (size_t)i * stride, lens[i]);
    *out = ds;
    API_END
}

extern "C" int gts_dataset_info(const gts_dataset *ds, int64_t *n, int *dim, int *metric)
{
    API_BEGIN
    if (!ds) fail(GTS_E_ARG, "null dataset");
    if (n) *n = ds->n;
    if (dim) *dim = ds->dim;
    if (metric) *metric = ds->metric;
    API_END
}

extern "C" int gts_dataset_rows(const gts_dataset *ds, float *out, int64_t cap)
{
    API_BEGIN
    if (!ds || !out) fail(GTS_E_ARG, "null argument");
    if (ds->metric == GTS_METRIC_EDIT) fail(GTS_E_ARG, "string dataset has no rows");
    if (cap < (int64_t)ds->vec.size()) fail(GTS_E_CAPACITY, "buffer too small");
    memcpy(out, ds->vec.data(), ds->vec.size() * sizeof(float));
    API_END
}

extern "C" void gts_dataset_free(gts_dataset *ds) { delete ds; }

extern "C" int gts_load_queries(const char *path, int32_t *ids, int64_t cap, int64_t *nq)
{
    API_BEGIN
    if (!path || !nq) fail(GTS_E_ARG, "null argument");
    std::string buf = read_file(path);
    size_t pos = 0;
    const char *lp;
    size_t ll;
    if (!next_line(buf, &pos, &lp, &ll)) fail(GTS_E_FORMAT, "empty query file");
    long long cnt = -1;
    if (sscanf(std::string(lp, ll).c_str(), "%lld", &cnt) != 1 || cnt < 0)
        fail(GTS_E_FORMAT, "bad query count");
    *nq = cnt;
    for (long long i = 0; i < cnt; i++) {
        if (!next_line(buf, &pos, &lp, &ll)) fail(GTS_E_FORMAT, "fewer query IDs than the count");
        long long v;
        if (sscanf(std::string(lp, ll).c_str(), "%lld", &v) != 1) fail(GTS_E_FORMAT, "bad query ID");
        if (ids && i < cap) ids[i] = (int32_t)v;
    }
    if (ids && cap < cnt) fail(GTS_E_CAPACITY, "query buffer too small");
    API_END
}

extern "C" int gts_load_updates(const char *path, gts_update_op *ops, int64_t cap, int64_t *nops)
{
    API_BEGIN
    if (!path || !nops) fail(GTS_E_ARG, "null argument");
    std::string buf = read_file(path);
    size_t pos = 0;
    const char *lp;
    size_t ll;
    if (!next_line(buf, &pos, &lp, &ll)) fail(GTS_E_FORMAT, "empty update file");
    long long cnt = -1;
    if (sscanf(std::string(lp, ll).c_str(), "%lld", &cnt) != 1 || cnt < 0)
        fail(GTS_E_FORMAT, "bad update count");
    *nops = cnt;
    for (long long i = 0; i < cnt; i++) {
        if (!next_line(buf, &pos, &lp, &ll)) fail(GTS_E_FORMAT, "fewer update ops than the count");
        int f, v;
        if (sscanf(std::string(lp, ll).c_str(), "%d %d", &f, &v) != 2) fail(GTS_E_FORMAT, "bad update op");
        if (ops && i < cap) ops[i] = gts_update_op{f, v};
    }
    if (ops && cap < cnt) fail(GTS_E_CAPACITY, "update buffer too small");
    API_END
}

// ------------------------------------------------------------------ index
extern "C" int gts_build(const gts_dataset *ds, const gts_build_params *params, gts_index **out)
{
    gts_index *idx = nullptr;
    API_BEGIN
    if (!ds || !out) fail(GTS_E_ARG, "null argument");
    gts_build_params p = params ? *params : gts_build_params{10, 20, 0, 0};
    if (p.order < 2 || p.order > 32) fail(GTS_E_ARG, "order must be in [2, 32]");
    if (p.max_size < 1) fail(GTS_E_ARG, "max_size must be >= 1");
    if (p.max_h < 0) fail(GTS_E_ARG, "max_h must be >= 0");
    if (ds->metric == GTS_METRIC_ANGULAR) {
        // the angle of a zero vector is undefined (the reference prints "Error!!!", tree.cuh:116-119)
        for (int64_t i = 0; i < ds->n; i++) {
            double s = 0;
            for (int j = 0; j < ds->dim; j++) s += (double)ds->vec[i * ds->dim + j] * ds->vec[i * ds->dim + j];
            if (!(s > 0)) fail(GTS_E_ARG, "angular metric: object with a zero vector");
        }
    }
    if (ds->n > (int64_t)INT32_MAX - 1) fail(GTS_E_ARG, "too many objects");
    int ndev = 0;
    GTS_HIP(hipGetDeviceCount(&ndev));
    if (p.device < 0 || p.device >= ndev) fail(GTS_E_ARG, "no such device");
    GTS_HIP(hipSetDevice(p.device));
    idx = new gts_index();
    idx->device = p.device;
    idx->T = p.order;
    idx->max_size = p.max_size;
    idx->max_h = p.max_h;
    GTS_HIP(hipStreamCreateWithFlags(&idx->st, hipStreamNonBlocking));
    {
        // the index's own host copy of the objects (kept for rebuilds on
        // update) is made on a host thread while the device builds from the
        // caller's dataset
        std::thread copy([idx, ds] { idx->data = *ds; });
        try {
            gts::build_tree(idx, ds);
        } catch (...) {
            copy.join();
            throw;
        }
        copy.join();
    }
    *out = idx;
    idx = nullptr;
    }
    catch (const Error &e)
    {
        g_err = e.msg;
        if (idx) gts_index_free(idx);
        return e.code;
    }
    catch (const std::bad_alloc &)
    {
        g_err = "host allocation failed";
        if (idx) gts_index_free(idx);
        return GTS_E_OOM;
    }
    return GTS_OK;
}

extern "C" void gts_index_free(gts_index *idx)
{
    if (!idx) return;
    (void)hipSetDevice(idx->device);
    if (idx->st) (void)hipStreamSynchronize(idx->st);
    gts::release_tree(idx);
    gts::release_scratch(idx);
    gts::graph_reset(idx);
    for (hipEvent_t &e : idx->graph_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t *e : {&idx->call_ev[0], &idx->call_ev[1], &idx->timing_ev[0], &idx->timing_ev[1]})
        if (*e) (void)hipEventDestroy(*e);
    if (idx->hpin) (void)hipHostFree(idx->hpin);
    if (idx->st) {
        gts::prims_release_stream(idx->st);
        (void)hipStreamDestroy(idx->st);
    }
    delete idx;
}

extern "C" int gts_index_info(const gts_index *idx, int *height, int64_t *num_nodes, int64_t *n)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    if (height) *height = idx->height;
    if (num_nodes) *num_nodes = idx->num_nodes;
    if (n) *n = idx->n;
    API_END
}

extern "C" int gts_index_export(const gts_index *idx, int32_t *perm, int32_t *lid, int32_t *size,
                                int32_t *pivot, float *min_dis, int32_t *flags)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    GTS_HIP(hipSetDevice(idx->device));
    hipStream_t st = idx->st;
    const int64_t nn = idx->num_nodes, n = idx->n;
    if (perm) GTS_HIP(hipMemcpyAsync(perm, idx->perm, n * 4, hipMemcpyDeviceToHost, st));
    if (lid) GTS_HIP(hipMemcpyAsync(lid, idx->nlid, nn * 4, hipMemcpyDeviceToHost, st));
    if (size) GTS_HIP(hipMemcpyAsync(size, idx->nsize, nn * 4, hipMemcpyDeviceToHost, st));
    if (min_dis) GTS_HIP(hipMemcpyAsync(min_dis, idx->nmin, nn * 4, hipMemcpyDeviceToHost, st));
    std::vector<int32_t> pp(pivot ? nn : 0);
    std::vector<uint8_t> fl(flags ? nn : 0);
    if (pivot) GTS_HIP(hipMemcpyAsync(pp.data(), idx->npiv, nn * 4, hipMemcpyDeviceToHost, st));
    if (flags) GTS_HIP(hipMemcpyAsync(fl.data(), idx->nflag, nn, hipMemcpyDeviceToHost, st));
    std::vector<int32_t> pm(pivot ? n : 0);
    if (pivot) GTS_HIP(hipMemcpyAsync(pm.data(), idx->perm, n * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    if (pivot)
        for (int64_t i = 0; i < nn; i++) pivot[i] = pp[i] >= 0 ? pm[pp[i]] : -1;
    if (flags)
        for (int64_t i = 0; i < nn; i++) flags[i] = fl[i];
    API_END
}

extern "C" int gts_last_stats(const gts_index *idx, gts_stats *out)
{
    API_BEGIN
    if (!idx || !out) fail(GTS_E_ARG, "null argument");
    // the last kNN call's event timings are read now (its events are complete)
    if (idx->pend_timing) {
        gts_index *m = const_cast<gts_index *>(idx);
        GTS_HIP(hipSetDevice(m->device));
        gts::finalize_timing(m);
    }
    *out = idx->stats;
    API_END
}

// ------------------------------------------------------------------ queries
// k is at most the per-query candidate capacity of the top-k (GTS_MAX_K)
static void check_k_cap(int64_t k)
{
    if (k > gts::topk_cap_limit())
        fail(GTS_E_ARG, "k above " + std::to_string(gts::topk_cap_limit()) + " (GTS_MAX_K)");
}

static void check_ids_host(const gts_index *idx, const int32_t *qids, int64_t nq)
{
    for (int64_t i = 0; i < nq; i++)
        if (qids[i] < 0 || qids[i] >= idx->n)
            fail(GTS_E_ARG, "query ID " + std::to_string(qids[i]) + " out of range");
}

extern "C" int gts_range_query(gts_index *idx, const int32_t *qids, int64_t nq, float r,
                               int64_t *counts, int64_t *offsets, int32_t *ids, float *dists,
                               int64_t cap)
{
    API_BEGIN
    if (!idx || (nq > 0 && (!qids || !counts))) fail(GTS_E_ARG, "null argument");
    if (nq < 0) fail(GTS_E_ARG, "negative query count");
    if (ids && !offsets) fail(GTS_E_ARG, "offsets required with ids");
    if (nq == 0) return GTS_OK;
    check_ids_host(idx, qids, nq);
    GTS_HIP(hipSetDevice(idx->device));
    gts::RangeOut ro;
    gts::range_run(idx, qids, nq, r, ids != nullptr, &ro);
    memcpy(counts, ro.counts.data(), nq * sizeof(int64_t));
    if (ids) {
        offsets[0] = 0;
        for (int64_t i = 0; i < nq; i++) offsets[i + 1] = offsets[i] + ro.counts[i];
        if (offsets[nq] > cap) fail(GTS_E_CAPACITY, "result buffer holds " + std::to_string(cap) +
                                                        " pairs, " + std::to_string(offsets[nq]) + " needed");
        memcpy(ids, ro.ids.data(), offsets[nq] * sizeof(int32_t));
        if (dists) memcpy(dists, ro.dists.data(), offsets[nq] * sizeof(float));
    }
    API_END
}

extern "C" int gts_knn_query_var(gts_index *idx, const int32_t *qids, int64_t nq, const int32_t *ks,
                                 int kstride, int32_t *ids, float *dists, float *kth)
{
    API_BEGIN
    if (!idx || (nq > 0 && (!qids || !ks))) fail(GTS_E_ARG, "null argument");
    if (nq <= 0) return nq == 0 ? GTS_OK : GTS_E_ARG;
    check_ids_host(idx, qids, nq);
    int kmax = 0;
    for (int64_t i = 0; i < nq; i++) {
        if (ks[i] < 1) fail(GTS_E_ARG, "k must be >= 1");
        kmax = std::max(kmax, ks[i]);
    }
    if (kstride < kmax) fail(GTS_E_ARG, "kstride below the largest k");
    check_k_cap(kmax);
    GTS_HIP(hipSetDevice(idx->device));
    hipStream_t st = idx->st;
    int32_t *dq = idx->dqid.ensure(nq), *dk = idx->dks.ensure(nq);
    int32_t *di = idx->dids.ensure((size_t)nq * kstride);
    float *dd = idx->ddist.ensure((size_t)nq * kstride), *dt = idx->dkth.ensure(nq);
    GTS_HIP(hipMemcpyAsync(dq, qids, nq * 4, hipMemcpyHostToDevice, st));
    GTS_HIP(hipMemcpyAsync(dk, ks, nq * 4, hipMemcpyHostToDevice, st));
    gts::knn_run(idx, dq, nq, dk, 0, kstride, di, dd, dt, 0, GTS_VERIFY_FP32);
    if (ids) GTS_HIP(hipMemcpyAsync(ids, di, (size_t)nq * kstride * 4, hipMemcpyDeviceToHost, st));
    if (dists) GTS_HIP(hipMemcpyAsync(dists, dd, (size_t)nq * kstride * 4, hipMemcpyDeviceToHost, st));
    if (kth) GTS_HIP(hipMemcpyAsync(kth, dt, nq * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    API_END
}

extern "C" int gts_knn_query(gts_index *idx, const int32_t *qids, int64_t nq, int k, int32_t *ids,
                             float *dists, float *kth)
{
    if (k < 1) {
        g_err = "k must be >= 1";
        return GTS_E_ARG;
    }
    std::vector<int32_t> ks(nq > 0 ? nq : 0, k);
    return gts_knn_query_var(idx, qids, nq, ks.data(), k, ids, dists, kth);
}

extern "C" int gts_knn_query_dev(gts_index *idx, const int32_t *d_qids, int64_t nq,
                                 const int32_t *d_ks, int kfix, int kstride, int32_t *d_ids,
                                 float *d_dists, float *d_kth, int learned, int verify)
{
    API_BEGIN
    if (!idx || nq < 0) fail(GTS_E_ARG, "bad argument");
    if (nq == 0) return GTS_OK;
    if (!d_ks && kfix < 1) fail(GTS_E_ARG, "k must be >= 1");
    if (kstride < 1 || (!d_ks && kstride < kfix)) fail(GTS_E_ARG, "kstride below k");
    if (!d_ks) check_k_cap(kfix);
    if (!d_qids || !d_ids || !d_dists || !d_kth) fail(GTS_E_ARG, "null argument");
    if (verify != GTS_VERIFY_FP32 && verify != GTS_VERIFY_BF16X3) fail(GTS_E_ARG, "unknown verification form");
    if (learned && idx->radius_src == GTS_RADIUS_NETWORK && !idx->model.trained)
        fail(GTS_E_ARG, "the radius network is not trained");
    GTS_HIP(hipSetDevice(idx->device));
    gts::knn_run(idx, d_qids, nq, d_ks, kfix, kstride, d_ids, d_dists, d_kth, learned, verify);
    API_END
}

extern "C" int gts_update(gts_index *idx, const gts_update_op *ops, int64_t nops, float r,
                          int cache_cap, const float *fresh, int64_t nfresh, int64_t *counts,
                          int64_t cap, int64_t *nqueries)
{
    API_BEGIN
    if (!idx || (nops > 0 && !ops)) fail(GTS_E_ARG, "null argument");
    if (cache_cap < 1) fail(GTS_E_ARG, "cache capacity must be >= 1");
    GTS_HIP(hipSetDevice(idx->device));
    gts::update_run(idx, ops, nops, r, cache_cap, fresh, nfresh, counts, cap, nqueries);
    API_END
}

extern "C" int gts_learned_train(gts_index *idx, const int32_t *train_qids, int64_t ntrain, int kmax,
                                 int epochs, float lr, uint64_t seed, float *loss)
{
    API_BEGIN
    if (!idx || !train_qids || ntrain < 1 || kmax < 1 || epochs < 1) fail(GTS_E_ARG, "bad argument");
    check_ids_host(idx, train_qids, ntrain);
    GTS_HIP(hipSetDevice(idx->device));
    float l = gts::learned_train(idx, train_qids, ntrain, kmax, epochs, lr, seed);
    if (loss) *loss = l;
    API_END
}

extern "C" int gts_mlp_num_params(void) { return gts::mlp_num_params(); }

extern "C" int gts_mlp_loss_grad(gts_index *idx, const float *X, const float *t, int64_t nb,
                                 const float *w, float *grad, float *loss)
{
    API_BEGIN
    if (!idx || !X || !t || !w || !grad || nb < 1) fail(GTS_E_ARG, "bad argument");
    GTS_HIP(hipSetDevice(idx->device));
    float l = gts::mlp_loss_grad(idx, X, t, nb, w, grad);
    if (loss) *loss = l;
    API_END
}

// result writers of file.cuh (saveK, saveR): "%d " / "%f " per entry, one
// line per query
static void save_rows(const char *path, const char *mode, int64_t nq, const std::function<int64_t(int64_t)> &len,
                      const std::function<std::string(int64_t, int64_t)> &item)
{
    if (!path) return;
    FILE *f = fopen(path, mode);
    if (!f) fail(GTS_E_IO, std::string("cannot open ") + path);
    std::string line;
    for (int64_t i = 0; i < nq; i++) {
        line.clear();
        for (int64_t j = 0, m = len(i); j < m; j++) line += item(i, j);
        line += '\n';
        if (fwrite(line.data(), 1, line.size(), f) != line.size()) {
            fclose(f);
            fail(GTS_E_IO, std::string("write failed: ") + path);
        }
    }
    if (fclose(f) != 0) fail(GTS_E_IO, std::string("write failed: ") + path);
}

static std::string fmt_id(int32_t v)
{
    char b[24];
    snprintf(b, sizeof b, "%d ", v);
    return b;
}

static std::string fmt_dist(float v)
{
    char b[64];
    snprintf(b, sizeof b, "%f ", v);
    return b;
}

extern "C" int gts_save_knn(const char *ids_path, const char *dists_path, const int32_t *ids, const float *dists,
                            int k, int64_t nq)
{
    API_BEGIN
    if (k < 1 || nq < 0 || (ids_path && !ids) || (dists_path && !dists)) fail(GTS_E_ARG, "bad argument");
    auto len = [&](int64_t) { return (int64_t)k; };
    save_rows(ids_path, "w", nq, len, [&](int64_t i, int64_t j) { return fmt_id(ids[i * k + j]); });
    save_rows(dists_path, "w", nq, len, [&](int64_t i, int64_t j) { return fmt_dist(dists[i * k + j]); });
    API_END
}

extern "C" int gts_save_range(const char *ids_path, const char *dists_path, const int32_t *ids, const float *dists,
                              const int64_t *counts, const int64_t *offsets, int64_t nq)
{
    API_BEGIN
    if (nq < 0 || (nq > 0 && (!counts || !offsets)) || (ids_path && !ids) || (dists_path && !dists))
        fail(GTS_E_ARG, "bad argument");
    auto len = [&](int64_t i) { return counts[i]; };
    save_rows(ids_path, "a", nq, len, [&](int64_t i, int64_t j) { return fmt_id(ids[offsets[i] + j]); });
    save_rows(dists_path, "a", nq, len, [&](int64_t i, int64_t j) { return fmt_dist(dists[offsets[i] + j]); });
    API_END
}

extern "C" int gts_mlp_loss_grad_noalloc(gts_index *idx, const float *d_X, const float *d_t, int64_t nb,
                                         const float *d_w, float *d_grad, float *d_loss)
{
    API_BEGIN
    if (!idx || !d_X || !d_t || !d_w || !d_grad || !d_loss) fail(GTS_E_ARG, "null argument");
    GTS_HIP(hipSetDevice(idx->device));
    gts::mlp_loss_grad_noalloc(idx, d_X, d_t, nb, d_w, d_grad, d_loss);
    API_END
}

extern "C" int gts_knn_query_ex(gts_index *idx, const int32_t *qids, int64_t nq, const int32_t *ks,
                                int kfix, int kstride, int32_t *ids, float *dists, float *kth,
                                int learned, int verify)
{
    API_BEGIN
    if (!idx || nq < 0 || (nq > 0 && !qids)) fail(GTS_E_ARG, "bad argument");
    if (nq == 0) return GTS_OK;
    check_ids_host(idx, qids, nq);
    int kmax = kfix;
    if (ks) {
        kmax = 0;
        for (int64_t i = 0; i < nq; i++) {
            if (ks[i] < 1) fail(GTS_E_ARG, "k must be >= 1");
            kmax = std::max(kmax, ks[i]);
        }
    } else if (kfix < 1) {
        fail(GTS_E_ARG, "k must be >= 1");
    }
    if (kstride < kmax) fail(GTS_E_ARG, "kstride below the largest k");
    check_k_cap(kmax);
    if (verify != GTS_VERIFY_FP32 && verify != GTS_VERIFY_BF16X3) fail(GTS_E_ARG, "unknown verification form");
    if (learned && idx->radius_src == GTS_RADIUS_NETWORK && !idx->model.trained)
        fail(GTS_E_ARG, "the radius network is not trained");
    GTS_HIP(hipSetDevice(idx->device));
    hipStream_t st = idx->st;
    int32_t *dq = idx->dqid.ensure(nq), *dk = ks ? idx->dks.ensure(nq) : nullptr;
    int32_t *di = idx->dids.ensure((size_t)nq * kstride);
    float *dd = idx->ddist.ensure((size_t)nq * kstride), *dt = idx->dkth.ensure(nq);
    GTS_HIP(hipMemcpyAsync(dq, qids, nq * 4, hipMemcpyHostToDevice, st));
    if (ks) GTS_HIP(hipMemcpyAsync(dk, ks, nq * 4, hipMemcpyHostToDevice, st));
    gts::knn_run(idx, dq, nq, dk, kfix, kstride, di, dd, dt, learned, verify);
    if (ids) GTS_HIP(hipMemcpyAsync(ids, di, (size_t)nq * kstride * 4, hipMemcpyDeviceToHost, st));
    if (dists) GTS_HIP(hipMemcpyAsync(dists, dd, (size_t)nq * kstride * 4, hipMemcpyDeviceToHost, st));
    if (kth) GTS_HIP(hipMemcpyAsync(kth, dt, nq * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    API_END
}

extern "C" int gts_device_alloc(gts_index *idx, int64_t bytes, void **ptr)
{
    API_BEGIN
    if (!idx || !ptr || bytes < 0) fail(GTS_E_ARG, "bad argument");
    GTS_HIP(hipSetDevice(idx->device));
    if (hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 1) != hipSuccess) fail(GTS_E_OOM, "device allocation");
    API_END
}

extern "C" int gts_device_free(gts_index *idx, void *ptr)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    GTS_HIP(hipSetDevice(idx->device));
    GTS_HIP(hipStreamSynchronize(idx->st));
    if (ptr) GTS_HIP(hipFree(ptr));
    API_END
}

extern "C" int gts_copy_to_device(gts_index *idx, void *dst, const void *src, int64_t bytes)
{
    API_BEGIN
    if (!idx || !dst || !src || bytes < 0) fail(GTS_E_ARG, "bad argument");
    GTS_HIP(hipSetDevice(idx->device));
    GTS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, idx->st));
    GTS_HIP(hipStreamSynchronize(idx->st));
    API_END
}

extern "C" int gts_copy_to_host(gts_index *idx, void *dst, const void *src, int64_t bytes)
{
    API_BEGIN
    if (!idx || !dst || !src || bytes < 0) fail(GTS_E_ARG, "bad argument");
    GTS_HIP(hipSetDevice(idx->device));
    GTS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, idx->st));
    GTS_HIP(hipStreamSynchronize(idx->st));
    API_END
}

extern "C" int gts_copy_device(gts_index *idx, void *dst, const void *src, int64_t bytes)
{
    API_BEGIN
    if (!idx || !dst || !src || bytes < 0) fail(GTS_E_ARG, "bad argument");
    GTS_HIP(hipSetDevice(idx->device));
    GTS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, idx->st));
    GTS_HIP(hipStreamSynchronize(idx->st));
    API_END
}

extern "C" int gts_set_knn_plan(gts_index *idx, int plan)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    if (plan != GTS_PLAN_AUTO && plan != GTS_PLAN_PRUNED && plan != GTS_PLAN_EXHAUSTIVE && plan != GTS_PLAN_CELLS &&
        plan != GTS_PLAN_PERSISTENT)
        fail(GTS_E_ARG, "unknown kNN plan");
    idx->knn_plan = plan;
    gts::graph_reset(idx);
    API_END
}

extern "C" int gts_set_fp32_mfma(gts_index *idx, int on)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    idx->fp32_mfma = on < 0 ? -1 : (on != 0 ? 1 : 0);
    API_END
}

extern "C" int gts_set_rebuild_period(gts_index *idx, int64_t nops)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    if (nops < 0) fail(GTS_E_ARG, "negative rebuild period");
    idx->rebuild_period = nops;
    API_END
}

extern "C" int gts_set_memory_budget(gts_index *idx, int64_t bytes)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    if (bytes < 0) fail(GTS_E_ARG, "negative memory budget");
    idx->mem_budget = bytes;
    API_END
}

extern "C" int gts_set_radius_source(gts_index *idx, int source)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    if (source != GTS_RADIUS_NETWORK && source != GTS_RADIUS_PIVOTS) fail(GTS_E_ARG, "unknown radius source");
    if (source == GTS_RADIUS_PIVOTS && idx->metric == GTS_METRIC_EDIT)
        fail(GTS_E_ARG, "the pivot radius needs a vector metric");
    idx->radius_src = source;
    idx->plan_memo.clear();
    gts::graph_reset(idx);
    API_END
}

extern "C" int gts_set_cell_count(gts_index *idx, int ncell)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    if (ncell < 0) fail(GTS_E_ARG, "cell count must be >= 0");
    GTS_HIP(hipSetDevice(idx->device));
    GTS_HIP(hipStreamSynchronize(idx->st));
    gts::cells_release(idx);
    idx->cells.want = ncell;
    gts::graph_reset(idx);
    API_END
}

extern "C" int gts_set_radius_quantile(gts_index *idx, double q)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    if (!(q > 0.0 && q <= 1.0)) fail(GTS_E_ARG, "quantile must be in (0, 1]");
    idx->model.quantile = (float)q;
    gts::learned_apply_quantile(idx->model);
    gts::graph_reset(idx);
    API_END
}

extern "C" int gts_set_autotune(gts_index *idx, int on)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    idx->autotune = on != 0;
    gts::graph_reset(idx);
    API_END
}

extern "C" int gts_set_dense_schedule(gts_index *idx, int persistent, int rows)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    if (rows < 0 || rows % 32) fail(GTS_E_ARG, "rows must be a non-negative multiple of 32");
    idx->forced_sched = gts::DenseSched{persistent != 0, rows};
    gts::graph_reset(idx);
    API_END
}

extern "C" int gts_set_graph(gts_index *idx, int on)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    idx->use_graph = on != 0;
    gts::graph_reset(idx);
    API_END
}

extern "C" int gts_set_timing(gts_index *idx, int on)
{
    API_BEGIN
    if (!idx) fail(GTS_E_ARG, "null index");
    idx->timing = on != 0;   // part of every graph key: graphs of either setting stay cached
    API_END
}

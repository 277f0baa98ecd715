// serve.cpp -- a continuous-batching request loop over the batch engine (no
// counterpart in the reference, whose callers invoke one GpuParser method at
// a time; crates/oneline-gpu/src/parser.rs:15-19 serialises them).
//
// Requests (parse, or a dictionary search against one of the resident
// dictionaries) queue in arrival order.  Each of `slots` worker threads owns a
// context (its own HIP stream and scratch) and repeatedly takes the oldest
// pending requests that belong together -- parse requests, or searches with
// the same dictionary -- as far as the HBM budget, the batch limits and the
// admission order allow, runs them as one batch (ol_batch_*), copies each
// request's results to the host and completes it.  Requests that arrive while
// other batches are on the device join the next batch a slot forms, so
// several batches overlap on the GPU; a request larger than the whole budget
// is refused at submission.  Dictionaries are compiled once and stay in HBM.
#include "../../include/oneline_hip.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

double now_s()
{
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    int kind = 0;            // OL_REQ_PARSE / OL_REQ_SEARCH
    uint32_t dict = 0;
    const uint8_t *doc = nullptr;
    size_t len = 0;
    size_t footprint = 0;    // HBM bytes charged while in flight
    bool keep = true;        // keep the results for ol_server_result_*
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = 1;          // 1 pending, 0 done, < 0 error
    std::string error;
    std::vector<uint32_t> offsets, parent, se, fc, ns, m_off, m_id;
    std::vector<uint8_t> classes;
    uint64_t out_bytes = 0;  // result bytes produced
};

// Device bytes a request occupies while its batch runs: the text (plus
// padding), per-line outputs (offset, class, key, four links: 22 bytes for a
// line of at least one byte, bounded by the text) and match lists.
size_t footprint_of(size_t len, int kind)
{
    const size_t lines = len / 16 + 64;
    return len + 64 * 1024 + lines * 22 + (kind == OL_REQ_SEARCH ? lines * 24 : 0);
}

}  // namespace

struct ol_server {
    ol_server_config cfg{};
    std::vector<ol_ctx *> ctx;
    std::vector<std::thread> workers;
    std::vector<ol_dict *> dicts;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<uint64_t, Request *> reqs;
    size_t in_flight = 0;
    uint64_t next_id = 1;
    bool stop = false;
    std::atomic<uint64_t> batches{0}, batched_requests{0};
};

namespace {

thread_local std::string s_err;

int sfail(int code, const std::string &msg)
{
    s_err = msg;
    return code;
}

// Takes the next batch: the oldest pending request and the following ones of
// the same kind and dictionary, in order, while the budget and limits allow;
// a request that does not fit stops the batch (no overtaking, so admission
// follows arrival order).
std::vector<Request *> take_batch(ol_server *s)
{
    std::vector<Request *> out;
    if (s->pending.empty())
        return out;
    Request *head = s->pending.front();
    if (s->in_flight + head->footprint > s->cfg.hbm_budget)
        return out;
    size_t bytes = 0;
    auto it = s->pending.begin();
    while (it != s->pending.end() && out.size() < s->cfg.max_batch_docs) {
        Request *r = *it;
        if (r->kind != head->kind || (r->kind == OL_REQ_SEARCH && r->dict != head->dict))
            break;
        if (!out.empty() && (bytes + r->len > s->cfg.max_batch_bytes ||
                             s->in_flight + r->footprint > s->cfg.hbm_budget))
            break;
        s->in_flight += r->footprint;
        bytes += r->len;
        out.push_back(r);
        it = s->pending.erase(it);
    }
    return out;
}

void run_batch(ol_server *s, ol_ctx *c, std::vector<Request *> &batch)
{
    const double t0 = now_s();
    for (Request *r : batch)
        r->t_start = t0;
    std::vector<const uint8_t *> ptrs(batch.size());
    std::vector<size_t> lens(batch.size());
    for (size_t i = 0; i < batch.size(); i++) {
        ptrs[i] = batch[i]->doc;
        lens[i] = batch[i]->len;
    }
    const int kind = batch[0]->kind;
    const ol_dict *dict = kind == OL_REQ_SEARCH ? s->dicts[batch[0]->dict] : nullptr;
    ol_batch *b = nullptr;
    int rc = ol_batch_create(c, ptrs.data(), lens.data(), batch.size(), &b);
    if (!rc)
        rc = ol_batch_run(c, b, dict, kind == OL_REQ_SEARCH ? OL_BATCH_SEARCH_ONLY : 0);
    std::vector<uint32_t> st, cnt;
    if (!rc && kind == OL_REQ_PARSE) {
        st.resize(batch.size());
        cnt.resize(batch.size());
        rc = ol_batch_lines(c, b, nullptr, st.data(), cnt.data());
    }
    const std::string err = rc ? std::string(ol_last_error()) : std::string();
    for (size_t i = 0; i < batch.size() && !rc; i++) {
        Request *r = batch[i];
        if (kind == OL_REQ_PARSE) {
            const size_t n = cnt[i];
            r->offsets.resize(n);
            r->classes.resize(n);
            r->parent.resize(n);
            r->se.resize(n);
            r->fc.resize(n);
            r->ns.resize(n);
            rc = ol_batch_download_doc(c, b, i, r->offsets.data(), r->classes.data(), r->parent.data(),
                                       r->se.data(), r->fc.data(), r->ns.data());
            r->out_bytes = (uint64_t)n * 21;
        } else {
            uint64_t m = 0;
            rc = ol_batch_matches(c, b, i, &m, nullptr, nullptr, 0);
            if (!rc) {
                r->m_off.resize(m);
                r->m_id.resize(m);
                rc = ol_batch_matches(c, b, i, &m, r->m_off.data(), r->m_id.data(), m);
            }
            r->out_bytes = m * 8;
        }
        if (!r->keep) {
            std::vector<uint32_t>().swap(r->offsets);
            std::vector<uint8_t>().swap(r->classes);
            std::vector<uint32_t>().swap(r->parent);
            std::vector<uint32_t>().swap(r->se);
            std::vector<uint32_t>().swap(r->fc);
            std::vector<uint32_t>().swap(r->ns);
            std::vector<uint32_t>().swap(r->m_off);
            std::vector<uint32_t>().swap(r->m_id);
        }
    }
    const std::string err2 = (rc && err.empty()) ? std::string(ol_last_error()) : err;
    ol_batch_free(b);
    const double t1 = now_s();
    s->batches++;
    s->batched_requests += batch.size();
    std::lock_guard<std::mutex> g(s->mu);
    for (Request *r : batch) {
        s->in_flight -= r->footprint;
        r->t_done = t1;
        r->status = rc ? rc : 0;
        if (rc)
            r->error = err2;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
}

void worker(ol_server *s, ol_ctx *c)
{
    while (true) {
        std::vector<Request *> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                if (s->stop)
                    return true;
                if (s->pending.empty())
                    return false;
                return s->in_flight + s->pending.front()->footprint <= s->cfg.hbm_budget;
            });
            if (s->stop && s->pending.empty())
                return;
            batch = take_batch(s);
        }
        if (!batch.empty())
            run_batch(s, c, batch);
    }
}

}  // namespace

extern "C" {

const char *ol_server_last_error(void) { return s_err.c_str(); }

int ol_server_create(const ol_server_config *cfg, ol_server **out)
{
    if (!out)
        return sfail(OL_ERR_INVALID, "null argument");
    *out = nullptr;
    ol_server *s = new (std::nothrow) ol_server();
    if (!s)
        return sfail(OL_ERR_NOMEM, "out of host memory");
    ol_server_config d{};
    d.hbm_budget = (size_t)32 << 30;
    d.slots = 4;
    d.max_batch_docs = 256;
    d.max_batch_bytes = (size_t)256 << 20;
    if (cfg) {
        if (cfg->hbm_budget) d.hbm_budget = cfg->hbm_budget;
        if (cfg->slots) d.slots = cfg->slots;
        if (cfg->max_batch_docs) d.max_batch_docs = cfg->max_batch_docs;
        if (cfg->max_batch_bytes) d.max_batch_bytes = cfg->max_batch_bytes;
    }
    if (d.slots > 16)
        d.slots = 16;
    s->cfg = d;
    for (uint32_t i = 0; i < d.slots; i++) {
        ol_ctx *c = nullptr;
        const int rc = ol_ctx_create(&c);
        if (rc) {
            const std::string m = ol_last_error();
            for (ol_ctx *x : s->ctx)
                ol_ctx_destroy(x);
            delete s;
            return sfail(rc, m);
        }
        // scan tiles from a ticket: batches of several slots share the device
        (void)ol_tune_set(c, "scan_ticket", 1);
        (void)ol_tune_set(c, "tree_rounds", 4);
        s->ctx.push_back(c);
    }
    for (ol_ctx *c : s->ctx)
        s->workers.emplace_back(worker, s, c);
    *out = s;
    return OL_OK;
}

void ol_server_destroy(ol_server *s)
{
    if (!s)
        return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers)
        t.join();
    for (auto &kv : s->reqs)
        delete kv.second;
    for (ol_dict *d : s->dicts)
        ol_dict_free(d);
    for (ol_ctx *c : s->ctx)
        ol_ctx_destroy(c);
    delete s;
}

int ol_server_add_dict(ol_server *s, const uint8_t *const *patterns, const size_t *lens, size_t count,
                       uint32_t *dict_id)
{
    if (!s || !dict_id)
        return sfail(OL_ERR_INVALID, "null argument");
    ol_dict *d = nullptr;
    const int rc = ol_dict_compile(s->ctx[0], nullptr, patterns, lens, count, &d);
    if (rc)
        return sfail(rc, ol_last_error());
    std::lock_guard<std::mutex> g(s->mu);
    *dict_id = (uint32_t)s->dicts.size();
    s->dicts.push_back(d);
    return OL_OK;
}

int ol_server_submit(ol_server *s, int kind, uint32_t dict_id, const uint8_t *doc, size_t len, int keep_results,
                     uint64_t *req_id)
{
    if (!s || !req_id || (len && !doc))
        return sfail(OL_ERR_INVALID, "null argument");
    if (kind != OL_REQ_PARSE && kind != OL_REQ_SEARCH)
        return sfail(OL_ERR_INVALID, "unknown request kind");
    Request *r = new (std::nothrow) Request();
    if (!r)
        return sfail(OL_ERR_NOMEM, "out of host memory");
    r->kind = kind;
    r->dict = dict_id;
    r->doc = doc;
    r->len = len;
    r->keep = keep_results != 0;
    r->footprint = footprint_of(len, kind);
    r->t_submit = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    if (kind == OL_REQ_SEARCH && dict_id >= s->dicts.size()) {
        delete r;
        return sfail(OL_ERR_INVALID, "unknown dictionary " + std::to_string(dict_id));
    }
    if (r->footprint > s->cfg.hbm_budget) {
        delete r;
        return sfail(OL_ERR_NOMEM, "request needs " + std::to_string(footprint_of(len, kind)) +
                                       " bytes of HBM, more than the whole budget of " +
                                       std::to_string(s->cfg.hbm_budget));
    }
    r->id = s->next_id++;
    s->reqs[r->id] = r;
    s->pending.push_back(r);
    *req_id = r->id;
    s->cv_work.notify_one();
    return OL_OK;
}

int ol_server_wait(ol_server *s, uint64_t req, double timeout_s)
{
    if (!s)
        return sfail(OL_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end())
        return sfail(OL_ERR_INVALID, "unknown request");
    Request *r = it->second;
    const auto until = Clock::now() + std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9);
    if (!s->cv_done.wait_until(lk, until, [&] { return r->status != 1; }))
        return sfail(OL_ERR_DEVICE, "request still running at the time limit");
    if (r->status < 0)
        return sfail(r->status, r->error);
    return OL_OK;
}

int ol_server_result_lines(ol_server *s, uint64_t req, size_t *lines, uint32_t *offsets, uint8_t *classes,
                           uint32_t *parent, uint32_t *subtree_end, uint32_t *first_child, uint32_t *next_sibling)
{
    if (!s || !lines)
        return sfail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end() || it->second->status != 0 || it->second->kind != OL_REQ_PARSE)
        return sfail(OL_ERR_INVALID, "no parse result for this request");
    Request *r = it->second;
    const size_t n = r->offsets.size();
    *lines = n;
    if (offsets) std::memcpy(offsets, r->offsets.data(), n * 4);
    if (classes) std::memcpy(classes, r->classes.data(), n);
    if (parent) std::memcpy(parent, r->parent.data(), n * 4);
    if (subtree_end) std::memcpy(subtree_end, r->se.data(), n * 4);
    if (first_child) std::memcpy(first_child, r->fc.data(), n * 4);
    if (next_sibling) std::memcpy(next_sibling, r->ns.data(), n * 4);
    return OL_OK;
}

int ol_server_result_matches(ol_server *s, uint64_t req, size_t *count, uint32_t *offsets, uint32_t *ids)
{
    if (!s || !count)
        return sfail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end() || it->second->status != 0 || it->second->kind != OL_REQ_SEARCH)
        return sfail(OL_ERR_INVALID, "no search result for this request");
    Request *r = it->second;
    *count = r->m_off.size();
    if (offsets) std::memcpy(offsets, r->m_off.data(), r->m_off.size() * 4);
    if (ids) std::memcpy(ids, r->m_id.data(), r->m_id.size() * 4);
    return OL_OK;
}

int ol_server_timing(ol_server *s, uint64_t req, double *submit, double *start, double *done)
{
    if (!s)
        return sfail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end())
        return sfail(OL_ERR_INVALID, "unknown request");
    if (submit) *submit = it->second->t_submit;
    if (start) *start = it->second->t_start;
    if (done) *done = it->second->t_done;
    return OL_OK;
}

int ol_server_release(ol_server *s, uint64_t req)
{
    if (!s)
        return sfail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end())
        return sfail(OL_ERR_INVALID, "unknown request");
    if (it->second->status == 1)
        return sfail(OL_ERR_INVALID, "request still pending");
    delete it->second;
    s->reqs.erase(it);
    return OL_OK;
}

int ol_server_load(ol_server *s, const uint8_t *const *docs, const size_t *lens, size_t ndocs, double qps,
                   size_t nreq, double search_fraction, uint64_t seed, ol_serve_stats *out)
{
    if (!s || !out || !ndocs || !docs || !lens || qps <= 0)
        return sfail(OL_ERR_INVALID, "bad argument");
    uint32_t ndict;
    {
        std::lock_guard<std::mutex> g(s->mu);
        ndict = (uint32_t)s->dicts.size();
    }
    const uint64_t batches0 = s->batches.load(), batched0 = s->batched_requests.load();
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(qps);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<uint64_t> ids;
    std::vector<double> arrival;
    ids.reserve(nreq);
    arrival.reserve(nreq);
    uint64_t bytes = 0, nsearch = 0, rejected = 0;
    const double t0 = now_s() + 0.01;
    double t = t0;
    for (size_t i = 0; i < nreq; i++) {
        t += gap(rng);
        const size_t di = i % ndocs;
        const bool search = ndict && u(rng) < search_fraction;
        const uint32_t dict = search ? (uint32_t)(rng() % ndict) : 0;
        // open loop: submit at the scheduled arrival time, whatever is in flight
        while (true) {
            const double w = t - now_s();
            if (w <= 0)
                break;
            if (w > 2e-4)
                std::this_thread::sleep_for(std::chrono::duration<double>(w - 1e-4));
        }
        uint64_t id = 0;
        if (ol_server_submit(s, search ? OL_REQ_SEARCH : OL_REQ_PARSE, dict, docs[di], lens[di], 0, &id)) {
            rejected++;
            continue;
        }
        ids.push_back(id);
        arrival.push_back(t);
        bytes += lens[di];
        nsearch += search ? 1 : 0;
    }
    std::vector<double> lat;
    lat.reserve(ids.size());
    uint64_t out_bytes = 0;
    double t_end = t0;
    int first_err = OL_OK;
    std::string err;
    for (size_t k = 0; k < ids.size(); k++) {
        const int rc = ol_server_wait(s, ids[k], 600.0);
        if (rc && !first_err) {
            first_err = rc;
            err = s_err;
        }
        std::lock_guard<std::mutex> g(s->mu);
        Request *r = s->reqs[ids[k]];
        lat.push_back(r->t_done - arrival[k]);
        t_end = std::max(t_end, r->t_done);
        out_bytes += r->out_bytes;
        delete r;
        s->reqs.erase(ids[k]);
    }
    if (first_err)
        return sfail(first_err, err);
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) -> double {
        if (lat.empty())
            return 0.0;
        size_t i = (size_t)std::ceil(p * lat.size()) - 1;
        return lat[std::min(i, lat.size() - 1)] * 1e3;
    };
    std::memset(out, 0, sizeof *out);
    const double wall = t_end - t0;
    out->requests = ids.size();
    out->search_requests = nsearch;
    out->parse_requests = ids.size() - nsearch;
    out->rejected = rejected;
    out->bytes = bytes;
    out->seconds = wall;
    out->qps_offered = qps;
    out->qps_achieved = wall > 0 ? ids.size() / wall : 0;
    out->gbps = wall > 0 ? (double)(bytes + out_bytes) / wall / 1e9 : 0;
    out->p50_ms = pct(0.50);
    out->p90_ms = pct(0.90);
    out->p99_ms = pct(0.99);
    out->max_ms = lat.empty() ? 0 : lat.back() * 1e3;
    double sum = 0;
    for (double x : lat)
        sum += x;
    out->mean_ms = lat.empty() ? 0 : sum / lat.size() * 1e3;
    out->batches = s->batches.load() - batches0;
    out->batched_requests = s->batched_requests.load() - batched0;
    return OL_OK;
}

}  // extern "C"

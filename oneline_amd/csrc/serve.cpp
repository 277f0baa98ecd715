// serve.cpp -- a continuous-batching request loop over serving slots (no
// counterpart in the reference, whose callers invoke one GpuParser method at
// a time; crates/oneline-gpu/src/parser.rs:15-19 serialises them).
//
// Requests (parse, or a dictionary search against one of the resident
// dictionaries, on bytes the caller passes or on a resident document) queue in
// arrival order.  Each of `slots` worker threads owns a slot (slot.h): a
// context with its own HIP stream and a device arena allocated once, here.  A
// worker repeatedly takes the oldest pending requests that belong together --
// parse requests, or searches with the same dictionary -- in order, as far as
// its arena and the batch limits allow, runs them as one batch carved from
// its arena and completes them.  Requests that arrive while other batches are
// on the device join the next batch a slot forms, so up to `slots` batches
// overlap on the GPU.  Dictionaries and resident documents stay in HBM; the
// budget is charged with the arenas, the dictionaries and the resident
// documents at their actual device sizes.
#include "../../include/oneline_hip.h"
#include "slot.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
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

// A document kept in HBM with its lines and tree.
struct Resident {
    ol_doc *doc = nullptr;
    ol_tree *tree = nullptr;
    size_t bytes = 0;
    ~Resident()
    {
        ol_tree_free(tree);
        ol_doc_free(doc);
    }
};

struct Request {
    uint64_t id = 0;
    int kind = 0;            // OL_REQ_PARSE / OL_REQ_SEARCH
    uint32_t dict = 0;
    const uint8_t *doc = nullptr;
    size_t len = 0;
    std::shared_ptr<Resident> resident;   // a request on a resident document
    bool keep = true;        // keep the results for ol_server_result_*
    double t_submit = 0, t_start = 0, t_done = 0;
    uint32_t slot = 0;
    int status = 1;          // 1 pending, 0 done, < 0 error
    std::string error;
    SlotResult res;
    uint64_t out_bytes = 0;  // result bytes produced
};

}  // namespace

struct ol_server {
    ol_server_config cfg{};
    size_t arena = 0;                      // device bytes of each slot's arena
    std::vector<ol_slot *> slots;
    std::vector<std::thread> workers;
    ol_ctx *admin = nullptr;               // builds dictionaries and resident documents
    std::mutex admin_mu;
    std::deque<ol_dict *> dicts;           // stable addresses
    std::unordered_map<uint64_t, std::shared_ptr<Resident>> residents;
    uint64_t next_doc = 1;
    size_t used = 0;                       // budget charged: arenas + dictionaries + resident documents
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<uint64_t, Request *> reqs;
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

// Arena bytes that hold a batch of max_batch_bytes of text as any number of
// documents up to max_batch_docs: the larger of the parse and the search plan
// (the search's match list gets what is left).
size_t arena_for(const ol_server_config &cfg)
{
    std::vector<size_t> lens(std::max<uint32_t>(1, cfg.max_batch_docs), 0);
    lens[0] = cfg.max_batch_bytes;
    const size_t parse = ol_slot_parse_bytes(lens.data(), lens.size());
    // the search plan, plus match lists of up to one match per 4 text bytes
    const size_t search = ol_slot_search_bytes(lens.data(), lens.size()) + cfg.max_batch_bytes * 3;
    return std::max(parse, search) + (1u << 20);
}

// Takes the next batch: the oldest pending request and the following ones of
// the same kind and dictionary, in order, while the slot's arena and the batch
// limits allow (no overtaking, so admission follows arrival order).  A request
// on a resident document runs alone.
std::vector<Request *> take_batch(ol_server *s)
{
    std::vector<Request *> out;
    if (s->pending.empty())
        return out;
    Request *head = s->pending.front();
    if (head->resident) {
        out.push_back(head);
        s->pending.pop_front();
        return out;
    }
    std::vector<size_t> lens;
    size_t bytes = 0;
    auto it = s->pending.begin();
    while (it != s->pending.end() && out.size() < s->cfg.max_batch_docs) {
        Request *r = *it;
        if (r->resident || r->kind != head->kind || (r->kind == OL_REQ_SEARCH && r->dict != head->dict))
            break;
        if (!out.empty()) {
            if (bytes + r->len > s->cfg.max_batch_bytes)
                break;
            lens.push_back(r->len);
            const size_t need = r->kind == OL_REQ_PARSE ? ol_slot_parse_bytes(lens.data(), lens.size())
                                                        : ol_slot_search_bytes(lens.data(), lens.size());
            lens.pop_back();
            if (need > s->arena)
                break;
        }
        lens.push_back(r->len);
        bytes += r->len;
        out.push_back(r);
        it = s->pending.erase(it);
    }
    return out;
}

// Runs the batch; rcs[i] / errs[i] get each request's status and message.
void run_requests(ol_slot *slot, std::vector<Request *> &batch, const ol_dict *dict, std::vector<int> &rcs,
                  std::vector<std::string> &errs)
{
    const size_t k = batch.size();
    std::vector<const uint8_t *> ptrs(k);
    std::vector<size_t> lens(k);
    for (size_t i = 0; i < k; i++) {
        ptrs[i] = batch[i]->doc;
        lens[i] = batch[i]->len;
    }
    std::vector<SlotResult> res(k);
    rcs.assign(k, OL_OK);
    errs.assign(k, std::string());
    const int rc = batch[0]->kind == OL_REQ_PARSE ? ol_slot_parse(slot, ptrs.data(), lens.data(), k, res.data())
                                                  : ol_slot_search(slot, ptrs.data(), lens.data(), k, dict, res.data());
    if (rc == OL_OK) {
        for (size_t i = 0; i < k; i++)
            batch[i]->res = std::move(res[i]);
        return;
    }
    if (k == 1) {
        rcs[0] = rc;
        errs[0] = ol_last_error();
        return;
    }
    // a document nested too deep, or a match list too long for the arena:
    // run each request alone, so only the one it concerns fails
    for (size_t i = 0; i < k; i++) {
        std::vector<Request *> one{batch[i]};
        std::vector<int> r1;
        std::vector<std::string> e1;
        run_requests(slot, one, dict, r1, e1);
        rcs[i] = r1[0];
        errs[i] = e1[0];
    }
}

int run_resident(ol_slot *slot, Request *r, const ol_dict *dict)
{
    const Resident &d = *r->resident;
    if (r->kind == OL_REQ_SEARCH)
        return ol_slot_search_doc(slot, d.doc, dict, &r->res);
    const size_t n = ol_doc_line_count(d.doc);
    SlotResult &x = r->res;
    x.offsets.resize(n);
    x.classes.resize(n);
    x.parent.resize(n);
    x.se.resize(n);
    x.fc.resize(n);
    x.ns.resize(n);
    ol_ctx *c = ol_slot_ctx(slot);
    int rc = ol_doc_download(c, d.doc, x.offsets.data(), x.classes.data(), nullptr);
    if (!rc)
        rc = ol_tree_download(c, d.tree, x.parent.data(), x.se.data(), x.fc.data(), x.ns.data());
    return rc;
}

void run_batch(ol_server *s, uint32_t slot_index, std::vector<Request *> &batch)
{
    ol_slot *slot = s->slots[slot_index];
    const double t0 = now_s();
    const ol_dict *dict = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (batch[0]->kind == OL_REQ_SEARCH)
            dict = s->dicts[batch[0]->dict];
    }
    for (Request *r : batch) {
        r->t_start = t0;
        r->slot = slot_index;
    }
    std::vector<int> rcs;
    std::vector<std::string> errs;
    if (batch[0]->resident) {
        const int rc = run_resident(slot, batch[0], dict);
        rcs.assign(1, rc);
        errs.assign(1, rc ? std::string(ol_last_error()) : std::string());
    } else {
        run_requests(slot, batch, dict, rcs, errs);
    }
    for (Request *r : batch) {
        r->out_bytes = r->kind == OL_REQ_PARSE ? (uint64_t)r->res.offsets.size() * 21 : (uint64_t)r->res.m_off.size() * 8;
        if (!r->keep)
            r->res = SlotResult();
    }
    const double t1 = now_s();
    s->batches++;
    s->batched_requests += batch.size();
    std::lock_guard<std::mutex> g(s->mu);
    for (size_t i = 0; i < batch.size(); i++) {
        Request *r = batch[i];
        r->t_done = t1;
        r->status = rcs[i] ? rcs[i] : 0;
        if (rcs[i])
            r->error = errs[i];
    }
    s->cv_done.notify_all();
}

void worker(ol_server *s, uint32_t slot_index)
{
    while (true) {
        std::vector<Request *> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
            if (s->stop && s->pending.empty())
                return;
            batch = take_batch(s);
        }
        if (!batch.empty())
            run_batch(s, slot_index, batch);
    }
}

Request *find_done(ol_server *s, uint64_t req, int kind)
{
    auto it = s->reqs.find(req);
    if (it == s->reqs.end() || it->second->status != 0 || it->second->kind != kind)
        return nullptr;
    return it->second;
}

// The request sequence of an open-loop load (kinds, documents, dictionaries,
// arrival gaps), drawn from one seeded generator.
struct LoadPlan {
    std::vector<double> gap;
    std::vector<size_t> doc;
    std::vector<int> search;
    std::vector<uint32_t> dict;
};
LoadPlan load_plan(size_t nreq, size_t ndocs, double qps, double search_fraction, uint32_t ndict, uint64_t seed)
{
    LoadPlan p;
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(qps > 0 ? qps : 1.0);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (size_t i = 0; i < nreq; i++) {
        p.gap.push_back(gap(rng));
        p.doc.push_back(i % ndocs);
        const bool search = ndict && u(rng) < search_fraction;
        p.search.push_back(search ? 1 : 0);
        p.dict.push_back(search ? (uint32_t)(rng() % ndict) : 0u);
    }
    return p;
}

void latency_stats(std::vector<double> lat, ol_serve_stats *out)
{
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) -> double {
        if (lat.empty())
            return 0.0;
        const size_t i = (size_t)std::ceil(q * lat.size()) - 1;
        return lat[std::min(i, lat.size() - 1)] * 1e3;
    };
    out->p50_ms = pct(0.50);
    out->p90_ms = pct(0.90);
    out->p99_ms = pct(0.99);
    out->max_ms = lat.empty() ? 0 : lat.back() * 1e3;
    double sum = 0;
    for (double x : lat)
        sum += x;
    out->mean_ms = lat.empty() ? 0 : sum / lat.size() * 1e3;
}

}  // namespace

extern "C" {

const char *ol_server_last_error(void) { return s_err.c_str(); }

int ol_server_create(const ol_server_config *cfg, ol_server **out)
{
    if (!out)
        return sfail(OL_ERR_INVALID, "null argument");
    *out = nullptr;
    ol_server_config d{};
    d.hbm_budget = (size_t)64 << 30;
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
    const size_t arena = arena_for(d);
    if ((size_t)d.slots * arena > d.hbm_budget)
        return sfail(OL_ERR_NOMEM, std::to_string(d.slots) + " slot arenas of " + std::to_string(arena) +
                                       " bytes (batches of " + std::to_string(d.max_batch_bytes) +
                                       " bytes) exceed the HBM budget of " + std::to_string(d.hbm_budget));
    ol_server *s = new (std::nothrow) ol_server();
    if (!s)
        return sfail(OL_ERR_NOMEM, "out of host memory");
    s->cfg = d;
    s->arena = arena;
    auto bail = [&](int rc, const std::string &m) {
        for (ol_slot *x : s->slots)
            ol_slot_destroy(x);
        ol_ctx_destroy(s->admin);
        delete s;
        return sfail(rc, m);
    };
    int rc = ol_ctx_create(&s->admin);
    if (rc)
        return bail(rc, ol_last_error());
    for (uint32_t i = 0; i < d.slots; i++) {
        ol_slot *x = nullptr;
        rc = ol_slot_create(arena, &x);
        if (rc)
            return bail(rc, ol_last_error());
        s->slots.push_back(x);
    }
    s->used = (size_t)d.slots * arena;
    for (uint32_t i = 0; i < d.slots; i++)
        s->workers.emplace_back(worker, s, i);
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
    s->residents.clear();
    for (ol_dict *d : s->dicts)
        ol_dict_free(d);
    for (ol_slot *x : s->slots)
        ol_slot_destroy(x);
    ol_ctx_destroy(s->admin);
    delete s;
}

size_t ol_server_arena_bytes(const ol_server *s) { return s ? s->arena : 0; }

size_t ol_server_hbm_used(ol_server *s)
{
    if (!s)
        return 0;
    std::lock_guard<std::mutex> g(s->mu);
    return s->used;
}

int ol_server_add_dict(ol_server *s, const uint8_t *const *patterns, const size_t *lens, size_t count,
                       uint32_t *dict_id)
{
    if (!s || !dict_id)
        return sfail(OL_ERR_INVALID, "null argument");
    ol_dict *d = nullptr;
    {
        std::lock_guard<std::mutex> g(s->admin_mu);
        const int rc = ol_dict_compile(s->admin, nullptr, patterns, lens, count, &d);
        if (rc)
            return sfail(rc, ol_last_error());
    }
    const size_t bytes = ol_dict_device_bytes(d);
    std::lock_guard<std::mutex> g(s->mu);
    if (s->used + bytes > s->cfg.hbm_budget) {
        ol_dict_free(d);
        return sfail(OL_ERR_NOMEM, "the dictionary's " + std::to_string(bytes) + " device bytes exceed the HBM budget");
    }
    s->used += bytes;
    *dict_id = (uint32_t)s->dicts.size();
    s->dicts.push_back(d);
    return OL_OK;
}

int ol_server_put_doc(ol_server *s, const uint8_t *doc, size_t len, uint64_t *doc_id)
{
    if (!s || !doc_id || (len && !doc))
        return sfail(OL_ERR_INVALID, "null argument");
    auto r = std::make_shared<Resident>();
    {
        std::lock_guard<std::mutex> g(s->admin_mu);
        int rc = ol_doc_upload_lean(s->admin, doc, len, &r->doc);
        if (!rc)
            rc = ol_tree_build(s->admin, r->doc, nullptr, &r->tree);
        if (rc)
            return sfail(rc, ol_last_error());
    }
    r->bytes = ol_doc_device_bytes(r->doc) + ol_tree_device_bytes(r->tree);
    std::lock_guard<std::mutex> g(s->mu);
    if (s->used + r->bytes > s->cfg.hbm_budget)
        return sfail(OL_ERR_NOMEM, "the document's " + std::to_string(r->bytes) + " device bytes exceed the HBM budget");
    s->used += r->bytes;
    *doc_id = s->next_doc++;
    s->residents[*doc_id] = r;
    return OL_OK;
}

int ol_server_drop_doc(ol_server *s, uint64_t doc_id)
{
    if (!s)
        return sfail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->residents.find(doc_id);
    if (it == s->residents.end())
        return sfail(OL_ERR_INVALID, "unknown document " + std::to_string(doc_id));
    s->used -= it->second->bytes;
    s->residents.erase(it);   // freed when the last request on it completes
    return OL_OK;
}

static int submit(ol_server *s, Request *r, uint64_t *req_id)
{
    std::lock_guard<std::mutex> g(s->mu);
    if (r->kind == OL_REQ_SEARCH && r->dict >= s->dicts.size()) {
        delete r;
        return sfail(OL_ERR_INVALID, "unknown dictionary " + std::to_string(r->dict));
    }
    r->id = s->next_id++;
    s->reqs[r->id] = r;
    s->pending.push_back(r);
    *req_id = r->id;
    s->cv_work.notify_one();
    return OL_OK;
}

int ol_server_submit(ol_server *s, int kind, uint32_t dict_id, const uint8_t *doc, size_t len, int keep_results,
                     uint64_t *req_id)
{
    if (!s || !req_id || (len && !doc))
        return sfail(OL_ERR_INVALID, "null argument");
    if (kind != OL_REQ_PARSE && kind != OL_REQ_SEARCH)
        return sfail(OL_ERR_INVALID, "unknown request kind");
    if (len > s->cfg.max_batch_bytes)
        return sfail(OL_ERR_NOMEM, "a request of " + std::to_string(len) + " bytes is larger than a batch (" +
                                       std::to_string(s->cfg.max_batch_bytes) + " bytes, what a slot's arena holds)");
    Request *r = new (std::nothrow) Request();
    if (!r)
        return sfail(OL_ERR_NOMEM, "out of host memory");
    r->kind = kind;
    r->dict = dict_id;
    r->doc = doc;
    r->len = len;
    r->keep = keep_results != 0;
    r->t_submit = now_s();
    return submit(s, r, req_id);
}

int ol_server_submit_doc(ol_server *s, int kind, uint32_t dict_id, uint64_t doc_id, int keep_results, uint64_t *req_id)
{
    if (!s || !req_id)
        return sfail(OL_ERR_INVALID, "null argument");
    if (kind != OL_REQ_PARSE && kind != OL_REQ_SEARCH)
        return sfail(OL_ERR_INVALID, "unknown request kind");
    Request *r = new (std::nothrow) Request();
    if (!r)
        return sfail(OL_ERR_NOMEM, "out of host memory");
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->residents.find(doc_id);
        if (it == s->residents.end()) {
            delete r;
            return sfail(OL_ERR_INVALID, "unknown document " + std::to_string(doc_id));
        }
        r->resident = it->second;
    }
    r->kind = kind;
    r->dict = dict_id;
    r->len = ol_doc_text_len(r->resident->doc);
    r->keep = keep_results != 0;
    r->t_submit = now_s();
    return submit(s, r, req_id);
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
    Request *r = find_done(s, req, OL_REQ_PARSE);
    if (!r)
        return sfail(OL_ERR_INVALID, "no parse result for this request");
    const SlotResult &x = r->res;
    const size_t n = x.offsets.size();
    *lines = n;
    if (offsets) std::memcpy(offsets, x.offsets.data(), n * 4);
    if (classes) std::memcpy(classes, x.classes.data(), n);
    if (parent) std::memcpy(parent, x.parent.data(), n * 4);
    if (subtree_end) std::memcpy(subtree_end, x.se.data(), n * 4);
    if (first_child) std::memcpy(first_child, x.fc.data(), n * 4);
    if (next_sibling) std::memcpy(next_sibling, x.ns.data(), n * 4);
    return OL_OK;
}

int ol_server_result_matches(ol_server *s, uint64_t req, size_t *count, uint32_t *offsets, uint32_t *ids)
{
    if (!s || !count)
        return sfail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    Request *r = find_done(s, req, OL_REQ_SEARCH);
    if (!r)
        return sfail(OL_ERR_INVALID, "no search result for this request");
    *count = r->res.m_off.size();
    if (offsets) std::memcpy(offsets, r->res.m_off.data(), r->res.m_off.size() * 4);
    if (ids) std::memcpy(ids, r->res.m_id.data(), r->res.m_id.size() * 4);
    return OL_OK;
}

int ol_server_timing(ol_server *s, uint64_t req, double *submit_t, double *start, double *done)
{
    if (!s)
        return sfail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end())
        return sfail(OL_ERR_INVALID, "unknown request");
    if (submit_t) *submit_t = it->second->t_submit;
    if (start) *start = it->second->t_start;
    if (done) *done = it->second->t_done;
    return OL_OK;
}

int ol_server_slot_of(ol_server *s, uint64_t req, uint32_t *slot)
{
    if (!s || !slot)
        return sfail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end() || it->second->status == 1)
        return sfail(OL_ERR_INVALID, "unknown or pending request");
    *slot = it->second->slot;
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
    const LoadPlan plan = load_plan(nreq, ndocs, qps, search_fraction, ndict, seed);
    const uint64_t batches0 = s->batches.load(), batched0 = s->batched_requests.load();
    std::vector<uint64_t> ids;
    std::vector<double> arrival;
    ids.reserve(nreq);
    arrival.reserve(nreq);
    uint64_t bytes = 0, nsearch = 0, rejected = 0;
    const double t0 = now_s() + 0.01;
    double t = t0;
    for (size_t i = 0; i < nreq; i++) {
        t += plan.gap[i];
        const size_t di = plan.doc[i];
        const bool search = plan.search[i] != 0;
        // open loop: submit at the scheduled arrival time, whatever is in flight
        while (true) {
            const double w = t - now_s();
            if (w <= 0)
                break;
            if (w > 2e-4)
                std::this_thread::sleep_for(std::chrono::duration<double>(w - 1e-4));
        }
        uint64_t id = 0;
        if (ol_server_submit(s, search ? OL_REQ_SEARCH : OL_REQ_PARSE, plan.dict[i], docs[di], lens[di], 0, &id)) {
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
    latency_stats(lat, out);
    out->batches = s->batches.load() - batches0;
    out->batched_requests = s->batched_requests.load() - batched0;
    return OL_OK;
}

int ol_server_load_serial(ol_server *s, const uint8_t *const *docs, const size_t *lens, size_t ndocs, size_t nreq,
                          double search_fraction, uint64_t seed, ol_serve_stats *out)
{
    if (!s || !out || !ndocs || !docs || !lens)
        return sfail(OL_ERR_INVALID, "bad argument");
    uint32_t ndict;
    std::vector<const ol_dict *> dicts;
    {
        std::lock_guard<std::mutex> g(s->mu);
        ndict = (uint32_t)s->dicts.size();
        dicts.assign(s->dicts.begin(), s->dicts.end());
    }
    const LoadPlan plan = load_plan(nreq, ndocs, 1.0, search_fraction, ndict, seed);
    ol_ctx *c = nullptr;
    int rc = ol_ctx_create(&c);
    if (rc)
        return sfail(rc, ol_last_error());
    std::vector<double> lat;
    uint64_t bytes = 0, out_bytes = 0, nsearch = 0;
    std::vector<uint32_t> a32, b32, c32, d32, e32, counts;
    std::vector<uint8_t> cls;
    const double t0 = now_s();
    for (size_t i = 0; i < nreq && rc == OL_OK; i++) {
        const double r0 = now_s();
        const uint8_t *doc = docs[plan.doc[i]];
        const size_t len = lens[plan.doc[i]];
        ol_doc *d = nullptr;
        rc = ol_doc_upload_lean(c, doc, len, &d);
        if (rc)
            break;
        if (plan.search[i]) {
            const ol_dict *dict = dicts[plan.dict[i]];
            counts.resize(ol_dict_len(dict));
            uint64_t total = 0;
            ol_many_hit *hits = nullptr;
            size_t nh = 0;
            rc = ol_search_many(c, d, dict, (size_t)0xFFFFFFFFu, counts.data(), &total, &hits, &nh);
            ol_free(hits);
            out_bytes += total * 8;
            nsearch++;
        } else {
            ol_tree *t = nullptr;
            rc = ol_tree_build(c, d, nullptr, &t);
            const size_t n = ol_doc_line_count(d);
            for (auto *v : {&a32, &b32, &c32, &d32, &e32})
                v->resize(n);
            cls.resize(n);
            if (!rc)
                rc = ol_doc_download(c, d, a32.data(), cls.data(), nullptr);
            if (!rc)
                rc = ol_tree_download(c, t, b32.data(), c32.data(), d32.data(), e32.data());
            ol_tree_free(t);
            out_bytes += (uint64_t)n * 21;
        }
        ol_doc_free(d);
        bytes += len;
        lat.push_back(now_s() - r0);
    }
    const double wall = now_s() - t0;
    const std::string err = rc ? std::string(ol_last_error()) : std::string();
    ol_ctx_destroy(c);
    if (rc)
        return sfail(rc, err);
    std::memset(out, 0, sizeof *out);
    out->requests = nreq;
    out->search_requests = nsearch;
    out->parse_requests = nreq - nsearch;
    out->bytes = bytes;
    out->seconds = wall;
    out->qps_achieved = wall > 0 ? nreq / wall : 0;
    out->gbps = wall > 0 ? (double)(bytes + out_bytes) / wall / 1e9 : 0;
    latency_stats(lat, out);
    return OL_OK;
}

}  // extern "C"

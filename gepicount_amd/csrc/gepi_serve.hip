/*
 * gepi_serve.hip -- the serving loop (gepi_server_*): a worker thread drives
 * the batched engine round after round, admitting queued requests between
 * rounds while their device footprint fits the budget.  Panels are resident
 * and keyed by a content hash.  No counterpart in the reference, which runs
 * one scan per process (src/gEpiCount.cpp:84-87).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gepicount.h"
#include "gepi_engine.h"
#include "gepi_host.h"

using namespace gepi_host;

namespace {

/* Two independent 64-bit hashes of the genotype bytes and shape (a 128-bit
 * content key): multiply-rotate mixing over 8-byte words, then a final
 * avalanche (the finaliser of MurmurHash3). */
struct Key128 {
    uint64_t a, b;
    bool operator<(const Key128 &o) const { return a < o.a || (a == o.a && b < o.b); }
};

inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

inline uint64_t fmix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

Key128 content_key(const uint8_t *p, size_t n, int64_t n_variants, int32_t n_samples) {
    uint64_t h1 = 0x9e3779b97f4a7c15ull ^ (uint64_t)n_variants, h2 = 0x632be59bd9b4e019ull ^ (uint64_t)n_samples;
    const uint64_t c1 = 0x87c37b91114253d5ull, c2 = 0x4cf5ad432745937full;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, p + i, 8);
        h1 = rotl(h1 ^ (w * c1), 31) * c2;
        h2 = rotl(h2 ^ (rotl(w, 29) * c2), 27) * c1 + 0x52dce729;
    }
    uint64_t t = 0;
    for (size_t j = 0; i + j < n; j++) t |= (uint64_t)p[i + j] << (8 * j);
    h1 = rotl(h1 ^ (t * c1), 31) * c2 ^ (uint64_t)n;
    h2 = rotl(h2 ^ (rotl(t, 29) * c2), 27) * c1 ^ (uint64_t)n;
    return Key128{fmix(h1 + h2), fmix(h2 + 3 * h1)};
}

struct Resident {
    gepi_panel *panel = nullptr;
    Key128 key{};
    bool keyed = false;        /* by_key maps key to this panel */
    std::vector<uint8_t> bed;  /* the payload, compared on a key hit */
    int64_t n_variants = 0;
    int32_t n_samples = 0;
    int64_t refs = 0; /* owners (add_panel calls) + requests in flight */

    bool same(const uint8_t *p, size_t n, int64_t nv, int32_t ns) const {
        return nv == n_variants && ns == n_samples && n == bed.size() &&
               (n == 0 || memcmp(p, bed.data(), n) == 0);
    }
};

struct Request {
    ScanItem item;
    std::vector<uint64_t> m0, m1;
    uint64_t panel_id = 0;
    std::chrono::steady_clock::time_point t_submit;
    double latency_ms = 0.0;
    bool done = false;
    bool claimed = false; /* a waiter has taken the ticket */
};

} // namespace

struct gepi_server {
    gepi_server_params par{};
    Engine eng;
    std::thread worker;
    std::mutex mu; /* guards everything below, and the engine's bookkeeping */
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::map<uint64_t, Resident> panels;
    std::map<Key128, uint64_t> by_key;
    std::deque<Request *> queue;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_ticket = 1;
    int64_t budget = 0;
    /* statistics */
    int64_t submitted = 0, completed = 0, failed = 0, evals = 0, rounds = 0;
    int32_t max_active = 0;
    double busy_ms = 0.0;
    std::vector<double> latencies;
    int device_error = GEPI_OK;

    void complete_locked(Request *r) {
        r->done = true;
        r->latency_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                                  r->t_submit)
                            .count();
        latencies.push_back(r->latency_ms);
        completed++;
        if (r->item.status != GEPI_OK) failed++;
        evals += r->item.stats.evals;
        auto it = panels.find(r->panel_id);
        if (it != panels.end() && --it->second.refs == 0) drop_panel_locked(it);
        cv_done.notify_all();
    }

    void drop_panel_locked(std::map<uint64_t, Resident>::iterator it) {
        if (it->second.keyed) by_key.erase(it->second.key);
        gepi_panel_destroy(it->second.panel);
        panels.erase(it);
    }

    /* The worker: admission by device footprint between rounds (a request
     * larger than the whole budget runs when nothing else does), then one
     * round of the engine, which only this thread touches.                 */
    void run() {
        hipSetDevice(par.device);
        std::vector<ScanItem *> fin;
        std::vector<Request *> admitted;
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !queue.empty() || !eng.idle(); });
                if (stop && queue.empty() && eng.idle()) return;
                size_t planned = eng.bytes_in_use();
                while (!queue.empty()) {
                    Request *r = queue.front();
                    const size_t fp = Engine::estimate_footprint(r->item.panel, r->item.topmost);
                    const bool alone = eng.idle() && admitted.empty();
                    if (!alone && planned + fp > (size_t)budget) break;
                    planned += fp;
                    queue.pop_front();
                    admitted.push_back(r);
                }
            }
            for (Request *r : admitted) {
                eng.add(&r->item);
                if (r->item.done) {
                    std::lock_guard<std::mutex> lk(mu);
                    complete_locked(r);
                }
            }
            admitted.clear();
            if (eng.idle()) continue;
            const auto t0 = std::chrono::steady_clock::now();
            const int32_t active = (int32_t)eng.active_count();
            fin.clear();
            const int rc = eng.round(fin);
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            std::lock_guard<std::mutex> lk(mu);
            busy_ms += ms;
            rounds++;
            max_active = std::max(max_active, active);
            for (ScanItem *it : fin) complete_locked(static_cast<Request *>(it->tag));
            if (rc != GEPI_OK) {
                /* a device error ends the worker: fail everything pending */
                device_error = rc;
                eng.release();
                for (auto &kv : requests)
                    if (!kv.second->done) {
                        kv.second->item.status = rc;
                        complete_locked(kv.second.get());
                    }
                queue.clear();
                stop = true;
                return;
            }
        }
    }
};

extern "C" int gepi_server_create(const gepi_server_params *par, gepi_server **out) {
    if (!out) return gepi_fail(GEPI_ERR_ARG, "null out");
    *out = nullptr;
    gepi_server_params p{};
    if (par) p = *par;
    int rc;
    if ((rc = set_device(p.device)) != GEPI_OK) return rc;
    std::unique_ptr<gepi_server> s(new gepi_server());
    s->par = p;
    if ((rc = s->eng.init(p.device)) != GEPI_OK) {
        s->eng.release();
        return rc;
    }
    s->eng.round_wg_budget = p.round_wg_budget > 0 ? p.round_wg_budget : 262144;
    if (p.autotune && (rc = s->eng.autotune(2)) != GEPI_OK) {
        s->eng.release();
        return rc;
    }
    s->budget = p.hbm_budget;
    if (s->budget <= 0) {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        s->budget = (int64_t)(fr / 4 * 3);
    }
    s->worker = std::thread([srv = s.get()] { srv->run(); });
    *out = s.release();
    return GEPI_OK;
}

extern "C" int gepi_server_destroy(gepi_server *s) {
    if (!s) return GEPI_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    s->eng.release();
    for (auto &kv : s->panels) gepi_panel_destroy(kv.second.panel);
    delete s;
    return GEPI_OK;
}

extern "C" int gepi_server_add_panel_bed(gepi_server *s, const uint8_t *bed_payload,
                                         int64_t n_variants, int32_t n_samples, uint64_t *panel_id) {
    if (!s || !panel_id) return gepi_fail(GEPI_ERR_ARG, "null argument");
    if (n_variants < 0 || n_samples < 0) return gepi_fail(GEPI_ERR_ARG, "negative panel size");
    const size_t bytes = (size_t)n_variants * (size_t)((n_samples + 3) / 4);
    if (bytes && !bed_payload) return gepi_fail(GEPI_ERR_ARG, "null payload");
    Key128 key = content_key(bed_payload, bytes, n_variants, n_samples);
    /* test hook: every payload gets one key, so that distinct payloads collide */
    if (const char *env = getenv("GEPI_SERVE_KEY_COLLIDE"); env && atoi(env) == 1) key = Key128{1, 1};
    /* A key hit is reused only when the payload is byte-identical: two
     * payloads with one key (a hash collision) get separate panels. */
    auto reuse_locked = [&](uint64_t *id) {
        auto it = s->by_key.find(key);
        if (it == s->by_key.end()) return false;
        Resident &r = s->panels[it->second];
        if (!r.same(bed_payload, bytes, n_variants, n_samples)) return false;
        r.refs++;
        *id = it->second;
        return true;
    };
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return gepi_fail(GEPI_ERR_DEVICE, "server stopped (error %d)", s->device_error);
        if (reuse_locked(panel_id)) return GEPI_OK;
    }
    gepi_panel *p = nullptr;
    int rc = gepi_panel_create_bed(bed_payload, n_variants, n_samples, s->par.device, &p);
    if (rc != GEPI_OK) return rc;
    Resident r;
    r.bed.assign(bed_payload, bed_payload + bytes);
    std::lock_guard<std::mutex> lk(s->mu);
    if (reuse_locked(panel_id)) { /* added by another thread meanwhile */
        gepi_panel_destroy(p);
        return GEPI_OK;
    }
    uint64_t id = key.a ? key.a : 1;
    while (s->panels.count(id)) id++; /* keep ids unique if the first halves collide */
    r.panel = p;
    r.key = key;
    r.keyed = s->by_key.count(key) == 0;
    r.n_variants = n_variants;
    r.n_samples = n_samples;
    r.refs = 1;
    if (r.keyed) s->by_key[key] = id;
    s->panels[id] = std::move(r);
    *panel_id = id;
    return GEPI_OK;
}

extern "C" int gepi_server_release_panel(gepi_server *s, uint64_t panel_id) {
    if (!s) return gepi_fail(GEPI_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->panels.find(panel_id);
    if (it == s->panels.end()) return gepi_fail(GEPI_ERR_ARG, "unknown panel %llu", (unsigned long long)panel_id);
    if (--it->second.refs == 0) s->drop_panel_locked(it);
    return GEPI_OK;
}

extern "C" int gepi_server_submit(gepi_server *s, uint64_t panel_id, const uint64_t *mask_ctrl,
                                  const uint64_t *mask_case, int64_t topmost, int32_t ffun,
                                  uint64_t *ticket) {
    if (!s || !ticket) return gepi_fail(GEPI_ERR_ARG, "null argument");
    std::unique_ptr<Request> r(new Request());
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return gepi_fail(GEPI_ERR_DEVICE, "server stopped (error %d)", s->device_error);
    auto it = s->panels.find(panel_id);
    if (it == s->panels.end())
        return gepi_fail(GEPI_ERR_ARG, "unknown panel %llu", (unsigned long long)panel_id);
    gepi_panel *p = it->second.panel;
    const size_t words = (size_t)std::max(p->w64, 1);
    r->m0.assign(words, 0);
    r->m1.assign(words, 0);
    if (mask_ctrl) memcpy(r->m0.data(), mask_ctrl, sizeof(uint64_t) * (size_t)p->w64);
    if (mask_case) memcpy(r->m1.data(), mask_case, sizeof(uint64_t) * (size_t)p->w64);
    r->item.panel = p;
    r->item.mask_ctrl = r->m0.data();
    r->item.mask_case = r->m1.data();
    r->item.topmost = topmost;
    r->item.ffun = ffun;
    r->item.tag = r.get();
    r->panel_id = panel_id;
    r->t_submit = std::chrono::steady_clock::now();
    it->second.refs++;
    const uint64_t t = s->next_ticket++;
    s->queue.push_back(r.get());
    s->requests[t] = std::move(r);
    s->submitted++;
    *ticket = t;
    s->cv_work.notify_one();
    return GEPI_OK;
}

extern "C" int gepi_server_done(gepi_server *s, uint64_t ticket, int32_t *done) {
    if (!s || !done) return gepi_fail(GEPI_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end())
        return gepi_fail(GEPI_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    *done = it->second->done ? 1 : 0;
    return GEPI_OK;
}

extern "C" int gepi_server_wait(gepi_server *s, uint64_t ticket, gepi_result *tops, int64_t capacity,
                                int64_t *n_tops, gepi_scan_stats *stats) {
    if (!s) return gepi_fail(GEPI_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end())
        return gepi_fail(GEPI_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    Request *r = it->second.get();
    if (r->claimed)
        return gepi_fail(GEPI_ERR_ARG, "ticket %llu is already waited for", (unsigned long long)ticket);
    r->claimed = true;
    s->cv_done.wait(lk, [&] { return r->done; });
    const int64_t n = std::min<int64_t>((int64_t)r->item.results.size(), std::max<int64_t>(capacity, 0));
    if (n > 0 && tops) memcpy(tops, r->item.results.data(), sizeof(gepi_result) * (size_t)n);
    if (n_tops) *n_tops = tops ? n : 0;
    if (stats) {
        *stats = r->item.stats;
        stats->ms_scan = r->latency_ms;
    }
    const int status = r->item.status;
    s->requests.erase(it);
    if (status != GEPI_OK) gepi_fail(status, "request %llu failed (status %d)", (unsigned long long)ticket, status);
    return status;
}

extern "C" int gepi_server_get_stats(gepi_server *s, gepi_server_stats *out) {
    if (!s || !out) return gepi_fail(GEPI_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->evals = s->evals;
    out->rounds = s->rounds;
    out->max_active = s->max_active;
    out->panels = (int32_t)s->panels.size();
    out->busy_ms = s->busy_ms;
    if (!s->latencies.empty()) {
        std::vector<double> l = s->latencies;
        std::sort(l.begin(), l.end());
        out->p50_ms = l[l.size() / 2];
        out->p99_ms = l[std::min(l.size() - 1, (size_t)(0.99 * (double)l.size()))];
        double sum = 0.0;
        for (double v : l) sum += v;
        out->mean_ms = sum / (double)l.size();
    }
    return GEPI_OK;
}

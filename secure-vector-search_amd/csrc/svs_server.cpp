// svs_server.cpp — the serving loop: continuous batching of BN-IVF queries
// over resident indexes (one or more tenants sharing a context).
//
// The reference serves one query per `score` call (scorer-bntm/src/ivf.rs:
// 746-822, eval.rs:2339-2344) and batches only when the caller hands it a
// chunk (eval.rs:2428-2433). Here requests are queued as they arrive; a
// dispatcher thread packs whatever is pending into a batch whenever one of
// the server's slots (device scratch + stream) is free, subject to a
// per-batch memory budget, and launches it asynchronously. Batches on
// different slots overlap on the device; results are copied back into
// pinned memory and handed to the waiting callers when the slot's event
// fires. A batch's k is the largest k of its requests: top-k lists are
// prefixes of each other under the (score desc, position asc) order, so
// each request keeps the first k of its own list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "svs_internal.h"

using namespace svs;

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t ticket = 0;
    uint32_t tenant = 0;
    std::vector<uint64_t> v_q;
    uint8_t seed[32];
    std::vector<uint32_t> probes;
    uint8_t verify = 0;
    uint32_t k = 0;
    // raw query (routed and quantised on the device) instead of v_q + probes
    bool raw = false;
    std::vector<float> x;
    uint32_t nprobe = 0;
    uint64_t qq = 0;
    double t_start = 0;  // submit time, or the scheduled arrival
    // results
    bool done = false;
    int rc = SVS_OK;
    std::string err;
    std::vector<uint32_t> ids;
    std::vector<int64_t> scores;
    uint32_t count = 0;
    int32_t status = 0;
    uint32_t batch_size = 0;
    double t_done = 0;
};

// Pinned host buffer that grows geometrically.
template <typename T>
struct Pinned {
    T* p = nullptr;
    size_t cap = 0;
    int ensure(size_t n) {
        if (n <= cap) return SVS_OK;
        size_t want = std::max(n, cap + cap / 2);
        if (p) hipHostFree(p);
        p = nullptr;
        cap = 0;
        HIP_TRY(hipHostMalloc((void**)&p, sizeof(T) * want, hipHostMallocDefault));
        cap = want;
        return SVS_OK;
    }
    void release() {
        if (p) hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct Slot {
    svs_slot sl;
    hipEvent_t done_ev = nullptr;
    bool busy = false;
    uint64_t seq = 0;  // launch order
    std::vector<Request*> reqs;
    BatchPlan plan;
    PlanView view;
    uint32_t k = 0;
    Pinned<uint64_t> h_vq;
    Pinned<uint8_t> h_seeds;
    // device-planned batches: probe_ptr, probe lists, verify flags, raw queries
    Pinned<uint32_t> h_pptr, h_plist;
    Pinned<uint8_t> h_ver;
    Pinned<float> h_x;
    Pinned<uint32_t> h_ids, h_count, h_flag;
    Pinned<int64_t> h_scores;
    Pinned<int32_t> h_status;
};

}  // namespace

struct svs_server {
    svs_ctx* ctx = nullptr;
    std::vector<svs_index*> tenants;
    // per tenant: sum of its n largest list sizes (the row bound of a
    // routed request with nprobe n); indexes do not change while serving
    std::vector<std::vector<uint64_t>> top_rows;
    // cross-tenant batching: every tenant's lists in one table (a view
    // index, no keys of its own) and each tenant's first list in it
    svs_index* combined = nullptr;
    std::vector<uint32_t> list_base;
    svs_server_config cfg{};
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> by_ticket;
    uint64_t next_ticket = 1, next_seq = 1;
    bool stop = false;
    std::thread loop;
    // counters (under mu)
    uint64_t submitted = 0, completed = 0, batches = 0, batched_queries = 0;
    uint32_t max_batch_seen = 0, max_in_flight = 0;

    void run();
    void launch(Slot& s, std::vector<Request*>& batch);
    void retire(Slot& s);
    uint64_t batch_rows(const Request& r) const;
    uint64_t batch_bytes(uint64_t rows, uint32_t nq, uint32_t k) const;
};

// Candidate rows a request adds to a batch (for a routed request, the
// bound its device plan is sized for: its nprobe largest lists).
uint64_t svs_server::batch_rows(const Request& r) const {
    const svs_index* idx = tenants[r.tenant];
    uint64_t rows = 0;
    if (r.raw) {
        rows = top_rows[r.tenant][r.nprobe];
    } else {
        for (uint32_t c : r.probes)
            if (c < idx->lists.size()) rows += idx->lists[c].m;
    }
    return rows;
}

// Device scratch of a batch: per candidate row of its (bucketed) capacity
// the product, its verification recompute, the score and the dense term
// (4 x 8 B) plus verification partials; per query the encode state (v_q,
// v_enc, [g ; lt], T, byte planes, ...) of about 30 KiB and its top-k.
uint64_t svs_server::batch_bytes(uint64_t rows, uint32_t nq, uint32_t k) const {
    const uint64_t cap = ctx->graphs ? cap_bucket(rows) : rows;
    return cap * 36 + (uint64_t)nq * (30 * 1024 + (uint64_t)k * 16);
}

void svs_server::launch(Slot& s, std::vector<Request*>& batch) {
    const uint32_t nq = (uint32_t)batch.size();
    // one tenant's index, or the combined list table of a cross-tenant
    // batch (requests sorted by tenant: one segment per tenant)
    svs_index* idx = combined ? combined : tenants[batch[0]->tenant];
    const bool raw = batch[0]->raw;  // batches are of one kind
    s.sl.segs.clear();
    std::vector<uint64_t> seg_key;
    if (combined) {
        for (uint32_t q = 0; q < nq;) {
            const uint32_t t = batch[q]->tenant;
            uint32_t e = q;
            while (e < nq && batch[e]->tenant == t) ++e;
            s.sl.segs.push_back(TenantSeg{q, e - q, list_base[t], tenants[t]});
            seg_key.push_back(((uint64_t)t << 32) | (e - q));
            q = e;
        }
    }
    uint32_t k = 0;
    for (Request* r : batch) k = std::max(k, r->k);
    int rc = SVS_OK;
    uint32_t n_items = 0;
    uint64_t rows = 0;
    for (Request* r : batch) {
        n_items += raw ? r->nprobe : (uint32_t)r->probes.size();
        if (raw) {
            rows += top_rows[r->tenant][r->nprobe];
        } else {
            const svs_index* ti = tenants[r->tenant];
            for (uint32_t c : r->probes)
                if (c < ti->lists.size()) rows += ti->lists[c].m;
        }
    }
    const bool dplan = cfg.mode != SVS_MODE_PLAIN && ctx->planner &&
                       device_plan_ok(idx, n_items) && rows < 0xFFFFFFFFull;
    // MFMA tiles, or the VALU kernels: a batch of one query has no rows to
    // share between queries, and the row-streaming product with the
    // per-candidate decode runs one query 10-20 % faster than the tiles
    // (nprobe >= 16 on 550k-row lists; equal at nprobe 8)
    const int kernel = cfg.mode == SVS_MODE_BATCHED && nq > 1 ? 1 : 0;
    int variant = 0, hkernel = 0;
    const uint32_t d = idx->cent_d;
    if (raw && !dplan)
        rc = fail(SVS_ERR_ARG, "a raw batch needs device planning (%u items)", n_items);
    if (rc == SVS_OK) rc = s.h_pptr.ensure(nq + 1);
    if (rc == SVS_OK) rc = s.h_plist.ensure(std::max<uint32_t>(n_items, 1));
    if (rc == SVS_OK) rc = s.h_ver.ensure(nq);
    if (rc == SVS_OK && raw) rc = s.h_x.ensure((size_t)nq * d);
    if (rc == SVS_OK && !raw) rc = s.h_vq.ensure((size_t)nq * N);
    if (rc == SVS_OK) rc = s.h_seeds.ensure((size_t)nq * 32);
    if (rc == SVS_OK) rc = s.h_ids.ensure((size_t)nq * k);
    if (rc == SVS_OK) rc = s.h_scores.ensure((size_t)nq * k);
    if (rc == SVS_OK) rc = s.h_count.ensure(nq);
    if (rc == SVS_OK) rc = s.h_status.ensure(nq);
    if (rc == SVS_OK) rc = s.h_flag.ensure(2);
    if (rc == SVS_OK) {
        s.h_pptr.p[0] = 0;
        for (uint32_t q = 0; q < nq; ++q) {
            const Request* r = batch[q];
            const uint32_t np = raw ? r->nprobe : (uint32_t)r->probes.size();
            if (!raw && !combined) {
                std::copy(r->probes.begin(), r->probes.end(), s.h_plist.p + s.h_pptr.p[q]);
            } else if (!raw) {
                // a tenant's list c sits at list_base + c of the combined
                // table; ids past the tenant's lists stay out of range
                const uint32_t nl = (uint32_t)tenants[r->tenant]->lists.size();
                const uint32_t total = (uint32_t)combined->lists.size();
                uint32_t* dst = s.h_plist.p + s.h_pptr.p[q];
                for (size_t i = 0; i < r->probes.size(); ++i)
                    dst[i] = r->probes[i] < nl ? list_base[r->tenant] + r->probes[i] : total;
            }
            s.h_pptr.p[q + 1] = s.h_pptr.p[q] + np;
            s.h_ver.p[q] = r->verify;
            std::memcpy(s.h_seeds.p + (size_t)q * 32, r->seed, 32);
            if (raw)
                std::memcpy(s.h_x.p + (size_t)q * d, r->x.data(), sizeof(float) * d);
            else
                std::memcpy(s.h_vq.p + (size_t)q * N, r->v_q.data(), sizeof(uint64_t) * N);
        }
        s.h_flag.p[1] = 0;
        if (!dplan)
            rc = plan_for_mode(idx, cfg.mode, nq, s.h_pptr.p, s.h_plist.p, s.h_ver.p, s.plan,
                               hkernel, variant);
    }
    if (rc == SVS_OK) {
        std::lock_guard<std::mutex> g(ctx->mu);
        rc = hipSetDevice(ctx->device) == hipSuccess ? SVS_OK : SVS_ERR_HIP;
        hipStream_t st = s.sl.stream;
        if (dplan) {
            variant = ctx->tile_variant[nq <= 16 ? 0 : nq <= 64 ? 1 : nq <= 256 ? 2 : 3];
            PlanView& v = s.view;
            const double qscale = raw ? (double)batch[0]->qq : 0.0;
            // sizing first (host work, allocations), then the batch as one
            // enqueue of copies from pinned memory and kernels, replayed as a
            // captured graph once its shape has been seen
            // (candidate capacity in buckets, so that batches of similar
            // size share a capture)
            const uint64_t cap = ctx->graphs ? cap_bucket(rows) : rows;
            if (rc == SVS_OK) rc = plan_setup(idx, s.sl, nq, n_items, s.h_ver.p, cap, variant, v);
            if (rc == SVS_OK && raw) rc = route_setup(idx, s.sl, nq);
            if (rc == SVS_OK) rc = stage_queries(ctx, s.sl, nq, nullptr, nullptr);
            if (rc == SVS_OK) rc = view_reserve(ctx, s.sl, v, k);
            uint64_t qbits;
            std::memcpy(&qbits, &qscale, 8);
            std::vector<uint64_t> key{
                3, idx->gen, raw, nq, n_items, k, (uint64_t)kernel,
                (uint64_t)variant, v.any_verify, v.planes_ok, cap, qbits,
                (uint64_t)(uintptr_t)s.h_pptr.p, (uint64_t)(uintptr_t)s.h_plist.p,
                (uint64_t)(uintptr_t)s.h_ver.p, (uint64_t)(uintptr_t)s.h_x.p,
                (uint64_t)(uintptr_t)s.h_vq.p, (uint64_t)(uintptr_t)s.h_seeds.p,
                (uint64_t)(uintptr_t)s.h_ids.p, (uint64_t)(uintptr_t)s.h_scores.p,
                (uint64_t)(uintptr_t)s.h_count.p, (uint64_t)(uintptr_t)s.h_status.p,
                (uint64_t)(uintptr_t)s.h_flag.p, (uint64_t)seg_key.size()};
            key.insert(key.end(), seg_key.begin(), seg_key.end());
            if (rc == SVS_OK)
                rc = run_graphed(ctx, s.sl, key, [&]() -> int {
                    SVS_TRY(plan_stage(s.sl, v, s.h_pptr.p, raw ? nullptr : s.h_plist.p,
                                       s.h_ver.p));
                    if (raw)
                        HIP_TRY(hipMemcpyAsync(s.sl.q_raw.ptr, s.h_x.p,
                                               sizeof(float) * (size_t)nq * d,
                                               hipMemcpyHostToDevice, st));
                    else
                        HIP_TRY(hipMemcpyAsync(s.sl.q_vq.ptr, s.h_vq.p,
                                               sizeof(uint64_t) * (size_t)nq * N,
                                               hipMemcpyHostToDevice, st));
                    HIP_TRY(hipMemcpyAsync(s.sl.q_seeds.ptr, s.h_seeds.p, (size_t)nq * 32,
                                           hipMemcpyHostToDevice, st));
                    if (raw) SVS_TRY(route_launch(idx, s.sl, nq, n_items, qscale));
                    SVS_TRY(plan_launch(idx, s.sl, v, variant));
                    SVS_TRY(run_view(idx, s.sl, v, k, false, kernel, true, variant));
                    HIP_TRY(hipMemcpyAsync(s.h_flag.p + 1, v.hdr + PH_OVERFLOW, sizeof(uint32_t),
                                           hipMemcpyDeviceToHost, st));
                    return collect_async(s.sl, nq, k, n_items, s.h_ids.p, s.h_scores.p,
                                         s.h_count.p, s.h_status.p, s.h_flag.p);
                });
        } else {
            if (rc == SVS_OK) rc = stage_queries(ctx, s.sl, nq, s.h_vq.p, s.h_seeds.p);
            if (rc == SVS_OK) rc = run_plan(idx, s.sl, s.plan, k, false, hkernel, true, variant);
            if (rc == SVS_OK)
                rc = collect_async(s.sl, nq, k, s.plan.n_items, s.h_ids.p, s.h_scores.p,
                                   s.h_count.p, s.h_status.p, s.h_flag.p);
        }
        if (rc == SVS_OK && hipEventRecord(s.done_ev, st) != hipSuccess)
            rc = fail(SVS_ERR_HIP, "hipEventRecord failed");
    }
    s.reqs = batch;
    s.k = k;
    if (rc != SVS_OK) {
        char buf[512];
        svs_last_error(buf, sizeof(buf));
        const double t = now_s();
        std::lock_guard<std::mutex> g(mu);
        for (Request* r : batch) {
            r->rc = rc;
            r->err = buf;
            r->batch_size = nq;
            r->t_done = t;
            r->done = true;
            ++completed;
        }
        s.reqs.clear();
        cv_done.notify_all();
        return;
    }
    s.busy = true;
    s.seq = next_seq++;
    std::lock_guard<std::mutex> g(mu);
    ++batches;
    batched_queries += nq;
    max_batch_seen = std::max(max_batch_seen, nq);
    uint32_t in_flight = 0;
    for (auto& x : slots) in_flight += x->busy ? 1 : 0;
    max_in_flight = std::max(max_in_flight, in_flight);
}

void svs_server::retire(Slot& s) {
    const hipError_t e = hipEventSynchronize(s.done_ev);
    int rc = SVS_OK;
    std::string err;
    if (e != hipSuccess) {
        rc = fail(SVS_ERR_HIP, "batch failed on the device: %s", hipGetErrorString(e));
    } else {
        rc = check_rng_flag(s.h_flag.p[0]);
        if (rc == SVS_OK && s.h_flag.p[1])
            rc = fail(SVS_ERR_STATE, "device plan exceeded its bounds");
    }
    if (rc != SVS_OK) {
        char buf[512];
        svs_last_error(buf, sizeof(buf));
        err = buf;
    }
    const double t = now_s();
    const uint32_t nq = (uint32_t)s.reqs.size();
    {
        std::lock_guard<std::mutex> g(mu);
        for (uint32_t q = 0; q < nq; ++q) {
            Request* r = s.reqs[q];
            r->rc = rc;
            r->err = err;
            r->batch_size = nq;
            if (rc == SVS_OK) {
                r->count = std::min(s.h_count.p[q], r->k);
                r->status = s.h_status.p[q];
                r->ids.assign(s.h_ids.p + (size_t)q * s.k, s.h_ids.p + (size_t)q * s.k + r->count);
                r->scores.assign(s.h_scores.p + (size_t)q * s.k,
                                 s.h_scores.p + (size_t)q * s.k + r->count);
            }
            r->t_done = t;
            r->done = true;
            ++completed;
        }
    }
    s.reqs.clear();
    s.busy = false;
    cv_done.notify_all();
}

void svs_server::run() {
    hipSetDevice(ctx->device);
    for (;;) {
        bool progressed = false;
        // retire finished batches
        for (auto& sp : slots)
            if (sp->busy && hipEventQuery(sp->done_ev) != hipErrorNotReady) {
                retire(*sp);
                progressed = true;
            }
        Slot* free_slot = nullptr;
        for (auto& sp : slots)
            if (!sp->busy) {
                free_slot = sp.get();
                break;
            }
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            bool any_busy = false;
            for (auto& sp : slots) any_busy |= sp->busy;
            if (stop && pending.empty() && !any_busy) return;
            if (free_slot && !pending.empty()) {
                // FIFO over requests; a batch holds requests of one kind (raw
                // with one quantisation scale, or routed) and, unless the
                // server batches across tenants, of one tenant
                const uint32_t tenant = pending.front()->tenant;
                const bool raw = pending.front()->raw;
                const uint64_t qq = pending.front()->qq;
                uint64_t rows = 0;
                uint32_t items = 0, kmax = 0;
                for (auto it = pending.begin(); it != pending.end() && batch.size() < cfg.max_batch;) {
                    Request* r = *it;
                    if ((!combined && r->tenant != tenant) || r->raw != raw ||
                        (raw && r->qq != qq)) {
                        ++it;
                        continue;
                    }
                    const uint64_t rr = batch_rows(*r);
                    const uint32_t kk = std::max(kmax, r->k);
                    if (!batch.empty() &&
                        batch_bytes(rows + rr, (uint32_t)batch.size() + 1, kk) > cfg.slot_bytes)
                        break;
                    // a raw batch stays within what the device planner takes
                    if (raw && !batch.empty() && items + r->nprobe > PLAN_MAX_ITEMS) break;
                    rows += rr;
                    kmax = kk;
                    items += raw ? r->nprobe : 0;
                    batch.push_back(r);
                    it = pending.erase(it);
                }
            } else if (!progressed) {
                if (!any_busy) {
                    cv_submit.wait(lk, [&] { return stop || !pending.empty(); });
                    continue;
                }
                if (!free_slot) {
                    // every slot busy: block on the oldest batch outside the lock
                    lk.unlock();
                    Slot* oldest = nullptr;
                    for (auto& sp : slots)
                        if (sp->busy && (!oldest || sp->seq < oldest->seq)) oldest = sp.get();
                    if (oldest) hipEventSynchronize(oldest->done_ev);
                    continue;
                }
                cv_submit.wait_for(lk, std::chrono::microseconds(50));
                continue;
            }
        }
        if (!batch.empty()) {
            // a cross-tenant batch runs in tenant segments (arrival order
            // kept within a tenant)
            if (combined)
                std::stable_sort(batch.begin(), batch.end(), [](const Request* a, const Request* b) {
                    return a->tenant < b->tenant;
                });
            launch(*free_slot, batch);
        }
    }
}

extern "C" int svs_server_create(svs_index* const* tenants, uint32_t n_tenants,
                                 const svs_server_config* cfg, svs_server** out) {
    if (!tenants || n_tenants == 0 || !cfg || !out) return fail(SVS_ERR_ARG, "null argument");
    if (cfg->max_batch == 0 || cfg->max_batch > 4096)
        return fail(SVS_ERR_ARG, "max_batch = %u must be in 1..4096", cfg->max_batch);
    if (cfg->slots == 0 || cfg->slots > 8)
        return fail(SVS_ERR_ARG, "slots = %u must be in 1..8", cfg->slots);
    if (cfg->mode != SVS_MODE_BATCHED && cfg->mode != SVS_MODE_BATCHED_VALU)
        return fail(SVS_ERR_ARG, "mode must be SVS_MODE_BATCHED or SVS_MODE_BATCHED_VALU");
    svs_ctx* ctx = nullptr;
    for (uint32_t t = 0; t < n_tenants; ++t) {
        if (!tenants[t]) return fail(SVS_ERR_ARG, "tenant %u is null", t);
        if (ctx && tenants[t]->ctx != ctx) return fail(SVS_ERR_ARG, "tenants on different contexts");
        ctx = tenants[t]->ctx;
        // every served batch decodes, so every list needs its M_plain (as
        // svs_score_batch and svs_score_raw_batch require)
        for (size_t c = 0; c < tenants[t]->lists.size(); ++c) {
            const ListHost& L = tenants[t]->lists[c];
            if (L.m && !L.d.mp_t32 && !L.d.mp_t64)
                return fail(SVS_ERR_STATE, "tenant %u list %zu was uploaded without M_plain", t, c);
        }
    }
    auto s = std::make_unique<svs_server>();
    s->ctx = ctx;
    s->tenants.assign(tenants, tenants + n_tenants);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        for (svs_index* t : s->tenants) {
            std::vector<uint64_t> tr(t->lists.size() + 1);
            for (size_t n = 0; n < tr.size(); ++n) tr[n] = top_rows(t, (uint32_t)n);
            s->top_rows.push_back(std::move(tr));
        }
    }
    if (cfg->cross_tenant && n_tenants > 1) {
        // one list table over every tenant's lists (device copies of their
        // list descriptors; the lists stay the tenants')
        uint32_t d = 0;
        for (svs_index* t : s->tenants)
            if (t->cent_dev) {
                if (d && t->cent_d != d)
                    return fail(SVS_ERR_ARG, "cross-tenant batching needs one centroid dimension");
                d = t->cent_d;
            }
        auto* u = new svs_index();
        u->ctx = ctx;
        u->view = true;
        u->gen = next_index_gen();
        u->cent_d = d;
        for (svs_index* t : s->tenants) {
            s->list_base.push_back((uint32_t)u->lists.size());
            for (const ListHost& L : t->lists) {
                ListHost c;
                c.m = L.m;
                c.d = L.d;
                u->lists.push_back(c);
            }
        }
        if (u->lists.size() >= PLAN_MAX_LISTS) {
            delete u;
            return fail(SVS_ERR_ARG, "cross-tenant batching takes up to %u lists in all",
                        PLAN_MAX_LISTS - 1);
        }
        std::lock_guard<std::mutex> g(ctx->mu);
        HIP_TRY(hipSetDevice(ctx->device));
        int rc = dev_alloc(ctx, (void**)&u->lists_dev, sizeof(ListDev) * u->lists.size());
        for (size_t t = 0; rc == SVS_OK && t < s->tenants.size(); ++t)
            if (!s->tenants[t]->lists.empty() &&
                hipMemcpy(u->lists_dev + s->list_base[t], s->tenants[t]->lists_dev,
                          sizeof(ListDev) * s->tenants[t]->lists.size(),
                          hipMemcpyDeviceToDevice) != hipSuccess)
                rc = fail(SVS_ERR_HIP, "copying list descriptors failed");
        if (rc != SVS_OK) {
            if (u->lists_dev) dev_free(ctx, u->lists_dev, sizeof(ListDev) * u->lists.size());
            delete u;
            return rc;
        }
        s->combined = u;
    }
    s->cfg = *cfg;
    if (s->cfg.slot_bytes == 0)
        s->cfg.slot_bytes = std::max<uint64_t>(1ull << 30, ctx->budget_bytes / 16);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        HIP_TRY(hipSetDevice(ctx->device));
        for (uint32_t i = 0; i < cfg->slots; ++i) {
            auto sl = std::make_unique<Slot>();
            HIP_TRY(hipStreamCreateWithFlags(&sl->sl.stream, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&sl->done_ev, hipEventDisableTiming));
            s->slots.push_back(std::move(sl));
        }
    }
    ctx_retain(ctx);
    svs_server* raw = s.release();
    raw->loop = std::thread([raw] { raw->run(); });
    *out = raw;
    return SVS_OK;
}

extern "C" int svs_server_destroy(svs_server* s) {
    if (!s) return SVS_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_submit.notify_all();
    if (s->loop.joinable()) s->loop.join();
    svs_ctx* ctx = s->ctx;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        for (auto& sp : s->slots) {
            hipStreamSynchronize(sp->sl.stream);
            release_graphs(sp->sl);
            for (DevBuf* b : sp->sl.slot_scratch()) b->release(ctx);
            if (sp->sl.prof.created)
                for (auto& e : sp->sl.prof.ev) hipEventDestroy(e);
            hipEventDestroy(sp->done_ev);
            hipStreamDestroy(sp->sl.stream);
            sp->sl.segs.clear();
            sp->h_vq.release();
            sp->h_seeds.release();
            sp->h_pptr.release();
            sp->h_plist.release();
            sp->h_ver.release();
            sp->h_x.release();
            sp->h_ids.release();
            sp->h_count.release();
            sp->h_flag.release();
            sp->h_scores.release();
            sp->h_status.release();
        }
    }
    if (s->combined) {
        std::lock_guard<std::mutex> g(ctx->mu);
        dev_free(ctx, s->combined->lists_dev, sizeof(ListDev) * s->combined->lists.size());
        delete s->combined;
    }
    for (auto& kv : s->by_ticket) delete kv.second;
    delete s;
    ctx_release(ctx);
    return SVS_OK;
}

extern "C" int svs_server_reserve(svs_server* s, uint32_t n_queries, uint32_t n_items,
                                  uint64_t n_candidates, uint32_t k) {
    if (!s || n_queries == 0 || k == 0 || k > 4096) return fail(SVS_ERR_ARG, "bad argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->submitted)
            return fail(SVS_ERR_STATE, "svs_server_reserve must come before the first request");
    }
    std::lock_guard<std::mutex> g(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    for (auto& sp : s->slots) {
        SVS_TRY(reserve_slot(s->ctx, sp->sl, n_queries, n_items,
                             s->ctx->graphs ? cap_bucket(n_candidates) : n_candidates, k));
        SVS_TRY(sp->h_vq.ensure((size_t)n_queries * N));
        SVS_TRY(sp->h_seeds.ensure((size_t)n_queries * 32));
        SVS_TRY(sp->h_ids.ensure((size_t)n_queries * k));
        SVS_TRY(sp->h_scores.ensure((size_t)n_queries * k));
        SVS_TRY(sp->h_count.ensure(n_queries));
        SVS_TRY(sp->h_status.ensure(n_queries));
        SVS_TRY(sp->h_flag.ensure(1));
    }
    return SVS_OK;
}

extern "C" int svs_server_submit(svs_server* s, const svs_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket || !r->q_seed || (!r->query && !r->v_q) ||
        (!r->query && r->n_probes && !r->probes))
        return fail(SVS_ERR_ARG, "null argument");
    if (r->tenant >= s->tenants.size()) return fail(SVS_ERR_ARG, "tenant %u out of range", r->tenant);
    if (r->k == 0 || r->k > 4096) return fail(SVS_ERR_ARG, "k = %u must be in 1..4096", r->k);
    auto q = std::make_unique<Request>();
    q->tenant = r->tenant;
    if (r->query) {
        const svs_index* idx = s->tenants[r->tenant];
        if (!idx->cent_dev)
            return fail(SVS_ERR_STATE, "tenant %u has no centroids for routing", r->tenant);
        if (idx->lists.size() > (size_t)RT_SELECT_MAX)
            return fail(SVS_ERR_ARG, "device routing takes up to %d lists", RT_SELECT_MAX);
        if (s->cfg.mode == SVS_MODE_PLAIN)
            return fail(SVS_ERR_ARG, "raw requests need a batched server mode");
        q->raw = true;
        q->x.assign(r->query, r->query + idx->cent_d);
        q->nprobe = std::min<uint32_t>(r->nprobe, (uint32_t)idx->lists.size());
        q->qq = r->quantisation_q;
    } else {
        for (int i = 0; i < N; ++i)
            if (r->v_q[i] >= P) return fail(SVS_ERR_ARG, "v_q[%d] is not canonical", i);
        q->v_q.assign(r->v_q, r->v_q + N);
        q->probes.assign(r->probes, r->probes + r->n_probes);
    }
    std::memcpy(q->seed, r->q_seed, 32);
    q->verify = r->verify ? 1 : 0;
    q->k = r->k;
    q->t_start = r->arrival_s > 0 ? r->arrival_s : now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return fail(SVS_ERR_STATE, "server is shutting down");
        q->ticket = s->next_ticket++;
        *ticket = q->ticket;
        s->by_ticket[q->ticket] = q.get();
        s->pending.push_back(q.release());
        ++s->submitted;
    }
    s->cv_submit.notify_one();
    return SVS_OK;
}

extern "C" int svs_server_wait(svs_server* s, uint64_t ticket, svs_result* out) {
    if (!s || !out) return fail(SVS_ERR_ARG, "null argument");
    Request* r = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->by_ticket.find(ticket);
        if (it == s->by_ticket.end()) return fail(SVS_ERR_ARG, "unknown ticket %llu",
                                                  (unsigned long long)ticket);
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->by_ticket.erase(it);
    }
    std::unique_ptr<Request> hold(r);
    out->count = r->count;
    out->status = r->status;
    out->rc = r->rc;
    out->batch_size = r->batch_size;
    out->latency_ms = (r->t_done - r->t_start) * 1e3;
    if (out->ids) std::copy(r->ids.begin(), r->ids.end(), out->ids);
    if (out->scores) std::copy(r->scores.begin(), r->scores.end(), out->scores);
    if (r->rc != SVS_OK) return fail(r->rc, "%s", r->err.c_str());
    return SVS_OK;
}

extern "C" int svs_server_stats(svs_server* s, svs_server_counters* out) {
    if (!s || !out) return fail(SVS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->batches = s->batches;
    out->batched_queries = s->batched_queries;
    out->max_batch = s->max_batch_seen;
    out->max_in_flight = s->max_in_flight;
    out->pending = (uint32_t)s->pending.size();
    out->graph_launches = out->graph_captures = 0;
    for (auto& sp : s->slots) {
        out->graph_launches += sp->sl.graph_launches;
        out->graph_captures += sp->sl.graph_captures;
    }
    return SVS_OK;
}

// Open-loop Poisson load: request i is submitted at its scheduled arrival
// t0 + sum of exponential gaps (rate qps, seeded), whatever the state of
// earlier requests; latency runs from the scheduled arrival to completion.
extern "C" int svs_server_run_poisson(svs_server* s, const svs_request* reqs, uint32_t n,
                                      double qps, uint64_t seed, svs_result* results,
                                      double* wall_s) {
    if (!s || (n && (!reqs || !results)) || !(qps > 0)) return fail(SVS_ERR_ARG, "bad argument");
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(qps);
    std::vector<double> at(n);
    const double t0 = now_s() + 1e-3;
    double t = t0;
    for (uint32_t i = 0; i < n; ++i) {
        t += gap(rng);
        at[i] = t;
    }
    std::vector<uint64_t> tickets(n, 0);
    std::vector<int> src(n, SVS_OK);
    std::thread submitter([&] {
        for (uint32_t i = 0; i < n; ++i) {
            const double wait = at[i] - now_s();
            if (wait > 2e-4)
                std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
            while (now_s() < at[i]) {
            }
            svs_request r = reqs[i];
            r.arrival_s = at[i];
            uint64_t tk = 0;
            src[i] = svs_server_submit(s, &r, &tk);
            {
                std::lock_guard<std::mutex> g(s->mu);
                tickets[i] = tk;
            }
            s->cv_done.notify_all();
        }
    });
    int rc = SVS_OK;
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t tk = 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_done.wait(lk, [&] { return tickets[i] != 0 || src[i] != SVS_OK; });
            tk = tickets[i];
        }
        if (src[i] != SVS_OK) {
            results[i].rc = src[i];
            if (rc == SVS_OK) rc = src[i];
            continue;
        }
        const int w = svs_server_wait(s, tk, &results[i]);
        if (w != SVS_OK && rc == SVS_OK) rc = w;
    }
    submitter.join();
    if (wall_s) *wall_s = now_s() - t0;
    return rc;
}

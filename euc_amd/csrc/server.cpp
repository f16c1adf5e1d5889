// server.cpp -- the serving loop of include/euchip.h (euchip_server_*).
//
// Continuous batching on top of the batch entry points: requests queue in
// submission order; every poll retires the batches that finished, then fills
// each free slot with as many queued requests as the device-memory budget
// admits (targets from a pool of freed ones, plus the slot's batch scratch)
// and launches it on the slot's stream.  With two or more slots a batch
// admitted while another one runs overlaps with it on the device.
#include <algorithm>
#include <cmath>
#include <array>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <utility>
#include <vector>

#include "../../include/euchip.h"

namespace euchip {
// runtime.cpp
size_t request_scratch_estimate(const euchip_pipeline_desc *d, const euchip_mesh *m, uint32_t w, uint32_t h);
size_t batch_fixed_bytes();
void set_last_error(euchip_ctx *ctx, const char *msg);
}  // namespace euchip

namespace {

// A request's targets: owning allocations of a size class (so freed targets
// serve later requests of other sizes), and views of the request's size.
struct Target {
    euchip_buffer *pix_store = nullptr, *dep_store = nullptr;  // class_texels(cls) texels each
    euchip_buffer *pixel = nullptr, *depth = nullptr;          // w x h views of the stores
    uint32_t w = 0, h = 0;
    uint32_t cls = 0;
};

struct Queued {
    uint64_t ticket;
    const euchip_mesh *mesh;
    const euchip_buffer *texture, *shadow;
    std::array<float, EUCHIP_UNIFORM_FLOATS> u;
    uint32_t w, h;
    bool own = false;             // rendered with `desc` instead of the server's
    euchip_pipeline_desc desc{};
};

struct Slot {
    euchip_batch *batch = nullptr;
    bool busy = false;
    std::vector<Queued> reqs;
    std::vector<Target> targets;
};

// Size classes of target storage: four per doubling, from 256 texels.
size_t class_texels(uint32_t c) { return (size_t)std::ceil(256.0 * std::exp2(c / 4.0)); }

uint32_t class_of(uint32_t w, uint32_t h) {
    const size_t n = (size_t)w * h;
    uint32_t c = 0;
    while (class_texels(c) < n) c++;
    return c;
}

size_t target_bytes(uint32_t w, uint32_t h) { return class_texels(class_of(w, h)) * 8; }

}  // namespace

struct euchip_server {
    euchip_ctx *ctx = nullptr;
    euchip_pipeline_desc desc{};
    size_t budget = 0;
    uint32_t max_batch = 1;
    uint32_t batch_cap = 1;  // lowered after a batch failed to allocate, raised again as batches succeed
    std::vector<Slot> slots;
    std::deque<Queued> queue;
    std::deque<euchip_completion> done;
    std::map<uint64_t, Target> out;  // targets of completed requests until released
    std::multimap<uint32_t, Target> pool;  // free target storage by size class
    uint64_t next_ticket = 1;
    euchip_server_stats st{};
};

namespace {

void drop_views(Target &t) {
    if (t.pixel) euchip_buffer_destroy(t.pixel);  // views own nothing
    if (t.depth) euchip_buffer_destroy(t.depth);
    t.pixel = t.depth = nullptr;
}

void destroy_target(Target &t) {
    drop_views(t);
    if (t.pix_store) euchip_buffer_destroy(t.pix_store);
    if (t.dep_store) euchip_buffer_destroy(t.dep_store);
    t.pix_store = t.dep_store = nullptr;
}

void to_pool(euchip_server *s, Target t) {
    drop_views(t);
    if (t.pix_store && t.dep_store) s->pool.emplace(t.cls, t);
    else destroy_target(t);
}

bool take_pooled(euchip_server *s, uint32_t w, uint32_t h, Target &t) {
    auto it = s->pool.find(class_of(w, h));
    if (it == s->pool.end()) return false;
    t = it->second;
    s->pool.erase(it);
    return true;
}

// Storage of a fresh target of class t.cls (accounted against the budget).
euchip_status alloc_store(euchip_server *s, Target &t) {
    const uint32_t n = (uint32_t)class_texels(t.cls);
    euchip_status st = euchip_buffer_create(s->ctx, EUCHIP_FORMAT_RGBA8, n, 1, &t.pix_store);
    if (!st) st = euchip_buffer_create(s->ctx, EUCHIP_FORMAT_F32, n, 1, &t.dep_store);
    return st;
}

// The request-sized views over a target's storage.
euchip_status make_views(euchip_server *s, Target &t) {
    euchip_status st = euchip_buffer_wrap(s->ctx, EUCHIP_FORMAT_RGBA8, t.w, t.h, euchip_buffer_device_ptr(t.pix_store),
                                          &t.pixel);
    if (!st)
        st = euchip_buffer_wrap(s->ctx, EUCHIP_FORMAT_F32, t.w, t.h, euchip_buffer_device_ptr(t.dep_store), &t.depth);
    return st;
}

// Free pooled targets until `bytes` more fit the budget (or the pool is empty).
bool make_room(euchip_server *s, size_t bytes) {
    if (!s->budget) return true;
    while (euchip_ctx_bytes_in_use(s->ctx) + bytes > s->budget) {
        if (s->pool.empty()) return false;
        auto it = s->pool.begin();
        destroy_target(it->second);
        s->pool.erase(it);
    }
    return true;
}

size_t slot_scratch(const Slot &sl) { return euchip_batch_scratch_bytes(sl.batch); }

void note_peak(euchip_server *s) {
    s->st.bytes_in_use = euchip_ctx_bytes_in_use(s->ctx);
    s->st.peak_bytes = std::max(s->st.peak_bytes, s->st.bytes_in_use);
}

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

void complete(euchip_server *s, uint64_t ticket, euchip_status st, Target t) {
    euchip_completion c{};
    c.ticket = ticket;
    c.done_ns = now_ns();
    c.status = st;
    c.pixel = t.pixel;
    c.depth = t.depth;
    if (t.pixel || t.depth) s->out[ticket] = t;
    s->done.push_back(c);
    s->st.completed++;
}

// Finished batches: their requests complete (a batch wait re-runs a launch whose
// bins overflowed, so results are whole).
void retire(euchip_server *s) {
    for (Slot &sl : s->slots) {
        if (!sl.busy || euchip_batch_query(sl.batch) == EUCHIP_NOT_READY) continue;
        const euchip_status st = euchip_batch_wait(sl.batch);
        for (size_t i = 0; i < sl.reqs.size(); i++) complete(s, sl.reqs[i].ticket, st, sl.targets[i]);
        sl.reqs.clear();
        sl.targets.clear();
        sl.busy = false;
        s->st.in_flight = 0;
        for (const Slot &o : s->slots)
            if (o.busy) s->st.in_flight += (uint32_t)o.reqs.size();
        if (st == EUCHIP_OK && s->batch_cap < s->max_batch) s->batch_cap = std::min(s->max_batch, s->batch_cap * 2);
    }
}

bool any_busy(const euchip_server *s) {
    for (const Slot &sl : s->slots)
        if (sl.busy) return true;
    return false;
}

// Give back memory nothing is using: pooled targets, and the scratch of idle
// slots other than `keep` (their batches are recreated empty).
void free_transient(euchip_server *s, const Slot *keep) {
    for (auto &kv : s->pool) destroy_target(kv.second);
    s->pool.clear();
    for (Slot &sl : s->slots) {
        if (&sl == keep || sl.busy || euchip_batch_scratch_bytes(sl.batch) == 0) continue;
        euchip_batch *fresh = nullptr;
        if (euchip_batch_create(s->ctx, &fresh) != EUCHIP_OK) continue;
        euchip_batch_destroy(sl.batch);
        sl.batch = fresh;
    }
}

// The head request cannot be admitted now.  It waits while memory will come
// back (a batch running, targets the caller has not released); otherwise idle
// memory is freed once and, if it still does not fit, it is rejected.
bool head_blocked(euchip_server *s, Slot &sl, bool &freed, euchip_status why) {
    if (any_busy(s) || !s->out.empty() || !s->done.empty()) return false;
    if (!freed) {
        free_transient(s, &sl);
        freed = true;
        return true;  // try again
    }
    complete(s, s->queue.front().ticket, why, Target());
    s->queue.pop_front();
    s->st.rejected++;
    return true;
}

// Fill one free slot from the queue; false when nothing more can be admitted now.
bool admit(euchip_server *s, Slot &sl, bool &freed) {
    std::vector<Queued> take;
    std::vector<Target> tg;
    std::vector<bool> fresh;  // target must be allocated
    size_t scratch = euchip::batch_fixed_bytes();
    size_t new_bytes = 0;
    const size_t have = slot_scratch(sl);
    auto need_of = [&](size_t sc) { return sc + sc / 2 > have ? sc + sc / 2 - have : 0; };
    while (!s->queue.empty() && take.size() < s->batch_cap) {
        const Queued &q = s->queue.front();
        Target t;
        const bool pooled = take_pooled(s, q.w, q.h, t);
        const size_t sc = scratch + euchip::request_scratch_estimate(q.own ? &q.desc : &s->desc, q.mesh, q.w, q.h);
        const size_t add = pooled ? 0 : target_bytes(q.w, q.h);
        const size_t need = new_bytes + add + need_of(sc);
        if (s->budget && !make_room(s, need)) {
            if (pooled) to_pool(s, t);
            break;
        }
        take.push_back(q);
        s->queue.pop_front();  // q is gone: use take.back()
        t.w = take.back().w, t.h = take.back().h, t.cls = class_of(t.w, t.h);
        tg.push_back(t);
        fresh.push_back(!pooled);
        scratch = sc;
        new_bytes += add;
    }
    if (take.empty()) return !s->queue.empty() && head_blocked(s, sl, freed, EUCHIP_OVER_BUDGET);
    euchip_status st = EUCHIP_OK;
    for (size_t i = 0; i < tg.size() && !st; i++) {
        if (fresh[i]) st = alloc_store(s, tg[i]);
        if (!st) st = make_views(s, tg[i]);
    }
    std::vector<euchip_request> reqs(take.size());
    if (!st) {
        for (size_t i = 0; i < take.size(); i++) {
            reqs[i].mesh = take[i].mesh;
            reqs[i].pixel = tg[i].pixel;
            reqs[i].depth = tg[i].depth;
            reqs[i].uniforms = take[i].u.data();
            reqs[i].texture = take[i].texture;
            reqs[i].shadow = take[i].shadow;
        }
        bool own = false;
        for (const Queued &q : take) own = own || q.own;
        if (!own) {
            st = euchip_batch_set(sl.batch, &s->desc, reqs.data(), (uint32_t)reqs.size());
        } else {  // requests with pipelines of their own: one part per distinct desc
            std::vector<euchip_pipeline_desc> descs(take.size());
            for (size_t i = 0; i < take.size(); i++) descs[i] = take[i].own ? take[i].desc : s->desc;
            st = euchip_batch_set_multi(sl.batch, descs.data(), reqs.data(), (uint32_t)reqs.size());
        }
    }
    if (!st) st = euchip_batch_launch(sl.batch, 0);
    if (st == EUCHIP_OVER_BUDGET || st == EUCHIP_OOM) {
        // memory ran out after all: hand everything back and admit fewer next time
        for (Target &t : tg) to_pool(s, t);
        for (size_t i = take.size(); i-- > 0;) s->queue.push_front(take[i]);
        if (s->batch_cap > 1) {
            s->batch_cap = std::max<uint32_t>(1, s->batch_cap / 2);
            return !any_busy(s);  // retry now only when nothing will free memory
        }
        return head_blocked(s, sl, freed, st);
    }
    if (st) {  // a request the batch refuses (invalid arguments): fail the whole group
        for (size_t i = 0; i < take.size(); i++) {
            destroy_target(tg[i]);
            complete(s, take[i].ticket, st, Target());
        }
        return true;
    }
    sl.reqs = std::move(take);
    sl.targets = std::move(tg);
    sl.busy = true;
    s->st.admitted += sl.reqs.size();
    s->st.batches++;
    s->st.in_flight += (uint32_t)sl.reqs.size();
    note_peak(s);
    return true;
}

// The server owns its targets and hands out pooled storage that still holds an
// earlier request's pixels and depth: every request must start from its desc's
// clear values, or it would see (and depth-test against) another's.
euchip_status check_server_desc(euchip_ctx *ctx, const euchip_pipeline_desc *desc) {
    if (desc->primitives != EUCHIP_PRIMS_TRIANGLE_LIST) {
        euchip::set_last_error(ctx, "servers draw triangle lists");
        return EUCHIP_INVALID_ARG;
    }
    if (!desc->clear_color_enable || !desc->clear_depth_enable) {
        euchip::set_last_error(ctx, "server pipelines must enable both clears (clear_color_enable, clear_depth_enable): "
                                   "pooled targets hold earlier requests' results");
        return EUCHIP_INVALID_ARG;
    }
    return EUCHIP_OK;
}

}  // namespace

extern "C" {

euchip_status euchip_server_create(euchip_ctx *ctx, const euchip_pipeline_desc *desc, size_t budget_bytes,
                                   uint32_t max_batch, uint32_t slots, euchip_server **out) {
    if (!ctx || !desc || !out || max_batch == 0 || slots == 0) return EUCHIP_INVALID_ARG;
    *out = nullptr;
    if (const euchip_status st = check_server_desc(ctx, desc)) return st;
    euchip_server *s = new euchip_server();
    s->ctx = ctx;
    s->desc = *desc;
    s->budget = budget_bytes;
    s->max_batch = s->batch_cap = max_batch;
    s->st.budget = budget_bytes;
    s->slots.resize(slots);
    for (Slot &sl : s->slots) {
        const euchip_status st = euchip_batch_create(ctx, &sl.batch);
        if (st) {
            euchip_server_destroy(s);
            return st;
        }
    }
    euchip_ctx_set_budget(ctx, budget_bytes);
    note_peak(s);
    *out = s;
    return EUCHIP_OK;
}

euchip_status euchip_server_destroy(euchip_server *s) {
    if (!s) return EUCHIP_INVALID_ARG;
    for (Slot &sl : s->slots) {
        if (!sl.batch) continue;
        if (sl.busy) (void)euchip_batch_wait(sl.batch);
        for (Target &t : sl.targets) destroy_target(t);
        euchip_batch_destroy(sl.batch);
    }
    for (auto &kv : s->out) destroy_target(kv.second);
    for (auto &kv : s->pool) destroy_target(kv.second);
    delete s;
    return EUCHIP_OK;
}

euchip_status euchip_server_submit_desc(euchip_server *s, const euchip_serve_request *r,
                                       const euchip_pipeline_desc *desc, uint64_t *ticket) {
    if (!s || !r || !r->mesh || !r->uniforms || !r->width || !r->height) return EUCHIP_INVALID_ARG;
    if (desc) {
        if (const euchip_status st = check_server_desc(s->ctx, desc)) return st;
    }
    Queued q;
    if (desc) q.own = true, q.desc = *desc;
    q.ticket = s->next_ticket++;
    q.mesh = r->mesh;
    q.texture = r->texture;
    q.shadow = r->shadow;
    memcpy(q.u.data(), r->uniforms, sizeof(float) * EUCHIP_UNIFORM_FLOATS);
    q.w = r->width;
    q.h = r->height;
    s->queue.push_back(q);
    s->st.submitted++;
    if (ticket) *ticket = q.ticket;
    return EUCHIP_OK;
}

euchip_status euchip_server_submit(euchip_server *s, const euchip_serve_request *r, uint64_t *ticket) {
    return euchip_server_submit_desc(s, r, nullptr, ticket);
}

euchip_status euchip_server_submit_n(euchip_server *s, const euchip_serve_request *reqs, uint32_t n,
                                     uint64_t *tickets) {
    if (!s || (n && !reqs)) return EUCHIP_INVALID_ARG;
    for (uint32_t i = 0; i < n; i++) {
        const euchip_status st = euchip_server_submit(s, &reqs[i], tickets ? &tickets[i] : nullptr);
        if (st) return st;
    }
    return EUCHIP_OK;
}

euchip_status euchip_server_poll(euchip_server *s, euchip_completion *out, uint32_t max, uint32_t *n_out) {
    if (!s) return EUCHIP_INVALID_ARG;
    retire(s);
    bool freed = false;
    for (Slot &sl : s->slots) {
        while (!sl.busy && !s->queue.empty())
            if (!admit(s, sl, freed)) break;
    }
    uint32_t n = 0;
    while (n < max && out && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    if (n_out) *n_out = n;
    s->st.queued = (uint32_t)s->queue.size();
    note_peak(s);
    return EUCHIP_OK;
}

euchip_status euchip_server_release(euchip_server *s, uint64_t ticket) {
    if (!s) return EUCHIP_INVALID_ARG;
    auto it = s->out.find(ticket);
    if (it == s->out.end()) return EUCHIP_INVALID_ARG;
    to_pool(s, it->second);
    s->out.erase(it);
    return EUCHIP_OK;
}

euchip_status euchip_server_release_n(euchip_server *s, const uint64_t *tickets, uint32_t n) {
    if (!s || (n && !tickets)) return EUCHIP_INVALID_ARG;
    euchip_status bad = EUCHIP_OK;
    for (uint32_t i = 0; i < n; i++)
        if (euchip_server_release(s, tickets[i])) bad = EUCHIP_INVALID_ARG;
    return bad;
}

uint64_t euchip_clock_ns(void) { return now_ns(); }

euchip_status euchip_server_get_stats(const euchip_server *s, euchip_server_stats *out) {
    if (!s || !out) return EUCHIP_INVALID_ARG;
    *out = s->st;
    out->queued = (uint32_t)s->queue.size();
    out->bytes_in_use = euchip_ctx_bytes_in_use(s->ctx);
    return EUCHIP_OK;
}

}  // extern "C"

// Device-cache residency policy: which pattern blocks occupy which slots of a
// pattern store, as host C++ behind the C ABI (no device call).
//
// The policy is Rocket's (R/cache/AbstractCache.java:19-409 over the buffer
// stack of R/cache/DeviceCache.java:19-62): every entry has a reference count
// and a status (WRITING: one holder fills it; READONLY: filled, shared;
// INVALID: a writer let go without filling it).  A transaction asks for
// several keys at once; transactions are served first come first served, and
// one that does not fit waits at the head of the queue, keeping what it has,
// while unreferenced entries are evicted least recently used first.  A
// released readable entry goes to the back of that order, a new or invalid one
// to the front.  Slots are handed out from a stack: the highest-numbered free
// slot first, a slot freed by an eviction reused first.
//
// What differs from the reference is the shape of the calls, not the policy:
// a transaction is a number that the caller polls (the reference completes a
// future), and an entry is addressed by its key (the reference hands out the
// entry object; a referenced entry cannot be evicted, so the key names it
// until its last release).  One mutex serialises every call on a cache.
#include <cstdint>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/prnu_api.h"

namespace prnu {
void set_error(const char *fmt, ...);
}

namespace {

enum Status { kWriting = 0, kReadonly = 1, kInvalid = 2 };

struct Txn;

struct Entry {
    int64_t key = 0;
    int64_t slot = 0;
    int64_t meta = 0;
    int64_t refcount = 0;
    int status = kInvalid;
    std::deque<Txn *> waiting;  // transactions pending on the writer
};

struct Txn {
    int64_t id = 0;
    std::vector<int64_t> keys;    // duplicates removed, first occurrences in order
    std::vector<Entry *> entries;
    int64_t pending = 0;
    bool complete = false;
    void trigger() {
        if (--pending == 0) complete = true;
    }
};

}  // namespace

struct prnu_slotcache {
    std::mutex mu;
    int64_t nslots = 0;
    std::vector<int64_t> free;  // stack, top at the back
    std::unordered_map<int64_t, std::unique_ptr<Entry>> entries;
    std::list<int64_t> lru;     // unreferenced keys, front evicted first
    std::unordered_map<int64_t, std::list<int64_t>::iterator> lru_pos;
    std::deque<Txn *> queue;    // transactions still collecting entries
    std::unordered_map<int64_t, std::unique_ptr<Txn>> txns;
    int64_t next_id = 1;
    int64_t hits = 0, misses = 0, evictions = 0;

    void lru_add_front(int64_t k) { lru_pos[k] = lru.insert(lru.begin(), k); }
    void lru_add_back(int64_t k) { lru_pos[k] = lru.insert(lru.end(), k); }
    void lru_remove(int64_t k) {
        auto it = lru_pos.find(k);
        lru.erase(it->second);
        lru_pos.erase(it);
    }

    // AbstractCache.Entry.acquire (AbstractCache.java:137-177)
    void acquire_entry(Entry *e, Txn *t) {
        Txn *trig = t;
        if (e->status == kWriting) {
            e->waiting.push_back(t);
            trig = nullptr;
        } else if (e->status == kInvalid) {
            e->status = kWriting;
        }
        if (++e->refcount == 1) lru_remove(e->key);
        if (trig) trig->trigger();
    }

    // AbstractCache.evictOne with DeviceCache.destroyBuffer (:351-378)
    bool evict_one() {
        if (lru.empty()) return false;
        const int64_t k = lru.front();
        lru_remove(k);
        auto it = entries.find(k);
        free.push_back(it->second->slot);
        entries.erase(it);
        ++evictions;
        return true;
    }

    // AbstractCache.makeProgress (:311-349)
    void make_progress() {
        bool done = false;
        while (!done && !queue.empty()) {
            Txn *t = queue.front();
            queue.pop_front();
            bool success = true;
            for (size_t i = 0; i < t->keys.size(); ++i) {
                if (t->entries[i]) continue;
                const int64_t k = t->keys[i];
                auto it = entries.find(k);
                if (it != entries.end()) {
                    t->entries[i] = it->second.get();
                    acquire_entry(t->entries[i], t);
                    continue;
                }
                if (!free.empty()) {
                    auto e = std::make_unique<Entry>();
                    e->key = k;
                    e->slot = free.back();
                    free.pop_back();
                    Entry *p = e.get();
                    entries.emplace(k, std::move(e));
                    lru_add_front(k);  // a new entry starts at the front (AbstractCache.java:52)
                    t->entries[i] = p;
                    acquire_entry(p, t);
                    continue;
                }
                success = false;
            }
            if (!success) {
                queue.push_front(t);
                done = !evict_one();
            }
        }
    }
};

namespace {

int bad(const char *what) {
    prnu::set_error("%s", what);
    return PRNU_EINVAL;
}

}  // namespace

PRNU_EXPORT int prnu_slotcache_slots_for(size_t entry_bytes, int64_t requested, size_t free_bytes,
                                         int64_t *nslots) {
    if (entry_bytes == 0 || !nslots) return bad("slotcache_slots_for: entry size must be positive");
    // RocketLauncher.java:371,391-398 and DeviceCache.java:27-31: the total is
    // the request, or free memory less the margin when the request is not
    // positive, at most 5000 entries; one more buffer while the allocated
    // bytes plus that buffer stay below the total.
    const int64_t eb = (int64_t)entry_bytes;
    int64_t total = requested > 0 ? requested : (int64_t)free_bytes + requested;
    if (total > 5000 * eb) total = 5000 * eb;
    *nslots = total > eb ? (total - 1) / eb : 0;
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_create(int64_t nslots, prnu_slotcache **out) {
    if (!out) return bad("slotcache_create: null output");
    // DeviceWorker.java:87-89: the worker needs two entries (one pair)
    if (nslots < 2) return bad("slotcache_create: cache requires at least two entries");
    auto c = new prnu_slotcache();
    c->nslots = nslots;
    c->free.reserve((size_t)nslots);
    for (int64_t s = 0; s < nslots; ++s) c->free.push_back(s);  // buffers.push(alloc) in order
    *out = c;
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_destroy(prnu_slotcache *c) {
    delete c;
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_acquire(prnu_slotcache *c, const int64_t *keys, int n, int64_t *txn) {
    if (!c || !txn || n < 0 || (n > 0 && !keys)) return bad("slotcache_acquire: invalid arguments");
    std::lock_guard<std::mutex> g(c->mu);
    auto t = std::make_unique<Txn>();
    t->id = c->next_id++;
    for (int i = 0; i < n; ++i) {  // removeDuplicates (AbstractCache.java:391-408)
        bool seen = false;
        for (int64_t k : t->keys) seen |= (k == keys[i]);
        if (!seen) t->keys.push_back(keys[i]);
    }
    t->entries.assign(t->keys.size(), nullptr);
    t->pending = (int64_t)t->keys.size();
    t->complete = t->keys.empty();
    Txn *p = t.get();
    c->txns.emplace(p->id, std::move(t));
    c->queue.push_back(p);
    c->make_progress();
    *txn = p->id;
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_poll(prnu_slotcache *c, int64_t txn, int *done, int64_t *keys, int64_t *slots,
                                    int32_t *writer, int cap, int *n) {
    if (!c || !done || !n) return bad("slotcache_poll: invalid arguments");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->txns.find(txn);
    if (it == c->txns.end()) {
        prnu::set_error("slotcache_poll: unknown transaction %lld", (long long)txn);
        return PRNU_EINVAL;
    }
    Txn *t = it->second.get();
    *n = (int)t->keys.size();
    *done = t->complete ? 1 : 0;
    if (!t->complete || (!keys && !slots && !writer)) return PRNU_OK;  // waiting, or a count only
    if (cap < *n || (*n > 0 && (!keys || !slots || !writer))) {
        prnu::set_error("slotcache_poll: %d entries exceed the capacity %d", *n, cap);
        return PRNU_EINVAL;
    }
    // DeviceWorker.java:208-226: an entry still being written is the caller's
    // to fill (a miss); any other is a hit
    for (int i = 0; i < *n; ++i) {
        const Entry *e = t->entries[i];
        keys[i] = t->keys[i];
        slots[i] = e->slot;
        writer[i] = e->status == kWriting;
        if (writer[i]) ++c->misses;
        else ++c->hits;
    }
    c->txns.erase(it);
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_ready(prnu_slotcache *c, int64_t key, int64_t meta) {
    if (!c) return bad("slotcache_ready: null cache");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->entries.find(key);
    if (it == c->entries.end() || it->second->status != kWriting)
        return bad("slotcache_ready: cannot downgrade cache entry, entry is not writable");
    // Entry.downgradeToReader (AbstractCache.java:103-121)
    Entry *e = it->second.get();
    e->meta = meta;
    e->status = kReadonly;
    std::deque<Txn *> w;
    w.swap(e->waiting);
    for (Txn *t : w) t->trigger();
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_release(prnu_slotcache *c, int64_t key) {
    if (!c) return bad("slotcache_release: null cache");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->entries.find(key);
    if (it == c->entries.end() || it->second->refcount <= 0)
        return bad("slotcache_release: cannot release cache entry since refcount is zero");
    // Entry.release (AbstractCache.java:179-232)
    Entry *e = it->second.get();
    Txn *handover = nullptr;
    if (e->status == kWriting) {
        if (!e->waiting.empty()) {  // the first pending reader becomes the writer
            handover = e->waiting.front();
            e->waiting.pop_front();
        } else {
            e->status = kInvalid;
        }
    }
    if (--e->refcount == 0) {
        if (e->status == kReadonly) c->lru_add_back(key);
        else c->lru_add_front(key);
        c->make_progress();
    }
    if (handover) handover->trigger();
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_try_acquire(prnu_slotcache *c, int64_t key, int64_t *slot) {
    if (!c || !slot) return bad("slotcache_try_acquire: invalid arguments");
    std::lock_guard<std::mutex> g(c->mu);
    // tryAcquireImmediately (AbstractCache.java:288-295, :123-135)
    *slot = -1;
    auto it = c->entries.find(key);
    if (it == c->entries.end() || it->second->status != kReadonly) return PRNU_OK;
    Entry *e = it->second.get();
    if (++e->refcount == 1) c->lru_remove(key);
    *slot = e->slot;
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_meta(prnu_slotcache *c, int64_t key, int64_t *meta) {
    if (!c || !meta) return bad("slotcache_meta: invalid arguments");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->entries.find(key);
    if (it == c->entries.end() || it->second->status != kReadonly)
        return bad("slotcache_meta: cannot obtain size data while entry is writable");
    *meta = it->second->meta;
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_entry(prnu_slotcache *c, int64_t key, int64_t *slot, int *status,
                                     int64_t *refcount) {
    if (!c || !slot || !status || !refcount) return bad("slotcache_entry: invalid arguments");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->entries.find(key);
    if (it == c->entries.end()) {
        *slot = -1;
        *status = -1;
        *refcount = 0;
        return PRNU_OK;
    }
    *slot = it->second->slot;
    *status = it->second->status;
    *refcount = it->second->refcount;
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_lru(prnu_slotcache *c, int64_t *keys, int64_t cap, int64_t *n) {
    if (!c || !n) return bad("slotcache_lru: invalid arguments");
    std::lock_guard<std::mutex> g(c->mu);
    *n = (int64_t)c->lru.size();
    if (!keys) return PRNU_OK;
    if (cap < *n) return bad("slotcache_lru: capacity too small");
    int64_t i = 0;
    for (int64_t k : c->lru) keys[i++] = k;
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_stats(prnu_slotcache *c, int64_t *stats) {
    if (!c || !stats) return bad("slotcache_stats: invalid arguments");
    std::lock_guard<std::mutex> g(c->mu);
    stats[0] = c->hits;
    stats[1] = c->misses;
    stats[2] = c->evictions;
    stats[3] = (int64_t)c->entries.size();
    stats[4] = (int64_t)c->free.size();
    stats[5] = (int64_t)c->queue.size();
    stats[6] = c->nslots;
    stats[7] = (int64_t)c->lru.size();
    return PRNU_OK;
}

PRNU_EXPORT int prnu_slotcache_cleanup(prnu_slotcache *c) {
    if (!c) return bad("slotcache_cleanup: null cache");
    std::lock_guard<std::mutex> g(c->mu);
    // AbstractCache.cleanup (:380-389)
    while (c->evict_one()) {
    }
    if (!c->entries.empty()) {
        prnu::set_error("slotcache_cleanup: failed to clean up cache, %lld entries are still in use!",
                        (long long)c->entries.size());
        return PRNU_EINVAL;
    }
    return PRNU_OK;
}

// A device memory pool (lbn_pool): the buffers a request holds (its resident witness, the fold's
// byte layout) are taken from free lists keyed by size and given back to them, never hipFree'd
// while the pool lives. The reference recycles its workspace for the same reason
// (crates/pcs/src/key.rs:185-191, scheme/prover.rs:116-131): first-touch allocation costs more
// than the work; on a GPU, hipFree also waits for the whole device, which would stall every
// stream in flight each time a request finishes.
#include <map>
#include <mutex>
#include <unordered_map>

#include "internal.hpp"

struct lbn_pool {
    struct Block {
        void *p = nullptr;
        uint64_t size = 0;  // rounded
        uint64_t tag = 0;   // what the last holder laid out in it (0: nothing in particular)
    };
    std::mutex mu;
    int device = 0;
    uint64_t cap = 0;        // bytes the pool may obtain from the device
    uint64_t obtained = 0;   // bytes it has (held until the pool is released)
    uint64_t in_use = 0, peak_in_use = 0;
    uint64_t mallocs = 0, reuses = 0, trims = 0;
    std::multimap<uint64_t, Block> free_blocks;
    std::unordered_map<void *, Block> used;
};

namespace lbn {

static uint64_t pool_round(uint64_t bytes) {
    const uint64_t g = bytes < (64ull << 20) ? (1ull << 20) : (16ull << 20);  // 1 MiB, or 16 MiB past 64 MiB
    return (std::max<uint64_t>(bytes, 1) + g - 1) / g * g;
}

int pool_alloc(lbn_pool *pool, uint64_t bytes, uint64_t tag, void **out, bool *same_tag) {
    const uint64_t sz = pool_round(bytes);
    std::lock_guard<std::mutex> lk(pool->mu);
    auto range = pool->free_blocks.equal_range(sz);
    auto pick = range.second;
    for (auto it = range.first; it != range.second; ++it)
        if (pick == range.second || it->second.tag == tag) {
            pick = it;
            if (it->second.tag == tag) break;
        }
    lbn_pool::Block b;
    if (pick != range.second) {
        b = pick->second;
        pool->free_blocks.erase(pick);
        pool->reuses++;
        *same_tag = b.tag == tag && tag != 0;
    } else {
        // make room by giving back free blocks of other sizes (hipFree waits for the device, so
        // only when the cap demands it)
        while (pool->obtained + sz > pool->cap && !pool->free_blocks.empty()) {
            auto it = std::prev(pool->free_blocks.end());
            (void)hipFree(it->second.p);
            pool->obtained -= it->second.size;
            pool->free_blocks.erase(it);
            pool->trims++;
        }
        if (pool->obtained + sz > pool->cap) return fail(LBN_ERR_OOM, "the device pool's cap is reached");
        const hipError_t e = hipMalloc(&b.p, sz);
        if (e != hipSuccess) return fail(LBN_ERR_OOM, std::string("pool hipMalloc: ") + hipGetErrorString(e));
        b.size = sz;
        pool->obtained += sz;
        pool->mallocs++;
        *same_tag = false;
    }
    b.tag = tag;
    pool->used[b.p] = b;
    pool->in_use += b.size;
    pool->peak_in_use = std::max(pool->peak_in_use, pool->in_use);
    *out = b.p;
    return LBN_OK;
}

void pool_free(lbn_pool *pool, void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(pool->mu);
    auto it = pool->used.find(p);
    if (it == pool->used.end()) return;
    pool->in_use -= it->second.size;
    pool->free_blocks.emplace(it->second.size, it->second);
    pool->used.erase(it);
}

uint64_t pool_rounded(uint64_t bytes) { return pool_round(bytes); }

void pool_set_cap(lbn_pool *pool, uint64_t cap) {
    std::lock_guard<std::mutex> lk(pool->mu);
    pool->cap = cap;
}

}  // namespace lbn

using namespace lbn;

extern "C" {

int lbn_pool_create(uint64_t cap_bytes, lbn_pool **out) {
    if (!out) return fail(LBN_ERR_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(LBN_ERR_NO_DEVICE, "no HIP device");
    lbn_pool *p = new lbn_pool();
    p->cap = cap_bytes;
    (void)hipGetDevice(&p->device);
    *out = p;
    return LBN_OK;
}

void lbn_pool_release(lbn_pool *pool) {
    if (!pool) return;
    for (auto &kv : pool->free_blocks) (void)hipFree(kv.second.p);
    for (auto &kv : pool->used) (void)hipFree(kv.second.p);
    delete pool;
}

int lbn_pool_stats(lbn_pool *pool, uint64_t *obtained, uint64_t *in_use, uint64_t *peak_in_use, uint64_t *mallocs,
                   uint64_t *reuses) {
    if (!pool) return fail(LBN_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(pool->mu);
    if (obtained) *obtained = pool->obtained;
    if (in_use) *in_use = pool->in_use;
    if (peak_in_use) *peak_in_use = pool->peak_in_use;
    if (mallocs) *mallocs = pool->mallocs;
    if (reuses) *reuses = pool->reuses;
    return LBN_OK;
}

int lbn_prover_set_pool(lbn_prover *pr, lbn_pool *pool) {
    if (!pr) return fail(LBN_ERR_ARG, "null argument");
    if (pool && pool->device != pr->pp->device) return fail(LBN_ERR_SHAPE, "the pool is on another device");
    // layouts recycled by the prover so far go back to where they came from
    for (uint8_t *b : pr->free_bytes_f) {
        if (pr->pool)
            pool_free(pr->pool, b);
        else
            (void)hipFree(b);
    }
    pr->free_bytes_f.clear();
    pr->pool = pool;
    return LBN_OK;
}

}  // extern "C"

// BLAKE3 from the public specification (see blake3.hpp).
//
// The tree is kept as a binary counter of complete subtrees: after chunk number c (0-based) is
// closed, its chaining value is carried upwards through every occupied level exactly as a binary
// increment carries, so level l holds a complete subtree of 2^l chunks. At finalisation the
// pending chunk is the rightmost leaf and the occupied levels are merged into it from the lowest
// upwards, which is the spec's rule that a left subtree holds the largest power-of-two number of
// chunks.
#include "blake3.hpp"
#include "threads.hpp"

#include <cstring>
#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <vector>

namespace lbn {

namespace {

const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                        0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
const uint8_t SCHEDULE[7][16] = {
    // the message word order of each round: the permutation (2 6 3 10 7 0 4 13 1 11 12 5 9 14 15 8)
    // applied r times to the identity
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
    {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8},
    {3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1},
    {10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6},
    {12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4},
    {9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7},
    {11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13},
};
enum : uint32_t {
    CHUNK_START = 1,
    CHUNK_END = 2,
    PARENT = 4,
    ROOT = 8,
    KEYED_HASH = 16,
    DERIVE_KEY_CONTEXT = 32,
    DERIVE_KEY_MATERIAL = 64
};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

inline void mix(uint32_t v[16], int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] += v[b] + x;
    v[d] = rotr(v[d] ^ v[a], 16);
    v[c] += v[d];
    v[b] = rotr(v[b] ^ v[c], 12);
    v[a] += v[b] + y;
    v[d] = rotr(v[d] ^ v[a], 8);
    v[c] += v[d];
    v[b] = rotr(v[b] ^ v[c], 7);
}

void compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter, uint32_t len,
              uint32_t flags, uint32_t out[16]) {
    uint32_t v[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      IV[0], IV[1], IV[2], IV[3], (uint32_t)counter, (uint32_t)(counter >> 32),
                      len,   flags};
    for (int r = 0; r < 7; r++) {
        const uint8_t *s = SCHEDULE[r];
        mix(v, 0, 4, 8, 12, m[s[0]], m[s[1]]);
        mix(v, 1, 5, 9, 13, m[s[2]], m[s[3]]);
        mix(v, 2, 6, 10, 14, m[s[4]], m[s[5]]);
        mix(v, 3, 7, 11, 15, m[s[6]], m[s[7]]);
        mix(v, 0, 5, 10, 15, m[s[8]], m[s[9]]);
        mix(v, 1, 6, 11, 12, m[s[10]], m[s[11]]);
        mix(v, 2, 7, 8, 13, m[s[12]], m[s[13]]);
        mix(v, 3, 4, 9, 14, m[s[14]], m[s[15]]);
    }
    for (int i = 0; i < 8; i++) {
        out[i] = v[i] ^ v[i + 8];
        out[i + 8] = v[i + 8] ^ cv[i];
    }
}

void load_words(const uint8_t *b, uint32_t *w, int n) {
    for (int i = 0; i < n; i++) {
        uint32_t x;
        std::memcpy(&x, b + 4 * i, 4);  // little-endian host
        w[i] = x;
    }
}

// The chaining value of one whole (non-root) chunk of 1024 bytes.
void chunk_cv(const uint32_t key[8], uint32_t flags, const uint8_t *p, uint64_t counter, uint32_t out[8]) {
    uint32_t cv[8], m[16], o[16];
    std::memcpy(cv, key, 32);
    for (int b = 0; b < 16; b++) {
        load_words(p + 64 * b, m, 16);
        compress(cv, m, counter, 64, flags | (b == 0 ? CHUNK_START : 0) | (b == 15 ? CHUNK_END : 0), o);
        std::memcpy(cv, o, 32);
    }
    std::memcpy(out, cv, 32);
}

void parent_cv(const uint32_t key[8], uint32_t flags, const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
    uint32_t m[16], o[16];
    std::memcpy(m, l, 32);
    std::memcpy(m + 8, r, 32);
    compress(key, m, 0, 64, flags | PARENT, o);
    std::memcpy(out, o, 32);
}

// The chaining value of a complete (non-root) subtree of n = 2^k chunks starting at chunk
// `counter`; the spec's tree splits it into two halves of 2^(k-1) chunks.
void subtree_cv(const uint32_t key[8], uint32_t flags, const uint8_t *p, uint64_t counter, uint64_t n,
                uint32_t out[8]) {
    if (n == 1) return chunk_cv(key, flags, p, counter, out);
    uint32_t l[8], r[8];
    subtree_cv(key, flags, p, counter, n / 2, l);
    subtree_cv(key, flags, p + n / 2 * 1024, counter + n / 2, n / 2, r);
    parent_cv(key, flags, l, r, out);
}

constexpr uint64_t SUBTREE_MIN = 16;         // chunks hashed as one subtree when the input allows
constexpr size_t PARALLEL_MIN = 1 << 20;     // bytes of one update from which subtrees go to threads
constexpr uint64_t LEAF = 32;                // chunks of one thread's task

struct Sub {
    const uint8_t *p;
    uint64_t counter, n;
};

// The chaining values of complete subtrees, all of them at once: each is split into leaves of at
// most LEAF chunks (both are powers of two), the leaves are hashed on up to 16 threads, and each
// subtree's leaves are merged pairwise as the tree does.
void subtree_cvs(const uint32_t key[8], uint32_t flags, const std::vector<Sub> &subs,
                 std::vector<std::array<uint32_t, 8>> &out) {
    std::vector<Sub> leaves;
    std::vector<size_t> first(subs.size());
    for (size_t i = 0; i < subs.size(); i++) {
        const uint64_t per = std::min(subs[i].n, LEAF);
        first[i] = leaves.size();
        for (uint64_t o = 0; o < subs[i].n; o += per) leaves.push_back({subs[i].p + o * 1024, subs[i].counter + o, per});
    }
    std::vector<std::array<uint32_t, 8>> cv(leaves.size());
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t k; (k = next.fetch_add(1)) < leaves.size();)
            subtree_cv(key, flags, leaves[k].p, leaves[k].counter, leaves[k].n, cv[k].data());
    };
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(16, hw), leaves.size());
    lbn::run_ranges(nt, nt, [&](size_t, size_t) { work(); });  // hashing allocates nothing
    out.resize(subs.size());
    for (size_t i = 0; i < subs.size(); i++) {
        std::array<uint32_t, 8> *c = cv.data() + first[i];
        for (uint64_t w = subs[i].n / std::min(subs[i].n, LEAF); w > 1; w /= 2)
            for (uint64_t j = 0; j < w / 2; j++) parent_cv(key, flags, c[2 * j].data(), c[2 * j + 1].data(), c[j].data());
        out[i] = c[0];
    }
}

}  // namespace

// A complete subtree of 2^level chunks enters the binary counter at `level`: chunk_ is a multiple of
// 2^level, so every lower level is empty, and the carry proceeds as when a chunk closes.
void Blake3::push_subtree(const uint32_t cv_in[8], int level) {
    uint32_t cv[8];
    std::memcpy(cv, cv_in, 32);
    while (has_ >> level & 1) {
        parent_cv(key_, flags_, subtree_[level], cv, cv);
        has_ &= ~(1ull << level);
        level++;
    }
    std::memcpy(subtree_[level], cv, 32);
    has_ |= 1ull << level;
}

Blake3::Blake3(const uint32_t key[8], uint32_t flags) {
    std::memcpy(key_, key, sizeof key_);
    flags_ = flags;
    std::memcpy(cv_, key, sizeof cv_);
    chunk_ = 0;
    buf_len_ = 0;
    blocks_ = 0;
    has_ = 0;
}

Blake3 Blake3::hasher() { return Blake3(IV, 0); }

Blake3 Blake3::keyed(const uint8_t key[32]) {
    uint32_t k[8];
    load_words(key, k, 8);
    return Blake3(k, KEYED_HASH);
}

Blake3 Blake3::derive_key(const std::string &context) {
    Blake3 c(IV, DERIVE_KEY_CONTEXT);
    c.update(context.data(), context.size());
    uint8_t ck[32];
    c.finalize().read(0, ck, 32);
    uint32_t k[8];
    load_words(ck, k, 8);
    return Blake3(k, DERIVE_KEY_MATERIAL);
}

void Blake3::close_chunk() {
    // the buffered (full) block is the chunk's last
    uint32_t m[16], o[16];
    load_words(buf_, m, 16);
    compress(cv_, m, chunk_, 64, flags_ | (blocks_ == 0 ? CHUNK_START : 0) | CHUNK_END, o);
    uint32_t cv[8];
    std::memcpy(cv, o, sizeof cv);
    int level = 0;
    while (has_ >> level & 1) {
        uint32_t pm[16];
        std::memcpy(pm, subtree_[level], 32);
        std::memcpy(pm + 8, cv, 32);
        compress(key_, pm, 0, 64, flags_ | PARENT, o);
        std::memcpy(cv, o, sizeof cv);
        has_ &= ~(1ull << level);
        level++;
    }
    std::memcpy(subtree_[level], cv, sizeof cv);
    has_ |= 1ull << level;
    chunk_++;
    std::memcpy(cv_, key_, sizeof cv_);
    buf_len_ = 0;
    blocks_ = 0;
}

// A large update runs its loop twice. The first run, on a copy, only records the complete
// subtrees the loop takes; their chaining values are then computed together on threads; the
// second run is the real one and takes each subtree's value from that list. The byte-wise work
// (chunks before the input reaches an aligned subtree, and the tail) stays sequential.
struct Blake3::Plan {
    bool record = true;
    std::vector<Sub> subs;
    std::vector<std::array<uint32_t, 8>> cvs;
    size_t next = 0;
};

void Blake3::update(const void *data, size_t len) {
    const uint8_t *p = (const uint8_t *)data;
    if (len >= PARALLEL_MIN) {
        try {
            Plan plan;
            {
                Blake3 dry = *this;
                dry.update_loop(p, len, &plan);
            }
            subtree_cvs(key_, flags_, plan.subs, plan.cvs);
            plan.record = false;
            update_loop(p, len, &plan);  // takes the planned values only: allocates nothing
            return;
        } catch (const std::exception &) {
            // no memory for the plan: the state is untouched (the plan ran on a copy), hash in order
        }
    }
    update_loop(p, len, nullptr);
}

void Blake3::update_loop(const uint8_t *p, size_t len, Plan *plan) {
    while (len) {
        if (buf_len_ == 64) {
            // more input follows the buffered block, so it is not the chunk's final one
            if (blocks_ == 15) {
                close_chunk();
            } else {
                uint32_t m[16], o[16];
                load_words(buf_, m, 16);
                compress(cv_, m, chunk_, 64, flags_ | (blocks_ == 0 ? CHUNK_START : 0), o);
                std::memcpy(cv_, o, sizeof cv_);
                blocks_++;
                buf_len_ = 0;
            }
        }
        // whole aligned subtrees straight from the input, while more input follows them (the last
        // chunk stays pending: finalize may have to make it the root)
        if (buf_len_ == 0 && blocks_ == 0 && len > SUBTREE_MIN * 1024) {
            uint64_t n = 1;
            while (n * 2 * 1024 < len) n *= 2;
            while (n > 1 && (chunk_ & (n - 1))) n /= 2;
            if (n >= SUBTREE_MIN) {
                uint32_t cv[8] = {};
                if (plan && plan->record)
                    plan->subs.push_back({p, chunk_, n});  // a placeholder value on the dry run
                else if (plan)
                    std::memcpy(cv, plan->cvs[plan->next++].data(), sizeof cv);
                else
                    subtree_cv(key_, flags_, p, chunk_, n, cv);
                int level = 0;
                while ((1ull << level) < n) level++;
                push_subtree(cv, level);
                chunk_ += n;
                p += n * 1024;
                len -= n * 1024;
                continue;
            }
        }
        size_t take = 64 - buf_len_;
        if (take > len) take = len;
        std::memcpy(buf_ + buf_len_, p, take);
        buf_len_ += (uint32_t)take;
        p += take;
        len -= take;
    }
}

void Blake3::update_u64(uint64_t x) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(x >> (8 * i));
    update(b, 8);
}

Blake3::Output Blake3::finalize() const {
    Output o;
    std::memcpy(o.cv, cv_, 32);
    uint8_t blk[64] = {0};
    std::memcpy(blk, buf_, buf_len_);
    load_words(blk, o.block, 16);
    o.counter = chunk_;
    o.len = buf_len_;
    o.flags = flags_ | (blocks_ == 0 ? CHUNK_START : 0) | CHUNK_END;
    for (int level = 0; level < 64; level++) {
        if (!(has_ >> level & 1)) continue;
        uint32_t w[16];
        compress(o.cv, o.block, o.counter, o.len, o.flags, w);
        std::memcpy(o.block, subtree_[level], 32);
        std::memcpy(o.block + 8, w, 32);
        std::memcpy(o.cv, key_, 32);
        o.counter = 0;
        o.len = 64;
        o.flags = flags_ | PARENT;
    }
    return o;
}

void Blake3::Output::read(uint64_t pos, uint8_t *out, size_t n) const {
    while (n) {
        uint32_t w[16];
        compress(cv, block, pos / 64, len, flags | ROOT, w);
        uint8_t b[64];
        std::memcpy(b, w, 64);  // little-endian host
        size_t off = pos % 64, take = 64 - off;
        if (take > n) take = n;
        std::memcpy(out, b + off, take);
        out += take;
        pos += take;
        n -= take;
    }
}

}  // namespace lbn

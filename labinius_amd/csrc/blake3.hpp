// BLAKE3 (hash, keyed_hash, derive_key, XOF) from the public specification, as the reference's
// transcript and seed expansions use the `blake3` crate (crates/pcs/src/challenge.rs:66-160,
// scheme/params.rs:321, scheme/witness.rs:66-71). Incremental, cheaply clonable (the transcript
// clones its state on every derivation).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace lbn {

class Blake3 {
   public:
    static Blake3 hasher();
    static Blake3 keyed(const uint8_t key[32]);
    static Blake3 derive_key(const std::string &context);

    void update(const void *data, size_t len);
    void update_u64(uint64_t x);

    // The root node, from which any number of output bytes can be read.
    struct Output {
        uint32_t cv[8];
        uint32_t block[16];
        uint64_t counter;
        uint32_t len;
        uint32_t flags;
        void read(uint64_t pos, uint8_t *out, size_t n) const;
    };
    Output finalize() const;

   private:
    struct Plan;
    Blake3(const uint32_t key[8], uint32_t flags);
    void update_loop(const uint8_t *p, size_t len, Plan *plan);
    void close_chunk();
    void push_subtree(const uint32_t cv[8], int level);
    uint32_t key_[8];
    uint32_t flags_;
    // the chunk being filled
    uint32_t cv_[8];
    uint64_t chunk_;
    uint8_t buf_[64];
    uint32_t buf_len_;
    uint32_t blocks_;
    // merged subtrees: subtree[level] holds the CV of a complete subtree of 2^level chunks waiting
    // for its right sibling; has_ bit `level` says whether it is occupied.
    uint32_t subtree_[64][8];
    uint64_t has_;
};

// A sequential reader over an Output (the OutputReader the reference fills buffers from).
class Xof {
   public:
    explicit Xof(const Blake3::Output &o) : o_(o), pos_(0) {}
    void fill(uint8_t *out, size_t n) {
        o_.read(pos_, out, n);
        pos_ += n;
    }

   private:
    Blake3::Output o_;
    uint64_t pos_;
};

}  // namespace lbn

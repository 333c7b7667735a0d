// A share of a range per thread, for the host's parallel loops (BLAKE3 subtrees, the bit-dropped
// digits, the challenge sampler, the staged witness uploads).
#pragma once

#include <atomic>
#include <cstddef>
#include <thread>
#include <vector>

namespace lbn {

// Runs work(k) for the shares k < nt on up to nt threads: this thread takes share 0, and any
// share that no thread can be had for. No exception leaves a thread or this call; returns false
// when a share threw (its work is then unfinished).
template <class F>
bool run_shares(size_t nt, F &&work) {
    std::atomic<bool> ok{true};
    auto share = [&](size_t k) {
        try {
            work(k);
        } catch (...) {
            ok = false;
        }
    };
    if (nt <= 1) {
        nt = 1;
        share(0);
        return ok;
    }
    std::vector<std::thread> th;
    try {
        th.reserve(nt - 1);
    } catch (...) {
    }
    for (size_t k = 1; k < nt; k++) {
        try {
            th.emplace_back(share, k);
        } catch (...) {
            share(k);
        }
    }
    share(0);
    for (auto &x : th) x.join();
    return ok;
}

// Runs work(lo, hi) over nt contiguous shares of [0, n), as run_shares.
template <class F>
bool run_ranges(size_t n, size_t nt, F &&work) {
    return run_shares(nt, [&](size_t k) { work(n * k / nt, n * (k + 1) / nt); });
}

}  // namespace lbn

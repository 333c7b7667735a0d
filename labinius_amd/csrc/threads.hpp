// A share of a range per thread, for the host's parallel loops (BLAKE3 subtrees, the bit-dropped
// digits, the challenge sampler).
#pragma once

#include <atomic>
#include <cstddef>
#include <thread>
#include <vector>

namespace lbn {

// Runs work(lo, hi) over nt contiguous shares of [0, n) on up to nt threads: this thread takes the
// first share, and any share that no thread can be had for. No exception leaves a thread or this
// call; returns false when a share threw (its range is then unfinished).
template <class F>
bool run_ranges(size_t n, size_t nt, F &&work) {
    std::atomic<bool> ok{true};
    auto share = [&](size_t k) {
        try {
            work(n * k / nt, n * (k + 1) / nt);
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

}  // namespace lbn

// The wire forms (wire/mod.rs, wire/rans.rs): the bit-packed commitment, bit-dropped digits and
// row evaluation, and the rANS-coded folded witness. The prover's side runs on the device over
// what the commit and the fold left in HBM (only the packed bytes come back); the verifier's side
// (unpacking, decoding) runs on the host, as the verifier does.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "launch.hpp"

namespace lbn {

void wire_free(lbn_wire_ws &w) {
    void *bufs[] = {w.d_in,     w.d_staged, w.d_escw, w.d_cnt, w.d_offs,
                    w.d_stream, w.d_pack,   w.d_stat, w.d_hist, w.d_sym};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t &e : w.ev)
        if (e) (void)hipEventDestroy(e);
    w = lbn_wire_ws{};
}

namespace {

constexpr uint32_t SCALE_BITS = 12, SCALE = 1u << SCALE_BITS;  // wire/mod.rs:56-58
constexpr size_t HEADER = 16;                                  // wire/mod.rs:61
constexpr uint32_t MAX_ELEMENTS = 1u << 24;                    // wire/mod.rs:64
constexpr uint32_t NHIST = (1u << 16) + 1;                     // every i16 symbol and the escape

// wire/mod.rs:77-79
uint32_t field_bits(uint32_t q) { return 32 - (uint32_t)__builtin_clz(q - 1); }

// bd/mod.rs:58-64
uint32_t top_bound(uint32_t q0, uint32_t d) { return ((q0 - 1) + (1u << (d - 1))) >> d; }
uint32_t top_field(uint32_t q0, uint32_t d) { return 32 - (uint32_t)__builtin_clz(top_bound(q0, d)); }

// rans.rs:220-223
uint32_t raw_bits(uint32_t length) { return length <= 1 ? 0 : 32 - (uint32_t)__builtin_clz(length - 1); }

void put_le32(uint8_t *p, uint32_t v) {
    for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i));
}
uint32_t get_le32(const uint8_t *p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

// Fields least significant bit first with nothing between them, the last byte zero-padded
// (wire/mod.rs:86-138).
struct BitSink {
    std::vector<uint8_t> &out;
    uint64_t acc = 0;
    uint32_t n = 0;
    explicit BitSink(std::vector<uint8_t> &o) : out(o) {}
    void put(uint64_t v, uint32_t bits) {  // bits <= 56
        acc |= (v & ((1ull << bits) - 1)) << n;
        n += bits;
        for (; n >= 8; n -= 8, acc >>= 8) out.push_back((uint8_t)acc);
    }
    void gamma(uint64_t v) {  // v >= 1: k zeros, the leading one, the k bits below it
        const uint32_t k = 63 - (uint32_t)__builtin_clzll(v);
        put(0, k);
        put(1, 1);
        put(v, k);
    }
    void finish() {
        if (n) out.push_back((uint8_t)acc);
        acc = 0;
        n = 0;
    }
};

// A cursor over a bit string (wire/mod.rs:141-218). A read fails (Truncated) when fewer bits
// remain than it asks for; bytes consumed = ceil(bits read / 8).
struct BitSource {
    const uint8_t *b;
    size_t len;
    uint64_t bit = 0;
    int get(uint32_t bits, uint64_t &v) {  // bits <= 32
        if (bit + bits > 8 * (uint64_t)len) return LBN_WIRE_TRUNCATED;
        const size_t at = (size_t)(bit >> 3);
        uint64_t w = 0;
        for (size_t i = 0; i < 5 && at + i < len; i++) w |= (uint64_t)b[at + i] << (8 * i);
        v = (w >> (bit & 7)) & ((1ull << bits) - 1);
        bit += bits;
        return LBN_OK;
    }
    int gamma(uint64_t &v) {
        uint32_t k = 0;
        for (;;) {
            uint64_t one;
            if (int e = get(1, one)) return e;
            if (one) break;
            if (++k > 32) return LBN_WIRE_MALFORMED;
        }
        uint64_t lo;
        if (int e = get(k, lo)) return e;
        v = (1ull << k) | lo;
        return LBN_OK;
    }
    size_t consumed() const { return (size_t)((bit + 7) / 8); }
};

int truncated() { return fail(LBN_WIRE_TRUNCATED, "the byte string ran out before the object did"); }
int malformed() { return fail(LBN_WIRE_MALFORMED, "not a coding of an object of the expected shape"); }
int wire_error(int e) { return e == LBN_WIRE_TRUNCATED ? truncated() : malformed(); }

int copy_out(const std::vector<uint8_t> &b, uint8_t *out, size_t cap, size_t *len) {
    if (len) *len = b.size();
    if (!out) return LBN_OK;
    if (cap < b.size()) return fail(LBN_ERR_SHAPE, "output buffer too small");
    if (!b.empty()) std::memcpy(out, b.data(), b.size());  // an empty vector's data() may be null
    return LBN_OK;
}

// The commitment matrix's record: 4 x r entries (VerticallyAlignedMatrix::iter order, column by
// column, rows 0..3), each limb's 162 slots as residues (wire/mod.rs:226-243).
PackDesc commitment_desc(const Shape &s) {
    PackDesc d{};
    const uint32_t limbs = (uint32_t)s.primes.size();
    d.nseg = limbs;
    uint32_t bit = 0;
    for (uint32_t l = 0; l < limbs; l++) {
        PackSeg &g = d.seg[l];
        g.count = N162;
        g.bits = field_bits(s.primes[l]);
        g.q = s.primes[l];
        g.bit0 = bit;
        g.off = (uint64_t)l * N162;
        g.rec_stride = (uint64_t)limbs * N162;
        g.stride = 1;
        bit += N162 * g.bits;
    }
    d.rec_bits = bit;
    d.total_bits = (uint64_t)bit * 4 * s.r;
    return d;
}

// The bit-dropped digits' record: per coefficient the top, then each digit (wire/mod.rs:386-397).
PackDesc dropped_desc(const Shape &s) {
    PackDesc d{};
    const uint32_t limbs = (uint32_t)s.primes.size();
    const uint64_t count = (uint64_t)s.r * N;
    d.nseg = limbs;
    uint32_t bit = 0;
    for (uint32_t l = 0; l < limbs; l++) {
        PackSeg &g = d.seg[l];
        g.count = 1;
        g.bits = l == 0 ? top_field(s.primes[0], s.dropped_bits) : field_bits(s.primes[l]);
        g.q = 0;
        g.bit0 = bit;
        g.off = l * count;
        g.rec_stride = 1;
        g.stride = 0;
        bit += g.bits;
    }
    d.rec_bits = bit;
    d.total_bits = (uint64_t)bit * count;
    return d;
}

template <typename T>
int grow(T **p, size_t &cap, size_t need) {
    if (need <= cap) return LBN_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    int rc = dalloc(p, need);
    if (rc) return rc;
    cap = need;
    return LBN_OK;
}

int wire_alloc(lbn_wire_ws &w) {
    int rc;
    if ((rc = dalloc(&w.d_stat, RANS_STAT))) return rc;
    if ((rc = dalloc(&w.d_hist, NHIST))) return rc;
    if ((rc = dalloc(&w.d_sym, NHIST))) return rc;
    for (hipEvent_t &e : w.ev) HIP_TRY(hipEventCreate(&e));
    return LBN_OK;
}

// The encoder's fixed scratch, allocated at its first use (all of it, or none).
int wire_ready(lbn_wire_ws &w) {
    if (w.d_stat && w.d_hist && w.d_sym && w.ev[4]) return LBN_OK;
    const int rc = wire_alloc(w);
    if (rc) wire_free(w);
    return rc;
}

// A packed object on the device from src, back to the host.
int pack_device(lbn_prover *pr, const PackDesc &d, const uint16_t *src, uint8_t *out, size_t cap, size_t *len) {
    const size_t bytes = (size_t)((d.total_bits + 7) / 8), words = (size_t)((d.total_bits + 31) / 32);
    if (len) *len = bytes;
    if (!out) return LBN_OK;
    if (cap < bytes) return fail(LBN_ERR_SHAPE, "output buffer too small");
    lbn_wire_ws &w = pr->wire;
    int rc = grow(&w.d_pack, w.cap_pack, words);
    if (rc) return rc;
    HIP_TRY(launch_wire_pack(pr->stream, d, src, w.d_pack));
    HIP_TRY(hipMemcpyAsync(out, w.d_pack, bytes, hipMemcpyDeviceToHost, pr->stream));
    HIP_TRY(hipStreamSynchronize(pr->stream));
    return LBN_OK;
}

// rans.rs:236-271: the counts (the last one the escape's, zero) scaled to a total of SCALE by
// largest remainder, every occupied symbol floored at 1; past SCALE - 1 occupied symbols the
// rarest are folded into the escape.
std::vector<uint32_t> quantize(std::vector<uint64_t> w) {
    const size_t esc = w.size() - 1;
    std::vector<uint32_t> f(w.size(), 0);
    std::vector<size_t> occ;
    for (size_t i = 0; i < esc; i++)
        if (w[i]) occ.push_back(i);
    if (occ.size() > SCALE - 1) {
        std::sort(occ.begin(), occ.end(), [&](size_t a, size_t b) { return w[a] != w[b] ? w[a] > w[b] : a < b; });
        for (size_t j = SCALE - 1; j < occ.size(); j++) {
            w[esc] += w[occ[j]];
            w[occ[j]] = 0;
        }
    }
    uint64_t total = 0;
    occ.clear();
    for (size_t i = 0; i < w.size(); i++) {
        total += w[i];
        if (w[i]) occ.push_back(i);
    }
    if (occ.empty()) return f;
    const uint64_t spare = SCALE - occ.size();
    uint64_t used = 0;
    std::vector<std::pair<uint64_t, size_t>> rem;
    rem.reserve(occ.size());
    for (size_t i : occ) {
        const uint64_t share = w[i] * spare / total;
        f[i] = 1 + (uint32_t)share;
        used += share;
        rem.emplace_back(w[i] * spare - share * total, i);
    }
    std::sort(rem.begin(), rem.end(), [](const std::pair<uint64_t, size_t> &a, const std::pair<uint64_t, size_t> &b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    for (uint64_t j = 0; j < spare - used; j++) f[rem[j].second]++;
    return f;
}

// rans.rs:18-99 over `count` ring elements at d_v (device), on the prover's stream. The min / max
// and the counts are taken on the device; the quantised table is built here (it is at most 65537
// entries) and goes back as the encoder's symbols; one wavefront runs the 64 interleaved coders;
// the stream is reversed on the device and only the coded bytes come back.
int encode_device(lbn_prover *pr, const int16_t *d_v, size_t count, int ntt, uint16_t base_q, uint8_t *out,
                  size_t cap, size_t *len) {
    if (count > MAX_ELEMENTS) return fail(LBN_ERR_SHAPE, "more elements than the wire format carries");
    std::vector<uint8_t> head(HEADER, 0);
    head[0] = 1;
    head[1] = ntt ? 1 : 0;
    head[2] = (uint8_t)base_q;
    head[3] = (uint8_t)(base_q >> 8);
    put_le32(&head[4], (uint32_t)count);
    const uint64_t coefs = (uint64_t)count * N;
    if (coefs == 0) return copy_out(head, out, cap, len);

    lbn_wire_ws &w = pr->wire;
    hipStream_t st = pr->stream;
    int rc = wire_ready(w);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(w.ev[0], st));
    HIP_TRY(launch_wire_stats(st, d_v, coefs, (int *)w.d_stat, w.d_hist, NHIST));
    HIP_TRY(hipEventRecord(w.ev[1], st));
    int32_t mm[2];
    HIP_TRY(hipMemcpyAsync(mm, w.d_stat, sizeof(mm), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const int32_t low = mm[0];
    const uint32_t length = (uint32_t)(mm[1] - mm[0] + 1);
    put_le32(&head[8], (uint32_t)low);
    put_le32(&head[12], length);
    std::vector<uint64_t> counts(length + 1, 0);
    HIP_TRY(hipMemcpyAsync(counts.data(), w.d_hist, length * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));

    const std::vector<uint32_t> freq = quantize(counts);
    BitSink bs(head);
    for (uint32_t f : freq) bs.gamma((uint64_t)f + 1);
    bs.finish();
    std::vector<RansSym> sym(length + 1);
    uint32_t start = 0;
    uint64_t escaped = 0;
    for (uint32_t i = 0; i <= length; i++) {
        const uint32_t f = freq[i];
        if (f == 0) {
            sym[i] = RansSym{0.0, 0, RANS_ESCAPED};
            if (i < length) escaped += counts[i];
        } else {
            // 1/f rounded up: exact for a power of two, else the double above the nearest one
            const double rcp = (f & (f - 1)) == 0 ? 1.0 / f : std::nextafter(1.0 / f, 2.0);
            sym[i] = RansSym{rcp, f == SCALE ? 0xffffffffu : (f << 20) - 1, (SCALE - f) | start << 16};
        }
        start += f;
    }
    const uint64_t nwords = coefs + escaped + 1, groups = (coefs + RANS_LANES - 1) / RANS_LANES;
    if (groups > w.cap_groups) {
        size_t c1 = w.cap_groups * RANS_LANES, c2 = c1, c3 = w.cap_groups, c4 = c3;
        w.cap_groups = 0;
        if ((rc = grow(&w.d_staged, c1, groups * RANS_LANES))) return rc;
        if ((rc = grow(&w.d_escw, c2, groups * RANS_LANES))) return rc;
        if ((rc = grow(&w.d_cnt, c3, groups + (groups + 63) / 64))) return rc;
        if ((rc = grow(&w.d_offs, c4, (groups + 63) / 64))) return rc;
        w.cap_groups = groups;
    }
    if ((rc = grow(&w.d_stream, w.cap_words, nwords))) return rc;
    HIP_TRY(hipMemcpyAsync(w.d_sym, sym.data(), sym.size() * sizeof(RansSym), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(w.ev[2], st));
    HIP_TRY(launch_rans_encode(st, d_v, coefs, w.d_sym, length + 1, low, raw_bits(length), w.d_staged, w.d_escw,
                               w.d_stat));
    HIP_TRY(hipEventRecord(w.ev[3], st));
    HIP_TRY(launch_rans_scatter(st, w.d_staged, w.d_escw, w.d_cnt, w.d_offs, w.d_stat, groups, w.d_stream));
    HIP_TRY(hipEventRecord(w.ev[4], st));
    uint32_t tail[RANS_STAT - 2];
    HIP_TRY(hipMemcpyAsync(tail, w.d_stat + 2, sizeof(tail), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const uint64_t top = (uint64_t)tail[0] | (uint64_t)tail[1] << 32;
    if (top > nwords) return fail(LBN_ERR_HIP, "the encoder pushed more words than it can");
    const size_t base = head.size(), total = base + 4 * RANS_LANES + 2 * top;
    w.ms[0] = elapsed(w.ev[0], w.ev[4]);
    w.ms[1] = elapsed(w.ev[0], w.ev[1]);
    w.ms[2] = elapsed(w.ev[2], w.ev[3]);
    if (len) *len = total;
    if (!out) return LBN_OK;
    if (cap < total) return fail(LBN_ERR_SHAPE, "output buffer too small");
    std::memcpy(out, head.data(), base);
    for (uint32_t k = 0; k < RANS_LANES; k++) put_le32(out + base + 4 * k, tail[2 + k]);
    if (top) {
        HIP_TRY(hipMemcpyAsync(out + base + 4 * RANS_LANES, w.d_stream, 2 * top, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return LBN_OK;
}

}  // namespace
}  // namespace lbn

using namespace lbn;

extern "C" {

int lbn_commitment_wire_bytes(const lbn_params *params, size_t *len) {
    if (!params || !len) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if (s.opening == LBN_OPENING_RECURSIVE) return fail(LBN_ERR_UNSUPPORTED, "a recursive commitment travels as T_Y");
    const PackDesc d = s.dropped_bits ? dropped_desc(s) : commitment_desc(s);
    *len = (size_t)((d.total_bits + 7) / 8);
    return LBN_OK;
}

int lbn_commit_pack(lbn_prover *pr, uint8_t *out, size_t cap, size_t *len) {
    if (!pr) return fail(LBN_ERR_ARG, "null argument");
    const Shape &s = pr->pp->shape;
    if (s.opening == LBN_OPENING_RECURSIVE) return fail(LBN_ERR_UNSUPPORTED, "a recursive commitment travels as T_Y");
    if (!pr->committed) return fail(LBN_ERR_ARG, "the prover has made no commitment");
    if (s.dropped_bits == 0) return pack_device(pr, commitment_desc(s), (const uint16_t *)pr->d_comp, out, cap, len);
    if (s.primes.size() > (size_t)PACK_SEGS) return fail(LBN_ERR_SHAPE, "too many moduli");
    if (!pr->bd_ready) {
        int rc = bd_digits_device(pr, s, pr->d_raw);
        if (rc) return rc;
        pr->bd_ready = true;
    }
    return pack_device(pr, dropped_desc(s), pr->d_bd_out, out, cap, len);
}

int lbn_pack_commitment(const lbn_params *params, const int16_t *components, uint8_t *out, size_t cap, size_t *len) {
    if (!params || !components) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if (s.dropped_bits || s.opening == LBN_OPENING_RECURSIVE)
        return fail(LBN_ERR_SHAPE, "not a commitment matrix (lbn_pack_dropped packs the bit-dropped digits)");
    const size_t limbs = s.primes.size();
    std::vector<uint8_t> b;
    b.reserve((size_t)((commitment_desc(s).total_bits + 7) / 8));
    BitSink bs(b);
    for (size_t e = 0; e < 4 * s.r; e++)
        for (size_t k = 0; k < limbs; k++) {
            const int32_t q = (int32_t)s.primes[k];
            const uint32_t bits = field_bits(s.primes[k]);
            const int16_t *v = components + (e * limbs + k) * N162;
            for (int i = 0; i < N162; i++) bs.put((uint64_t)(((int32_t)v[i] % q + q) % q), bits);
        }
    bs.finish();
    return copy_out(b, out, cap, len);
}

int lbn_unpack_commitment(const lbn_params *params, const uint8_t *bytes, size_t len, int16_t *components) {
    if (!params || (!bytes && len) || !components) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if (s.dropped_bits) return fail(LBN_ERR_SHAPE, "bit-dropped parameters: lbn_unpack_dropped");
    if (s.opening == LBN_OPENING_RECURSIVE || len != (size_t)((commitment_desc(s).total_bits + 7) / 8))
        return malformed();
    const size_t limbs = s.primes.size();
    BitSource r{bytes, len};
    for (size_t e = 0; e < 4 * s.r; e++)
        for (size_t k = 0; k < limbs; k++) {
            const uint32_t q = s.primes[k], bits = field_bits(q), half = (q - 1) / 2;
            int16_t *v = components + (e * limbs + k) * N162;
            for (int i = 0; i < N162; i++) {
                uint64_t x;
                if (int err = r.get(bits, x)) return wire_error(err);
                if (x >= q) return malformed();
                v[i] = (int16_t)(x > half ? (int32_t)x - (int32_t)q : (int32_t)x);
            }
        }
    return LBN_OK;
}

int lbn_unpack_dropped(const lbn_params *params, const uint8_t *bytes, size_t len, uint16_t *top, uint16_t *digits) {
    if (!params || (!bytes && len) || !top || !digits) return fail(LBN_ERR_ARG, "null argument");
    Shape s;
    int rc = make_shape(params, &s);
    if (rc) return rc;
    if (s.opening == LBN_OPENING_RECURSIVE || s.dropped_bits == 0 || s.primes.size() < 2 ||
        len != (size_t)((dropped_desc(s).total_bits + 7) / 8))
        return malformed();
    const size_t limbs = s.primes.size(), count = s.r * N;
    const uint32_t tb = top_field(s.primes[0], s.dropped_bits), bound = top_bound(s.primes[0], s.dropped_bits);
    BitSource r{bytes, len};
    for (size_t i = 0; i < count; i++) {
        uint64_t x;
        if (int err = r.get(tb, x)) return wire_error(err);
        if (x > bound) return malformed();
        top[i] = (uint16_t)x;
        for (size_t k = 1; k < limbs; k++) {
            if (int err = r.get(field_bits(s.primes[k]), x)) return wire_error(err);
            if (x >= s.primes[k]) return malformed();
            digits[(k - 1) * count + i] = (uint16_t)x;
        }
    }
    return LBN_OK;
}

int lbn_pack_f162(const lbn_f162 *x, size_t n, uint8_t *out, size_t cap, size_t *len) {
    if (!x && n) return fail(LBN_ERR_ARG, "null argument");
    std::vector<uint8_t> b;
    b.reserve((n * 162 + 7) / 8);
    BitSink bs(b);
    for (size_t i = 0; i < n; i++) {
        bs.put(x[i].w[0], 32);
        bs.put(x[i].w[0] >> 32, 32);
        bs.put(x[i].w[1], 32);
        bs.put(x[i].w[1] >> 32, 32);
        bs.put(x[i].w[2], 34);
    }
    bs.finish();
    return copy_out(b, out, cap, len);
}

int lbn_unpack_f162(const uint8_t *bytes, size_t len, size_t n, lbn_f162 *x) {
    if ((!bytes && len) || (!x && n)) return fail(LBN_ERR_ARG, "null argument");
    if (len != (n * 162 + 7) / 8) return malformed();
    BitSource r{bytes, len};
    for (size_t i = 0; i < n; i++) {
        uint64_t a, b;
        for (int j = 0; j < 2; j++) {
            if (int err = r.get(32, a)) return wire_error(err);
            if (int err = r.get(32, b)) return wire_error(err);
            x[i].w[j] = a | b << 32;
        }
        if (int err = r.get(32, a)) return wire_error(err);
        if (int err = r.get(2, b)) return wire_error(err);
        x[i].w[2] = a | b << 32;
    }
    return LBN_OK;
}

size_t lbn_wire_encode_bound(size_t count) {
    // header, at most 25 bits of gamma code a symbol, the states, at most two words a coefficient
    return HEADER + ((size_t)NHIST * 25 + 7) / 8 + 4 * RANS_LANES + 2 * (2 * count * N + 1);
}

int lbn_fold_encode(lbn_prover *pr, uint16_t base_q, uint8_t *out, size_t cap, size_t *len) {
    if (!pr) return fail(LBN_ERR_ARG, "null argument");
    if (!pr->folded) return fail(LBN_ERR_ARG, "the prover has folded nothing");
    return encode_device(pr, pr->d_v, pr->pp->shape.len_ring, 0, base_q, out, cap, len);
}

int lbn_wire_encode_device(lbn_prover *pr, const int16_t *v, size_t count, int ntt, uint16_t base_q, uint8_t *out,
                           size_t cap, size_t *len) {
    if (!pr || (!v && count)) return fail(LBN_ERR_ARG, "null argument");
    if (count > MAX_ELEMENTS) return fail(LBN_ERR_SHAPE, "more elements than the wire format carries");
    lbn_wire_ws &w = pr->wire;
    int rc = grow(&w.d_in, w.cap_in, count * N);
    if (rc) return rc;
    if (count) HIP_TRY(hipMemcpyAsync(w.d_in, v, count * N * sizeof(int16_t), hipMemcpyHostToDevice, pr->stream));
    return encode_device(pr, w.d_in, count, ntt, base_q, out, cap, len);
}

int lbn_wire_decode(const uint8_t *bytes, size_t len, int16_t *v, size_t cap_count, uint32_t *count_out, int *ntt) {
    if ((!bytes && len) || (!v && cap_count)) return fail(LBN_ERR_ARG, "null argument");
    if (len < HEADER) return truncated();
    if (bytes[0] != 1 || bytes[1] > 1) return malformed();
    const uint32_t count = get_le32(bytes + 4), length = get_le32(bytes + 12);
    const int32_t low = (int32_t)get_le32(bytes + 8);
    if (ntt) *ntt = bytes[1];
    if (count_out) *count_out = count;
    if (count == 0) return len == HEADER && length == 0 ? LBN_OK : malformed();
    if (length == 0 || length > (1u << 16) || count > MAX_ELEMENTS || low < -32768 ||
        (int64_t)low + length - 1 > 32767)
        return malformed();
    if (count > cap_count) return fail(LBN_ERR_SHAPE, "the message has more elements than the buffer holds");

    BitSource r{bytes + HEADER, len - HEADER};
    std::vector<uint32_t> freq(length + 1);
    uint64_t total = 0;
    for (uint32_t i = 0; i <= length; i++) {
        uint64_t g;
        if (int err = r.gamma(g)) return wire_error(err);
        total += g - 1;
        if (total > SCALE) return malformed();
        freq[i] = (uint32_t)(g - 1);
    }
    if (total != SCALE) return malformed();
    // the slot table: frequency | (slot - start) << 16, and the symbol
    std::vector<uint32_t> fb(SCALE), sy(SCALE);
    for (uint32_t i = 0, start = 0; i <= length; start += freq[i], i++)
        for (uint32_t j = 0; j < freq[i]; j++) {
            fb[start + j] = freq[i] | j << 16;
            sy[start + j] = i;
        }
    const size_t used = HEADER + r.consumed();
    if (len - used < 4 * RANS_LANES) return truncated();
    uint32_t x[RANS_LANES];
    for (uint32_t k = 0; k < RANS_LANES; k++) {
        x[k] = get_le32(bytes + used + 4 * k);
        if (x[k] < RANS_LOW) return malformed();
    }
    const uint8_t *stream = bytes + used + 4 * RANS_LANES;
    const size_t stream_len = len - used - 4 * RANS_LANES;
    if (stream_len % 2) return malformed();
    const size_t nwords = stream_len / 2;
    const uint32_t raw = raw_bits(length);
    size_t pos = 0;
    auto next_word = [&]() -> uint32_t {
        return pos < nwords ? (uint32_t)stream[2 * pos] | (uint32_t)stream[2 * pos + 1] << 8 : 0;
    };
    const uint64_t coefs = (uint64_t)count * N;
    for (uint64_t g = 0; g * RANS_LANES < coefs; g++) {
        const uint64_t lanes = std::min<uint64_t>(RANS_LANES, coefs - g * RANS_LANES);
        for (uint64_t k = 0; k < lanes; k++) {
            const uint32_t slot = x[k] & (SCALE - 1);
            uint32_t s = sy[slot];
            uint32_t y = (fb[slot] & 0xffff) * (x[k] >> SCALE_BITS) + (fb[slot] >> 16);
            if (y < RANS_LOW) {
                x[k] = y << 16 | next_word();
                pos++;
            } else {
                x[k] = y;
            }
            if (s == length) {  // the escape: the symbol's raw bits follow
                s = 0;
                if (raw) {
                    s = x[k] & ((1u << raw) - 1);
                    y = x[k] >> raw;
                    if (y < RANS_LOW) {
                        x[k] = y << 16 | next_word();
                        pos++;
                    } else {
                        x[k] = y;
                    }
                }
                if (s >= length) return malformed();
            }
            v[g * RANS_LANES + k] = (int16_t)(low + (int32_t)s);
        }
        if (pos > nwords) return truncated();
    }
    if (2 * pos != stream_len) return malformed();
    for (uint32_t k = 0; k < RANS_LANES; k++)
        if (x[k] != RANS_LOW) return malformed();
    return LBN_OK;
}

int lbn_wire_timings(const lbn_prover *pr, float *ms, size_t n) {
    if (!pr || !ms) return fail(LBN_ERR_ARG, "null argument");
    for (size_t i = 0; i < n && i < 3; i++) ms[i] = pr->wire.ms[i];
    return LBN_OK;
}

}  // extern "C"

"""The wire forms (wire/mod.rs, wire/rans.rs) on the CPU.

The oracle's restatement is pinned by the reference's own wire tests (tests/wire.rs), restated
case for case: exact sizes (162 bits an F162, ceil(log2 q) bits a slot, the 16 + 5 + 4 * 64 + 2
bytes of a message whose words all come from one lane), exact round trips on honest rounds and on
adversarial messages (the escape symbol included), and which error a broken stream gets. Known
answers derived from the format's description (least significant bit first, the largest-remainder
quantisation worked by hand) and an independent decoder written from the rANS definition (slot
search, no tables) pin the byte layout beyond what round trips can. The host side of the product
(unpack, decode, the host packers) is checked against the oracle byte for byte.
"""
import ctypes

import numpy as np
import pytest

N = 648
MODULI = [2917, 3889, 4861, 9721, 12637, 17497, 19441]


def second(base):
    return 9721 if base == 3889 else 3889


def _rng(seed):
    return np.random.default_rng(seed)


def _honest_round(orc, wl, cl, primes, seed=7):
    """tests/wire.rs:25-43 through the oracle: commit, point, row evaluation, challenges, fold."""
    mseed = bytes([21]) * 32
    params = orc.make_params(wl, cl, primes, mseed)
    L, cols = (1 << (wl - cl)) // 4, 1 << cl
    a = orc.key_expand(primes, L, orc.matrix_seed(mseed))
    w = orc.witness_random(bytes([23 + seed]) * 32, 1 << wl)
    raw, comp = orc.commit(primes, L, cols, a, w)
    t = orc.Transcript(b"labinius/test/wire")
    p0, p1 = orc.derive_evaluation_point(t, params, comp)
    u = orc.row_evaluate(w, p0)
    ch = orc.derive_folding_challenges(t, u)
    v = orc.fold(primes[0], L, cols, w, ch)
    return comp, u, v, cols


# ---- the bit writer's layout: known answers --------------------------------------------------------

def test_f162_bit_layout_known_answers(orc):
    # bits least significant first, element i at bit 162 i (wire/mod.rs:99-110, 445-453)
    assert orc.pack_f162([[1, 0, 0]]) == bytes([1]) + bytes(20)
    assert orc.pack_f162([[0, 0, 1 << 33]]) == bytes(20) + bytes([2])
    assert orc.pack_f162([[0, 0, 0], [1, 0, 0]]) == bytes(20) + bytes([4]) + bytes(20)
    assert orc.pack_f162([[0, 1 << 63, 0]]) == bytes(15) + bytes([0x80]) + bytes(5)
    # the top word keeps its 34 bits only
    assert orc.pack_f162([[0, 0, (1 << 34) | 5]]) == bytes(16) + bytes([5]) + bytes(4)


def test_commitment_bit_layout_known_answer(orc):
    # one column, primes [3889 (12 bits), 9721 (14 bits)]: entry 0's limb-0 slot 0 at bit 0, slot 1
    # at bit 12, limb 1 from bit 12 * 162; -1 is q - 1 (rem_euclid)
    primes, cols = [3889, 9721], 1
    comp = np.zeros((cols, 4, 2, 162), np.int16)
    comp[0, 0, 0, 0] = 5
    comp[0, 0, 0, 1] = -1
    comp[0, 0, 1, 0] = 3
    b = orc.pack_commitment(primes, cols, comp)
    assert len(b) == (4 * 162 * 26 + 7) // 8
    bits = int.from_bytes(b, "little")
    assert bits & 0xFFF == 5
    assert (bits >> 12) & 0xFFF == 3888
    assert (bits >> (12 * 162)) & 0x3FFF == 3
    assert bits >> (12 * 162 + 14) == 0


# ---- bit-packed uniform objects (tests/wire.rs:61-177) ---------------------------------------------

def test_row_evaluation_packs_at_162_bits(orc):
    rng = _rng(0xF162)
    for count in [0, 1, 2, 7, 256, 1024]:
        v = rng.integers(0, 1 << 63, (count, 3), dtype=np.uint64) * 2 + rng.integers(0, 2, (count, 3),
                                                                                       dtype=np.uint64)
        v[:, 2] &= np.uint64((1 << 34) - 1)
        b = orc.pack_f162(v)
        assert len(b) == (count * 162 + 7) // 8
        err, back = orc.unpack_f162(b, count)
        assert err == orc.WIRE_OK and np.array_equal(back, v)
    b = orc.pack_f162(rng.integers(0, 1 << 34, (256, 3), dtype=np.uint64))
    assert len(b) == 5184
    assert orc.unpack_f162(b[:5183], 256)[0] == orc.WIRE_MALFORMED


@pytest.mark.parametrize("base", MODULI)
def test_commitment_packs_at_the_residue_width(orc, base):
    primes = [base, second(base)]
    comp, _, _, cols = _honest_round(orc, 11, 3, primes)
    b = orc.pack_commitment(primes, cols, comp)
    width = sum(orc.residue_bits(q) for q in primes)
    assert len(b) == (4 * cols * 162 * width + 7) // 8 == orc.commitment_bytes(primes, cols)
    err, back = orc.unpack_commitment(primes, cols, b)
    assert err == orc.WIRE_OK and np.array_equal(back, comp)


def test_commitment_packer_takes_the_extremes(orc):
    primes, cols = [19441, 2917], 8
    for pick in range(3):
        comp = np.zeros((cols, 4, 2, 162), np.int16)
        for c in range(cols):
            for t in range(4):
                for k, q in enumerate(primes):
                    half = (q - 1) // 2
                    comp[c, t, k, :] = [0, half, -half][(pick + 4 * c + t) % 3]
        b = orc.pack_commitment(primes, cols, comp)
        err, back = orc.unpack_commitment(primes, cols, b)
        assert err == orc.WIRE_OK and np.array_equal(back, comp)


def _out_of_range_cases(cols):
    last = 4 * cols - 1
    return [(0, 0, 0), (0, 0, 160), (0, 1, 161), (1, 0, 17), (last, 1, 161), (last, 0, 161)]


def _set_field(b, bit, width):
    bad = bytearray(b)
    for i in range(bit, bit + width):
        bad[i // 8] |= 1 << (i % 8)
    return bytes(bad)


def test_out_of_range_residue_is_refused(orc):
    primes = [3889, 9721]
    comp, _, _, cols = _honest_round(orc, 11, 3, primes)
    b = orc.pack_commitment(primes, cols, comp)
    widths = [orc.residue_bits(q) for q in primes]
    per = sum(162 * w for w in widths)
    for element, limb, slot in _out_of_range_cases(cols):
        bit = element * per + sum(162 * w for w in widths[:limb]) + slot * widths[limb]
        assert orc.unpack_commitment(primes, cols, _set_field(b, bit, widths[limb]))[0] == orc.WIRE_MALFORMED
    assert orc.unpack_commitment(primes, cols, b[:-1])[0] == orc.WIRE_MALFORMED
    assert orc.unpack_commitment(primes, cols, b + b"\0")[0] == orc.WIRE_MALFORMED


def test_dropped_round_trips_and_refuses(orc):
    # wire/mod.rs:386-431: top (top_bits) then each digit; a top above its bound or a digit at its
    # modulus is refused
    primes, d, cols = [3889, 9721, 2917], 3, 4
    rng = _rng(5)
    bound = ((primes[0] - 1) + (1 << (d - 1))) >> d
    top = rng.integers(0, bound + 1, cols * N).astype(np.uint16)
    digits = np.stack([rng.integers(0, q, cols * N) for q in primes[1:]]).astype(np.uint16)
    b = orc.pack_dropped(primes, cols, d, top, digits)
    err, t2, d2 = orc.unpack_dropped(primes, cols, d, b)
    assert err == orc.WIRE_OK and np.array_equal(t2, top) and np.array_equal(d2, digits)
    tb = orc.top_bits(primes[0], d)
    per = tb + sum(orc.residue_bits(q) for q in primes[1:])
    bad = _set_field(b, 5 * per, tb)  # 2^tb - 1 > bound for these primes
    assert (1 << tb) - 1 > bound
    assert orc.unpack_dropped(primes, cols, d, bad)[0] == orc.WIRE_MALFORMED
    bad = _set_field(b, 7 * per + tb, orc.residue_bits(primes[1]))
    assert orc.unpack_dropped(primes, cols, d, bad)[0] == orc.WIRE_MALFORMED
    assert orc.unpack_dropped(primes, cols, d, b[:-1])[0] == orc.WIRE_MALFORMED


# ---- the quantiser ---------------------------------------------------------------------------------

def test_quantize_known_answer():
    import oracle as orc
    # 645 zeros and 3 ones: spare 4094; shares 4075 and 18 (remainders 30 and 618 of 648), so
    # 1 + share each and the one slot left to the larger remainder
    assert list(orc.rans_quantize([645, 3, 0])) == [4076, 20, 0]
    assert list(orc.rans_quantize([0, 0, 0])) == [0, 0, 0]
    assert list(orc.rans_quantize([7, 0])) == [4096, 0]
    # equal remainders: the smaller index first
    assert list(orc.rans_quantize([1, 1, 1, 0])) == [1366, 1365, 1365, 0]


def test_quantize_properties(orc):
    rng = _rng(3)
    for trial in range(20):
        n = int(rng.integers(2, 3000))
        counts = np.zeros(n + 1, np.uint64)
        occ = rng.choice(n, size=min(n, int(rng.integers(1, 400))), replace=False)
        counts[occ] = rng.integers(1, 10 ** int(rng.integers(1, 7)), len(occ))
        f = orc.rans_quantize(counts)
        assert f.sum() == 4096
        assert np.all((f[:-1] > 0) == (counts[:-1] > 0)) and f[-1] == 0
        # largest remainder: every frequency within one slot of its exact share
        spare = 4096 - len(occ)
        exact = 1 + counts.astype(np.float64) * spare / counts.sum()
        assert np.all(np.abs(f[counts > 0] - exact[counts > 0]) < 1.0 + 1e-9)


def test_quantize_folds_the_rarest_into_the_escape(orc):
    # 5000 occupied symbols: the 4095 most frequent keep slots (ties to the smaller index), the
    # rest go to the escape (rans.rs:243-250)
    counts = np.zeros(5001, np.uint64)
    counts[:5000] = 1
    counts[100:200] = 3
    f = orc.rans_quantize(counts)
    assert f.sum() == 4096
    assert np.all(f[100:200] > 0)
    # kept: the 100 threes, then the ones by index: 0..99 and 200..4094
    assert np.array_equal(np.flatnonzero(f[:5000]), np.arange(4095))
    # 4096 occupied slots with the escape: every frequency is 1
    assert np.all(f[:4095] == 1) and f[5000] == 1


# ---- an independent decoder (the rANS definition, no slot tables) --------------------------------

def _gamma_read(bits, pos):
    k = 0
    while bits[pos] == 0:
        k += 1
        pos += 1
    pos += 1
    value = 1 << k
    for i in range(k):
        value |= bits[pos + i] << i
    return value, pos + k


def _decode_by_definition(data):
    """Decode from the format description (wire/mod.rs:22-41): a state x gives slot x mod 4096,
    the symbol s whose [start, start + f) holds it, x <- f (x >> 12) + slot - start, then 16-bit
    refills from the word stream while x < 2^16."""
    count = int.from_bytes(data[4:8], "little")
    low = int.from_bytes(data[8:12], "little", signed=True)
    length = int.from_bytes(data[12:16], "little")
    if count == 0:
        return np.zeros((0, N), np.int16)
    bits = [(byte >> i) & 1 for byte in data[16:] for i in range(8)]
    pos, freq = 0, []
    for _ in range(length + 1):
        g, pos = _gamma_read(bits, pos)
        freq.append(g - 1)
    start = np.concatenate([[0], np.cumsum(freq)])
    at = 16 + (pos + 7) // 8
    x = [int.from_bytes(data[at + 4 * k:at + 4 * k + 4], "little") for k in range(64)]
    words = data[at + 256:]
    wpos = 0
    raw = (length - 1).bit_length() if length > 1 else 0
    out = []
    n = count * N
    for i in range(n):
        k = i % 64
        slot = x[k] & 4095
        s = int(np.searchsorted(start, slot, side="right") - 1)
        x[k] = freq[s] * (x[k] >> 12) + slot - int(start[s])
        if x[k] < 1 << 16:
            x[k] = (x[k] << 16) | int.from_bytes(words[2 * wpos:2 * wpos + 2], "little")
            wpos += 1
        if s == length:
            s = x[k] & ((1 << raw) - 1)
            x[k] >>= raw
            if x[k] < 1 << 16:
                x[k] = (x[k] << 16) | int.from_bytes(words[2 * wpos:2 * wpos + 2], "little")
                wpos += 1
        out.append(low + s)
    assert 2 * wpos == len(words) and all(v == 1 << 16 for v in x)
    return np.array(out, np.int16).reshape(count, N)


def test_encoder_matches_the_definition_decoder(orc):
    rng = _rng(11)
    v = np.clip(np.rint(rng.normal(0, 30, (2, N))), -200, 200).astype(np.int16)
    assert np.array_equal(_decode_by_definition(orc.rans_encode(v, 3889)), v)
    # with escapes: 5000 distinct values
    w = ((np.arange(10 * N) * 7) % 5000 - 2500).astype(np.int16).reshape(10, N)
    assert np.array_equal(_decode_by_definition(orc.rans_encode(w, 19441)), w)


# ---- the entropy coder (tests/wire.rs:179-440) -----------------------------------------------------

@pytest.mark.parametrize("base", MODULI)
def test_honest_fold_round_trips(orc, base):
    for wl, cl in [(11, 3), (12, 2)]:
        _, _, v, _ = _honest_round(orc, wl, cl, [base, second(base)])
        b = orc.rans_encode(v, base)
        err, back, ntt = orc.rans_decode(b)
        assert err == orc.WIRE_OK and not ntt and np.array_equal(back, v)
        assert len(b) < v.size * 2


def _adversarial_cases(q):
    half = (q - 1) // 2
    cases = [
        np.zeros((1, N), np.int16),
        np.full((1, N), half, np.int16),
        np.full((1, N), -half, np.int16),
        np.array([[half if i % 2 == 0 else -half for i in range(N)]], np.int16),
        np.stack([np.array([(i % (2 * half + 1)) - half for i in range(N)], np.int16), np.full(N, half, np.int16),
                  np.full(N, -half, np.int16)]),
    ]
    state = 0x2545F4914F6CDD1D
    m = (1 << 64) - 1
    rows = []
    for _ in range(16):
        row = []
        for _ in range(N):
            state ^= (state << 13) & m
            state ^= state >> 7
            state ^= (state << 17) & m
            row.append(state % (2 * half + 1) - half)
        rows.append(row)
    cases.append(np.array(rows, np.int16))
    return cases


def test_adversarial_fold_round_trips(orc):
    b = orc.rans_encode(np.zeros((0, N), np.int16), 3889)
    assert len(b) == 16
    err, back, _ = orc.rans_decode(b)
    assert err == orc.WIRE_OK and len(back) == 0
    for q in MODULI:
        for case, v in enumerate(_adversarial_cases(q)):
            err, back, _ = orc.rans_decode(orc.rans_encode(v, q))
            assert err == orc.WIRE_OK and np.array_equal(back, v), (q, case)


def _any_i16():
    state = 0x9E3779B97F4A7C15
    m = (1 << 64) - 1
    v = np.zeros((8, N), np.int16)
    for e in range(8):
        for i in range(N):
            state = (state * 6364136223846793005 + 1) & m
            x = (state >> 32) & 0xFFFF
            v[e, i] = x - 65536 if x >= 32768 else x
    v[0, 0], v[0, 1] = -32768, 32767
    return v


def test_coder_takes_any_i16(orc):
    v = _any_i16()
    err, back, _ = orc.rans_decode(orc.rans_encode(v, 3889))
    assert err == orc.WIRE_OK and np.array_equal(back, v)


def test_broken_stream_is_refused(orc):
    _, _, v, _ = _honest_round(orc, 11, 3, [3889, 9721])
    b = orc.rans_encode(v, 3889)
    for cut in [0, 1, 15, 16, 20, len(b) // 2, len(b) - 1]:
        assert orc.rans_decode(b[:cut])[0] != orc.WIRE_OK, cut
    assert orc.rans_decode(b + b"\0\0")[0] == orc.WIRE_MALFORMED
    assert orc.rans_decode(b"\x02" + b[1:])[0] == orc.WIRE_MALFORMED


def _lcg_message(count, state):
    m = (1 << 64) - 1
    v = np.zeros((count, N), np.int16)
    for e in range(count):
        for i in range(N):
            state = (state * 6364136223846793005 + 1442695040888963407) & m
            v[e, i] = sum((state >> (10 * j)) & 0x3F for j in range(6)) - 189
    return v, state


def test_interleaved_fold_round_trips_at_every_length(orc):
    state = 0x123456789ABCDEF1
    for count in [1, 2, 3, 5, 7, 8]:
        v, state = _lcg_message(count, state)
        b = orc.rans_encode(v, 3889)
        err, back, _ = orc.rans_decode(b)
        assert err == orc.WIRE_OK and np.array_equal(back, v), count
        assert len(b) <= orc.entropy_bytes(v) * 1.02 + 4.0 * 64 + 200.0


def _one_lane_message(k):
    v = np.zeros((1, N), np.int16)
    for i in list(range(k, N, 64))[:3]:
        v[0, i] = 1
    return v


@pytest.mark.parametrize("k", [0, 1, 17, 31, 32, 63])
def test_stream_short_or_long_by_one_word_on_any_lane_is_refused(orc, k):
    v = _one_lane_message(k)
    b = orc.rans_encode(v, 3889)
    assert len(b) == 16 + 5 + 4 * 64 + 2
    err, back, _ = orc.rans_decode(b)
    assert err == orc.WIRE_OK and np.array_equal(back, v)
    assert orc.rans_decode(b[:-2])[0] == orc.WIRE_TRUNCATED
    assert orc.rans_decode(b[:-3])[0] == orc.WIRE_TRUNCATED
    assert orc.rans_decode(b + b"\0\0")[0] == orc.WIRE_MALFORMED
    assert orc.rans_decode(b[:16 + 5 + 4 * k + 2])[0] == orc.WIRE_TRUNCATED
    low = bytearray(b)
    low[16 + 5 + 4 * k + 2] = 0
    low[16 + 5 + 4 * k + 3] = 0
    assert orc.rans_decode(bytes(low))[0] == orc.WIRE_MALFORMED


def test_every_lane_emits_its_word(orc):
    for k in range(64):
        v = _one_lane_message(k)
        b = orc.rans_encode(v, 3889)
        assert len(b) == 16 + 5 + 4 * 64 + 2, k
        assert np.array_equal(orc.rans_decode(b)[1], v)


def test_coded_size_approaches_the_entropy(orc):
    rng = _rng(17)
    v = np.rint(rng.normal(0, 56, (64, N))).astype(np.int16)
    b = orc.rans_encode(v, 3889)
    h = orc.entropy_bytes(v)
    assert h < len(b) < h * 1.01 + 4 * 64 + 600


# ---- the host side of the product against the oracle ----------------------------------------------

def _params(lbn, wl, cl, primes, opening=None, d=0):
    extra = [lbn.Modulus(q) for q in primes[1:]]
    op = lbn.Opening.CLEAR if not d else lbn.Opening.bit_dropped(d)
    return lbn.Params.with_base(wl, cl, lbn.Modulus(primes[0]), extra, op)


def test_host_unpack_commitment_matches_oracle(lbn, orc):
    from labinius_amd import wire
    primes = [3889, 9721]
    comp, _, _, cols = _honest_round(orc, 11, 3, primes)
    params = _params(lbn, 11, 3, primes)
    b = orc.pack_commitment(primes, cols, comp)
    assert wire.commitment_bytes(params) == len(b)
    assert np.array_equal(wire.unpack_commitment(params, b), comp)
    assert wire.pack_commitment_host(params, comp) == b
    widths = [orc.residue_bits(q) for q in primes]
    per = sum(162 * w for w in widths)
    for element, limb, slot in _out_of_range_cases(cols):
        bit = element * per + sum(162 * w for w in widths[:limb]) + slot * widths[limb]
        with pytest.raises(wire.WireError) as e:
            wire.unpack_commitment(params, _set_field(b, bit, widths[limb]))
        assert e.value == wire.Malformed
    with pytest.raises(wire.WireError) as e:
        wire.unpack_commitment(params, b[:-1])
    assert e.value == wire.Malformed


def test_host_unpack_dropped_matches_oracle(lbn, orc):
    from labinius_amd import wire
    primes, d = [3889, 9721], 3
    params = _params(lbn, 11, 3, primes, d=d)
    cols = params.columns()
    rng = _rng(9)
    bound = ((primes[0] - 1) + (1 << (d - 1))) >> d
    top = rng.integers(0, bound + 1, cols * N).astype(np.uint16)
    digits = rng.integers(0, primes[1], (1, cols * N)).astype(np.uint16)
    b = orc.pack_dropped(primes, cols, d, top, digits)
    assert wire.commitment_bytes(params) == len(b)
    got = wire.unpack_commitment(params, b)
    assert np.array_equal(got.top, top) and np.array_equal(got.digits, digits)
    with pytest.raises(wire.WireError):
        wire.unpack_commitment(params, b + b"\0")


def test_host_f162_matches_oracle(lbn, orc):
    from labinius_amd import wire
    rng = _rng(4)
    for count in [0, 1, 3, 256]:
        v = rng.integers(0, 1 << 62, (count, 3), dtype=np.uint64)
        v[:, 2] &= np.uint64((1 << 34) - 1)
        b = wire.pack_f162(v)
        assert b == orc.pack_f162(v)
        assert np.array_equal(wire.unpack_f162(b, count), v)
    with pytest.raises(wire.WireError) as e:
        wire.unpack_f162(b[:-1], 256)
    assert e.value == wire.Malformed


def test_host_decoder_matches_oracle(lbn, orc):
    from labinius_amd import wire
    msgs = [np.zeros((0, N), np.int16), _any_i16(), _lcg_message(3, 7)[0]]
    msgs += _adversarial_cases(19441)
    msgs.append(_honest_round(orc, 11, 3, [9721, 3889])[2])
    for v in msgs:
        b = orc.rans_encode(v, 3889)
        assert np.array_equal(wire.decode(b).elements, v)


def test_host_decoder_refuses_as_the_reference(lbn, orc):
    from labinius_amd import wire
    _, _, v, _ = _honest_round(orc, 11, 3, [3889, 9721])
    b = orc.rans_encode(v, 3889)
    cases = [b[:cut] for cut in [0, 1, 15, 16, 20, len(b) // 2, len(b) - 1]]
    cases += [b + b"\0\0", b"\x02" + b[1:]]
    for k in [0, 33, 63]:
        one = orc.rans_encode(_one_lane_message(k), 3889)
        low = bytearray(one)
        low[16 + 5 + 4 * k + 2:16 + 5 + 4 * k + 4] = b"\0\0"
        cases += [one[:-2], one[:-3], one + b"\0\0", one[:16 + 5 + 4 * k + 2], bytes(low)]
    for c in cases:
        err = orc.rans_decode(c)[0]
        assert err != orc.WIRE_OK
        with pytest.raises(wire.WireError) as e:
            wire.decode(c)
        assert e.value.kind == {orc.WIRE_TRUNCATED: "Truncated", orc.WIRE_MALFORMED: "Malformed"}[err]


def test_f64_reciprocal_quotient_is_exact():
    # The device encoder divides by a frequency f as floor(x * rcp) in f64, rcp = 1/f rounded up
    # (exact for a power of two): exact for every state x < f 2^20 it divides. Checked at the
    # multiples of f (where a short product would floor one low), one below and f - 1 above them,
    # for every f and a spread of quotients up to 2^20 - 1.
    f = np.arange(1, 4097, dtype=np.float64)[:, None]
    fi = f.astype(np.int64)
    pow2 = (fi & (fi - 1)) == 0
    rcp = np.where(pow2, 1.0 / f, np.nextafter(1.0 / f, 2.0))
    rng = _rng(29)
    n = np.concatenate([np.arange(1, 64), rng.integers(64, 1 << 20, 400), [(1 << 20) - 1]])[None, :]
    for dx in (-1, 0):
        x = n * fi + dx
        q = np.floor(x.astype(np.float64) * rcp).astype(np.int64)
        assert np.array_equal(q, x // fi)
    x = n * fi + (fi - 1)
    ok = x < fi * (1 << 20)
    q = np.floor(x.astype(np.float64) * rcp).astype(np.int64)
    assert np.array_equal(q[ok], (x // fi)[ok])
    # The encoder also divides a state it will shed (x > f 2^20 - 1, any x < 2^32) and uses the
    # quotient's top: x / f >> 16 = (x >> 16) / f. Both checked up to 2^32 - 1.
    top = (1 << 32) - 1
    n = np.concatenate([rng.integers(1 << 20, 1 << 32, 300), [top]])[None, :] // fi
    for dx in (-1, 0, fi - 1):
        x = np.minimum(n * fi + dx, top)
        q = np.floor(x.astype(np.float64) * rcp).astype(np.int64)
        assert np.array_equal(q, x // fi)
        assert np.array_equal(q >> 16, (x >> 16) // fi)

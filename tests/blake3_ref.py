"""A pure-Python BLAKE3 (hash mode) from the specification, in the recursive tree form: a node
over n > 1 chunks splits into a left subtree of the largest power of two of chunks below n and
the rest. Independent of both C implementations' incremental stack; used to check their tree
handling across chunk boundaries. Slow: small inputs only."""
IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19]
PERM = [2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8]
CHUNK_START, CHUNK_END, PARENT, ROOT = 1, 2, 4, 8
M32 = 0xFFFFFFFF


def _rotr(x, n):
    return ((x >> n) | (x << (32 - n))) & M32


def _g(v, a, b, c, d, x, y):
    v[a] = (v[a] + v[b] + x) & M32
    v[d] = _rotr(v[d] ^ v[a], 16)
    v[c] = (v[c] + v[d]) & M32
    v[b] = _rotr(v[b] ^ v[c], 12)
    v[a] = (v[a] + v[b] + y) & M32
    v[d] = _rotr(v[d] ^ v[a], 8)
    v[c] = (v[c] + v[d]) & M32
    v[b] = _rotr(v[b] ^ v[c], 7)


def _compress(cv, block, counter, blen, flags):
    m = [int.from_bytes(block[4 * i:4 * i + 4], "little") for i in range(16)]
    v = list(cv) + IV[:4] + [counter & M32, counter >> 32, blen, flags]
    for r in range(7):
        _g(v, 0, 4, 8, 12, m[0], m[1])
        _g(v, 1, 5, 9, 13, m[2], m[3])
        _g(v, 2, 6, 10, 14, m[4], m[5])
        _g(v, 3, 7, 11, 15, m[6], m[7])
        _g(v, 0, 5, 10, 15, m[8], m[9])
        _g(v, 1, 6, 11, 12, m[10], m[11])
        _g(v, 2, 7, 8, 13, m[12], m[13])
        _g(v, 3, 4, 9, 14, m[14], m[15])
        m = [m[PERM[i]] for i in range(16)]
    return [v[i] ^ v[i + 8] for i in range(8)] + [v[i + 8] ^ cv[i] for i in range(8)]


def _chunk_node(chunk, index):
    """(cv, block, counter, len, flags) of the chunk's last block, before its final compression."""
    cv = list(IV)
    blocks = [chunk[i:i + 64] for i in range(0, len(chunk), 64)] or [b""]
    for bi, blk in enumerate(blocks[:-1]):
        cv = _compress(cv, blk, index, 64, CHUNK_START if bi == 0 else 0)[:8]
    last = blocks[-1]
    flags = (CHUNK_START if len(blocks) == 1 else 0) | CHUNK_END
    return cv, last.ljust(64, b"\0"), index, len(last), flags


def _node(data, first_chunk):
    nchunks = max(1, (len(data) + 1023) // 1024)
    if nchunks == 1:
        return _chunk_node(data, first_chunk)
    left = 1
    while left * 2 < nchunks:
        left *= 2
    lcv = _cv(_node(data[:1024 * left], first_chunk))
    rcv = _cv(_node(data[1024 * left:], first_chunk + left))
    block = b"".join(x.to_bytes(4, "little") for x in lcv + rcv)
    return IV, block, 0, 64, PARENT


def _cv(node):
    cv, block, counter, blen, flags = node
    return _compress(cv, block, counter, blen, flags)[:8]


def blake3_hash(data, out_len=32):
    cv, block, _, blen, flags = _node(bytes(data), 0)
    out = b""
    i = 0
    while len(out) < out_len:
        w = _compress(cv, block, i, blen, flags | ROOT)
        out += b"".join(x.to_bytes(4, "little") for x in w)
        i += 1
    return out[:out_len]

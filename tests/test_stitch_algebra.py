"""Window starts stitched from per-segment hash pieces (walk_tiles in
gt_kernels.hip) equal CyclicHash rolled over the K bytes
(cyclichash.h:85-108: fw = XOR_k rotl(T[b_k], K-1-k); the reverse-complement
hash rc = XOR_k rotl(T[comp(b_k)], k)), for every K the library takes and the
segments the walker stitches from (4 bytes, as the first binned pass uses
them, and 8 bytes, which the walker also supports), with the rotation amounts
reduced as the kernel reduces them (an unsigned 32-bit value masked to 6
bits)."""
import random

import pytest

M64 = (1 << 64) - 1


def rotl(x, r):
    r &= 63
    return ((x << r) | (x >> (64 - r))) & M64 if r else x


def seg_values(by, u, L, T, TR, PER):
    pf = pr = qf = qr = 0
    for i in range(PER):
        c = by[u * PER + i]
        pf = rotl(pf, 1) ^ T[c]
        pr ^= rotl(TR[c], i)
        if i == L - 1:
            qf, qr = pf, pr
    return pf, pr, qf, qr


CASES = ([(8, K) for K in list(range(1, 34)) + [63, 64, 65, 127, 128, 200, 254, 255]]
         + [(4, K) for K in list(range(1, 34)) + [63, 64, 65, 127, 128, 131, 132, 133, 200, 254, 255]])


@pytest.mark.parametrize('PER,K', CASES)
def test_stitched_start_equals_rolled(PER, K):
    rng = random.Random(K)
    T = [rng.getrandbits(64) for _ in range(256)]
    TR = [rng.getrandbits(64) for _ in range(256)]
    n_seg = (K + PER - 1) // PER
    L = K - (n_seg - 1) * PER
    n_threads = 40
    by = [rng.randrange(256) for _ in range(PER * (n_threads + n_seg))]
    assert n_seg - 1 <= 64          # the kernel's segment scratch holds 64 extra segments
    S = [seg_values(by, u, L, T, TR, PER) for u in range(n_threads + n_seg - 1)]
    for t in range(n_threads):
        li = t * PER
        fw = rc = 0
        for k in range(K):
            fw = rotl(fw, 1) ^ T[by[li + k]]
        for k in range(K - 1, -1, -1):
            rc = rotl(rc, 1) ^ TR[by[li + k]]
        f2 = r2 = 0
        for j in range(n_seg - 1):
            f2 ^= rotl(S[t + j][0], (K - (j + 1) * PER) & 0xFFFFFFFF)
            r2 ^= rotl(S[t + j][1], j * PER)
        f2 ^= S[t + n_seg - 1][2]
        r2 ^= rotl(S[t + n_seg - 1][3], (n_seg - 1) * PER)
        assert (f2, r2) == (fw, rc)

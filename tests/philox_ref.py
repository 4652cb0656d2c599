"""Pure-Python Philox (Salmon et al., SC'11) for any word size, used to
cross-check the oracle's Philox4x32-10: the same code reproduces numpy's
Philox4x64-10 bit generator and the published 4x32 known-answer vectors."""

CONST = {
    32: dict(M=(0xD2511F53, 0xCD9E8D57), W=(0x9E3779B9, 0xBB67AE85)),
    64: dict(M=(0xD2E7470EE14C6C93, 0xCA5A826395121157), W=(0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B)),
}


def philox4(ctr, key, bits=32, rounds=10):
    mask = (1 << bits) - 1
    M0, M1 = CONST[bits]["M"]
    W0, W1 = CONST[bits]["W"]
    c = [int(x) & mask for x in ctr]
    k0, k1 = int(key[0]) & mask, int(key[1]) & mask
    for r in range(rounds):
        if r:
            k0 = (k0 + W0) & mask
            k1 = (k1 + W1) & mask
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        c = [((p1 >> bits) ^ c[1] ^ k0) & mask, p1 & mask, ((p0 >> bits) ^ c[3] ^ k1) & mask, p0 & mask]
    return c

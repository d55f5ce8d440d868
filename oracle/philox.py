"""Counter-based Philox4x32-10 streams, vectorised in numpy (oracle side).

TEST INFRASTRUCTURE ONLY: imported by tests/, bench.py's CPU-baseline leg and
``__graft_entry__.smoke()`` as the checker. Product code never imports oracle/.

Philox4x32-10 is the counter-based generator of Salmon, Moraes, Dror and Shaw,
"Parallel random numbers: as easy as 1, 2, 3" (SC'11): ten rounds of two
32x32->64 multiplies with the published multipliers M0/M1 and Weyl key
increments W0/W1. Every draw of the simulation is a pure function of
(seed, counter), so the GPU kernels and this restatement produce the same
numbers regardless of thread or launch order.

Stream layout (counter word c2 selects the purpose; the same table is used by
the HIP kernels in covasim_amd/csrc/cvk_rng.h):

=========================  ==============================  =====================
purpose                    counter (c0, c1, c2, c3)        words used
=========================  ==============================  =====================
transmission, edge e       (e_lo, t, TX | v<<8 | l, e_hi)   (w0,w1)->dir 0, (w2,w3)->dir 1
infection outcome slot s   (agent, t, INFECT, s)           (w0,w1)->u0, (w2,w3)->u1
seed-infection choice      (agent, 0, CHOOSE, 0)           w0 = selection key
variant import choice      (agent, t, IMPORT | v, 0)       w0 = selection key
dynamic-layer resample     (q_lo, t, RESAMPLE | l, q_hi)   edge 2q: w0->p1, w1->p2;
  (pair q = e >> 1)                                         edge 2q+1: w2->p1, w3->p2
resample retry k of edge e (e_lo, t, RETRY | l, e_hi<<8|k)  w0->p1, w1->p2
test of agent i (op s)     (i, t, TEST | s, 0)             (w0,w1)->tested, (w2,w3)->positive
                           (i, t, TEST | s, 1)             (w0,w1)->not lost
trace of contact i, layer  (i, t, TRACE | s<<8 | l, 0)     (w0,w1)->traced
=========================  ==============================  =====================

Resampled endpoints are unbiased bounded integers (Lemire's multiply-and-reject):
m = w * n, accepted when (m mod 2^32) >= (2^32 - n) mod n, value m >> 32; a rejected
endpoint takes the same-position word of retry k = 0, 1, ... (at most 255 retries; the
256th word is accepted as is, which happens with probability below 10^-600).
"""

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

STREAM_TX = 0x10000000
STREAM_INFECT = 0x20000000
STREAM_CHOOSE = 0x30000000
STREAM_IMPORT = 0x40000000
STREAM_RESAMPLE = 0x50000000
STREAM_RESAMPLE_RETRY = 0x60000000
RESAMPLE_MAX_RETRY = 255
STREAM_TEST = 0x90000000   # | slot
STREAM_PARTIAL = 0xB0000000  # | layer: edge choice of a partial resample
STREAM_TRACE = 0xA0000000  # | slot << 8 | layer
STREAM_ILI = 0xC0000000    # | slot: the agents with influenza-like symptoms (test_prob ili_prev)
STREAM_TRACE_CAP = 0xD0000000  # | slot: the cases traced under a capacity (contact_tracing)

# Fixed per-infection draw slots (see People.infect restatement in covasim_oracle.py)
SLOT_EXP2INF = 0
SLOT_BERN_SYMP_SEV = 1   # u0: symptomatic, u1: severe
SLOT_BERN_CRIT_DEATH = 2  # u0: critical,    u1: death
SLOT_DUR_A = 3  # asym2rec | inf2sym
SLOT_DUR_B = 4  # mild2rec | sym2sev
SLOT_DUR_C = 5  # sev2rec  | sev2crit
SLOT_DUR_D = 6  # crit2rec | crit2die
SLOT_NAB = 7


def _u32(x):
    return np.asarray(x, dtype=np.uint64) & MASK32


def philox4x32(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 block function on broadcastable uint32-valued arrays.

    Returns four uint64 arrays holding the 32-bit output words.
    """
    c0, c1, c2, c3 = np.broadcast_arrays(_u32(c0), _u32(c1), _u32(c2), _u32(c3))
    c0 = c0.copy(); c1 = c1.copy(); c2 = c2.copy(); c3 = c3.copy()
    k0 = np.uint64(int(k0) & 0xFFFFFFFF)
    k1 = np.uint64(int(k1) & 0xFFFFFFFF)
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0 = p0 >> np.uint64(32)
        lo0 = p0 & MASK32
        hi1 = p1 >> np.uint64(32)
        lo1 = p1 & MASK32
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0), lo1, (hi0 ^ c3 ^ k1), lo0
        k0 = (k0 + np.uint64(W0)) & MASK32
        k1 = (k1 + np.uint64(W1)) & MASK32
    return c0, c1, c2, c3


def seed_key(seed):
    """Split a 64-bit sim seed into the Philox key pair."""
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return s & 0xFFFFFFFF, s >> 32


def u01(a, b):
    """53-bit double in [0, 1) from two 32-bit words: ((a>>5)*2^26 + (b>>6)) / 2^53."""
    a = np.asarray(a, dtype=np.uint64) >> np.uint64(5)
    b = np.asarray(b, dtype=np.uint64) >> np.uint64(6)
    return (a.astype(np.float64) * 67108864.0 + b.astype(np.float64)) / 9007199254740992.0


def below(a, b, n):
    """Uniform integer in [0, n) from two words: floor(((a<<32)|b) * n / 2^64), n < 2^32."""
    n = np.uint64(n)
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    lo = (b * n) >> np.uint64(32)
    return ((a * n + lo) >> np.uint64(32)).astype(np.int64)


def edge_uniforms(seed, t, v, l, edges, direction):
    """Per-edge transmission uniforms for (variant v, layer l, day t)."""
    k0, k1 = seed_key(seed)
    e = np.asarray(edges, dtype=np.int64).astype(np.uint64)
    w = philox4x32(e & MASK32, t, STREAM_TX | (v << 8) | l, e >> np.uint64(32), k0, k1)
    if direction == 0:
        return u01(w[0], w[1])
    return u01(w[2], w[3])


def slot_uniforms(seed, t, agents, slot):
    """(u0, u1) for an infection-outcome slot of each agent on day t."""
    k0, k1 = seed_key(seed)
    w = philox4x32(np.asarray(agents, dtype=np.int64), t, STREAM_INFECT, slot, k0, k1)
    return u01(w[0], w[1]), u01(w[2], w[3])


def stream_uniforms(seed, t, agents, stream, c3=0):
    """(u0, u1) of the draw (agent, t, stream, c3) of each agent: words 0,1 and 2,3."""
    k0, k1 = seed_key(seed)
    w = philox4x32(np.asarray(agents, dtype=np.int64), t, stream, c3, k0, k1)
    return u01(w[0], w[1]), u01(w[2], w[3])


def selection_keys(seed, t, stream, agents):
    """64-bit selection keys (philox word << 32 | agent): distinct, so the k smallest form a uniform k-subset."""
    k0, k1 = seed_key(seed)
    agents = np.asarray(agents, dtype=np.int64)
    w = philox4x32(agents, t, stream, 0, k0, k1)
    return (w[0] << np.uint64(32)) | agents.astype(np.uint64)


def choose_k(seed, t, stream, eligible, k):
    """k distinct agents, uniform over eligible (ascending index order returned)."""
    eligible = np.asarray(eligible, dtype=np.int64)
    if k >= len(eligible):
        return np.sort(eligible)
    keys = selection_keys(seed, t, stream, eligible)
    sel = np.argpartition(keys, k - 1)[:k] if k > 0 else np.empty(0, dtype=np.int64)
    return np.sort(eligible[sel])


def bounded(w, n):
    """(value, accepted) of Lemire's multiply-and-reject for 32-bit words w and 0 < n < 2^32."""
    n = np.uint64(n)
    m = np.asarray(w, dtype=np.uint64) * n
    thr = (np.uint64(1 << 32) - n) % n
    return (m >> np.uint64(32)).astype(np.int64), (m & MASK32) >= thr


def resample_edges(seed, t, l, n_edges, n_agents):
    """Dynamic-layer resample: iid uniform endpoints for every edge (Layer.update with frac=1)."""
    k0, k1 = seed_key(seed)
    q = np.arange((n_edges + 1) // 2, dtype=np.uint64)
    w = philox4x32(q & MASK32, t, STREAM_RESAMPLE | l, q >> np.uint64(32), k0, k1)
    words1 = np.empty(2 * len(q), dtype=np.uint64)
    words2 = np.empty(2 * len(q), dtype=np.uint64)
    words1[0::2], words2[0::2] = w[0], w[1]
    words1[1::2], words2[1::2] = w[2], w[3]
    out = []
    for j, words in enumerate((words1[:n_edges], words2[:n_edges])):
        val, ok = bounded(words, n_agents)
        for k in range(RESAMPLE_MAX_RETRY):
            bad = np.flatnonzero(~ok)
            if len(bad) == 0:
                break
            e = bad.astype(np.uint64)
            wr = philox4x32(e & MASK32, t, STREAM_RESAMPLE_RETRY | l, ((e >> np.uint64(32)) << np.uint64(8)) | np.uint64(k),
                            k0, k1)
            v2, ok2 = bounded(wr[j], n_agents)
            val[bad] = v2
            last = k == RESAMPLE_MAX_RETRY - 1
            ok[bad] = ok2 | last
        out.append(val.astype(np.int32))
    return out[0], out[1]


def partial_resample(seed, t, l, p1, p2, beta, n_agents, frac):
    """Layer.update(people, frac) of base.py:1866-1875: k = round(n * frac) edges (np.round, half to
    even), a uniform k-subset, get new uniform endpoints and weight 1; the others keep theirs.
    The subset is the k smallest keys (Philox word 0 of (edge, t, STREAM_PARTIAL | l) << 32 | edge);
    a chosen edge's endpoints are those the full resample of day t gives it. Returns new arrays."""
    n = len(p1)
    k = int(np.round(n * frac))
    p1, p2, beta = np.array(p1, dtype=np.int32), np.array(p2, dtype=np.int32), np.array(beta, dtype=np.float32)
    if k == 0:
        return p1, p2, beta
    k0, k1 = seed_key(seed)
    e = np.arange(n, dtype=np.uint64)
    w = philox4x32(e & MASK32, t, STREAM_PARTIAL | l, e >> np.uint64(32), k0, k1)
    keys = (w[0] << np.uint64(32)) | (e & MASK32)
    sel = np.argpartition(keys, k - 1)[:k] if k < n else np.arange(n)
    q1, q2 = resample_edges(seed, t, l, n, n_agents)
    p1[sel], p2[sel], beta[sel] = q1[sel], q2[sel], 1.0
    return p1, p2, beta


def box_muller(u0, u1):
    """Standard normal from two uniforms: sqrt(-2 ln(1-u0)) * cos(2 pi u1)."""
    return np.sqrt(-2.0 * np.log1p(-u0)) * np.cos(2.0 * np.pi * u1)

"""The method the device octree build uses to fold a long quantize_error or
total_color chain in parallel, restated in Python and checked against the
serial left-to-right double sum the reference performs (src/rd_octree.cpp:
305-340 adds each run's term to its node in run order).

All terms are finite and non-negative, so the running value only grows.  On
one binade's grid every addition is an integer addition of the term rounded on
its own; the exceptions -- a tie, and the step that leaves the binade -- are
taken as plain IEEE additions.  The restatement keeps the device's structure
(serial prologue, windows of T x K terms, a clean test on the window total,
rounds that resolve one special step per chain, a serial remainder after too
many rounds) with small T and K so that every branch is exercised.
"""
import math
import random

import pytest

SAT = 2.0 ** 54
MAXD = 1.7976931348623157e308


def _ldx(v, k):
    try:
        return math.ldexp(v, k)
    except OverflowError:
        return math.inf


def _rint(v):
    return float(round(v)) if math.isfinite(v) else v


def _fmin(a, b):
    if a != a:
        return b
    if b != b:
        return a
    return min(a, b)


def _grid(s):
    return -1022 if s < 2.0 ** -1022 else math.frexp(s)[1] - 1


def exact_sums(s, chains, T, K, PRO, ROUNDS):
    """Restates exact_sums (stipple_amd/csrc/sa_octree.hip) for NC chains."""
    NC, e, win = len(chains), len(chains[0]), T * K
    s = list(s)
    pro = min(e, PRO)
    for c in range(NC):
        for i in range(pro):
            s[c] += chains[c][i]
    pos = [pro] * NC
    while True:
        for c in range(NC):
            if not s[c] <= MAXD:
                pos[c] = e
        W = min(pos)
        if W >= e:
            return s
        wend = min(e, W + win)
        x = [_grid(v) for v in s]
        S = [_ldx(s[c], 52 - x[c]) for c in range(NC)]
        clean, total = True, [0.0] * NC
        for c in range(NC):
            tot, tie = 0.0, False
            for i in range(max(W, pos[c]), wend):
                q = _ldx(chains[c][i], 52 - x[c])
                r = _rint(q)
                tie |= math.isfinite(q) and abs(q - r) == 0.5
                tot += r
            total[c] = _fmin(tot, SAT)
            if tie or not S[c] + tot < 2.0 ** 53 - 1:
                clean = False
        if clean:
            for c in range(NC):
                if pos[c] < wend:
                    s[c], pos[c] = _ldx(S[c] + total[c], x[c] - 52), wend
            continue
        for rnd in range(ROUNDS + 1):
            if all(p >= wend for p in pos):
                break
            for c in range(NC):
                if pos[c] >= wend:
                    continue
                if rnd == ROUNDS:  # dense special steps: the rest serially
                    for i in range(pos[c], wend):
                        s[c] += chains[c][i]
                    pos[c] = wend
                    continue
                x = _grid(s[c])
                Sc, R, tot, first = _ldx(s[c], 52 - x), None, 0.0, None
                R = Sc
                for i in range(pos[c], wend):
                    tot = _fmin(tot + _rint(_fmin(_ldx(chains[c][i], 52 - x), SAT)), SAT)
                for i in range(pos[c], wend):
                    q = _ldx(chains[c][i], 52 - x)
                    if (math.isfinite(q) and q - math.floor(q) == 0.5) or not q < 2.0 ** 53 - R:
                        first = i
                        break
                    R += _rint(_fmin(q, SAT))
                if first is None:
                    s[c], pos[c] = _ldx(_fmin(Sc + tot, SAT), x - 52), wend
                else:
                    s[c] = _ldx(R, x - 52) + chains[c][first]
                    pos[c] = first + 1
                    if not s[c] <= MAXD:
                        pos[c] = max(pos[c], wend)


def _serial(s, terms):
    for t in terms:
        s += t
    return s


def _chain(kind, n, rng):
    if kind == "rand":
        return 0.0, [rng.random() * 100 for _ in range(n)]
    if kind == "ints":
        return 0.0, [float(rng.randint(0, 9)) for _ in range(n)]
    if kind == "zeros":
        return 0.0, [0.0] * (n // 2) + [1.5] * (n - n // 2)
    if kind == "ties":  # every term a tie on the grid of 2^52
        return 2.0 ** 52, [0.5 * (2 * rng.randint(0, 3) + 1) for _ in range(n)]
    if kind == "halfulp":
        return 1.0, [2.0 ** -53] * n
    if kind == "subnormal":
        return 0.0, [5e-324 * rng.randint(0, 7) for _ in range(n)]
    if kind == "overflow":
        return 0.0, ([rng.random() for _ in range(n // 2)] + [1e308] * 3 + [1.0] * n)[:n]
    if kind == "jumps":
        return 7.0, [10.0 ** rng.randint(-20, 20) * rng.random() for _ in range(n)]
    if kind == "boundary":  # walks onto 2^53 and stays (ties after it)
        return 2.0 ** 53 - 64.0, [1.0] * n
    if kind == "sparse_ties":
        return 2.0 ** 40, [0.5 if rng.random() < 0.02 else rng.random() * 3 for _ in range(n)]
    raise ValueError(kind)


KINDS = ["rand", "ints", "zeros", "ties", "halfulp", "subnormal", "overflow", "jumps", "boundary",
         "sparse_ties"]


@pytest.mark.parametrize("cfg", [(4, 3, 0, 2), (8, 4, 5, 3), (2, 2, 0, 1), (16, 2, 9, 8)])
@pytest.mark.parametrize("n", [1, 37, 300])
def test_exact_sums_equal_serial_sums(cfg, n):
    rng = random.Random(hash((cfg, n)) & 0xFFFF)
    for kind in KINDS:
        kinds = [kind] + [rng.choice(KINDS) for _ in range(3)]
        gens = [_chain(k, n, rng) for k in kinds]
        for nc in (1, 4):
            got = exact_sums([g[0] for g in gens[:nc]], [g[1] for g in gens[:nc]], *cfg)
            for c in range(nc):
                exp = _serial(*gens[c])
                assert got[c] == exp or (got[c] != got[c] and exp != exp), (kinds[c], cfg, n)


def test_a_dropped_special_step_is_caught():
    """Planting a mistake (treating a tie as an ordinary step) breaks equality."""
    s0, terms = 2.0 ** 52 + 1.0, [0.5, 0.5, 0.5]  # the first tie rounds up (to even)
    naive = s0 + sum(_rint(t) for t in terms)  # rounds every tie on its own
    assert naive != _serial(s0, terms)
    assert exact_sums([s0], [terms], 2, 2, 0, 2)[0] == _serial(s0, terms)

"""Generate the min/max selection networks used by the HIP kernels.

Writes ``networks.inc`` next to this file. Each network is a straight-line
sequence of fmin/fmax on doubles built from small optimal sorting networks and
Batcher odd-even merges, then simplified:
  * rank pruning: in a sorted sub-list of m values taken from a window of N,
    entry i can only be the median if it has at most M=(N-1)/2 values below and
    above it, so the others are replaced by -inf/+inf (the window's M-th order
    statistic is unchanged, and lists stay sorted);
  * constant folding of +-inf, hash-consing of identical nodes, absorption
    (min(x, max(x, y)) = x) and dead-node elimination.
Strip networks compute T vertically adjacent k x k medians and share the merged
rows common to neighbouring outputs (the separable scheme of A. Adams, "Fast
Median Filters Using Separable Sorting Networks", ACM TOG 40(4), 2021).

Every network is checked here against numpy on random inputs with ties before
it is written. Run: python cusmic_amd/csrc/gen_networks.py
"""

from pathlib import Path

import numpy as np

INF, NINF = "INF", "NINF"

# Optimal sorting networks for n <= 7 (Knuth TAOCP 5.3.4).
SORT = {
    1: [],
    2: [(0, 1)],
    3: [(0, 2), (0, 1), (1, 2)],
    4: [(0, 1), (2, 3), (0, 2), (1, 3), (1, 2)],
    5: [(0, 3), (1, 4), (0, 2), (1, 3), (0, 1), (2, 4), (1, 2), (3, 4), (2, 3)],
    6: [(0, 5), (1, 3), (2, 4), (1, 2), (3, 4), (0, 3), (2, 5), (0, 1), (2, 3), (4, 5), (1, 2), (3, 4)],
    7: [(0, 6), (2, 3), (4, 5), (0, 2), (1, 4), (3, 6), (0, 1), (2, 5), (3, 4), (1, 2), (4, 6),
        (2, 3), (4, 5), (1, 2), (3, 4), (5, 6)],
}


class Net:
    """A hash-consed DAG of min/max nodes over named inputs."""

    def __init__(self):
        self.key = {}
        self.defs = []

    def leaf(self, name):
        k = ("in", name)
        if k not in self.key:
            self.key[k] = len(self.defs)
            self.defs.append(k)
        return self.key[k]

    def op(self, op, a, b):
        if a == b:
            return a
        lo_const, hi_const = (INF, NINF) if op == "min" else (NINF, INF)
        # min(x, INF) = x; min(x, NINF) = NINF; max symmetric
        if a == lo_const:
            return b
        if b == lo_const:
            return a
        if a == hi_const or b == hi_const:
            return hi_const
        if a > b:
            a, b = b, a
        for x, y in ((a, b), (b, a)):
            d = self.defs[y]
            if d[0] not in ("in", op) and x in d[1:]:
                return x
        k = (op, a, b)
        if k not in self.key:
            self.key[k] = len(self.defs)
            self.defs.append(k)
        return self.key[k]

    def ce(self, a, b):
        return self.op("min", a, b), self.op("max", a, b)

    def live(self, outs):
        seen, stack = set(), [o for o in outs if isinstance(o, int)]
        while stack:
            n = stack.pop()
            if n in seen:
                continue
            seen.add(n)
            d = self.defs[n]
            if d[0] != "in":
                stack += [c for c in d[1:] if isinstance(c, int)]
        return sorted(seen)

    def count(self, outs):
        return sum(1 for n in self.live(outs) if self.defs[n][0] != "in")

    def evaluate(self, outs, values):
        """values: dict input name -> np.ndarray (batched)."""
        val = {}
        for n in self.live(outs):
            d = self.defs[n]
            if d[0] == "in":
                val[n] = values[d[1]]
            else:
                f = np.minimum if d[0] == "min" else np.maximum
                val[n] = f(val[d[1]], val[d[2]])

        def get(o):
            if o == INF:
                return np.inf
            if o == NINF:
                return -np.inf
            return val[o]
        return [get(o) for o in outs]


def sortnet(net, xs):
    xs = list(xs)
    if len(xs) in SORT:
        for i, j in SORT[len(xs)]:
            xs[i], xs[j] = net.ce(xs[i], xs[j])
        return xs
    h = len(xs) // 2
    return merge(net, sortnet(net, xs[:h]), sortnet(net, xs[h:]))


def merge(net, a, b):
    """Batcher odd-even merge of two sorted lists of any lengths."""
    if not a:
        return list(b)
    if not b:
        return list(a)
    if len(a) == 1 and len(b) == 1:
        return list(net.ce(a[0], b[0]))
    even = merge(net, a[0::2], b[0::2])
    odd = merge(net, a[1::2], b[1::2])
    out = [even[0]]
    i = 0
    while i < len(odd) and i + 1 < len(even):
        out += list(net.ce(odd[i], even[i + 1]))
        i += 1
    return out + odd[i:] + even[i + 1:]


def prune(lst, n):
    mid = (n - 1) // 2
    m = len(lst)
    return [INF if i > mid else NINF if m - 1 - i > mid else v for i, v in enumerate(lst)]


def strip_median(k, t):
    """t vertically adjacent k x k medians; inputs r{row}_{col}, rows 0..t+k-2."""
    net = Net()
    n = k * k
    rows = [sortnet(net, [net.leaf(f"r{j}_{c}") for c in range(k)]) for j in range(t + k - 1)]
    memo = {}

    def rows_merged(rs):
        if rs not in memo:
            if len(rs) == 1:
                memo[rs] = prune(rows[rs[0]], n)
            else:
                h = len(rs) // 2
                memo[rs] = prune(merge(net, rows_merged(rs[:h]), rows_merged(rs[h:])), n)
        return memo[rs]

    outs = [None] * t

    def solve(a, b, lst, have):
        add = sorted(set(range(b - 1, a + k)) - have)
        groups = []
        for r in add:
            if groups and groups[-1][-1] == r - 1:
                groups[-1].append(r)
            else:
                groups.append([r])
        for g in groups:
            sub = rows_merged(tuple(g))
            lst = sub if lst is None else prune(merge(net, lst, sub), n)
        have = have | set(add)
        if b - a == 1:
            outs[a] = lst[(n - 1) // 2]
            return
        h = (a + b) // 2
        solve(a, h, lst, have)
        solve(h, b, lst, have)

    solve(0, t, None, frozenset())
    inputs = [f"r{j}_{c}" for j in range(t + k - 1) for c in range(k)]
    return net, outs, inputs


def window_median(n):
    """Median of n inputs v0..v{n-1} (one window)."""
    net = Net()
    xs = [net.leaf(f"v{i}") for i in range(n)]
    groups = []
    step = 7 if n % 7 == 0 else 5 if n % 5 == 0 else 3
    for i in range(0, n, step):
        groups.append(prune(sortnet(net, xs[i:i + step]), n))
    while len(groups) > 1:
        nxt = []
        for i in range(0, len(groups) - 1, 2):
            nxt.append(prune(merge(net, groups[i], groups[i + 1]), n))
        if len(groups) % 2:
            nxt.append(groups[-1])
        groups = nxt
    return net, [groups[0][(n - 1) // 2]], [f"v{i}" for i in range(n)]


def lower_half(n, keep):
    """Sorted positions 0..keep-1 of n inputs (the replacement selector)."""
    net = Net()
    xs = [net.leaf(f"v{i}") for i in range(n)]
    s = sortnet(net, xs)
    return net, s[:keep], [f"v{i}" for i in range(n)]


def median49_steps():
    """Median of 7 rows of 7 fed one row at a time: step k sorts row k and merges
    it into the rank-pruned list of the rows before; only the ranks that can
    still be the median of 49 are kept (at most 22 values live)."""
    steps = []

    def valid(m):
        return max(0, m - 25), min(m - 1, 24)

    for k in range(7):
        net = Net()
        row = sortnet(net, [net.leaf(f"r{i}") for i in range(7)])
        if k == 0:
            out = prune(row, 49)
            inputs = [f"r{i}" for i in range(7)]
        else:
            lo, hi = valid(7 * k)
            acc = [NINF if i < lo else INF if i > hi else net.leaf(f"a{i - lo}") for i in range(7 * k)]
            out = prune(merge(net, acc, row), 49)
            inputs = [f"a{i}" for i in range(hi - lo + 1)] + [f"r{i}" for i in range(7)]
        lo, hi = valid(7 * (k + 1))
        steps.append((net, out[lo:hi + 1], inputs))
    return steps


def check_median49_steps(steps, trials=20000, seed=1):
    rng = np.random.default_rng(seed)
    for kind in range(3):
        if kind == 0:
            x = rng.normal(size=(49, trials))
        else:
            x = rng.integers(0, 3 if kind == 1 else 2, size=(49, trials)).astype(float)
        acc = []
        for k, (net, outs, inputs) in enumerate(steps):
            vals = {f"r{i}": x[7 * k + i] for i in range(7)}
            vals.update({f"a{i}": a for i, a in enumerate(acc)})
            acc = net.evaluate(outs, vals)
        assert np.array_equal(np.array(acc[0]), np.sort(x, axis=0)[24]), "median49 steps"


def emit_step(name, net, outs, inputs):
    """Device function of one incremental step: acc (previous valid ranks) and
    row (7 values) in, new valid ranks out."""
    names = {}
    for n in inputs:
        kind, idx = n[0], int(n[1:])
        if ("in", n) in net.key:
            names[net.key[("in", n)]] = f"{'acc' if kind == 'a' else 'row'}[{idx}]"
    body, tmp = [], 0
    for node in net.live(outs):
        d = net.defs[node]
        if d[0] == "in":
            continue
        a = "INFINITY" if d[1] == INF else "-INFINITY" if d[1] == NINF else names[d[1]]
        b = "INFINITY" if d[2] == INF else "-INFINITY" if d[2] == NINF else names[d[2]]
        names[node] = f"t{tmp}"
        body.append(f"\tconst double t{tmp} = {'fmin' if d[0] == 'min' else 'fmax'}({a}, {b});")
        tmp += 1
    for i, o in enumerate(outs):
        body.append(f"\tout[{i}] = {names[o]};")
    head = (f"/* {len(outs)} ranks out, {tmp} min/max operations */\n"
            f"static __device__ __forceinline__ void\n{name}(const double *acc, const double *row, double *out)\n{{")
    return "\n".join([head] + body + ["}"]) + "\n"


def check(net, outs, inputs, expect, trials=20000, seed=0):
    rng = np.random.default_rng(seed)
    for kind in range(3):
        if kind == 0:
            x = rng.normal(size=(len(inputs), trials))
        elif kind == 1:
            x = rng.integers(0, 4, size=(len(inputs), trials)).astype(float)
        else:
            x = rng.integers(0, 2, size=(len(inputs), trials)).astype(float)
        vals = dict(zip(inputs, x))
        got = np.array(net.evaluate(outs, vals))
        want = expect(x)
        assert np.array_equal(got, want), f"network mismatch ({kind})"


def emit(name, net, outs, inputs, args, ret):
    """C++ device function: inputs come from an array argument, outputs written."""
    lines = []
    live = net.live(outs)
    names = {}
    for idx, n in enumerate(inputs):
        names[net.key[("in", n)]] = f"{args}[{idx}]" if ("in", n) in net.key else None
    body = []
    tmp = 0
    for node in live:
        d = net.defs[node]
        if d[0] == "in":
            continue
        a = "INFINITY" if d[1] == INF else "-INFINITY" if d[1] == NINF else names[d[1]]
        b = "INFINITY" if d[2] == INF else "-INFINITY" if d[2] == NINF else names[d[2]]
        names[node] = f"t{tmp}"
        body.append(f"\tconst double t{tmp} = {'fmin' if d[0] == 'min' else 'fmax'}({a}, {b});")
        tmp += 1
    for i, o in enumerate(outs):
        v = "INFINITY" if o == INF else "-INFINITY" if o == NINF else names[o]
        body.append(f"\t{ret}[{i}] = {v};")
    lines.append(f"/* {len(inputs)} inputs, {len(outs)} outputs, {tmp} min/max operations */")
    lines.append(f"static __device__ __forceinline__ void\n{name}(const double *{args}, double *{ret})\n{{")
    lines += body
    lines.append("}")
    return "\n".join(lines) + "\n"


def main():
    parts = ["/* Generated by gen_networks.py -- do not edit. */\n#pragma once\n#include <math.h>\n"]
    # single-window medians
    for n, name in ((9, "median9"), (25, "median25"), (49, "median49")):
        net, outs, inputs = window_median(n)
        check(net, outs, inputs, lambda x: [np.sort(x, axis=0)[(x.shape[0] - 1) // 2]])
        parts.append(emit(name, net, outs, inputs, "v", "out"))
    # strip medians: T outputs of a k x k window, rows streamed top to bottom
    for k, t in ((5, 4), (3, 4)):
        net, outs, inputs = strip_median(k, t)

        def expect(x, k=k, t=t):
            x = x.reshape(t + k - 1, k, -1)
            return [np.median(x[o:o + k].reshape(k * k, -1), axis=0) for o in range(t)]
        check(net, outs, inputs, expect)
        parts.append(emit(f"strip_median{k}x{t}", net, outs, inputs, "r", "out"))
    # replacement selectors: lowest 13 of the 24 neighbours of a 5x5 window,
    # lowest 25 of the 48 of a 7x7 window
    for n, keep in ((24, 13), (48, 25)):
        net, outs, inputs = lower_half(n, keep)
        check(net, outs, inputs, lambda x, keep=keep: list(np.sort(x, axis=0)[:keep]))
        parts.append(emit(f"sort_low{keep}_of{n}", net, outs, inputs, "v", "out"))
    # incremental median of 49 (fine structure), one 7-value row per step
    steps = median49_steps()
    check_median49_steps(steps)
    for k, (net, outs, inputs) in enumerate(steps):
        parts.append(emit_step(f"median49_step{k}", net, outs, inputs))
    Path(__file__).with_name("networks.inc").write_text("\n".join(parts))


if __name__ == "__main__":
    main()

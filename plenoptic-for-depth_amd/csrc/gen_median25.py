"""Generate lfpm_median25.h: a min/max selection network for the median of 25 values.

Batcher's odd-even merge sort on 32 wires, wires 25..31 held at +infinity: comparators that
touch a constant wire become no-ops or wire swaps, and every comparator output that cannot
reach sorted position 12 (the 13th smallest of 25, the median) is dropped.  The network is
checked exhaustively with the 0-1 principle by tests/test_median_network.py.

Usage: python gen_median25.py > lfpm_median25.h
"""


def batcher(n):
    cs = []
    p = 1
    while p < n:
        k = p
        while k >= 1:
            for j in range(k % p, n - k, 2 * k):
                for i in range(min(k, n - j - k)):
                    if (i + j) // (2 * p) == (i + j + k) // (2 * p):
                        cs.append((i + j, i + j + k))
            k //= 2
        p *= 2
    return cs


def median_network(nreal=25, nwire=32):
    inf = None
    state = ["v%d" % i if i < nreal else inf for i in range(nwire)]
    ops = []
    for a, b in batcher(nwire):
        x, y = state[a], state[b]
        if y is inf:
            continue
        if x is inf:
            state[a], state[b] = y, inf
            continue
        k = len(ops)
        ops.append((k, x, y))
        state[a], state[b] = "n%d" % k, "x%d" % k
    target = state[nreal // 2]
    need, keep = {target}, []
    for k, x, y in reversed(ops):
        mn, mx = "n%d" % k in need, "x%d" % k in need
        if mn or mx:
            keep.append((k, x, y, mn, mx))
            need.update((x, y))
    keep.reverse()
    return keep, target


def main():
    keep, target = median_network()
    nops = sum(int(mn) + int(mx) for _, _, _, mn, mx in keep)
    print("// Generated by gen_median25.py -- do not edit.  Median of 25 values as a min/max")
    print("// selection network (%d comparators, %d min/max operations): Batcher's odd-even merge"
          % (len(keep), nops))
    print("// sort on 32 wires with 7 wires at +inf, pruned to the outputs that reach sorted")
    print("// position 12.  T is any type with elementwise MIN and MAX (scalars, packed lanes).")
    print("#pragma once")
    print("template <typename T, typename Min, typename Max>")
    print("__device__ __forceinline__ T median25(const T (&v)[25], Min vmin, Max vmax) {")
    for i in range(25):
        print("    const T v%d = v[%d];" % (i, i))
    for k, x, y, mn, mx in keep:
        if mn:
            print("    const T n%d = vmin(%s, %s);" % (k, x, y))
        if mx:
            print("    const T x%d = vmax(%s, %s);" % (k, x, y))
    print("    return %s;" % target)
    print("}")


if __name__ == "__main__":
    main()

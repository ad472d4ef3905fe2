"""The SPME spread's home-brick ranges (pme.hip, home_bricks): the closed
form (bricks before the wrap, then those after it that precede the first)
equals the brute-force walk over every base point whose B-spline stencil
reaches a brick, for all grid sizes up to 300 and orders 4 and 5; never more
than 3 bricks per axis."""


def _walk(lo, hi, K, order):
    out = []
    for d in range(min(hi - lo + order, K)):
        b = ((lo + d) % K) // 8
        if b not in out:
            out.append(b)
    return out


def _closed(lo, hi, K, order):
    ln = min(hi - lo + order, K)
    last = lo + ln - 1
    b0 = lo // 8
    n1 = min(last, K - 1) // 8 - b0 + 1
    e2 = last - K
    n = n1 + ((min(e2 // 8, b0 - 1) + 1) if (e2 >= 0 and b0 > 0) else 0)
    return [b0 + j if j < n1 else j - n1 for j in range(n)]


def test_home_bricks_closed_form():
    worst = 0
    for K in range(4, 301):
        for order in (4, 5):
            if K < order:
                continue
            for lo in range(0, K, 8):
                hi = min(lo + 8, K) - 1
                a = _walk(lo, hi, K, order)
                assert _closed(lo, hi, K, order) == a, (K, order, lo)
                worst = max(worst, len(a))
    assert worst <= 3

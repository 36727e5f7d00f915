"""Host-side models of the device sweep schedules (fluid-2d_amd/csrc/fs2d_sweeps.hip).

Each model replays the kernel's own decomposition -- tiles, launch levels,
halo loads, the in-tile evaluation order, ring handling, buffer ping-pong --
one cell at a time in float32 on the CPU, and asserts what the kernel relies
on for correctness:

* every read and write stays inside the R x R grid;
* tile-sweeps launched together (one level of the Gauss-Seidel hyperplane)
  never read a cell another of them writes, so they may run in any order;
* within a Gauss-Seidel tile, lane r at step s consumes only values produced
  at step s - 1 or loaded before the sweep.

The tests compare the models with the oracle bit for bit, so a flaw in the
schedule itself (a wrong level formula, a missing halo, a stale ring) shows
up on the CPU before the kernels run on a GPU.
"""
from __future__ import annotations

import numpy as np

F = np.float32
ROWS = 64  # Gauss-Seidel tile rows (one per lane)


def _relax(kind, b, n, s, w, e, a, denom):
    if kind == 0:
        return F(F(b + F(a * F(F(F(n + s) + w) + e))) / denom)
    return F(F(F(F(F(b + e) + w) + s) + n) / F(4.0))


def _sgn(neg, x):
    return F(-x) if neg else F(x)


def gs_levels(R, K, TW):
    GI = -(-(R - 2) // ROWS)
    GJ = -(-(R - 2) // TW)
    nlev = (GI - 1) + (GJ - 1) + 2 * (K - 1) + 1 if K > 0 else 0
    levels = [[] for _ in range(nlev)]
    for k in range(K):
        for I in range(GI):
            for J in range(GJ):
                levels[I + J + 2 * k].append((I, J, k))
    return levels


def gs_tile_model(x, b, K, kind, a, denom, neg_tb, neg_lr, zero_start, TW=64):
    """In-place Gauss-Seidel solve by the tile-hyperplane schedule of k_gs_tile."""
    R = x.shape[0]
    a, denom = F(a), F(denom)
    for level in gs_levels(R, K, TW):
        snap = x.copy()
        reads, writes, updates = [], [], []
        for (I, J, k) in level:
            rd, wr = set(), set()
            gi0, gj0 = 1 + ROWS * I, 1 + TW * J
            rows, cols = min(ROWS, R - 1 - gi0), min(TW, R - 1 - gj0)
            first = k == 0
            T = np.full((ROWS + 2, TW + 2), np.nan, np.float32)

            def rd_(i, j):
                assert 0 <= i < R and 0 <= j < R, (i, j)
                rd.add((i, j))
                return snap[i, j]

            for r in range(-1, rows + 1):
                gi = gi0 + r
                ring_i = gi == 0 or gi == R - 1
                for c in range(-1, cols + 1):
                    if r in (-1, rows) and c in (-1, cols):
                        continue
                    gj = gj0 + c
                    ring_j = gj == 0 or gj == R - 1
                    if first and zero_start:
                        val = rd_(gi, gj) if ((r == -1 and not ring_i) or (c == -1 and not ring_j)) else F(0)
                    elif not first and ring_i:
                        val = _sgn(neg_tb, rd_(1 if gi == 0 else R - 2, gj))
                    elif not first and ring_j:
                        val = _sgn(neg_lr, rd_(gi, 1 if gj == 0 else R - 2))
                    else:
                        val = rd_(gi, gj)
                    T[r + 1, c + 1] = val
            Bm = np.array([[b[gi0 + r, gj0 + c] for c in range(cols)] for r in range(rows)], np.float32).reshape(rows, cols)
            # skewed sweep: step s, lane r relaxes column s - r
            xprev = np.zeros(ROWS, np.float32)
            produced = {}  # (r, c) -> step
            for s in range(rows + cols - 1):
                from_up = np.concatenate([[F(0)], xprev[:-1]])  # wave_shr:1
                new_prev = xprev.copy()
                for lane in range(ROWS):
                    c = s - lane
                    if not (lane < rows and 0 <= c < cols):
                        continue
                    if lane == 0:
                        n = T[0, c + 1]
                    else:
                        assert produced.get((lane - 1, c)) == s - 1
                        n = from_up[lane]
                    if c == 0:
                        w = T[lane + 1, 0]
                    else:
                        assert produced.get((lane, c - 1)) == s - 1
                        w = xprev[lane]
                    # south and east must still hold sweep k-1 values
                    assert (lane + 1, c) not in produced and (lane, c + 1) not in produced
                    so, e = T[lane + 2, c + 1], T[lane + 1, c + 2]
                    assert not (np.isnan(so) or np.isnan(e) or np.isnan(n) or np.isnan(w))
                    v = _relax(kind, Bm[lane, c], n, so, w, e, a, denom)
                    T[lane + 1, c + 1] = v
                    new_prev[lane] = v
                    produced[(lane, c)] = s
                xprev = new_prev
            out = {}
            for r in range(rows):
                for c in range(cols):
                    out[(gi0 + r, gj0 + c)] = T[r + 1, c + 1]
            if k == K - 1:
                top, bottom = gi0 == 1, gi0 + rows == R - 1
                left, right = gj0 == 1, gj0 + cols == R - 1
                for c in range(cols):
                    if top:
                        out[(0, gj0 + c)] = _sgn(neg_tb, T[1, c + 1])
                    if bottom:
                        out[(R - 1, gj0 + c)] = _sgn(neg_tb, T[rows, c + 1])
                for r in range(rows):
                    if left:
                        out[(gi0 + r, 0)] = _sgn(neg_lr, T[r + 1, 1])
                    if right:
                        out[(gi0 + r, R - 1)] = _sgn(neg_lr, T[r + 1, cols])
                corner = lambda v: F(F(0.5) * F(_sgn(neg_tb, v) + _sgn(neg_lr, v)))
                if top and left:
                    out[(0, 0)] = corner(T[1, 1])
                if top and right:
                    out[(0, R - 1)] = corner(T[1, cols])
                if bottom and left:
                    out[(R - 1, 0)] = corner(T[rows, 1])
                if bottom and right:
                    out[(R - 1, R - 1)] = corner(T[rows, cols])
            for (i, j) in out:
                assert 0 <= i < R and 0 <= j < R
                wr.add((i, j))
            reads.append(rd)
            writes.append(wr)
            updates.append(out)
        # concurrency: no tile of this level reads or writes what another writes
        for t in range(len(level)):
            for u in range(len(level)):
                if t != u:
                    assert not (reads[t] & writes[u]), ("read/write overlap", level[t], level[u])
                    assert not (writes[t] & writes[u]), ("write/write overlap", level[t], level[u])
        for out in updates:
            for (i, j), val in out.items():
                x[i, j] = val
    return x


def jacobi_block_model(x, b, K, kind, a, denom, neg_tb, neg_lr, zero_start, M, JX=64, JY=32):
    """Jacobi solve by k_jacobi_block launches of M sweeps (ping-pong buffers).
    Returns the result array (the kernel's x or scratch)."""
    R = x.shape[0]
    a, denom = F(a), F(denom)
    src = x.copy()
    zero = bool(zero_start)
    done = 0
    while done < K:
        m = min(M, K - done)
        dst = np.full((R, R), np.nan, np.float32)
        W, H = JX + 2 * m, JY + 2 * m
        for by in range(-(-(R - 2) // JY)):
            for bx in range(-(-(R - 2) // JX)):
                gi0, gj0 = 1 + JY * by - m, 1 + JX * bx - m
                buf = [np.zeros((H, W), np.float32), np.full((H, W), np.nan, np.float32)]
                valid = np.zeros((H, W), bool)
                for r in range(H):
                    for c in range(W):
                        gi, gj = gi0 + r, gj0 + c
                        if 0 <= gi < R and 0 <= gj < R:
                            buf[0][r, c] = F(0) if zero else src[gi, gj]
                            valid[r, c] = True
                for s in range(1, m + 1):
                    e = m - s
                    cur, nxt = buf[(s - 1) & 1], buf[s & 1]
                    newvalid = np.zeros((H, W), bool)
                    for r in range(m - e, m + JY + e):
                        for c in range(m - e, m + JX + e):
                            gi, gj = gi0 + r, gj0 + c
                            if not (1 <= gi <= R - 2 and 1 <= gj <= R - 2):
                                continue
                            assert 1 <= r <= H - 2 and 1 <= c <= W - 2
                            for (rr, cc) in ((r - 1, c), (r + 1, c), (r, c - 1), (r, c + 1)):
                                assert valid[rr, cc], ("stale read", s, gi, gj)
                            v = _relax(kind, b[gi, gj], cur[r - 1, c], cur[r + 1, c], cur[r, c - 1], cur[r, c + 1], a, denom)
                            tb, lr = _sgn(neg_tb, v), _sgn(neg_lr, v)
                            cn = F(F(0.5) * F(tb + lr))
                            tgt = []
                            tgt.append((r, c, v))
                            if gi == 1: tgt.append((r - 1, c, tb))
                            if gi == R - 2: tgt.append((r + 1, c, tb))
                            if gj == 1: tgt.append((r, c - 1, lr))
                            if gj == R - 2: tgt.append((r, c + 1, lr))
                            if gi == 1 and gj == 1: tgt.append((r - 1, c - 1, cn))
                            if gi == 1 and gj == R - 2: tgt.append((r - 1, c + 1, cn))
                            if gi == R - 2 and gj == 1: tgt.append((r + 1, c - 1, cn))
                            if gi == R - 2 and gj == R - 2: tgt.append((r + 1, c + 1, cn))
                            for (rr, cc, val) in tgt:
                                if s < m:
                                    assert 0 <= rr < H and 0 <= cc < W
                                    nxt[rr, cc] = val
                                    newvalid[rr, cc] = True
                                else:
                                    dst[gi0 + rr, gj0 + cc] = val
                    valid = newvalid
        assert not np.isnan(dst).any(), "unwritten cells after a Jacobi launch"
        src = dst
        zero = False
        done += m
    return src


# ---------------------------------------------------------------------------
# Gauss-Seidel, temporally blocked: k_gs_unit (m sweeps per tile in skewed
# coordinates i' = i + k, j' = j + k; wave w of a unit relaxes sweep
# kappa*m + w, all waves on the same skewed anti-diagonal per step).
def gs_unit_plan(R, K, TW, M):
    n = R - 2
    GI = -(-(n + K - 1) // ROWS)
    GJ = -(-(n + K - 1) // TW)
    GK = -(-K // M)
    levels = {}
    for kap in range(GK):
        k_lo, k_hi = kap * M, min(kap * M + M, K) - 1
        for I in range(GI):
            for J in range(GJ):
                ok = False
                for k in range(k_lo, k_hi + 1):
                    if ROWS * I + 1 <= n + k and ROWS * I + ROWS >= 1 + k and \
                       TW * J + 1 <= n + k and TW * J + TW >= 1 + k:
                        ok = True
                        break
                if ok:
                    levels.setdefault(I + J + kap, []).append((I, J, kap))
    return [levels[L] for L in sorted(levels)]


def _unit_computed(gi, gj, I, J, k_lo, k_hi, n, TW):
    lo = max(k_lo, ROWS * I + 1 - gi, TW * J + 1 - gj)
    hi = min(k_hi, ROWS * I + ROWS - gi, TW * J + TW - gj)
    inside = 1 <= gi <= n and 1 <= gj <= n
    return inside and lo <= hi, hi


def gs_unit_model(x, b, K, kind, a, denom, neg_tb, neg_lr, zero_start, M, TW=64):
    """In-place Gauss-Seidel solve by the temporally blocked unit schedule."""
    R = x.shape[0]
    n = R - 2
    a, denom = F(a), F(denom)
    XH, XW = ROWS + 1 + M, TW + 1 + M
    for level in gs_unit_plan(R, K, TW, M):
        snap = x.copy()
        reads, writes, updates = [], [], []
        for (I, J, kap) in level:
            rd, wr = set(), set()
            k_lo, k_hi, k_top = kap * M, min(kap * M + M, K) - 1, kap * M + M - 1
            gi_lo, gj_lo = ROWS * I - k_top, TW * J - k_top
            X = np.zeros((XH, XW), np.float32)
            for li in range(XH):
                for lj in range(XW):
                    X[li, lj] = snap[min(max(gi_lo + li, 0), R - 1), min(max(gj_lo + lj, 0), R - 1)]

            def use(li, lj):
                gi, gj = gi_lo + li, gj_lo + lj
                assert 0 <= li < XH and 0 <= lj < XW and 0 <= gi < R and 0 <= gj < R, (li, lj, gi, gj)
                rd.add((gi, gj))
                return X[li, lj]

            xprev = np.zeros((M, ROWS), np.float32)
            written_at = {}
            for s in range(ROWS + TW - 1):
                pending = []
                step_reads = set()
                for w in range(M):
                    k = kap * M + w
                    if k >= K:
                        continue
                    from_up = np.concatenate([[F(0)], xprev[w, :-1]])
                    new_prev = xprev[w].copy()
                    for r in range(ROWS):
                        c = s - r
                        i = ROWS * I + 1 + r - k
                        j = TW * J + 1 + c - k
                        if not (0 <= c < TW and 1 <= i <= n and 1 <= j <= n):
                            continue
                        li, lj = i - gi_lo, j - gj_lo
                        zero0 = zero_start and k == 0
                        self_old = use(li, lj)
                        if i == 1:
                            N = _sgn(neg_tb, self_old) if k > 0 else (F(0) if zero_start else use(li - 1, lj))
                        elif r == 0:
                            N = use(li - 1, lj)
                        else:
                            assert written_at.get((w, r - 1, c)) == s - 1
                            N = from_up[r]
                        if j == 1:
                            Wv = _sgn(neg_lr, self_old) if k > 0 else (F(0) if zero_start else use(li, lj - 1))
                        elif c == 0:
                            Wv = use(li, lj - 1)
                        else:
                            assert written_at.get((w, r, c - 1)) == s - 1
                            Wv = xprev[w, r]
                        if i == n:
                            S = _sgn(neg_tb, self_old) if k > 0 else (F(0) if zero_start else use(li + 1, lj))
                        else:
                            S = F(0) if zero0 else use(li + 1, lj)
                        if j == n:
                            E = _sgn(neg_lr, self_old) if k > 0 else (F(0) if zero_start else use(li, lj + 1))
                        else:
                            E = F(0) if zero0 else use(li, lj + 1)
                        for (pi, pj) in ((li - 1, lj), (li + 1, lj), (li, lj - 1), (li, lj + 1)):
                            step_reads.add((pi, pj))
                        bb = b[i, j]
                        v = _relax(kind, bb, N, S, Wv, E, a, denom)
                        pending.append((li, lj, v))
                        new_prev[r] = v
                        written_at[(w, r, c)] = s
                    xprev[w] = new_prev
                # one barrier per step: no wave may write what another reads in
                # the same step
                wset = {(li, lj) for (li, lj, _) in pending}
                assert len(wset) == len(pending)
                assert not (step_reads & wset), ("in-step read/write overlap", s)
                for (li, lj, v) in pending:
                    X[li, lj] = v
            out = {}
            for li in range(XH):
                for lj in range(XW):
                    gi, gj = gi_lo + li, gj_lo + lj
                    ok, khi = _unit_computed(gi, gj, I, J, k_lo, k_hi, n, TW)
                    if not ok:
                        continue
                    out[(gi, gj)] = X[li, lj]
                    if khi == K - 1:
                        v = X[li, lj]
                        tb, lr = _sgn(neg_tb, v), _sgn(neg_lr, v)
                        cn = F(F(0.5) * F(tb + lr))
                        if gi == 1: out[(0, gj)] = tb
                        if gi == n: out[(n + 1, gj)] = tb
                        if gj == 1: out[(gi, 0)] = lr
                        if gj == n: out[(gi, n + 1)] = lr
                        if gi == 1 and gj == 1: out[(0, 0)] = cn
                        if gi == 1 and gj == n: out[(0, n + 1)] = cn
                        if gi == n and gj == 1: out[(n + 1, 0)] = cn
                        if gi == n and gj == n: out[(n + 1, n + 1)] = cn
            wr.update(out.keys())
            reads.append(rd)
            writes.append(wr)
            updates.append(out)
        for t in range(len(level)):
            for u in range(len(level)):
                if t != u:
                    assert not (reads[t] & writes[u]), ("read/write overlap", level[t], level[u])
                    assert not (writes[t] & writes[u]), ("write/write overlap", level[t], level[u])
        for out in updates:
            for (i, j), val in out.items():
                x[i, j] = val
    return x


def gs_unit_lag_model(x, b, K, kind, a, denom, neg_tb, neg_lr, zero_start, M, L, TW=64):
    """k_gs_unit with lagged waves: wave w runs its local step sigma at tick
    sigma + L*w; a barrier opens every interval of L ticks, after which each
    wave reads from LDS every operand of its L steps at once and then relaxes
    them.  Asserts that no operand a wave prefetches in an interval is written
    by another wave in the same interval, and replays the unit with the values
    present at the interval's start."""
    R = x.shape[0]
    n = R - 2
    a, denom = F(a), F(denom)
    XH, XW = ROWS + 1 + M, TW + 1 + M
    STEPS = ROWS + TW - 1
    TICKS = STEPS + L * (M - 1)
    for level in gs_unit_plan(R, K, TW, M):
        snap = x.copy()
        reads, writes, updates = [], [], []
        for (I, J, kap) in level:
            rd = set()
            k_lo, k_hi, k_top = kap * M, min(kap * M + M, K) - 1, kap * M + M - 1
            gi_lo, gj_lo = ROWS * I - k_top, TW * J - k_top
            X = np.zeros((XH, XW), np.float32)
            for li in range(XH):
                for lj in range(XW):
                    X[li, lj] = snap[min(max(gi_lo + li, 0), R - 1), min(max(gj_lo + lj, 0), R - 1)]
            xprev = np.zeros((M, ROWS), np.float32)
            for b0 in range(0, TICKS, L):
                # prefetch: every wave captures its operands for ticks b0 .. b0+L-1
                cap = {}
                cap_cells = {w: set() for w in range(M)}
                for w in range(M):
                    k = kap * M + w
                    for q in range(L):
                        s = b0 + q - L * w
                        for r in range(ROWS):
                            c = s - r
                            i, j = ROWS * I + 1 + r - k, TW * J + 1 + c - k
                            if k >= K or not (0 <= c < TW and 1 <= i <= n and 1 <= j <= n):
                                continue
                            li, lj = i - gi_lo, j - gj_lo
                            ops = {}
                            for nm, (pi, pj) in (("self", (li, lj)), ("n", (li - 1, lj)), ("s", (li + 1, lj)),
                                                 ("w", (li, lj - 1)), ("e", (li, lj + 1))):
                                assert 0 <= pi < XH and 0 <= pj < XW
                                ops[nm] = X[pi, pj]
                                cap_cells[w].add((pi, pj))
                                rd.add((gi_lo + pi, gj_lo + pj))
                            cap[(w, q, r)] = ops
                # relax
                wrote = {w: set() for w in range(M)}
                for q in range(L):
                    for w in range(M):
                        k = kap * M + w
                        s = b0 + q - L * w
                        from_up = np.concatenate([[F(0)], xprev[w, :-1]])
                        new_prev = xprev[w].copy()
                        for r in range(ROWS):
                            if (w, q, r) not in cap:
                                continue
                            c = s - r
                            i, j = ROWS * I + 1 + r - k, TW * J + 1 + c - k
                            li, lj = i - gi_lo, j - gj_lo
                            o = cap[(w, q, r)]
                            zero0 = zero_start and k == 0
                            so = o["self"]
                            if i == 1:
                                N = _sgn(neg_tb, so) if k > 0 else (F(0) if zero_start else o["n"])
                            else:
                                N = o["n"] if r == 0 else from_up[r]
                            if j == 1:
                                Wv = _sgn(neg_lr, so) if k > 0 else (F(0) if zero_start else o["w"])
                            else:
                                Wv = o["w"] if c == 0 else xprev[w, r]
                            if i == n:
                                S = _sgn(neg_tb, so) if k > 0 else (F(0) if zero_start else o["s"])
                            else:
                                S = F(0) if zero0 else o["s"]
                            if j == n:
                                E = _sgn(neg_lr, so) if k > 0 else (F(0) if zero_start else o["e"])
                            else:
                                E = F(0) if zero0 else o["e"]
                            v = _relax(kind, b[i, j], N, S, Wv, E, a, denom)
                            X[li, lj] = v
                            wrote[w].add((li, lj))
                            new_prev[r] = v
                        xprev[w] = new_prev
                for w in range(M):
                    for w2 in range(M):
                        if w != w2:
                            assert not (cap_cells[w] & wrote[w2]), ("interval race", b0, w, w2)
            out = {}
            for li in range(XH):
                for lj in range(XW):
                    gi, gj = gi_lo + li, gj_lo + lj
                    ok, khi = _unit_computed(gi, gj, I, J, k_lo, k_hi, n, TW)
                    if not ok:
                        continue
                    out[(gi, gj)] = X[li, lj]
                    if khi == K - 1:
                        v = X[li, lj]
                        tb, lr = _sgn(neg_tb, v), _sgn(neg_lr, v)
                        cn = F(F(0.5) * F(tb + lr))
                        if gi == 1: out[(0, gj)] = tb
                        if gi == n: out[(n + 1, gj)] = tb
                        if gj == 1: out[(gi, 0)] = lr
                        if gj == n: out[(gi, n + 1)] = lr
                        if gi == 1 and gj == 1: out[(0, 0)] = cn
                        if gi == 1 and gj == n: out[(0, n + 1)] = cn
                        if gi == n and gj == 1: out[(n + 1, 0)] = cn
                        if gi == n and gj == n: out[(n + 1, n + 1)] = cn
            reads.append(rd)
            writes.append(set(out.keys()))
            updates.append(out)
        for t in range(len(level)):
            for u in range(len(level)):
                if t != u:
                    assert not (reads[t] & writes[u]), ("read/write overlap", level[t], level[u])
                    assert not (writes[t] & writes[u]), ("write/write overlap", level[t], level[u])
        for out in updates:
            for (i, j), val in out.items():
                x[i, j] = val
    return x


def gs_unit_ring_model(x, b, K, kind, a, denom, neg_tb, neg_lr, M, L, TW=64, lag=None, prefetch=False):
    """k_gs_unit, final form: lagged waves (as gs_unit_lag_model) with the ring
    kept as data instead of logic.  Each relaxed edge cell also writes its
    sign-adjusted ring cell -- in LDS for the next sweep of the same unit, and
    to the grid at write-back for the next unit -- so every neighbour is read
    the same way.  A zero-start solve (pressure) is run on an iterate that the
    caller zero-filled, ring included.

    lag: ticks by which wave w + 1 trails wave w (default L).  prefetch: each
    wave reads the operands of an interval right after the barrier that opens
    the interval before it (so the reads overlap that interval's relaxations);
    asserted race-free over both intervals."""
    R = x.shape[0]
    n = R - 2
    a, denom = F(a), F(denom)
    XH, XW = ROWS + 1 + M, TW + 1 + M
    STEPS = ROWS + TW - 1
    D = L if lag is None else lag
    TICKS = STEPS + D * (M - 1)
    for level in gs_unit_plan(R, K, TW, M):
        snap = x.copy()
        reads, writes, updates = [], [], []
        for (I, J, kap) in level:
            rd = set()
            k_lo, k_hi, k_top = kap * M, min(kap * M + M, K) - 1, kap * M + M - 1
            gi_lo, gj_lo = ROWS * I - k_top, TW * J - k_top
            X = np.zeros((XH, XW), np.float32)
            for li in range(XH):
                for lj in range(XW):
                    X[li, lj] = snap[min(max(gi_lo + li, 0), R - 1), min(max(gj_lo + lj, 0), R - 1)]
            xprev = np.zeros((M, ROWS), np.float32)

            def capture(b0):
                cap, cap_cells = {}, {w: set() for w in range(M)}
                for w in range(M):
                    k = kap * M + w
                    for q in range(L):
                        s = b0 + q - D * w
                        for r in range(ROWS):
                            c = s - r
                            i, j = ROWS * I + 1 + r - k, TW * J + 1 + c - k
                            if k >= K or not (0 <= c < TW and 1 <= i <= n and 1 <= j <= n):
                                continue
                            li, lj = i - gi_lo, j - gj_lo
                            ops = {}
                            for nm, (pi, pj) in (("n", (li - 1, lj)), ("s", (li + 1, lj)), ("w", (li, lj - 1)),
                                                 ("e", (li, lj + 1))):
                                assert 0 <= pi < XH and 0 <= pj < XW
                                ops[nm] = X[pi, pj]
                                cap_cells[w].add((pi, pj))
                                rd.add((gi_lo + pi, gj_lo + pj))
                            cap[(w, q, r)] = ops
                return cap, cap_cells

            pending = capture(0) if prefetch else None
            wrote_before = {w: set() for w in range(M)}
            for b0 in range(0, TICKS, L):
                if prefetch:
                    cap, cap_cells = pending
                    pending = capture(b0 + L)
                else:
                    cap, cap_cells = capture(b0)
                wrote = {w: set() for w in range(M)}
                for q in range(L):
                    for w in range(M):
                        k = kap * M + w
                        s = b0 + q - D * w
                        from_up = np.concatenate([[F(0)], xprev[w, :-1]])
                        new_prev = xprev[w].copy()
                        for r in range(ROWS):
                            if (w, q, r) not in cap:
                                continue
                            c = s - r
                            i, j = ROWS * I + 1 + r - k, TW * J + 1 + c - k
                            li, lj = i - gi_lo, j - gj_lo
                            o = cap[(w, q, r)]
                            N = o["n"] if (r == 0 or i == 1) else from_up[r]
                            Wv = o["w"] if (c == 0 or j == 1) else xprev[w, r]
                            v = _relax(kind, b[i, j], N, o["s"], Wv, o["e"], a, denom)
                            X[li, lj] = v
                            wrote[w].add((li, lj))
                            tb, lr = _sgn(neg_tb, v), _sgn(neg_lr, v)
                            for cond, (pi, pj), val in ((i == 1, (li - 1, lj), tb), (i == n, (li + 1, lj), tb),
                                                        (j == 1, (li, lj - 1), lr), (j == n, (li, lj + 1), lr)):
                                if cond:
                                    X[pi, pj] = val
                                    wrote[w].add((pi, pj))
                            new_prev[r] = v
                        xprev[w] = new_prev
                # no other wave writes an operand between its read and its use
                for w in range(M):
                    for w2 in range(M):
                        if w != w2:
                            assert not (cap_cells[w] & wrote[w2]), ("interval race", b0, w, w2)
                            if prefetch:
                                assert not (cap_cells[w] & wrote_before[w2]), ("prefetch race", b0, w, w2)
                wrote_before = wrote
            out = {}
            for li in range(XH):
                for lj in range(XW):
                    gi, gj = gi_lo + li, gj_lo + lj
                    ok, khi = _unit_computed(gi, gj, I, J, k_lo, k_hi, n, TW)
                    if not ok:
                        continue
                    v = X[li, lj]
                    out[(gi, gj)] = v
                    tb, lr = _sgn(neg_tb, v), _sgn(neg_lr, v)
                    cn = F(F(0.5) * F(tb + lr))
                    if gi == 1: out[(0, gj)] = tb
                    if gi == n: out[(n + 1, gj)] = tb
                    if gj == 1: out[(gi, 0)] = lr
                    if gj == n: out[(gi, n + 1)] = lr
                    if gi == 1 and gj == 1: out[(0, 0)] = cn
                    if gi == 1 and gj == n: out[(0, n + 1)] = cn
                    if gi == n and gj == 1: out[(n + 1, 0)] = cn
                    if gi == n and gj == n: out[(n + 1, n + 1)] = cn
            reads.append(rd)
            writes.append(set(out.keys()))
            updates.append(out)
        for t in range(len(level)):
            for u in range(len(level)):
                if t != u:
                    assert not (reads[t] & writes[u]), ("read/write overlap", level[t], level[u])
                    assert not (writes[t] & writes[u]), ("write/write overlap", level[t], level[u])
        for out in updates:
            for (i, j), val in out.items():
                x[i, j] = val
    return x


def jacobi_wave_model(x, b, K, kind, a, denom, neg_tb, neg_lr, zero_start, M, TY=64):
    """Jacobi solve by k_jacobi_wave launches: one 64-lane wave per tile, lane =
    column of a 64-column region (output 64 - 2M columns), rows held in
    registers; per sweep, the region shrinks by one cell on each side, W/E
    come from the neighbouring lanes and N/S from the neighbouring rows, and
    ring cells inside the region are rewritten from the new interior.
    Returns the result array (the kernel's x or scratch)."""
    R = x.shape[0]
    a, denom = F(a), F(denom)
    src = x.copy()
    zero = bool(zero_start)
    done = 0
    while done < K:
        m = min(M, K - done)  # a remainder launch runs with its own margin
        OX = 64 - 2 * m
        dst = np.full((R, R), np.nan, np.float32)
        H = TY + 2 * m
        for ty in range(-(-(R - 2) // TY)):
            for tx in range(-(-(R - 2) // OX)):
                gi0, gj0 = 1 + TY * ty - m, 1 + OX * tx - m  # region origin
                gis = [gi0 + q for q in range(H)]
                gjs = [gj0 + l for l in range(64)]
                cl = lambda v, lo, hi: min(max(v, lo), hi)
                X = np.array([[F(0) if zero else src[cl(gi, 0, R - 1), cl(gj, 0, R - 1)] for gj in gjs] for gi in gis],
                             np.float32)
                B = np.array([[b[cl(gi, 1, R - 2), cl(gj, 1, R - 2)] for gj in gjs] for gi in gis], np.float32)
                valid = np.array([[0 <= gi < R and 0 <= gj < R for gj in gjs] for gi in gis])
                for s in range(1, m + 1):
                    new = X.copy()
                    for q in range(s, H - s):
                        for l in range(s, 64 - s):
                            gi, gj = gis[q], gjs[l]
                            if not (1 <= gi <= R - 2 and 1 <= gj <= R - 2):
                                continue
                            nb = (X[q - 1, l], X[q + 1, l], X[q, l - 1], X[q, l + 1])
                            for (qq, ll) in ((q - 1, l), (q + 1, l), (q, l - 1), (q, l + 1)):
                                assert valid[qq, ll]
                            new[q, l] = _relax(kind, B[q, l], nb[0], nb[1], nb[2], nb[3], a, denom)
                    # ring cells of the region, from the new interior (rows, then columns)
                    for q in range(s, H - s):
                        for l in range(s, 64 - s):
                            gi, gj = gis[q], gjs[l]
                            if gi == 0 and 1 <= gj <= R - 2: new[q, l] = _sgn(neg_tb, new[q + 1, l])
                            if gi == R - 1 and 1 <= gj <= R - 2: new[q, l] = _sgn(neg_tb, new[q - 1, l])
                    for q in range(s, H - s):
                        for l in range(s, 64 - s):
                            gi, gj = gis[q], gjs[l]
                            if gj == 0 and 1 <= gi <= R - 2: new[q, l] = _sgn(neg_lr, new[q, l + 1])
                            if gj == R - 1 and 1 <= gi <= R - 2: new[q, l] = _sgn(neg_lr, new[q, l - 1])
                    X = new
                # output: interior cells of the tile and the ring next to them
                for q in range(m, m + TY):
                    for l in range(m, m + OX):
                        gi, gj = gis[q], gjs[l]
                        if not (1 <= gi <= R - 2 and 1 <= gj <= R - 2):
                            continue
                        v = X[q, l]
                        dst[gi, gj] = v
                        tb, lr = _sgn(neg_tb, v), _sgn(neg_lr, v)
                        cn = F(F(0.5) * F(tb + lr))
                        if gi == 1: dst[0, gj] = tb
                        if gi == R - 2: dst[R - 1, gj] = tb
                        if gj == 1: dst[gi, 0] = lr
                        if gj == R - 2: dst[gi, R - 1] = lr
                        if gi == 1 and gj == 1: dst[0, 0] = cn
                        if gi == 1 and gj == R - 2: dst[0, R - 1] = cn
                        if gi == R - 2 and gj == 1: dst[R - 1, 0] = cn
                        if gi == R - 2 and gj == R - 2: dst[R - 1, R - 1] = cn
        assert not np.isnan(dst).any(), "unwritten cells after a Jacobi launch"
        src = dst
        zero = False
        done += m
    return src


# ---------------------------------------------------------------------------
# Two solves pipelined in the same level launches: the pressure solve of a
# projection (right-hand side the divergence of the diffused velocity, computed
# where each unit loads it) trailing the diffuse by `offset` levels.
def gs_unit_writes(R, I, J, kap, K, TW, M):
    """Cells (interior and ring) a unit of the diffuse plan writes back: every
    cell it relaxed, and for a unit meeting the ring the ring beside its edge
    cells (corners beside corner cells)."""
    n = R - 2
    k_lo, k_hi = kap * M, min(kap * M + M, K) - 1
    k_top = k_lo + M - 1
    gi = np.arange(ROWS * I - k_top, ROWS * I - k_top + ROWS + 1 + M)[:, None]
    gj = np.arange(TW * J - k_top, TW * J - k_top + TW + 1 + M)[None, :]
    lo = np.maximum(k_lo, np.maximum(ROWS * I + 1 - gi, TW * J + 1 - gj))
    hi = np.minimum(k_hi, np.minimum(ROWS * I + ROWS - gi, TW * J + TW - gj))
    relaxed = (gi >= 1) & (gi <= n) & (gj >= 1) & (gj <= n) & (lo <= hi)
    ii, jj = np.nonzero(relaxed)
    cells = set(zip((ii + gi[0, 0]).tolist(), (jj + gj[0, 0]).tolist()))
    ring = set()
    for (i, j) in cells:
        if i == 1:
            ring.add((0, j))
        if i == n:
            ring.add((R - 1, j))
        if j == 1:
            ring.add((i, 0))
        if j == n:
            ring.add((i, R - 1))
        if i in (1, n) and j in (1, n):
            ring.add((0 if i == 1 else R - 1, 0 if j == 1 else R - 1))
    return cells | ring


def gs_pipeline_min_offset(R, Kd, Kp, TW, M):
    """Brute force: the smallest offset such that every pressure unit runs at a
    later level than every diffuse unit that writes a velocity cell its
    divergence reads (the four neighbours of each right-hand-side cell it
    loads, clamped into the interior as the kernel clamps them)."""
    n = R - 2
    last = np.full((R, R), -1, dtype=np.int64)
    for units in gs_unit_plan(R, Kd, TW, M):
        for (I, J, kap) in units:
            lvl = I + J + kap
            for (i, j) in gs_unit_writes(R, I, J, kap, Kd, TW, M):
                last[i, j] = max(last[i, j], lvl)
    need = 0
    for units in gs_unit_plan(R, Kp, TW, M):
        for (I, J, kap) in units:
            k_top = kap * M + M - 1
            bi = np.clip(np.arange(ROWS * I - k_top + 1, ROWS * I - k_top + ROWS + M), 1, n)
            bj = np.clip(np.arange(TW * J - k_top + 1, TW * J - k_top + TW + M), 1, n)
            # the four neighbours: (i +- 1, j) and (i, j +- 1) of the region
            m = max(last[np.ix_(np.unique(np.concatenate([bi - 1, bi + 1])), np.unique(bj))].max(),
                    last[np.ix_(np.unique(bi), np.unique(np.concatenate([bj - 1, bj + 1])))].max())
            need = max(need, int(m) + 1 - (I + J + kap))
    return need


# ---- sheared units (a candidate schedule for the joint launches) -----------
# In the coordinates i' = i + k, j'' = i + j + 2k every Gauss-Seidel dependence
# of sweep k's cell (i, j) -- N (i - 1, j) and W (i, j - 1) of sweep k, S
# (i + 1, j) and E (i, j + 1) of sweep k - 1 -- sits at i' - 1 or i', and at
# j'' - 1.  So tiles of RW rows x TW columns x M sweeps in (i', j'', k) are
# legal, units with equal I + J + kappa are independent, and inside a unit lane
# r (row i' = RW I + 1 + r) relaxes j'' = TW J + 1 + s at step s: every lane is
# active in every step (the rectangles in (i + k, j + k) leave lane r idle for
# r of their RW + TW - 1 steps).  N comes from lane r - 1's previous step, W
# from the lane's own; at step 0 lane r - 1's held west halo is lane r's north.

def gs_shear_plan(R, K, TW, M, RW=64):
    n = R - 2
    GI = -(-(n + K - 1) // RW)
    GK = -(-K // M)
    levels = {}
    for kap in range(GK):
        k_lo, k_hi = kap * M, min(kap * M + M, K) - 1
        for I in range(GI):
            # j'' of band I's cells: (i' + 1 - k) + ... spans [RW I + 2, RW I + RW + n + 2 k_hi]
            for J in range((RW * I + 1) // TW, (RW * I + RW + n + 2 * k_hi) // TW + 1):
                if any(1 <= RW * I + 1 + r - k <= n and 1 <= TW * J + 1 + s - (RW * I + 1 + r) - k <= n
                       for k in range(k_lo, k_hi + 1) for r in range(RW) for s in range(TW)):
                    levels.setdefault(I + J + kap, []).append((I, J, kap))
    return [levels[L] for L in sorted(levels)]


def gs_shear_model(x, b, K, kind, a, denom, neg_tb, neg_lr, zero_start, M, TW=64, RW=64):
    """In-place Gauss-Seidel solve by sheared units, waves in lockstep (one
    barrier per step); asserts the level and in-step conditions the unit
    kernels rely on."""
    R = x.shape[0]
    n = R - 2
    a, denom = F(a), F(denom)
    for level in gs_shear_plan(R, K, TW, M, RW):
        snap = x.copy()
        reads, writes, updates = [], [], []
        for (I, J, kap) in level:
            k_lo, k_hi = kap * M, min(kap * M + M, K) - 1
            lds = {}
            rd = set()

            def use(gi, gj):
                assert 0 <= gi < R and 0 <= gj < R, (gi, gj)
                if (gi, gj) not in lds:
                    lds[(gi, gj)] = snap[gi, gj]
                    rd.add((gi, gj))
                return lds[(gi, gj)]

            def cell(w, r, s):
                k = k_lo + w
                i = RW * I + 1 + r - k
                return i, TW * J + 1 + s - (RW * I + 1 + r) - k

            # each lane holds its west halo (step 0's W) before the first step
            xprev = np.zeros((M, RW), np.float32)
            held = np.ones((M, RW), bool)
            for w in range(k_hi - k_lo + 1):
                for r in range(RW):
                    i, j = cell(w, r, 0)
                    if 1 <= i <= n and 2 <= j <= n + 1:
                        xprev[w, r] = use(i, j - 1)
            last = {}
            for s in range(TW):
                pending, step_reads = [], set()
                for w in range(k_hi - k_lo + 1):
                    k = k_lo + w
                    up = np.concatenate([[F(0)], xprev[w, :-1]])
                    new_prev = xprev[w].copy()
                    for r in range(RW):
                        i, j = cell(w, r, s)
                        if not (1 <= i <= n and 1 <= j <= n):
                            continue
                        zero0 = zero_start and k == 0
                        self_old = use(i, j)
                        if i == 1:
                            N = _sgn(neg_tb, self_old) if k > 0 else (F(0) if zero_start else use(0, j))
                        elif r == 0:
                            N = use(i - 1, j)
                        else:
                            # lane r - 1's previous step (its held west halo at step 0)
                            pi, pj = cell(w, r - 1, s - 1)
                            assert (pi, pj) == (i - 1, j)
                            N = up[r]
                        if j == 1:
                            Wv = _sgn(neg_lr, self_old) if k > 0 else (F(0) if zero_start else use(i, 0))
                        else:
                            Wv = xprev[w, r]
                        if i == n:
                            S = _sgn(neg_tb, self_old) if k > 0 else (F(0) if zero_start else use(n + 1, j))
                        else:
                            S = F(0) if zero0 else use(i + 1, j)
                        if j == n:
                            E = _sgn(neg_lr, self_old) if k > 0 else (F(0) if zero_start else use(i, n + 1))
                        else:
                            E = F(0) if zero0 else use(i, j + 1)
                        step_reads.update({(i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)})
                        v = _relax(kind, b[i, j], N, S, Wv, E, a, denom)
                        pending.append((i, j, v))
                        new_prev[r] = v
                        last[(i, j)] = k
                    xprev[w] = new_prev
                wset = {(i, j) for (i, j, _) in pending}
                assert len(wset) == len(pending)
                assert not (step_reads & wset), ("in-step read/write overlap", s)
                for (i, j, v) in pending:
                    lds[(i, j)] = v
            out = {}
            for (gi, gj), khi in last.items():
                v = lds[(gi, gj)]
                out[(gi, gj)] = v
                if khi == K - 1:
                    tb, lr = _sgn(neg_tb, v), _sgn(neg_lr, v)
                    cn = F(F(0.5) * F(tb + lr))
                    if gi == 1: out[(0, gj)] = tb
                    if gi == n: out[(n + 1, gj)] = tb
                    if gj == 1: out[(gi, 0)] = lr
                    if gj == n: out[(gi, n + 1)] = lr
                    if gi == 1 and gj == 1: out[(0, 0)] = cn
                    if gi == 1 and gj == n: out[(0, n + 1)] = cn
                    if gi == n and gj == 1: out[(n + 1, 0)] = cn
                    if gi == n and gj == n: out[(n + 1, n + 1)] = cn
            reads.append(rd)
            writes.append(set(out.keys()))
            updates.append(out)
        for t in range(len(level)):
            for u in range(len(level)):
                if t != u:
                    assert not (reads[t] & writes[u]), ("read/write overlap", level[t], level[u])
                    assert not (writes[t] & writes[u]), ("write/write overlap", level[t], level[u])
        for out in updates:
            for (i, j), val in out.items():
                x[i, j] = val
    return x


def gs_shear_lag_model(x, b, K, kind, a, denom, neg_tb, neg_lr, zero_start, M, L, TW=64, RW=64):
    """Sheared units with lagged waves, as k_gs_unit runs rectangles: wave w
    runs its step s at tick s + L w, a barrier opens every interval of L ticks,
    each wave reads every operand of its L steps right after it, and the
    lane's held west halo is read once at the unit's start.  Asserts that no
    operand a wave reads in an interval is written by another wave in the same
    interval, and that no wave writes a held halo before its lane's step 0."""
    R = x.shape[0]
    n = R - 2
    a, denom = F(a), F(denom)
    TICKS = TW + L * (M - 1)
    for level in gs_shear_plan(R, K, TW, M, RW):
        snap = x.copy()
        reads, writes, updates = [], [], []
        for (I, J, kap) in level:
            k_lo, k_hi = kap * M, min(kap * M + M, K) - 1
            lds, rd = {}, set()

            def use(gi, gj):
                assert 0 <= gi < R and 0 <= gj < R, (gi, gj)
                if (gi, gj) not in lds:
                    lds[(gi, gj)] = snap[gi, gj]
                    rd.add((gi, gj))
                return lds[(gi, gj)]

            def cell(w, r, s):
                k = k_lo + w
                return RW * I + 1 + r - k, TW * J + 1 + s - (RW * I + 1 + r) - k

            nw = k_hi - k_lo + 1
            xprev = np.zeros((M, RW), np.float32)
            halo_until = {}  # held west halo cell -> tick of the lane's step 0
            for w in range(nw):
                for r in range(RW):
                    i, j = cell(w, r, 0)
                    if 1 <= i <= n and 2 <= j <= n + 1:
                        xprev[w, r] = use(i, j - 1)
                        halo_until[(i, j - 1)] = max(halo_until.get((i, j - 1), -1), L * w)
            last, written_tick = {}, {}
            for b0 in range(0, TICKS, L):
                cap, cap_cells = {}, {w: set() for w in range(nw)}
                for w in range(nw):
                    for q in range(L):
                        s = b0 + q - L * w
                        if not 0 <= s < TW:
                            continue
                        for r in range(RW):
                            i, j = cell(w, r, s)
                            if not (1 <= i <= n and 1 <= j <= n):
                                continue
                            ops = {}
                            for nm, (pi, pj) in (("self", (i, j)), ("n", (i - 1, j)), ("s", (i + 1, j)),
                                                 ("w", (i, j - 1)), ("e", (i, j + 1))):
                                ops[nm] = use(pi, pj)
                                cap_cells[w].add((pi, pj))
                            cap[(w, q, r)] = ops
                wrote = {w: set() for w in range(nw)}
                for q in range(L):
                    for w in range(nw):
                        k = k_lo + w
                        s = b0 + q - L * w
                        if not 0 <= s < TW:
                            continue
                        up = np.concatenate([[F(0)], xprev[w, :-1]])
                        new_prev = xprev[w].copy()
                        for r in range(RW):
                            if (w, q, r) not in cap:
                                continue
                            i, j = cell(w, r, s)
                            o = cap[(w, q, r)]
                            zero0 = zero_start and k == 0
                            so = o["self"]
                            if i == 1:
                                N = _sgn(neg_tb, so) if k > 0 else (F(0) if zero_start else o["n"])
                            else:
                                N = o["n"] if r == 0 else up[r]
                            if j == 1:
                                Wv = _sgn(neg_lr, so) if k > 0 else (F(0) if zero_start else o["w"])
                            else:
                                Wv = xprev[w, r]
                            if i == n:
                                S = _sgn(neg_tb, so) if k > 0 else (F(0) if zero_start else o["s"])
                            else:
                                S = F(0) if zero0 else o["s"]
                            if j == n:
                                E = _sgn(neg_lr, so) if k > 0 else (F(0) if zero_start else o["e"])
                            else:
                                E = F(0) if zero0 else o["e"]
                            v = _relax(kind, b[i, j], N, S, Wv, E, a, denom)
                            tick = s + L * w
                            assert tick >= halo_until.get((i, j), -1), ("held halo overwritten", (i, j))
                            lds[(i, j)] = v
                            wrote[w].add((i, j))
                            last[(i, j)] = k
                            new_prev[r] = v
                        xprev[w] = new_prev
                for w in range(nw):
                    for w2 in range(nw):
                        if w != w2:
                            assert not (cap_cells[w] & wrote[w2]), ("interval race", b0, w, w2)
            out = {}
            for (gi, gj), khi in last.items():
                v = lds[(gi, gj)]
                out[(gi, gj)] = v
                if khi == K - 1:
                    tb, lr = _sgn(neg_tb, v), _sgn(neg_lr, v)
                    cn = F(F(0.5) * F(tb + lr))
                    if gi == 1: out[(0, gj)] = tb
                    if gi == n: out[(n + 1, gj)] = tb
                    if gj == 1: out[(gi, 0)] = lr
                    if gj == n: out[(gi, n + 1)] = lr
                    if gi == 1 and gj == 1: out[(0, 0)] = cn
                    if gi == 1 and gj == n: out[(0, n + 1)] = cn
                    if gi == n and gj == 1: out[(n + 1, 0)] = cn
                    if gi == n and gj == n: out[(n + 1, n + 1)] = cn
            reads.append(rd)
            writes.append(set(out.keys()))
            updates.append(out)
        for t in range(len(level)):
            for u in range(len(level)):
                if t != u:
                    assert not (reads[t] & writes[u]), ("read/write overlap", level[t], level[u])
                    assert not (writes[t] & writes[u]), ("write/write overlap", level[t], level[u])
        for out in updates:
            for (i, j), val in out.items():
                x[i, j] = val
    return x

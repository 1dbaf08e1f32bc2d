"""Lane-exact CPU emulation of the batched SW kernel's schedule
(clustkit_amd/csrc/sw_wave.hip: sw_wave_kernel<S>) for one wavefront of up
to 64/S pairs: S-lane segments, stripes of S rows, per-stripe position
windows, the DPP row_shr:1 exchange, the LDS row buffer between stripes
(initialised only where the kernel initialises it; stale values elsewhere),
and the ramp-up / steady / ramp-down step phases with their reset, band-edge,
mask and write rules.  Keys are the kernel's packed int32 keys.

Test infrastructure: it lets the CPU suite check the kernel's schedule
against the oracle on cases that would need a GPU otherwise.  It must be
kept in step with the kernel."""
import numpy as np

PREF_F, PREF_E, PREF_D, PREF_0 = 1 << 13, 2 << 13, 3 << 13, 4 << 13
POISON = 31
WAVE = 64


class Layout:
    """Narrow keys (score << 16, 3-bit preference, zero as a key) or wide
    keys (score << 15, 2-bit preference, zero by a compare)."""

    def __init__(self, wide: bool):
        self.wide = wide
        self.shift = 15 if wide else 16
        self.one = 1 << self.shift
        self.low = self.one - 1
        self.pmask = (3 if wide else 7) << 13
        self.sat = 65500 if wide else 32700
        self.hn0 = self.key(-11, 0, 0)
        self.enn = self.key(-30000, PREF_E, 0)
        self.fnn = self.key(-30000, PREF_F, 0)
        self.ene = self.key(-12, PREF_E, 0)
        self.k0 = self.key(-1, PREF_0, 0)
        self.zero_m1 = self.key(-1, 0, 0)

    def key(self, s, p, m):
        return s * self.one + p + m


def table_of(sub, lay, add=10):
    t = np.zeros((32, 32), np.int64)
    for a in range(32):
        for b in range(32):
            m = 0
            if a == 31 or b == 31:
                s = -16000
            else:
                s = int(sub[a, b]) if (a < 20 and b < 20) else -4
                m = 1 if a == b else 0
            t[a, b] = lay.key(s + add, PREF_D, m)
    return t


def geom(A, B, hint, band):
    la, lb = len(A), len(B)
    sh = min(la, lb)
    if hint > sh or hint < -sh:
        hint = 0
    if la > lb:
        A, B, la, lb, hint = B, A, lb, la, -hint
    bw = lb if (band <= 0 or lb <= 50) else band
    return A, B, la, lb, bw, hint


def run_wave(items, band, sub, rng=None, S=16, wide=False):
    """items: list (<= 64/S) of (A uint8, B uint8, hint); S lanes per pair;
    wide: the wide-score key layout.  Returns [(score, matches)], or None
    for a pair the pass hands on (score at the layout's limit)."""
    lay = Layout(wide)
    HN_ZERO, EN_NEG, FN_NEG, EN_EDGE, K0 = lay.hn0, lay.enn, lay.fnn, lay.ene, lay.k0
    KEY_ONE, PREF_MASK, SH = lay.one, lay.pmask, lay.shift
    tab = table_of(sub, lay)
    rng = rng or np.random.default_rng(0)
    NS = WAVE // S
    RAMP = 2 * (S - 1)
    G = []
    for s in range(NS):
        if s < len(items) and items[s] is not None:
            G.append(geom(*items[s][:2], items[s][2], band))
        else:
            G.append(None)
    W = np.array([2 * g[4] + 1 if g else 0 for g in G])
    nstr = np.array([(g[2] + S - 1) // S if g else 0 for g in G])
    wmax, smax = int(W.max()), int(nstr.max())
    RB = wmax + 64 + 8
    rowbuf = rng.integers(-2**30, 2**30, size=(NS, RB, 2)).astype(np.int64)  # stale garbage
    rowbuf[:, :wmax + RAMP + 1, 0] = HN_ZERO
    rowbuf[:, :wmax + RAMP + 1, 1] = EN_NEG
    seg_best = np.full(NS, HN_ZERO, np.int64)
    lane = np.arange(64)
    seg = lane // S
    lr = lane % S
    for s in range(smax):
        live = np.array([s < nstr[q] for q in range(NS)])
        pos0 = np.zeros(NS, np.int64); npos = np.zeros(NS, np.int64); edge = np.full(NS, -1)
        for q in range(NS):
            if not live[q]:
                continue
            A, B, la, lb, bw, h = G[q]
            c0 = s * S + 1 + h - bw
            kend = min(W[q] - 1, lb - c0)
            p0 = max(0, 1 - (c0 + S - 1))
            n = max(0, kend - p0 + 1)
            if n == 0:
                p0 = 0
            pos0[q], npos[q] = p0, n
            edge[q] = n - 1 if (n > 0 and kend == W[q] - 1) else -1
        tmax = int(npos.max())
        tmin = int(min([npos[q] for q in range(NS) if live[q]], default=0x7FFFFFFF))
        T = tmax + RAMP
        sb = [(npos[q] - 1 if edge[q] >= 0 else npos[q]) for q in range(NS) if live[q]]
        ramp_steps = RAMP + 2
        fast = tmin >= ramp_steps + 2
        wsteady = min(sb) if sb else 0
        # staging
        bst = np.zeros((NS, T + S), np.int64)
        acode = np.full(64, POISON)
        for q in range(NS):
            if G[q] is None:
                bst[q, :] = POISON
                continue
            A, B, la, lb, bw, h = G[q]
            c0 = s * S + 1 + h - bw
            cb = c0 + pos0[q] - (S - 1)
            for x in range(T + S):
                j = cb + x
                bst[q, x] = B[j - 1] if (live[q] and 1 <= j <= lb) else POISON
            for l in range(S):
                i = s * S + l + 1
                if live[q] and i <= la:
                    acode[q * S + l] = A[i - 1]
        hu = rowbuf[seg, pos0[seg], 0].copy()
        hout = np.full(64, HN_ZERO, np.int64); eout = np.full(64, EN_NEG, np.int64)
        hl = np.full(64, HN_ZERO, np.int64); fl = np.full(64, FN_NEG, np.int64)
        best = np.full(64, HN_ZERO, np.int64)
        n_l = npos[seg]; e_l = edge[seg]; p_l = pos0[seg]

        def step(t, mode, clamp):
            nonlocal hu, hout, eout, hl, fl, best
            kp = t - 2 * lr
            idx = np.minimum(t, n_l) if clamp else np.full(64, t)
            rbx = rowbuf[seg, p_l + 1 + idx, 0]; rby = rowbuf[seg, p_l + 1 + idx, 1]
            hd = hu
            nh = np.where(lr == 0, rbx, np.roll(hout, 1))
            eu = np.where(lr == 0, rby, np.roll(eout, 1))
            hu = nh
            if "R" in mode:
                r = kp == 0
                hl = np.where(r, HN_ZERO, hl); fl = np.where(r, FN_NEG, fl)
            tv = tab[acode, bst[seg, S - 1 - lr + t]]
            d = hd + tv
            e = (np.maximum(eu, hu) | PREF_E) - KEY_ONE
            if "E" in mode:
                e = np.where(kp == e_l, EN_EDGE, e)
            f = (np.maximum(fl, hl) | PREF_F) - KEY_ONE
            if wide:
                hh = np.maximum(np.maximum(d, e), f)
                hh = np.where(hh < 0, lay.zero_m1, hh)
            else:
                hh = np.maximum(np.maximum(d, e), np.maximum(f, K0))
            hn = (hh & ~PREF_MASK) - 10 * KEY_ONE
            upd = (hn >> SH) > (best >> SH)
            if "M" in mode:
                upd &= (kp >= 0) & (kp < n_l)
            best = np.where(upd, hn, best)
            hl = hn; fl = f; hout = hn; eout = e
            if "W" in mode:
                wr = (lr == S - 1)
                if mode != "W":  # guarded writer
                    wr &= (t - RAMP >= 0) & (t - RAMP < n_l)
                for q in range(NS):
                    L = q * S + S - 1
                    if wr[L]:
                        ix = p_l[L] - RAMP + t
                        assert 0 <= ix < wmax + RAMP + 1, (ix, wmax)
                        rowbuf[q, ix, 0] = hn[L]; rowbuf[q, ix, 1] = e[L]

        t = 0
        if fast:
            while t < ramp_steps:
                step(t, "RMW", False); t += 1
            while t + 8 <= wsteady:
                for u in range(8):
                    step(t + u, "W", False)
                t += 8
            while t < T:
                step(t, "EMW", True); t += 1
        else:
            while t < T:
                step(t, "REMW", True); t += 1
        for q in range(NS):
            if not live[q]:
                continue
            bl = best[q * S:(q + 1) * S]
            sc = bl >> SH
            w = int(np.argmax(sc))          # first (lowest lane) of the max score
            if (bl[w] >> SH) > (seg_best[q] >> SH):
                seg_best[q] = bl[w]
    out = []
    for q in range(len(items)):
        if G[q] is None:
            out.append(None); continue
        score = int(seg_best[q] >> SH) + 11
        m = int(seg_best[q] & 8191)
        if score >= lay.sat:
            out.append(None)
            continue
        if score <= 0:
            score, m = 0, 0
        out.append((score, m))
    return out


def run_wave_pipe(items, band, sub, rng=None, S=16, wide=False, stats=None):
    """Stripe-pipelined schedule (sw_pipe_kernel): 64 / S segments of S
    lanes (S = 16: four pairs, S = 8: eight), one pair each; every lane
    walks its stripes back to back, lane l two steps behind lane l-1, so
    lane 0 starts stripe g+1 while lanes 1..S-1 finish stripe g.  Stripe
    groups (stripe g of all pairs) last L_g = max(2S, longest window) steps
    for lane 0.  In the 2S steps after a
    group start, lane l switches to the new group at step 2l (row code,
    staging buffer, window, state reset, its running maximum saved); then
    steady steps while every live window holds, then tail steps with masks.
    Lane 0 takes the row above from the LDS row buffer of its own segment
    (written by lane 15 one group earlier) and, at its first cell of a group,
    the diagonal from that buffer too.  S=64: one pair per wavefront, the
    64 lanes one chain (a wavefront-wide DPP shift), transitions of 128
    steps.  wide: the wide-score key layout.

    Narrow keys carry a row offset: every value of row i holds i << 16 on
    top of its score, so E = max(E_up - 1, H_up - 11) needs no subtraction
    (the diagonal addend is s + 11, the zero candidate and the reset values
    of a lane follow its row, the row buffer's entries left of column 1 are
    refilled with the next stripe's H = 0 before it reads them, and the
    maxima lose the offset in the reduction; a pair whose offset maximum
    reaches the limit is handed on).  Transitions whose previous windows
    all span their group and whose new windows all exceed 2S positions run
    the kernel's reduced rule set (no masks; the band-edge rule only at a
    lane's last previous cell); the emulator checks it against the full
    rules at every such step and counts those groups in ``stats``; in
    groups of at least 2S + 7 steps (the kernel reads a block's row-buffer
    entries before its stores) it also checks that no block reads an entry
    stored within the block.

    Returns [(score, matches)] or None at the layout's limit."""
    NS, RAMP, XW = WAVE // S, 2 * (S - 1), 2 * S
    PAD = XW + S
    lay = Layout(wide)
    HN_ZERO, EN_NEG, FN_NEG, EN_EDGE, K0 = lay.hn0, lay.enn, lay.fnn, lay.ene, lay.k0
    KEY_ONE, PREF_MASK, SH = lay.one, lay.pmask, lay.shift
    roff = not wide
    tab = table_of(sub, lay, 11 if roff else 10)
    rng = rng or np.random.default_rng(0)
    G = [geom(*it[:2], it[2], band) if (q < len(items) and items[q] is not None) else None
         for q, it in enumerate(list(items) + [None] * (NS - len(items)))]
    W = np.array([2 * g[4] + 1 if g else 0 for g in G])
    nstr = np.array([(g[2] + S - 1) // S if g else 0 for g in G])
    wmax, smax = int(W.max()), int(nstr.max())
    RB = wmax + 2 * S + 8
    rowbuf = rng.integers(-2**30, 2**30, size=(NS, RB, 2)).astype(np.int64)
    rowbuf[:, :wmax + RAMP + 2, 0] = HN_ZERO
    rowbuf[:, :wmax + RAMP + 2, 1] = EN_NEG
    wtime = np.full((NS, RB), -1, np.int64)   # step of each entry's last store
    lane = np.arange(64)
    seg = lane // S
    lr = lane % S

    def coff(g):
        """Row offset of each lane in group(s) g (scalar or per lane)."""
        return (np.asarray(g) * S + lr + 1) * KEY_ONE if roff else np.zeros(64, np.int64)

    sat = np.zeros(NS, bool)

    def window(q, s):
        """(live, pos0, npos, edge, c0) of stripe s of pair q."""
        if G[q] is None or s >= nstr[q]:
            return False, 0, 0, -1, 0
        A, B, la, lb, bw, h = G[q]
        c0 = s * S + 1 + h - bw
        kend = min(W[q] - 1, lb - c0)
        p0 = max(0, 1 - (c0 + S - 1))
        n = max(0, kend - p0 + 1)
        if n == 0:
            p0 = 0
        return True, p0, n, (n - 1 if (n > 0 and kend == W[q] - 1) else -1), c0

    # per group (0 .. smax-1, then an empty group smax for the drain)
    grp = []
    start = 0
    for g in range(smax + 1):
        wins = [window(q, g) for q in range(NS)]
        L = max([XW] + [w[2] for w in wins])
        if g == smax:
            L = XW
        grp.append(dict(wins=wins, start=start, L=L))
        start += L
    # staging: buffer per group of codes for x in [0, L + 46)
    def stage(g):
        d = grp[g]
        buf = np.full((NS, d["L"] + PAD), POISON, np.int64)
        ac = np.full(64, POISON)
        for q in range(NS):
            live, p0, n, e, c0 = d["wins"][q]
            if not live:
                continue
            A, B, la, lb, bw, h = G[q]
            cb = c0 + p0 - (S - 1)
            for x in range(buf.shape[1]):
                j = cb + x
                buf[q, x] = B[j - 1] if 1 <= j <= lb else POISON
            for l in range(S):
                i = g * S + l + 1
                if i <= la:
                    ac[q * S + l] = A[i - 1]
        return buf, ac

    staged = {0: stage(0)}
    if stats is None:
        stats = {}
    stats.update(groups=0, fast_groups=0)
    # per-lane state
    hu = np.full(64, HN_ZERO, np.int64)
    hout = np.full(64, HN_ZERO, np.int64); eout = np.full(64, EN_NEG, np.int64)
    hl = np.full(64, HN_ZERO, np.int64); fl = np.full(64, FN_NEG, np.int64)
    best = np.full(64, HN_ZERO, np.int64); pbest = np.full(64, HN_ZERO, np.int64)
    lg = np.full(64, -1)                      # group each lane is in (-1: none yet)
    seg_best = np.full(NS, HN_ZERO, np.int64)

    def params(g):
        """per-lane (pos0, npos, edge, start) arrays of group g (g < 0: empty)."""
        if g < 0 or g > smax:
            return (np.zeros(64, np.int64), np.zeros(64, np.int64), np.full(64, -1),
                    np.full(64, -10**9))
        d = grp[g]
        p0 = np.array([d["wins"][q][1] for q in range(NS)])[seg]
        n = np.array([d["wins"][q][2] for q in range(NS)])[seg]
        e = np.array([d["wins"][q][3] for q in range(NS)])[seg]
        return p0, n, e, np.full(64, d["start"])

    def step(t, mode, fast=None, xu=None):
        nonlocal hu, hout, eout, hl, fl, best, pbest
        # lane groups at this step
        P = {}
        cur = lg.copy()
        for l in range(64):
            g = cur[l]
            if g + 1 <= smax and t >= grp[g + 1]["start"] + 2 * lr[l]:
                cur[l] = g + 1
        sw = cur != lg
        if sw.any():
            assert "T" in mode, (t, mode)
        # switching lanes: save best, reset state
        pbest = np.where(sw, best, pbest)
        best = np.where(sw, HN_ZERO + coff(cur), best)
        hl = np.where(sw, HN_ZERO + coff(cur), hl); fl = np.where(sw, FN_NEG, fl)
        in_prev = cur == lg
        lg[:] = cur
        p0 = np.zeros(64, np.int64); n = np.zeros(64, np.int64); e = np.full(64, -1)
        st = np.zeros(64, np.int64)
        for l in range(64):
            pp = params(int(cur[l]))
            p0[l], n[l], e[l], st[l] = pp[0][l], pp[1][l], pp[2][l], pp[3][l]
        kp = t - st - 2 * lr
        # lane 0's row-buffer read: its own group's window
        rbx = np.zeros(64, np.int64); rby = np.zeros(64, np.int64)
        for q in range(NS):
            L0 = q * S
            idx = min(kp[L0], n[L0]) if "C" in mode else kp[L0]
            assert idx >= -1
            if fast is not None and fast[2] is not None:
                # the kernel reads a block's entries before its stores
                assert wtime[q, p0[L0] + 1 + idx] < fast[2], (t, "fast block reads its own store")
            rbx[L0 : L0 + S] = rowbuf[q, p0[L0] + 1 + idx, 0]
            rby[L0 : L0 + S] = rowbuf[q, p0[L0] + 1 + idx, 1]
        hd = hu
        hu = np.where(lr == 0, rbx, np.roll(hout, 1))
        eu = np.where(lr == 0, rby, np.roll(eout, 1))
        codes = np.zeros(64, np.int64); ac = np.zeros(64, np.int64)
        for l in range(64):
            buf, a = staged[int(cur[l])] if cur[l] >= 0 else (None, None)
            if buf is None:
                codes[l] = POISON; ac[l] = POISON
            else:
                x = S - 1 - lr[l] + (t - st[l])
                codes[l] = buf[seg[l], x]; ac[l] = a[l]
        tv = tab[ac, codes]
        d = hd + tv
        ee = np.maximum(eu, hu) | PREF_E
        if not roff:
            ee = ee - KEY_ONE
        if fast is not None:
            # the reduced rules: every cell inside its window, the band edge
            # only at the last cell of the previous group (step 2l - 1)
            u, edge_p, blk0 = fast
            assert np.all((kp >= 0) & (kp < n)), (t, "fast transition: cell outside window")
            edge_fast = in_prev & (u == 2 * lr - 1) & (edge_p[seg] >= 0)
            assert np.array_equal(kp == e, edge_fast), (t, "fast transition: edge rule")
        if "E" in mode:
            ee = np.where(kp == e, EN_EDGE, ee)
        if roff and xu is not None:
            # The cell above a lane's last cell of the previous group (step
            # 2l - 1) comes from lane l - 1, already in the new group, S - 1
            # rows further down: its offset would lift E.  That cell is the
            # band edge, right of the matrix or outside the window, so its E
            # is the edge value in every case.
            ee = np.where(in_prev & (xu == 2 * lr - 1), EN_EDGE, ee)
        f = (np.maximum(fl, hl) | PREF_F) - KEY_ONE
        if wide:
            hh = np.maximum(np.maximum(d, ee), f)
            hh = np.where(hh < 0, lay.zero_m1, hh)
        else:
            hh = np.maximum(np.maximum(d, ee), np.maximum(f, K0 + coff(cur)))
        hn = (hh & ~PREF_MASK) - 10 * KEY_ONE
        upd = (hn >> SH) > (best >> SH)
        if "M" in mode:
            upd &= (kp >= 0) & (kp < n)
        else:
            assert np.all((kp >= 0) & (kp < n) | ~np.isin(seg, live_q)), t
        best = np.where(upd, hn, best)
        hl = hn; fl = f; hout = hn; eout = ee
        for q in range(NS):
            L15 = q * S + S - 1
            ok = (kp[L15] >= 0) and (kp[L15] < n[L15])
            if "M" not in mode:
                ok = True
            if ok:
                ix = p0[L15] + kp[L15]
                assert 0 <= ix < wmax + RAMP + 1, (ix, wmax)
                rowbuf[q, ix, 0] = hn[L15]; rowbuf[q, ix, 1] = ee[L15]
                wtime[q, ix] = t

    def reduce(g):
        if g < 0:
            return
        for q in range(NS):
            if not grp[g]["wins"][q][0]:
                continue
            po = pbest[q * S:(q + 1) * S]
            if roff:
                sat[q] |= bool(np.any(po >= lay.key(lay.sat - 11, 0, 0)))
            bl = po - coff(g)[q * S:(q + 1) * S]
            w = int(np.argmax(bl >> SH))
            if (bl[w] >> SH) > (seg_best[q] >> SH):
                seg_best[q] = bl[w]

    for g in range(smax + 1):
        d = grp[g]
        s0 = d["start"]
        # lane 0's diagonal at its first cell of the group
        for q in range(NS):
            p0 = d["wins"][q][1]
            hu[q * S] = rowbuf[q, p0, 0]
        live_q = [q for q in range(NS) if d["wins"][q][0]]
        fast = None
        if g >= 1:
            pw, Lp = grp[g - 1]["wins"], grp[g - 1]["L"]
            if all(pw[q][2] == Lp for q in range(NS)) and \
                    all(d["wins"][q][2] > XW for q in range(NS)):
                fast = np.array([pw[q][3] for q in range(NS)])
                batched = Lp >= XW + 7
                stats["fast_groups"] += 1
        stats["groups"] += 1
        for u in range(XW):
            step(s0 + u, "TEM",
                 None if fast is None else (u, fast, s0 + u - u % 8 if batched else None), u)
        reduce(g - 1)
        if g + 1 <= smax:
            staged[g + 1] = stage(g + 1) if g + 1 < smax else (None, None)
            if g + 1 == smax:
                staged[g + 1] = (np.full((NS, XW + PAD), POISON, np.int64), np.full(64, POISON))
            elif roff:
                # entries left of column 1 for the next stripe's first row:
                # H = 0 of row (g + 1) * S, with its offset
                # (all of them when this stripe's window is empty: then the
                # stripe wrote nothing, being left of column 1 or right of the
                # matrix, where the next stripe's values do not matter)
                for q in range(NS):
                    live, p0n = grp[g + 1]["wins"][q][0], grp[g + 1]["wins"][q][1]
                    if live:
                        p0c = d["wins"][q][1] if d["wins"][q][2] > 0 else wmax + RAMP + 2
                        for x in range(p0n, p0c):
                            rowbuf[q, x, 0] = HN_ZERO + (g + 1) * S * KEY_ONE
                            rowbuf[q, x, 1] = EN_NEG
        if g == smax:
            break
        ws = [(w[2] - 1 if w[3] >= 0 else w[2]) for w in d["wins"] if w[0]]
        wsteady = min(ws) if ws else 0
        u = XW
        while u + 8 <= wsteady:
            for v in range(8):
                step(s0 + u + v, "")
            u += 8
        while u < d["L"]:
            step(s0 + u, "CEM"); u += 1
    out = []
    for q in range(len(items)):
        if G[q] is None:
            out.append(None); continue
        score = int(seg_best[q] >> SH) + 11
        m = int(seg_best[q] & 8191)
        if score >= lay.sat or sat[q]:
            out.append(None)
            continue
        if score <= 0:
            score, m = 0, 0
        out.append((score, m))
    return out


# ---------------------------------------------------------------------------
# Global alignment (NW identity): the 16-lane stripe schedule of
# clustkit_amd/csrc/nw_wave.hip (nw_wave_kernel), one pair.
# ---------------------------------------------------------------------------

NW_NEG = -1000000


def nw_wave_pair(a, b, band_width, threshold, sub=None, w_cap=479, S=16):
    """Emulates one S-lane segment (S = 16 or 8) of nw_wave_kernel on one
    pair (a, b in the pair's order), asserting every LDS index against the
    buffers the kernel allocates.  Returns (identity f32, score, mask)."""
    import numpy as np
    f32 = np.float32
    scored = sub is not None
    lx, ly = len(a), len(b)
    sh, lg = min(lx, ly), max(lx, ly)
    diff = lg - sh
    if (lg > 0 and f32(sh) / f32(lg) < f32(threshold)) or diff > band_width:
        return f32(0.0), NW_NEG, False
    if sh == 0:
        return f32(0.0), 0, f32(0.0) >= f32(threshold)
    ab = min(band_width, max(10, diff + 10))
    A, B = (a, b) if lx <= ly else (b, a)
    la, lb = sh, lg
    bw = lg if (ab <= 0 or lg <= 50) else ab
    W = 2 * bw + 1
    assert W <= w_cap
    go, ge = (-11, -1) if scored else (-10, -1)
    rb_len, sb_len = w_cap + 1, ((w_cap + 48 + 3) & ~3)

    def row0(k):
        j = k - bw
        h = NW_NEG
        if j == 0:
            h = 0
        elif 1 <= j <= bw and j <= lb:
            h = go - (j - 1)
        return [h, 0, NW_NEG, 0]

    rb = [None] * rb_len
    for k in range(0, W + 1):
        rb[k] = row0(k)
    fh, fm, stopped = NW_NEG, 0, False
    L = range(S)
    for s in range((la + S - 1) // S):
        i0 = S * s + 1
        jb = i0 - bw - (S - 1)
        sb = [None] * sb_len
        for x in range(W + 3 * S - 1):
            assert x < sb_len
            j = jb + x
            c = int(B[j - 1]) if 1 <= j <= lb else 0xFF
            sb[x] = c | (min(c, 20) << 8)
        i = [i0 + l for l in L]
        real = [i[l] <= la for l in L]
        xa = [int(A[i[l] - 1]) if real[l] else 0xFE for l in L]
        kr = [bw - i[l] if i[l] <= bw else -1 for l in L]
        hr = [go + ge * (i[l] - 1) if i[l] <= bw + 1 else NW_NEG for l in L]
        klo = [max(0, bw + 1 - i[l]) if real[l] else 1 for l in L]
        khi = [min(W - 1, lb - i[l] + bw) if real[l] else 0 for l in L]
        wl = min(S - 1, la - i0)
        hout = list(hr); mout = [0] * S; eout = [NW_NEG] * S; emout = [0] * S
        F = [NW_NEG] * S; FM = [0] * S; rmax = [0] * S
        hu = [rb[0][0]] * S; mu = [rb[0][1]] * S
        T = W + 2 * (S - 1)
        for t in range(T):
            old = (list(hout), list(mout), list(eout), list(emout))
            writes = []
            idx = min(t + 1, W)
            assert 0 <= idx < rb_len and rb[idx] is not None
            r = rb[idx]
            for l in L:
                k = t - 2 * l
                hd, md = hu[l], mu[l]
                if l == 0:
                    hu[l], mu[l], eu, emu = r
                else:
                    hu[l], mu[l], eu, emu = (old[0][l - 1], old[1][l - 1], old[2][l - 1],
                                             old[3][l - 1])
                x = t - l + S - 1
                assert 0 <= x < W + 3 * S - 1 and sb[x] is not None
                assert t + S - l < sb_len          # the next step's prefetch stays inside
                v = sb[x]
                mt = 1 if (v & 0xFF) == xa[l] else 0
                if scored:
                    ra, cb = min(xa[l], 20), v >> 8
                    sc = int(sub[ra, cb]) if ra < 20 and cb < 20 else -4
                else:
                    sc = 2 if mt else -1
                d, dm = hd + sc, md + mt
                eext, eop = eu + ge, hu[l] + go
                e = eext if eext >= eop else eop
                em = emu if eext >= eop else mu[l]
                if k == W - 1:
                    e = NW_NEG
                fext, fop = F[l] + ge, old[0][l] + go
                if fext >= fop:
                    F[l] = fext
                else:
                    F[l], FM[l] = fop, old[1][l]
                h, m = d, dm
                if e > h:
                    h, m = e, em
                if F[l] > h:
                    h, m = F[l], FM[l]
                if k == kr[l]:
                    h, m, F[l] = hr[l], 0, NW_NEG
                hout[l], mout[l], eout[l], emout[l] = h, m, e, em
                if klo[l] <= k <= khi[l]:
                    rmax[l] = max(rmax[l], m)
                if l == wl and 0 <= k < W:
                    writes.append((k, [h, m, e, em]))
            for k, val in writes:
                rb[k] = val
        if any(real[l] and f32(rmax[l] + la - i[l]) / f32(la) < f32(threshold) for l in L):
            stopped = True
            break
    if stopped:
        return f32(0.0), NW_NEG, f32(0.0) >= f32(threshold)
    if lb - la <= bw:
        fh, fm = rb[lb - la + bw][0], rb[lb - la + bw][1]
    ident = f32(fm) / f32(la)
    return ident, fh, ident >= f32(threshold)


def _i32(x):
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x >= (1 << 31) else x


def nw_pack_pair(a, b, band_width, threshold, sub=None, w_cap=479, S=16):
    """Emulates one S-lane segment of nw_pack_kernel (packed keys) on one pair
    with int32 wrap-around, asserting every LDS index against the kernel's
    buffers.  The caller keeps the pair inside the packed range (the prep
    kernel's test).  Returns (identity f32, score, mask)."""
    import numpy as np
    f32 = np.float32
    scored = sub is not None
    lx, ly = len(a), len(b)
    sh, lg = min(lx, ly), max(lx, ly)
    diff = lg - sh
    if (lg > 0 and f32(sh) / f32(lg) < f32(threshold)) or diff > band_width:
        return f32(0.0), NW_NEG, False
    if sh == 0:
        return f32(0.0), 0, f32(0.0) >= f32(threshold)
    ab = min(band_width, max(10, diff + 10))
    A, B = (a, b) if lx <= ly else (b, a)
    la, lb = sh, lg
    bw = lg if (ab <= 0 or lg <= 50) else ab
    W = 2 * bw + 1
    assert W <= w_cap and la <= 8191
    go, ge = (-11, -1) if scored else (-10, -1)
    ONE, PD, PE, PF, PM, MM, NEG = 1 << 16, 3 << 13, 2 << 13, 1 << 13, 3 << 13, (1 << 13) - 1, -32000

    def key(score, pref, m):
        return _i32((score << 16) + pref + m)

    def tabv(ra, cb):
        sc = int(sub[ra, cb]) if ra < 20 and cb < 20 else -4
        return key(sc - go, PD, 0)

    add_match, add_miss = key(2 - go, PD, 1), key(-1 - go, PD, 0)
    ek_edge, fn_reset = key(NEG, PE, 0), key(NEG - 1, PF, 0)
    rb_len, sb_len = w_cap + 1, ((w_cap + 48 + 3) & ~3)

    def row0(k):
        j = k - bw
        h = NEG
        if j == 0:
            h = 0
        elif 1 <= j <= bw and j <= lb:
            h = go - (j - 1)
        return (key(h + go, 0, 0), key(NEG - 1, PE, 0))

    rb = [None] * rb_len
    for k in range(W + 1):
        rb[k] = row0(k)
    stopped = False
    L = range(S)
    for s in range((la + S - 1) // S):
        i0 = S * s + 1
        jb = i0 - bw - (S - 1)
        sb = [None] * sb_len
        for x in range(W + 3 * S - 1):
            assert x < sb_len
            j = jb + x
            c = int(B[j - 1]) if 1 <= j <= lb else 0xFF
            sb[x] = c | (min(c, 20) << 8)
        i = [i0 + l for l in L]
        real = [i[l] <= la for l in L]
        xa = [int(A[i[l] - 1]) if real[l] else 0xFE for l in L]
        kr = [bw - i[l] if i[l] <= bw else -1 for l in L]
        hn_r = [key((go + ge * (i[l] - 1) if i[l] <= bw + 1 else NEG) + go, 0, 0) for l in L]
        klo = [max(0, bw + 1 - i[l]) for l in L]
        khi = [min(W - 1, lb - i[l] + bw) for l in L]
        wl = min(S - 1, la - i0)
        hout = list(hn_r); eout = [key(NEG - 1, PE, 0)] * S; fn = [fn_reset] * S
        rmax = [0] * S
        hu = [rb[0][0]] * S
        for t in range(W + 2 * (S - 1)):
            old = (list(hout), list(eout))
            r = rb[min(t + 1, W)]
            assert r is not None
            writes = []
            for l in L:
                k = t - 2 * l
                hd = hu[l]
                if l == 0:
                    hu[l], eu = r
                else:
                    hu[l], eu = old[0][l - 1], old[1][l - 1]
                x = t - l + S - 1
                assert 0 <= x < W + 3 * S - 1 and t + S - l < sb_len
                v = sb[x]
                mt = 1 if (v & 0xFF) == xa[l] else 0
                if scored:
                    add = _i32(tabv(min(xa[l], 20), v >> 8) + mt)
                else:
                    add = add_match if mt else add_miss
                d = _i32(hd + add)
                ek = max(eu, hu[l]) | PE
                if k == W - 1:
                    ek = ek_edge
                fk = max(fn[l], old[0][l]) | PF
                h = max(d, ek, fk)
                hout[l] = _i32((h & ~PM) + go * ONE)
                eout[l] = _i32(ek - ONE)
                fn[l] = _i32(fk - ONE)
                if k == kr[l]:
                    hout[l], fn[l] = hn_r[l], fn_reset
                if real[l] and klo[l] <= k <= khi[l]:
                    rmax[l] = max(rmax[l], h & MM)
                if l == wl and 0 <= k < W:
                    writes.append((k, (hout[l], eout[l])))
            for k, val in writes:
                rb[k] = val
        if any(real[l] and f32(rmax[l] + la - i[l]) / f32(la) < f32(threshold) for l in L):
            stopped = True
            break
    if stopped:
        return f32(0.0), NW_NEG, f32(0.0) >= f32(threshold)
    fh, fm = NW_NEG, 0
    if lb - la <= bw:
        hn = rb[lb - la + bw][0]
        fh, fm = (hn >> 16) - go, hn & MM
    ident = f32(fm) / f32(la)
    return ident, fh, ident >= f32(threshold)


def nw_pipe_pair(a, b, band_width, threshold, sub=None, w_cap=479, S=16, L_extra=0,
                 g_extra=0, t_hi_other=None, t_lo_other=None, general_other=False,
                 wave_break=True):
    """Emulates one S-lane segment of the stripe-pipelined packed NW kernel
    (nw_pipe_pairs) lane by lane, with int32 wrap-around for the keys and
    every LDS index checked against the kernel's buffers.

    Every lane walks its stripes back to back: stripe g of the segment forms
    group g of L >= max(2S + 1, W) steps; lane l works on group g from its
    step 2l to step 2l + L - 1 (its cells at positions 0 .. L-1; positions
    >= W lie outside the band and are thrown away), so during the first 2S
    steps of group g (the transition) lanes l > tau/2 still finish group
    g - 1.  At its switch step 2l a lane checks the early stop of its
    previous row (max matches + remaining rows against the threshold, as an
    integer bound), clears its row maximum and takes its new row.  The last
    row of a group goes through the row buffer to lane 0 of the next group
    (the last real row of the pair's last group instead, for the result).
    Codes of b are staged per group into one of two buffers.

    Transitions whose rows (old and new) all lie inside the band's matrix
    region run a reduced rule set ("fast": no column-0 cell, the row maximum
    and the row-buffer write limited by the band end only, the band edge
    only for old rows); the others run every rule.  After the transition,
    steps in [t_lo, t_hi) -- every real row's lane at a position inside its
    row's cell range, short of the band's last column -- run without rules.

    L_extra / g_extra emulate a wavefront whose other segments need longer
    groups / more groups; t_lo_other / t_hi_other narrow the rule-free range
    and general_other forces full transitions as other segments would;
    wave_break=False keeps running after the segment is done.

    Keys carry a diagonal offset: every value of cell (i, j) holds (i + j)
    << 16 on top of its score, and the stored HN, EN and FN sit in the
    domain of the cells that read them (i + j + 1), so E and F extensions
    need no subtraction; the diagonal addend carries + 1 << 16, row 0 and
    the column-0 boundary carry their cells' offsets, and the result drops
    la + lb + 1.  Returns (identity f32, score, mask)."""
    import numpy as np
    f32 = np.float32
    scored = sub is not None
    lx, ly = len(a), len(b)
    sh, lg = min(lx, ly), max(lx, ly)
    diff = lg - sh
    if (lg > 0 and f32(sh) / f32(lg) < f32(threshold)) or diff > band_width:
        return f32(0.0), NW_NEG, False
    if sh == 0:
        return f32(0.0), 0, f32(0.0) >= f32(threshold)
    ab = min(band_width, max(10, diff + 10))
    A, B = (a, b) if lx <= ly else (b, a)
    la, lb = sh, lg
    bw = lg if (ab <= 0 or lg <= 50) else ab
    W = 2 * bw + 1
    cap = max(w_cap, 2 * S + 1)
    L = max(2 * S + 1, W) + L_extra
    assert W <= w_cap and L <= cap and la <= 8191
    go, ge = (-11, -1) if scored else (-10, -1)
    ONE, PD, PE, PF, PM, MM, NEG = 1 << 16, 3 << 13, 2 << 13, 1 << 13, 3 << 13, (1 << 13) - 1, -32000
    BIG = 1 << 30

    def key(score, pref, m):
        return _i32((score << 16) + pref + m)

    add_match, add_miss = key(2 - go, PD, 1), key(-1 - go, PD, 0)
    ek_edge, fn_reset = key(NEG, PE, 0), key(NEG - 1, PF, 0)
    rb_len = cap + 1
    sb_len = cap + 2 * S - 1

    def row0(k):
        j = k - bw
        h = NEG
        if j == 0:
            h = 0
        elif 1 <= j <= bw and j <= lb:
            h = go - (j - 1)
        return (key(h + go + (j + 1 if h != NEG else 0), 0, 0), key(NEG - 1, PE, 0))

    # x*: the smallest x with fl32(x / la) >= t, from a window around t * la
    # (the kernel's search); a row stops the pair when its maximum match
    # count + the rows after it is below x*.
    xs = next((x for x in range(0, 2 * la + 2) if f32(x) / f32(la) >= f32(threshold)),
              2 * la + 2)
    fx = int(np.floor(f32(threshold) * f32(la))) if threshold > 0 else 0
    cand = [x for x in range(max(0, fx - 2), fx + 3) if f32(x) / f32(la) >= f32(threshold)]
    assert threshold <= 0 or threshold > 2 or (cand and cand[0] == xs), (threshold, la)
    nstr = (la + S - 1) // S
    wl = la - 1 - (nstr - 1) * S

    class P:
        pass

    def rowp(g, l):
        p = P()
        live = 0 <= g < nstr
        i = g * S + l + 1
        p.real = live and i <= la
        p.xa = int(A[i - 1]) if p.real else 0xFE
        p.kr = bw - i if i <= bw else -1
        p.hn_r = key((go + ge * (i - 1) + i + 1 if i <= bw + 1 else NEG) + go, 0, 0)
        p.fsw = max(fn_reset, p.hn_r) | PF
        p.klo = max(0, bw + 1 - i)
        p.khi = min(W - 1, lb - i + bw)
        p.kspan = (p.khi - p.klo) & 0xFFFFFFFF
        p.need = xs - la + i if p.real else -BIG
        p.writer = live and l == (wl if g == nstr - 1 else S - 1)
        p.inner_row = (not p.real) or (i > bw and i <= lb - bw)
        return p

    def stage(g):
        jb = g * S + 1 - bw - (S - 1)
        buf = [None] * sb_len
        for x in range(L + 2 * S - 1):
            j = jb + x
            c = int(B[j - 1]) if (0 <= g < nstr and 1 <= j <= lb) else 0xFF
            buf[x] = c
        return buf

    def tabv(ra, cb):
        sc = int(sub[min(ra, 20), min(cb, 20)]) if ra < 20 and cb < 20 else -4
        return key(sc - go, PD, 0)

    rb = [None] * rb_len
    for k in range(L + 1):
        rb[k] = row0(k)
    junk_h, junk_e = key(NEG - 1, 0, 0), key(NEG - 1, PE, 0)
    hu = [junk_h] * S
    hout = [junk_h] * S
    eout = [junk_e] * S
    fn = [fn_reset] * S
    rmax = [0] * S
    stop = [False] * S
    sbuf = [stage(0), [0xFF] * sb_len]
    prev = [rowp(-1, l) for l in range(S)]
    smax = nstr + g_extra
    rn = None

    def step(gi, tau, cur, mode):
        nonlocal rn
        old_h, old_e = list(hout), list(eout)
        r = rn
        rn = rb[min(tau + 2, L)]
        writes = []
        for l in range(S):
            in_c = tau >= 2 * l
            p = cur[l] if in_c else prev[l]
            k = tau - 2 * l if in_c else L + tau - 2 * l
            tp = tau if in_c else L + tau
            buf = sbuf[gi & 1] if in_c else sbuf[(gi - 1) & 1]
            x = tp - l + S - 1
            assert 0 <= x < L + 2 * S - 1 <= sb_len
            v = buf[x]
            sw = mode in ("general", "fast") and tau == 2 * l
            if sw:
                stop[l] = stop[l] or rmax[l] < prev[l].need
                rmax[l] = 0
            hd = hu[l]
            if l == 0:
                hu[l], eu = r
            else:
                hu[l], eu = old_h[l - 1], old_e[l - 1]
            mt = 1 if v == p.xa else 0
            add = _i32((tabv(p.xa, v) + mt if scored else (add_match if mt else add_miss)) + ONE)
            d = _i32(hd + add)
            ek = max(eu, hu[l]) | PE
            # the full rule set, and the reduced one where it applies
            edge_g = k == W - 1
            kr_g = k == p.kr
            rng_g = ((k - p.klo) & 0xFFFFFFFF) <= p.kspan
            wr_g = p.writer and 0 <= k < W
            if mode == "fast":
                c_l = W - 1 - L + 2 * l
                valid = in_c or tau <= c_l
                edge, krr, rng_, wr = (not in_c) and tau == c_l, False, valid, p.writer and valid
                if p.real:
                    assert (edge, krr, rng_, wr) == (edge_g, kr_g, rng_g, wr_g), (gi, tau, l)
                else:
                    assert (wr, krr) == (wr_g, False) or not p.writer
            elif mode == "steady":
                edge, krr, rng_, wr = False, False, True, p.writer
                if p.real:
                    assert (edge_g, kr_g, rng_g, wr_g) == (False, False, True, p.writer)
                assert not p.writer or 0 <= k < W
            else:
                edge, krr, rng_, wr = edge_g, kr_g, rng_g, wr_g
            if edge:
                ek = ek_edge
            fk = p.fsw if sw else (max(fn[l], old_h[l]) | PF)
            h = max(d, ek, fk)
            hout[l] = _i32((h & ~PM) + (go + 1) * ONE)
            eout[l] = ek
            fn[l] = fk
            if krr:
                hout[l], fn[l] = p.hn_r, fn_reset
            if rng_:
                rmax[l] = max(rmax[l], h & MM)
            if wr:
                writes.append((k, (hout[l], eout[l])))
        for k, val in writes:
            rb[k] = val

    for gi in range(smax + 1):
        cur = [rowp(gi, l) for l in range(S)]
        hu[0] = rb[0][0]
        rn = rb[1]
        fast = (W >= 2 * S + 1 and all(p.inner_row for p in prev + cur)
                and not general_other)
        for tau in range(2 * S):
            step(gi, tau, cur, "fast" if fast else "general")
        done = any(stop) or gi >= nstr
        if gi == smax or (wave_break and done):
            break
        sbuf[(gi + 1) & 1] = stage(gi + 1)
        reals = [(l, p) for l, p in enumerate(cur) if p.real]
        t_lo = max([2 * S] + [p.klo + 2 * l for l, p in reals])
        t_hi = min([W - 1] + [p.khi + 2 * l + 1 for l, p in reals]) if reals else BIG
        if t_lo_other is not None:
            t_lo = max(t_lo, t_lo_other)
        if t_hi_other is not None:
            t_hi = min(t_hi, t_hi_other)
        t_hi = min(t_hi, L)
        for tau in range(2 * S, L):
            step(gi, tau, cur, "steady" if t_lo <= tau < t_hi else "general-tail")
        prev = cur
    if any(stop):
        return f32(0.0), NW_NEG, f32(0.0) >= f32(threshold)
    fh, fm = NW_NEG, 0
    if lb - la <= bw:
        hn = rb[lb - la + bw][0]
        fh, fm = (hn >> 16) - go - (la + lb + 1), hn & MM
    ident = f32(fm) / f32(la)
    return ident, fh, ident >= f32(threshold)

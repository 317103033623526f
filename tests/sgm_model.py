"""Independent Python models used by the tests to pin the C oracle.

``schedule`` enumerates the visits of DenseLayer.Compute's sheared-tile wavefront
(DenseStereoSgm.cs:206-222) in order. ``label_eval`` evaluates a uniform-range layer from that order
with a *label* model of the line buffers: each (half, slot) holds the path states of the last visit that
wrote it, the two margin slots never change and the guard words between path segments stay 2000. It
shares no code with the oracle's literal buffer arithmetic (LIdx offsets, stale cleanup, Rd), so their
agreement pins that arithmetic. The brute-force census and box filter are written from the definitions.
"""
import numpy as np

INIT = 2000


def schedule(W, H, T, xsw=False):
    """Yield (dir, x, y, cur) for every Pixel call, in the reference's order."""
    for d in (1, -1):
        gwN, ghN = (W + 2 * T - 1) // T, (H + T - 1) // T
        for j in range(ghN):
            for i in range(gwN):
                x0 = i * T if d > 0 else W - 1 - i * T
                y0 = j * T if d > 0 else H - 1 - j * T
                yEnd = max(-1, min(H, y0 + d * T))
                r = 0
                while True:
                    y = y0 + d * r
                    if y == yEnd:
                        break
                    xs = max(0, min(x0 - d * r, W if (xsw and d > 0) else W - 1))
                    xe = max(-1, min(W, x0 + d * (T - r)))
                    if xs != xe:
                        cur = 1 if (r & 1) == 0 else 0
                        x = xs
                        while x != xe:
                            yield d, x, y, cur
                            x += d
                    r += 1


def step(raw, prev, pmin, p1, p2):
    """One path step (DenseStereoSgm.cs:288-298) on int64 vectors; guards at d=-1 and d=cap are 2000."""
    cap = raw.size
    lm = np.concatenate(([INIT], prev[:-1]))
    lp = np.concatenate((prev[1:], [INIT]))
    p2p = min(p2 + pmin, 65535)
    m = np.minimum(np.minimum(np.minimum(lm + p1, 65535), prev), np.minimum(np.minimum(lp + p1, 65535), p2p))
    lc = np.maximum(np.minimum(raw.astype(np.int64) + m, 65535) - pmin, 0)
    assert lc.size == cap
    return lc, int(lc.min())


def label_eval(raw, T, p1, p2, xsw=False, return_pass1=False, return_visits=False):
    """raw: uint8 [H, W, cap] (uniform full range). Returns uint16 agg [H, W, cap] (and pass-1 agg)."""
    H, W, cap = raw.shape
    init = (np.full(cap, INIT, np.int64), INIT)
    slots = {}  # (half, slot) -> list of 4 (vec, min)
    agg = np.zeros((H, W, cap), np.int64)
    agg1 = None
    visits = {}
    last_dir = 1
    for d, x, y, cur in schedule(W, H, T, xsw):
        if d != last_dir:
            agg1 = agg.copy()
            last_dir = d
        visits[(d, x, y)] = visits.get((d, x, y), 0) + 1
        other = 1 - cur
        slot = x + 1
        states = []
        for p in range(4):
            ns = slot + (0 if p == 2 else (d if p == 3 else -d))
            half = cur if p == 0 else other
            st = slots.get((half, ns))
            prev, pmin = init if (st is None or ns in (0, W + 1)) else st[p]
            lc, mn = step(raw[y, x], prev, pmin, p1, p2)
            states.append((lc, mn))
            agg[y, x] += lc
        slots[(cur, slot)] = states
    if agg1 is None:
        agg1 = agg.copy()
    out = [np.minimum(agg, 65535).astype(np.uint16)]
    if return_pass1:
        out.append(np.minimum(agg1, 65535).astype(np.uint16))
    if return_visits:
        out.append(visits)
    return out[0] if len(out) == 1 else tuple(out)


def step_ranged(raw, prev, pmin, sp, s, p1, p2):
    """One path step for a cell whose planes start at s, from a predecessor vector `prev` stored for planes
    [sp, sp + len(prev)) (2000 outside it, the line buffer after the stale cleanup)."""
    cap = raw.size
    full = lambda d: prev[d - sp] if 0 <= d - sp < prev.size else INIT  # noqa: E731
    lm = np.array([full(s + j - 1) for j in range(cap)], np.int64)
    l0 = np.array([full(s + j) for j in range(cap)], np.int64)
    lp = np.array([full(s + j + 1) for j in range(cap)], np.int64)
    p2p = min(p2 + pmin, 65535)
    m = np.minimum(np.minimum(np.minimum(lm + p1, 65535), l0), np.minimum(np.minimum(lp + p1, 65535), p2p))
    lc = np.maximum(np.minimum(raw.astype(np.int64) + m, 65535) - pmin, 0)
    return lc, (int(lc.min()) if cap else 0xFFFF)


def label_eval_ranges(raws, start, count, align, T, p1, p2, xsw=False):
    """Per-pixel ranges: raws[i] is pixel i's raw vector (cap_i bytes). A slot holds, per path, the state of
    the last visit that wrote it in that visit's own plane frame. Returns (agg list, pass-1 agg list, wta)."""
    H, W = start.shape
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    init = (np.zeros(0, np.int64), INIT, 0)
    slots = {}
    agg = [np.zeros(int(caps[y, x]), np.int64) for y in range(H) for x in range(W)]
    agg1 = None
    last_dir = 1
    for d, x, y, cur in schedule(W, H, T, xsw):
        if d != last_dir:
            agg1 = [a.copy() for a in agg]
            last_dir = d
        other = 1 - cur
        slot = x + 1
        i = y * W + x
        s = int(start[y, x])
        states = []
        for p in range(4):
            ns = slot + (0 if p == 2 else (d if p == 3 else -d))
            half = cur if p == 0 else other
            st = slots.get((half, ns))
            prev, pmin, sp = init if (st is None or ns in (0, W + 1)) else st[p]
            lc, mn = step_ranged(raws[i], prev, pmin, sp, s, p1, p2)
            states.append((lc, mn, s))
            agg[i] += lc
        slots[(cur, slot)] = states
    if agg1 is None:
        agg1 = [a.copy() for a in agg]
    wta = np.zeros((H, W), np.uint16)
    for y in range(H):
        for x in range(W):
            a = np.minimum(agg[y * W + x], 65535)
            c = int(count[y, x])
            wta[y, x] = int(start[y, x]) + (int(np.argmin(a[:c])) if c else 0)
    pack = lambda v: np.concatenate([np.minimum(a, 65535) for a in v]).astype(np.uint16)  # noqa: E731
    return pack(agg), pack(agg1), wta


def random_ranges(rng, W, H, planes, cmax, align, smooth=True):
    """Per-pixel (start, count) whose padding stays inside the line-buffer segments (start - MinLo + cap <=
    MaxHi): a smooth field with jumps of both parities, one pixel at plane 0 and one reaching the top."""
    count = rng.integers(0, cmax + 1, (H, W)).astype(np.int64)
    count[rng.random((H, W)) < 0.7] = cmax
    caps = ((count + align - 1) // align) * align
    top = planes - ((cmax + align - 1) // align) * align
    if smooth:
        base = np.cumsum(rng.integers(-3, 4, (H, W)), axis=1) + np.cumsum(rng.integers(-2, 3, (H, 1)), axis=0)
        base = base - base.min()
        start = np.clip(base, 0, top)
    else:
        start = rng.integers(0, top + 1, (H, W))
    start = np.minimum(start, planes - caps)
    start[0, 0], count[0, 0] = 0, cmax
    ys, xs = H - 1, W - 1
    capm = ((cmax + align - 1) // align) * align
    start[ys, xs], count[ys, xs] = planes - capm, capm
    return start.astype(np.uint16), count.astype(np.uint16)


def census_bruteforce(img, radius):
    H, W = img.shape
    out = np.zeros((H, W), np.uint64)
    for y in range(H):
        for x in range(W):
            c = int(img[y, x])
            code, bit = 0, 0
            for dy in range(-radius, radius + 1):
                for dx in range(-radius, radius + 1):
                    if dx == 0 and dy == 0:
                        continue
                    v = int(img[min(max(y + dy, 0), H - 1), min(max(x + dx, 0), W - 1)])
                    if v < c:
                        code |= 1 << bit
                    bit += 1
            out[y, x] = code
    return out


def cost_bruteforce(cl, cr, D, cap):
    H, W = cl.shape
    raw = np.full((H, W, cap), 255, np.uint8)
    for y in range(H):
        for x in range(W):
            for d in range(D):
                raw[y, x, d] = bin(int(cl[y, x]) ^ int(cr[y, max(x - d, 0)])).count("1")
    return raw


def box3_bruteforce(img, maxval):
    H, W = img.shape
    out = np.zeros_like(img)
    for y in range(H):
        for x in range(W):
            ys = range(max(y - 1, 0), min(y + 1, H - 1) + 1)
            xs = range(max(x - 1, 0), min(x + 1, W - 1) + 1)
            s = sum(int(img[yy, xx]) for yy in ys for xx in xs)
            inv = np.float32(1.0) / (np.float32(len(xs)) * np.float32(len(ys)))
            v = np.float32(s) * inv
            out[y, x] = min(max(int(np.rint(v)), 0), maxval)
    return out


def wta_sub_lr_model(agg, D, tol=1):
    """Straight restatement of the WTA / parabola / L-R definitions, one pixel at a time."""
    H, W, cap = agg.shape
    disp = np.zeros((H, W), np.uint16)
    sub = np.zeros((H, W), np.float32)
    valid = np.zeros((H, W), np.uint8)
    for y in range(H):
        dl = []
        for x in range(W):
            a = [int(v) for v in agg[y, x, :D]]
            arg = a.index(min(a))
            dl.append(arg)
            disp[y, x] = arg
            v = np.float32(arg)
            if 0 < arg < D - 1:
                l, c, r = a[arg - 1], a[arg], a[arg + 1]
                den = l - 2 * c + r
                if den > 0:
                    v = np.float32(arg) + (np.float32(0.5) * np.float32(l - r)) / np.float32(den)
            sub[y, x] = v
        dr = []
        for xr in range(W):
            cand = [int(agg[y, xr + d, d]) for d in range(D) if xr + d < W]
            dr.append(cand.index(min(cand)))
        for x in range(W):
            xr = x - dl[x]
            valid[y, x] = 1 if (xr >= 0 and abs(dl[x] - dr[xr]) <= tol) else 0
    return disp, sub, valid


# ------------------------------------------------------------------ the line-buffer-literal schedule
# A restatement of literal.hip's row sequences and waits (lit_rev, lit_before, lit_need and the reach of each visit),
# so the tests can check on the CPU that every order the waits allow gives the reference's results.
LIT_REACH = 2


def literal_flags(start, count, align):
    """Per cell [H][W] (image coordinates): pe, whether its path 3 reads past its segment into the next slot
    (start - MinLo + cap >= MaxHi + 1), and sp, whether its path-3 write spills into the next slot (> MaxHi + 1)."""
    caps = ((count.astype(np.int64) + align - 1) // align) * align
    minlo = int(start.min())
    maxhi = int(-(-int((start.astype(np.int64) + count).max()) // align) * align)
    q = start.astype(np.int64) - minlo
    return q + caps >= maxhi + 1, q + caps > maxhi + 1


def literal_reach(T, d, row, row2, x, y, flags):
    """Columns past a visit (in pass order) whose visits of row2 it waits for: literal.hip's reach1 / reach2."""
    if flags is None or T % 2:
        return LIT_REACH
    if d < 0:
        return 1
    pe, sp = flags
    if abs(row2 - row) == 1:
        return 1 + int(pe[y, x])
    return 1 + int(sp[y, x] or (y >= 2 and sp[y - 2, x]))


def literal_rows(W, H, T, xsw, d):
    """Per row in pass order: the list of (x, y, cur, key) the kernel's wavefront visits, key = (j, i, r)."""
    gwN = (W + 2 * T - 1) // T
    rows = []
    for row in range(H):
        j, r = divmod(row, T)
        y = row if d > 0 else H - 1 - row
        cur = 1 if r % 2 == 0 else 0
        rev = (d < 0 or not xsw) and (gwN - 1) * T - r >= W
        seq = []
        for pos in range(W + int(rev)):
            c = pos if pos < W else W - 1
            i = (pos + r) // T if pos < W else gwN - 1
            seq.append((c if d > 0 else W - 1 - c, y, cur, (j, i, r)))
        rows.append(seq)
    return rows


def literal_need(W, H, T, xsw, d, row, pos, row2, reach=None):
    """How many of row2's visits the kernel waits for before row's visit pos (reach: LIT_REACH when None)."""
    reach = LIT_REACH if reach is None else reach
    gwN = (W + 2 * T - 1) // T
    j, r = divmod(row, T)
    jj, rr = divmod(row2, T)
    c = pos if pos < W else W - 1
    i = (pos + r) // T if pos < W else gwN - 1
    rev2 = int((d < 0 or not xsw) and (gwN - 1) * T - rr >= W)
    if jj < j:
        before = W + rev2
    elif jj > j:
        before = 0
    elif rr < r and i == gwN - 1:
        before = W + rev2
    else:
        before = min(max(((i + 1) if rr < r else i) * T - rr, 0), W)
    mp = min(c + reach, W - 1)
    if rev2 and c + reach >= W - 1:
        mp = W
    return min(before, mp + 1)


def literal_depth(W, T):
    return T + 4


def literal_wait_rows(W, H, T, row, pos):
    """The rows a visit waits on: rows row-2 .. row+2, and for the second visit of the border column (pos W) every
    row of its band, since the buffers it reads were last written by whichever rows of the band came before."""
    rows = {r2 for r2 in (row - 2, row - 1, row + 1, row + 2) if 0 <= r2 < H}
    if pos >= W:
        j = row // T
        rows |= set(range(j * T, min(H, j * T + T)))
        rows.discard(row)
    return sorted(rows)


def literal_interleave(W, H, T, xsw, rng, resident, flags=None):
    """One order of both passes that the kernel's waits allow: `resident` wavefronts take rows by ticket and a
    random runnable one advances. flags: literal_flags of the layer (the kernel's per-visit reach), or None for
    LIT_REACH everywhere. Returns the visits as (x, y, cur) rows, and n_plus. Raises on a deadlock."""
    out = []
    n_plus = 0
    for d in (1, -1):
        rows = literal_rows(W, H, T, xsw, d)
        prog = [0] * H
        nxt = 0
        active = []
        while nxt < H and len(active) < resident:
            active.append(nxt)
            nxt += 1
        while active:
            ready = []
            for row in active:
                pos = prog[row]
                x, y = rows[row][min(pos, len(rows[row]) - 1)][:2]
                # the second visit of the border column keeps LIT_REACH (it waits on its whole band)
                if all(prog[r2] >= literal_need(W, H, T, xsw, d, row, pos, r2,
                                                None if pos >= W else literal_reach(T, d, row, r2, x, y, flags))
                       for r2 in literal_wait_rows(W, H, T, row, pos)):
                    ready.append(row)
            if not ready:
                raise RuntimeError(f"deadlock: W={W} H={H} T={T} d={d} resident={resident} active={active}")
            row = ready[int(rng.integers(len(ready)))]
            x, y, cur, _ = rows[row][prog[row]]
            out.append((x, y, cur))
            prog[row] += 1
            if prog[row] == len(rows[row]):
                active.remove(row)
                if nxt < H:
                    active.append(nxt)
                    nxt += 1
        if d > 0:
            n_plus = len(out)
    return np.array(out, np.int32), n_plus


def spill_ranges(rng, W, H, planes, cmax, align):
    """Per-pixel (start, count) whose padding spills past the line-buffer segments: MinLo = 0, many ranges ending
    at the top plane with counts that are not multiples of the alignment, and ranges starting at plane 0 that read
    the guard and low words the spills write."""
    count = rng.integers(1, cmax + 1, (H, W)).astype(np.int64)
    start = (rng.random((H, W)) * (planes - count + 1)).astype(np.int64)
    u = rng.random((H, W))
    start[u < 0.3] = planes - count[u < 0.3]
    start[u > 0.8] = 0
    start[0, 0] = 0
    count[H - 1, W - 1] = max(1, cmax - (cmax % align == 0))
    start[H - 1, W - 1] = planes - count[H - 1, W - 1]
    return start.astype(np.uint16), count.astype(np.uint16)


class _RefIndexError(IndexError):
    """Where the reference's bounds-checked C# array access would throw."""


_E = [np.array(b, np.uint32).view(np.float32)[()] for b in (0x3d9fcb52, 0x3e677e26, 0x3f322226, 0x3f7ffb19)]


def p2_fast_exp2(p2_base, g, q, gw):
    """Pixel's adaptive P2 (DenseStereoSgm.cs:266-273) in float32, one rounded operation at a time: the weighted
    guidance differences, S = (|d0| + |d2|) + |d1|, e = clamp(-S, -126, 128), the cubic in the fraction (highest power
    first, ExpPoly3..0), the exponent added to its bits, the truncating conversion of p2_base * E."""
    f32 = np.float32
    d = [(f32(g[c]) - f32(q[c])) * f32(gw[c]) for c in range(3)]
    S = (abs(d[0]) + abs(d[2])) + abs(d[1])
    e = max(min(-S, f32(128.0)), f32(-126.0))
    ei = int(e) + (-1 if e < 0 else 0)
    f = e - f32(ei)
    poly = ((_E[0] * f + _E[1]) * f + _E[2]) * f + _E[3]
    E = np.array(int(np.array(poly, np.float32).view(np.int32)) + (ei << 23), np.int64).astype(np.uint32).view(np.float32)
    return int(f32(p2_base) * f32(E))


def word_layer(raws, start, count, align, T, p1, p2, xsw=False, skip=None, guidance=None, gw=None, p2_scale=None,
               past_end=0):
    """DenseLayer.Compute over the reference's line buffers, word for word, written from DenseStereoSgm.cs without the
    C oracle: AllocLineBuffers and LIdx (:158-166), the tile schedule (:206-222, via ``schedule``), Pixel's stale
    cleanup (:254-264), neighbour reads through Rd (:318-324: below 0 -> 2000, past the end -> 0), the scalar
    recurrence in its 8-lane blocks (:300-313), the range store (:314-316) and the first-minimum WTA (:225-231).
    P2 is the constant p2, or with ``guidance`` ([H, W, 4] RGBA8), ``gw`` and ``p2_scale`` the adaptive one against the
    guidance stored with the slot read (_Pix, :239, 0 where nothing was stored). ``past_end`` is the value a read past
    the buffer's end returns (the reference's 0; another value shows the rule decides results). The vector form
    (:280-299) is not modelled: its
    conditions (whole blocks, every access inside the buffer, written range apart from the range read) make it give
    the scalar form's words.

    raws: list of per-pixel uint8 vectors (cap_i bytes, raster order); start/count: [H, W]; skip: [H, W] non-zero =
    skipped (raw 0 over the count, 255 over the padding, :242) or None. Returns (agg, agg_pass1, wta, events):
    agg/agg_pass1 packed like raws (uint16), wta [H, W] (start + arg), and a dict counting what the layer did with
    the words past a range: writes to a segment's last word ('guard'), into the next path's segment of the slot
    ('next_segment'), path 3's writes into the next slot's lead word ('next_lead') and path-0 segment
    ('next_slot_path0'), reads of a word last written by another path or slot ('foreign_read'), reads past the end of
    the buffer ('past_end'), and reads below its start ('below_start').
    Raises _RefIndexError where the reference's array accesses go out of bounds (a write past the buffer, or an
    aggregate lane past cap: the scalar blocks run to roundup(cap, 8) and cap is then not a multiple of 8)."""
    H, W = start.shape
    st = [int(v) for v in start.reshape(-1)]
    cn = [int(v) for v in count.reshape(-1)]
    caps = [((c + align - 1) // align) * align for c in cn]
    sk = [0] * (W * H) if skip is None else [int(v) for v in np.asarray(skip).reshape(-1)]
    min_lo = min(st)
    max_hi = max(s + c for s, c in zip(st, cn))
    M = ((max_hi + align - 1) // align) * align  # MaxHi, rounded (:51, :61)
    stride = 4 * M + 5
    slots = W + 2
    half_l, half_m = slots * stride, slots * 4
    L = [INIT] * (2 * half_l)
    owner = [None] * (2 * half_l)  # (path, half, slot) of the last write of each word
    Min = [INIT] * (2 * half_m)
    Pix = [(0.0, 0.0, 0.0, 0.0)] * (2 * slots)
    gimg = None if guidance is None else np.asarray(guidance).reshape(W * H, 4)
    p2_base = None if guidance is None else np.float32(p1) * np.float32(p2_scale)
    RS = [0] * (2 * slots)
    RC = [align] * (2 * slots)
    ev = dict(guard=0, next_segment=0, next_lead=0, next_slot_path0=0, foreign_read=0, past_end=0, below_start=0)
    agg = [[0] * c for c in caps]
    agg1 = None

    def lidx(half, slot, path):
        return half * half_l + slot * stride + 1 + path * (M + 1)

    def rd(i, path, half, slot):
        if i < 0:
            ev["below_start"] += 1
            return INIT
        if i >= len(L):
            ev["past_end"] += 1
            return past_end
        o = owner[i]
        if o is not None and o != (path, half, slot):
            ev["foreign_read"] += 1
        return L[i]

    last_dir = 1
    for d, x, y, cur in schedule(W, H, T, xsw):
        if d != last_dir:
            agg1 = [list(a) for a in agg]
            last_dir = d
        other = 1 - cur
        slot = x + 1
        pi = y * W + x
        g = None if gimg is None else tuple(float(v) for v in gimg[pi])
        if g is not None:
            Pix[cur * slots + slot] = g
        s0, c0, cap = st[pi], cn[pi], caps[pi]
        if sk[pi]:
            raw = [0] * c0 + [255] * (cap - c0)
        else:
            raw = [int(v) for v in raws[pi]]
        raw = raw + [0] * 8  # the scalar blocks read up to 7 bytes past cap of the caller's scratch (:245-246)
        sp, cp = RS[cur * slots + slot], RC[cur * slots + slot]
        for p in range(4):
            base = lidx(cur, slot, p)
            q0 = sp - min_lo
            k0, k1 = max(0, -q0), min(cp, len(L) - base - q0)
            for k in range(k0, k1):  # stale cleanup of the slot's previous range (:261-264)
                L[base + q0 + k] = INIT
                owner[base + q0 + k] = None
            ns = slot + (0 if p == 2 else (d if p == 3 else -d))
            half = cur if p == 0 else other
            min_prev = Min[half * half_m + ns * 4 + p]
            P2 = p2 if g is None else p2_fast_exp2(p2_base, g, Pix[half * slots + ns], gw)
            p2p = min(P2 + min_prev, 65535)
            if cap == 0:
                Min[cur * half_m + slot * 4 + p] = 0xFFFF
                continue
            nb = lidx(half, ns, p)
            run = 0xFFFF
            q = s0 - min_lo
            for b in range(0, cap, 8):
                for l in range(8):
                    qi = q + b + l
                    lm = rd(nb + qi - 1, p, half, ns)
                    lp = rd(nb + qi + 1, p, half, ns)
                    l0 = rd(nb + qi, p, half, ns)
                    m = min(min(min(lm + p1, 65535), l0), min(min(lp + p1, 65535), p2p))
                    lc = max(min(raw[b + l] + m, 65535) - min_prev, 0)
                    wi = base + qi
                    if wi < 0 or wi >= len(L):
                        raise _RefIndexError(f"line-buffer write at {wi} of {len(L)}")
                    if qi == M:
                        ev["guard"] += 1
                    elif qi > M:
                        if p < 3:
                            ev["next_segment"] += 1
                        elif qi == M + 1:
                            ev["next_lead"] += 1
                        else:
                            ev["next_slot_path0"] += 1
                    L[wi] = lc
                    owner[wi] = (p, cur, slot)
                    run = min(run, lc)
                    if b + l >= cap:
                        raise _RefIndexError(f"aggregate lane {b + l} of cap {cap}")
                    agg[pi][b + l] = min(agg[pi][b + l] + lc, 65535)
            Min[cur * half_m + slot * 4 + p] = run
        RS[cur * slots + slot] = s0
        RC[cur * slots + slot] = cap
    if agg1 is None:
        agg1 = [list(a) for a in agg]
    wta = np.zeros((H, W), np.uint16)
    for i in range(W * H):
        best, arg = 0xFFFF, 0
        for k in range(cn[i]):
            if agg[i][k] < best:
                best, arg = agg[i][k], k
        wta.flat[i] = st[i] + arg
    pack = lambda v: np.array([w for a in v for w in a], np.uint16)  # noqa: E731
    return pack(agg), pack(agg1), wta, ev

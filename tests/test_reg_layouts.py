"""CPU model of the register-resident NTT stages (pi_compute_amd/csrc/ntt_regs.h)
and of the lane layouts of k_row_reg (rows of 2^10, 2^11 and 2^12 points), k_col_fwd_reg and k_col_inv_reg.

The model moves element labels through a wave exactly as the two gfx950 lane
swaps do (v_permlane32_swap: lanes 32..63 of the first register trade with
lanes 0..31 of the second; v_permlane16_swap: odd 16-lane rows of the first
with even rows of the second -- the lane layout measured on an MI355X), then
checks against the textbook in-place Cooley-Tukey network:
* every butterfly of the six-layer stage pairs the two positions that
  layer pairs, and uses the twiddle of their block;
* the post-swap layout is the one the kernels address (post_t);
* every kernel layout covers each position of its tile exactly once;
* every LDS access hits 32 distinct banks in each half-wave (32 banks,
  measured), and the global accesses of a wave are 16-word runs.
"""
import itertools

import numpy as np


def entry_t(i, lane):
    return (i << 2) | (lane >> 4)


def post_t(r, lane, sb=1):
    """Element of post-layout register r: v_permlane32_swap traded register bit
    3 with lane bit 5, v_permlane16_swap register bit sb with lane bit 4."""
    i3, i2, i1, i0 = (r >> 3) & 1, (r >> 2) & 1, (r >> 1) & 1, r & 1
    l5, l4 = (lane >> 5) & 1, (lane >> 4) & 1
    if sb == 1:
        return (l5 << 5) | (i2 << 4) | (l4 << 3) | (i0 << 2) | (i3 << 1) | i1
    return (l5 << 5) | (l4 << 4) | (i1 << 3) | (i0 << 2) | (i3 << 1) | i2


def post5_t(r, lane):
    """Five-layer stage: entry t = (i3 i2 i1 i0 L5), post t = (L5 i2 i1 i0 i3)."""
    return (((lane >> 5) & 1) << 4) | ((r & 7) << 1) | ((r >> 3) & 1)


def permlane32_swap(x, y):
    x, y = x.copy(), y.copy()
    nx = np.concatenate([x[:32], y[:32]])
    ny = np.concatenate([x[32:], y[32:]])
    return nx, ny


def permlane16_swap(x, y):
    nx, ny = x.copy(), y.copy()
    for row in (0, 2):
        lo, hi = slice(16 * row, 16 * row + 16), slice(16 * row + 16, 16 * row + 32)
        nx[hi] = y[lo]
        ny[lo] = x[hi]
    return nx, ny


def lane_swaps(v, sb=1, with16=True):
    v = [c.copy() for c in v]
    for i in range(8):
        v[i], v[i | 8] = permlane32_swap(v[i], v[i | 8])
    if with16:
        for i in range(16):
            if i & (1 << sb):
                continue
            v[i], v[i | (1 << sb)] = permlane16_swap(v[i], v[i | (1 << sb)])
    return v


def test_swap_model_matches_measured_lane_layout():
    a, b = np.arange(64), 100 + np.arange(64)
    x, y = permlane32_swap(a, b)
    assert list(x) == list(range(32)) + list(range(100, 132))
    assert list(y) == list(range(32, 64)) + list(range(132, 164))
    x, y = permlane16_swap(a, b)
    assert list(x[:32]) == list(range(16)) + list(range(100, 116))
    assert list(y[32:]) == list(range(48, 64)) + list(range(148, 164))


def test_swaps_give_post_layout_and_undo_themselves():
    lanes = np.arange(64)
    v = [np.array([entry_t(i, l) for l in lanes]) for i in range(16)]
    for sb in (1, 2):
        w = lane_swaps(v, sb)
        for r in range(16):
            assert list(w[r]) == [post_t(r, l, sb) for l in lanes]
        back = lane_swaps(w, sb)
        for i in range(16):
            assert list(back[i]) == list(v[i])
    # Five-layer stage: groups on lanes l and l + 32, entry t = (i L5).
    v5 = [np.array([(i << 1) | ((l >> 5) & 1) for l in lanes]) for i in range(16)]
    w5 = lane_swaps(v5, with16=False)
    for r in range(16):
        assert list(w5[r]) == [post5_t(r, l) for l in lanes]


def ct6_pairs(sb):
    """(layout, register a, register b, layer, block) of every butterfly of
    ct6<SB = sb>, as the kernel forms them: layers 0..3 pair registers 8 >> l
    apart (block = the top l register bits); layer 4 pairs r and r | 8 for r
    with bits 3 and sb clear and their twin r | 2^sb, with the block
    post_t(r) >> 2; layer 5 pairs r and r | 2^sb with the block post_t(r) >> 1."""
    out = []
    for l in range(4):
        h = 8 >> l
        for sub in range(1 << l):
            for x in range(h):
                out.append(("entry", sub * 2 * h + x, sub * 2 * h + x + h, l, lambda lane, sub=sub: sub))
    for r in range(16):
        if r & 8 or r & (1 << sb):
            continue
        for q in (0, 1):
            rr = r | (q << sb)
            out.append(("post", rr, rr | 8, 4, lambda lane, r=r: post_t(r, lane, sb) >> 2))
    for r in range(16):
        if r & (1 << sb):
            continue
        out.append(("post", r, r | (1 << sb), 5, lambda lane, r=r: post_t(r, lane, sb) >> 1))
    return out


def test_ct6_butterflies_follow_the_textbook_network():
    # In-place CT over a 64-element group: layer m pairs t and t + 2^(5-m)
    # (bit 5-m of t clear) with the twiddle of block t >> (6-m).
    for sb in (1, 2):
        pairs = ct6_pairs(sb)
        assert len(pairs) == 6 * 8
        for lane in range(64):
            for kind, ra, rb, m, blk in pairs:
                ta = entry_t(ra, lane) if kind == "entry" else post_t(ra, lane, sb)
                tb = entry_t(rb, lane) if kind == "entry" else post_t(rb, lane, sb)
                d = 1 << (5 - m)
                assert ta & d == 0 and tb == ta + d, (sb, lane, ra, rb, m)
                assert blk(lane) == ta >> (6 - m), (sb, lane, ra, m)


def test_ct5_butterflies_follow_the_textbook_network():
    # 32-element groups: layers 0..3 on registers (entry t = (i L5)), layer 4
    # pairs r and r | 8 after the swap with the block (L5 i2 i1 i0).
    for lane in range(64):
        l5 = (lane >> 5) & 1
        for l in range(4):
            h = 8 >> l
            for sub in range(1 << l):
                for x in range(h):
                    ra, rb = sub * 2 * h + x, sub * 2 * h + x + h
                    ta, tb = (ra << 1) | l5, (rb << 1) | l5
                    d = 1 << (4 - l)
                    assert ta & d == 0 and tb == ta + d and ta >> (5 - l) == sub
        for r in range(8):
            ta, tb = post5_t(r, lane), post5_t(r | 8, lane)
            assert ta & 1 == 0 and tb == ta + 1 and ta >> 1 == (l5 << 3) | r


def _banks_ok(addrs):
    """addrs: 64 LDS word addresses of one wave instruction."""
    for half in (slice(0, 32), slice(32, 64)):
        if len(set(a % 32 for a in addrs[half])) != 32:
            return False
    return True


ROW_GEOMETRY = {  # LC: (swap-B register bit, LDS pad shift, pad words) as in RegRows
    10: (2, 4, 1),
    11: (2, 5, 1),
    12: (1, 6, 2),
}


def _row_positions(LC, ell, reg):
    """Positions of register `reg` of row-local thread ell in each layout of
    k_row_reg<LC> (RegRows in ntt_rows.hip)."""
    sb = ROW_GEOMETRY[LC][0]
    GA = LC - 6
    lane, wr = ell & 63, ell >> 6
    j = (wr << 4) | (lane & 15)
    lanepos = ((lane >> 4) << GA) | j
    if LC == 12:
        u, tb = j, (reg << 2) | (lane >> 4)
    elif LC == 11:
        u, tb = (wr << 5) | (lane & 31), (reg << 1) | (lane >> 5)
    else:
        u, tb = lane, reg
    if LC == 12:
        tpost = post_t(reg, lane, 1)
    elif LC == 11:
        tpost = post5_t(reg, lane)
    else:
        tpost = reg
    return {"a_entry": (reg << (LC - 4)) | lanepos, "a_post": (post_t(reg, lane, sb) << GA) | j,
            "b_entry": (u << GA) | tb, "b_post": (u << GA) | tpost, "rho": (reg << (LC - 4)) | ell,
            "a_group": j, "b_group": u}


def test_row_reg_layouts():
    for LC, (sb, psh, pw) in ROW_GEOMETRY.items():
        L = 1 << LC
        lds = lambda k: k + pw * (k >> psh)
        names = ("a_entry", "a_post", "b_entry", "b_post", "rho")
        seen = {name: [] for name in names}
        for wr in range(L // 16 // 64):  # waves of one row
            for reg in range(16):
                acc = {name: [] for name in names}
                for lane in range(64):
                    pos = _row_positions(LC, (wr << 6) | lane, reg)
                    for name in names:
                        acc[name].append(pos[name])
                for name in names:
                    seen[name] += acc[name]
                # LDS: stage A post (write/read) and stage B entry (read/write).
                assert _banks_ok([lds(k) for k in acc["a_post"]]), (LC, "stage A post")
                assert _banks_ok([lds(k) for k in acc["b_entry"]]), (LC, "stage B entry")
                # Global: stage A entry and register-order b rows, 16-word runs.
                for name in ("a_entry", "rho"):
                    ks = acc[name]
                    for g in range(4):
                        run = ks[16 * g:16 * g + 16]
                        assert run == list(range(run[0], run[0] + 16)), (LC, name)
        for name, ks in seen.items():
            assert sorted(ks) == list(range(L)), (LC, name)
        # Groups: stage A's are the positions j + 2^(LC-6) t (t = the six layers'
        # bits), stage B's the 2^(LC-6) positions 2^(LC-6) u + t'.
        GA = LC - 6
        for ell, reg in itertools.product(range(L // 16), range(16)):
            pos = _row_positions(LC, ell, reg)
            for key in ("a_entry", "a_post"):
                assert pos[key] & ((1 << GA) - 1) == pos["a_group"]
            for key in ("b_entry", "b_post"):
                assert pos[key] >> GA == pos["b_group"]
        assert max(lds(k) for k in range(L)) < L + pw * (L >> psh)


def _col_model(LR):
    """LDS address, stage B group bits (hi) and stage B rows of the register
    column kernels (RegCols in ntt_cols.hip)."""
    S, PSH = {7: (4, 6), 8: (4, 6), 9: (2, 6), 10: (1, 6), 11: (1, 7)}[LR]
    NT = 1 << (LR - 1)
    lds = lambda r, w: ((r << 3) | w) + S * (((r << 3) | w) >> PSH)

    def hi_bits(lane, wave):
        l3, l4, l5 = (lane >> 3) & 1, (lane >> 4) & 1, (lane >> 5) & 1
        if LR == 11:
            return (wave & 3) | (l3 << 2) | (l4 << 3) | ((wave >> 2) << 4)
        if LR == 10:
            return l5 | ((wave & 1) << 1) | (l3 << 2) | (l4 << 3) | ((wave >> 1) << 4)
        if LR == 9:
            return l5 | (l3 << 1) | (l4 << 2) | (wave << 3)
        return l3 | (l4 << 1) | (l5 << 2) | (wave << 3)

    def row_b(i, lane, hi, post=False):
        if LR == 11:
            return (hi << 5) | (post5_t(i, lane) if post else (i << 1) | ((lane >> 5) & 1))
        return (hi << 4) | i

    return S, NT, lds, hi_bits, row_b


def test_col_reg_layouts():
    for LR in (7, 8, 9, 10, 11):
        S, NT, lds, hi_bits, row_b = _col_model(LR)
        sb = 2 if LR == 7 else 1  # RegCols::SB
        R = 1 << LR
        a_entry, a_post, b_rows, b_post = [], [], [], []
        for wave in range(NT // 64):
            for reg in range(16):
                ae, ap, br, bp = [], [], [], []
                for lane in range(64):
                    g = (wave << 4) | (lane & 15)
                    w, rlow = g & 7, g >> 3
                    ae.append(((entry_t(reg, lane) << (LR - 6)) | rlow, w))
                    ap.append(((post_t(reg, lane, sb) << (LR - 6)) | rlow, w))
                    hi = hi_bits(lane, wave)
                    br.append((row_b(reg, lane, hi), lane & 7))
                    bp.append((row_b(reg, lane, hi, True), lane & 7))
                assert _banks_ok([lds(r, w) for r, w in ap]), (LR, "stage A post")
                assert _banks_ok([lds(r, w) for r, w in br]), (LR, "stage B")
                # Global rows: 8 consecutive columns of one row per 8 lanes.
                for seq in (ae, bp):
                    for g8 in range(8):
                        run = seq[8 * g8:8 * g8 + 8]
                        assert len({r for r, _ in run}) == 1 and [w for _, w in run] == list(range(8))
                a_entry += ae
                a_post += ap
                b_rows += br
                b_post += bp
        full = sorted((r, w) for r in range(R) for w in range(8))
        for name, seq in (("a_entry", a_entry), ("a_post", a_post), ("b", b_rows), ("b_post", b_post)):
            assert sorted(seq) == full, (LR, name)
        # Stage B groups: the rows sharing r >> (LR - 6) in one column, 2^(LR-6)
        # of them, on one lane (LR = 10), part of a lane (LR = 7 .. 9, the
        # register bits above LR - 6) or lanes l and l + 32 (LR = 11).
        groups = {}
        for wave, lane, reg in itertools.product(range(NT // 64), range(64), range(16)):
            r = row_b(reg, lane, hi_bits(lane, wave))
            groups.setdefault((r >> (LR - 6), lane & 7), set()).add(r)
            if LR <= 10:
                # in-lane stage: register reg = (g << KB) | t holds row
                # (outer << KB) | t, outer = TwColB's block at layer 6
                kb = LR - 6
                g, t = reg >> kb, reg & ((1 << kb) - 1)
                hi = hi_bits(lane, wave)
                outer = hi if kb >= 4 else (hi << (4 - kb)) | g
                assert r == (outer << kb) | t, (LR, wave, lane, reg)
        assert len(groups) == 64 * 8 and all(len(v) == 1 << (LR - 6) for v in groups.values())
        # Stage A groups: the 64 rows sharing r's low LR - 6 bits, in column w.
        for wave, lane in itertools.product(range(NT // 64), range(64)):
            g = (wave << 4) | (lane & 15)
            rows = {(entry_t(reg, l2) << (LR - 6)) | (g >> 3) for reg in range(16)
                    for l2 in (lane & 15, (lane & 15) | 16, (lane & 15) | 32, (lane & 15) | 48)}
            assert rows == {(t << (LR - 6)) | (g >> 3) for t in range(64)}


# Inverse bounds (narrow set): values below 4p, a butterfly may skip the
# reduction of its sum only when both inputs are below 2p.  The predicates
# below are the ones gs6 / gs5 / gs_lane evaluate while unrolling; the model
# carries each (register, lane) value's bound (in units of p) through the
# butterflies and the lane swaps and checks every lazy choice.
def _gs_butterfly(bnd, ra, rb, lazy):
    a, b = bnd[ra], bnd[rb]
    assert (a <= 4).all() and (b <= 4).all()
    if lazy:
        assert (a <= 2).all() and (b <= 2).all(), (ra, rb)
        bnd[ra] = a + b
    else:
        bnd[ra] = np.full(64, 4)
    bnd[rb] = np.full(64, 2)
    return 1 if lazy else 0


def _gs_post_swap_layers(bnd):
    lazy = 0
    for l in range(3, -1, -1):
        h = 8 >> l
        for sub in range(1 << l):
            for x in range(h):
                lazy += _gs_butterfly(bnd, sub * 2 * h + x, sub * 2 * h + x + h, l < 3 and (x & (h >> 1)) != 0)
    return lazy


def _gs6_bounds(sb, fresh):
    bnd = [np.full(64, 2 if fresh else 4) for _ in range(16)]
    lazy = 0
    for r in range(16):
        if r & (1 << sb):
            continue
        lazy += _gs_butterfly(bnd, r, r | (1 << sb), fresh)
    for r in range(16):
        if r & 8 or r & (1 << sb):
            continue
        for q in (0, 1):
            rr = r | (q << sb)
            lazy += _gs_butterfly(bnd, rr, rr | 8, q == 1)
    bnd = lane_swaps(bnd, sb)
    lazy += _gs_post_swap_layers(bnd)
    return bnd, lazy


def _gs5_bounds(fresh):
    bnd = [np.full(64, 2 if fresh else 4) for _ in range(16)]
    lazy = 0
    for r in range(8):
        lazy += _gs_butterfly(bnd, r, r | 8, fresh)
    bnd = lane_swaps(bnd, with16=False)
    lazy += _gs_post_swap_layers(bnd)
    return bnd, lazy


def _gs_lane_bounds(K, fresh):
    bnd = [np.full(64, 2 if fresh else 4) for _ in range(16)]
    lazy = 0
    for g in range(16 >> K):
        for l in range(K - 1, -1, -1):
            h = (1 << K) >> (l + 1)
            for sub in range(1 << l):
                for x in range(h):
                    ra = (g << K) + sub * 2 * h + x
                    lazy += _gs_butterfly(bnd, ra, ra + h, fresh if l == K - 1 else (x & (h >> 1)) != 0)
    return bnd, lazy


def test_lazy_inverse_choices_hold_their_bounds():
    for sb in (1, 2):
        for fresh in (False, True):
            bnd, lazy = _gs6_bounds(sb, fresh)
            assert all((b <= 4).all() for b in bnd)
            assert lazy == (24 if fresh else 16), (sb, fresh, lazy)
    for fresh in (False, True):
        bnd, lazy = _gs5_bounds(fresh)
        assert all((b <= 4).all() for b in bnd)
        assert lazy == (20 if fresh else 12)
    for K in (1, 2, 3, 4):
        for fresh in (False, True):
            bnd, lazy = _gs_lane_bounds(K, fresh)
            assert all((b <= 4).all() for b in bnd)


def test_lazy_inverse_model_catches_a_wrong_choice():
    """The first layer after a lane swap mixes sums and differences of one
    register across lanes: calling it lazy must fail the bound check."""
    bnd = [np.full(64, 2) for _ in range(16)]
    for r in range(16):
        if r & 2:
            continue
        _gs_butterfly(bnd, r, r | 2, True)
    for r in range(16):
        if r & 8 or r & 2:
            continue
        for q in (0, 1):
            _gs_butterfly(bnd, r | (q << 1), r | (q << 1) | 8, q == 1)
    bnd = lane_swaps(bnd, 1)
    try:
        _gs_butterfly(bnd, 0, 1, True)
    except AssertionError:
        return
    raise AssertionError("a lazy butterfly on mixed bounds went unnoticed")

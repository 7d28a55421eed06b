"""CPU model of the register-resident NTT stages (pi_compute_amd/csrc/ntt_regs.h)
and of the lane layouts of k_row_reg12, k_col_fwd_reg and k_col_inv_reg.

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


def post_t(r, lane):
    i3, i2, i1, i0 = (r >> 3) & 1, (r >> 2) & 1, (r >> 1) & 1, r & 1
    return (((lane >> 5) & 1) << 5) | (i2 << 4) | (((lane >> 4) & 1) << 3) | (i0 << 2) | (i3 << 1) | i1


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


def lane_swaps(v):
    v = [c.copy() for c in v]
    for i in range(8):
        v[i], v[i | 8] = permlane32_swap(v[i], v[i | 8])
    for i in range(16):
        if i & 2:
            continue
        v[i], v[i | 2] = permlane16_swap(v[i], v[i | 2])
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
    w = lane_swaps(v)
    for r in range(16):
        assert list(w[r]) == [post_t(r, l) for l in lanes]
    back = lane_swaps(w)
    for i in range(16):
        assert list(back[i]) == list(v[i])


def ct6_pairs():
    """(register a, register b, layer, block) of every butterfly of ct6, per lane,
    in the post-swap part evaluated with the lane's own bits."""
    out = []
    for l in range(4):
        h = 8 >> l
        for sub in range(1 << l):
            for x in range(h):
                out.append(("entry", sub * 2 * h + x, sub * 2 * h + x + h, l, lambda lane, sub=sub: sub))
    for i2, i0 in itertools.product((0, 1), repeat=2):
        for i1 in (0, 1):
            r = (i2 << 2) | (i1 << 1) | i0
            out.append(("post", r, r | 8, 4,
                        lambda lane, i2=i2, i0=i0: (((lane >> 5) & 1) << 3) | (i2 << 2) | (((lane >> 4) & 1) << 1) | i0))
    for i3, i2, i0 in itertools.product((0, 1), repeat=3):
        r = (i3 << 3) | (i2 << 2) | i0
        out.append(("post", r, r | 2, 5,
                    lambda lane, i3=i3, i2=i2, i0=i0:
                    (((lane >> 5) & 1) << 4) | (i2 << 3) | (((lane >> 4) & 1) << 2) | (i0 << 1) | i3))
    return out


def test_ct6_butterflies_follow_the_textbook_network():
    # In-place CT over a 64-element group: layer m pairs t and t + 2^(5-m)
    # (bit 5-m of t clear) with the twiddle of block t >> (6-m).
    for lane in range(64):
        for kind, ra, rb, m, blk in ct6_pairs():
            f = entry_t if kind == "entry" else post_t
            ta, tb = f(ra, lane), f(rb, lane)
            d = 1 << (5 - m)
            assert ta & d == 0 and tb == ta + d, (lane, ra, rb, m)
            assert blk(lane) == ta >> (6 - m), (lane, ra, m)


def _banks_ok(addrs):
    """addrs: 64 LDS word addresses of one wave instruction."""
    for half in (slice(0, 32), slice(32, 64)):
        if len(set(a % 32 for a in addrs[half])) != 32:
            return False
    return True


def test_row_reg12_layouts():
    L = 4096
    lds = lambda k: k + 2 * (k >> 6)
    seen = {name: [] for name in ("a_entry", "a_post", "b_entry", "b_post", "rho")}
    for wave in range(4):
        for reg in range(16):
            acc = {name: [] for name in seen}
            for lane in range(64):
                j = (wave << 4) | (lane & 15)
                lanepos = ((lane >> 4) << 6) | j
                acc["a_entry"].append((reg << 8) | lanepos)
                acc["a_post"].append((post_t(reg, lane) << 6) | j)
                acc["b_entry"].append((j << 6) | (reg << 2) | (lane >> 4))
                acc["b_post"].append((j << 6) | post_t(reg, lane))
                acc["rho"].append((reg << 8) | lanepos)
            for name in seen:
                seen[name] += acc[name]
            # LDS: stage A post (write/read) and stage B entry (read/write).
            assert _banks_ok([lds(k) for k in acc["a_post"]])
            assert _banks_ok([lds(k) for k in acc["b_entry"]])
            # Global: stage A entry and register-order b rows, 16-word runs.
            for name in ("a_entry", "rho"):
                ks = acc[name]
                for g in range(4):
                    run = ks[16 * g:16 * g + 16]
                    assert run == list(range(run[0], run[0] + 16))
    for name, ks in seen.items():
        assert sorted(ks) == list(range(L)), name
    # Stage A groups are positions j + 64 t; stage B groups 64 u + t'.
    for wave, lane, reg in itertools.product(range(4), range(64), range(16)):
        j = (wave << 4) | (lane & 15)
        assert ((reg << 8) | ((lane >> 4) << 6) | j) == (entry_t(reg, lane) << 6) | j
        assert ((j << 6) | (reg << 2) | (lane >> 4)) == (j << 6) | entry_t(reg, lane)
    assert max(lds(k) for k in range(L)) < L + 2 * (L >> 6)


def _col_model(LR):
    S = 1 if LR == 10 else 2
    NT = 1 << (LR - 1)
    lds = lambda r, w: ((r << 3) | w) + S * (((r << 3) | w) >> 6)

    def hi_bits(lane, wave):
        l3, l4, l5 = (lane >> 3) & 1, (lane >> 4) & 1, (lane >> 5) & 1
        if LR == 10:
            return l5 | ((wave & 1) << 1) | (l3 << 2) | (l4 << 3) | ((wave >> 1) << 4)
        return l5 | (l3 << 1) | (l4 << 2) | (wave << 3)

    return S, NT, lds, hi_bits


def test_col_reg_layouts():
    for LR in (9, 10):
        S, NT, lds, hi_bits = _col_model(LR)
        R = 1 << LR
        a_entry, a_post, b_rows = [], [], []
        for wave in range(NT // 64):
            for reg in range(16):
                ae, ap, br = [], [], []
                for lane in range(64):
                    g = (wave << 4) | (lane & 15)
                    w, rlow = g & 7, g >> 3
                    ae.append(((entry_t(reg, lane) << (LR - 6)) | rlow, w))
                    ap.append(((post_t(reg, lane) << (LR - 6)) | rlow, w))
                    br.append(((hi_bits(lane, wave) << 4) | reg, lane & 7))
                assert _banks_ok([lds(r, w) for r, w in ap]), (LR, "stage A post")
                assert _banks_ok([lds(r, w) for r, w in br]), (LR, "stage B")
                # Global rows: 8 consecutive columns of one row per 8 lanes.
                for seq in (ae, br):
                    for g8 in range(8):
                        run = seq[8 * g8:8 * g8 + 8]
                        assert len({r for r, _ in run}) == 1 and [w for _, w in run] == list(range(8))
                a_entry += ae
                a_post += ap
                b_rows += br
        full = sorted((r, w) for r in range(R) for w in range(8))
        for name, seq in (("a_entry", a_entry), ("a_post", a_post), ("b", b_rows)):
            assert sorted(seq) == full, (LR, name)
        # Stage B groups: a lane's registers are consecutive rows of one column,
        # 16 rows (LR = 10) or two runs of 8 (LR = 9, register bit 3 = row bit 3).
        for wave, lane in itertools.product(range(NT // 64), range(64)):
            rows = [(hi_bits(lane, wave) << 4) | reg for reg in range(16)]
            assert rows == list(range(rows[0], rows[0] + 16))
        # Stage A groups: the 64 rows sharing r's low LR - 6 bits, in column w.
        for wave, lane in itertools.product(range(NT // 64), range(64)):
            g = (wave << 4) | (lane & 15)
            rows = {(entry_t(reg, l2) << (LR - 6)) | (g >> 3) for reg in range(16)
                    for l2 in (lane & 15, (lane & 15) | 16, (lane & 15) | 32, (lane & 15) | 48)}
            assert rows == {(t << (LR - 6)) | (g >> 3) for t in range(64)}

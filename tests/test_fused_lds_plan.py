"""Host emulation of the fused training kernel's LDS plan and index math
(openntc_amd/csrc/fused_body.h), for every layout the library compiles.

- The plan: the BF16 array `smem` (weight tables | X^T | a union region that
  holds X and the staged footprint boxes during the gather, the
  decoder-gradient staging rows during the rounds) and the separate static
  arrays dX (`dxs`), target words, axis tables and tile origins. Its byte
  total, restated from the kernel's constants, must equal the group segment
  size the compiler recorded for each kernel in the built code object, and
  fit the 160 KiB of LDS a workgroup can have.
- The indices: every LDS element each lane of each wave reads or writes
  (weight-table fragments, X rows and X^T rows, staging rows, dX rows,
  staged boxes, target words), over every loop variable, evaluated with the
  kernel's own index expressions and checked against the bounds of the
  region it belongs to; global indices of the partial writes and of the
  decoder update's reduction against their buffers.

No GPU is used. `_plan` and the index expressions restate the kernel's; a
change to a stride there that is not made here fails the size check.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from tests.lds_banks import MIN_CYCLES, cycles

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BODY = os.path.join(ROOT, "openntc_amd", "csrc", "fused_body.h")
OBJ = os.path.join(ROOT, "openntc_amd", "_build", "kernels_fused.hip.o")
LLVM = "/opt/rocm/llvm/bin"
LDS_LIMIT = 160 * 1024
THREADS = 256
CORNERS, BILINEAR = 0, 1


def _layouts():
    """(H, FP, G, C0, C1, G0MODE, PEC, LODS, ACT) of every compiled layout,
    read from the NTC_FUSED_LAYOUTS list."""
    src = open(BODY).read()
    block = src[src.index("#define NTC_FUSED_LAYOUTS"):]
    block = block[:block.index("\n\n")]
    out = []
    for m in re.finditer(r"X2?\((\d+), (\d+), (\d+), (\d+), (\d+), NTC_G0_(\w+), (\d), (\d), NTC_ACT_(\w+)\)", block):
        H, FP, G, C0, C1 = (int(m.group(i)) for i in range(1, 6))
        mode = CORNERS if m.group(6) == "CORNERS" else BILINEAR
        act = 1 if m.group(9) == "RELU" else 0
        out.append((H, FP, G, C0, C1, mode, int(m.group(7)), int(m.group(8)), act))
    assert len(out) >= 9
    return out


def _nt(H):
    return 1 if H >= 128 else 2


def _plan(H, FP, G, C0, C1, mode, NT):
    """The kernel's Cfg constants and LDS byte total."""
    PASS = 64 * NT
    BH = PASS // 16
    WT = H * FP + H * H + 16 * H
    XS = FP + 8
    SH = H + 16
    STAGE = 4 * 64 * SH + 64 * 16
    WMAX = 18
    Q0, Q1 = (C0 + 3) // 4, (C1 + 3) // 4
    CP0, CP1 = 4 * Q0, 4 * Q1
    G1B = 4 * C0 if mode == CORNERS else C0
    GSLOTS = G1B + C1
    GS_BANK = G + (16 - G) % 64

    def lds_bytes(sw, sv, gs, stage=STAGE):
        u = 2 * sw * sv * (CP0 + CP1)
        return 2 * (WT + PASS * XS + max(u, stage)) + 4 * PASS * gs + 4 * 192 + 4 * 128 + 16 * PASS

    cands = [(WMAX, BH + 3, GS_BANK), (WMAX, BH + 3, G + 4), (10, BH // 2 + 3, GS_BANK), (10, BH // 2 + 3, G + 4)]
    sel = next((i for i, c in enumerate(cands) if lds_bytes(*c) <= LDS_LIMIT), 4)
    assert sel < 4, "layout does not fit"
    SW, SV, GS = cands[sel]
    UNION = max(2 * SW * SV * (CP0 + CP1), STAGE)
    return dict(PASS=PASS, BH=BH, WT=WT, XS=XS, UNION=UNION, SW=SW, SV=SV, GS=GS, CP0=CP0, CP1=CP1,
                Q0=Q0, Q1=Q1, G1B=G1B, GSLOTS=GSLOTS, MT=H // 16, KF=FP // 32, KH=H // 32, GT=G // 16,
                RW=H // 4, MTW=H // 64, NTO=H // 64, SBOX0=SW * SV * CP0,
                DXS=PASS * GS, SH=SH, STAGE=STAGE, total=lds_bytes(SW, SV, GS))


def swz(row, col, wd):
    nc = wd >> 3
    if nc & (nc - 1):
        return col
    r = row & 15
    if nc >= 16:
        key = (r & 1) | ((r >> 2) & 2) | ((r << 1) & 4) | ((r << 1) & 8)
        return (((col >> 3) ^ key) << 3) | (col & 7)
    if nc == 8:
        key = (r & 2) | ((((r >> 2) ^ (r >> 3)) & 1) << 2)
        return (((col >> 3) ^ key) << 3) | (col & 7)
    per = 16 // nc
    return (((col >> 3) ^ ((row // per) & (nc - 1))) << 3) | (col & 7)


def kperm(j, g, e):
    return 32 * j + (4 * g + e if e < 4 else 16 + 4 * g + (e - 4))


def kperm_pos(pos):
    j, r = pos >> 5, pos & 31
    return kperm(j, r >> 3, r & 7)


def kperm_col(k):
    r = k & 31
    q = r & 15
    return (k & ~31) + 8 * (q >> 2) + (q & 3) + 4 * (r >> 4)


def tr_row(j, h, g, i):
    return 32 * j + 16 * h + 4 * g + (i >> 2)


def tr_read(table, addrs):
    """ds_read_b64_tr_b16 over 64 lanes: within each 16-lane group, lane i
    receives element (i & 3) of the 8 bytes lane 4k + (i >> 2) addresses, for
    k = 0..3 (the instruction's lane mapping, established on the GPU with a
    probe that read a table of distinct values)."""
    out = np.zeros((64, 4), table.dtype)
    for lane in range(64):
        base, i = lane & ~15, lane & 15
        for k in range(4):
            src = base + 4 * k + (i >> 2)
            out[lane, k] = table[addrs[src] + (i & 3)]
    return out


class Region:
    """Element bounds of one LDS array (or sub-range); records every index."""

    def __init__(self, name, size):
        self.name, self.size, self.lo, self.hi = name, size, 1 << 60, -1

    def touch(self, idx, width=1):
        idx = np.asarray(idx, np.int64)
        self.lo = min(self.lo, int(idx.min()))
        self.hi = max(self.hi, int(idx.max()) + width - 1)

    def check(self):
        assert self.hi >= 0, f"{self.name}: never touched"
        assert 0 <= self.lo and self.hi < self.size, f"{self.name}: [{self.lo}, {self.hi}] outside [0, {self.size})"


def _emulate(H, FP, G, C0, C1, mode, PEC, LODS, NT):
    P = _plan(H, FP, G, C0, C1, mode, NT)
    PASS, XS = P["PASS"], P["XS"]
    lanes = np.arange(64)
    n16, g = lanes & 15, lanes >> 4
    wt = Region("weight tables", P["WT"])
    xnf = Region("X rows", PASS * XS)
    sbox = Region("staged boxes (floats)", P["UNION"] // 2)
    SH = P["SH"]
    stA, stB, stC, stD = (Region(n, 64 * SH) for n in ("stA", "stB", "stC", "stD"))
    stY = Region("stY", 64 * 16)
    dxs = Region("dX", P["DXS"])
    tgt = Region("target words", PASS * 4)
    # the union holds X + boxes, then stA | stB | stY
    assert 2 * P["SW"] * P["SV"] * (P["CP0"] + P["CP1"]) <= P["UNION"]
    assert P["STAGE"] <= P["UNION"]
    oW0, oW1 = 0, H * FP
    oWo = oW1 + H * H
    assert oWo + 16 * H == P["WT"]
    vswz = np.vectorize(swz)
    vkcol = np.vectorize(kperm_col)
    for wave in range(4):
        for nt in range(NT):
            tloc = wave * 16 * NT + nt * 16 + n16
            for j in range(P["KF"]):  # layer 0: X rows and W0
                xnf.touch(tloc * XS + 32 * j + 8 * g, 8)
                for r in range(P["MT"]):
                    row = 16 * r + n16
                    wt.touch(oW0 + row * FP + vswz(row, 32 * j + 8 * g, FP), 8)
            for j in range(P["KH"]):
                for r in range(P["MT"]):
                    row = 16 * r + n16
                    wt.touch(oW1 + row * H + vswz(row, 32 * j + 8 * g, H), 8)
                    # W1^T by transposing reads of W1 (8 bytes a lane)
                    c = vkcol(16 * r + 4 * (n16 & 3))
                    for h in range(2):
                        tr = tr_row(j, h, g, n16)
                        wt.touch(oW1 + tr * H + vswz(tr, c, H), 4)
                wt.touch(oWo + n16 * H + vswz(n16, 32 * j + 8 * g, H), 8)
                for q in range(P["GT"]):  # W0^T by transposing reads of W0
                    c = 16 * q + 4 * (n16 & 3)
                    for h in range(2):
                        tr = tr_row(j, h, g, n16)
                        wt.touch(oW0 + tr * FP + vswz(tr, c, FP), 4)
            for r in range(P["MT"]):  # Wout^T by transposing reads of Wout
                o = 4 * g + (n16 >> 2)
                wt.touch(oWo + o * H + vswz(o, vkcol(16 * r + 4 * (n16 & 3)), H), 4)
            for r in range(P["GT"]):  # dX rows: 16-byte or single writes of slots < GSLOTS
                s0 = 16 * r + 4 * g
                ok = s0 < P["GSLOTS"]
                if ok.any():
                    dxs.touch((tloc * P["GS"] + s0)[ok], 4)
            tgt.touch(tloc * 4 + g)
        # decoder-gradient rounds: texel-row stores, transposing reads
        row = wave * 16 + n16
        for j in range(P["KH"]):
            for st in (stA, stB, stC, stD):
                st.touch(row * SH + 32 * j + 8 * g, 8)
        stY.touch(row * 16 + 4 * g, 4)
        trr0, tq = 8 * g + (n16 >> 2), 4 * (n16 & 3)

        def frag_st(reg, stride, kc, c0):
            for h in range(2):
                reg.touch((32 * kc + trr0 + 4 * h) * stride + c0 + tq, 4)
        for kc in range(2):
            for mt in range(P["MTW"]):
                frag_st(stB, SH, kc, P["RW"] * wave + 16 * mt)  # dW1 A operand
                frag_st(stC, SH, kc, P["RW"] * wave + 16 * mt)  # dW0 A operand
            for c in range(P["MT"]):
                frag_st(stA, SH, kc, 16 * c)
            frag_st(stY, 16, kc, 0)
            for c in range(P["NTO"]):
                frag_st(stD, SH, kc, (wave * P["NTO"] + c) * 16)
        for rnd in range(NT):  # X^T by transposing reads of the X rows
            for kc in range(2):
                for h in range(2):
                    p = 32 * kc + 8 * g + 4 * h + (n16 >> 2)
                    trow = (p >> 4) * 16 * NT + 16 * rnd + (p & 15)
                    for c in range(FP // 16):
                        xnf.touch(trow * XS + 16 * c + 4 * (n16 & 3), 4)
    # staged gather: put4 (8-byte X-row writes) and single puts for every
    # slot < FP of every texel; box reads for boxes up to SW x SV cells
    t = np.arange(PASS)
    for slot0 in range(0, FP, 4):
        xnf.touch(t * XS + slot0, 4)
    for slot in range(FP):
        xnf.touch(t * XS + slot)
    for g1, (Q, CP) in enumerate(((P["Q0"], P["CP0"]), (P["Q1"], P["CP1"]))):
        base = P["SBOX0"] if g1 else 0
        bw, bh = P["SW"], P["SV"]
        ay, ax = np.meshgrid(np.arange(bh), np.arange(bw), indexing="ij")
        for q in range(Q):
            sbox.touch(base + 4 * q + (ay * bw + ax) * CP, 4)  # float4 reads
            cb = np.arange(bw * bh)
            sbox.touch(base + cb * CP + 4 * q, 4)  # prefetch_commit writes
    # box gradients: dX reads of slot0 + c*C + ch, ch < C, texel 4kk + q
    for slot0, C, nc in ((0, C0, 4 if mode == CORNERS else 1), (P["G1B"], C1, 1)):
        tex = np.arange(PASS)
        for c in range(nc):
            for ch in range(C):
                dxs.touch(tex * P["GS"] + slot0 + c * C + ch)
    for r in (wt, xnf, sbox, stA, stB, stC, stD, stY, dxs, tgt):
        r.check()
    # every dX slot stays inside its row (no row reaches into the next texel's)
    assert P["GSLOTS"] <= G <= P["GS"]
    return P


@pytest.mark.parametrize("lay", _layouts(), ids=lambda l: "H%d-F%d-G%d-c%d-%d-m%d-a%d" % (l[0], l[1], l[2], l[3], l[4], l[5], l[8]))
def test_fused_lds_indices_in_bounds(lay):
    H, FP, G, C0, C1, mode, PEC, LODS, act = lay
    P = _emulate(H, FP, G, C0, C1, mode, PEC, LODS, _nt(H))
    assert P["total"] <= LDS_LIMIT


def _code_object_lds():
    """{mangled kernel name: group segment bytes} of the built fused kernels."""
    if not os.path.exists(OBJ) or not os.path.exists(os.path.join(LLVM, "llvm-readelf")):
        pytest.skip("library not built here (openntc_amd/_build) or no LLVM tools")
    with tempfile.TemporaryDirectory() as d:
        fat, co, obj = os.path.join(d, "fat.bin"), os.path.join(d, "k.co"), os.path.join(d, "k.o")
        # objcopy rewrites its input: work on a copy, so the build's object
        # keeps its time stamp (an incremental build compares it with sources)
        shutil.copyfile(OBJ, obj)
        subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section", f".hip_fatbin={fat}", obj],
                       check=True, capture_output=True)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={fat}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True,
                       capture_output=True)
        notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", co], check=True,
                               capture_output=True, text=True).stdout
    out, size = {}, None
    for ln in notes.splitlines():
        m = re.match(r"\s*\.group_segment_fixed_size:\s+(\d+)", ln)
        if m:
            size = int(m.group(1))
        m = re.match(r"\s*\.name:\s+(_Z\S+)", ln)
        if m and size is not None:
            out[m.group(1)] = size
            size = None
    return out


def test_lds_plan_matches_code_object():
    sizes = _code_object_lds()
    seen = 0
    for name, size in sizes.items():
        m = re.search(r"k_fused_train(?:_sets)?I((?:Li-?\d+E)+)E", name)
        if not m:
            continue
        H, FP, G, C0, C1, mode, PEC, LODS, act, NT = (int(v) for v in re.findall(r"Li(-?\d+)E", m.group(1)))
        P = _plan(H, FP, G, C0, C1, mode, NT)
        assert size == P["total"], (name, size, P["total"])
        assert size <= LDS_LIMIT
        seen += 1
    assert seen >= 2 * len(_layouts()) - 4  # k_fused_train and k_fused_train_sets instances


def _reduction_reads(blocks, n_dec):
    """Highest partial index k_decoder_update reads (blocks of 32 loads, then
    the rest one by one)."""
    hi = -1
    b = 0
    while b + 32 <= blocks:
        hi = max(hi, (b + 31) * n_dec + n_dec - 1)
        b += 32
    while b < blocks:
        hi = max(hi, b * n_dec + n_dec - 1)
        b += 1
    return hi


@pytest.mark.parametrize("blocks", [1, 31, 32, 33, 64, 255, 256, 257, 512])
def test_decoder_reduction_reads_in_bounds(blocks):
    H, F = 128, 40
    n_dec = H * F + H * H + 16 * H
    assert _reduction_reads(blocks, n_dec) == blocks * n_dec - 1


@pytest.mark.parametrize("H,FP", [(128, 64), (64, 96), (64, 32)])
def test_transposing_reads_form_transposed_fragments(H, FP):
    """The backward's A fragments of W1^T, W0^T and Wout^T, assembled from the
    forward tables (swizzled columns, kperm column order of W1 and Wout) with
    the transposing reads' lane mapping, equal the fragments of the plain
    transposed matrices in the K order of the B operands (kperm: the hidden
    unit at K position p is kperm_pos(p); outputs 4g..4g+3 then zeros)."""
    rng = np.random.default_rng(H + FP)
    G = 32
    w0 = rng.integers(1, 60000, (H, FP)).astype(np.int64)
    w1 = rng.integers(1, 60000, (H, H)).astype(np.int64)
    wo = rng.integers(1, 60000, (16, H)).astype(np.int64)
    # the forward tables as the pack writes them
    t0 = np.zeros(H * FP, np.int64)
    t1 = np.zeros(H * H, np.int64)
    to = np.zeros(16 * H, np.int64)
    for m in range(H):
        for c in range(FP):
            t0[m * FP + swz(m, c, FP)] = w0[m, c]
        for k in range(H):
            t1[m * H + swz(m, kperm_col(k), H)] = w1[m, k]
    for o in range(16):
        for h in range(H):
            to[o * H + swz(o, kperm_col(h), H)] = wo[o, h]
    lanes = np.arange(64)
    n16, g = lanes & 15, lanes >> 4
    for j in range(H // 32):
        for r in range(H // 16):
            c = np.array([kperm_col(16 * r + 4 * (i & 3)) for i in n16])
            parts = []
            for h in range(2):
                tr = tr_row(j, h, g, n16)
                parts.append(tr_read(t1, tr * H + np.array([swz(a, b, H) for a, b in zip(tr, c)])))
            frag = np.concatenate(parts, axis=1)
            for lane in range(64):
                want = [w1[kperm_pos(32 * j + 8 * g[lane] + e), 16 * r + n16[lane]] for e in range(8)]
                assert list(frag[lane]) == want, ("W1^T", j, r, lane)
        for q in range(G // 16):
            c = 16 * q + 4 * (n16 & 3)
            parts = []
            for h in range(2):
                tr = tr_row(j, h, g, n16)
                parts.append(tr_read(t0, tr * FP + np.array([swz(a, b, FP) for a, b in zip(tr, c)])))
            frag = np.concatenate(parts, axis=1)
            for lane in range(64):
                want = [w0[kperm_pos(32 * j + 8 * g[lane] + e), 16 * q + n16[lane]] for e in range(8)]
                assert list(frag[lane]) == want, ("W0^T", j, q, lane)
    for r in range(H // 16):
        o = 4 * g + (n16 >> 2)
        c = np.array([kperm_col(16 * r + 4 * (i & 3)) for i in n16])
        frag = tr_read(to, o * H + np.array([swz(a, b, H) for a, b in zip(o, c)]))
        for lane in range(64):
            want = [wo[4 * g[lane] + e, 16 * r + n16[lane]] for e in range(4)]
            assert list(frag[lane]) == want, ("Wout^T", r, lane)


def _banks(byte_addrs, width):
    """Distinct 4-byte LDS banks (of 64) touched by one access of `width`
    bytes per lane at each address."""
    out = []
    for a in byte_addrs:
        out += [((a + 4 * w) // 4) % 64 for w in range(width // 4)]
    return out


@pytest.mark.parametrize("j,r", [(j, r) for j in range(4) for r in range(8)])
def test_weight_table_reads_take_fewest_lds_cycles(j, r):
    """gfx950 LDS cycles (tests/lds_banks.py) of the weight-table reads of one
    fragment: the forward 16-byte reads of W1 (H = 128) and W0 (FP = 64)
    take 4 cycles, free of bank conflicts; the transposing reads of W0^T
    take 2; those of W1^T and Wout^T take 4, the least for reads whose 32
    lanes all take the same 8-byte half of their 16-byte chunks (the kperm
    column order), i.e. each bank holds two dwords and no more."""
    H, FP = 128, 64
    lanes = np.arange(64)
    n16, g = lanes & 15, lanes >> 4
    fwd1 = [2 * ((16 * r + i) * H + swz(16 * r + i, 32 * j + 8 * gg, H)) for i, gg in zip(n16, g)]
    assert cycles("read16", fwd1) == MIN_CYCLES["read16"]
    if j < FP // 32:
        fwd0 = [2 * ((16 * r + i) * FP + swz(16 * r + i, 32 * j + 8 * gg, FP)) for i, gg in zip(n16, g)]
        assert cycles("read16", fwd0) == MIN_CYCLES["read16"]
    for h in range(2):
        t1 = []
        for i, gg in zip(n16, g):
            row, col = tr_row(j, h, gg, i), kperm_col(16 * r + 4 * (i & 3))
            t1.append(2 * (row * H + swz(row, col, H)))
        assert cycles("read8", t1) == 2 * MIN_CYCLES["read8"]
        if r < 2:
            t0 = []
            for i, gg in zip(n16, g):
                row, col = tr_row(j, h, gg, i), 16 * r + 4 * (i & 3)
                t0.append(2 * (row * FP + swz(row, col, FP)))
            assert cycles("read8", t0) == MIN_CYCLES["read8"]
    to = []
    for i, gg in zip(n16, g):
        o, col = 4 * gg + (i >> 2), kperm_col(16 * r + 4 * (i & 3))
        to.append(2 * (16 * 0 + o * H + swz(o, col, H)))
    assert cycles("read8", to) == 2 * MIN_CYCLES["read8"]


@pytest.mark.parametrize("FP,NT", [(64, 1), (32, 2), (64, 2), (96, 2)])
def test_x_transposing_reads_form_xt_fragments(FP, NT):
    """The dW0 B fragments (K = the round's 32 texels of a K chunk, N = 16
    feature slots) read from the X rows with transposing reads equal X^T in
    the staged texel order, for every round, K chunk and slot tile; within a
    16-lane group the reads are bank-conflict free for the 64-slot rows."""
    PASS, XS = 64 * NT, FP + 8
    X = np.arange(PASS * XS, dtype=np.int64)  # X[t][f] = t * XS + f
    lanes = np.arange(64)
    n16, g = lanes & 15, lanes >> 4
    for rnd in range(NT):
        for kc in range(2):
            for c in range(FP // 16):
                parts = []
                for h in range(2):
                    p = 32 * kc + 8 * g + 4 * h + (n16 >> 2)
                    trow = (p >> 4) * 16 * NT + 16 * rnd + (p & 15)
                    addrs = trow * XS + 16 * c + 4 * (n16 & 3)
                    parts.append(tr_read(X, addrs))
                    if FP == 64:
                        for g0 in range(4):
                            b = _banks([2 * a for a in addrs[16 * g0:16 * g0 + 16]], 8)
                            assert len(set(b)) == 32
                frag = np.concatenate(parts, axis=1)
                for lane in range(64):
                    want = []
                    for e in range(8):
                        p = 32 * kc + 8 * g[lane] + e
                        t = (p >> 4) * 16 * NT + 16 * rnd + (p & 15)
                        want.append(t * XS + 16 * c + n16[lane])
                    assert list(frag[lane]) == want, (rnd, kc, c, lane)


@pytest.mark.parametrize("H", [64, 128])
def test_texel_row_staging_fragments_and_banks(H):
    """The decoder-gradient operands read back from the texel-row staging
    with transposing reads are the fragments of the [unit][texel] matrices
    (K = the round's texels), and each 16-lane group's reads hit 32 distinct
    banks (rows of H + 16 elements; dY rows of 16)."""
    SH = H + 16
    S = np.arange(64 * SH, dtype=np.int64)  # S[t][p] = t * SH + p
    lanes = np.arange(64)
    n16, g = lanes & 15, lanes >> 4
    trr0, tq = 8 * g + (n16 >> 2), 4 * (n16 & 3)
    for stride in (SH, 16):
        for kc in range(2):
            for c0 in range(0, (SH - 16 if stride == SH else 16), 16):
                parts = []
                for h in range(2):
                    addrs = (32 * kc + trr0 + 4 * h) * stride + c0 + tq
                    parts.append(tr_read(S if stride == SH else np.arange(64 * 16), addrs))
                    for g0 in range(4):
                        assert len(set(_banks([2 * a for a in addrs[16 * g0:16 * g0 + 16]], 8))) == 32
                frag = np.concatenate(parts, axis=1)
                for lane in range(64):
                    want = [(32 * kc + 8 * g[lane] + e) * stride + c0 + n16[lane] for e in range(8)]
                    assert list(frag[lane]) == want

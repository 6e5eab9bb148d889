"""Host emulation of the three-launch training step's LDS plans and index
math (openntc_amd/csrc/mlp_body.h, kernels_split.hip), for the compiled
layout (H = 128, FP = 64, G = 32, 16 + 16 bilinear grid channels, two
positional-encoding periods, ReLU).

- The plans: the MLP launch's BF16 array (weight tables | X rows | the
  decoder-gradient staging rows) with the tile origins and target words; the
  gather's staged boxes, axis tables and tile origins; the scatter's dX rows,
  axis tables and tile origins. Each byte total, restated from the kernels'
  constants, must equal the group segment size the compiler recorded for
  the kernel in the built code object, and fit the 160 KiB of LDS.
- The indices: every LDS element each lane of each of the 8 MLP waves reads
  or writes (weight-table fragments and their transposing reads, X rows and
  X^T rows, staging rows in both rounds, dY rows, target words, the
  feature-row commit), every element the gather and the scatter touch in
  their LDS arrays, and the global rows (feature rows, dX rows, partials)
  against their buffers.

No GPU is used.
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
from tests.test_fused_lds_plan import LLVM, Region, kperm_col, kperm_pos, swz, tr_read, tr_row

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "openntc_amd", "_build", "kernels_split.hip.o")
LDS_LIMIT = 160 * 1024
PASS = 128
W = 8                      # MLP waves per workgroup
H, FP, G, C0, C1, PEC, LODS = 128, 64, 32, 16, 16, 2, 0
F = C0 + C1 + 4 * PEC + LODS
FXS = (F + 7) // 8 * 8


def _mlp_plan():
    MT, KF, KH, GT = H // 16, FP // 32, H // 32, G // 16
    RW = H // W
    T1 = MT * MT // W
    BN1 = 4 if T1 >= 16 else (2 if T1 >= 4 else 1)
    BM1 = T1 // BN1
    WT = H * FP + H * H + 16 * H
    XS = FP if FP == 64 else FP + 8
    SH = H if H >= 128 and H & (H - 1) == 0 else H + 16
    STAGE = 4 * 64 * SH + 64 * 16
    smem = WT + PASS * XS + STAGE
    total = 2 * smem + 4 * 128 + 4 * PASS * 4
    return dict(MT=MT, KF=KF, KH=KH, GT=GT, RW=RW, MTW=RW // 16, NTO=H // (16 * W), BM1=BM1, BN1=BN1,
                NCW=(F + 15) // 16, WT=WT, XS=XS, SH=SH, STAGE=STAGE, total=total)


def x_col(t, col):
    """Column position in X row t (128-byte rows, chunk XOR key)."""
    key = (t & 2) | ((((t >> 2) ^ (t >> 3)) & 1) << 2)
    return col ^ (key << 3)


def st_col(r, col):
    """Column position in staging row r (256-byte rows, chunk XOR key)."""
    return col ^ ((((r & 3) << 1) | (r & 8)) << 3)


def y_slot(r):
    return r ^ ((r >> 1) & 4)


def _geo():
    BH = PASS // 16
    SW, SV = 18, BH + 3
    Q0, Q1 = (C0 + 3) // 4, (C1 + 3) // 4
    CP0, CP1 = 4 * Q0, 4 * Q1
    GS = G + (16 - G) % 32
    return dict(BH=BH, SW=SW, SV=SV, Q0=Q0, Q1=Q1, CP0=CP0, CP1=CP1, SBOX0=SW * SV * CP0,
                SBOXF=SW * SV * (CP0 + CP1), GS=GS, G1B=C0, GSLOTS=C0 + C1)


def test_scatter_row_stride_spreads_banks():
    """dX rows of GS floats with GS = 16 or 48 mod 64 (the smallest such >= G):
    the four texel rows of one FP32 MFMA B-operand read (16 consecutive
    slots each) land on 64 distinct banks."""
    gs = _geo()["GS"]
    assert gs >= G and gs % 64 in (16, 48)
    banks = {(t * gs + s) % 64 for t in range(4) for s in range(16)}
    assert len(banks) == 64


def test_mlp_lds_indices_in_bounds():
    P = _mlp_plan()
    XS, SH = P["XS"], P["SH"]
    lanes = np.arange(64)
    n16, g = lanes & 15, lanes >> 4
    wt = Region("weight tables", P["WT"])
    xnf = Region("X rows", PASS * XS)
    stA, stB, stC, stD = (Region(n, 64 * SH) for n in ("stA", "stB", "stC", "stD"))
    stY = Region("stY", 64 * 16)
    tgt = Region("target words", PASS * 4)
    oW0, oW1 = 0, H * FP
    oWo = oW1 + H * H
    vswz = np.vectorize(swz)
    vkcol = np.vectorize(kperm_col)
    assert (H // W) % 16 == 0 and P["MTW"] == 1 and P["NTO"] == 1
    assert (P["MT"] // P["BM1"]) * (P["MT"] // P["BN1"]) == W
    for wave in range(W):
        tloc = wave * 16 + n16
        for j in range(P["KF"]):
            xnf.touch(tloc * XS + x_col(tloc, 32 * j + 8 * g), 8)
            for r in range(P["MT"]):
                row = 16 * r + n16
                wt.touch(oW0 + row * FP + vswz(row, 32 * j + 8 * g, FP), 8)
        for j in range(P["KH"]):
            for r in range(P["MT"]):
                row = 16 * r + n16
                wt.touch(oW1 + row * H + vswz(row, 32 * j + 8 * g, H), 8)
                c = vkcol(16 * r + 4 * (n16 & 3))
                for h in range(2):
                    tr = tr_row(j, h, g, n16)
                    wt.touch(oW1 + tr * H + vswz(tr, c, H), 4)
            wt.touch(oWo + n16 * H + vswz(n16, 32 * j + 8 * g, H), 8)
            for q in range(P["GT"]):
                c = 16 * q + 4 * (n16 & 3)
                for h in range(2):
                    tr = tr_row(j, h, g, n16)
                    wt.touch(oW0 + tr * FP + vswz(tr, c, FP), 4)
        for r in range(P["MT"]):
            o = 4 * g + (n16 >> 2)
            wt.touch(oWo + o * H + vswz(o, vkcol(16 * r + 4 * (n16 & 3)), H), 4)
        tgt.touch(tloc * 4 + g)
        trr0, tq = 8 * g + (n16 >> 2), 4 * (n16 & 3)

        def frag_st(reg, kc, c0):
            for h in range(2):
                r = 32 * kc + trr0 + 4 * h
                reg.touch(r * SH + st_col(r, c0 + tq), 4)
        mb1, nb1 = wave // (P["MT"] // P["BN1"]), wave % (P["MT"] // P["BN1"])
        for rnd in range(PASS // 64):
            if wave >> 2 == rnd:  # this wave's tile staged in round rnd
                row = (wave & 3) * 16 + n16
                for j in range(P["KH"]):
                    for st in (stA, stB, stC, stD):
                        st.touch(row * SH + st_col(row, 32 * j + 8 * g), 8)
                stY.touch(y_slot(row) * 16 + 4 * g, 4)
            for kc in range(2):
                for mt in range(P["BM1"]):
                    frag_st(stB, kc, 16 * (P["BM1"] * mb1 + mt))
                for c in range(P["BN1"]):
                    frag_st(stA, kc, 16 * (P["BN1"] * nb1 + c))
                for mt in range(P["MTW"]):
                    frag_st(stC, kc, P["RW"] * wave + 16 * mt)
                for h in range(2):
                    p = 32 * kc + 8 * g + 4 * h + (n16 >> 2)
                    t = 64 * rnd + p
                    for c in range(P["NCW"]):
                        xnf.touch(t * XS + x_col(t, 16 * c + 4 * (n16 & 3)), 4)
                for h in range(2):
                    stY.touch(y_slot(32 * kc + trr0 + 4 * h) * 16 + tq, 4)
                for c in range(P["NTO"]):
                    frag_st(stD, kc, (wave * P["NTO"] + c) * 16)
    # feature-row commit: chunk c of PASS * FXS / 8 -> row c / xchunks, 8 q
    xch = FXS // 8
    c = np.arange(PASS * xch)
    xnf.touch((c // xch) * XS + x_col(c // xch, 8 * (c % xch)), 8)
    for r in (wt, xnf, stA, stB, stC, stD, stY, tgt):
        r.check()
    assert P["total"] <= LDS_LIMIT
    # the SSE reduction reuses the staging rows as 8 * 64 doubles
    assert 8 * 64 * W <= 2 * P["STAGE"]


def _x_rows():
    """X rows as the MLP launch lays them out: logical X[t][f] = 1000 t + f + 1
    for the FXS stored features, zero padding up to FP, at x_col positions."""
    XS = _mlp_plan()["XS"]
    phys = np.full(PASS * XS, -1, np.int64)
    xch = FXS // 8
    for c in range(PASS * xch):
        t, q = divmod(c, xch)
        for e in range(8):
            phys[t * XS + x_col(t, 8 * q) + e] = 1000 * t + 8 * q + e + 1
    for i in range(PASS * XS):
        t = i // XS
        if x_col(t, i - t * XS) >= FXS:
            phys[i] = 0
    return phys


def _want_x(t, f):
    return 1000 * t + f + 1 if f < FXS else 0


def test_mlp_swizzled_rows_form_the_operand_fragments():
    """With the chunk-permuted X rows, staging rows and dY slots, every
    operand fragment the MLP launch reads equals the logical one: layer-0 B
    (X rows), the dW0 X^T fragments (transposing reads), the dW1 / dW0 /
    dWout operands from the staging rows and the dY rows."""
    P = _mlp_plan()
    XS, SH = P["XS"], P["SH"]
    assert XS == 64 and SH == H
    phys = _x_rows()
    assert (phys >= 0).all()  # every element written or zeroed
    lanes = np.arange(64)
    n16, g = lanes & 15, lanes >> 4
    for wave in range(W):
        tloc = wave * 16 + n16
        for j in range(P["KF"]):
            for lane in range(64):
                t, c0 = tloc[lane], 32 * j + 8 * g[lane]
                got = phys[t * XS + x_col(t, c0): t * XS + x_col(t, c0) + 8]
                assert list(got) == [_want_x(t, c0 + e) for e in range(8)]
    for rnd in range(PASS // 64):
        for kc in range(2):
            for c in range(P["NCW"]):
                parts = []
                for h in range(2):
                    pp = 32 * kc + 8 * g + 4 * h + (n16 >> 2)
                    t = 64 * rnd + pp
                    parts.append(tr_read(phys, t * XS + np.array([x_col(a, 16 * c + 4 * (b & 3))
                                                                  for a, b in zip(t, n16)])))
                frag = np.concatenate(parts, axis=1)
                for lane in range(64):
                    want = [_want_x(64 * rnd + 32 * kc + 8 * g[lane] + e, 16 * c + n16[lane]) for e in range(8)]
                    assert list(frag[lane]) == want, (rnd, kc, c, lane)
    # staging: logical S[row][unit] = 1000 row + unit + 1 written by the waves
    st = np.full(64 * SH, -1, np.int64)
    for row in range(64):
        for j in range(P["KH"]):
            for gg in range(4):
                for e in range(8):
                    st[row * SH + st_col(row, 32 * j + 8 * gg) + e] = 1000 * row + 32 * j + 8 * gg + e + 1
    assert (st >= 0).all()
    trr0, tq = 8 * g + (n16 >> 2), 4 * (n16 & 3)
    for kc in range(2):
        for c0 in range(0, H, 16):
            r0 = 32 * kc + trr0
            a0 = r0 * SH + np.array([st_col(r, c0 + q) for r, q in zip(r0, tq)])
            a1 = (r0 + 4) * SH + np.array([st_col(r + 4, c0 + q) for r, q in zip(r0, tq)])
            frag = np.concatenate([tr_read(st, a0), tr_read(st, a1)], axis=1)
            for lane in range(64):
                want = [1000 * (32 * kc + 8 * g[lane] + e) + c0 + n16[lane] + 1 for e in range(8)]
                assert list(frag[lane]) == want
    y = np.full(64 * 16, -1, np.int64)
    for row in range(64):
        for gg in range(4):
            for e in range(4):
                y[y_slot(row) * 16 + 4 * gg + e] = 100 * row + 4 * gg + e + 1
    assert (y >= 0).all()
    for kc in range(2):
        r0 = 32 * kc + trr0
        frag = np.concatenate([tr_read(y, np.array([y_slot(r) for r in r0]) * 16 + tq),
                               tr_read(y, np.array([y_slot(r + 4) for r in r0]) * 16 + tq)], axis=1)
        for lane in range(64):
            assert list(frag[lane]) == [100 * (32 * kc + 8 * g[lane] + e) + n16[lane] + 1 for e in range(8)]


def test_mlp_lds_accesses_take_fewest_cycles():
    """gfx950 LDS cycles (tests/lds_banks.py) of every per-wave access of the
    MLP launch: X-row reads (layer-0 B) 4, X^T transposing reads 2, staging
    writes 8, staging and dY transposing reads 2 — no bank conflicts. (The
    weight-table reads are checked in test_fused_lds_plan.py.)"""
    P = _mlp_plan()
    XS, SH = P["XS"], P["SH"]
    lanes = np.arange(64)
    n16, g = lanes & 15, lanes >> 4
    trr0, tq = 8 * g + (n16 >> 2), 4 * (n16 & 3)
    for wave in range(W):
        tloc = wave * 16 + n16
        for j in range(P["KF"]):
            a = [2 * (t * XS + x_col(t, 32 * j + 8 * gg)) for t, gg in zip(tloc, g)]
            assert cycles("read16", a) == MIN_CYCLES["read16"]
        row = (wave & 3) * 16 + n16
        for j in range(P["KH"]):
            a = [2 * (r * SH + st_col(r, 32 * j + 8 * gg)) for r, gg in zip(row, g)]
            assert cycles("write16", a) == MIN_CYCLES["write16"]
    for rnd in range(PASS // 64):
        for kc in range(2):
            for h in range(2):
                pp = 32 * kc + 8 * g + 4 * h + (n16 >> 2)
                t = 64 * rnd + pp
                for c in range(P["NCW"]):
                    a = [2 * (tt * XS + x_col(tt, 16 * c + 4 * (i & 3))) for tt, i in zip(t, n16)]
                    assert cycles("read8", a) == MIN_CYCLES["read8"], (rnd, kc, h, c)
                r = 32 * kc + trr0 + 4 * h
                for c0 in range(0, H, 16):
                    a = [2 * (rr * SH + st_col(rr, c0 + q)) for rr, q in zip(r, tq)]
                    assert cycles("read8", a) == MIN_CYCLES["read8"], (kc, h, c0)
                a = [2 * (y_slot(rr) * 16 + q) for rr, q in zip(r, tq)]
                assert cycles("read8", a) == MIN_CYCLES["read8"]


def test_mlp_global_rows_in_bounds():
    """Feature-row loads, dX-row stores and partial writes of every workgroup
    stay inside their buffers for batches of several sizes and grids."""
    P = _mlp_plan()
    n_dec = H * F + H * H + 16 * H
    for B, blocks in ((524288, 256), (4096, 32), (12, 1), (200, 2)):
        npass = -(-B // PASS)
        xch = FXS // 8
        kxl = (PASS * 6 + 64 * W - 1) // (64 * W)
        for p in (0, npass - 1):
            c = np.arange(kxl * 64 * W)
            src = np.where(c < PASS * xch, c, 0)
            assert (p * PASS * FXS // 8 + src).max() < npass * PASS * FXS // 8
            tl = np.arange(PASS)
            for r in range(P["GT"]):
                hi = (p * PASS + tl) * G + 16 * r + 4 * 3 + 3
                assert hi.max() < npass * PASS * G
        # the dW1 blocks of the W waves tile the whole matrix once
        nmb, nnb = P["MT"] // P["BM1"], P["MT"] // P["BN1"]
        assert nmb * nnb == W
        rows = [kperm_pos(16 * (P["BM1"] * mb + mt) + 4 * gg + i) for mb in range(nmb) for mt in range(P["BM1"])
                for gg in range(4) for i in range(4)]
        cols = [kperm_pos(16 * (P["BN1"] * nb + c) + n) for nb in range(nnb) for c in range(P["BN1"])
                for n in range(16)]
        assert len(rows) == H and len(cols) == H
        assert sorted(set(rows)) == list(range(H)) and sorted(set(cols)) == list(range(H))
        assert (blocks - 1) * n_dec + H * F + H * H + 15 * H + H - 1 < blocks * n_dec


def test_gather_and_scatter_lds_indices_in_bounds():
    Gm = _geo()
    sbox = Region("staged boxes", Gm["SBOXF"])
    for g1, (Q, CP) in enumerate(((Gm["Q0"], Gm["CP0"]), (Gm["Q1"], Gm["CP1"]))):
        base = Gm["SBOX0"] if g1 else 0
        bw, bh = Gm["SW"], Gm["SV"]
        cb = np.arange(bw * bh)
        for q in range(Q):
            sbox.touch(base + cb * CP + 4 * q, 4)   # staged writes
            ay, ax = np.meshgrid(np.arange(bh), np.arange(bw), indexing="ij")
            sbox.touch(base + (ay * bw + ax) * CP + 4 * q, 4)  # corner reads
    sbox.check()
    dxs = Region("dX rows", PASS * Gm["GS"])
    c = np.arange(PASS * G // 4)
    dxs.touch((c // (G // 4)) * Gm["GS"] + 4 * (c % (G // 4)), 4)  # row loads
    tex = np.arange(PASS)
    for slot0, C in ((0, C0), (Gm["G1B"], C1)):
        for ch in range(C):
            dxs.touch(tex * Gm["GS"] + slot0 + ch)  # box-gradient B operands
    dxs.check()


def _code_object_lds():
    if not os.path.exists(OBJ) or not os.path.exists(os.path.join(LLVM, "llvm-readelf")):
        pytest.skip("library not built here (openntc_amd/_build) or no LLVM tools")
    with tempfile.TemporaryDirectory() as d:
        fat, co, obj = os.path.join(d, "fat.bin"), os.path.join(d, "k.co"), os.path.join(d, "k.o")
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


def test_plans_match_code_object():
    sizes = _code_object_lds()
    Gm = _geo()
    want = {
        "k_mlp_train": _mlp_plan()["total"],
        "k_split_gather": 4 * Gm["SBOXF"] + 2 * 192 + 4 * 128,
        "k_split_scatter": 4 * PASS * Gm["GS"] + 2 * 192 + 4 * 128,
    }
    seen = set()
    for name, size in sizes.items():
        for k, v in want.items():
            if k in name:
                assert size == v, (k, size, v)
                assert size <= LDS_LIMIT
                seen.add(k)
    assert seen == set(want), seen

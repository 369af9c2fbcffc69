"""CPU check of the FP32 panel multiply's generated asm block
(gcoospdm_amd/csrc/gcoo_panel_asm.h).

A small interpreter runs the block's instructions for one lane: scalar and
vector registers, SCC, M0 and the VGPR-index mode, with list entries in
memory and LDS reads recorded symbolically.  Every fmaf step is logged as
(accumulator register written, value SGPR contents, D word read).  For each
entry count, the log must be exactly the list's entries in order, 4 steps
each (one per column component), each on accumulator v[16c + slot] with the
entry's value and the D row at the entry's LDS offset -- the per-row chains
of the multiply.  The staging loads are logged too and checked against the
requested count.
"""
import os
import re

import pytest

HDR = os.path.join(os.path.dirname(__file__), "..", "gcoospdm_amd", "csrc", "gcoo_panel_asm.h")


def _instructions():
    text = open(HDR).read()
    body = text.split("#define GCOO_PANEL_CHUNK_ASM", 1)[1].split("#define GCOO_PANEL_CHUNK_CLOBBERS", 1)[0]
    lines = re.findall(r'"([^"]*)\\n\\t"', body)
    return [ln.strip() for ln in lines if ln.strip()]


class Lane:
    def __init__(self, prog, entries, n, ops):
        self.prog = prog
        self.labels = {ln[:-1]: i for i, ln in enumerate(prog) if ln.endswith(":")}
        self.s = {}
        self.v = {}
        self.scc = 0
        self.m0 = 0x1234
        self.idx_on = False
        self.idx = 0
        self.entries = entries
        self.ops = ops  # operand values
        self.fmas = []
        self.stage = []
        self.s["n"] = n

    def sval(self, tok):
        tok = tok.strip()
        if tok == "m0":
            return self.m0
        if tok.startswith("%["):
            return self.ops[tok[2:-1]]
        if tok.startswith("0x"):
            return int(tok, 16)
        if re.fullmatch(r"-?\d+", tok):
            return int(tok)
        m = re.fullmatch(r"s(\d+)", tok)
        assert m, tok
        return self.s[int(m.group(1))]

    def setS(self, tok, val):
        tok = tok.strip()
        if tok == "m0":
            self.m0 = val & 0xFFFFFFFF
            return
        self.s[int(re.fullmatch(r"s(\d+)", tok).group(1))] = val & 0xFFFFFFFF

    def run(self, limit=200000):
        pc = 0
        steps = 0
        while pc < len(self.prog):
            steps += 1
            assert steps < limit, "runaway"
            ln = self.prog[pc]
            pc += 1
            if ln.endswith(":"):
                continue
            op, _, rest = ln.partition(" ")
            args = [a.strip() for a in rest.split(",")] if rest else []
            if op in ("s_waitcnt", "s_barrier", "s_setprio", "s_nop"):
                continue
            if op == "s_mov_b32":
                self.setS(args[0], self.sval(args[1]))
            elif op == "s_movk_i32":
                self.setS(args[0], self.sval(args[1]))
            elif op in ("s_add_u32", "s_sub_u32"):
                a, b = self.sval(args[1]), self.sval(args[2])
                self.setS(args[0], a + b if op == "s_add_u32" else a - b)
            elif op == "s_lshr_b32":
                self.setS(args[0], self.sval(args[1]) >> self.sval(args[2]))
            elif op.startswith("s_cmp_"):
                a, b = self.sval(args[0]), self.sval(args[1])
                rel = op[6:8]
                self.scc = int({"eq": a == b, "lg": a != b, "lt": a < b, "le": a <= b,
                                "gt": a > b, "ge": a >= b}[rel])
            elif op == "s_cbranch_scc0":
                if not self.scc:
                    pc = self.labels[args[0]]
            elif op == "s_cbranch_scc1":
                if self.scc:
                    pc = self.labels[args[0]]
            elif op == "s_branch":
                pc = self.labels[args[0]]
            elif op == "s_load_dwordx16":
                lo, hi = map(int, re.fullmatch(r"s\[(\d+):(\d+)\]", args[0]).groups())
                assert hi - lo == 15
                byte_off = self.sval(args[2])
                assert byte_off % 8 == 0
                first = byte_off // 8
                for u in range(8):
                    x, val = self.entries[first + u]  # entries past the list are padding
                    self.s[lo + 2 * u] = x
                    self.s[lo + 2 * u + 1] = val
            elif op == "s_set_gpr_idx_on":
                assert not self.idx_on
                self.idx_on = True
                self.idx = self.sval(args[0]) & 0xFF
            elif op == "s_set_gpr_idx_idx":
                assert self.idx_on
                self.idx = self.sval(args[0]) & 0xFF
            elif op == "s_set_gpr_idx_off":
                self.idx_on = False
            elif op == "v_readlane_b32":
                raise AssertionError("unexpected")
            elif op == "v_add_u32_e32":
                assert not self.idx_on
                d = int(args[0][1:])
                self.v[d] = ("addr", self.sval(args[1]) + self.ops["base"])
            elif op == "ds_read_b128":
                lo, hi = map(int, re.fullmatch(r"v\[(\d+):(\d+)\]", args[0]).groups())
                kind, addr = self.v[int(args[1][1:])]
                assert kind == "addr"
                for c in range(4):
                    self.v[lo + c] = ("d", addr, c)
            elif op == "buffer_load_dwordx4":
                self.stage.append((self.m0, self.sval(args[2].split()[0])))
            elif op == "v_fma_f32":
                assert self.idx_on, "fma outside index mode"
                d = int(args[0][1:]) + self.idx
                s2 = int(args[3][1:]) + self.idx
                assert d == s2
                src = self.v[int(args[2][1:])]
                self.fmas.append((d, self.sval(args[1]), src))
            else:
                raise AssertionError("unhandled " + ln)
        assert not self.idx_on
        return self


PROG = None


def _prog():
    global PROG
    if PROG is None:
        PROG = _instructions()
    return PROG


def _run(n, nld=0, stg=0, seed=0):
    import random
    rng = random.Random(seed * 1000 + n)
    base = 0x100
    # the wave's entries: rows in order, k ascending inside a row
    ents = []
    rows = sorted(rng.randrange(16) for _ in range(n))
    for slot in rows:
        k = rng.randrange(75)
        ents.append(((k * 1024) << 8 | slot, rng.getrandbits(32)))
    pad = [((rng.randrange(75) * 1024) << 8 | rng.randrange(16), rng.getrandbits(32)) for _ in range(24)]
    ops = {"n": n, "lp": 0, "base": base, "stg": stg, "nld": nld, "soff": 5000, "dst": 0x9000,
           "voff": 0, "rsrc": 0, "rowb": 32000}
    lane = Lane(_prog(), ents + pad, n, ops).run()
    return lane, ents, base


@pytest.mark.parametrize("n", list(range(0, 49)))
def test_entry_steps_follow_the_list(n):
    lane, ents, base = _run(n)
    expect = []
    for x, val in ents:
        slot, off = x & 0xFF, x >> 8
        for c in range(4):
            expect.append((16 * c + slot, val, ("d", base + off, c)))
    got = lane.fmas
    # the steps of one entry may be in any component order, but entries and
    # their per-row chains must be in list order
    assert len(got) == len(expect)
    for i in range(0, len(expect), 4):
        assert sorted(got[i:i + 4]) == sorted(expect[i:i + 4]), f"entry {i // 4} of {n}"


@pytest.mark.parametrize("nld", [0, 1, 7, 15])
def test_staging_loads(nld):
    lane, _, _ = _run(5, nld=nld, stg=1)
    assert lane.stage == [(0x9000 + 1024 * i, 5000 + 32000 * i) for i in range(nld)]
    assert lane.m0 == 0x1234  # restored

"""CPU check of the FP32 panel multiply's generated asm block
(gcoospdm_amd/csrc/gcoo_panel_asm.h): a segment of consecutive chunks.

A small interpreter runs the block's instructions for one lane: scalar and
vector registers, SCC, M0 and the VGPR-index mode, with the wave's list
entries in memory and LDS reads recorded symbolically.  Every fmaf step is
logged as (accumulator register written, value SGPR contents, D word read),
tagged with the chunk (barrier) it ran under.  For each chunk the log must be
exactly that chunk's entries in list order, 4 steps each (one per column
component), each on accumulator v[16c + slot] with the entry's value and the
D row at the entry's LDS offset in the chunk's buffer -- the per-row chains
of the multiply.  The staging loads are logged and checked against the rows
below K, and the block's final state (segment end, staging offset, M0)
against what the kernel carries to the next segment.

The interpreter also tracks the loads in flight under the LGKM counter: LDS
reads complete in order, scalar loads in any order, so after
s_waitcnt lgkmcnt(n) an LDS read is known complete only with n later LDS
reads issued, and a scalar load only for n = 0.  Reading or overwriting a
register that a load in flight may still write fails the test, and so does
leaving a load in flight at the block's end.
"""
import os
import random
import re

import pytest

HDR = os.path.join(os.path.dirname(__file__), "..", "gcoospdm_amd", "csrc", "gcoo_panel_asm.h")
CHUNKB = 75 * 1024
ROWB = 32000


ROWS = (16, 21, 28, 42)


ROWS_BF16 = (16, 21, 28)


def _instructions(rows=16, bf16=False):
    text = open(HDR).read()
    tag = f"{'BF16_' if bf16 else ''}R{rows}"
    body = text.split(f"#define GCOO_PANEL_SEGMENT_ASM_{tag} ", 1)[1].split(
        f"#define GCOO_PANEL_SEGMENT_CLOBBERS_{tag} ", 1)[0]
    lines = re.findall(r'"([^"]*)\\n\\t"', body)
    return [ln.strip() for ln in lines if ln.strip()]


def _regs(tok):
    """Register numbers named by an operand like s5, v[64:67], s[36:51]."""
    m = re.fullmatch(r"([sv])\[(\d+):(\d+)\]", tok)
    if m:
        return m.group(1), list(range(int(m.group(2)), int(m.group(3)) + 1))
    m = re.fullmatch(r"([sv])(\d+)", tok)
    if m:
        return m.group(1), [int(m.group(2))]
    return None, []


class Lane:
    def __init__(self, prog, entries, ops, vlanes):
        self.prog = prog
        self.labels = {ln[:-1]: i for i, ln in enumerate(prog) if ln.endswith(":")}
        self.s = {}
        self.v = {}
        self.scc = 0
        self.m0 = 0x1234
        self.idx_on = False
        self.idx = 0
        self.entries = entries     # the list from %[pbase], 8 bytes each
        self.ops = dict(ops)       # operand values (in/out ones updated)
        self.vlanes = vlanes       # per-lane values of the window VGPRs
        self.fmas = []
        self.stage = []
        self.barriers = 0
        self.inflight = []         # ("smem"|"lds", set of ("s"|"v", reg))

    # -- loads in flight --
    def _check(self, kind, regs, what):
        for _, dest in self.inflight:
            for r in regs:
                assert (kind, r) not in dest, f"{what} {kind}{r} while a load to it is in flight"

    def wait_lgkm(self, n):
        lds = [op for op in self.inflight if op[0] == "lds"]
        keep = set(map(id, lds[len(lds) - n:])) if n else set()
        self.inflight = [op for op in self.inflight
                         if (op[0] == "smem" and n > 0) or (op[0] == "lds" and id(op) in keep)]

    # -- operands --
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
        r = int(m.group(1))
        self._check("s", [r], "read")
        return self.s[r]

    def setS(self, tok, val):
        tok = tok.strip()
        val &= 0xFFFFFFFF
        if tok == "m0":
            self.m0 = val
            return
        if tok.startswith("%["):
            self.ops[tok[2:-1]] = val
            return
        r = int(re.fullmatch(r"s(\d+)", tok).group(1))
        self._check("s", [r], "write")
        self.s[r] = val

    def vread(self, tok):
        r = int(tok[1:])
        self._check("v", [r], "read")
        return self.v[r]

    def run(self, limit=400000):
        for _ in self.execute(limit):
            pass
        return self

    # hooks of the multi-wave model (tests/test_panel_asm_waves.py); a single
    # lane has no other waves and no memory timing
    def on_dma(self, m0, soff, voff):
        pass

    def on_lds_read(self, addr, nbytes):
        pass

    def execute(self, limit=400000):
        """Run the block; yields ("barrier",) at each s_barrier (before
        passing it), ("vmcnt", n) at each vector-memory wait, and ("step",)
        after every other instruction, so a scheduler can interleave waves."""
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
            if op in ("s_setprio", "s_nop"):
                continue
            if op == "s_barrier":
                yield ("barrier",)
                self.barriers += 1
            elif op == "s_waitcnt":
                m = re.fullmatch(r"(vmcnt|lgkmcnt)\((\d+)\)", rest.strip())
                assert m, ln
                if m.group(1) == "lgkmcnt":
                    self.wait_lgkm(int(m.group(2)))
                else:
                    yield ("vmcnt", int(m.group(2)))
            elif op in ("s_mov_b32", "s_movk_i32"):
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
                kind, regs = _regs(args[0])
                assert kind == "s" and len(regs) == 16 and args[1] == "%[pbase]"
                byte_off = self.sval(args[2])
                assert byte_off % 8 == 0
                self._check("s", regs, "load into")
                first = byte_off // 8
                for u in range(8):
                    x, val = self.entries[first + u]  # entries past the list are padding
                    self.s[regs[2 * u]] = x
                    self.s[regs[2 * u + 1]] = val
                self.inflight.append(("smem", {("s", r) for r in regs}))
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
                assert not self.idx_on
                lane = self.sval(args[2])
                assert 0 <= lane < 64
                self.setS(args[0], self.vlanes[args[1][2:-1]][lane])
            elif op == "v_mad_u32_u16":
                # D = S0.u16 * S1.u16 + S2.u32; op_sel:[1,0,0,0] takes S0's
                # high half
                assert not self.idx_on
                d = int(args[0][1:])
                assert ln.endswith(" op_sel:[1,0,0,0]"), ln
                src2 = args[3].split()[0]
                self._check("v", [d], "write")
                self.v[d] = ("addr", ((self.sval(args[1]) >> 16) & 0xFFFF)
                             * (self.sval(args[2]) & 0xFFFF) + self.sval(src2))
            elif op == "ds_read_b128":
                kind, regs = _regs(args[0])
                akind, addr = self.vread(args[1])
                assert akind == "addr"
                self._check("v", regs, "load into")
                self.on_lds_read(addr, 16)
                for c in range(4):
                    self.v[regs[c]] = ("d", addr, c)
                self.inflight.append(("lds", {("v", r) for r in regs}))
            elif op == "ds_read_b64":
                # two dwords of BF16 pairs: dword i holds columns 2i (low
                # half) and 2i + 1 (high half)
                kind, regs = _regs(args[0])
                assert len(regs) == 2 and regs[0] % 2 == 0
                akind, addr = self.vread(args[1])
                assert akind == "addr"
                self._check("v", regs, "load into")
                self.on_lds_read(addr, 8)
                for i in range(2):
                    self.v[regs[i]] = ("d64", addr, i)
                self.inflight.append(("lds", {("v", r) for r in regs}))
            elif op == "ds_read_u16_d16_hi":
                # BF16: 16 bits into the register's high half, the low half
                # kept; the result is the exact FP32 value only if that low
                # half is zero (a zeroed register, or one only such reads
                # have written).  Byte offset 2 j is column j of the lane.
                d = int(args[0][1:])
                a, _, off = args[1].partition(" ")
                off = int(off.split(":")[1]) if off else 0
                assert off in (0, 2, 4, 6), ln
                akind, addr = self.vread(a)
                assert akind == "addr"
                self._check("v", [d], "load into")
                prev = self.v.get(d)
                assert prev == ("zero",) or (prev is not None and prev[0] == "d16"), \
                    f"16-bit read into v{d} whose low half is not known zero"
                self.on_lds_read(addr + off, 2)
                self.v[d] = ("d16", addr, off // 2)
                self.inflight.append(("lds", {("v", d)}))
            elif op in ("v_and_b32_e32", "v_lshlrev_b32_e32"):
                # BF16 widening: mask keeps the high half (odd column), a
                # shift by 16 moves the low half up (even column)
                assert not self.idx_on, "widening in index mode"
                d = int(args[0][1:])
                kind, addr, i = self.vread(args[2])
                assert kind == "d64", ln
                if op == "v_and_b32_e32":
                    assert args[1] == "0xffff0000", ln
                    col = 2 * i + 1
                else:
                    assert args[1] == "16", ln
                    col = 2 * i
                self._check("v", [d], "write")
                self.v[d] = ("d", addr, col)
            elif op == "v_mov_b32":
                assert not self.idx_on and args[1] == "0", ln
                d = int(args[0][1:])
                self._check("v", [d], "write")
                self.v[d] = ("zero",)
            elif op == "buffer_load_dwordx4":
                assert args[0] in ("%[voff]", "%[voffz]") and args[1] == "%[rsrc]"
                soff = self.sval(args[2].split()[0])
                self.stage.append((self.barriers, self.m0, soff, args[0][2:-1]))
                self.on_dma(self.m0, soff, args[0][2:-1])
            elif op == "v_fma_f32":
                assert self.idx_on, "fma outside index mode"
                d = int(args[0][1:]) + self.idx
                s2 = int(args[3][1:]) + self.idx
                assert d == s2
                src = self.vread(args[2])
                if src[0] == "d16":
                    src = ("d",) + src[1:]
                assert src[0] == "d", f"step on {src}"
                self._check("v", [d], "accumulate into")
                self.fmas.append((self.barriers, d, self.sval(args[1]), src))
            else:
                raise AssertionError("unhandled " + ln)
            yield ("step",)
        assert not self.idx_on
        assert not self.inflight, "loads in flight at the block's end"


PROG = {}


def word(offset, slot):
    """A list entry's first word: (LDS byte offset / 16) << 16 | row slot."""
    assert offset % 16 == 0 and offset // 16 < 65536 and slot < 256
    return (offset // 16) << 16 | slot


def _prog(rows=16, bf16=False):
    if (rows, bf16) not in PROG:
        PROG[rows, bf16] = _instructions(rows, bf16)
    return PROG[rows, bf16]


def _segment(counts, c0=0, cw=0, stager=True, K=8000, seed=0, rows=16, bf16=False):
    """Run one segment: chunks c0 .. c0 + len(counts) - 1 of the window at
    cw, with counts[i] entries in chunk c0 + i (the window's other chunks get
    random counts).  Returns the lane, the per-chunk entries and the inputs."""
    rng = random.Random(seed)
    # chunk rows and staged bytes per D row: FP32 75 x 1 KB, BF16 150 x 512 B
    tk, rowd = (150, 512) if bf16 else (75, 1024)
    nchunks = c0 - cw + len(counts)
    assert nchunks <= 64
    per = [rng.randrange(20) for _ in range(c0 - cw)] + list(counts)
    per += [rng.randrange(20) for _ in range(64 - len(per))]
    entries, vpbo, vpn, chunks = [], [], [], []
    for i, n in enumerate(per):
        vpbo.append(len(entries) * 8)
        vpn.append(n)
        slots = sorted(rng.randrange(rows) for _ in range(n))
        # offsets from buffer 0: chunk c's rows are in buffer c & 1
        buf = ((cw + i) & 1) * CHUNKB
        ch = [(word(buf + rng.randrange(tk) * rowd, slot), rng.getrandbits(32)) for slot in slots]
        chunks.append(ch)
        entries += ch
    entries += [(word(rng.randrange(tk) * rowd, rng.randrange(rows)), rng.getrandbits(32))
                for _ in range(24)]
    wave = 2
    rpl = 2 if bf16 else 1                 # D rows per wave load
    rowb = rpl * ROWB                      # D bytes a load advances
    kb = K * ROWB if stager else 0
    lim = max(kb - 15 * rowb, 0) if stager else 0
    sso = ((c0 + 1) * tk + wave * 15 * rpl) * ROWB
    ops = {"pbase": 0, "sli": c0 - cw, "liend": c0 - cw + len(counts), "par": c0 & 1,
           "vb0": 0x40, "sdst": wave * 15 * 1024, "sso": sso,
           "lim": lim, "kb": kb, "rowb": rowb, "row1": ROWB, "tkrowb": tk * ROWB}
    lane = Lane(_prog(rows, bf16), entries, ops, {"vpbo": vpbo, "vpn": vpn}).run()
    return lane, chunks[c0 - cw:c0 - cw + len(counts)], ops


def _check_steps(lane, chunks, c0, ops, rows=16):
    for i, ch in enumerate(chunks):
        base = ops["vb0"]   # the list offsets include the buffer
        expect = []
        for x, val in ch:
            slot, off = x & 0xFF, (x >> 16) * 16
            for c in range(4):
                expect.append((rows * c + slot, val, ("d", base + off, c)))
        got = [f[1:] for f in lane.fmas if f[0] == i + 1]
        assert len(got) == len(expect), f"chunk {c0 + i}: {len(got) // 4} entries"
        # the 4 steps of one entry may be in any component order, but entries
        # (and so every row's chain) must be in list order
        for j in range(0, len(expect), 4):
            assert sorted(got[j:j + 4]) == sorted(expect[j:j + 4]), f"chunk {c0 + i} entry {j // 4}"


@pytest.mark.parametrize("rows", ROWS)
@pytest.mark.parametrize("n", list(range(0, 49)))
@pytest.mark.parametrize("c0", [0, 1])
def test_one_chunk_steps_follow_the_list(n, c0, rows):
    lane, chunks, ops = _segment([n], c0=c0, cw=0, seed=n, rows=rows)
    assert lane.barriers == 1
    _check_steps(lane, chunks, c0, ops, rows)


@pytest.mark.parametrize("rows", ROWS)
@pytest.mark.parametrize("seed", range(12))
def test_segment_steps_follow_the_list(seed, rows):
    rng = random.Random(100 + seed)
    cw = rng.choice([0, 64, 128])
    c0 = cw + rng.randrange(8)
    counts = [rng.choice([0, 1, 3, 4, 5, 8, 9, 12, 15, 16, 17, 23, 24, 31, 40])
              for _ in range(rng.randrange(1, 14))]
    lane, chunks, ops = _segment(counts, c0=c0, cw=cw, seed=seed, rows=rows)
    assert lane.barriers == len(counts)
    _check_steps(lane, chunks, c0, ops, rows)
    assert lane.ops["sli"] == c0 - cw + len(counts)
    assert lane.m0 == 0x1234  # restored


@pytest.mark.parametrize("rows", ROWS)
@pytest.mark.parametrize("K", [8000, 300, 161, 75 * 7 + 33])
def test_staging_loads(K, rows):
    counts = [5, 0, 17, 2, 9, 30]
    c0 = 0
    lane, _, ops = _segment(counts, c0=c0, K=K, seed=K, rows=rows)
    wave = 2
    want = []
    for i in range(len(counts)):
        c1 = c0 + i + 1                      # the chunk this barrier's staging loads
        k0 = c1 * 75 + wave * 15
        dst = ops["sdst"] + (c1 % 2) * CHUNKB
        for r in range(15):
            if k0 + r < K:
                want.append((i + 1, dst + 1024 * r, (k0 + r) * ROWB, "voff"))
    assert lane.stage == want
    assert lane.ops["sso"] == ops["sso"] + len(counts) * 75 * ROWB


@pytest.mark.parametrize("rows", ROWS_BF16)
@pytest.mark.parametrize("K", [8000, 8001, 300, 301, 161, 150 * 3 + 33])
def test_staging_loads_bf16(K, rows):
    """BF16: a wave load holds two D rows; a pair with its first row below K
    is loaded, and with its second row at K or beyond it takes the
    out-of-range lane offsets (odd K)."""
    counts = [5, 0, 17, 2, 9, 30]
    lane, _, ops = _segment(counts, c0=0, K=K, seed=K, rows=rows, bf16=True)
    wave = 2
    want = []
    for i in range(len(counts)):
        c1 = i + 1
        k0 = c1 * 150 + wave * 30
        dst = ops["sdst"] + (c1 % 2) * CHUNKB
        for r in range(15):
            k = k0 + 2 * r
            if k < K:
                want.append((i + 1, dst + 1024 * r, k * ROWB, "voff" if k + 1 < K else "voffz"))
    assert lane.stage == want
    assert lane.ops["sso"] == ops["sso"] + len(counts) * 150 * ROWB


@pytest.mark.parametrize("rows", ROWS_BF16)
@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 7, 8, 11, 12, 15, 16, 17, 23, 24, 25, 31, 40, 47])
def test_bf16_steps_follow_the_list(n, rows):
    for c0 in (0, 1):
        lane, chunks, ops = _segment([n], c0=c0, cw=0, seed=n, rows=rows, bf16=True)
        assert lane.barriers == 1
        _check_steps(lane, chunks, c0, ops, rows)


@pytest.mark.parametrize("rows", ROWS_BF16)
@pytest.mark.parametrize("seed", range(6))
def test_bf16_segment_steps_follow_the_list(seed, rows):
    rng = random.Random(200 + seed)
    cw = rng.choice([0, 64])
    c0 = cw + rng.randrange(8)
    counts = [rng.choice([0, 1, 3, 4, 9, 16, 17, 24, 31, 40]) for _ in range(rng.randrange(1, 10))]
    lane, chunks, ops = _segment(counts, c0=c0, cw=cw, seed=seed, rows=rows, bf16=True)
    assert lane.barriers == len(counts)
    _check_steps(lane, chunks, c0, ops, rows)
    assert lane.m0 == 0x1234


@pytest.mark.parametrize("rows", ROWS)
def test_non_staging_wave_loads_nothing(rows):
    lane, chunks, ops = _segment([3, 20, 0, 8], c0=5, cw=0, stager=False, seed=7, rows=rows)
    assert lane.stage == []
    _check_steps(lane, chunks, 5, ops, rows)


def test_interpreter_catches_an_early_read():
    # the hazard model itself: a wait that is one read short must fail
    prog = [ln.replace("lgkmcnt(4)", "lgkmcnt(5)") for ln in _prog()]
    lane_cls = Lane
    rng = random.Random(3)
    ents = [(word(rng.randrange(75) * 1024, i % 16), i) for i in range(40)]
    ops = {"pbase": 0, "sli": 0, "liend": 1, "par": 0, "vb0": 0, "sdst": 0,
           "sso": 75 * ROWB, "lim": 0, "kb": 0, "rowb": ROWB, "tkrowb": 75 * ROWB}
    with pytest.raises(AssertionError, match="in flight"):
        lane_cls(prog, ents, ops, {"vpbo": [0] * 64, "vpn": [12] + [0] * 63}).run()


def test_interpreter_catches_a_bf16_read_without_a_zero_low_half():
    # BF16 column 0 is exact only if its register's low half is zero: a block
    # that does not zero those registers first must fail
    prog = [ln for ln in _prog(21, True) if not re.fullmatch(r"v_mov_b32 v\d+, 0", ln)]
    assert len(prog) < len(_prog(21, True))
    rng = random.Random(4)
    ents = [(word(rng.randrange(150) * 512, i % 21), i) for i in range(40)]
    ops = {"pbase": 0, "sli": 0, "liend": 1, "par": 0, "vb0": 0, "sdst": 0,
           "sso": 150 * ROWB, "lim": 0, "kb": 0, "rowb": 2 * ROWB, "row1": ROWB,
           "tkrowb": 150 * ROWB}
    with pytest.raises(AssertionError, match="low half"):
        Lane(prog, ents, ops, {"vpbo": [0] * 64, "vpn": [12] + [0] * 63}).run()


def test_committed_asm_header_is_the_generator_output():
    """gcoo_panel_asm.h is exactly what gen_panel_asm.py writes."""
    import subprocess
    import sys
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "gcoospdm_amd", "csrc")
    out = subprocess.run([sys.executable, os.path.join(here, "gen_panel_asm.py")],
                         capture_output=True, text=True, check=True, timeout=120).stdout
    with open(os.path.join(here, "gcoo_panel_asm.h")) as f:
        assert out == f.read()

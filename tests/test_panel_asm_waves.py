"""CPU model of the panel multiply's chunk loop across the waves of a
workgroup: the generated asm block (gcoospdm_amd/csrc/gcoo_panel_asm.h) run
by several waves at once, with the workgroup barrier and the buffer-to-LDS
staging loads (LDS DMA) modelled.

Five staging waves each load their 15 slots (1 KB each) of the next chunk
into the other LDS buffer; one more wave only computes.  A scheduler
interleaves the waves instruction by instruction at random; a wave blocks at
s_barrier until every wave has arrived, and at s_waitcnt vmcnt(n) until at
most n of its own staging loads are outstanding.  A staging load's LDS write
lands at a random later moment -- in issue order per wave, or in any order
(out-of-order completion; vmcnt(n) then only bounds the number outstanding).
Every LDS read of an entry's D row must find its slot holding the chunk the
wave is in, completely written: a read of a slot whose load has not landed,
or that a later chunk's load has already claimed, fails the test.

The shipped schedule -- each staging wave waits for all its loads
(vmcnt(0)), then the barrier, then the staging of the next chunk into the
buffer every wave has just finished with -- passes under both completion
orders.  Planted mistakes fail: no wait before the barrier, or staging issued
before the barrier.  An abstract four-buffer schedule with partial waits
(the next chunk's loads left in flight at each barrier) passes with in-order
completion and fails with out-of-order completion.
"""
import random

import pytest

from tests import test_panel_asm as T

CHUNKB = T.CHUNKB
NSTAGE = 5


class WaveLane(T.Lane):
    """A lane of one wave, reporting staging loads and LDS reads to the
    shared model."""

    def __init__(self, model, wave, *a, **kw):
        super().__init__(*a, **kw)
        self.model = model
        self.wave = wave

    def on_dma(self, m0, soff, voff):
        self.model.dma(self, m0)

    def on_lds_read(self, addr, nbytes):
        self.model.read(self, addr)


class Model:
    def __init__(self, c0, bf16, ooo, rng):
        self.c0, self.bf16, self.ooo, self.rng = c0, bf16, ooo, rng
        # slot state per (buffer, 1 KB slot): (chunk, landed)
        self.slots = {}
        for b in (0, 1):
            for s in range(75):
                # chunk c0 is in its buffer before the segment; the other
                # buffer holds the chunk before it
                self.slots[b, s] = (c0 if b == c0 % 2 else c0 - 1, True)
        self.outstanding = {}   # wave -> list of (buffer, slot, chunk)
        self.errors = []

    def chunk_of(self, lane):
        return self.c0 + lane.barriers - 1

    def dma(self, lane, m0):
        b, s = m0 // CHUNKB, (m0 % CHUNKB) // 1024
        # the load is for the next chunk that lives in buffer b (the body of
        # chunk c stages c + 1)
        chunk = self.chunk_of(lane) + 1
        if chunk % 2 != b:
            chunk += 1
        self.slots[b, s] = (chunk, False)
        self.outstanding.setdefault(lane.wave, []).append((b, s, chunk))

    def land(self, wave, pick_any):
        q = self.outstanding.get(wave, [])
        if not q:
            return False
        i = self.rng.randrange(len(q)) if pick_any else 0
        b, s, chunk = q.pop(i)
        cur = self.slots[b, s]
        if cur[0] == chunk:
            self.slots[b, s] = (chunk, True)
        return True

    def read(self, lane, addr):
        off = addr - 0x40
        b, o = off // CHUNKB, off % CHUNKB
        row = o // (512 if self.bf16 else 1024)
        s = row // 2 if self.bf16 else row
        want = self.chunk_of(lane)
        got = self.slots[b, s]
        if got != (want, True):
            self.errors.append(f"wave {lane.wave} in chunk {want} read buffer {b} slot {s} "
                               f"holding {got}")


def _simulate(prog, counts, c0=2, bf16=False, ooo=False, seed=0, nwaves=6, rows=16, K=8000,
              p_land=0.05, slow=None):
    """`slow`: a wave picked 20 times less often than the others when runnable."""
    rng = random.Random(seed)
    model = Model(c0, bf16, ooo, rng)
    tk, rowd = (150, 512) if bf16 else (75, 1024)
    rpl = 2 if bf16 else 1
    lanes = []
    for w in range(nwaves):
        wr = random.Random(seed * 131 + w)
        entries, vpbo, vpn = [], [], []
        for c, n in enumerate([wr.randrange(20) for _ in range(c0)] + list(counts[w]) + [0] * 8):
            vpbo.append(len(entries) * 8)
            vpn.append(n)
            # offsets from buffer 0: chunk c's rows are in buffer c & 1
            entries += [(T.word((c & 1) * CHUNKB + wr.randrange(tk) * rowd, wr.randrange(rows)),
                         wr.getrandbits(32)) for _ in range(n)]
        vpbo += [0] * (64 - len(vpbo))
        vpn += [0] * (64 - len(vpn))
        entries += [(T.word(0, 0), 0)] * 24
        stager = w < NSTAGE
        kb = K * T.ROWB if stager else 0
        rowb = rpl * T.ROWB
        ops = {"pbase": 0, "sli": c0, "liend": c0 + len(counts[w]), "par": c0 & 1,
               "vb0": 0x40, "sdst": w * 15 * 1024,
               "sso": ((c0 + 1) * tk + w * 15 * rpl) * T.ROWB,
               "lim": max(kb - 15 * rowb, 0) if stager else 0, "kb": kb, "rowb": rowb,
               "row1": T.ROWB, "tkrowb": tk * T.ROWB}
        lanes.append(WaveLane(model, w, prog, entries, ops, {"vpbo": vpbo, "vpn": vpn}))
    gens = [ln.execute() for ln in lanes]
    state = [None] * nwaves        # None runnable, ("barrier",), ("vmcnt", n), "done"
    while True:
        # release a barrier every wave has reached
        if all(st == ("barrier",) or st == "done" for st in state) and \
                any(st == ("barrier",) for st in state):
            assert all(st == ("barrier",) for st in state), "waves pass different barrier counts"
            state = [None] * nwaves
        # land some staging loads: rarely on their own (p_land per step), and
        # with a finite latency for a wave that waits for them
        for w in range(nwaves):
            waiting = isinstance(state[w], tuple) and state[w][0] == "vmcnt"
            if model.outstanding.get(w) and rng.random() < (0.05 if waiting else p_land):
                model.land(w, ooo)
        run = [w for w in range(nwaves) if state[w] is None or
               (isinstance(state[w], tuple) and state[w][0] == "vmcnt" and
                len(model.outstanding.get(w, [])) <= state[w][1])]
        if not run:
            if all(st == "done" for st in state):
                break
            blocked = [w for w in range(nwaves) if isinstance(state[w], tuple)
                       and state[w][0] == "vmcnt"]
            if blocked:
                model.land(rng.choice(blocked), ooo)
                continue
            if all(st in ("done", ("barrier",)) for st in state):
                continue
            raise AssertionError(f"deadlock: {state}")
        if slow in run and len(run) > 1 and rng.random() < 0.95:
            run.remove(slow)
        w = rng.choice(run)
        state[w] = None
        try:
            ev = next(gens[w])
        except StopIteration:
            state[w] = "done"
            continue
        if ev[0] in ("barrier", "vmcnt"):
            state[w] = ev
    return model, lanes


def _counts(rng, nwaves, nchunks):
    return [[rng.choice([0, 1, 3, 5, 8, 12, 16, 17, 24]) for _ in range(nchunks)]
            for _ in range(nwaves)]


@pytest.mark.parametrize("bf16,rows", [(False, 16), (False, 21), (False, 28), (False, 42),
                                       (True, 16), (True, 21), (True, 28)])
@pytest.mark.parametrize("ooo", [False, True])
@pytest.mark.parametrize("p_land", [0.05, 0.0005])
@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("slow", [None, 0, 5])
def test_shipped_schedule_reads_only_landed_chunks(bf16, rows, ooo, p_land, seed, slow):
    """Staging loads landing quickly or only when waited for; all waves at
    one pace, or a staging wave (0) or the other wave (5) far behind."""
    prog = T._prog(rows, bf16)
    counts = _counts(random.Random(seed), 6, 5)
    model, lanes = _simulate(prog, counts, bf16=bf16, ooo=ooo, seed=seed, rows=rows,
                             p_land=p_land, slow=slow)
    assert not model.errors, model.errors[:3]
    assert all(ln.barriers == 5 for ln in lanes)


def _mutations():
    prog = T._prog(16, False)
    # 1. no wait for the staging loads before the barrier
    no_wait = [ln for ln in prog if ln != "s_waitcnt vmcnt(0)"]
    # 2. the next chunk staged before the barrier: move each stage sequence
    # (from the m0 set-up to its end label) in front of the barrier
    early = list(prog)
    for p in (0, 1):
        b = early.index("s_barrier", early.index(f".Lc{p}_%=:"))
        s0 = next(i for i in range(b, len(early)) if early[i] == "s_cmp_le_u32 %[sso], %[lim]")
        s1 = next(i for i in range(s0, len(early)) if early[i].startswith(
            "s_add_u32 %[sso], %[sso], %[tkrowb]")) + 1
        block = early[s0:s1]
        del early[s0:s1]
        early[b:b] = block
    return {"no_wait": no_wait, "stage_before_barrier": early}


@pytest.mark.parametrize("name", ["no_wait", "stage_before_barrier"])
def test_planted_schedule_mistakes_are_caught(name):
    prog = _mutations()[name]
    caught = False
    for seed in range(6):
        counts = _counts(random.Random(seed), 6, 5)
        model, _ = _simulate(prog, counts, seed=seed, p_land=0.0005, slow=seed % 2 * 5)
        caught |= bool(model.errors)
    assert caught, name


def _abstract_four_buffer(ooo, seed, nwaves=4, nchunks=12, loads=3):
    """Four buffers, chunk c in buffer c % 4, staged two chunks ahead: after
    barrier c each staging wave issues its loads of chunk c + 2, then works
    on chunk c, then waits vmcnt(loads) -- chunk c + 1's loads complete,
    chunk c + 2's may stay in flight -- before barrier c + 1."""
    rng = random.Random(seed)
    landed = {}                      # (chunk, wave, i) -> landed
    outstanding = {w: [] for w in range(nwaves)}
    errors = []
    for c in (0, 1):
        for w in range(nwaves):
            for i in range(loads):
                landed[c, w, i] = True
    for c in range(nchunks):
        # barrier c passed; each wave stages chunk c + 2 then reads chunk c
        order = list(range(nwaves))
        rng.shuffle(order)
        for w in order:
            for i in range(loads):
                landed[c + 2, w, i] = False
                outstanding[w].append((c + 2, w, i))
        for w in order:
            for _ in range(rng.randrange(3)):
                q = outstanding[rng.randrange(nwaves)]
                if q:
                    key = q.pop(rng.randrange(len(q)) if ooo else 0)
                    landed[key] = True
            for v in range(nwaves):
                for i in range(loads):
                    if not landed[c, v, i]:
                        errors.append((c, w, v, i))
        # each wave: s_waitcnt vmcnt(loads), then the barrier
        for w in range(nwaves):
            while len(outstanding[w]) > loads:
                key = outstanding[w].pop(rng.randrange(len(outstanding[w])) if ooo else 0)
                landed[key] = True
    return errors


def test_four_buffer_partial_waits_need_in_order_completion():
    assert not any(_abstract_four_buffer(False, s) for s in range(20))
    assert any(_abstract_four_buffer(True, s) for s in range(20))

#!/usr/bin/env python3
"""Writes gcoo_panel_asm.h: the chunk loop of the FP32 panel multiply
(spdm_panel_kernel, gcoo_spdm.hip) for one wave, as one inline-asm body that
runs a segment of consecutive chunks, for each supported number R of rows
per wave (ROWS below): macros GCOO_PANEL_SEGMENT_ASM_R<R>, its clobbers
GCOO_PANEL_SEGMENT_CLOBBERS_R<R>, and GCOO_PANEL_ACC_OPERANDS_R<R>(a), the
operands pinning the accumulators (an array of 4 R / 8 f32x8) to v0..v(4R-1);
and for BF16 D (ROWS_BF16) GCOO_PANEL_SEGMENT_ASM_BF16_R<R> with its
clobbers.  The BF16 block stages 150-row chunks of 512-byte D rows (two rows
per 1 KB wave load, %[rowb] the two rows' bytes, %[row1] one row's) and
reads an entry's 4 values with ds_read_b64, widening columns 1..3 exactly
to FP32 (the BF16 bits as the high half: 2 ANDs, 1 shift), and column 0
with a second, 16-bit read into the high half of a register whose low half
the block zeroes once per segment (so that column costs an LDS read instead
of a vector instruction); the list values are already widened.

    python3 gen_panel_asm.py > gcoo_panel_asm.h

A segment is chunks c .. c + m - 1 of one 64-chunk window: lane i of the
VGPRs %[vpbo] / %[vpn] holds the wave's list of chunk cw + i (byte offset of
its first entry from %[pbase], entry count), and the segment is lanes
%[sli] .. %[liend] - 1.  Chunk c uses LDS buffer c & 1 and scalar bank c & 1
(a bank is two sets of 8 entries, A and B), so the body is written twice,
once per parity, and the chunks alternate between the two copies.

Per chunk, in order:
  * the staging waves (%[kb] != 0) wait for their loads of this chunk
    (vmcnt(0)), then the workgroup barrier: the chunk is in LDS and every
    wave is done with the other buffer;
  * the wait for this chunk's first 16 entries (lgkmcnt(0)): loaded during
    the previous chunk, or before the barrier for a segment's first chunk;
  * the next chunk's bounds (v_readlane) and the scalar loads of its first
    16 entries into the other bank, which land while this chunk runs;
  * the staging of the next chunk into the other buffer: a wave's 15 loads
    of 1 KB (one D row each: LDS destination M0 + 1 KB per load, D offset
    %[sso] + %[rowb] per load, lane offset %[voff]).  When all 15 rows are
    below K (%[sso] <= %[lim]) the loads are unrolled without tests,
    otherwise each is issued only while its offset is below %[kb];
  * entries 0..15 in groups of 4 over two D buffers: the LDS reads of group
    g + 1 are issued before the steps of group g, and lgkmcnt(4) waits for
    group g (LDS reads complete in order, so with 4 later reads issued the
    count is reached only once group g is in, whatever scalar loads are
    still outstanding; BF16 issues 8 reads per group and waits for
    lgkmcnt(8)).  A last group of 1..3 entries reads and steps only
    its valid entries, after lgkmcnt(0);
  * entries from 16 on: the batch loop, 8 entries per batch: 8 LDS reads,
    the wait for them, the scalar load of the next batch's 8 entries (run
    under this batch's steps), 8 steps, the wait for the next entries.  Two
    register sets alternate; the last 1..7 entries take one guarded round.
Entry u is two words in SGPRs: x = (LDS byte offset of its D row from
buffer 0, so buffer 1's rows at + 75 KB, / 16) << 16 | row slot, and the
FP32 bits of its value.  The LDS address of an entry is one vector
multiply-add of x's high half (v_mad_u32_u16 with op_sel: hi16 * 16 + the
lane's address in buffer 0), so no scalar instruction goes into it; two
address VGPRs serve all reads in turn (a read takes its address at issue).  A
step is acc[slot] = fmaf(value, d, acc[slot]) for the 4 columns, in VGPR-index
mode (s_set_gpr_idx_on/_idx take the slot from bits 7:0 of x).  The wave's
issue priority is 3 for a chunk of 3R/4 or more entries (2 below: R rows
hold 3R/4 entries of a chunk at 1 % density), 3 / 2 before each batch by
whether R or more entries are left, and 0 at the chunk's end (n = 8000,
s = 0.99, 16 x 21: 403 us against 410 us with the thresholds of 16 rows).

Operands: %[pbase] SGPR pair; %[vpbo], %[vpn] VGPRs as above; %[sli]
(in/out: lane of the current chunk, the segment's end on exit) and %[liend]
SGPRs; %[par] SGPR, parity of the first chunk; %[vb0] VGPR, this lane's
LDS byte address in buffer 0; %[sdst] SGPR, the wave's staging
destination in buffer 0; %[sso] (in/out: D byte offset of the wave's first
staging row of the next chunk, advanced by %[tkrowb] per chunk), %[lim],
%[kb], %[rowb], %[tkrowb] SGPRs; %[voff] VGPR; %[rsrc] the buffer resource
of D; the accumulators v0..v(4R-1) (component c of row slot q in
v[R c + q]).  Scratch: v(4R)..v(4R+33), s24..s30 and s36..s99 (s32..s34 are
the stack, frame and base pointers).  M0 is saved and restored.  The scalar loads
read up to 7 entries past the end of a chunk's list, which the list capacity
covers; no LDS read is issued for an entry past the end.  Every wave runs
the body once per chunk (one barrier each).  When the block exits, the only
loads of it still in flight are the staging loads, which the next barrier's
vmcnt(0) covers.
"""
ST, SREM, SOFF, SCNT, SM0, PBN, SNN = 24, 25, 26, 27, 28, 29, 30
BANKS = ((list(range(36, 52)), list(range(52, 68))),
         (list(range(68, 84)), list(range(84, 100))))
CHUNKB = 75 * 1024
ROWS = (16, 21, 28, 42)   # rows per wave of the generated FP32 blocks
ROWS_BF16 = (16, 21, 28)  # ... and BF16 blocks
NSCR = 34                 # scratch VGPRs above the accumulators
PRIO_CHUNK, PRIO_BATCH = 12, 16  # entries left for issue priority 3 (see configure)
R = 16                    # rows per wave of the block being generated
BF16 = False              # the block reads BF16 D (8 bytes per lane) and widens it
RPE = 1                   # LDS read instructions per entry (BF16: 2)
DV, AV = 64, 96           # D values of the batch loop; two address VGPRs, used in turn
G_D = (64, 80)            # D buffers of the rolling phase
L = []


def configure(rows, bf16=False):
    """Register layout for `rows` rows per wave: accumulators v0..v(4 rows - 1),
    scratch from v(4 rows) on."""
    global R, DV, AV, G_D, BF16, RPE, PRIO_CHUNK, PRIO_BATCH
    R = rows
    BF16 = bf16
    RPE = 2 if bf16 else 1
    PRIO_CHUNK = (3 * rows + 2) // 4
    PRIO_BATCH = rows
    base = 4 * rows
    DV, AV = base, base + 32
    G_D = (base, base + 16)


def av(u):
    """Address VGPR of the u-th LDS read of a group or batch: two, used in
    turn (a read takes its address when it issues, so the next-but-one
    address may overwrite it)."""
    return AV + u % 2


def e(x):
    L.append(x)


class Body:
    """Code of one parity p: bank p, LDS buffer %[vb<p>], labels suffixed p."""

    def __init__(self, p):
        self.p = p
        self.A, self.B = BANKS[p]
        self.base = "%[vb0]"   # list offsets include the buffer

    # ---- D reads and their widening ----
    @staticmethod
    def dread(d, a):
        """The LDS reads of an entry's 4 D values (columns 0..3 of the lane)
        into v[d..d+3] from the address in v{a}.  BF16: the 8 bytes into
        v[d+2], v[d+3] (two values per dword, low half first), and column 0
        again with a 16-bit read into the high half of v[d], whose low half
        stays zero: its exact FP32 value."""
        if BF16:
            e(f"ds_read_b64 v[{d+2}:{d+3}], v{a}")
            e(f"ds_read_u16_d16_hi v{d}, v{a}")
        else:
            e(f"ds_read_b128 v[{d}:{d+3}], v{a}")

    @staticmethod
    def widen(d):
        """BF16: columns 1..3 of an entry widened exactly to FP32 (the BF16
        bits as the high half) from the pairs in v[d+2], v[d+3]; column 0 came
        widened from its 16-bit read.  Outside VGPR-index mode, which would
        offset the destinations."""
        if BF16:
            e(f"v_and_b32_e32 v{d+1}, 0xffff0000, v{d+2}")
            e(f"v_lshlrev_b32_e32 v{d+2}, 16, v{d+3}")
            e(f"v_and_b32_e32 v{d+3}, 0xffff0000, v{d+3}")

    # ---- batch loop (entries from 16 on) ----
    def reads(self, S, k):
        for u in range(k):
            e(f"v_mad_u32_u16 v{av(u)}, s{S[2*u]}, 16, {self.base} op_sel:[1,0,0,0]")
            self.dread(DV + 4 * u, av(u))

    @staticmethod
    def entry(S, u):
        for c in range(4):
            e(f"v_fma_f32 v{R*c}, s{S[2*u+1]}, v{DV+4*u+c}, v{R*c}")

    def full_steps(self, S):
        for h in range(2):
            for q in range(4):
                self.widen(DV + 4 * (4 * h + q))
            e(f"s_set_gpr_idx_on s{S[8*h]}, gpr_idx(SRC2,DST)")
            for q in range(4):
                if q:
                    e(f"s_set_gpr_idx_idx s{S[2*(4*h+q)]}")
                self.entry(S, 4 * h + q)
            e("s_set_gpr_idx_off")

    def tail(self, S, t):
        p = self.p
        e(f"v_mad_u32_u16 v{av(0)}, s{S[0]}, 16, {self.base} op_sel:[1,0,0,0]")
        self.dread(DV, av(0))
        for u in range(1, 7):
            e(f"s_cmp_gt_u32 s{SREM}, {u}")
            e(f"s_cbranch_scc0 .Lgw{t}{p}_%=")
            e(f"v_mad_u32_u16 v{av(u)}, s{S[2*u]}, 16, {self.base} op_sel:[1,0,0,0]")
            self.dread(DV + 4 * u, av(u))
        e(f".Lgw{t}{p}_%=:")
        e("s_waitcnt lgkmcnt(0)")
        if BF16:
            self.widen(DV)
            for u in range(1, 7):
                e(f"s_cmp_gt_u32 s{SREM}, {u}")
                e(f"s_cbranch_scc0 .Lgv{t}{p}_%=")
                self.widen(DV + 4 * u)
            e(f".Lgv{t}{p}_%=:")
        e(f"s_set_gpr_idx_on s{S[0]}, gpr_idx(SRC2,DST)")
        self.entry(S, 0)
        for u in range(1, 7):
            e(f"s_cmp_gt_u32 s{SREM}, {u}")
            e(f"s_cbranch_scc0 .Lge{t}{p}_%=")
            e(f"s_set_gpr_idx_idx s{S[2*u]}")
            self.entry(S, u)
        e(f".Lge{t}{p}_%=:")
        e("s_set_gpr_idx_off")
        e(f"s_branch .Lce{p}_%=")

    def batch(self, S, N, t, nxt):
        p = self.p
        e(f".Lb{t}{p}_%=:")
        e(f"s_cmp_lt_u32 s{SREM}, 8")
        e(f"s_cbranch_scc1 .Lt{t}{p}_%=")
        e(f"s_cmp_ge_u32 s{SREM}, {PRIO_BATCH}")
        e(f"s_cbranch_scc0 .Lp{t}{p}_%=")
        e("s_setprio 3")
        e(f"s_branch .Lq{t}{p}_%=")
        e(f".Lp{t}{p}_%=:")
        e("s_setprio 2")
        e(f".Lq{t}{p}_%=:")
        self.reads(S, 8)
        e("s_waitcnt lgkmcnt(0)")
        e(f"s_cmp_le_u32 s{SREM}, 8")
        e(f"s_cbranch_scc1 .Ln{t}{p}_%=")
        e(f"s_load_dwordx16 s[{N[0]}:{N[15]}], %[pbase], s{SOFF}")
        e(f".Ln{t}{p}_%=:")
        self.full_steps(S)
        e(f"s_sub_u32 s{SREM}, s{SREM}, 8")
        e(f"s_add_u32 s{SOFF}, s{SOFF}, 64")
        e(f"s_cmp_eq_u32 s{SREM}, 0")
        e(f"s_cbranch_scc1 .Lce{p}_%=")
        e("s_waitcnt lgkmcnt(0)")
        if nxt:
            e(f"s_branch .Lb{nxt}{p}_%=")

    # ---- rolling phase (entries 0..15) ----
    def group_entry(self, g, q):
        """(SGPR of x, SGPR of the value) of entry q of group g (0..15 overall)."""
        S = self.A if g < 2 else self.B
        u = 4 * (g % 2) + q
        return S[2 * u], S[2 * u + 1]

    def group_reads(self, g):
        b = g % 2
        for q in range(4):
            sx, _ = self.group_entry(g, q)
            e(f"v_mad_u32_u16 v{av(q)}, s{sx}, 16, {self.base} op_sel:[1,0,0,0]")
            self.dread(G_D[b] + 4 * q, av(q))

    def group_entry_steps(self, g, q):
        _, sv = self.group_entry(g, q)
        b = g % 2
        for c in range(4):
            e(f"v_fma_f32 v{R*c}, s{sv}, v{G_D[b]+4*q+c}, v{R*c}")

    def group_steps_full(self, g):
        for q in range(4):
            self.widen(G_D[g % 2] + 4 * q)
        sx, _ = self.group_entry(g, 0)
        e(f"s_set_gpr_idx_on s{sx}, gpr_idx(SRC2,DST)")
        for q in range(4):
            if q:
                e(f"s_set_gpr_idx_idx s{self.group_entry(g, q)[0]}")
            self.group_entry_steps(g, q)
        e("s_set_gpr_idx_off")

    def group_steps_guarded(self, g, tag):
        # s{SCNT} = valid entries of this group (1..4)
        if BF16:
            self.widen(G_D[g % 2])
            for q in range(1, 4):
                e(f"s_cmp_gt_u32 s{SCNT}, {q}")
                e(f"s_cbranch_scc0 .Lgv{tag}{self.p}_%=")
                self.widen(G_D[g % 2] + 4 * q)
            e(f".Lgv{tag}{self.p}_%=:")
        sx, _ = self.group_entry(g, 0)
        e(f"s_set_gpr_idx_on s{sx}, gpr_idx(SRC2,DST)")
        self.group_entry_steps(g, 0)
        for q in range(1, 4):
            e(f"s_cmp_gt_u32 s{SCNT}, {q}")
            e(f"s_cbranch_scc0 .Lgg{tag}{self.p}_%=")
            e(f"s_set_gpr_idx_idx s{self.group_entry(g, q)[0]}")
            self.group_entry_steps(g, q)
        e(f".Lgg{tag}{self.p}_%=:")
        e("s_set_gpr_idx_off")

    def group_reads_guarded(self, g, tag):
        # s{SCNT} = valid entries of the group (1..3): reads for those only
        b = g % 2
        for q in range(4):
            sx, _ = self.group_entry(g, q)
            if q:
                e(f"s_cmp_gt_u32 s{SCNT}, {q}")
                e(f"s_cbranch_scc0 .L{tag}{self.p}_%=")
            e(f"v_mad_u32_u16 v{av(q)}, s{sx}, 16, {self.base} op_sel:[1,0,0,0]")
            self.dread(G_D[b] + 4 * q, av(q))
        e(f".L{tag}{self.p}_%=:")

    def roll16(self):
        p = self.p
        e(f"s_cmp_ge_u32 s{SREM}, 4")
        e(f"s_cbranch_scc1 .Lf0{p}_%=")
        e(f"s_mov_b32 s{SCNT}, s{SREM}")
        self.group_reads_guarded(0, "p0r")
        e("s_waitcnt lgkmcnt(0)")
        self.group_steps_guarded(0, "p0s")
        e(f"s_branch .Lce{p}_%=")
        e(f".Lf0{p}_%=:")
        self.group_reads(0)
        for g in range(4):
            # group g is full and its reads are in flight
            if g == 3:
                e("s_waitcnt lgkmcnt(0)")
                self.group_steps_full(3)
                break
            e(f"s_cmp_ge_u32 s{SREM}, {4*g+8}")
            e(f"s_cbranch_scc0 .Lq{g}{p}_%=")
            self.group_reads(g + 1)
            # group g is in once at most the next group's reads are out
            e(f"s_waitcnt lgkmcnt({4 * RPE})")
            self.group_steps_full(g)
            e(f"s_branch .Lf{g+1}{p}_%=")
            e(f".Lq{g}{p}_%=:")
            e(f"s_cmp_gt_u32 s{SREM}, {4*g+4}")
            e(f"s_cbranch_scc0 .Lz{g}{p}_%=")
            e(f"s_sub_u32 s{SCNT}, s{SREM}, {4*g+4}")
            self.group_reads_guarded(g + 1, f"p{g+1}r")
            e("s_waitcnt lgkmcnt(0)")
            self.group_steps_full(g)
            self.group_steps_guarded(g + 1, f"p{g+1}s")
            e(f"s_branch .Lce{p}_%=")
            e(f".Lz{g}{p}_%=:")
            e("s_waitcnt lgkmcnt(0)")
            self.group_steps_full(g)
            e(f"s_branch .Lce{p}_%=")
            e(f".Lf{g+1}{p}_%=:")

    # ---- the chunk ----
    @staticmethod
    def load_first16(A, B, skip):
        """Scalar loads of the first 16 entries of the chunk whose list is at
        %[pbase] + s{PBN} (s{SNN} entries) into A and B."""
        e(f"s_cmp_eq_u32 s{SNN}, 0")
        e(f"s_cbranch_scc1 {skip}")
        e(f"s_load_dwordx16 s[{A[0]}:{A[15]}], %[pbase], s{PBN}")
        e(f"s_cmp_le_u32 s{SNN}, 8")
        e(f"s_cbranch_scc1 {skip}")
        e(f"s_add_u32 s{ST}, s{PBN}, 0x40")
        e(f"s_load_dwordx16 s[{B[0]}:{B[15]}], %[pbase], s{ST}")

    def stage(self):
        p = self.p
        # the next chunk goes to the other buffer
        dst = "s_mov_b32 m0, %[sdst]" if p == 1 else f"s_add_u32 m0, %[sdst], {hex(CHUNKB)}"
        e("s_cmp_le_u32 %[sso], %[lim]")
        e(f"s_cbranch_scc0 .Lsg{p}_%=")
        e(dst)
        e(f"s_mov_b32 s{ST}, %[sso]")
        for i in range(15):
            e(f"buffer_load_dwordx4 %[voff], %[rsrc], s{ST} offen lds")
            if i < 14:
                e("s_add_u32 m0, m0, 0x400")
                e(f"s_add_u32 s{ST}, s{ST}, %[rowb]")
        e(f"s_branch .Lsd{p}_%=")
        e(f".Lsg{p}_%=:")
        e("s_cmp_eq_u32 %[kb], 0")
        e(f"s_cbranch_scc1 .Lsd{p}_%=")
        e(dst)
        e(f"s_mov_b32 s{ST}, %[sso]")
        for i in range(15):
            e(f"s_cmp_lt_u32 s{ST}, %[kb]")
            e(f"s_cbranch_scc0 .Lsd{p}_%=")
            if BF16:
                # a load holds two D rows (lanes 32..63 the second): when
                # the second is past K those lanes take %[voffz], out of
                # the buffer's range, and read zeros
                e(f"s_add_u32 s{SCNT}, s{ST}, %[row1]")
                e(f"s_cmp_lt_u32 s{SCNT}, %[kb]")
                e(f"s_cbranch_scc0 .Lsz{i}{p}_%=")
                e(f"buffer_load_dwordx4 %[voff], %[rsrc], s{ST} offen lds")
                e(f"s_branch .Lsn{i}{p}_%=")
                e(f".Lsz{i}{p}_%=:")
                e(f"buffer_load_dwordx4 %[voffz], %[rsrc], s{ST} offen lds")
                e(f".Lsn{i}{p}_%=:")
            else:
                e(f"buffer_load_dwordx4 %[voff], %[rsrc], s{ST} offen lds")
            e("s_add_u32 m0, m0, 0x400")
            e(f"s_add_u32 s{ST}, s{ST}, %[rowb]")
        e(f".Lsd{p}_%=:")
        e("s_add_u32 %[sso], %[sso], %[tkrowb]")

    def emit(self):
        p = self.p
        A2, B2 = BANKS[1 - p]
        e(f".Lc{p}_%=:")
        # this chunk's list: its first 16 entries are in bank p (or on the way)
        e(f"s_mov_b32 s{SREM}, s{SNN}")
        e(f"s_add_u32 s{SOFF}, s{PBN}, 0x80")
        e("s_cmp_eq_u32 %[kb], 0")
        e(f"s_cbranch_scc1 .Lw{p}_%=")
        e("s_waitcnt vmcnt(0)")
        e(f".Lw{p}_%=:")
        e("s_barrier")
        e("s_waitcnt lgkmcnt(0)")
        # the next chunk's bounds and first 16 entries (other bank)
        e(f"s_add_u32 s{ST}, %[sli], 1")
        e(f"s_cmp_lt_u32 s{ST}, %[liend]")
        e(f"s_cbranch_scc0 .Lnp{p}_%=")
        e(f"v_readlane_b32 s{PBN}, %[vpbo], s{ST}")
        e(f"v_readlane_b32 s{SNN}, %[vpn], s{ST}")
        self.load_first16(A2, B2, f".Lnp{p}_%=")
        e(f".Lnp{p}_%=:")
        self.stage()
        e(f"s_cmp_eq_u32 s{SREM}, 0")
        e(f"s_cbranch_scc1 .Lce{p}_%=")
        e(f"s_cmp_ge_u32 s{SREM}, {PRIO_CHUNK}")
        e(f"s_cbranch_scc0 .Lpr2{p}_%=")
        e("s_setprio 3")
        e(f"s_branch .Lpr{p}_%=")
        e(f".Lpr2{p}_%=:")
        e("s_setprio 2")
        e(f".Lpr{p}_%=:")
        self.roll16()
        # entries from 16 on: the batch loop, first batch into A
        e(f"s_cmp_le_u32 s{SREM}, 16")
        e(f"s_cbranch_scc1 .Lce{p}_%=")
        e(f"s_sub_u32 s{SREM}, s{SREM}, 16")
        e(f"s_load_dwordx16 s[{self.A[0]}:{self.A[15]}], %[pbase], s{SOFF}")
        e(f"s_add_u32 s{SOFF}, s{SOFF}, 0x40")
        e("s_waitcnt lgkmcnt(0)")
        self.batch(self.A, self.B, "a", None)      # falls through to the other set
        self.batch(self.B, self.A, "b", "a")
        e(f".Lta{p}_%=:")
        self.tail(self.A, "a")
        e(f".Ltb{p}_%=:")
        self.tail(self.B, "b")
        e(f".Lce{p}_%=:")
        e("s_setprio 0")
        e("s_add_u32 %[sli], %[sli], 1")
        e("s_cmp_lt_u32 %[sli], %[liend]")
        e("s_cbranch_scc0 .Lx_%=")


def generate(rows, bf16=False):
    """The asm lines of the segment block for `rows` rows per wave."""
    configure(rows, bf16)
    L.clear()
    body0, body1 = Body(0), Body(1)
    # prologue: the segment's first chunk, its entries loaded before its barrier
    e(f"s_mov_b32 s{SM0}, m0")
    if BF16:
        # column 0 registers: 16-bit reads write their high halves only
        for i in range(0, 32, 4):
            e(f"v_mov_b32 v{DV + i}, 0")
    e("s_nop 4")  # %[sli] may have just been written by a vector instruction
    e(f"v_readlane_b32 s{PBN}, %[vpbo], %[sli]")
    e(f"v_readlane_b32 s{SNN}, %[vpn], %[sli]")
    e("s_cmp_eq_u32 %[par], 0")
    e("s_cbranch_scc0 .Lo1_%=")
    Body.load_first16(*BANKS[0], ".Lc0_%=")
    e("s_branch .Lc0_%=")
    e(".Lo1_%=:")
    Body.load_first16(*BANKS[1], ".Lc1_%=")
    e("s_branch .Lc1_%=")
    body0.emit()    # falls through to the odd body
    body1.emit()
    e("s_branch .Lc0_%=")
    e(".Lx_%=:")
    e(f"s_mov_b32 m0, s{SM0}")
    return list(L)


def main():
    print("// Generated by gen_panel_asm.py (same directory); see its docstring.")
    print("// A segment of chunks of the FP32 panel multiply for one wave: per chunk the")
    print("// barrier, the next chunk's entry loads and staging, and the entry loop.")
    print("#pragma once")
    for rows, bf16 in [(r, False) for r in ROWS] + [(r, True) for r in ROWS_BF16]:
        lines = generate(rows, bf16)
        tag = f"{'BF16_' if bf16 else ''}R{rows}"
        print(f"// {'BF16' if bf16 else 'FP32'} D, {rows} rows per wave: accumulators "
              f"v0..v{4 * rows - 1}, component c of row slot q")
        print(f"// in v[{rows} c + q]; scratch v{4 * rows}..v{4 * rows + NSCR - 1}")
        print(f"#define GCOO_PANEL_SEGMENT_ASM_{tag} \\")
        for i, x in enumerate(lines):
            end = " \\" if i + 1 < len(lines) else ""
            print(f'    "{x}\\n\\t"{end}')
        print(f"#define GCOO_PANEL_SEGMENT_CLOBBERS_{tag} \\")
        regs = ([f'"v{i}"' for i in range(4 * rows, 4 * rows + NSCR)] +
                [f'"s{i}"' for i in range(24, 31)] +
                [f'"s{i}"' for i in range(36, 100)] + ['"scc"', '"memory"'])
        for i in range(0, len(regs), 12):
            end = ", \\" if i + 12 < len(regs) else ""
            print("    " + ", ".join(regs[i:i+12]) + end)
        if bf16:
            continue
        # f32x8 blocks of accumulators, f32x4 blocks for an odd row count
        w = 8 if rows % 2 == 0 else 4
        ops = [f'"+{{v[{w * i}:{w * i + w - 1}]}}"(a[{i}])' for i in range(4 * rows // w)]
        print(f"#define GCOO_PANEL_ACC_OPERANDS_R{rows}(a) \\")
        for i in range(0, len(ops), 4):
            end = ", \\" if i + 4 < len(ops) else ""
            print("    " + ", ".join(ops[i:i+4]) + end)


if __name__ == "__main__":
    main()

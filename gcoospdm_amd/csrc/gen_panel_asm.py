#!/usr/bin/env python3
"""Writes gcoo_panel_asm.h: one chunk of the FP32 panel multiply
(spdm_panel_kernel, gcoo_spdm.hip) for one wave, as one inline-asm body --
the chunk barrier, the staging of the next chunk and the entry loop.

    python3 gen_panel_asm.py > gcoo_panel_asm.h

In order:
  * the scalar loads of the chunk's first 16 list entries (two sets of 8),
    issued before the barrier so that their latency passes while the wave
    waits there;
  * the staging waves wait for their loads of this chunk (vmcnt(0)), then the
    workgroup barrier: chunk c is in LDS and every wave is done with the
    buffer of chunk c - 1;
  * the staging waves issue their %[nld] buffer-to-LDS loads of chunk c + 1
    into that buffer (unrolled; LDS destination %[dst] + 1 KB per load, D row
    offset %[soff] + %[rowb] per load, lane offset %[voff]);
  * entries 0..15 in groups of 4 over two D buffers: the LDS reads of group
    g + 1 are issued before the steps of group g, and lgkmcnt(4) waits for
    group g only (no scalar load is outstanding then, and LDS reads complete
    in order).  A last group of 1..3 entries reads and steps only its valid
    entries, after lgkmcnt(0);
  * entries from 16 on: the batch loop, 8 entries per batch.  Per batch: 8
    LDS reads of the entries' D rows (16 bytes per lane); the wait for them
    (LDS only); the scalar load of the next batch's 8 entries, issued now so
    that it runs under this batch's steps (issued any earlier, the wait for
    the LDS reads would also wait for it: both count in lgkmcnt); 8 steps;
    the wait for the next batch's entries.  Two register sets alternate
    between batches.  The last 1..7 entries take one guarded round.
Entry u is two words in SGPRs: x = (LDS byte offset of its D row << 8) | row
slot, and the FP32 bits of its value.  A step is acc[slot] = fmaf(value, d,
acc[slot]) for the 4 columns, in VGPR-index mode (s_set_gpr_idx_on/_idx take
the slot from bits 7:0 of x).  The wave's issue priority is set from the
entries it has left (3 from 12 on before the groups and from 16 on before a
batch, else 2).
Entries are the wave's rows in row order with k ascending, so every row's
chain is unchanged.

Operands: %[lp] SGPR pair, address of the first entry; %[n] SGPR, the number
of entries; %[base] VGPR, this lane's LDS byte address of the chunk buffer;
%[stg] SGPR, 1 on a staging wave; %[nld], %[dst], %[soff], %[rowb] SGPRs and
%[voff] VGPR as above; %[rsrc] the buffer resource of D (SGPR quad); the
accumulators v0..v63 (component c of row slot q in v[16c + q]).  Scratch:
v64..v103 and s40..s76.  M0 is saved and restored.  The scalar loads read up
to 7 entries past the end of the list, which the list capacity covers; no LDS
read is issued for an entry past the end.  Every
wave runs the body exactly once per chunk (one barrier each).
"""
A = list(range(40, 56))
B = list(range(56, 72))
ST, SREM, SOFF, SM0 = 72, 73, 74, 75
DV, AV = 64, 96
L = []


def e(x):
    L.append(x)


def reads(S, k):
    for u in range(k):
        e(f"s_lshr_b32 s{ST}, s{S[2*u]}, 8")
        e(f"v_add_u32_e32 v{AV+u}, s{ST}, %[base]")
        e(f"ds_read_b128 v[{DV+4*u}:{DV+4*u+3}], v{AV+u}")


def entry(S, u):
    for c in range(4):
        e(f"v_fma_f32 v{16*c}, s{S[2*u+1]}, v{DV+4*u+c}, v{16*c}")


def full_steps(S):
    for h in range(2):
        e(f"s_set_gpr_idx_on s{S[8*h]}, gpr_idx(SRC2,DST)")
        for q in range(4):
            if q:
                e(f"s_set_gpr_idx_idx s{S[2*(4*h+q)]}")
            entry(S, 4 * h + q)
        e("s_set_gpr_idx_off")


def tail(S, t):
    e(f"s_lshr_b32 s{ST}, s{S[0]}, 8")
    e(f"v_add_u32_e32 v{AV}, s{ST}, %[base]")
    e(f"ds_read_b128 v[{DV}:{DV+3}], v{AV}")
    for u in range(1, 7):
        e(f"s_cmp_gt_u32 s{SREM}, {u}")
        e(f"s_cbranch_scc0 .Lgw{t}_%=")
        e(f"s_lshr_b32 s{ST}, s{S[2*u]}, 8")
        e(f"v_add_u32_e32 v{AV+u}, s{ST}, %[base]")
        e(f"ds_read_b128 v[{DV+4*u}:{DV+4*u+3}], v{AV+u}")
    e(f".Lgw{t}_%=:")
    e("s_waitcnt lgkmcnt(0)")
    e(f"s_set_gpr_idx_on s{S[0]}, gpr_idx(SRC2,DST)")
    entry(S, 0)
    for u in range(1, 7):
        e(f"s_cmp_gt_u32 s{SREM}, {u}")
        e(f"s_cbranch_scc0 .Lge{t}_%=")
        e(f"s_set_gpr_idx_idx s{S[2*u]}")
        entry(S, u)
    e(f".Lge{t}_%=:")
    e("s_set_gpr_idx_off")
    e(".Lend_%=:" if t == "b" else "s_branch .Lend_%=")


# ---- first 16 entries: both sets loaded before the barrier, LDS reads of
# groups of 4 rolled over two D buffers with partial lgkmcnt waits ----
G_D = (64, 80)            # D buffers of the rolling phase
G_AV = (96, 100)          # their address VGPRs
SCNT = 76                 # valid entries of a partial group


def group_entry(g, q):
    """(SGPR of x, SGPR of the value) of entry q of group g (0..15 overall)."""
    S = A if g < 2 else B
    u = 4 * (g % 2) + q
    return S[2 * u], S[2 * u + 1]


def group_reads(g):
    b = g % 2
    for q in range(4):
        sx, _ = group_entry(g, q)
        e(f"s_lshr_b32 s{ST}, s{sx}, 8")
        e(f"v_add_u32_e32 v{G_AV[b]+q}, s{ST}, %[base]")
        e(f"ds_read_b128 v[{G_D[b]+4*q}:{G_D[b]+4*q+3}], v{G_AV[b]+q}")


def group_entry_steps(g, q):
    _, sv = group_entry(g, q)
    b = g % 2
    for c in range(4):
        e(f"v_fma_f32 v{16*c}, s{sv}, v{G_D[b]+4*q+c}, v{16*c}")


def group_steps_full(g):
    sx, _ = group_entry(g, 0)
    e(f"s_set_gpr_idx_on s{sx}, gpr_idx(SRC2,DST)")
    for q in range(4):
        if q:
            e(f"s_set_gpr_idx_idx s{group_entry(g, q)[0]}")
        group_entry_steps(g, q)
    e("s_set_gpr_idx_off")


def group_steps_guarded(g, tag):
    # s{SCNT} = valid entries of this group (1..4)
    sx, _ = group_entry(g, 0)
    e(f"s_set_gpr_idx_on s{sx}, gpr_idx(SRC2,DST)")
    group_entry_steps(g, 0)
    for q in range(1, 4):
        e(f"s_cmp_gt_u32 s{SCNT}, {q}")
        e(f"s_cbranch_scc0 .Lgg{tag}_%=")
        e(f"s_set_gpr_idx_idx s{group_entry(g, q)[0]}")
        group_entry_steps(g, q)
    e(f".Lgg{tag}_%=:")
    e("s_set_gpr_idx_off")


def group_reads_guarded(g, tag):
    # s{SCNT} = valid entries of the group (1..3): reads for those only
    b = g % 2
    for q in range(4):
        sx, _ = group_entry(g, q)
        if q:
            e(f"s_cmp_gt_u32 s{SCNT}, {q}")
            e(f"s_cbranch_scc0 .L{tag}_%=")
        e(f"s_lshr_b32 s{ST}, s{sx}, 8")
        e(f"v_add_u32_e32 v{G_AV[b]+q}, s{ST}, %[base]")
        e(f"ds_read_b128 v[{G_D[b]+4*q}:{G_D[b]+4*q+3}], v{G_AV[b]+q}")
    e(f".L{tag}_%=:")


def roll16():
    """Entries 0..min(n,16)-1 in groups of 4 over two D buffers: the next
    group's reads are issued before this group's steps, and lgkmcnt(4) waits
    for this group only (no scalar load is outstanding here).  A last group
    of 1..3 entries takes guarded reads and steps after lgkmcnt(0)."""
    e(f"s_cmp_ge_u32 s{SREM}, 4")
    e(f"s_cbranch_scc1 .Lf0_%=")
    e(f"s_mov_b32 s{SCNT}, s{SREM}")
    group_reads_guarded(0, "p0r")
    e("s_waitcnt lgkmcnt(0)")
    group_steps_guarded(0, "p0s")
    e("s_branch .Lrdone_%=")
    e(".Lf0_%=:")
    group_reads(0)
    for g in range(4):
        # group g is full and its reads are in flight
        if g == 3:
            e("s_waitcnt lgkmcnt(0)")
            group_steps_full(3)
            break
        e(f"s_cmp_ge_u32 s{SREM}, {4*g+8}")
        e(f"s_cbranch_scc0 .Lq{g}_%=")
        group_reads(g + 1)
        e("s_waitcnt lgkmcnt(4)")
        group_steps_full(g)
        e(f"s_branch .Lf{g+1}_%=")
        e(f".Lq{g}_%=:")
        e(f"s_cmp_gt_u32 s{SREM}, {4*g+4}")
        e(f"s_cbranch_scc0 .Lz{g}_%=")
        e(f"s_sub_u32 s{SCNT}, s{SREM}, {4*g+4}")
        group_reads_guarded(g + 1, f"p{g+1}r")
        e("s_waitcnt lgkmcnt(0)")
        group_steps_full(g)
        group_steps_guarded(g + 1, f"p{g+1}s")
        e("s_branch .Lrdone_%=")
        e(f".Lz{g}_%=:")
        e("s_waitcnt lgkmcnt(0)")
        group_steps_full(g)
        e("s_branch .Lrdone_%=")
        e(f".Lf{g+1}_%=:")
    e(".Lrdone_%=:")


def batch(S, N, t, nxt):
    e(f".Lb{t}_%=:")
    e(f"s_cmp_lt_u32 s{SREM}, 8")
    e(f"s_cbranch_scc1 .Lt{t}_%=")
    e(f"s_cmp_ge_u32 s{SREM}, 16")
    e(f"s_cbranch_scc0 .Lp{t}_%=")
    e("s_setprio 3")
    e(f"s_branch .Lq{t}_%=")
    e(f".Lp{t}_%=:")
    e("s_setprio 2")
    e(f".Lq{t}_%=:")
    reads(S, 8)
    e("s_waitcnt lgkmcnt(0)")
    e(f"s_cmp_le_u32 s{SREM}, 8")
    e(f"s_cbranch_scc1 .Ln{t}_%=")
    e(f"s_load_dwordx16 s[{N[0]}:{N[15]}], %[lp], s{SOFF}")
    e(f".Ln{t}_%=:")
    full_steps(S)
    e(f"s_sub_u32 s{SREM}, s{SREM}, 8")
    e(f"s_add_u32 s{SOFF}, s{SOFF}, 64")
    e(f"s_cmp_eq_u32 s{SREM}, 0")
    e(f"s_cbranch_scc1 .Lend_%=")
    e("s_waitcnt lgkmcnt(0)")
    if nxt:
        e(f"s_branch .Lb{nxt}_%=")


e(f"s_mov_b32 s{SM0}, m0")
e(f"s_mov_b32 s{SREM}, %[n]")
# the first 16 entries' loads go out before the barrier
e(f"s_cmp_eq_u32 s{SREM}, 0")
e(f"s_cbranch_scc1 .Lnl_%=")
e(f"s_load_dwordx16 s[{A[0]}:{A[15]}], %[lp], 0x0")
e(f"s_cmp_le_u32 s{SREM}, 8")
e(f"s_cbranch_scc1 .Lnl_%=")
e(f"s_load_dwordx16 s[{B[0]}:{B[15]}], %[lp], 0x40")
e(".Lnl_%=:")
# the staging waves' loads of this chunk have landed, every wave is done
# with the previous chunk
e("s_cmp_eq_u32 %[stg], 0")
e("s_cbranch_scc1 .Lsw_%=")
e("s_waitcnt vmcnt(0)")
e(".Lsw_%=:")
e("s_barrier")
# the next chunk's %[nld] loads (staging waves), unrolled
e(f"s_mov_b32 s{SOFF}, %[soff]")
e(f"s_mov_b32 s{ST}, %[dst]")
for i in range(15):
    e(f"s_cmp_le_u32 %[nld], {i}")
    e("s_cbranch_scc1 .Lsd_%=")
    e(f"s_mov_b32 m0, s{ST}")
    e(f"s_add_u32 s{ST}, s{ST}, 0x400")
    e("buffer_load_dwordx4 %[voff], %[rsrc], s" + str(SOFF) + " offen lds")
    e(f"s_add_u32 s{SOFF}, s{SOFF}, %[rowb]")
e(".Lsd_%=:")
e(f"s_cmp_eq_u32 s{SREM}, 0")
e(f"s_cbranch_scc1 .Lend_%=")
e("s_waitcnt lgkmcnt(0)")
e(f"s_cmp_ge_u32 s{SREM}, 12")
e(f"s_cbranch_scc0 .Lpr2_%=")
e("s_setprio 3")
e(f"s_branch .Lpr_%=")
e(".Lpr2_%=:")
e("s_setprio 2")
e(".Lpr_%=:")
roll16()
# entries from 16 on: the batch loop, first batch into A
e(f"s_cmp_le_u32 s{SREM}, 16")
e(f"s_cbranch_scc1 .Lend_%=")
e(f"s_sub_u32 s{SREM}, s{SREM}, 16")
e(f"s_load_dwordx16 s[{A[0]}:{A[15]}], %[lp], 0x80")
e(f"s_movk_i32 s{SOFF}, 0xc0")
e("s_waitcnt lgkmcnt(0)")
batch(A, B, "a", None)      # falls through to the other register set
batch(B, A, "b", "a")
e(".Lta_%=:")
tail(A, "a")
e(".Ltb_%=:")
tail(B, "b")
e(f"s_mov_b32 m0, s{SM0}")

print("// Generated by gen_panel_asm.py (same directory); see its docstring.")
print("// One chunk of the FP32 panel multiply for one wave: barrier, staging, entry loop.")
print("#pragma once")
print("#define GCOO_PANEL_CHUNK_ASM \\")
for i, x in enumerate(L):
    end = " \\" if i + 1 < len(L) else ""
    print(f'    "{x}\\n\\t"{end}')
print("#define GCOO_PANEL_CHUNK_CLOBBERS \\")
regs = [f'"v{i}"' for i in range(64, 104)] + [f'"s{i}"' for i in range(40, 77)] + ['"scc"', '"memory"']
for i in range(0, len(regs), 12):
    end = ", \\" if i + 12 < len(regs) else ""
    print("    " + ", ".join(regs[i:i+12]) + end)

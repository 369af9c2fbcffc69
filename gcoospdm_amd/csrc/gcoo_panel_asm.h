// Generated by gen_panel_asm.py (same directory); see its docstring.
// One chunk of the FP32 panel multiply for one wave: barrier, staging, entry loop.
#pragma once
#define GCOO_PANEL_CHUNK_ASM \
    "s_mov_b32 s75, m0\n\t" \
    "s_mov_b32 s73, %[n]\n\t" \
    "s_cmp_eq_u32 s73, 0\n\t" \
    "s_cbranch_scc1 .Lnl_%=\n\t" \
    "s_load_dwordx16 s[40:55], %[lp], 0x0\n\t" \
    "s_cmp_le_u32 s73, 8\n\t" \
    "s_cbranch_scc1 .Lnl_%=\n\t" \
    "s_load_dwordx16 s[56:71], %[lp], 0x40\n\t" \
    ".Lnl_%=:\n\t" \
    "s_cmp_eq_u32 %[stg], 0\n\t" \
    "s_cbranch_scc1 .Lsw_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lsw_%=:\n\t" \
    "s_barrier\n\t" \
    "s_mov_b32 s74, %[soff]\n\t" \
    "s_mov_b32 s72, %[dst]\n\t" \
    "s_cmp_le_u32 %[nld], 0\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 1\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 2\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 3\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 4\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 5\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 6\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 7\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 8\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 9\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 10\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 11\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 12\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 13\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    "s_cmp_le_u32 %[nld], 14\n\t" \
    "s_cbranch_scc1 .Lsd_%=\n\t" \
    "s_mov_b32 m0, s72\n\t" \
    "s_add_u32 s72, s72, 0x400\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s74 offen lds\n\t" \
    "s_add_u32 s74, s74, %[rowb]\n\t" \
    ".Lsd_%=:\n\t" \
    "s_cmp_eq_u32 s73, 0\n\t" \
    "s_cbranch_scc1 .Lend_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_ge_u32 s73, 12\n\t" \
    "s_cbranch_scc0 .Lpr2_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr_%=\n\t" \
    ".Lpr2_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr_%=:\n\t" \
    "s_cmp_ge_u32 s73, 4\n\t" \
    "s_cbranch_scc1 .Lf0_%=\n\t" \
    "s_mov_b32 s76, s73\n\t" \
    "s_lshr_b32 s72, s40, 8\n\t" \
    "v_add_u32_e32 v96, s72, %[base]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s76, 1\n\t" \
    "s_cbranch_scc0 .Lp0r_%=\n\t" \
    "s_lshr_b32 s72, s42, 8\n\t" \
    "v_add_u32_e32 v97, s72, %[base]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s76, 2\n\t" \
    "s_cbranch_scc0 .Lp0r_%=\n\t" \
    "s_lshr_b32 s72, s44, 8\n\t" \
    "v_add_u32_e32 v98, s72, %[base]\n\t" \
    "ds_read_b128 v[72:75], v98\n\t" \
    "s_cmp_gt_u32 s76, 3\n\t" \
    "s_cbranch_scc0 .Lp0r_%=\n\t" \
    "s_lshr_b32 s72, s46, 8\n\t" \
    "v_add_u32_e32 v99, s72, %[base]\n\t" \
    "ds_read_b128 v[76:79], v99\n\t" \
    ".Lp0r_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s40, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s41, v64, v0\n\t" \
    "v_fma_f32 v16, s41, v65, v16\n\t" \
    "v_fma_f32 v32, s41, v66, v32\n\t" \
    "v_fma_f32 v48, s41, v67, v48\n\t" \
    "s_cmp_gt_u32 s76, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v68, v0\n\t" \
    "v_fma_f32 v16, s43, v69, v16\n\t" \
    "v_fma_f32 v32, s43, v70, v32\n\t" \
    "v_fma_f32 v48, s43, v71, v48\n\t" \
    "s_cmp_gt_u32 s76, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v72, v0\n\t" \
    "v_fma_f32 v16, s45, v73, v16\n\t" \
    "v_fma_f32 v32, s45, v74, v32\n\t" \
    "v_fma_f32 v48, s45, v75, v48\n\t" \
    "s_cmp_gt_u32 s76, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v76, v0\n\t" \
    "v_fma_f32 v16, s47, v77, v16\n\t" \
    "v_fma_f32 v32, s47, v78, v32\n\t" \
    "v_fma_f32 v48, s47, v79, v48\n\t" \
    ".Lggp0s_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lrdone_%=\n\t" \
    ".Lf0_%=:\n\t" \
    "s_lshr_b32 s72, s40, 8\n\t" \
    "v_add_u32_e32 v96, s72, %[base]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_lshr_b32 s72, s42, 8\n\t" \
    "v_add_u32_e32 v97, s72, %[base]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_lshr_b32 s72, s44, 8\n\t" \
    "v_add_u32_e32 v98, s72, %[base]\n\t" \
    "ds_read_b128 v[72:75], v98\n\t" \
    "s_lshr_b32 s72, s46, 8\n\t" \
    "v_add_u32_e32 v99, s72, %[base]\n\t" \
    "ds_read_b128 v[76:79], v99\n\t" \
    "s_cmp_ge_u32 s73, 8\n\t" \
    "s_cbranch_scc0 .Lq0_%=\n\t" \
    "s_lshr_b32 s72, s48, 8\n\t" \
    "v_add_u32_e32 v100, s72, %[base]\n\t" \
    "ds_read_b128 v[80:83], v100\n\t" \
    "s_lshr_b32 s72, s50, 8\n\t" \
    "v_add_u32_e32 v101, s72, %[base]\n\t" \
    "ds_read_b128 v[84:87], v101\n\t" \
    "s_lshr_b32 s72, s52, 8\n\t" \
    "v_add_u32_e32 v102, s72, %[base]\n\t" \
    "ds_read_b128 v[88:91], v102\n\t" \
    "s_lshr_b32 s72, s54, 8\n\t" \
    "v_add_u32_e32 v103, s72, %[base]\n\t" \
    "ds_read_b128 v[92:95], v103\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s40, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s41, v64, v0\n\t" \
    "v_fma_f32 v16, s41, v65, v16\n\t" \
    "v_fma_f32 v32, s41, v66, v32\n\t" \
    "v_fma_f32 v48, s41, v67, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v68, v0\n\t" \
    "v_fma_f32 v16, s43, v69, v16\n\t" \
    "v_fma_f32 v32, s43, v70, v32\n\t" \
    "v_fma_f32 v48, s43, v71, v48\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v72, v0\n\t" \
    "v_fma_f32 v16, s45, v73, v16\n\t" \
    "v_fma_f32 v32, s45, v74, v32\n\t" \
    "v_fma_f32 v48, s45, v75, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v76, v0\n\t" \
    "v_fma_f32 v16, s47, v77, v16\n\t" \
    "v_fma_f32 v32, s47, v78, v32\n\t" \
    "v_fma_f32 v48, s47, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf1_%=\n\t" \
    ".Lq0_%=:\n\t" \
    "s_cmp_gt_u32 s73, 4\n\t" \
    "s_cbranch_scc0 .Lz0_%=\n\t" \
    "s_sub_u32 s76, s73, 4\n\t" \
    "s_lshr_b32 s72, s48, 8\n\t" \
    "v_add_u32_e32 v100, s72, %[base]\n\t" \
    "ds_read_b128 v[80:83], v100\n\t" \
    "s_cmp_gt_u32 s76, 1\n\t" \
    "s_cbranch_scc0 .Lp1r_%=\n\t" \
    "s_lshr_b32 s72, s50, 8\n\t" \
    "v_add_u32_e32 v101, s72, %[base]\n\t" \
    "ds_read_b128 v[84:87], v101\n\t" \
    "s_cmp_gt_u32 s76, 2\n\t" \
    "s_cbranch_scc0 .Lp1r_%=\n\t" \
    "s_lshr_b32 s72, s52, 8\n\t" \
    "v_add_u32_e32 v102, s72, %[base]\n\t" \
    "ds_read_b128 v[88:91], v102\n\t" \
    "s_cmp_gt_u32 s76, 3\n\t" \
    "s_cbranch_scc0 .Lp1r_%=\n\t" \
    "s_lshr_b32 s72, s54, 8\n\t" \
    "v_add_u32_e32 v103, s72, %[base]\n\t" \
    "ds_read_b128 v[92:95], v103\n\t" \
    ".Lp1r_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s40, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s41, v64, v0\n\t" \
    "v_fma_f32 v16, s41, v65, v16\n\t" \
    "v_fma_f32 v32, s41, v66, v32\n\t" \
    "v_fma_f32 v48, s41, v67, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v68, v0\n\t" \
    "v_fma_f32 v16, s43, v69, v16\n\t" \
    "v_fma_f32 v32, s43, v70, v32\n\t" \
    "v_fma_f32 v48, s43, v71, v48\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v72, v0\n\t" \
    "v_fma_f32 v16, s45, v73, v16\n\t" \
    "v_fma_f32 v32, s45, v74, v32\n\t" \
    "v_fma_f32 v48, s45, v75, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v76, v0\n\t" \
    "v_fma_f32 v16, s47, v77, v16\n\t" \
    "v_fma_f32 v32, s47, v78, v32\n\t" \
    "v_fma_f32 v48, s47, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s48, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s49, v80, v0\n\t" \
    "v_fma_f32 v16, s49, v81, v16\n\t" \
    "v_fma_f32 v32, s49, v82, v32\n\t" \
    "v_fma_f32 v48, s49, v83, v48\n\t" \
    "s_cmp_gt_u32 s76, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v84, v0\n\t" \
    "v_fma_f32 v16, s51, v85, v16\n\t" \
    "v_fma_f32 v32, s51, v86, v32\n\t" \
    "v_fma_f32 v48, s51, v87, v48\n\t" \
    "s_cmp_gt_u32 s76, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s_%=\n\t" \
    "s_set_gpr_idx_idx s52\n\t" \
    "v_fma_f32 v0, s53, v88, v0\n\t" \
    "v_fma_f32 v16, s53, v89, v16\n\t" \
    "v_fma_f32 v32, s53, v90, v32\n\t" \
    "v_fma_f32 v48, s53, v91, v48\n\t" \
    "s_cmp_gt_u32 s76, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v92, v0\n\t" \
    "v_fma_f32 v16, s55, v93, v16\n\t" \
    "v_fma_f32 v32, s55, v94, v32\n\t" \
    "v_fma_f32 v48, s55, v95, v48\n\t" \
    ".Lggp1s_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lrdone_%=\n\t" \
    ".Lz0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s40, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s41, v64, v0\n\t" \
    "v_fma_f32 v16, s41, v65, v16\n\t" \
    "v_fma_f32 v32, s41, v66, v32\n\t" \
    "v_fma_f32 v48, s41, v67, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v68, v0\n\t" \
    "v_fma_f32 v16, s43, v69, v16\n\t" \
    "v_fma_f32 v32, s43, v70, v32\n\t" \
    "v_fma_f32 v48, s43, v71, v48\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v72, v0\n\t" \
    "v_fma_f32 v16, s45, v73, v16\n\t" \
    "v_fma_f32 v32, s45, v74, v32\n\t" \
    "v_fma_f32 v48, s45, v75, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v76, v0\n\t" \
    "v_fma_f32 v16, s47, v77, v16\n\t" \
    "v_fma_f32 v32, s47, v78, v32\n\t" \
    "v_fma_f32 v48, s47, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lrdone_%=\n\t" \
    ".Lf1_%=:\n\t" \
    "s_cmp_ge_u32 s73, 12\n\t" \
    "s_cbranch_scc0 .Lq1_%=\n\t" \
    "s_lshr_b32 s72, s56, 8\n\t" \
    "v_add_u32_e32 v96, s72, %[base]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_lshr_b32 s72, s58, 8\n\t" \
    "v_add_u32_e32 v97, s72, %[base]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_lshr_b32 s72, s60, 8\n\t" \
    "v_add_u32_e32 v98, s72, %[base]\n\t" \
    "ds_read_b128 v[72:75], v98\n\t" \
    "s_lshr_b32 s72, s62, 8\n\t" \
    "v_add_u32_e32 v99, s72, %[base]\n\t" \
    "ds_read_b128 v[76:79], v99\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s48, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s49, v80, v0\n\t" \
    "v_fma_f32 v16, s49, v81, v16\n\t" \
    "v_fma_f32 v32, s49, v82, v32\n\t" \
    "v_fma_f32 v48, s49, v83, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v84, v0\n\t" \
    "v_fma_f32 v16, s51, v85, v16\n\t" \
    "v_fma_f32 v32, s51, v86, v32\n\t" \
    "v_fma_f32 v48, s51, v87, v48\n\t" \
    "s_set_gpr_idx_idx s52\n\t" \
    "v_fma_f32 v0, s53, v88, v0\n\t" \
    "v_fma_f32 v16, s53, v89, v16\n\t" \
    "v_fma_f32 v32, s53, v90, v32\n\t" \
    "v_fma_f32 v48, s53, v91, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v92, v0\n\t" \
    "v_fma_f32 v16, s55, v93, v16\n\t" \
    "v_fma_f32 v32, s55, v94, v32\n\t" \
    "v_fma_f32 v48, s55, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf2_%=\n\t" \
    ".Lq1_%=:\n\t" \
    "s_cmp_gt_u32 s73, 8\n\t" \
    "s_cbranch_scc0 .Lz1_%=\n\t" \
    "s_sub_u32 s76, s73, 8\n\t" \
    "s_lshr_b32 s72, s56, 8\n\t" \
    "v_add_u32_e32 v96, s72, %[base]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s76, 1\n\t" \
    "s_cbranch_scc0 .Lp2r_%=\n\t" \
    "s_lshr_b32 s72, s58, 8\n\t" \
    "v_add_u32_e32 v97, s72, %[base]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s76, 2\n\t" \
    "s_cbranch_scc0 .Lp2r_%=\n\t" \
    "s_lshr_b32 s72, s60, 8\n\t" \
    "v_add_u32_e32 v98, s72, %[base]\n\t" \
    "ds_read_b128 v[72:75], v98\n\t" \
    "s_cmp_gt_u32 s76, 3\n\t" \
    "s_cbranch_scc0 .Lp2r_%=\n\t" \
    "s_lshr_b32 s72, s62, 8\n\t" \
    "v_add_u32_e32 v99, s72, %[base]\n\t" \
    "ds_read_b128 v[76:79], v99\n\t" \
    ".Lp2r_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s48, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s49, v80, v0\n\t" \
    "v_fma_f32 v16, s49, v81, v16\n\t" \
    "v_fma_f32 v32, s49, v82, v32\n\t" \
    "v_fma_f32 v48, s49, v83, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v84, v0\n\t" \
    "v_fma_f32 v16, s51, v85, v16\n\t" \
    "v_fma_f32 v32, s51, v86, v32\n\t" \
    "v_fma_f32 v48, s51, v87, v48\n\t" \
    "s_set_gpr_idx_idx s52\n\t" \
    "v_fma_f32 v0, s53, v88, v0\n\t" \
    "v_fma_f32 v16, s53, v89, v16\n\t" \
    "v_fma_f32 v32, s53, v90, v32\n\t" \
    "v_fma_f32 v48, s53, v91, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v92, v0\n\t" \
    "v_fma_f32 v16, s55, v93, v16\n\t" \
    "v_fma_f32 v32, s55, v94, v32\n\t" \
    "v_fma_f32 v48, s55, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s56, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s57, v64, v0\n\t" \
    "v_fma_f32 v16, s57, v65, v16\n\t" \
    "v_fma_f32 v32, s57, v66, v32\n\t" \
    "v_fma_f32 v48, s57, v67, v48\n\t" \
    "s_cmp_gt_u32 s76, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v68, v0\n\t" \
    "v_fma_f32 v16, s59, v69, v16\n\t" \
    "v_fma_f32 v32, s59, v70, v32\n\t" \
    "v_fma_f32 v48, s59, v71, v48\n\t" \
    "s_cmp_gt_u32 s76, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v72, v0\n\t" \
    "v_fma_f32 v16, s61, v73, v16\n\t" \
    "v_fma_f32 v32, s61, v74, v32\n\t" \
    "v_fma_f32 v48, s61, v75, v48\n\t" \
    "s_cmp_gt_u32 s76, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v76, v0\n\t" \
    "v_fma_f32 v16, s63, v77, v16\n\t" \
    "v_fma_f32 v32, s63, v78, v32\n\t" \
    "v_fma_f32 v48, s63, v79, v48\n\t" \
    ".Lggp2s_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lrdone_%=\n\t" \
    ".Lz1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s48, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s49, v80, v0\n\t" \
    "v_fma_f32 v16, s49, v81, v16\n\t" \
    "v_fma_f32 v32, s49, v82, v32\n\t" \
    "v_fma_f32 v48, s49, v83, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v84, v0\n\t" \
    "v_fma_f32 v16, s51, v85, v16\n\t" \
    "v_fma_f32 v32, s51, v86, v32\n\t" \
    "v_fma_f32 v48, s51, v87, v48\n\t" \
    "s_set_gpr_idx_idx s52\n\t" \
    "v_fma_f32 v0, s53, v88, v0\n\t" \
    "v_fma_f32 v16, s53, v89, v16\n\t" \
    "v_fma_f32 v32, s53, v90, v32\n\t" \
    "v_fma_f32 v48, s53, v91, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v92, v0\n\t" \
    "v_fma_f32 v16, s55, v93, v16\n\t" \
    "v_fma_f32 v32, s55, v94, v32\n\t" \
    "v_fma_f32 v48, s55, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lrdone_%=\n\t" \
    ".Lf2_%=:\n\t" \
    "s_cmp_ge_u32 s73, 16\n\t" \
    "s_cbranch_scc0 .Lq2_%=\n\t" \
    "s_lshr_b32 s72, s64, 8\n\t" \
    "v_add_u32_e32 v100, s72, %[base]\n\t" \
    "ds_read_b128 v[80:83], v100\n\t" \
    "s_lshr_b32 s72, s66, 8\n\t" \
    "v_add_u32_e32 v101, s72, %[base]\n\t" \
    "ds_read_b128 v[84:87], v101\n\t" \
    "s_lshr_b32 s72, s68, 8\n\t" \
    "v_add_u32_e32 v102, s72, %[base]\n\t" \
    "ds_read_b128 v[88:91], v102\n\t" \
    "s_lshr_b32 s72, s70, 8\n\t" \
    "v_add_u32_e32 v103, s72, %[base]\n\t" \
    "ds_read_b128 v[92:95], v103\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s56, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s57, v64, v0\n\t" \
    "v_fma_f32 v16, s57, v65, v16\n\t" \
    "v_fma_f32 v32, s57, v66, v32\n\t" \
    "v_fma_f32 v48, s57, v67, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v68, v0\n\t" \
    "v_fma_f32 v16, s59, v69, v16\n\t" \
    "v_fma_f32 v32, s59, v70, v32\n\t" \
    "v_fma_f32 v48, s59, v71, v48\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v72, v0\n\t" \
    "v_fma_f32 v16, s61, v73, v16\n\t" \
    "v_fma_f32 v32, s61, v74, v32\n\t" \
    "v_fma_f32 v48, s61, v75, v48\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v76, v0\n\t" \
    "v_fma_f32 v16, s63, v77, v16\n\t" \
    "v_fma_f32 v32, s63, v78, v32\n\t" \
    "v_fma_f32 v48, s63, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf3_%=\n\t" \
    ".Lq2_%=:\n\t" \
    "s_cmp_gt_u32 s73, 12\n\t" \
    "s_cbranch_scc0 .Lz2_%=\n\t" \
    "s_sub_u32 s76, s73, 12\n\t" \
    "s_lshr_b32 s72, s64, 8\n\t" \
    "v_add_u32_e32 v100, s72, %[base]\n\t" \
    "ds_read_b128 v[80:83], v100\n\t" \
    "s_cmp_gt_u32 s76, 1\n\t" \
    "s_cbranch_scc0 .Lp3r_%=\n\t" \
    "s_lshr_b32 s72, s66, 8\n\t" \
    "v_add_u32_e32 v101, s72, %[base]\n\t" \
    "ds_read_b128 v[84:87], v101\n\t" \
    "s_cmp_gt_u32 s76, 2\n\t" \
    "s_cbranch_scc0 .Lp3r_%=\n\t" \
    "s_lshr_b32 s72, s68, 8\n\t" \
    "v_add_u32_e32 v102, s72, %[base]\n\t" \
    "ds_read_b128 v[88:91], v102\n\t" \
    "s_cmp_gt_u32 s76, 3\n\t" \
    "s_cbranch_scc0 .Lp3r_%=\n\t" \
    "s_lshr_b32 s72, s70, 8\n\t" \
    "v_add_u32_e32 v103, s72, %[base]\n\t" \
    "ds_read_b128 v[92:95], v103\n\t" \
    ".Lp3r_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s56, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s57, v64, v0\n\t" \
    "v_fma_f32 v16, s57, v65, v16\n\t" \
    "v_fma_f32 v32, s57, v66, v32\n\t" \
    "v_fma_f32 v48, s57, v67, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v68, v0\n\t" \
    "v_fma_f32 v16, s59, v69, v16\n\t" \
    "v_fma_f32 v32, s59, v70, v32\n\t" \
    "v_fma_f32 v48, s59, v71, v48\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v72, v0\n\t" \
    "v_fma_f32 v16, s61, v73, v16\n\t" \
    "v_fma_f32 v32, s61, v74, v32\n\t" \
    "v_fma_f32 v48, s61, v75, v48\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v76, v0\n\t" \
    "v_fma_f32 v16, s63, v77, v16\n\t" \
    "v_fma_f32 v32, s63, v78, v32\n\t" \
    "v_fma_f32 v48, s63, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s64, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s65, v80, v0\n\t" \
    "v_fma_f32 v16, s65, v81, v16\n\t" \
    "v_fma_f32 v32, s65, v82, v32\n\t" \
    "v_fma_f32 v48, s65, v83, v48\n\t" \
    "s_cmp_gt_u32 s76, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v84, v0\n\t" \
    "v_fma_f32 v16, s67, v85, v16\n\t" \
    "v_fma_f32 v32, s67, v86, v32\n\t" \
    "v_fma_f32 v48, s67, v87, v48\n\t" \
    "s_cmp_gt_u32 s76, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s_%=\n\t" \
    "s_set_gpr_idx_idx s68\n\t" \
    "v_fma_f32 v0, s69, v88, v0\n\t" \
    "v_fma_f32 v16, s69, v89, v16\n\t" \
    "v_fma_f32 v32, s69, v90, v32\n\t" \
    "v_fma_f32 v48, s69, v91, v48\n\t" \
    "s_cmp_gt_u32 s76, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v92, v0\n\t" \
    "v_fma_f32 v16, s71, v93, v16\n\t" \
    "v_fma_f32 v32, s71, v94, v32\n\t" \
    "v_fma_f32 v48, s71, v95, v48\n\t" \
    ".Lggp3s_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lrdone_%=\n\t" \
    ".Lz2_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s56, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s57, v64, v0\n\t" \
    "v_fma_f32 v16, s57, v65, v16\n\t" \
    "v_fma_f32 v32, s57, v66, v32\n\t" \
    "v_fma_f32 v48, s57, v67, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v68, v0\n\t" \
    "v_fma_f32 v16, s59, v69, v16\n\t" \
    "v_fma_f32 v32, s59, v70, v32\n\t" \
    "v_fma_f32 v48, s59, v71, v48\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v72, v0\n\t" \
    "v_fma_f32 v16, s61, v73, v16\n\t" \
    "v_fma_f32 v32, s61, v74, v32\n\t" \
    "v_fma_f32 v48, s61, v75, v48\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v76, v0\n\t" \
    "v_fma_f32 v16, s63, v77, v16\n\t" \
    "v_fma_f32 v32, s63, v78, v32\n\t" \
    "v_fma_f32 v48, s63, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lrdone_%=\n\t" \
    ".Lf3_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s64, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s65, v80, v0\n\t" \
    "v_fma_f32 v16, s65, v81, v16\n\t" \
    "v_fma_f32 v32, s65, v82, v32\n\t" \
    "v_fma_f32 v48, s65, v83, v48\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v84, v0\n\t" \
    "v_fma_f32 v16, s67, v85, v16\n\t" \
    "v_fma_f32 v32, s67, v86, v32\n\t" \
    "v_fma_f32 v48, s67, v87, v48\n\t" \
    "s_set_gpr_idx_idx s68\n\t" \
    "v_fma_f32 v0, s69, v88, v0\n\t" \
    "v_fma_f32 v16, s69, v89, v16\n\t" \
    "v_fma_f32 v32, s69, v90, v32\n\t" \
    "v_fma_f32 v48, s69, v91, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v92, v0\n\t" \
    "v_fma_f32 v16, s71, v93, v16\n\t" \
    "v_fma_f32 v32, s71, v94, v32\n\t" \
    "v_fma_f32 v48, s71, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    ".Lrdone_%=:\n\t" \
    "s_cmp_le_u32 s73, 16\n\t" \
    "s_cbranch_scc1 .Lend_%=\n\t" \
    "s_sub_u32 s73, s73, 16\n\t" \
    "s_load_dwordx16 s[40:55], %[lp], 0x80\n\t" \
    "s_movk_i32 s74, 0xc0\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba_%=:\n\t" \
    "s_cmp_lt_u32 s73, 8\n\t" \
    "s_cbranch_scc1 .Lta_%=\n\t" \
    "s_cmp_ge_u32 s73, 16\n\t" \
    "s_cbranch_scc0 .Lpa_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa_%=\n\t" \
    ".Lpa_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa_%=:\n\t" \
    "s_lshr_b32 s72, s40, 8\n\t" \
    "v_add_u32_e32 v96, s72, %[base]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_lshr_b32 s72, s42, 8\n\t" \
    "v_add_u32_e32 v97, s72, %[base]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_lshr_b32 s72, s44, 8\n\t" \
    "v_add_u32_e32 v98, s72, %[base]\n\t" \
    "ds_read_b128 v[72:75], v98\n\t" \
    "s_lshr_b32 s72, s46, 8\n\t" \
    "v_add_u32_e32 v99, s72, %[base]\n\t" \
    "ds_read_b128 v[76:79], v99\n\t" \
    "s_lshr_b32 s72, s48, 8\n\t" \
    "v_add_u32_e32 v100, s72, %[base]\n\t" \
    "ds_read_b128 v[80:83], v100\n\t" \
    "s_lshr_b32 s72, s50, 8\n\t" \
    "v_add_u32_e32 v101, s72, %[base]\n\t" \
    "ds_read_b128 v[84:87], v101\n\t" \
    "s_lshr_b32 s72, s52, 8\n\t" \
    "v_add_u32_e32 v102, s72, %[base]\n\t" \
    "ds_read_b128 v[88:91], v102\n\t" \
    "s_lshr_b32 s72, s54, 8\n\t" \
    "v_add_u32_e32 v103, s72, %[base]\n\t" \
    "ds_read_b128 v[92:95], v103\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s73, 8\n\t" \
    "s_cbranch_scc1 .Lna_%=\n\t" \
    "s_load_dwordx16 s[56:71], %[lp], s74\n\t" \
    ".Lna_%=:\n\t" \
    "s_set_gpr_idx_on s40, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s41, v64, v0\n\t" \
    "v_fma_f32 v16, s41, v65, v16\n\t" \
    "v_fma_f32 v32, s41, v66, v32\n\t" \
    "v_fma_f32 v48, s41, v67, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v68, v0\n\t" \
    "v_fma_f32 v16, s43, v69, v16\n\t" \
    "v_fma_f32 v32, s43, v70, v32\n\t" \
    "v_fma_f32 v48, s43, v71, v48\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v72, v0\n\t" \
    "v_fma_f32 v16, s45, v73, v16\n\t" \
    "v_fma_f32 v32, s45, v74, v32\n\t" \
    "v_fma_f32 v48, s45, v75, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v76, v0\n\t" \
    "v_fma_f32 v16, s47, v77, v16\n\t" \
    "v_fma_f32 v32, s47, v78, v32\n\t" \
    "v_fma_f32 v48, s47, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s48, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s49, v80, v0\n\t" \
    "v_fma_f32 v16, s49, v81, v16\n\t" \
    "v_fma_f32 v32, s49, v82, v32\n\t" \
    "v_fma_f32 v48, s49, v83, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v84, v0\n\t" \
    "v_fma_f32 v16, s51, v85, v16\n\t" \
    "v_fma_f32 v32, s51, v86, v32\n\t" \
    "v_fma_f32 v48, s51, v87, v48\n\t" \
    "s_set_gpr_idx_idx s52\n\t" \
    "v_fma_f32 v0, s53, v88, v0\n\t" \
    "v_fma_f32 v16, s53, v89, v16\n\t" \
    "v_fma_f32 v32, s53, v90, v32\n\t" \
    "v_fma_f32 v48, s53, v91, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v92, v0\n\t" \
    "v_fma_f32 v16, s55, v93, v16\n\t" \
    "v_fma_f32 v32, s55, v94, v32\n\t" \
    "v_fma_f32 v48, s55, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s73, s73, 8\n\t" \
    "s_add_u32 s74, s74, 64\n\t" \
    "s_cmp_eq_u32 s73, 0\n\t" \
    "s_cbranch_scc1 .Lend_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb_%=:\n\t" \
    "s_cmp_lt_u32 s73, 8\n\t" \
    "s_cbranch_scc1 .Ltb_%=\n\t" \
    "s_cmp_ge_u32 s73, 16\n\t" \
    "s_cbranch_scc0 .Lpb_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb_%=\n\t" \
    ".Lpb_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb_%=:\n\t" \
    "s_lshr_b32 s72, s56, 8\n\t" \
    "v_add_u32_e32 v96, s72, %[base]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_lshr_b32 s72, s58, 8\n\t" \
    "v_add_u32_e32 v97, s72, %[base]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_lshr_b32 s72, s60, 8\n\t" \
    "v_add_u32_e32 v98, s72, %[base]\n\t" \
    "ds_read_b128 v[72:75], v98\n\t" \
    "s_lshr_b32 s72, s62, 8\n\t" \
    "v_add_u32_e32 v99, s72, %[base]\n\t" \
    "ds_read_b128 v[76:79], v99\n\t" \
    "s_lshr_b32 s72, s64, 8\n\t" \
    "v_add_u32_e32 v100, s72, %[base]\n\t" \
    "ds_read_b128 v[80:83], v100\n\t" \
    "s_lshr_b32 s72, s66, 8\n\t" \
    "v_add_u32_e32 v101, s72, %[base]\n\t" \
    "ds_read_b128 v[84:87], v101\n\t" \
    "s_lshr_b32 s72, s68, 8\n\t" \
    "v_add_u32_e32 v102, s72, %[base]\n\t" \
    "ds_read_b128 v[88:91], v102\n\t" \
    "s_lshr_b32 s72, s70, 8\n\t" \
    "v_add_u32_e32 v103, s72, %[base]\n\t" \
    "ds_read_b128 v[92:95], v103\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s73, 8\n\t" \
    "s_cbranch_scc1 .Lnb_%=\n\t" \
    "s_load_dwordx16 s[40:55], %[lp], s74\n\t" \
    ".Lnb_%=:\n\t" \
    "s_set_gpr_idx_on s56, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s57, v64, v0\n\t" \
    "v_fma_f32 v16, s57, v65, v16\n\t" \
    "v_fma_f32 v32, s57, v66, v32\n\t" \
    "v_fma_f32 v48, s57, v67, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v68, v0\n\t" \
    "v_fma_f32 v16, s59, v69, v16\n\t" \
    "v_fma_f32 v32, s59, v70, v32\n\t" \
    "v_fma_f32 v48, s59, v71, v48\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v72, v0\n\t" \
    "v_fma_f32 v16, s61, v73, v16\n\t" \
    "v_fma_f32 v32, s61, v74, v32\n\t" \
    "v_fma_f32 v48, s61, v75, v48\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v76, v0\n\t" \
    "v_fma_f32 v16, s63, v77, v16\n\t" \
    "v_fma_f32 v32, s63, v78, v32\n\t" \
    "v_fma_f32 v48, s63, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s64, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s65, v80, v0\n\t" \
    "v_fma_f32 v16, s65, v81, v16\n\t" \
    "v_fma_f32 v32, s65, v82, v32\n\t" \
    "v_fma_f32 v48, s65, v83, v48\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v84, v0\n\t" \
    "v_fma_f32 v16, s67, v85, v16\n\t" \
    "v_fma_f32 v32, s67, v86, v32\n\t" \
    "v_fma_f32 v48, s67, v87, v48\n\t" \
    "s_set_gpr_idx_idx s68\n\t" \
    "v_fma_f32 v0, s69, v88, v0\n\t" \
    "v_fma_f32 v16, s69, v89, v16\n\t" \
    "v_fma_f32 v32, s69, v90, v32\n\t" \
    "v_fma_f32 v48, s69, v91, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v92, v0\n\t" \
    "v_fma_f32 v16, s71, v93, v16\n\t" \
    "v_fma_f32 v32, s71, v94, v32\n\t" \
    "v_fma_f32 v48, s71, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s73, s73, 8\n\t" \
    "s_add_u32 s74, s74, 64\n\t" \
    "s_cmp_eq_u32 s73, 0\n\t" \
    "s_cbranch_scc1 .Lend_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba_%=\n\t" \
    ".Lta_%=:\n\t" \
    "s_lshr_b32 s72, s40, 8\n\t" \
    "v_add_u32_e32 v96, s72, %[base]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s73, 1\n\t" \
    "s_cbranch_scc0 .Lgwa_%=\n\t" \
    "s_lshr_b32 s72, s42, 8\n\t" \
    "v_add_u32_e32 v97, s72, %[base]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s73, 2\n\t" \
    "s_cbranch_scc0 .Lgwa_%=\n\t" \
    "s_lshr_b32 s72, s44, 8\n\t" \
    "v_add_u32_e32 v98, s72, %[base]\n\t" \
    "ds_read_b128 v[72:75], v98\n\t" \
    "s_cmp_gt_u32 s73, 3\n\t" \
    "s_cbranch_scc0 .Lgwa_%=\n\t" \
    "s_lshr_b32 s72, s46, 8\n\t" \
    "v_add_u32_e32 v99, s72, %[base]\n\t" \
    "ds_read_b128 v[76:79], v99\n\t" \
    "s_cmp_gt_u32 s73, 4\n\t" \
    "s_cbranch_scc0 .Lgwa_%=\n\t" \
    "s_lshr_b32 s72, s48, 8\n\t" \
    "v_add_u32_e32 v100, s72, %[base]\n\t" \
    "ds_read_b128 v[80:83], v100\n\t" \
    "s_cmp_gt_u32 s73, 5\n\t" \
    "s_cbranch_scc0 .Lgwa_%=\n\t" \
    "s_lshr_b32 s72, s50, 8\n\t" \
    "v_add_u32_e32 v101, s72, %[base]\n\t" \
    "ds_read_b128 v[84:87], v101\n\t" \
    "s_cmp_gt_u32 s73, 6\n\t" \
    "s_cbranch_scc0 .Lgwa_%=\n\t" \
    "s_lshr_b32 s72, s52, 8\n\t" \
    "v_add_u32_e32 v102, s72, %[base]\n\t" \
    "ds_read_b128 v[88:91], v102\n\t" \
    ".Lgwa_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s40, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s41, v64, v0\n\t" \
    "v_fma_f32 v16, s41, v65, v16\n\t" \
    "v_fma_f32 v32, s41, v66, v32\n\t" \
    "v_fma_f32 v48, s41, v67, v48\n\t" \
    "s_cmp_gt_u32 s73, 1\n\t" \
    "s_cbranch_scc0 .Lgea_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v68, v0\n\t" \
    "v_fma_f32 v16, s43, v69, v16\n\t" \
    "v_fma_f32 v32, s43, v70, v32\n\t" \
    "v_fma_f32 v48, s43, v71, v48\n\t" \
    "s_cmp_gt_u32 s73, 2\n\t" \
    "s_cbranch_scc0 .Lgea_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v72, v0\n\t" \
    "v_fma_f32 v16, s45, v73, v16\n\t" \
    "v_fma_f32 v32, s45, v74, v32\n\t" \
    "v_fma_f32 v48, s45, v75, v48\n\t" \
    "s_cmp_gt_u32 s73, 3\n\t" \
    "s_cbranch_scc0 .Lgea_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v76, v0\n\t" \
    "v_fma_f32 v16, s47, v77, v16\n\t" \
    "v_fma_f32 v32, s47, v78, v32\n\t" \
    "v_fma_f32 v48, s47, v79, v48\n\t" \
    "s_cmp_gt_u32 s73, 4\n\t" \
    "s_cbranch_scc0 .Lgea_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v80, v0\n\t" \
    "v_fma_f32 v16, s49, v81, v16\n\t" \
    "v_fma_f32 v32, s49, v82, v32\n\t" \
    "v_fma_f32 v48, s49, v83, v48\n\t" \
    "s_cmp_gt_u32 s73, 5\n\t" \
    "s_cbranch_scc0 .Lgea_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v84, v0\n\t" \
    "v_fma_f32 v16, s51, v85, v16\n\t" \
    "v_fma_f32 v32, s51, v86, v32\n\t" \
    "v_fma_f32 v48, s51, v87, v48\n\t" \
    "s_cmp_gt_u32 s73, 6\n\t" \
    "s_cbranch_scc0 .Lgea_%=\n\t" \
    "s_set_gpr_idx_idx s52\n\t" \
    "v_fma_f32 v0, s53, v88, v0\n\t" \
    "v_fma_f32 v16, s53, v89, v16\n\t" \
    "v_fma_f32 v32, s53, v90, v32\n\t" \
    "v_fma_f32 v48, s53, v91, v48\n\t" \
    ".Lgea_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lend_%=\n\t" \
    ".Ltb_%=:\n\t" \
    "s_lshr_b32 s72, s56, 8\n\t" \
    "v_add_u32_e32 v96, s72, %[base]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s73, 1\n\t" \
    "s_cbranch_scc0 .Lgwb_%=\n\t" \
    "s_lshr_b32 s72, s58, 8\n\t" \
    "v_add_u32_e32 v97, s72, %[base]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s73, 2\n\t" \
    "s_cbranch_scc0 .Lgwb_%=\n\t" \
    "s_lshr_b32 s72, s60, 8\n\t" \
    "v_add_u32_e32 v98, s72, %[base]\n\t" \
    "ds_read_b128 v[72:75], v98\n\t" \
    "s_cmp_gt_u32 s73, 3\n\t" \
    "s_cbranch_scc0 .Lgwb_%=\n\t" \
    "s_lshr_b32 s72, s62, 8\n\t" \
    "v_add_u32_e32 v99, s72, %[base]\n\t" \
    "ds_read_b128 v[76:79], v99\n\t" \
    "s_cmp_gt_u32 s73, 4\n\t" \
    "s_cbranch_scc0 .Lgwb_%=\n\t" \
    "s_lshr_b32 s72, s64, 8\n\t" \
    "v_add_u32_e32 v100, s72, %[base]\n\t" \
    "ds_read_b128 v[80:83], v100\n\t" \
    "s_cmp_gt_u32 s73, 5\n\t" \
    "s_cbranch_scc0 .Lgwb_%=\n\t" \
    "s_lshr_b32 s72, s66, 8\n\t" \
    "v_add_u32_e32 v101, s72, %[base]\n\t" \
    "ds_read_b128 v[84:87], v101\n\t" \
    "s_cmp_gt_u32 s73, 6\n\t" \
    "s_cbranch_scc0 .Lgwb_%=\n\t" \
    "s_lshr_b32 s72, s68, 8\n\t" \
    "v_add_u32_e32 v102, s72, %[base]\n\t" \
    "ds_read_b128 v[88:91], v102\n\t" \
    ".Lgwb_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s56, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s57, v64, v0\n\t" \
    "v_fma_f32 v16, s57, v65, v16\n\t" \
    "v_fma_f32 v32, s57, v66, v32\n\t" \
    "v_fma_f32 v48, s57, v67, v48\n\t" \
    "s_cmp_gt_u32 s73, 1\n\t" \
    "s_cbranch_scc0 .Lgeb_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v68, v0\n\t" \
    "v_fma_f32 v16, s59, v69, v16\n\t" \
    "v_fma_f32 v32, s59, v70, v32\n\t" \
    "v_fma_f32 v48, s59, v71, v48\n\t" \
    "s_cmp_gt_u32 s73, 2\n\t" \
    "s_cbranch_scc0 .Lgeb_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v72, v0\n\t" \
    "v_fma_f32 v16, s61, v73, v16\n\t" \
    "v_fma_f32 v32, s61, v74, v32\n\t" \
    "v_fma_f32 v48, s61, v75, v48\n\t" \
    "s_cmp_gt_u32 s73, 3\n\t" \
    "s_cbranch_scc0 .Lgeb_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v76, v0\n\t" \
    "v_fma_f32 v16, s63, v77, v16\n\t" \
    "v_fma_f32 v32, s63, v78, v32\n\t" \
    "v_fma_f32 v48, s63, v79, v48\n\t" \
    "s_cmp_gt_u32 s73, 4\n\t" \
    "s_cbranch_scc0 .Lgeb_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v80, v0\n\t" \
    "v_fma_f32 v16, s65, v81, v16\n\t" \
    "v_fma_f32 v32, s65, v82, v32\n\t" \
    "v_fma_f32 v48, s65, v83, v48\n\t" \
    "s_cmp_gt_u32 s73, 5\n\t" \
    "s_cbranch_scc0 .Lgeb_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v84, v0\n\t" \
    "v_fma_f32 v16, s67, v85, v16\n\t" \
    "v_fma_f32 v32, s67, v86, v32\n\t" \
    "v_fma_f32 v48, s67, v87, v48\n\t" \
    "s_cmp_gt_u32 s73, 6\n\t" \
    "s_cbranch_scc0 .Lgeb_%=\n\t" \
    "s_set_gpr_idx_idx s68\n\t" \
    "v_fma_f32 v0, s69, v88, v0\n\t" \
    "v_fma_f32 v16, s69, v89, v16\n\t" \
    "v_fma_f32 v32, s69, v90, v32\n\t" \
    "v_fma_f32 v48, s69, v91, v48\n\t" \
    ".Lgeb_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    ".Lend_%=:\n\t" \
    "s_mov_b32 m0, s75\n\t"
#define GCOO_PANEL_CHUNK_CLOBBERS \
    "v64", "v65", "v66", "v67", "v68", "v69", "v70", "v71", "v72", "v73", "v74", "v75", \
    "v76", "v77", "v78", "v79", "v80", "v81", "v82", "v83", "v84", "v85", "v86", "v87", \
    "v88", "v89", "v90", "v91", "v92", "v93", "v94", "v95", "v96", "v97", "v98", "v99", \
    "v100", "v101", "v102", "v103", "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", \
    "s48", "s49", "s50", "s51", "s52", "s53", "s54", "s55", "s56", "s57", "s58", "s59", \
    "s60", "s61", "s62", "s63", "s64", "s65", "s66", "s67", "s68", "s69", "s70", "s71", \
    "s72", "s73", "s74", "s75", "s76", "scc", "memory"
